// Transpose-free QMR solve of (H - E S) X = B on the device, plain FP64 or FP32 with FP64 refinement.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/gfa.h"
#include "gfa_device.hpp"

namespace gfa {

int solve(gfa_plan* plan, gfa_solve_opts const& opts, double* rho, gfa_solve_stats* stats, std::string& err);
void solver_release(gfa_plan* plan);

// provided by gfa_capi.cpp
int capi_action(gfa_plan* p, int fp_bits, void* y, void const* x, hipStream_t stream, std::string& err);
Plan const& capi_host(gfa_plan const* p);
DevicePlan& capi_dev(gfa_plan* p);

// the random shadow residual: a SplitMix64 hash of (seed, element index) mapped to [-1, 1)
__host__ __device__ inline double shadow_value(uint64_t seed, uint64_t idx) {
    uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

}  // namespace gfa
