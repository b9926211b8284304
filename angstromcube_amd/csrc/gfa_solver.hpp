// Transpose-free QMR solve of (H - E S) X = B on the device, plain FP64 or FP32 with FP64 refinement.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/gfa.h"
#include "gfa_device.hpp"

namespace gfa {

int solve(gfa_plan* plan, gfa_solve_opts const& opts, double* rho, gfa_solve_stats* stats, std::string& err);
// frees the solver state of a plan (work vectors, solution, captured graph); the next solve allocates again
void solver_release(gfa_plan* plan);
// device bytes of the solver state of a solve with these options at its largest (FP32 solves: including the FP64
// work vectors of a column group that falls back to FP64)
uint64_t solve_bytes(gfa_plan* plan, int fp_bits, int method);
// device bytes the solver state of a plan holds now
uint64_t solver_held_bytes(gfa_plan* plan);

// provided by gfa_capi.cpp
int capi_action(gfa_plan* p, int fp_bits, void* y, void const* x, hipStream_t stream, std::string& err,
                int parts = ACTION_ALL, double* dot = nullptr);
int capi_autotune(gfa_plan* p, int fp_bits, void* y, void const* x, hipStream_t stream, int reps, double* best_ms,
                  std::string& err);
Plan const& capi_host(gfa_plan const* p);
DevicePlan& capi_dev(gfa_plan* p);

// The random shadow residual r0*: the murmur3 32-bit finaliser (Appleby, MurmurHash3) of a mix of the
// seed and the element index, its top 24 bits as a signed fraction in [-1, 1).  The values are multiples
// of 2^-23, exact in FP32 and FP64, so both precisions use the same shadow vector.
__host__ __device__ inline double shadow_value(uint64_t seed, uint64_t idx) {
    uint32_t h = uint32_t(idx) + 0x9E3779B9u * (uint32_t(idx >> 32) + 1u) + uint32_t(seed) * 0x27D4EB2Fu +
                 uint32_t(seed >> 32);
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return double(int32_t(h) >> 8) * (1.0 / 8388608.0);
}

}  // namespace gfa
