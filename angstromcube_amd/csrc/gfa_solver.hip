// Transpose-free QMR on the device.
//
// Algorithm: TFQMR of Freund (SIAM J. Sci. Comput. 14, 470 (1993)) in the form of Saad,
// "Iterative Methods for Sparse Linear Systems", 2nd ed., Algorithm 7.8, run independently for every
// right-hand-side column q = 64*iCol + j.  The reference delegates this to the external tfQMRgpu
// library (call site include/green_action.hxx:216, tfqmrgpu::solve(action, buffer, 1e-9, maxiter, 0, true)),
// which likewise keeps per-column scalars and FP64 inner products; its source is not available, so the
// iterate sequence is this file's own: a seeded random shadow residual, per-column convergence on the
// TFQMR residual bound tau*sqrt(m+1) <= tol*||b||, frozen columns once converged or broken down.
//
// All scalars live on the device; one half-step is a fixed sequence of launches with no host
// synchronisation, and the host polls the column status every few half-steps.  Inner products are
// reduced per block, then per column over fixed segments of blocks, so results are deterministic.
#include "gfa_solver.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace gfa {

namespace {

#define HIP_TRY(call)                                                              \
    do {                                                                           \
        hipError_t const e_ = (call);                                              \
        if (hipSuccess != e_) {                                                    \
            err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return GFA_ERR_DEVICE;                                                 \
        }                                                                          \
    } while (0)

constexpr int NBLK = 2 * 64 * 64;
constexpr int SEG = 32;  // blocks per reduction segment

// per-column scalar fields, stored field-major: S[field * ncol + q]
enum Field { RHO_RE, RHO_IM, ALPHA_RE, ALPHA_IM, THETA, TAU, ETA_RE, ETA_IM, CD_RE, CD_IM,
             BETA_RE, BETA_IM, BNORM, STATUS, TOL, NFIELDS };
enum Status { ACTIVE = 0, CONVERGED = 1, BREAKDOWN = 2 };

struct SolverWs {
    bool ready = false;
    uint64_t nnzb = 0;
    uint32_t ncol = 0;             // 64 * nCols
    uint32_t nseg = 0;
    // structure
    uint32_t* col_blocks = nullptr;  // blocks grouped by column
    uint32_t* seg = nullptr;         // [nseg][3]: column, begin, end into col_blocks
    uint32_t* col_seg = nullptr;     // [nCols+1] segment range per column
    uint8_t* diag = nullptr;         // [nnzb] 1 if the block is the diagonal block of its column
    // reductions
    double* part = nullptr;          // [nnzb][3][64]
    double* segpart = nullptr;       // [nseg][3][64]
    double* red = nullptr;           // [3][ncol]
    double* S = nullptr;             // [NFIELDS][ncol]
    double* S_host = nullptr;        // pinned copy of the status field
    double* res = nullptr;           // [ncol] true residual norms
    // vectors, per precision
    void* v64[12] = {nullptr};   // 0 x, 1-7 TFQMR work vectors, 8 x64, 9 y64, 10 rhs64
    void* v32[12] = {nullptr};   // 0 x, 1-7 TFQMR work vectors, 8 rhs32
    double* tol_abs = nullptr;   // [ncol] absolute per-column tolerances of refinement solves
    std::vector<void*> allocations;
};

template <typename real_t> struct Cplx { real_t re, im; };

// ------------------------------------------------------------------------------------------------
// reductions
// ------------------------------------------------------------------------------------------------
// combine the 4 partial sums of a column j held by threads j, j+64, j+128, j+192 and store them
template <int NC>
__device__ __forceinline__ void store_partials(double (&acc)[NC], double* __restrict__ part, uint64_t blk) {
    __shared__ double sh[NC][256];
    int const tid = threadIdx.x;
#pragma unroll
    for (int c = 0; c < NC; ++c) sh[c][tid] = acc[c];
    __syncthreads();
    if (tid < 64) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            double const s = (sh[c][tid] + sh[c][tid + 64]) + (sh[c][tid + 128] + sh[c][tid + 192]);
            part[(blk * 3 + c) * 64 + tid] = s;
        }
    }
}

__global__ __launch_bounds__(64) void k_seg_reduce(double* __restrict__ segpart, double const* __restrict__ part,
                                                   uint32_t const* __restrict__ seg, uint32_t const* __restrict__ col_blocks,
                                                   int nc) {
    uint32_t const s = blockIdx.x;
    uint32_t const b0 = seg[3 * s + 1], b1 = seg[3 * s + 2];
    int const j = threadIdx.x;
    for (int c = 0; c < nc; ++c) {
        double acc = 0;
        for (uint32_t b = b0; b < b1; ++b) acc += part[(uint64_t(col_blocks[b]) * 3 + c) * 64 + j];
        segpart[(uint64_t(s) * 3 + c) * 64 + j] = acc;
    }
}

__global__ __launch_bounds__(64) void k_col_reduce(double* __restrict__ red, double const* __restrict__ segpart,
                                                   uint32_t const* __restrict__ col_seg, uint32_t ncol, int nc) {
    uint32_t const col = blockIdx.x;
    int const j = threadIdx.x;
    for (int c = 0; c < nc; ++c) {
        double acc = 0;
        for (uint32_t s = col_seg[col]; s < col_seg[col + 1]; ++s) acc += segpart[(uint64_t(s) * 3 + c) * 64 + j];
        red[uint64_t(c) * ncol + col * 64 + j] = acc;
    }
}

// ------------------------------------------------------------------------------------------------
// element-wise kernels: one work-group per block, 256 threads, each thread 16 complex elements
// (i = 4k + tid/64, j = tid % 64) of one column q = 64*colindx[blk] + j
// ------------------------------------------------------------------------------------------------
#define ELEMENT_LOOP for (int k = 0; k < 16; ++k)
#define ELEM_RE(ptr) ptr[blk * NBLK + (4 * k + (tid >> 6)) * 64 + j]
#define ELEM_IM(ptr) ptr[blk * NBLK + 4096 + (4 * k + (tid >> 6)) * 64 + j]

__device__ __forceinline__ uint64_t elem_index(uint64_t blk, int p, int i, int j) {
    return ((blk * 2 + p) * 64 + i) * 64 + j;
}

// x = 0, d = 0, w = u = rhs; partials |rhs|^2, dot(rs, rhs)
template <typename real_t>
__global__ __launch_bounds__(256) void k_init(real_t* __restrict__ x, real_t* __restrict__ d, real_t* __restrict__ w,
                                              real_t* __restrict__ u, real_t const* __restrict__ rhs,
                                              uint8_t const* __restrict__ diag, double* __restrict__ part,
                                              uint64_t seed) {
    uint64_t const blk = blockIdx.x;
    int const tid = threadIdx.x, j = tid & 63;
    double acc[3] = {0, 0, 0};
    bool const dg = diag[blk];
    ELEMENT_LOOP {
        int const i = 4 * k + (tid >> 6);
        real_t br, bi;
        if (rhs) { br = ELEM_RE(rhs); bi = ELEM_IM(rhs); }
        else { br = (dg && i == j) ? real_t(1) : real_t(0); bi = 0; }
        ELEM_RE(x) = 0; ELEM_IM(x) = 0;
        ELEM_RE(d) = 0; ELEM_IM(d) = 0;
        ELEM_RE(w) = br; ELEM_IM(w) = bi;
        ELEM_RE(u) = br; ELEM_IM(u) = bi;
        double const sr = shadow_value(seed, elem_index(blk, 0, i, j)), si = shadow_value(seed, elem_index(blk, 1, i, j));
        acc[0] += double(br) * br + double(bi) * bi;
        acc[1] += sr * br + si * bi;   // conj(rs) * b
        acc[2] += sr * bi - si * br;
    }
    store_partials<3>(acc, part, blk);
}

// partial dot(rs, v)
template <typename real_t>
__global__ __launch_bounds__(256) void k_dot_shadow(real_t const* __restrict__ v, double* __restrict__ part, uint64_t seed) {
    uint64_t const blk = blockIdx.x;
    int const tid = threadIdx.x, j = tid & 63;
    double acc[2] = {0, 0};
    ELEMENT_LOOP {
        int const i = 4 * k + (tid >> 6);
        double const vr = ELEM_RE(v), vi = ELEM_IM(v);
        double const sr = shadow_value(seed, elem_index(blk, 0, i, j)), si = shadow_value(seed, elem_index(blk, 1, i, j));
        acc[0] += sr * vr + si * vi;
        acc[1] += sr * vi - si * vr;
    }
    store_partials<2>(acc, part, blk);
}

template <typename real_t>
__device__ __forceinline__ Cplx<real_t> coef(double const* __restrict__ S, uint32_t ncol, uint32_t q, int fre) {
    return Cplx<real_t>{real_t(S[uint64_t(fre) * ncol + q]), real_t(S[uint64_t(fre + 1) * ncol + q])};
}

// even half-step: un = u - alpha v; w -= alpha Au; partial |w|^2
template <typename real_t>
__global__ __launch_bounds__(256) void k_even(real_t* __restrict__ un, real_t const* __restrict__ u,
                                              real_t const* __restrict__ v, real_t* __restrict__ w,
                                              real_t const* __restrict__ Au, uint16_t const* __restrict__ colindx,
                                              double const* __restrict__ S, uint32_t ncol, double* __restrict__ part) {
    uint64_t const blk = blockIdx.x;
    int const tid = threadIdx.x, j = tid & 63;
    uint32_t const q = uint32_t(colindx[blk]) * 64 + j;
    auto const a = coef<real_t>(S, ncol, q, ALPHA_RE);
    double acc[1] = {0};
    ELEMENT_LOOP {
        real_t const ur = ELEM_RE(u), ui = ELEM_IM(u), vr = ELEM_RE(v), vi = ELEM_IM(v);
        ELEM_RE(un) = ur - (a.re * vr - a.im * vi);
        ELEM_IM(un) = ui - (a.re * vi + a.im * vr);
        real_t const yr = ELEM_RE(Au), yi = ELEM_IM(Au);
        real_t const wr = ELEM_RE(w) - (a.re * yr - a.im * yi);
        real_t const wi = ELEM_IM(w) - (a.re * yi + a.im * yr);
        ELEM_RE(w) = wr; ELEM_IM(w) = wi;
        acc[0] += double(wr) * wr + double(wi) * wi;
    }
    store_partials<1>(acc, part, blk);
}

// odd half-step, first part: w -= alpha Au; partials |w|^2 and dot(rs, w)
template <typename real_t>
__global__ __launch_bounds__(256) void k_odd_w(real_t* __restrict__ w, real_t const* __restrict__ Au,
                                               uint16_t const* __restrict__ colindx, double const* __restrict__ S,
                                               uint32_t ncol, double* __restrict__ part, uint64_t seed) {
    uint64_t const blk = blockIdx.x;
    int const tid = threadIdx.x, j = tid & 63;
    uint32_t const q = uint32_t(colindx[blk]) * 64 + j;
    auto const a = coef<real_t>(S, ncol, q, ALPHA_RE);
    double acc[3] = {0, 0, 0};
    ELEMENT_LOOP {
        int const i = 4 * k + (tid >> 6);
        real_t const yr = ELEM_RE(Au), yi = ELEM_IM(Au);
        real_t const wr = ELEM_RE(w) - (a.re * yr - a.im * yi);
        real_t const wi = ELEM_IM(w) - (a.re * yi + a.im * yr);
        ELEM_RE(w) = wr; ELEM_IM(w) = wi;
        acc[0] += double(wr) * wr + double(wi) * wi;
        double const sr = shadow_value(seed, elem_index(blk, 0, i, j)), si = shadow_value(seed, elem_index(blk, 1, i, j));
        acc[1] += sr * wr + si * wi;
        acc[2] += sr * wi - si * wr;
    }
    store_partials<3>(acc, part, blk);
}

// d = u + cd d; x += eta d
template <typename real_t>
__global__ __launch_bounds__(256) void k_dx(real_t* __restrict__ d, real_t* __restrict__ x, real_t const* __restrict__ u,
                                            uint16_t const* __restrict__ colindx, double const* __restrict__ S, uint32_t ncol) {
    uint64_t const blk = blockIdx.x;
    int const tid = threadIdx.x, j = tid & 63;
    uint32_t const q = uint32_t(colindx[blk]) * 64 + j;
    auto const c = coef<real_t>(S, ncol, q, CD_RE);
    auto const e = coef<real_t>(S, ncol, q, ETA_RE);
    ELEMENT_LOOP {
        real_t const dr0 = ELEM_RE(d), di0 = ELEM_IM(d);
        real_t const dr = ELEM_RE(u) + (c.re * dr0 - c.im * di0);
        real_t const di = ELEM_IM(u) + (c.re * di0 + c.im * dr0);
        ELEM_RE(d) = dr; ELEM_IM(d) = di;
        ELEM_RE(x) += e.re * dr - e.im * di;
        ELEM_IM(x) += e.re * di + e.im * dr;
    }
}

// un = w + beta u
template <typename real_t>
__global__ __launch_bounds__(256) void k_unext(real_t* __restrict__ un, real_t const* __restrict__ w, real_t const* __restrict__ u,
                                               uint16_t const* __restrict__ colindx, double const* __restrict__ S, uint32_t ncol) {
    uint64_t const blk = blockIdx.x;
    int const tid = threadIdx.x, j = tid & 63;
    uint32_t const q = uint32_t(colindx[blk]) * 64 + j;
    auto const b = coef<real_t>(S, ncol, q, BETA_RE);
    ELEMENT_LOOP {
        real_t const ur = ELEM_RE(u), ui = ELEM_IM(u);
        ELEM_RE(un) = ELEM_RE(w) + (b.re * ur - b.im * ui);
        ELEM_IM(un) = ELEM_IM(w) + (b.re * ui + b.im * ur);
    }
}

// v = Aun + beta (Au + beta v)
template <typename real_t>
__global__ __launch_bounds__(256) void k_vnext(real_t* __restrict__ v, real_t const* __restrict__ Aun, real_t const* __restrict__ Au,
                                               uint16_t const* __restrict__ colindx, double const* __restrict__ S, uint32_t ncol) {
    uint64_t const blk = blockIdx.x;
    int const tid = threadIdx.x, j = tid & 63;
    uint32_t const q = uint32_t(colindx[blk]) * 64 + j;
    auto const b = coef<real_t>(S, ncol, q, BETA_RE);
    ELEMENT_LOOP {
        real_t const vr = ELEM_RE(v), vi = ELEM_IM(v);
        real_t const tr = ELEM_RE(Au) + (b.re * vr - b.im * vi);
        real_t const ti = ELEM_IM(Au) + (b.re * vi + b.im * vr);
        ELEM_RE(v) = ELEM_RE(Aun) + (b.re * tr - b.im * ti);
        ELEM_IM(v) = ELEM_IM(Aun) + (b.re * ti + b.im * tr);
    }
}

// r = b - y (FP64); partial |r|^2; optionally rhs32 = float(r)
template <typename out_t>
__global__ __launch_bounds__(256) void k_residual(double const* __restrict__ y, uint8_t const* __restrict__ diag,
                                                  double* __restrict__ part, out_t* __restrict__ r_out) {
    uint64_t const blk = blockIdx.x;
    int const tid = threadIdx.x, j = tid & 63;
    bool const dg = diag[blk];
    double acc[1] = {0};
    ELEMENT_LOOP {
        int const i = 4 * k + (tid >> 6);
        double const rr = ((dg && i == j) ? 1.0 : 0.0) - ELEM_RE(y);
        double const ri = -ELEM_IM(y);
        acc[0] += rr * rr + ri * ri;
        if (r_out) { ELEM_RE(r_out) = out_t(rr); ELEM_IM(r_out) = out_t(ri); }
    }
    store_partials<1>(acc, part, blk);
}

// x64 += delta
template <typename real_t>
__global__ __launch_bounds__(256) void k_accumulate(double* __restrict__ x64, real_t const* __restrict__ delta) {
    uint64_t const blk = blockIdx.x;
    int const tid = threadIdx.x;
    for (int k = 0; k < 32; ++k) {
        uint64_t const e = blk * NBLK + k * 256 + tid;
        x64[e] += double(delta[e]);
    }
}

// ------------------------------------------------------------------------------------------------
// per-column scalar updates
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ double& F(double* S, uint32_t ncol, int f, uint32_t q) { return S[uint64_t(f) * ncol + q]; }

__device__ __forceinline__ void freeze(double* S, uint32_t ncol, uint32_t q) {
    F(S, ncol, ALPHA_RE, q) = 0; F(S, ncol, ALPHA_IM, q) = 0;
    F(S, ncol, ETA_RE, q) = 0; F(S, ncol, ETA_IM, q) = 0;
    F(S, ncol, CD_RE, q) = 0; F(S, ncol, CD_IM, q) = 0;
    F(S, ncol, BETA_RE, q) = 0; F(S, ncol, BETA_IM, q) = 0;
}

// after k_init: red = |rhs|^2, dot(rs, rhs)
__global__ void k_scalar_init(double* __restrict__ S, double const* __restrict__ red, uint32_t ncol,
                              double tol_rel, double const* __restrict__ tol_abs) {
    uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncol) return;
    double const bn = sqrt(red[q]);
    F(S, ncol, BNORM, q) = bn;
    F(S, ncol, TAU, q) = bn;
    F(S, ncol, THETA, q) = 0;
    F(S, ncol, RHO_RE, q) = red[ncol + q];
    F(S, ncol, RHO_IM, q) = red[2 * ncol + q];
    double const tol = fmax(tol_rel * bn, tol_abs ? tol_abs[q] : 0.0);
    F(S, ncol, TOL, q) = tol;
    freeze(S, ncol, q);
    // nothing to do for a column whose right-hand side already meets the tolerance
    F(S, ncol, STATUS, q) = (bn <= tol) ? CONVERGED : ACTIVE;
}

// even half-step: alpha = rho / dot(rs, v)
__global__ void k_scalar_alpha(double* __restrict__ S, double const* __restrict__ red, uint32_t ncol) {
    uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncol) return;
    if (F(S, ncol, STATUS, q) != ACTIVE) { freeze(S, ncol, q); return; }
    double const sr = red[q], si = red[ncol + q];
    double const den = sr * sr + si * si;
    if (!(den > 0)) { F(S, ncol, STATUS, q) = BREAKDOWN; freeze(S, ncol, q); return; }
    double const rr = F(S, ncol, RHO_RE, q), ri = F(S, ncol, RHO_IM, q);
    F(S, ncol, ALPHA_RE, q) = (rr * sr + ri * si) / den;
    F(S, ncol, ALPHA_IM, q) = (ri * sr - rr * si) / den;
}

// after the w update: theta, c, tau, eta, the d coefficient, convergence; for odd steps also rho, beta
__global__ void k_scalar_theta(double* __restrict__ S, double const* __restrict__ red, uint32_t ncol, int m, int odd) {
    uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncol) return;
    if (F(S, ncol, STATUS, q) != ACTIVE) { freeze(S, ncol, q); return; }
    double const ar = F(S, ncol, ALPHA_RE, q), ai = F(S, ncol, ALPHA_IM, q);
    double const th = F(S, ncol, THETA, q), er = F(S, ncol, ETA_RE, q), ei = F(S, ncol, ETA_IM, q);
    // cd = theta^2 eta / alpha with the previous theta, eta
    double const a2 = ar * ar + ai * ai;
    double cr = 0, ci = 0;
    if (a2 > 0) {
        double const tr = th * th * er, ti = th * th * ei;
        cr = (tr * ar + ti * ai) / a2;
        ci = (ti * ar - tr * ai) / a2;
    }
    double const tau = F(S, ncol, TAU, q);
    double const thn = sqrt(red[q]) / tau;
    double const c = 1.0 / sqrt(1.0 + thn * thn);
    double const taun = tau * thn * c;
    F(S, ncol, CD_RE, q) = cr; F(S, ncol, CD_IM, q) = ci;
    F(S, ncol, THETA, q) = thn;
    F(S, ncol, TAU, q) = taun;
    F(S, ncol, ETA_RE, q) = c * c * ar;
    F(S, ncol, ETA_IM, q) = c * c * ai;
    if (taun * sqrt(double(m + 2)) <= F(S, ncol, TOL, q)) F(S, ncol, STATUS, q) = CONVERGED;  // x still gets this step
    if (odd) {
        double const rnr = red[ncol + q], rni = red[2 * ncol + q];
        double const rr = F(S, ncol, RHO_RE, q), ri = F(S, ncol, RHO_IM, q);
        double const den = rr * rr + ri * ri;
        if (!(den > 0) || !(rnr * rnr + rni * rni > 0)) {
            if (F(S, ncol, STATUS, q) == ACTIVE) F(S, ncol, STATUS, q) = BREAKDOWN;
            F(S, ncol, BETA_RE, q) = 0; F(S, ncol, BETA_IM, q) = 0;
        } else {
            F(S, ncol, BETA_RE, q) = (rnr * rr + rni * ri) / den;
            F(S, ncol, BETA_IM, q) = (rni * rr - rnr * ri) / den;
        }
        F(S, ncol, RHO_RE, q) = rnr; F(S, ncol, RHO_IM, q) = rni;
    }
    if (!(taun == taun)) { F(S, ncol, STATUS, q) = BREAKDOWN; freeze(S, ncol, q); }
}

__global__ void k_sqrt_copy(double* __restrict__ out, double const* __restrict__ red, uint32_t n) {
    uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = sqrt(red[q]);
}

// rho[c][i] = G[subset[c]][*][i][i] / pi
__global__ void k_diag(double* __restrict__ out, double const* __restrict__ X, uint32_t const* __restrict__ subset) {
    uint32_t const c = blockIdx.x;
    int const i = threadIdx.x;
    uint64_t const b = subset[c];
    double const f = 0.31830988618379067154;  // 1/pi
    out[(uint64_t(c) * 64 + i) * 2 + 0] = X[(b * 2 + 0) * 4096 + i * 64 + i] * f;
    out[(uint64_t(c) * 64 + i) * 2 + 1] = X[(b * 2 + 1) * 4096 + i * 64 + i] * f;
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------------
struct SolverState {
    SolverWs ws;
    uint32_t* subset = nullptr;
};

namespace {

std::vector<std::pair<gfa_plan*, SolverState*>> g_states;

SolverState* state_of(gfa_plan* p, bool create) {
    for (auto& e : g_states) if (e.first == p) return e.second;
    if (!create) return nullptr;
    auto* s = new SolverState();
    g_states.emplace_back(p, s);
    return s;
}

template <typename T>
int walloc(SolverWs& ws, T*& ptr, size_t n, std::string& err) {
    void* q = nullptr;
    HIP_TRY(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    ws.allocations.push_back(q);
    ptr = (T*)q;
    return 0;
}

int setup(gfa_plan* plan, SolverState& st, std::string& err) {
    auto& ws = st.ws;
    if (ws.ready) return 0;
    Plan const& h = capi_host(plan);
    ws.nnzb = h.nnzb;
    ws.ncol = 64 * h.nCols;
    // blocks grouped by column, in ascending block order, and segments of SEG blocks
    std::vector<uint32_t> count(h.nCols + 1, 0), cb(h.nnzb), segs, col_seg(1, 0);
    for (uint64_t b = 0; b < h.nnzb; ++b) ++count[h.colindx[b] + 1];
    for (uint32_t c = 0; c < h.nCols; ++c) count[c + 1] += count[c];
    std::vector<uint32_t> fill(count.begin(), count.end() - 1);
    for (uint64_t b = 0; b < h.nnzb; ++b) cb[fill[h.colindx[b]]++] = uint32_t(b);
    for (uint32_t c = 0; c < h.nCols; ++c) {
        for (uint32_t b = count[c]; b < count[c + 1]; b += SEG) {
            segs.push_back(c); segs.push_back(b); segs.push_back(std::min(count[c + 1], b + SEG));
        }
        col_seg.push_back(uint32_t(segs.size() / 3));
    }
    ws.nseg = uint32_t(segs.size() / 3);
    std::vector<uint8_t> diag(h.nnzb, 0);
    for (uint32_t c = 0; c < h.nCols; ++c) diag[h.subset[c]] = 1;
    int s;
    if ((s = walloc(ws, ws.col_blocks, cb.size(), err))) return s;
    if ((s = walloc(ws, ws.seg, segs.size(), err))) return s;
    if ((s = walloc(ws, ws.col_seg, col_seg.size(), err))) return s;
    if ((s = walloc(ws, ws.diag, diag.size(), err))) return s;
    if ((s = walloc(ws, st.subset, h.subset.size(), err))) return s;
    HIP_TRY(hipMemcpy(ws.col_blocks, cb.data(), cb.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ws.seg, segs.data(), segs.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ws.col_seg, col_seg.data(), col_seg.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ws.diag, diag.data(), diag.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(st.subset, h.subset.data(), h.subset.size() * 4, hipMemcpyHostToDevice));
    if ((s = walloc(ws, ws.part, size_t(h.nnzb) * 3 * 64, err))) return s;
    if ((s = walloc(ws, ws.segpart, size_t(ws.nseg) * 3 * 64, err))) return s;
    if ((s = walloc(ws, ws.red, size_t(3) * ws.ncol, err))) return s;
    if ((s = walloc(ws, ws.S, size_t(NFIELDS) * ws.ncol, err))) return s;
    if ((s = walloc(ws, ws.res, ws.ncol, err))) return s;
    if ((s = walloc(ws, ws.tol_abs, ws.ncol, err))) return s;
    HIP_TRY(hipHostMalloc((void**)&ws.S_host, size_t(ws.ncol) * sizeof(double), hipHostMallocDefault));
    ws.ready = true;
    return 0;
}

template <typename real_t>
void** vecs(SolverWs& ws);
template <> void** vecs<double>(SolverWs& ws) { return ws.v64; }
template <> void** vecs<float>(SolverWs& ws) { return ws.v32; }

// allocate vectors lo..hi-1 of one precision (each nnzb blocks)
template <typename real_t>
int ensure_vectors(SolverWs& ws, int lo, int hi, std::string& err) {
    void** v = vecs<real_t>(ws);
    for (int k = lo; k < hi; ++k) {
        if (v[k]) continue;
        real_t* p = nullptr;
        if (int s = walloc(ws, p, size_t(ws.nnzb) * NBLK, err)) return s;
        v[k] = p;
    }
    return 0;
}

int reduce(SolverWs& ws, int nc, hipStream_t stream, std::string& err) {
    hipLaunchKernelGGL(k_seg_reduce, dim3(ws.nseg), dim3(64), 0, stream, ws.segpart, ws.part, ws.seg, ws.col_blocks, nc);
    hipLaunchKernelGGL(k_col_reduce, dim3(ws.ncol / 64), dim3(64), 0, stream, ws.red, ws.segpart, ws.col_seg, ws.ncol, nc);
    HIP_TRY(hipGetLastError());
    return 0;
}

struct TfqmrResult { int halfsteps = 0; int actions = 0; };

// TFQMR on real_t vectors for the right-hand side rhs (nullptr: unit columns); solution into x
template <typename real_t>
int tfqmr(gfa_plan* plan, SolverState& st, real_t* x, real_t const* rhs, double tol_rel, double const* tol_abs,
          int maxit, int check_every, uint64_t seed, hipStream_t stream, TfqmrResult& res, std::string& err) {
    auto& ws = st.ws;
    if (int s = ensure_vectors<real_t>(ws, 1, 8, err)) return s;
    void** V = vecs<real_t>(ws);
    // V[0] x is the caller's x; buffers: w, u0, u1, v, d, Aodd, Aeven
    real_t* w = (real_t*)V[1];
    real_t* uu[2] = {(real_t*)V[2], (real_t*)V[3]};
    real_t* v = (real_t*)V[4];
    real_t* d = (real_t*)V[5];
    real_t* Aodd = (real_t*)V[6];
    real_t* Aeven = (real_t*)V[7];
    int const fp = sizeof(real_t) == 8 ? 64 : 32;
    uint32_t const ncol = ws.ncol;
    dim3 const gb(uint32_t(ws.nnzb)), tb(256), gs((ncol + 255) / 256), ts(256);
    auto const& dev = capi_dev(plan);
    uint16_t const* colindx = dev.colindx;
    int cur = 0;

    hipLaunchKernelGGL((k_init<real_t>), gb, tb, 0, stream, x, d, w, uu[0], rhs, ws.diag, ws.part, seed);
    if (int s = reduce(ws, 3, stream, err)) return s;
    hipLaunchKernelGGL(k_scalar_init, gs, ts, 0, stream, ws.S, ws.red, ncol, tol_rel, tol_abs);
    if (int s = capi_action(plan, fp, v, uu[0], stream, err)) return s;
    res.actions++;
    real_t* Au_even = v;  // A u_0 = v_0
    for (int m = 0; m < maxit; ++m) {
        if ((m & 1) == 0) {
            hipLaunchKernelGGL((k_dot_shadow<real_t>), gb, tb, 0, stream, v, ws.part, seed);
            if (int s = reduce(ws, 2, stream, err)) return s;
            hipLaunchKernelGGL(k_scalar_alpha, gs, ts, 0, stream, ws.S, ws.red, ncol);
            hipLaunchKernelGGL((k_even<real_t>), gb, tb, 0, stream, uu[1 - cur], uu[cur], v, w, Au_even, colindx, ws.S, ncol, ws.part);
            if (int s = reduce(ws, 1, stream, err)) return s;
            hipLaunchKernelGGL(k_scalar_theta, gs, ts, 0, stream, ws.S, ws.red, ncol, m, 0);
            hipLaunchKernelGGL((k_dx<real_t>), gb, tb, 0, stream, d, x, uu[cur], colindx, ws.S, ncol);
            cur = 1 - cur;
        } else {
            if (int s = capi_action(plan, fp, Aodd, uu[cur], stream, err)) return s;
            res.actions++;
            hipLaunchKernelGGL((k_odd_w<real_t>), gb, tb, 0, stream, w, Aodd, colindx, ws.S, ncol, ws.part, seed);
            if (int s = reduce(ws, 3, stream, err)) return s;
            hipLaunchKernelGGL(k_scalar_theta, gs, ts, 0, stream, ws.S, ws.red, ncol, m, 1);
            hipLaunchKernelGGL((k_dx<real_t>), gb, tb, 0, stream, d, x, uu[cur], colindx, ws.S, ncol);
            hipLaunchKernelGGL((k_unext<real_t>), gb, tb, 0, stream, uu[1 - cur], w, uu[cur], colindx, ws.S, ncol);
            if (int s = capi_action(plan, fp, Aeven, uu[1 - cur], stream, err)) return s;
            res.actions++;
            hipLaunchKernelGGL((k_vnext<real_t>), gb, tb, 0, stream, v, Aeven, Aodd, colindx, ws.S, ncol);
            Au_even = Aeven;
            cur = 1 - cur;
        }
        HIP_TRY(hipGetLastError());
        res.halfsteps = m + 1;
        if ((m + 1) % check_every == 0 || m + 1 == maxit) {
            HIP_TRY(hipMemcpyAsync(ws.S_host, ws.S + size_t(STATUS) * ncol, ncol * sizeof(double), hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            bool any = false;
            for (uint32_t q = 0; q < ncol && !any; ++q) any = (ws.S_host[q] == ACTIVE);
            if (!any) break;
        }
    }
    return 0;
}

// true residual r = b - A x64 with the FP64 action; per-column norms into ws.res; optionally r as rhs32
template <typename out_t>
int residual(gfa_plan* plan, SolverState& st, double const* x64, double* y64, out_t* r_out, hipStream_t stream,
             int& actions, std::string& err) {
    auto& ws = st.ws;
    if (int s = capi_action(plan, 64, y64, x64, stream, err)) return s;
    actions++;
    hipLaunchKernelGGL((k_residual<out_t>), dim3(uint32_t(ws.nnzb)), dim3(256), 0, stream, y64, ws.diag, ws.part, r_out);
    if (int s = reduce(ws, 1, stream, err)) return s;
    hipLaunchKernelGGL(k_sqrt_copy, dim3((ws.ncol + 255) / 256), dim3(256), 0, stream, ws.res, ws.red, ws.ncol);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

void solver_release(gfa_plan* plan) {
    for (size_t k = 0; k < g_states.size(); ++k) {
        if (g_states[k].first != plan) continue;
        auto* s = g_states[k].second;
        for (void* p : s->ws.allocations) (void)hipFree(p);
        if (s->ws.S_host) (void)hipHostFree(s->ws.S_host);
        delete s;
        g_states.erase(g_states.begin() + k);
        return;
    }
}

int solve(gfa_plan* plan, gfa_solve_opts const& o, double* rho, gfa_solve_stats* stats, std::string& err) {
    if (o.fp_bits != 64 && o.fp_bits != 32) { err = "fp_bits must be 32 or 64"; return GFA_ERR_ARGUMENT; }
    if (!(o.tolerance > 0)) { err = "tolerance must be positive"; return GFA_ERR_ARGUMENT; }
    if (o.max_iterations < 0 || o.max_refinements < 0) { err = "negative iteration limits"; return GFA_ERR_ARGUMENT; }
    auto const t0 = std::chrono::steady_clock::now();
    SolverState& st = *state_of(plan, true);
    if (int s = setup(plan, st, err)) return s;
    auto& ws = st.ws;
    hipStream_t const stream = (hipStream_t)o.stream;
    int const check_every = o.check_every > 0 ? o.check_every : 8;
    uint64_t const seed = o.shadow_seed ? o.shadow_seed : 0x5EEDA43Cull;
    gfa_solve_stats S;
    std::memset(&S, 0, sizeof(S));
    S.total_columns = int32_t(ws.ncol);
    if (int s = ensure_vectors<double>(ws, 8, 10, err)) return s;
    double* x64 = (double*)ws.v64[8];   // the FP64 solution
    double* y64 = (double*)ws.v64[9];   // FP64 action output
    std::vector<double> resid(ws.ncol);
    auto fetch_resid = [&]() -> int {
        HIP_TRY(hipMemcpyAsync(resid.data(), ws.res, ws.ncol * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return 0;
    };
    auto max_resid = [&]() { double mx = 0; for (double v : resid) mx = std::max(mx, v); return mx; };
    {   // refinement solves stop a column once it is below half the final tolerance
        std::vector<double> tabs(ws.ncol, o.tolerance * 0.5);
        HIP_TRY(hipMemcpyAsync(ws.tol_abs, tabs.data(), ws.ncol * sizeof(double), hipMemcpyHostToDevice, stream));
    }

    if (o.fp_bits == 64) {
        if (int s = ensure_vectors<double>(ws, 0, 11, err)) return s;
        double* x = (double*)ws.v64[0];
        double* rhs64 = (double*)ws.v64[10];
        TfqmrResult r;
        if (int s = tfqmr<double>(plan, st, x, nullptr, o.tolerance, nullptr, o.max_iterations, check_every, seed, stream, r, err)) return s;
        S.iterations += r.halfsteps; S.actions += r.actions;
        HIP_TRY(hipMemcpyAsync(x64, x, ws.nnzb * NBLK * sizeof(double), hipMemcpyDeviceToDevice, stream));
        if (int s = residual<double>(plan, st, x64, y64, rhs64, stream, S.actions, err)) return s;
        if (int s = fetch_resid()) return s;
        if (o.echo > 0) std::fprintf(stderr, "# gfa_solve fp64: %d half-steps, max residual %.3e\n", r.halfsteps, max_resid());
        // FP64 refinement steps if requested and needed: solve A delta = r, x += delta
        for (int k = 0; k < o.max_refinements && max_resid() > o.tolerance; ++k) {
            TfqmrResult r2;
            if (int s = tfqmr<double>(plan, st, x, rhs64, 1e-30, ws.tol_abs, o.max_iterations, check_every, seed + k + 1, stream, r2, err)) return s;
            S.iterations += r2.halfsteps; S.actions += r2.actions; S.refinements++;
            hipLaunchKernelGGL((k_accumulate<double>), dim3(uint32_t(ws.nnzb)), dim3(256), 0, stream, x64, x);
            if (int s = residual<double>(plan, st, x64, y64, rhs64, stream, S.actions, err)) return s;
            if (int s = fetch_resid()) return s;
        }
    } else {
        if (int s = ensure_vectors<float>(ws, 0, 9, err)) return s;
        float* x32 = (float*)ws.v32[0];
        float* rhs32 = (float*)ws.v32[8];
        HIP_TRY(hipMemsetAsync(x64, 0, ws.nnzb * NBLK * sizeof(double), stream));
        double const floor_rel = o.inner_tolerance > 0 ? o.inner_tolerance : 1e-5;  // gain asked of one FP32 solve
        for (int k = 0; k <= o.max_refinements; ++k) {
            TfqmrResult r;
            if (int s = tfqmr<float>(plan, st, x32, k == 0 ? nullptr : rhs32, floor_rel, ws.tol_abs, o.max_iterations,
                                     check_every, seed + k, stream, r, err)) return s;
            S.iterations += r.halfsteps; S.actions += r.actions;
            if (k > 0) S.refinements++;
            hipLaunchKernelGGL((k_accumulate<float>), dim3(uint32_t(ws.nnzb)), dim3(256), 0, stream, x64, x32);
            if (int s = residual<float>(plan, st, x64, y64, rhs32, stream, S.actions, err)) return s;
            if (int s = fetch_resid()) return s;
            if (o.echo > 0)
                std::fprintf(stderr, "# gfa_solve fp32 step %d: %d half-steps, max FP64 residual %.3e\n", k, r.halfsteps, max_resid());
            if (max_resid() <= o.tolerance) break;
        }
    }
    S.max_residual = 0;
    S.converged_columns = 0;
    for (double v : resid) {
        S.max_residual = std::max(S.max_residual, v);
        S.converged_columns += (v <= o.tolerance);
    }
    if (o.X_device) HIP_TRY(hipMemcpyAsync(o.X_device, x64, ws.nnzb * NBLK * sizeof(double), hipMemcpyDeviceToDevice, stream));
    if (rho) {
        double* drho = ws.red;  // 3*ncol doubles >= 2*ncol
        hipLaunchKernelGGL(k_diag, dim3(ws.ncol / 64), dim3(64), 0, stream, drho, x64, st.subset);
        HIP_TRY(hipMemcpyAsync(rho, drho, size_t(ws.ncol) * 2 * sizeof(double), hipMemcpyDeviceToHost, stream));
    }
    HIP_TRY(hipStreamSynchronize(stream));
    S.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    S.flops = double(S.actions) * reference_flops(capi_host(plan), o.fp_bits);
    if (stats) *stats = S;
    return 0;
}

}  // namespace gfa
