// Transpose-free QMR on the device.
//
// Algorithm: TFQMR of Freund (SIAM J. Sci. Comput. 14, 470 (1993)) in the form of Saad,
// "Iterative Methods for Sparse Linear Systems", 2nd ed., Algorithm 7.8, run independently for every
// right-hand-side column q = 64*iCol + j.  The reference delegates this to the external tfQMRgpu
// library (call site include/green_action.hxx:216, tfqmrgpu::solve(action, buffer, 1e-9, maxiter, 0, true)),
// which likewise keeps per-column scalars and FP64 inner products; its source is not available, so the
// iterate sequence is this file's own: a seeded random shadow residual, per-column convergence on the
// TFQMR residual bound tau*sqrt(m+2) <= tol*||b||, frozen columns once converged or broken down.
//
// Layout of the work: one full TFQMR step (an even and an odd half-step, two actions) is
//     alpha | K_even | theta_e | A un | K_odd | theta_o | K_dx2 | A u | K_vnext
// where the K_* kernels stream the vectors of one segment (<= SEG blocks of one column group) per
// work-group and leave their inner-product partials per segment; the scalar kernels fold the segments of
// a column in a fixed order, so every reduction is deterministic.  The two x updates of a step are
// deferred into one pass (K_dx2), the shadow product of v is fused into its update (K_vnext).
// The solution x is always accumulated in FP64, also when the vectors are FP32: the FP32 rounding of x
// would otherwise limit the true residual to about eps32 * ||A|| * ||x||.
//
// Retirement: a column group (one source block, 64 columns) whose columns have all converged leaves the
// active set; the vector kernels, reductions and the action then skip its blocks.  The iteration sequence
// of the active set is captured once as a HIP graph and replayed; the half-step counter lives on the
// device so that replays need no new arguments.  The host polls the column status every few half-steps.
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
#ifndef VEC_NT
#define VEC_NT 1
#endif
#ifndef SEG_BLOCKS
#define SEG_BLOCKS 16
#endif
constexpr int SEG = SEG_BLOCKS;  // blocks per segment (work-group of the vector kernels)

// per-column scalar fields, stored field-major: S[field * ncol + q]
enum Field { RHO_RE, RHO_IM, ALPHA_RE, ALPHA_IM, THETA, TAU, ETAE_RE, ETAE_IM, ETAO_RE, ETAO_IM,
             CDE_RE, CDE_IM, CDO_RE, CDO_IM, BETA_RE, BETA_IM, BNORM, STATUS, TOL, MSTEP,
             APREV_RE, APREV_IM,   // COCG: alpha of the last even iteration, whose x update is deferred to the odd one
             RN,                   // COCG: the recursive residual norm of the last iteration relative to |b|
             NFIELDS };
enum Status { ACTIVE = 0, CONVERGED = 1, BREAKDOWN = 2 };

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
template <typename T> struct Pk;
template <> struct Pk<float> { static constexpr int V = 4; typedef f4 t; };
template <> struct Pk<double> { static constexpr int V = 2; typedef d2 t; };

// thread mapping of the vector kernels: 256 threads, each V consecutive columns j0..j0+V-1 of rows
// i = pass*RPP + r, for both planes
template <typename T> struct Map {
    static constexpr int V = Pk<T>::V;
    static constexpr int JG = 64 / V;      // column groups
    static constexpr int RPP = 256 / JG;   // rows per pass
    static constexpr int NPASS = 64 / RPP;
};

// the vectors are streamed once per kernel: non-temporal 16-byte accesses (no reuse in L2 within a kernel)
template <typename T>
__device__ __forceinline__ void ld(T (&a)[Pk<T>::V], T const* p) {
#if VEC_NT
    typename Pk<T>::t const v = __builtin_nontemporal_load(reinterpret_cast<typename Pk<T>::t const*>(p));
#else
    typename Pk<T>::t const v = *reinterpret_cast<typename Pk<T>::t const*>(p);
#endif
#pragma unroll
    for (int t = 0; t < Pk<T>::V; ++t) a[t] = v[t];
}
template <typename T>
__device__ __forceinline__ void st(T* p, T const (&a)[Pk<T>::V]) {
    typename Pk<T>::t v;
#pragma unroll
    for (int t = 0; t < Pk<T>::V; ++t) v[t] = a[t];
#if VEC_NT
    __builtin_nontemporal_store(v, reinterpret_cast<typename Pk<T>::t*>(p));
#else
    *reinterpret_cast<typename Pk<T>::t*>(p) = v;
#endif
}
// V doubles (V = 2 or 4) as 16-byte pieces
template <int V>
__device__ __forceinline__ void ldd(double (&a)[V], double const* p) {
#pragma unroll
    for (int h = 0; h < V / 2; ++h) {
        d2 const v = *reinterpret_cast<d2 const*>(p + 2 * h);
        a[2 * h] = v[0]; a[2 * h + 1] = v[1];
    }
}
template <int V>
__device__ __forceinline__ void std_(double* p, double const (&a)[V]) {
#pragma unroll
    for (int h = 0; h < V / 2; ++h) {
        d2 v; v[0] = a[2 * h]; v[1] = a[2 * h + 1];
        *reinterpret_cast<d2*>(p + 2 * h) = v;
    }
}

struct SegArgs {
    uint32_t const* act;         // [grid] active segments (global segment numbers)
    uint32_t const* seg;         // [nseg][3]: column group, begin, end into col_blocks
    uint32_t const* col_blocks;  // blocks grouped by column group
    double* segpart;             // [nseg][3][64], indexed by global segment number
};

// fold the per-thread partials acc[c][t] (columns j0+t) of a work-group into segpart[s][c][j]
template <typename T, int NC>
__device__ __forceinline__ void store_partials(double (&acc)[NC][Pk<T>::V], double* __restrict__ segpart, uint32_t s) {
    using M = Map<T>;
    __shared__ double sh[NC][M::RPP][64];
    int const tid = threadIdx.x, r = tid / M::JG, j0 = (tid % M::JG) * M::V;
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int t = 0; t < M::V; ++t) sh[c][r][j0 + t] = acc[c][t];
    __syncthreads();
    if (tid < 64) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            double sum = 0;
            for (int k = 0; k < M::RPP; ++k) sum += sh[c][k][tid];
            segpart[(uint64_t(s) * 3 + c) * 64 + tid] = sum;
        }
    }
}

__device__ __forceinline__ uint64_t elem_index(uint64_t blk, int p, int i, int j) {
    return ((blk * 2 + p) * 64 + i) * 64 + j;
}

template <typename T> struct Cx { T re, im; };

template <typename T, int N>
__device__ __forceinline__ void coefs(Cx<T> (&c)[N], double const* __restrict__ S, uint32_t ncol, uint32_t q0, int fre) {
#pragma unroll
    for (int t = 0; t < N; ++t) {
        c[t].re = T(S[uint64_t(fre) * ncol + q0 + t]);
        c[t].im = T(S[uint64_t(fre + 1) * ncol + q0 + t]);
    }
}

#define SEG_PROLOGUE                                                                     \
    using M = Map<T>;                                                                    \
    constexpr int V = M::V;                                                              \
    uint32_t const s = g.act[blockIdx.x];                                                \
    uint32_t const col = g.seg[3 * s], b0 = g.seg[3 * s + 1], b1 = g.seg[3 * s + 2];     \
    int const tid = threadIdx.x, r = tid / M::JG, j0 = (tid % M::JG) * V;                \
    uint32_t const q0 = col * 64 + j0;                                                   \
    (void)q0; (void)r;
#define OFF(p, i) ((uint64_t(blk) * 2 + (p)) * 4096 + uint64_t(i) * 64 + j0)

// the right-hand side of an inner solve at the thread's V columns of row i: the unit columns on the diagonal blocks
// minus y = A x (the FP64 action on the current FP64 solution, so b - A x is the refinement residual), or the unit
// columns alone if y is null, rounded to T
template <typename T, int V>
__device__ __forceinline__ void rhs_values(T (&br)[V], T (&bi)[V], double const* __restrict__ y, uint64_t o_re,
                                           uint64_t o_im, bool dg, int i, int j0) {
    if (y) {
        double yr[V], yi[V];
        ldd<V>(yr, y + o_re); ldd<V>(yi, y + o_im);
#pragma unroll
        for (int t = 0; t < V; ++t) { br[t] = T(((dg && i == j0 + t) ? 1.0 : 0.0) - yr[t]); bi[t] = T(-yi[t]); }
    } else {
#pragma unroll
        for (int t = 0; t < V; ++t) { br[t] = (dg && i == j0 + t) ? T(1) : T(0); bi[t] = 0; }
    }
}

// x untouched; d = 0, w = u = b (see rhs_values); partials |b|^2 and (r0*, b)
template <typename T>
__global__ __launch_bounds__(256) void k_init(T* __restrict__ d, T* __restrict__ w, T* __restrict__ u,
                                              double const* __restrict__ y, uint8_t const* __restrict__ diag,
                                              SegArgs g, uint64_t seed) {
    SEG_PROLOGUE
    double acc[3][V] = {};
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
        bool const dg = diag[blk];
#pragma unroll 2
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            T br[V], bi[V], z[V];
            rhs_values<T, V>(br, bi, y, OFF(0, i), OFF(1, i), dg, i, j0);
#pragma unroll
            for (int t = 0; t < V; ++t) z[t] = 0;
            st(d + OFF(0, i), z); st(d + OFF(1, i), z);
            st(w + OFF(0, i), br); st(w + OFF(1, i), bi);
            st(u + OFF(0, i), br); st(u + OFF(1, i), bi);
#pragma unroll
            for (int t = 0; t < V; ++t) {
                double const sr = shadow_value(seed, elem_index(blk, 0, i, j0 + t));
                double const si = shadow_value(seed, elem_index(blk, 1, i, j0 + t));
                double const xr = br[t], xi = bi[t];
                acc[0][t] += xr * xr + xi * xi;
                acc[1][t] += sr * xr + si * xi;   // conj(r0*) b
                acc[2][t] += sr * xi - si * xr;
            }
        }
    }
    store_partials<T, 3>(acc, g.segpart, s);
}

// partial (r0*, v)
template <typename T>
__global__ __launch_bounds__(256) void k_dot_shadow(T const* __restrict__ v, SegArgs g, uint64_t seed) {
    SEG_PROLOGUE
    double acc[2][V] = {};
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
#pragma unroll 2
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            T vr[V], vi[V];
            ld(vr, v + OFF(0, i)); ld(vi, v + OFF(1, i));
#pragma unroll
            for (int t = 0; t < V; ++t) {
                double const sr = shadow_value(seed, elem_index(blk, 0, i, j0 + t));
                double const si = shadow_value(seed, elem_index(blk, 1, i, j0 + t));
                acc[0][t] += sr * vr[t] + si * vi[t];
                acc[1][t] += sr * vi[t] - si * vr[t];
            }
        }
    }
    store_partials<T, 2>(acc, g.segpart, s);
}

// even half-step: un = u - alpha v; w -= alpha Au; partial |w|^2
template <typename T>
__global__ __launch_bounds__(256) void k_even(T* __restrict__ un, T const* __restrict__ u, T const* __restrict__ v,
                                              T* __restrict__ w, T const* __restrict__ Au, double const* __restrict__ S,
                                              uint32_t ncol, SegArgs g) {
    SEG_PROLOGUE
    Cx<T> a[V];
    coefs(a, S, ncol, q0, ALPHA_RE);
    double acc[1][V] = {};
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
#pragma unroll 2
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            T ur[V], ui[V], vr[V], vi[V], wr[V], wi[V], yr[V], yi[V];
            ld(ur, u + OFF(0, i)); ld(ui, u + OFF(1, i));
            ld(vr, v + OFF(0, i)); ld(vi, v + OFF(1, i));
            ld(wr, w + OFF(0, i)); ld(wi, w + OFF(1, i));
            ld(yr, Au + OFF(0, i)); ld(yi, Au + OFF(1, i));
#pragma unroll
            for (int t = 0; t < V; ++t) {
                ur[t] = ur[t] - (a[t].re * vr[t] - a[t].im * vi[t]);
                ui[t] = ui[t] - (a[t].re * vi[t] + a[t].im * vr[t]);
                wr[t] = wr[t] - (a[t].re * yr[t] - a[t].im * yi[t]);
                wi[t] = wi[t] - (a[t].re * yi[t] + a[t].im * yr[t]);
                acc[0][t] += double(wr[t]) * wr[t] + double(wi[t]) * wi[t];
            }
            st(un + OFF(0, i), ur); st(un + OFF(1, i), ui);
            st(w + OFF(0, i), wr); st(w + OFF(1, i), wi);
        }
    }
    store_partials<T, 1>(acc, g.segpart, s);
}

// odd half-step: w -= alpha Au; partials |w|^2 and (r0*, w)
template <typename T>
__global__ __launch_bounds__(256) void k_odd(T* __restrict__ w, T const* __restrict__ Au, double const* __restrict__ S,
                                             uint32_t ncol, SegArgs g, uint64_t seed) {
    SEG_PROLOGUE
    Cx<T> a[V];
    coefs(a, S, ncol, q0, ALPHA_RE);
    double acc[3][V] = {};
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
#pragma unroll 2
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            T wr[V], wi[V], yr[V], yi[V];
            ld(wr, w + OFF(0, i)); ld(wi, w + OFF(1, i));
            ld(yr, Au + OFF(0, i)); ld(yi, Au + OFF(1, i));
#pragma unroll
            for (int t = 0; t < V; ++t) {
                wr[t] = wr[t] - (a[t].re * yr[t] - a[t].im * yi[t]);
                wi[t] = wi[t] - (a[t].re * yi[t] + a[t].im * yr[t]);
                double const sr = shadow_value(seed, elem_index(blk, 0, i, j0 + t));
                double const si = shadow_value(seed, elem_index(blk, 1, i, j0 + t));
                double const xr = wr[t], xi = wi[t];
                acc[0][t] += xr * xr + xi * xi;
                acc[1][t] += sr * xr + si * xi;
                acc[2][t] += sr * xi - si * xr;
            }
            st(w + OFF(0, i), wr); st(w + OFF(1, i), wi);
        }
    }
    store_partials<T, 3>(acc, g.segpart, s);
}

// both x updates of a step and the next u:
//   d1 = u + cdE d; x += etaE d1; d2 = un + cdO d1; x += etaO d2; d = d2; u = w + beta un
template <typename T>
__global__ __launch_bounds__(256) void k_dx2(T* __restrict__ d, double* __restrict__ x, T* __restrict__ u,
                                             T const* __restrict__ un, T const* __restrict__ w,
                                             double const* __restrict__ S, uint32_t ncol, SegArgs g) {
    SEG_PROLOGUE
    Cx<T> ce[V], co[V], be[V];
    Cx<double> ee[V], eo[V];
    coefs(ce, S, ncol, q0, CDE_RE);
    coefs(co, S, ncol, q0, CDO_RE);
    coefs(be, S, ncol, q0, BETA_RE);
    coefs(ee, S, ncol, q0, ETAE_RE);
    coefs(eo, S, ncol, q0, ETAO_RE);
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
#pragma unroll 2
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            T dr[V], di[V], ur[V], ui[V], nr[V], ni[V], wr[V], wi[V];
            double xr[V], xi[V];
            ld(dr, d + OFF(0, i)); ld(di, d + OFF(1, i));
            ld(ur, u + OFF(0, i)); ld(ui, u + OFF(1, i));
            ld(nr, un + OFF(0, i)); ld(ni, un + OFF(1, i));
            ld(wr, w + OFF(0, i)); ld(wi, w + OFF(1, i));
            ldd<V>(xr, x + OFF(0, i)); ldd<V>(xi, x + OFF(1, i));
#pragma unroll
            for (int t = 0; t < V; ++t) {
                T const d1r = ur[t] + (ce[t].re * dr[t] - ce[t].im * di[t]);
                T const d1i = ui[t] + (ce[t].re * di[t] + ce[t].im * dr[t]);
                xr[t] += ee[t].re * double(d1r) - ee[t].im * double(d1i);
                xi[t] += ee[t].re * double(d1i) + ee[t].im * double(d1r);
                T const d2r = nr[t] + (co[t].re * d1r - co[t].im * d1i);
                T const d2i = ni[t] + (co[t].re * d1i + co[t].im * d1r);
                xr[t] += eo[t].re * double(d2r) - eo[t].im * double(d2i);
                xi[t] += eo[t].re * double(d2i) + eo[t].im * double(d2r);
                dr[t] = d2r; di[t] = d2i;
                ur[t] = wr[t] + (be[t].re * nr[t] - be[t].im * ni[t]);
                ui[t] = wi[t] + (be[t].re * ni[t] + be[t].im * nr[t]);
            }
            st(d + OFF(0, i), dr); st(d + OFF(1, i), di);
            st(u + OFF(0, i), ur); st(u + OFF(1, i), ui);
            std_<V>(x + OFF(0, i), xr); std_<V>(x + OFF(1, i), xi);
        }
    }
}

// v = Au + beta (Aun_prev + beta v); partial (r0*, v)
template <typename T>
__global__ __launch_bounds__(256) void k_vnext(T* __restrict__ v, T const* __restrict__ Au, T const* __restrict__ Aodd,
                                               double const* __restrict__ S, uint32_t ncol, SegArgs g, uint64_t seed) {
    SEG_PROLOGUE
    Cx<T> be[V];
    coefs(be, S, ncol, q0, BETA_RE);
    double acc[2][V] = {};
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
#pragma unroll 2
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            T vr[V], vi[V], ar[V], ai[V], or_[V], oi[V];
            ld(vr, v + OFF(0, i)); ld(vi, v + OFF(1, i));
            ld(ar, Au + OFF(0, i)); ld(ai, Au + OFF(1, i));
            ld(or_, Aodd + OFF(0, i)); ld(oi, Aodd + OFF(1, i));
#pragma unroll
            for (int t = 0; t < V; ++t) {
                T const tr = or_[t] + (be[t].re * vr[t] - be[t].im * vi[t]);
                T const ti = oi[t] + (be[t].re * vi[t] + be[t].im * vr[t]);
                vr[t] = ar[t] + (be[t].re * tr - be[t].im * ti);
                vi[t] = ai[t] + (be[t].re * ti + be[t].im * tr);
                double const sr = shadow_value(seed, elem_index(blk, 0, i, j0 + t));
                double const si = shadow_value(seed, elem_index(blk, 1, i, j0 + t));
                acc[0][t] += sr * vr[t] + si * vi[t];
                acc[1][t] += sr * vi[t] - si * vr[t];
            }
            st(v + OFF(0, i), vr); st(v + OFF(1, i), vi);
        }
    }
    store_partials<T, 2>(acc, g.segpart, s);
}

// x = 0 over the blocks of the active segments
template <typename T>
__global__ __launch_bounds__(256) void k_zero_seg(double* __restrict__ x, SegArgs g) {
    SEG_PROLOGUE
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            double z[V];
#pragma unroll
            for (int t = 0; t < V; ++t) z[t] = 0;
            std_<V>(x + OFF(0, i), z); std_<V>(x + OFF(1, i), z);
        }
    }
}

// partial |b - y|^2 (FP64, b = unit columns on the diagonal blocks, y = A x)
template <typename T>
__global__ __launch_bounds__(256) void k_residual(double const* __restrict__ y, uint8_t const* __restrict__ diag, SegArgs g) {
    SEG_PROLOGUE
    double acc[1][V] = {};
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
        bool const dg = diag[blk];
#pragma unroll 2
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            double yr[V], yi[V];
            ldd<V>(yr, y + OFF(0, i)); ldd<V>(yi, y + OFF(1, i));
#pragma unroll
            for (int t = 0; t < V; ++t) {
                double const rr = ((dg && i == j0 + t) ? 1.0 : 0.0) - yr[t], ri = -yi[t];
                acc[0][t] += rr * rr + ri * ri;
            }
        }
    }
    store_partials<T, 1>(acc, g.segpart, s);
}

// ------------------------------------------------------------------------------------------------
// COCG (conjugate orthogonal CG for complex-symmetric A, van der Vorst & Melissen, IEEE Trans. Magn. 26, 706
// (1990)): CG with the bilinear form (a, b) = sum a_i b_i.  Vectors r, p, q = A p; the solution in FP64.
// ------------------------------------------------------------------------------------------------
// r = p = b (see rhs_values); partials |b|^2, Re (b, b), Im (b, b)
template <typename T>
__global__ __launch_bounds__(256) void k_cocg_init(T* __restrict__ rv, T* __restrict__ p, double const* __restrict__ y,
                                                   uint8_t const* __restrict__ diag, SegArgs g) {
    SEG_PROLOGUE
    double acc[3][V] = {};
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
        bool const dg = diag[blk];
#pragma unroll 2
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            T br[V], bi[V];
            rhs_values<T, V>(br, bi, y, OFF(0, i), OFF(1, i), dg, i, j0);
            st(rv + OFF(0, i), br); st(rv + OFF(1, i), bi);
            st(p + OFF(0, i), br); st(p + OFF(1, i), bi);
#pragma unroll
            for (int t = 0; t < V; ++t) {
                double const xr = br[t], xi = bi[t];
                acc[0][t] += xr * xr + xi * xi;
                acc[1][t] += xr * xr - xi * xi;
                acc[2][t] += 2 * xr * xi;
            }
        }
    }
    store_partials<T, 3>(acc, g.segpart, s);
}

// partials Re (p, q), Im (p, q)
template <typename T>
__global__ __launch_bounds__(256) void k_cocg_pq(T const* __restrict__ p, T const* __restrict__ q, SegArgs g) {
    SEG_PROLOGUE
    double acc[2][V] = {};
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
#pragma unroll 2
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            T pr[V], pi[V], qr[V], qi[V];
            ld(pr, p + OFF(0, i)); ld(pi, p + OFF(1, i));
            ld(qr, q + OFF(0, i)); ld(qi, q + OFF(1, i));
#pragma unroll
            for (int t = 0; t < V; ++t) {
                acc[0][t] += double(pr[t]) * qr[t] - double(pi[t]) * qi[t];
                acc[1][t] += double(pr[t]) * qi[t] + double(pi[t]) * qr[t];
            }
        }
    }
    store_partials<T, 2>(acc, g.segpart, s);
}

// r -= alpha q; partials |r|^2, Re (r, r), Im (r, r)
template <typename T>
__global__ __launch_bounds__(256) void k_cocg_r(T* __restrict__ rv, T const* __restrict__ q, double const* __restrict__ S,
                                                uint32_t ncol, SegArgs g) {
    SEG_PROLOGUE
    Cx<T> a[V];
    coefs(a, S, ncol, q0, ALPHA_RE);
    double acc[3][V] = {};
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
#pragma unroll 2
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            T qr[V], qi[V], rr[V], ri[V];
            ld(qr, q + OFF(0, i)); ld(qi, q + OFF(1, i));
            ld(rr, rv + OFF(0, i)); ld(ri, rv + OFF(1, i));
#pragma unroll
            for (int t = 0; t < V; ++t) {
                rr[t] = rr[t] - (a[t].re * qr[t] - a[t].im * qi[t]);
                ri[t] = ri[t] - (a[t].re * qi[t] + a[t].im * qr[t]);
                double const vr = rr[t], vi = ri[t];
                acc[0][t] += vr * vr + vi * vi;
                acc[1][t] += vr * vr - vi * vi;
                acc[2][t] += 2 * vr * vi;
            }
            st(rv + OFF(0, i), rr); st(rv + OFF(1, i), ri);
        }
    }
    store_partials<T, 3>(acc, g.segpart, s);
}

// COCG keeps two search directions P0, P1 and updates x every second iteration: an even iteration (p in P0) forms
// P1 = r + beta P0 and leaves x alone (its alpha is kept in APREV); the odd one (p in P1) adds both steps,
// x += alpha_prev P0 + alpha P1, and forms P0 = r + beta P1.  That reads and writes x once per two iterations: 4.5
// vector passes per iteration instead of 5.  x is the solution itself (FP64) or, for FP32 vectors, the FP32
// increment since the last fold into the FP64 solution.
template <typename T>
__global__ __launch_bounds__(256) void k_cocg_p_even(T* __restrict__ pn, T const* __restrict__ po,
                                                     T const* __restrict__ rv, double const* __restrict__ S,
                                                     uint32_t ncol, SegArgs g) {
    SEG_PROLOGUE
    Cx<T> be[V];
    coefs(be, S, ncol, q0, BETA_RE);
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
#pragma unroll 2
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            T pr[V], pi[V], rr[V], ri[V];
            ld(pr, po + OFF(0, i)); ld(pi, po + OFF(1, i));
            ld(rr, rv + OFF(0, i)); ld(ri, rv + OFF(1, i));
#pragma unroll
            for (int t = 0; t < V; ++t) {
                T const nr = rr[t] + (be[t].re * pr[t] - be[t].im * pi[t]);
                T const ni = ri[t] + (be[t].re * pi[t] + be[t].im * pr[t]);
                pr[t] = nr; pi[t] = ni;
            }
            st(pn + OFF(0, i), pr); st(pn + OFF(1, i), pi);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_cocg_px_odd(T* __restrict__ x, T* __restrict__ p0, T const* __restrict__ p1,
                                                     T const* __restrict__ rv, double const* __restrict__ S,
                                                     uint32_t ncol, SegArgs g) {
    SEG_PROLOGUE
    Cx<T> ap[V], a[V], be[V];
    coefs(ap, S, ncol, q0, APREV_RE);
    coefs(a, S, ncol, q0, ALPHA_RE);
    coefs(be, S, ncol, q0, BETA_RE);
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
#pragma unroll 2
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            T or_[V], oi[V], pr[V], pi[V], rr[V], ri[V], xr[V], xi[V];
            ld(or_, p0 + OFF(0, i)); ld(oi, p0 + OFF(1, i));
            ld(pr, p1 + OFF(0, i)); ld(pi, p1 + OFF(1, i));
            ld(rr, rv + OFF(0, i)); ld(ri, rv + OFF(1, i));
            ld(xr, x + OFF(0, i)); ld(xi, x + OFF(1, i));
#pragma unroll
            for (int t = 0; t < V; ++t) {
                xr[t] += ap[t].re * or_[t] - ap[t].im * oi[t];
                xi[t] += ap[t].re * oi[t] + ap[t].im * or_[t];
                xr[t] += a[t].re * pr[t] - a[t].im * pi[t];
                xi[t] += a[t].re * pi[t] + a[t].im * pr[t];
                T const nr = rr[t] + (be[t].re * pr[t] - be[t].im * pi[t]);
                T const ni = ri[t] + (be[t].re * pi[t] + be[t].im * pr[t]);
                or_[t] = nr; oi[t] = ni;
            }
            st(x + OFF(0, i), xr); st(x + OFF(1, i), xi);
            st(p0 + OFF(0, i), or_); st(p0 + OFF(1, i), oi);
        }
    }
}

// segment partials of (p, A p) from the per-block products the action computed: segpart[s][0, 1][j]
__global__ __launch_bounds__(64) void k_fold_blocks(double const* __restrict__ bpart, SegArgs g) {
    uint32_t const s = g.act[blockIdx.x];
    uint32_t const b0 = g.seg[3 * s + 1], b1 = g.seg[3 * s + 2];
    int const j = threadIdx.x;
    double re = 0, im = 0;
    for (uint32_t b = b0; b < b1; ++b) {
        uint64_t const blk = g.col_blocks[b];
        re += bpart[(blk * 2 + 0) * 64 + j];
        im += bpart[(blk * 2 + 1) * 64 + j];
    }
    g.segpart[(uint64_t(s) * 3 + 0) * 64 + j] = re;
    g.segpart[(uint64_t(s) * 3 + 1) * 64 + j] = im;
}

// x64 += x32, x32 = 0 (the FP32 increments of the solution since the last fold)
template <typename T>
__global__ __launch_bounds__(256) void k_fold_x(double* __restrict__ x64, T* __restrict__ x32, SegArgs g) {
    SEG_PROLOGUE
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
#pragma unroll 2
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
#pragma unroll
            for (int pl = 0; pl < 2; ++pl) {
                T v[V], z[V];
                double d[V];
                ld(v, x32 + OFF(pl, i));
                ldd<V>(d, x64 + OFF(pl, i));
#pragma unroll
                for (int t = 0; t < V; ++t) { d[t] += double(v[t]); z[t] = 0; }
                std_<V>(x64 + OFF(pl, i), d);
                st(x32 + OFF(pl, i), z);
            }
        }
    }
}

// x32 = 0 over the active segments
template <typename T>
__global__ __launch_bounds__(256) void k_zero_x(T* __restrict__ x32, SegArgs g) {
    SEG_PROLOGUE
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            T z[V];
#pragma unroll
            for (int t = 0; t < V; ++t) z[t] = 0;
            st(x32 + OFF(0, i), z); st(x32 + OFF(1, i), z);
        }
    }
}

// residual replacement: r = b - y with y = A x (FP64), b the unit columns; partials |r|^2, Re (r, r), Im (r, r)
template <typename T>
__global__ __launch_bounds__(256) void k_cocg_replace(T* __restrict__ rv, double const* __restrict__ y,
                                                      uint8_t const* __restrict__ diag, SegArgs g) {
    SEG_PROLOGUE
    double acc[3][V] = {};
    for (uint32_t b = b0; b < b1; ++b) {
        uint32_t const blk = g.col_blocks[b];
        bool const dg = diag[blk];
#pragma unroll 2
        for (int ps = 0; ps < M::NPASS; ++ps) {
            int const i = ps * M::RPP + r;
            double yr[V], yi[V];
            ldd<V>(yr, y + OFF(0, i)); ldd<V>(yi, y + OFF(1, i));
            T orr[V], oi[V];
#pragma unroll
            for (int t = 0; t < V; ++t) {
                double const vr = ((dg && i == j0 + t) ? 1.0 : 0.0) - yr[t], vi = -yi[t];
                orr[t] = T(vr); oi[t] = T(vi);
                double const cr = orr[t], ci = oi[t];
                acc[0][t] += cr * cr + ci * ci;
                acc[1][t] += cr * cr - ci * ci;
                acc[2][t] += 2 * cr * ci;
            }
            st(rv + OFF(0, i), orr); st(rv + OFF(1, i), oi);
        }
    }
    store_partials<T, 3>(acc, g.segpart, s);
}

// ------------------------------------------------------------------------------------------------
// per-column scalar updates; partials are folded over the column group's segments in a fixed order
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ double& F(double* S, uint32_t ncol, int f, uint32_t q) { return S[uint64_t(f) * ncol + q]; }

constexpr int FOLD_P = 8;  // waves folding the segments of one column group

// sums of the partial fields f[0..NC) of column group c over its segments: wave p adds segments
// p, p + FOLD_P, ...; the waves' sums are then added in wave order (deterministic).  Valid in threads
// tid < 64 (column j = tid).
template <int NC>
__device__ __forceinline__ void fold_group(double const* __restrict__ segpart, uint32_t const* __restrict__ col_seg,
                                           uint32_t c, int const (&f)[NC], double (&tot)[NC]) {
    __shared__ double sh[NC][FOLD_P][64];
    int const tid = threadIdx.x, j = tid & 63, part = tid >> 6;
    uint32_t const k0 = col_seg[c], k1 = col_seg[c + 1];
    double acc[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n) acc[n] = 0;
    for (uint32_t k = k0 + part; k < k1; k += FOLD_P)
#pragma unroll
        for (int n = 0; n < NC; ++n) acc[n] += segpart[(uint64_t(k) * 3 + f[n]) * 64 + j];
#pragma unroll
    for (int n = 0; n < NC; ++n) sh[n][part][j] = acc[n];
    __syncthreads();
    if (tid < 64) {
#pragma unroll
        for (int n = 0; n < NC; ++n) {
            double sum = 0;
            for (int p = 0; p < FOLD_P; ++p) sum += sh[n][p][j];
            tot[n] = sum;
        }
    }
}

__device__ __forceinline__ void freeze_step(double* S, uint32_t ncol, uint32_t q) {
    for (int f = ALPHA_RE; f <= ALPHA_IM; ++f) F(S, ncol, f, q) = 0;
    for (int f = ETAE_RE; f <= BETA_IM; ++f) F(S, ncol, f, q) = 0;
}

// The scalar kernels run one work-group of 64 * FOLD_P threads per column group; threads tid < 64
// update column q = 64 c + tid.

// after k_init: |b|^2, (r0*, b); columns of inactive groups start as converged
__global__ __launch_bounds__(64 * FOLD_P) void k_scalar_init(double* __restrict__ S, double const* __restrict__ segpart,
                                                          uint32_t const* __restrict__ col_seg,
                                                          uint8_t const* __restrict__ col_active, uint32_t ncol,
                                                          double tol_rel, double const* __restrict__ tol_abs) {
    uint32_t const c = blockIdx.x, q = 64 * c + threadIdx.x;
    if (!col_active[c]) {
        if (threadIdx.x < 64) {
            F(S, ncol, STATUS, q) = CONVERGED;
            F(S, ncol, MSTEP, q) = 0;
            freeze_step(S, ncol, q);
        }
        return;
    }
    int const f[3] = {0, 1, 2};
    double t[3];
    fold_group<3>(segpart, col_seg, c, f, t);
    if (threadIdx.x >= 64) return;
    double const bn = sqrt(t[0]);
    F(S, ncol, BNORM, q) = bn;
    F(S, ncol, TAU, q) = bn;
    F(S, ncol, THETA, q) = 0;
    F(S, ncol, RHO_RE, q) = t[1];
    F(S, ncol, RHO_IM, q) = t[2];
    double const tol = fmax(tol_rel * bn, tol_abs ? tol_abs[q] : 0.0);
    F(S, ncol, TOL, q) = tol;
    F(S, ncol, MSTEP, q) = 0;
    freeze_step(S, ncol, q);
    // nothing to do for a column whose right-hand side already meets the tolerance
    F(S, ncol, STATUS, q) = (bn <= tol) ? CONVERGED : ACTIVE;
}

// start of a step: alpha = rho / (r0*, v); a column that is no longer active gets zero coefficients
__global__ __launch_bounds__(64 * FOLD_P) void k_scalar_alpha(double* __restrict__ S, double const* __restrict__ segpart,
                                                           uint32_t const* __restrict__ col_seg,
                                                           uint8_t const* __restrict__ col_active, uint32_t ncol) {
    uint32_t const c = blockIdx.x, q = 64 * c + threadIdx.x;
    if (!col_active[c]) return;
    int const f[2] = {0, 1};
    double t[2];
    fold_group<2>(segpart, col_seg, c, f, t);
    if (threadIdx.x >= 64) return;
    if (F(S, ncol, STATUS, q) != ACTIVE) { freeze_step(S, ncol, q); return; }
    double const sr = t[0], si = t[1];
    double const den = sr * sr + si * si;
    if (!(den > 0)) { F(S, ncol, STATUS, q) = BREAKDOWN; freeze_step(S, ncol, q); return; }
    double const rr = F(S, ncol, RHO_RE, q), ri = F(S, ncol, RHO_IM, q);
    F(S, ncol, ALPHA_RE, q) = (rr * sr + ri * si) / den;
    F(S, ncol, ALPHA_IM, q) = (ri * sr - rr * si) / den;
}

// after a w update: theta, c, tau, eta, the d coefficient, convergence; on odd half-steps also rho, beta.
// Even half-steps write the E fields, odd ones the O fields (both are used by k_dx2).
template <int ODD>
__global__ __launch_bounds__(64 * FOLD_P) void k_scalar_theta(double* __restrict__ S, double const* __restrict__ segpart,
                                                           uint32_t const* __restrict__ col_seg,
                                                           uint8_t const* __restrict__ col_active, uint32_t ncol) {
    uint32_t const c = blockIdx.x, q = 64 * c + threadIdx.x;
    if (!col_active[c]) return;
    constexpr int NC = ODD ? 3 : 1;
    int f[NC];
    double t[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n) f[n] = n;
    fold_group<NC>(segpart, col_seg, c, f, t);
    if (threadIdx.x >= 64) return;
    int const FE = ODD ? ETAO_RE : ETAE_RE, FC = ODD ? CDO_RE : CDE_RE, FP = ODD ? ETAE_RE : ETAO_RE;
    if (F(S, ncol, STATUS, q) != ACTIVE) {
        F(S, ncol, FE, q) = 0; F(S, ncol, FE + 1, q) = 0;
        F(S, ncol, FC, q) = 0; F(S, ncol, FC + 1, q) = 0;
        if (ODD) { F(S, ncol, BETA_RE, q) = 0; F(S, ncol, BETA_IM, q) = 0; }
        return;
    }
    double const m = F(S, ncol, MSTEP, q);
    double const ar = F(S, ncol, ALPHA_RE, q), ai = F(S, ncol, ALPHA_IM, q);
    double const th = F(S, ncol, THETA, q), er = F(S, ncol, FP, q), ei = F(S, ncol, FP + 1, q);
    // cd = theta^2 eta / alpha with the previous theta, eta
    double const a2 = ar * ar + ai * ai;
    double cr = 0, ci = 0;
    if (a2 > 0) {
        double const tr = th * th * er, ti = th * th * ei;
        cr = (tr * ar + ti * ai) / a2;
        ci = (ti * ar - tr * ai) / a2;
    }
    double const tau = F(S, ncol, TAU, q);
    double const thn = sqrt(t[0]) / tau;
    double const cc = 1.0 / sqrt(1.0 + thn * thn);
    double const taun = tau * thn * cc;
    F(S, ncol, FC, q) = cr; F(S, ncol, FC + 1, q) = ci;
    F(S, ncol, THETA, q) = thn;
    F(S, ncol, TAU, q) = taun;
    F(S, ncol, FE, q) = cc * cc * ar;
    F(S, ncol, FE + 1, q) = cc * cc * ai;
    F(S, ncol, MSTEP, q) = m + 1;
    if (taun * sqrt(m + 2) <= F(S, ncol, TOL, q)) F(S, ncol, STATUS, q) = CONVERGED;  // x still gets this step
    if (ODD) {
        double const rnr = t[ODD ? 1 : 0], rni = t[ODD ? 2 : 0];
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
    if (!(taun == taun)) {
        F(S, ncol, STATUS, q) = BREAKDOWN;
        F(S, ncol, FE, q) = 0; F(S, ncol, FE + 1, q) = 0;
        F(S, ncol, FC, q) = 0; F(S, ncol, FC + 1, q) = 0;
    }
}

// COCG: after k_cocg_init: |b|^2, (b, b); tolerance and status as k_scalar_init
__global__ __launch_bounds__(64 * FOLD_P) void k_cocg_scalar_init(double* __restrict__ S, double const* __restrict__ segpart,
                                                               uint32_t const* __restrict__ col_seg,
                                                               uint8_t const* __restrict__ col_active, uint32_t ncol,
                                                               double tol_rel, double const* __restrict__ tol_abs) {
    uint32_t const c = blockIdx.x, q = 64 * c + threadIdx.x;
    if (!col_active[c]) {
        if (threadIdx.x < 64) {
            F(S, ncol, STATUS, q) = CONVERGED;
            F(S, ncol, MSTEP, q) = 0;
            F(S, ncol, RN, q) = 0;
            freeze_step(S, ncol, q);
        }
        return;
    }
    int const f[3] = {0, 1, 2};
    double t[3];
    fold_group<3>(segpart, col_seg, c, f, t);
    if (threadIdx.x >= 64) return;
    double const bn = sqrt(t[0]);
    F(S, ncol, BNORM, q) = bn;
    F(S, ncol, RHO_RE, q) = t[1];
    F(S, ncol, RHO_IM, q) = t[2];
    double const tol = fmax(tol_rel * bn, tol_abs ? tol_abs[q] : 0.0);
    F(S, ncol, TOL, q) = tol;
    F(S, ncol, MSTEP, q) = 0;
    freeze_step(S, ncol, q);
    F(S, ncol, STATUS, q) = (bn <= tol) ? CONVERGED : ACTIVE;
    F(S, ncol, RN, q) = (bn <= tol) ? 0.0 : 1.0;
}

// COCG: alpha = (r, r) / (p, A p)
__global__ __launch_bounds__(64 * FOLD_P) void k_cocg_alpha(double* __restrict__ S, double const* __restrict__ segpart,
                                                         uint32_t const* __restrict__ col_seg,
                                                         uint8_t const* __restrict__ col_active, uint32_t ncol) {
    uint32_t const c = blockIdx.x, q = 64 * c + threadIdx.x;
    if (!col_active[c]) return;
    int const f[2] = {0, 1};
    double t[2];
    fold_group<2>(segpart, col_seg, c, f, t);
    if (threadIdx.x >= 64) return;
    if (F(S, ncol, STATUS, q) != ACTIVE) { freeze_step(S, ncol, q); return; }
    double const mr = t[0], mi = t[1], den = mr * mr + mi * mi;
    if (!(den > 0)) { F(S, ncol, STATUS, q) = BREAKDOWN; freeze_step(S, ncol, q); return; }
    double const rr = F(S, ncol, RHO_RE, q), ri = F(S, ncol, RHO_IM, q);
    F(S, ncol, ALPHA_RE, q) = (rr * mr + ri * mi) / den;
    F(S, ncol, ALPHA_IM, q) = (ri * mr - rr * mi) / den;
}

// COCG: after the r update: convergence on the recursive residual |r| <= tol; beta = (r', r') / (r, r)
__global__ __launch_bounds__(64 * FOLD_P) void k_cocg_beta(double* __restrict__ S, double const* __restrict__ segpart,
                                                        uint32_t const* __restrict__ col_seg,
                                                        uint8_t const* __restrict__ col_active, uint32_t ncol,
                                                        int save_alpha) {
    uint32_t const c = blockIdx.x, q = 64 * c + threadIdx.x;
    if (!col_active[c]) return;
    int const f[3] = {0, 1, 2};
    double t[3];
    fold_group<3>(segpart, col_seg, c, f, t);
    if (threadIdx.x >= 64) return;
    if (save_alpha) {   // an even iteration: its x update is done by the next one (alpha 0 for stopped columns)
        F(S, ncol, APREV_RE, q) = F(S, ncol, ALPHA_RE, q);
        F(S, ncol, APREV_IM, q) = F(S, ncol, ALPHA_IM, q);
    }
    if (F(S, ncol, STATUS, q) != ACTIVE) { F(S, ncol, BETA_RE, q) = 0; F(S, ncol, BETA_IM, q) = 0; return; }
    F(S, ncol, MSTEP, q) = F(S, ncol, MSTEP, q) + 1;
    double const rn = sqrt(t[0]);
    F(S, ncol, RN, q) = rn / F(S, ncol, BNORM, q);   // relative to the right-hand side (> 0 for an active column)
    double const rr = F(S, ncol, RHO_RE, q), ri = F(S, ncol, RHO_IM, q), den = rr * rr + ri * ri;
    if (rn <= F(S, ncol, TOL, q)) {
        F(S, ncol, STATUS, q) = CONVERGED;   // x already holds this step
        F(S, ncol, BETA_RE, q) = 0; F(S, ncol, BETA_IM, q) = 0;
    } else if (!(den > 0) || !(rn == rn)) {
        F(S, ncol, STATUS, q) = BREAKDOWN;
        F(S, ncol, BETA_RE, q) = 0; F(S, ncol, BETA_IM, q) = 0;
    } else {
        F(S, ncol, BETA_RE, q) = (t[1] * rr + t[2] * ri) / den;
        F(S, ncol, BETA_IM, q) = (t[2] * rr - t[1] * ri) / den;
    }
    F(S, ncol, RHO_RE, q) = t[1];
    F(S, ncol, RHO_IM, q) = t[2];
}

// COCG after a residual replacement: rho = (r, r) of the replaced residual; convergence on its norm
__global__ __launch_bounds__(64 * FOLD_P) void k_cocg_rho(double* __restrict__ S, double const* __restrict__ segpart,
                                                       uint32_t const* __restrict__ col_seg,
                                                       uint8_t const* __restrict__ col_active, uint32_t ncol) {
    uint32_t const c = blockIdx.x, q = 64 * c + threadIdx.x;
    if (!col_active[c]) return;
    int const f[3] = {0, 1, 2};
    double t[3];
    fold_group<3>(segpart, col_seg, c, f, t);
    if (threadIdx.x >= 64) return;
    if (F(S, ncol, STATUS, q) != ACTIVE) return;
    F(S, ncol, RHO_RE, q) = t[1];
    F(S, ncol, RHO_IM, q) = t[2];
    if (sqrt(t[0]) <= F(S, ncol, TOL, q)) { F(S, ncol, STATUS, q) = CONVERGED; freeze_step(S, ncol, q); }
}

// residual norms of the active column groups (others keep their value)
__global__ __launch_bounds__(64 * FOLD_P) void k_col_norm(double* __restrict__ res, double const* __restrict__ segpart,
                                                       uint32_t const* __restrict__ col_seg,
                                                       uint8_t const* __restrict__ col_active) {
    uint32_t const c = blockIdx.x;
    if (!col_active[c]) return;
    int const f[1] = {0};
    double t[1];
    fold_group<1>(segpart, col_seg, c, f, t);
    if (threadIdx.x < 64) res[64 * c + threadIdx.x] = sqrt(t[0]);
}

// x = 0 over whole blocks (a fill kernel rather than a memset, so that it may sit in a captured graph)
__global__ __launch_bounds__(256) void k_zero(double* __restrict__ x, uint64_t n) {
    uint64_t const i = (uint64_t(blockIdx.x) * 256 + threadIdx.x) * 2;
    if (i < n) { d2 z; z[0] = 0; z[1] = 0; *reinterpret_cast<d2*>(x + i) = z; }
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
    bool ready = false;
    uint64_t nnzb = 0;
    uint32_t nCols = 0, ncol = 0;
    hipStream_t stream = nullptr;
    // static structure
    std::vector<uint32_t> h_cstart, h_cb;   // blocks of column group c: h_cb[h_cstart[c] .. h_cstart[c+1])
    uint32_t* col_blocks = nullptr;
    uint8_t* diag = nullptr;
    uint32_t* subset = nullptr;
    // segments of <= SEG blocks of one column group, numbered once for the whole solve
    uint32_t nseg_all = 0;
    uint32_t* seg = nullptr;                // [nseg_all][3]
    uint32_t* col_seg = nullptr;            // [nCols+1] segment range of each column group
    std::vector<uint32_t> h_col_seg;
    // active set
    std::vector<uint8_t> active;            // [nCols]
    uint32_t nseg = 0;                      // active segments
    uint32_t* act_seg = nullptr;            // [nseg_all] global numbers of the active segments
    uint32_t* act_order = nullptr;          // [nnzb] blocks of active groups in the action's order
    uint32_t* act_items = nullptr;          // [nnzb][head_words] their local-kernel headers
    uint64_t n_act_blocks = 0;
    uint8_t* col_active = nullptr;          // [nCols]
    std::vector<uint32_t> h_act_seg, h_act;
    // reductions and scalars
    double* segpart = nullptr;              // [nseg_all][3][64]
    double* S = nullptr;                    // [NFIELDS][ncol]
    double* S_host = nullptr;               // pinned copy of the status field
    double* res = nullptr;                  // [ncol] true residual norms
    double* tol_abs = nullptr;              // [ncol]
    double* bpart = nullptr;                // [nnzb][2][64] per-block (p, A p) from the action
    std::vector<double> col_halfsteps;      // [ncol] half-steps of every column in this solve
    // vectors
    double* x64 = nullptr;                  // the FP64 solution
    double* y64 = nullptr;                  // A x64 with the FP64 action (the refinement right-hand side is b - y64)
    void* arena = nullptr;                  // the inner solver's work vectors, sized for the largest method run so
    size_t arena_bytes = 0;                 // far; FP64 and FP32 inner solves run one after another and share it
    // captured step
    hipGraphExec_t graph = nullptr;
    int graph_fp = 0;
    int graph_steps = 0;                    // solver steps in the captured graph
    uint64_t graph_seed = 0;
    std::vector<void*> allocations;
    size_t held_bytes = 0;                  // device bytes of allocations[]
    // per solve: operator applications after which a column group leaves (empty or 0: no cap), and counters
    std::vector<int32_t> group_cap;
    // per FP32 solve: the application count of the progress check (0: none), the residual a group must be below
    // by then, the groups that failed it (they leave the solve at once: retire) and the groups that failed it
    // in this solve (lagging, kept for the precision decision after the solve)
    int progress_check = 0;
    double progress_floor = 0;
    std::vector<uint8_t> retire, lagging;
    int host_syncs = 0, graph_captures = 0;
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

// hipMalloc with the out-of-memory case reported as GFA_ERR_MEMORY
int dmalloc(void** q, size_t bytes, char const* what, std::string& err) {
    hipError_t const e = hipMalloc(q, bytes);
    if (e == hipSuccess) return 0;
    (void)hipGetLastError();
    *q = nullptr;
    err = std::string("solver ") + what + ": hipMalloc of " + std::to_string(bytes) + " bytes: " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? GFA_ERR_MEMORY : GFA_ERR_DEVICE;
}

template <typename T>
int walloc(SolverState& st, T*& ptr, size_t n, std::string& err) {
    void* q = nullptr;
    size_t const bytes = std::max<size_t>(n, 1) * sizeof(T);
    if (int r = dmalloc(&q, bytes, "state", err)) return r;
    st.allocations.push_back(q);
    st.held_bytes += bytes;
    ptr = (T*)q;
    return 0;
}

void drop_graph(SolverState& st) {
    if (st.graph) (void)hipGraphExecDestroy(st.graph);
    st.graph = nullptr;
}

int setup(gfa_plan* plan, SolverState& st, std::string& err) {
    if (st.ready) return 0;
    Plan const& h = capi_host(plan);
    st.nnzb = h.nnzb;
    st.nCols = h.nCols;
    st.ncol = 64 * h.nCols;
    HIP_TRY(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking));
    // blocks grouped by column group, ascending
    st.h_cstart.assign(h.nCols + 1, 0);
    st.h_cb.resize(h.nnzb);
    for (uint64_t b = 0; b < h.nnzb; ++b) ++st.h_cstart[h.colindx[b] + 1];
    for (uint32_t c = 0; c < h.nCols; ++c) st.h_cstart[c + 1] += st.h_cstart[c];
    {
        std::vector<uint32_t> fill(st.h_cstart.begin(), st.h_cstart.end() - 1);
        for (uint64_t b = 0; b < h.nnzb; ++b) st.h_cb[fill[h.colindx[b]]++] = uint32_t(b);
    }
    std::vector<uint32_t> segs;
    st.h_col_seg.assign(1, 0);
    for (uint32_t c = 0; c < h.nCols; ++c) {
        for (uint32_t b = st.h_cstart[c]; b < st.h_cstart[c + 1]; b += SEG) {
            segs.push_back(c); segs.push_back(b); segs.push_back(std::min(st.h_cstart[c + 1], b + SEG));
        }
        st.h_col_seg.push_back(uint32_t(segs.size() / 3));
    }
    st.nseg_all = uint32_t(segs.size() / 3);
    std::vector<uint8_t> diag(h.nnzb, 0);
    for (uint32_t c = 0; c < h.nCols; ++c) diag[h.subset[c]] = 1;
    int s;
    if ((s = walloc(st, st.col_blocks, h.nnzb, err))) return s;
    if ((s = walloc(st, st.diag, h.nnzb, err))) return s;
    if ((s = walloc(st, st.subset, h.nCols, err))) return s;
    if ((s = walloc(st, st.seg, size_t(st.nseg_all) * 3, err))) return s;
    if ((s = walloc(st, st.col_seg, size_t(h.nCols) + 1, err))) return s;
    if ((s = walloc(st, st.act_seg, st.nseg_all, err))) return s;
    if ((s = walloc(st, st.act_order, h.nnzb, err))) return s;
    if ((s = walloc(st, st.act_items, h.nnzb * uint64_t(capi_dev(plan).head_words), err))) return s;
    if ((s = walloc(st, st.col_active, h.nCols, err))) return s;
    HIP_TRY(hipMemcpy(st.col_blocks, st.h_cb.data(), h.nnzb * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(st.diag, diag.data(), h.nnzb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(st.subset, h.subset.data(), size_t(h.nCols) * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(st.seg, segs.data(), segs.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(st.col_seg, st.h_col_seg.data(), st.h_col_seg.size() * 4, hipMemcpyHostToDevice));
    if ((s = walloc(st, st.segpart, size_t(st.nseg_all) * 3 * 64, err))) return s;
    if ((s = walloc(st, st.S, size_t(NFIELDS) * st.ncol, err))) return s;
    if ((s = walloc(st, st.res, st.ncol, err))) return s;
    if ((s = walloc(st, st.tol_abs, st.ncol, err))) return s;
    if ((s = walloc(st, st.bpart, size_t(h.nnzb) * 128, err))) return s;
    HIP_TRY(hipHostMalloc((void**)&st.S_host, size_t(st.ncol) * sizeof(double), hipHostMallocDefault));
    if ((s = walloc(st, st.x64, h.nnzb * NBLK, err))) return s;
    if ((s = walloc(st, st.y64, h.nnzb * NBLK, err))) return s;
    st.ready = true;
    return 0;
}

// the active segments, the action's block list and the column flags for the active column groups;
// segment numbers stay fixed, so partial sums written before a retirement keep their meaning
int set_active(gfa_plan* plan, SolverState& st, std::vector<uint8_t> const& active, std::string& err) {
    st.active = active;
    st.h_act_seg.clear();
    for (uint32_t c = 0; c < st.nCols; ++c)
        if (active[c])
            for (uint32_t k = st.h_col_seg[c]; k < st.h_col_seg[c + 1]; ++k) st.h_act_seg.push_back(k);
    st.nseg = uint32_t(st.h_act_seg.size());
    DevicePlan const& dev = capi_dev(plan);
    Plan const& h = capi_host(plan);
    st.h_act.clear();
    for (uint32_t b : dev.h_order) if (active[h.colindx[b]]) st.h_act.push_back(b);
    st.n_act_blocks = st.h_act.size();
    if (st.nseg) HIP_TRY(hipMemcpyAsync(st.act_seg, st.h_act_seg.data(), st.nseg * 4, hipMemcpyHostToDevice, st.stream));
    if (st.n_act_blocks) {
        HIP_TRY(hipMemcpyAsync(st.act_order, st.h_act.data(), st.n_act_blocks * 4, hipMemcpyHostToDevice, st.stream));
        if (int r = device_gather_heads(dev, st.act_order, st.n_act_blocks, st.act_items, st.stream, err)) return r;
    }
    HIP_TRY(hipMemcpyAsync(st.col_active, active.data(), st.nCols, hipMemcpyHostToDevice, st.stream));
    HIP_TRY(hipStreamSynchronize(st.stream));  // the host vectors may change before the copies ran otherwise
    ++st.host_syncs;
    drop_graph(st);
    return 0;
}

// the action restricted to the active set while it is in scope
struct ActiveScope {
    DevicePlan& d;
    ActiveScope(DevicePlan& d_, SolverState const& st) : d(d_) {
        d.active_order = st.act_order; d.active_items = st.act_items; d.n_active = st.n_act_blocks;
        d.col_active = st.col_active;
    }
    ~ActiveScope() { d.active_order = nullptr; d.active_items = nullptr; d.n_active = 0; d.col_active = nullptr; }
};

// work vector k of precision T in the arena
template <typename T> T* work(SolverState& st, int k) { return reinterpret_cast<T*>(st.arena) + size_t(k) * st.nnzb * NBLK; }

// work vectors per inner solve: TFQMR w, u, un, v, d, A un, A u; COCG r, the two search directions, A p and, in
// FP32, the FP32 increment of x
template <typename T> int work_vectors(int method) { return method == 1 ? (sizeof(T) == 4 ? 5 : 4) : 7; }

// an arena of at least n work vectors of T; a larger one replaces the old (whose contents are not kept: every
// inner solve starts from its right-hand side), and a captured graph that points into it is dropped
template <typename T>
int ensure_work(SolverState& st, int n, std::string& err) {
    size_t const need = size_t(n) * st.nnzb * NBLK * sizeof(T);
    if (st.arena_bytes >= need) return 0;
    drop_graph(st);
    if (st.arena) { HIP_TRY(hipStreamSynchronize(st.stream)); HIP_TRY(hipFree(st.arena)); }
    st.arena = nullptr;
    st.arena_bytes = 0;
    if (int r = dmalloc(&st.arena, need, "work vectors", err)) return r;
    st.arena_bytes = need;
    return 0;
}

struct TfqmrResult { int halfsteps = 0; int actions = 0; };

// work vectors of one TFQMR run
template <typename T> struct Vecs { T *w, *u, *un, *v, *d, *Aodd, *Aeven; };

// one full step (even + odd half-step); all arguments are invariant between steps
template <typename T>
int enqueue_step(gfa_plan* plan, SolverState& st, Vecs<T> const& V, uint64_t seed, hipStream_t s, std::string& err) {
    int const fp = sizeof(T) == 8 ? 64 : 32;
    SegArgs const g{st.act_seg, st.seg, st.col_blocks, st.segpart};
    dim3 const gb(st.nseg), tb(256), gs(st.nCols), ts(64 * FOLD_P);
    hipLaunchKernelGGL(k_scalar_alpha, gs, ts, 0, s, st.S, st.segpart, st.col_seg, st.col_active, st.ncol);
    hipLaunchKernelGGL((k_even<T>), gb, tb, 0, s, V.un, V.u, V.v, V.w, V.Aeven, st.S, st.ncol, g);
    hipLaunchKernelGGL((k_scalar_theta<0>), gs, ts, 0, s, st.S, st.segpart, st.col_seg, st.col_active, st.ncol);
    if (int r = capi_action(plan, fp, V.Aodd, V.un, s, err)) return r;
    hipLaunchKernelGGL((k_odd<T>), gb, tb, 0, s, V.w, V.Aodd, st.S, st.ncol, g, seed);
    hipLaunchKernelGGL((k_scalar_theta<1>), gs, ts, 0, s, st.S, st.segpart, st.col_seg, st.col_active, st.ncol);
    hipLaunchKernelGGL((k_dx2<T>), gb, tb, 0, s, V.d, st.x64, V.u, V.un, V.w, st.S, st.ncol, g);
    if (int r = capi_action(plan, fp, V.Aeven, V.u, s, err)) return r;
    hipLaunchKernelGGL((k_vnext<T>), gb, tb, 0, s, V.v, V.Aeven, V.Aodd, st.S, st.ncol, g, seed);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Runs steps (each `halfsteps_per_step` half-steps = actions) until every column stopped or maxit half-steps:
// replays a captured HIP graph of the `steps_per_check` steps between two status polls (eager: plain launches),
// polls the column status, and retires finished column groups.  A retired group's columns are frozen on the
// device by their status, so the active set is compacted (and the graph captured again) only when the live blocks
// fall below 7/8 of the captured set, or when a group reaches its iteration cap (its status stays active).
struct NoHook { int operator()(int) { return 0; } };

template <typename StepFn, typename HookFn = NoHook>
int iterate(gfa_plan* plan, SolverState& st, ActiveScope& scope, int maxit, int steps_per_check, int halfsteps_per_step,
            int fp, uint64_t key, bool eager, StepFn&& step, TfqmrResult& res, std::string& err, HookFn hook = HookFn()) {
    hipStream_t const s = st.stream;
    int m = 0;
    while (m < maxit) {
        int const nsteps = std::min(steps_per_check, (maxit - m + halfsteps_per_step - 1) / halfsteps_per_step);
        // step k of the interval (intervals start at multiples of steps_per_check: COCG uses the step's parity)
        if (eager || nsteps != steps_per_check) {   // a shorter last interval runs its steps directly
            for (int k = 0; k < nsteps; ++k)
                if (int r = step(s, k)) return r;
        } else {
            // a full poll interval is one graph of nsteps steps
            int const gsteps = nsteps;
            if (!st.graph || st.graph_fp != fp || st.graph_seed != key || st.graph_steps != gsteps) {
                drop_graph(st);
                hipGraph_t gr = nullptr;
                HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                int r = 0;
                for (int k = 0; k < gsteps && !r; ++k) r = step(s, k);
                hipError_t const e = hipStreamEndCapture(s, &gr);
                if (r) { if (gr) (void)hipGraphDestroy(gr); return r; }
                if (e != hipSuccess) { err = std::string("graph capture: ") + hipGetErrorString(e); return GFA_ERR_DEVICE; }
                hipError_t const e2 = hipGraphInstantiate(&st.graph, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (e2 != hipSuccess) { st.graph = nullptr; err = std::string("graph instantiate: ") + hipGetErrorString(e2); return GFA_ERR_DEVICE; }
                st.graph_fp = fp;
                st.graph_seed = key;
                st.graph_steps = gsteps;
                ++st.graph_captures;
            }
            for (int k = 0; k < nsteps / gsteps; ++k) HIP_TRY(hipGraphLaunch(st.graph, s));
        }
        m += halfsteps_per_step * nsteps;
        res.actions += halfsteps_per_step * nsteps;
        res.halfsteps = m;
        if (int r = hook(m)) return r;   // e.g. a residual replacement, between graph replays
        HIP_TRY(hipMemcpyAsync(st.S_host, st.S + size_t(STATUS) * st.ncol, st.ncol * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        ++st.host_syncs;
        std::vector<uint8_t> act = st.active;
        bool any = false, changed = false, capped = false;
        uint64_t live_blocks = 0;
        for (uint32_t c = 0; c < st.nCols; ++c) {
            if (!act[c]) continue;
            bool live = false;
            for (int j = 0; j < 64 && !live; ++j) live = (st.S_host[64 * c + j] == ACTIVE);
            if (live && !st.group_cap.empty() && st.group_cap[c] > 0 && m >= st.group_cap[c]) { live = false; capped = true; }
            if (live && !st.retire.empty() && st.retire[c]) { live = false; capped = true; }
            if (live) { any = true; live_blocks += st.h_cstart[c + 1] - st.h_cstart[c]; }
            else { act[c] = 0; changed = true; }
        }
        if (!any) break;
        if (changed && (capped || 8 * live_blocks < 7 * st.n_act_blocks)) {
            if (int r = set_active(plan, st, act, err)) return r;
            scope.d.active_order = st.act_order; scope.d.active_items = st.act_items; scope.d.n_active = st.n_act_blocks;
        }
    }
    {   // per-column half-step counts of this run (columns of inactive groups count 0)
        std::vector<double> ms(st.ncol);
        HIP_TRY(hipMemcpyAsync(ms.data(), st.S + size_t(MSTEP) * st.ncol, st.ncol * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        ++st.host_syncs;
        for (uint32_t q = 0; q < st.ncol; ++q) st.col_halfsteps[q] += ms[q];
    }
    return 0;
}

// TFQMR on T vectors for the right-hand side b - rhs_y (rhs_y nullptr: b, the unit columns), accumulating into
// st.x64.
// Column groups whose columns all stopped are retired from the active set as the solve goes.
template <typename T>
int tfqmr(gfa_plan* plan, SolverState& st, double const* rhs_y, double tol_rel, double const* tol_abs, int maxit,
          int check_every, uint64_t seed, bool eager, bool autotune, TfqmrResult& res, std::string& err) {
    if (int s = ensure_work<T>(st, work_vectors<T>(0), err)) return s;
    Vecs<T> V{work<T>(st, 0), work<T>(st, 1), work<T>(st, 2), work<T>(st, 3), work<T>(st, 4), work<T>(st, 5), work<T>(st, 6)};
    int const fp = sizeof(T) == 8 ? 64 : 32;
    hipStream_t const s = st.stream;
    SegArgs const g{st.act_seg, st.seg, st.col_blocks, st.segpart};
    dim3 const gs(st.nCols), ts(64 * FOLD_P);
    DevicePlan& dev = capi_dev(plan);
    ActiveScope scope(dev, st);
    if (st.nseg == 0) return 0;

    hipLaunchKernelGGL((k_init<T>), dim3(st.nseg), dim3(256), 0, s, V.d, V.w, V.u, rhs_y, st.diag, g, seed);
    hipLaunchKernelGGL(k_scalar_init, gs, ts, 0, s, st.S, st.segpart, st.col_seg, st.col_active, st.ncol, tol_rel, tol_abs);
    // v = A u0 = Au_even of the first step
    if (int r = capi_action(plan, fp, V.v, V.u, s, err)) return r;
    res.actions++;
    HIP_TRY(hipMemcpyAsync(V.Aeven, V.v, st.nnzb * NBLK * sizeof(T), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL((k_dot_shadow<T>), dim3(st.nseg), dim3(256), 0, s, V.v, g, seed);
    HIP_TRY(hipGetLastError());
    if (autotune && !dev.tuned[sizeof(T) == 8 ? 0 : 1]) {  // tile shapes, before any graph is captured
        drop_graph(st);
        if (int r = capi_autotune(plan, fp, V.Aodd, V.u, s, 2, nullptr, err)) return r;
    }

    return iterate(plan, st, scope, maxit, std::max(1, check_every / 2), 2, fp, seed * 2 + 0, eager,
                   [&](hipStream_t ss, int) { return enqueue_step<T>(plan, st, V, seed, ss, err); }, res, err);
}

// one COCG iteration of parity `odd`: q = A p (p in P1 if odd, else P0), alpha, r, beta, then P1 = r + beta P0
// (even) or x += alpha_prev P0 + alpha P1, P0 = r + beta P1 (odd)
template <typename T>
int enqueue_cocg(gfa_plan* plan, SolverState& st, T* x, T* r, T* p0, T* p1, T* q, int odd, hipStream_t s,
                 std::string& err) {
    T* const p = odd ? p1 : p0;
    int const fp = sizeof(T) == 8 ? 64 : 32;
    SegArgs const g{st.act_seg, st.seg, st.col_blocks, st.segpart};
    dim3 const gb(st.nseg), tb(256), gs(st.nCols), ts(64 * FOLD_P);
    // (p, A p) from the action's epilogue when nothing is added to A p after the local kernel (no numax 6, 7
    // atoms, add-back on the vector ALUs), else from a pass over p and q
    DevicePlan const& dv = capi_dev(plan);
    bool const fused = dv.n_wide_blocks == 0 && dv.prj_engine[fp == 64 ? 0 : 1] == 0;
    if (int e = capi_action(plan, fp, q, p, s, err, ACTION_ALL, fused ? st.bpart : nullptr)) return e;
    if (fused) hipLaunchKernelGGL(k_fold_blocks, gb, dim3(64), 0, s, st.bpart, g);
    else hipLaunchKernelGGL((k_cocg_pq<T>), gb, tb, 0, s, p, q, g);
    hipLaunchKernelGGL(k_cocg_alpha, gs, ts, 0, s, st.S, st.segpart, st.col_seg, st.col_active, st.ncol);
    hipLaunchKernelGGL((k_cocg_r<T>), gb, tb, 0, s, r, q, st.S, st.ncol, g);
    hipLaunchKernelGGL(k_cocg_beta, gs, ts, 0, s, st.S, st.segpart, st.col_seg, st.col_active, st.ncol, odd ? 0 : 1);
    if (odd) hipLaunchKernelGGL((k_cocg_px_odd<T>), gb, tb, 0, s, x, p0, p1, r, st.S, st.ncol, g);
    else hipLaunchKernelGGL((k_cocg_p_even<T>), gb, tb, 0, s, p1, p0, r, st.S, st.ncol, g);
    HIP_TRY(hipGetLastError());
    return 0;
}

// COCG on T vectors for the right-hand side b - rhs_y (rhs_y nullptr: b, the unit columns), accumulating into
// st.x64; converges on the recursive residual |r| <= max(tol_rel |b|, tol_abs).  Requires a complex-symmetric
// operator.
template <typename T>
int cocg(gfa_plan* plan, SolverState& st, double const* rhs_y, double tol_rel, double const* tol_abs, int maxit,
         int check_every, bool eager, bool autotune, TfqmrResult& res, std::string& err) {
    if (int s = ensure_work<T>(st, work_vectors<T>(1), err)) return s;
    // r, the search directions P0 and P1, q; x = the FP32 increment of the solution (FP32) or the FP64 solution
    T *r = work<T>(st, 0), *p = work<T>(st, 1), *q = work<T>(st, 2);
    bool const f32 = sizeof(T) == 4;
    T* x = f32 ? work<T>(st, 3) : reinterpret_cast<T*>(st.x64);
    T* const p1 = work<T>(st, f32 ? 4 : 3);
    int const fp = f32 ? 32 : 64;
    hipStream_t const s = st.stream;
    DevicePlan& dev = capi_dev(plan);
    ActiveScope scope(dev, st);
    if (st.nseg == 0) return 0;
    std::vector<uint8_t> const started = st.active;
    {
        SegArgs const g{st.act_seg, st.seg, st.col_blocks, st.segpart};
        hipLaunchKernelGGL((k_cocg_init<T>), dim3(st.nseg), dim3(256), 0, s, r, p, rhs_y, st.diag, g);
        hipLaunchKernelGGL(k_cocg_scalar_init, dim3(st.nCols), dim3(64 * FOLD_P), 0, s, st.S, st.segpart, st.col_seg,
                           st.col_active, st.ncol, tol_rel, tol_abs);
        if (f32) hipLaunchKernelGGL((k_zero_x<T>), dim3(st.nseg), dim3(256), 0, s, x, g);
        HIP_TRY(hipGetLastError());
    }
    if (int e = capi_action(plan, fp, q, p, s, err, 0)) return e;   // allocates the action's workspace, launches nothing
    if (autotune && !dev.tuned[f32 ? 1 : 0]) {
        drop_graph(st);
        if (int e = capi_autotune(plan, fp, q, p, s, 2, nullptr, err)) return e;
    }
    auto fold = [&]() -> int {   // x64 += x32 over the active groups
        SegArgs const g{st.act_seg, st.seg, st.col_blocks, st.segpart};
        hipLaunchKernelGGL((k_fold_x<T>), dim3(st.nseg), dim3(256), 0, s, st.x64, x, g);
        HIP_TRY(hipGetLastError());
        return 0;
    };
    // FP32: every REPLACE iterations the FP32 increments are folded into the FP64 solution and the recursive
    // residual is replaced by the true one, b - A x with the FP64 action (reliable updates, van der Vorst & Ye,
    // SIAM J. Sci. Comput. 22, 835 (2000)), so the FP32 rounding of the recurrences cannot accumulate into a floor
    // of the true residual; b is the unit columns also in refinement solves, since x holds the whole solution.
    constexpr int REPLACE = 64;
    int last = 0;
    double* y64 = st.y64;   // A x64, replaced for the active groups (their right-hand side is no longer read)
    bool checked = false;
    auto hook = [&](int m) -> int {
        if (f32 && st.progress_check > 0 && !checked && m >= st.progress_check) {
            // the progress check: a group whose recursive residual (max over its columns) is still above the
            // floor leaves this solve at this poll and is marked for FP64
            checked = true;
            std::vector<double> rn(st.ncol);
            HIP_TRY(hipMemcpyAsync(rn.data(), st.S + size_t(RN) * st.ncol, st.ncol * sizeof(double), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            ++st.host_syncs;
            st.retire.assign(st.nCols, 0);
            for (uint32_t c = 0; c < st.nCols; ++c) {
                if (!st.active[c]) continue;
                double g = 0;
                for (int j = 0; j < 64; ++j) g = std::max(g, rn[64 * c + j]);
                if (g > st.progress_floor) st.retire[c] = st.lagging[c] = 1;
            }
        }
        if (!f32 || m - last < REPLACE) return 0;
        last = m;
        if (int e = fold()) return e;
        if (int e = capi_action(plan, 64, y64, st.x64, s, err)) return e;
        res.actions++;
        SegArgs const g2{st.act_seg, st.seg, st.col_blocks, st.segpart};
        hipLaunchKernelGGL((k_cocg_replace<T>), dim3(st.nseg), dim3(256), 0, s, r, y64, st.diag, g2);
        hipLaunchKernelGGL(k_cocg_rho, dim3(st.nCols), dim3(64 * FOLD_P), 0, s, st.S, st.segpart, st.col_seg,
                           st.col_active, st.ncol);
        HIP_TRY(hipGetLastError());
        return 0;
    };
    // an even number of iterations per poll interval and in all (an odd limit runs one iteration more), so that
    // every interval starts with an even iteration and ends with an odd one: no x update is left deferred when
    // the host polls, compacts the active set or returns
    int const per_check = std::max(2, (check_every + 1) & ~1);
    int const rc = iterate(plan, st, scope, maxit + (maxit & 1), per_check, 1, fp, 1, eager,
                           [&](hipStream_t ss, int k) { return enqueue_cocg<T>(plan, st, x, r, p, p1, q, k & 1, ss, err); },
                           res, err, hook);
    if (rc) return rc;
    if (f32) {   // the increments of every group that took part, also of those retired since the last fold
        if (int e = set_active(plan, st, started, err)) return e;
        scope.d.active_order = st.act_order; scope.d.active_items = st.act_items; scope.d.n_active = st.n_act_blocks;
        if (int e = fold()) return e;
    }
    return 0;
}

// the inner solver of `method` (0 TFQMR, 1 COCG) for the right-hand side b - rhs_y
template <typename T>
int krylov(int method, gfa_plan* plan, SolverState& st, double const* rhs_y, double tol_rel, double const* tol_abs,
           int maxit, int check_every, uint64_t seed, bool eager, bool autotune, TfqmrResult& res, std::string& err) {
    if (method == 1) return cocg<T>(plan, st, rhs_y, tol_rel, tol_abs, maxit, check_every, eager, autotune, res, err);
    return tfqmr<T>(plan, st, rhs_y, tol_rel, tol_abs, maxit, check_every, seed, eager, autotune, res, err);
}

// true residual b - A x64 with the FP64 action over the active groups: y64 = A x64 (the next refinement's
// right-hand side is b - y64), norms into st.res
int residual(gfa_plan* plan, SolverState& st, int& actions, std::string& err) {
    DevicePlan& dev = capi_dev(plan);
    ActiveScope scope(dev, st);
    if (st.nseg == 0) return 0;
    hipStream_t const s = st.stream;
    if (int r = capi_action(plan, 64, st.y64, st.x64, s, err)) return r;
    actions++;
    SegArgs const g{st.act_seg, st.seg, st.col_blocks, st.segpart};
    hipLaunchKernelGGL((k_residual<double>), dim3(st.nseg), dim3(256), 0, s, st.y64, st.diag, g);
    hipLaunchKernelGGL(k_col_norm, dim3(st.nCols), dim3(64 * FOLD_P), 0, s, st.res, st.segpart, st.col_seg,
                       st.col_active);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

uint64_t solver_held_bytes(gfa_plan* plan) {
    SolverState const* st = state_of(plan, false);
    return st ? uint64_t(st->held_bytes + st->arena_bytes) : 0;
}

uint64_t solve_bytes(gfa_plan* plan, int fp_bits, int method) {
    Plan const& h = capi_host(plan);
    uint64_t const nnzb = h.nnzb, ncol = 64ull * h.nCols;
    uint64_t nseg = 0;
    {
        std::vector<uint64_t> per(h.nCols, 0);
        for (uint64_t b = 0; b < nnzb; ++b) ++per[h.colindx[b]];
        for (uint32_t c = 0; c < h.nCols; ++c) nseg += (per[c] + SEG - 1) / SEG;
    }
    uint64_t const vec64 = nnzb * NBLK * 8;
    uint64_t bytes = 2 * vec64;                          // x64, y64
    bytes += nnzb * (4 + 1 + 4) + nnzb * 128 * 8;        // col_blocks, diag, act_order; the per-block (p, A p)
    {   // the active blocks' headers: 16 words, 32 with the 16th-order stencil
        bool wide = false;
        for (int d = 0; d < 3; ++d) wide = wide || h.kinetic[d].fd_range == 8;
        bytes += nnzb * 4 * uint64_t(wide ? 32 : 16);
    }
    bytes += nseg * (12 + 4) + nseg * 3 * 64 * 8;        // segments, active segments, segment partials
    bytes += ncol * 8 * (NFIELDS + 2) + uint64_t(h.nCols) * 13;
    int const m = method == 1 ? 1 : 0;
    uint64_t arena = uint64_t(fp_bits == 64 ? work_vectors<double>(m) * 2 : work_vectors<float>(m)) * (vec64 / 2);
    if (fp_bits == 32) arena = std::max(arena, uint64_t(work_vectors<double>(m)) * vec64);   // an FP64 fallback
    return bytes + arena;
}

void solver_release(gfa_plan* plan) {
    for (size_t k = 0; k < g_states.size(); ++k) {
        if (g_states[k].first != plan) continue;
        auto* s = g_states[k].second;
        drop_graph(*s);
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        for (void* p : s->allocations) (void)hipFree(p);
        if (s->arena) (void)hipFree(s->arena);
        if (s->S_host) (void)hipHostFree(s->S_host);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        delete s;
        g_states.erase(g_states.begin() + k);
        return;
    }
}

int solve(gfa_plan* plan, gfa_solve_opts const& o, double* rho, gfa_solve_stats* stats, std::string& err) {
    if (o.fp_bits != 64 && o.fp_bits != 32) { err = "fp_bits must be 32 or 64"; return GFA_ERR_ARGUMENT; }
    if (!(o.tolerance > 0)) { err = "tolerance must be positive"; return GFA_ERR_ARGUMENT; }
    if (o.max_iterations < 0 || o.max_refinements < 0) { err = "negative iteration limits"; return GFA_ERR_ARGUMENT; }
    if (o.method != 0 && o.method != 1) { err = "method must be 0 (tfQMR) or 1 (COCG)"; return GFA_ERR_ARGUMENT; }
    if (o.method == 1 && !capi_host(plan).complex_symmetric()) {
        err = "COCG needs a complex-symmetric operator (real Bloch phases: Gamma, k = 1/2 or a wrapped cell)";
        return GFA_ERR_ARGUMENT;
    }
    auto const t0 = std::chrono::steady_clock::now();
    SolverState& st = *state_of(plan, true);
    if (int s = setup(plan, st, err)) return s;
    // the caller's stream orders the inputs; the solve itself runs on the solver's own stream
    HIP_TRY(hipStreamSynchronize((hipStream_t)o.stream));
    hipStream_t const s = st.stream;
    bool const eager = o.eager != 0;
    int const check_every = o.check_every > 0 ? o.check_every : 8;
    uint64_t const seed = o.shadow_seed ? o.shadow_seed : 0x5EEDA43Cull;
    gfa_solve_stats S;
    std::memset(&S, 0, sizeof(S));
    S.total_columns = int32_t(st.ncol);
    std::vector<double> resid(st.ncol, 0.0);
    auto fetch_resid = [&]() -> int {
        HIP_TRY(hipMemcpyAsync(resid.data(), st.res, st.ncol * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        ++st.host_syncs;
        return 0;
    };
    auto max_resid = [&]() { double mx = 0; for (double v : resid) mx = std::max(mx, v); return mx; };
    // groups whose columns all meet the tolerance leave the active set of the next refinement
    auto unconverged = [&]() {
        std::vector<uint8_t> a(st.nCols, 0);
        for (uint32_t c = 0; c < st.nCols; ++c)
            for (int j = 0; j < 64; ++j) a[c] |= uint8_t(resid[64 * c + j] > o.tolerance);
        return a;
    };
    {   // refinement solves stop a column once it is below half the final tolerance
        std::vector<double> tabs(st.ncol, o.tolerance * 0.5);
        HIP_TRY(hipMemcpyAsync(st.tol_abs, tabs.data(), st.ncol * sizeof(double), hipMemcpyHostToDevice, s));
    }
    st.host_syncs = 0;
    st.graph_captures = 0;
    if (o.group_max_iterations) st.group_cap.assign(o.group_max_iterations, o.group_max_iterations + st.nCols);
    else st.group_cap.clear();
    std::vector<uint8_t> all(st.nCols, 1);
    st.col_halfsteps.assign(st.ncol, 0.0);
    int fp64_groups = 0;   // FP32 solves: column groups that went on in FP64
    if (int r = set_active(plan, st, all, err)) return r;
    {
        uint64_t const n = st.nnzb * NBLK;
        hipLaunchKernelGGL(k_zero, dim3(uint32_t((n / 2 + 255) / 256)), dim3(256), 0, s, st.x64, n);
        HIP_TRY(hipGetLastError());
    }

    if (o.fp_bits == 64) {
        TfqmrResult tr;
        if (int r = krylov<double>(o.method, plan, st, nullptr, o.tolerance, nullptr, o.max_iterations, check_every, seed, eager, o.autotune != 0, tr, err)) return r;
        S.iterations += tr.halfsteps; S.actions += tr.actions;
        if (int r = set_active(plan, st, all, err)) return r;
        if (int r = residual(plan, st, S.actions, err)) return r;
        if (int r = fetch_resid()) return r;
        if (o.echo > 0) std::fprintf(stderr, "# gfa_solve fp64: %d half-steps, max residual %.3e\n", tr.halfsteps, max_resid());
        // FP64 refinement steps if requested and needed: solve A delta = b - A x, x += delta
        for (int k = 0; k < o.max_refinements && max_resid() > o.tolerance; ++k) {
            auto const act = unconverged();
            if (int r = set_active(plan, st, act, err)) return r;
            TfqmrResult t2;
            if (int r = krylov<double>(o.method, plan, st, st.y64, 1e-30, st.tol_abs, o.max_iterations, check_every, seed + k + 1, eager, false, t2, err)) return r;
            S.iterations += t2.halfsteps; S.actions += t2.actions; S.refinements++;
            if (int r = set_active(plan, st, act, err)) return r;
            if (int r = residual(plan, st, S.actions, err)) return r;
            if (int r = fetch_resid()) return r;
        }
    } else {
        // FP32 inner solves inside FP64 refinement, per column group.  tfQMR: one FP32 solve (to inner_tolerance or
        // the FP32 budget), then a group that is not converged goes on in FP64 from its FP32 iterate; a group whose
        // FP32 solve reduced its true (FP64) residual less than tenfold (the FP32 recurrences of an ill-conditioned
        // instance stall) is solved in FP64 from zero instead.  Further FP32 refinement solves cost more than they
        // gain with tfQMR: each restarts the Krylov space.  COCG: a group that spends the FP32 budget without
        // converging continues in FP64 from its FP32 iterate (COCG with residual replacement does not stall, but in
        // FP32 it needs up to several times the FP64 iterations on ill-conditioned instances, and an FP64 iteration
        // costs about twice an FP32 one); otherwise it stays in FP32.  The FP64 work vectors take the arena of the
        // FP32 ones (the two never run at the same time).
        double const floor_rel = o.inner_tolerance > 0 ? o.inner_tolerance : 1e-4;  // gain asked of one FP32 solve
        int const budget32 = o.fp32_max_iterations > 0 ? o.fp32_max_iterations : o.max_iterations;
        std::vector<uint8_t> prec64(st.nCols, 0), done(st.nCols, 0);
        std::vector<double> gres(st.nCols, 1.0);  // max relative residual of each group (||b|| = 1)
        bool first32 = true;
        // the arena at the size of the FP64 inner solve from the start (which also holds the FP32 vectors), so that
        // a group going on in FP64 does not free and allocate it again in the middle of the solve
        if (int r = ensure_work<double>(st, work_vectors<double>(o.method), err)) return r;
        auto group_max = [&](uint32_t c) { double m = 0; for (int j = 0; j < 64; ++j) m = std::max(m, resid[64 * c + j]); return m; };
        for (int k = 0; k <= o.max_refinements; ++k) {
            std::vector<uint8_t> a32(st.nCols, 0), a64(st.nCols, 0), fresh64(st.nCols, 0);
            bool any32 = false, any64 = false;
            for (uint32_t c = 0; c < st.nCols; ++c) {
                if (done[c]) continue;
                if (prec64[c]) { a64[c] = 1; any64 = true; } else { a32[c] = 1; any32 = true; }
            }
            if (!any32 && !any64) break;
            if (any32) {
                if (int r = set_active(plan, st, a32, err)) return r;
                std::vector<double> const hs0 = st.col_halfsteps;   // applications before this FP32 solve
                TfqmrResult tr;
                st.lagging.assign(st.nCols, 0);
                st.retire.clear();
                st.progress_check = (o.method == 1 && first32) ? std::max(0, o.fp32_progress_check) : 0;
                st.progress_floor = std::cbrt(o.tolerance);
                int const rk = krylov<float>(o.method, plan, st, first32 ? nullptr : st.y64, floor_rel, st.tol_abs,
                                             budget32, check_every, seed + k, eager, o.autotune != 0, tr, err);
                st.progress_check = 0;
                st.retire.clear();
                if (rk) return rk;
                first32 = false;
                S.iterations += tr.halfsteps; S.actions += tr.actions;
                if (int r = set_active(plan, st, a32, err)) return r;
                if (int r = residual(plan, st, S.actions, err)) return r;
                if (int r = fetch_resid()) return r;
                double mx = 0;
                for (uint32_t c = 0; c < st.nCols; ++c) {
                    if (!a32[c]) continue;
                    double const g = group_max(c);
                    double used = 0;   // applications of this FP32 solve to the group
                    for (int j = 0; j < 64; ++j) used = std::max(used, st.col_halfsteps[64 * c + j] - hs0[64 * c + j]);
                    mx = std::max(mx, g);
                    if (g <= o.tolerance) done[c] = 1;
                    else if (o.method == 1 && (used >= budget32 || st.lagging[c])) { prec64[c] = 1; ++fp64_groups; }
                    else if (g > 0.1 * gres[c]) { prec64[c] = 1; fresh64[c] = 1; ++fp64_groups; }
                    else if (o.method == 0) { prec64[c] = 1; ++fp64_groups; }
                    gres[c] = g;
                }
                if (o.echo > 0)
                    std::fprintf(stderr, "# gfa_solve fp32 step %d: %d half-steps, max FP64 residual %.3e\n", k, tr.halfsteps, mx);
            }
            bool any_fresh = false;
            for (uint32_t c = 0; c < st.nCols; ++c) any_fresh |= fresh64[c] != 0;
            if (any64 || any_fresh) {
                if (any_fresh) {
                    // groups leaving FP32 start over in FP64 from x = 0 (continuing from the stalled FP32
                    // iterate converges more slowly than a fresh FP64 solve): y = A 0 = 0, rhs = b
                    if (int r = set_active(plan, st, fresh64, err)) return r;
                    SegArgs const g{st.act_seg, st.seg, st.col_blocks, st.segpart};
                    hipLaunchKernelGGL((k_zero_seg<double>), dim3(st.nseg), dim3(256), 0, s, st.x64, g);
                    hipLaunchKernelGGL((k_zero_seg<double>), dim3(st.nseg), dim3(256), 0, s, st.y64, g);
                    HIP_TRY(hipGetLastError());
                }
                std::vector<uint8_t> b64(st.nCols, 0);
                for (uint32_t c = 0; c < st.nCols; ++c) b64[c] = (!done[c] && prec64[c]) ? 1 : 0;
                if (int r = set_active(plan, st, b64, err)) return r;
                TfqmrResult t2;
                if (int r = krylov<double>(o.method, plan, st, st.y64, 1e-30, st.tol_abs, o.max_iterations, check_every,
                                          seed + 1000 + k, eager, o.autotune != 0, t2, err)) return r;
                S.iterations += t2.halfsteps; S.actions += t2.actions;
                if (int r = set_active(plan, st, b64, err)) return r;
                if (int r = residual(plan, st, S.actions, err)) return r;
                if (int r = fetch_resid()) return r;
                double mx = 0;
                for (uint32_t c = 0; c < st.nCols; ++c) {
                    if (!b64[c]) continue;
                    double const g = group_max(c);
                    mx = std::max(mx, g);
                    if (g <= o.tolerance) done[c] = 1;
                    gres[c] = g;
                }
                if (o.echo > 0)
                    std::fprintf(stderr, "# gfa_solve fp64 finish %d: %d half-steps, max residual %.3e\n", k, t2.halfsteps, mx);
            }
            if (k > 0) S.refinements++;
        }
    }
    if (o.halfsteps)
        for (uint32_t q = 0; q < st.ncol; ++q) o.halfsteps[q] = int32_t(st.col_halfsteps[q]);
    if (o.residuals) std::memcpy(o.residuals, resid.data(), st.ncol * sizeof(double));
    S.max_residual = 0;
    S.converged_columns = 0;
    for (double v : resid) {
        S.max_residual = std::max(S.max_residual, v);
        S.converged_columns += (v <= o.tolerance);
    }
    if (o.X_device) HIP_TRY(hipMemcpyAsync(o.X_device, st.x64, st.nnzb * NBLK * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (rho) {
        double* out = st.segpart;  // >= 192 doubles per column group, 128 needed
        hipLaunchKernelGGL(k_diag, dim3(st.nCols), dim3(64), 0, s, out, st.x64, st.subset);
        HIP_TRY(hipMemcpyAsync(rho, out, size_t(st.ncol) * 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    ++st.host_syncs;
    S.host_syncs = st.host_syncs;
    S.graph_captures = st.graph_captures;
    S.fp64_groups = fp64_groups;
    st.group_cap.clear();
    S.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    S.flops = double(S.actions) * reference_flops(capi_host(plan), o.fp_bits);
    if (stats) *stats = S;
    return 0;
}

}  // namespace gfa
