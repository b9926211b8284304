// HIP kernels of the Green-function action y = (H - E S) x on gfx950.
//
// Data: x, y real_t[nnzb][2][64][64] (block, re/im plane, target point i, right-hand side j).
// Columns j are independent in every term, so a thread owns one (j, plane) pair and all loads and
// stores of a wave are 512 B (FP64) / 256 B (FP32) contiguous rows.
//
// Three launches per action:
//   1. k_sho_project  c_img = sum over the blocks touched by an atom image of H(x)H(y)H(z) x
//                     (include/green_dyadic.hxx:112-288, SHOprj), sum-factorised, split into
//                     chunks of blocks whose partial sums are reduced in a fixed order;
//   2. k_sho_atoms    collect images -> atom with Bloch phases, multiply the atom matrix
//                     (h - E s) dVol, broadcast back with conjugate phases
//                     (include/green_dyadic.hxx:600-852, SHOsum + SHOmul + SHOsum);
//   3. k_local        per block: (V - E) x (include/green_potential.hxx:16-139) + the three
//                     finite-difference Laplacians (include/green_kinetic.hxx:24-329) + the SHO
//                     add-back of all atoms touching the block (include/green_dyadic.hxx:325-518),
//                     written once.  The block is staged in LDS; neighbour blocks come through L2.
#include "gfa_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace gfa {

namespace {

constexpr int NBLK = 2 * 64 * 64;  // elements of one block

#define HIP_TRY(call)                                                              \
    do {                                                                           \
        hipError_t const e_ = (call);                                              \
        if (hipSuccess != e_) {                                                    \
            err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return 2;                                                              \
        }                                                                          \
    } while (0)

template <int L> struct Sho { static constexpr int n3 = ((L + 1) * (L + 2) * (L + 3)) / 6, n2 = ((L + 1) * (L + 2)) / 2; };

// ---------------------------------------------------------------------------------------------
// 1D Hermite-Gauss functions on the 4 grid points of a block, per direction
// (include/green_dyadic.hxx:33-107): xi = ((cube*4 + i4 + 0.5) h - a) / sigma,
// H0 = exp(-xi^2/2) inside the projection radius, H_{n+1} = xi H_n - n/2 H_{n-1}.
// xi^2 is exported in float: the 3D sphere mask is evaluated in float as in the reference.
// Evaluated without contraction so that xi^2 rounds as on the host.
// ---------------------------------------------------------------------------------------------
template <typename real_t>
__device__ __forceinline__ void hermite_1d(real_t (*H)[3][4], float (*xi2)[4], int t, int lmax,
                                           double const* apos, float const* cpos, double const* hgrid) {
#pragma clang fp contract(off)
    int const dir = t >> 2, i4 = t & 3;
    double const h = hgrid[dir];
    double const cube = cpos[dir];
    double const a = apos[dir];
    double const sig_inv = apos[3] * apos[3];
    double const R2 = hgrid[3] * hgrid[3];
    double const xi = sig_inv * ((cube * 4 + i4 + 0.5) * h - a);
    double const x2 = xi * xi;
    xi2[dir][i4] = float(x2);
    double const H0 = (x2 < R2) ? exp(-0.5 * x2) : 0.0;
    H[0][dir][i4] = real_t(H0);
    double hn = H0, hm = 0;
    for (int nu = 0; nu < lmax; ++nu) {
        double const hp = xi * hn - (0.5 * nu) * hm;
        H[nu + 1][dir][i4] = real_t(hp);
        hm = hn; hn = hp;
    }
}

// ---------------------------------------------------------------------------------------------
// 1. projection: one work item = (atom image, column, chunk of blocks); 128 threads = (j, plane)
// ---------------------------------------------------------------------------------------------
template <typename real_t, int L>
__global__ __launch_bounds__(128) void k_sho_project(
    real_t* __restrict__ partial, real_t const* __restrict__ x, ProjChunk const* __restrict__ chunks,
    uint32_t const* __restrict__ bucket, uint32_t const* __restrict__ chunk_coef,
    uint32_t const* __restrict__ prj_blocks, uint32_t const* __restrict__ rowindx,
    float const* __restrict__ row_pos, double const* __restrict__ image_pos, double hx, double hy, double hz,
    double rproj) {
    __shared__ real_t H[L + 1][3][4];
    __shared__ float xi2[3][4];
    __shared__ double apos[4];
    __shared__ float cpos[4];
    __shared__ double hgrid[4];
    uint32_t const ich = bucket[blockIdx.x];
    ProjChunk const ch = chunks[ich];
    int const tid = threadIdx.x, j = tid & 63, p = tid >> 6;
    if (tid < 4) apos[tid] = image_pos[4 * size_t(ch.image) + tid];
    if (tid == 0) { hgrid[0] = hx; hgrid[1] = hy; hgrid[2] = hz; hgrid[3] = rproj; }
    constexpr int n3 = Sho<L>::n3, n2 = Sho<L>::n2;
    real_t czyx[n3];
#pragma unroll
    for (int s = 0; s < n3; ++s) czyx[s] = 0;
    float const R2 = float(rproj * rproj);
    for (uint32_t b = ch.begin; b < ch.end; ++b) {
        uint32_t const inzb = prj_blocks[b];
        __syncthreads();
        if (tid < 3) cpos[tid] = row_pos[4 * size_t(rowindx[inzb]) + tid];
        __syncthreads();
        if (tid < 12) hermite_1d<real_t>(H, xi2, tid, L, apos, cpos, hgrid);
        __syncthreads();
        real_t const* xb = x + size_t(inzb) * NBLK + p * 4096 + j;
#pragma unroll
        for (int z = 0; z < 4; ++z) {
            float const d2z = xi2[2][z] - R2;
            if (d2z < 0) {
                real_t byx[n2];
#pragma unroll
                for (int k = 0; k < n2; ++k) byx[k] = 0;
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    float const d2yz = xi2[1][y] + d2z;
                    if (d2yz < 0) {
                        real_t ax[L + 1];
#pragma unroll
                        for (int k = 0; k <= L; ++k) ax[k] = 0;
#pragma unroll
                        for (int xx = 0; xx < 4; ++xx) {
                            float const d2 = xi2[0][xx] + d2yz;
                            if (d2 < 0) {
                                real_t const ps = xb[64 * ((z * 4 + y) * 4 + xx)];
#pragma unroll
                                for (int k = 0; k <= L; ++k) ax[k] += ps * H[k][0][xx];
                            }
                        }
#pragma unroll
                        for (int iy = 0; iy <= L; ++iy) {
                            int const iyx0 = (iy * (2 * L + 3 - iy)) >> 1;
                            real_t const hy_ = H[iy][1][y];
#pragma unroll
                            for (int ix = 0; ix <= L - iy; ++ix) byx[iyx0 + ix] += ax[ix] * hy_;
                        }
                    }
                }
                int s = 0;
#pragma unroll
                for (int iz = 0; iz <= L; ++iz) {
                    real_t const hz_ = H[iz][2][z];
#pragma unroll
                    for (int iy = 0; iy <= L - iz; ++iy) {
                        int const iyx0 = (iy * (2 * L + 3 - iy)) >> 1;
#pragma unroll
                        for (int ix = 0; ix <= L - iz - iy; ++ix) czyx[s++] += byx[iyx0 + ix] * hz_;
                    }
                }
            }
        }
    }
    real_t* out = partial + size_t(chunk_coef[ich]) * 128 + p * 64 + j;
#pragma unroll
    for (int s = 0; s < n3; ++s) out[size_t(s) * 128] = czyx[s];
}

// ---------------------------------------------------------------------------------------------
// 2. per (atom, column): collect images with phase, multiply (h - E s) dVol, broadcast with
//    conjugate phase.  64 threads = j; both planes per thread; coefficients staged in LDS.
// ---------------------------------------------------------------------------------------------
template <typename real_t>
__global__ __launch_bounds__(64) void k_sho_atoms(
    real_t* __restrict__ dcoef, real_t const* __restrict__ partial, uint32_t const* __restrict__ chunk_coef,
    uint32_t const* __restrict__ pair_chunk, uint32_t const* __restrict__ sum_start,
    uint32_t const* __restrict__ sum_index, uint32_t const* __restrict__ image_starts,
    int8_t const* __restrict__ atom_lmax, double const* __restrict__ atom_mat,
    uint64_t const* __restrict__ atom_mat_off, double const* __restrict__ image_phase, uint32_t nCols,
    double* __restrict__ cscratch, uint32_t const* __restrict__ atom_starts) {
    extern __shared__ double lds[];
    uint32_t const a = blockIdx.x, c = blockIdx.y;
    int const j = threadIdx.x;
    int const L = atom_lmax[a];
    int const nc = ((L + 1) * (L + 2) * (L + 3)) / 6;
    // atom coefficients of this column, private to each thread: in LDS up to 64 coefficients
    // (numax <= 5), else in a global scratch row
    double* cre = (nc <= 64) ? lds : cscratch + (size_t(atom_starts[a]) * nCols + size_t(c) * nc) * 128;
    double* cim = cre + nc * 64;
    for (int s = 0; s < nc; ++s) { cre[s * 64 + j] = 0; cim[s * 64 + j] = 0; }
    for (uint32_t k = sum_start[a]; k < sum_start[a + 1]; ++k) {
        uint32_t const img = sum_index[k];
        double const ph_re = image_phase[2 * img], ph_im = image_phase[2 * img + 1];
        uint32_t const c0 = pair_chunk[size_t(img) * nCols + c], c1 = pair_chunk[size_t(img) * nCols + c + 1];
        for (int s = 0; s < nc; ++s) {
            double vr = 0, vi = 0;
            for (uint32_t ch = c0; ch < c1; ++ch) {
                real_t const* src = partial + (size_t(chunk_coef[ch]) + s) * 128 + j;
                vr += double(src[0]);
                vi += double(src[64]);
            }
            cre[s * 64 + j] += ph_re * vr - ph_im * vi;
            cim[s * 64 + j] += ph_re * vi + ph_im * vr;
        }
    }
    double const* M = atom_mat + atom_mat_off[a];  // [2][nc][nc]
    // d = M c, complex double accumulation in ascending column order (include/green_dyadic.hxx:800-819)
    for (int i = 0; i < nc; ++i) {
        double dr = 0, di = 0;
        for (int k = 0; k < nc; ++k) {
            double const mr = M[i * nc + k], mi = M[nc * nc + i * nc + k];
            double const cr = cre[k * 64 + j], ci = cim[k * 64 + j];
            dr += mr * cr - mi * ci;
            di += mr * ci + mi * cr;
        }
        for (uint32_t k = sum_start[a]; k < sum_start[a + 1]; ++k) {
            uint32_t const img = sum_index[k];
            double const ph_re = image_phase[2 * img], ph_im = -image_phase[2 * img + 1];
            real_t* dst = dcoef + ((size_t(image_starts[img]) + i) * nCols + c) * 128 + j;
            dst[0] = real_t(ph_re * dr - ph_im * di);
            dst[64] = real_t(ph_re * di + ph_im * dr);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// 3. local part of the action per block, 256 threads = (j, plane, z half)
// ---------------------------------------------------------------------------------------------
template <typename real_t>
struct LocalArgs {
    real_t* y;
    real_t const* x;
    real_t const* dcoef;
    int32_t const* veff;
    int32_t const* nbr;
    double const* V;
    uint32_t const* add_start;
    uint32_t const* add_index;
    uint16_t const* colindx;
    uint32_t const* rowindx;
    float const* row_pos;
    double const* image_pos;
    int8_t const* image_lmax;
    uint32_t const* image_starts;
    uint64_t nnzb;
    uint32_t nCols;
    uint32_t nblk_per_xcd;   // XCD-contiguous mapping of work-groups to blocks, 0: identity
    real_t E_re, E_im;
    real_t c8[3][5];         // 8th-order weights per direction
    real_t c16[3][9];        // 16th-order weights per direction
    int nh[3];               // neighbour blocks per side per direction (1 or 2)
    double phase[3][2][2];
    double hgrid[4];
};

// neighbour value at point i of block code `nb` (>= 0 block, -1 absent, <= -2 periodic image with phase)
template <typename real_t>
__device__ __forceinline__ real_t nbr_value(real_t const* __restrict__ x, int32_t nb, int i, int p, int j,
                                            double const ph[2]) {
    if (nb >= 0) return x[size_t(nb) * NBLK + p * 4096 + i * 64 + j];
    if (nb == -1) return real_t(0);
    size_t const img = size_t(-nb - 2);
    real_t const v = x[img * NBLK + p * 4096 + i * 64 + j];
    real_t const w = x[img * NBLK + (1 - p) * 4096 + i * 64 + j];
    real_t const pr = real_t(ph[0]), pi = real_t(ph[1]) * real_t(1 - 2 * p);
    return pr * v - pi * w;
}

// SHO add-back of one atom image for this thread's two z planes
template <typename real_t, int L>
__device__ __forceinline__ void sho_add_atom(real_t (&out)[32], real_t const (*H)[3][4], float const (*xi2)[4],
                                             float R2, real_t const* __restrict__ dsrc, uint32_t nCols, int hz) {
    int s = 0;
#pragma unroll
    for (int iz = 0; iz <= L; ++iz) {
        real_t byx[4][4];
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int xx = 0; xx < 4; ++xx) byx[y][xx] = 0;
#pragma unroll
        for (int iy = 0; iy <= L - iz; ++iy) {
            real_t ax[4] = {0, 0, 0, 0};
#pragma unroll
            for (int ix = 0; ix <= L - iz - iy; ++ix) {
                real_t const ca = dsrc[size_t(s) * nCols * 128];
                ++s;
#pragma unroll
                for (int xx = 0; xx < 4; ++xx) ax[xx] += H[ix][0][xx] * ca;
            }
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                real_t const hy_ = H[iy][1][y];
#pragma unroll
                for (int xx = 0; xx < 4; ++xx) byx[y][xx] += hy_ * ax[xx];
            }
        }
#pragma unroll
        for (int zz = 0; zz < 2; ++zz) {
            int const z = 2 * hz + zz;
            real_t const hz_ = H[iz][2][z];
            float const d2z = xi2[2][z] - R2;
            if (d2z < 0) {
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    float const d2yz = xi2[1][y] + d2z;
                    if (d2yz < 0) {
#pragma unroll
                        for (int xx = 0; xx < 4; ++xx) {
                            float const d2 = xi2[0][xx] + d2yz;
                            if (d2 < 0) out[(zz * 4 + y) * 4 + xx] += hz_ * byx[y][xx];
                        }
                    }
                }
            }
        }
    }
}

template <typename real_t, int NHMAX>
__global__ __launch_bounds__(256) void k_local(LocalArgs<real_t> a) {
    __shared__ real_t xs[NBLK];
    __shared__ real_t H[8][3][4];
    __shared__ float xi2[3][4];
    __shared__ double apos[4];
    __shared__ float cpos[4];
    __shared__ double hgrid[4];

    uint64_t blk = blockIdx.x;
    if (a.nblk_per_xcd) blk = uint64_t(blockIdx.x & 7) * a.nblk_per_xcd + (blockIdx.x >> 3);
    if (blk >= a.nnzb) return;  // uniform over the work-group
    int const tid = threadIdx.x, j = tid & 63, p = (tid >> 6) & 1, hz = tid >> 7;

    // stage the block in LDS with 16-byte loads
    {
        using vec_t = typename std::conditional<sizeof(real_t) == 8, double2, float4>::type;
        constexpr int VE = 16 / sizeof(real_t);
        vec_t const* src = reinterpret_cast<vec_t const*>(a.x + blk * NBLK);
        vec_t* dst = reinterpret_cast<vec_t*>(xs);
#pragma unroll
        for (int k = 0; k < NBLK / VE / 256; ++k) dst[k * 256 + tid] = src[k * 256 + tid];
    }
    if (tid < 4) hgrid[tid] = a.hgrid[tid];
    __syncthreads();

    real_t out[32];
    real_t const* xsp = xs + p * 4096 + j;
    // ---- local potential: (V - E) x ----
    {
        int32_t const iv = a.veff[blk];
        real_t const Vim = a.E_im * real_t(1 - 2 * p);
        real_t const* xso = xs + (1 - p) * 4096 + j;
#pragma unroll
        for (int il = 0; il < 32; ++il) {
            int const i = hz * 32 + il;
            real_t const Vd = (iv < 0) ? real_t(0) : real_t(a.V[size_t(iv) * 64 + i]);
            out[il] = (Vd - a.E_re) * xsp[i * 64] + Vim * xso[i * 64];
        }
    }

    // ---- kinetic energy, one direction at a time ----
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int const stride = (d == 0) ? 1 : ((d == 1) ? 4 : 16);
        int32_t nbl[2], nbr_[2];
#pragma unroll
        for (int k = 0; k < NHMAX; ++k) {
            nbl[k] = a.nbr[((d * 2 + 0) * NHMAX + k) * a.nnzb + blk];
            nbr_[k] = a.nbr[((d * 2 + 1) * NHMAX + k) * a.nnzb + blk];
        }
        bool const wide = (NHMAX == 2) && (a.nh[d] == 2);
        // 16 lines of 4 points in direction d; lines are indexed by the two other coordinates
#pragma unroll
        for (int line = 0; line < 16; ++line) {
            int i0;  // first point of the line
            if (d == 0) i0 = line * 4;                                  // line = z*4 + y
            else if (d == 1) i0 = (line >> 2) * 16 + (line & 3);        // line = z*4 + x
            else i0 = line;                                             // line = y*4 + x
            // which of this line's points belong to this thread
            int zsel;  // for d < 2: line belongs to the thread if its z is in the thread's half
            if (d < 2) { zsel = line >> 2; if ((zsel >> 1) != hz) continue; }
            real_t w[20];  // positions -8..11 relative to the block start along the line
#pragma unroll
            for (int q = 0; q < 20; ++q) w[q] = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) w[8 + q] = xsp[(i0 + q * stride) * 64];
            // for the z direction the 9-point stencil of this thread's two planes needs only part
            // of the neighbour planes
            int const qlo = (d == 2 && !wide) ? 2 * hz : 0;       // left neighbour planes qlo..3
            int const qhi = (d == 2 && !wide) ? 2 * hz + 1 : 3;   // right neighbour planes 0..qhi
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (q >= qlo) w[4 + q] = nbr_value(a.x, nbl[0], i0 + q * stride, p, j, a.phase[d][0]);
                if (q <= qhi) w[12 + q] = nbr_value(a.x, nbr_[0], i0 + q * stride, p, j, a.phase[d][1]);
                if (wide) {
                    if (d < 2 || q >= 2 * hz) w[q] = nbr_value(a.x, nbl[NHMAX - 1], i0 + q * stride, p, j, a.phase[d][0]);
                    if (d < 2 || q <= 2 * hz + 1) w[16 + q] = nbr_value(a.x, nbr_[NHMAX - 1], i0 + q * stride, p, j, a.phase[d][1]);
                }
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                int const i = i0 + k * stride;            // point in the block
                int const il = i - hz * 32;               // local output index
                if (d == 2 && (il < 0 || il >= 32)) continue;
                int const c = 8 + k;
                real_t t;
                if (wide) {
                    t = a.c16[d][0] * w[c];
#pragma unroll
                    for (int m = 1; m <= 8; ++m) t += a.c16[d][m] * (w[c - m] + w[c + m]);
                } else {
                    t = a.c8[d][0] * w[c];
#pragma unroll
                    for (int m = 1; m <= 4; ++m) t += a.c8[d][m] * (w[c - m] + w[c + m]);
                }
                out[il] += t;
            }
        }
    }

    // ---- SHO add-back of all atom images touching this block ----
    uint32_t const b0 = a.add_start[blk], b1 = a.add_start[blk + 1];
    if (b1 > b0) {
        uint32_t const col = a.colindx[blk];
        float const R2 = float(a.hgrid[3] * a.hgrid[3]);
        for (uint32_t b = b0; b < b1; ++b) {
            uint32_t const img = a.add_index[b];
            int const L = a.image_lmax[img];
            __syncthreads();
            if (tid < 4) apos[tid] = a.image_pos[4 * size_t(img) + tid];
            if (tid < 3) cpos[tid] = a.row_pos[4 * size_t(a.rowindx[blk]) + tid];
            __syncthreads();
            if (tid < 12) hermite_1d<real_t>(H, xi2, tid, L, apos, cpos, hgrid);
            __syncthreads();
            real_t const* dsrc = a.dcoef + (size_t(a.image_starts[img]) * a.nCols + col) * 128 + p * 64 + j;
            switch (L) {
                case 0: sho_add_atom<real_t, 0>(out, H, xi2, R2, dsrc, a.nCols, hz); break;
                case 1: sho_add_atom<real_t, 1>(out, H, xi2, R2, dsrc, a.nCols, hz); break;
                case 2: sho_add_atom<real_t, 2>(out, H, xi2, R2, dsrc, a.nCols, hz); break;
                case 3: sho_add_atom<real_t, 3>(out, H, xi2, R2, dsrc, a.nCols, hz); break;
                case 4: sho_add_atom<real_t, 4>(out, H, xi2, R2, dsrc, a.nCols, hz); break;
                case 5: sho_add_atom<real_t, 5>(out, H, xi2, R2, dsrc, a.nCols, hz); break;
                case 6: sho_add_atom<real_t, 6>(out, H, xi2, R2, dsrc, a.nCols, hz); break;
                default: sho_add_atom<real_t, 7>(out, H, xi2, R2, dsrc, a.nCols, hz); break;
            }
        }
    }

    real_t* yb = a.y + blk * NBLK + p * 4096 + j;
#pragma unroll
    for (int il = 0; il < 32; ++il) yb[(hz * 32 + il) * 64] = out[il];
}

template <typename T>
int dalloc(DevicePlan& d, T*& ptr, size_t n, std::string& err) {
    ptr = nullptr;
    if (n == 0) n = 1;
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, n * sizeof(T)));
    d.allocations.push_back(p);
    d.bytes += n * sizeof(T);
    ptr = (T*)p;
    return 0;
}

template <typename T>
int dupload(DevicePlan& d, T*& ptr, std::vector<T> const& v, std::string& err) {
    if (int s = dalloc(d, ptr, v.size(), err)) return s;
    if (!v.empty()) HIP_TRY(hipMemcpy(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

}  // namespace

void device_free(DevicePlan& d) {
    for (void* p : d.allocations) (void)hipFree(p);
    d = DevicePlan();
}

double reference_flops(Plan const& p, int fp_bits) {
    (void)fp_bits;
    double const nnzb = double(p.nnzb);
    double n = 0;
    // potential: (1 + 2 [complex E]) * nnzb * 64^2 * 2 (include/green_potential.hxx:171-177)
    n += (1.0 + 2.0 * (p.E.imag() != 0)) * nnzb * 4096 * 2;
    // kinetic: nnzb * (2 nFD + 1) * 2 * 64^2 * 2 per direction (include/green_kinetic.hxx:381)
    for (int d = 0; d < 3; ++d) n += nnzb * (2.0 * p.kinetic[d].fd_range + 1) * 2 * 4096 * 2;
    // dyadic (include/dyadic_plan.hxx:93-103, src/dyadic_plan.cxx:103-135), R1C2 = 2, Noco = 1
    if (p.nAtomImages > 0) {
        double add = 0, gen = 0, mul = 0, sum = 0;
        for (uint32_t k = 0; k < p.sho_add.index.size(); ++k) {
            int const L = p.image_lmax[p.sho_add.index[k]];
            add += 64.0 * 2 * (64 * (L + 1) + 16 * ((L + 1) * (L + 2)) / 2 + 4 * ((L + 1) * (L + 2) * (L + 3)) / 6);
            gen += 12.0 * (7 + 4 * L);
        }
        for (uint32_t i = 0; i < p.nAtoms; ++i) mul += double(nsho(p.image_lmax[i])) * nsho(p.image_lmax[i]);
        mul *= 2.0 * p.nCols * 64;
        if (p.nAtomImages > p.nAtoms) {
            for (uint32_t i = 0; i < p.nAtomImages; ++i) sum += nsho(p.image_lmax[i]);
            sum *= 2.0 * p.nCols * 64;
        }
        n += mul * 4 + 2 * sum * 4 + 2 * gen + 2 * add * 2;
    }
    return n;
}

int device_upload_plan(Plan const& p, DevicePlan& d, std::string& err) {
    if (d.uploaded) return 0;
    HIP_TRY(hipGetDevice(&d.device));
    d.nnzb = p.nnzb; d.nCols = p.nCols; d.nRows = p.nRows;
    d.nAtoms = p.nAtoms; d.nAtomImages = p.nAtomImages;
    d.natomcoeffs = p.natomcoeffs();
    d.nh_max = 1;
    for (int dd = 0; dd < 3; ++dd) {
        d.fd_range[dd] = p.kinetic[dd].fd_range;
        if (p.kinetic[dd].fd_range == 8) d.nh_max = 2;
    }
    int s = 0;
    if ((s = dupload(d, d.colindx, p.colindx, err))) return s;
    if ((s = dupload(d, d.rowindx, p.rowindx, err))) return s;
    if ((s = dupload(d, d.veff, p.veff_index, err))) return s;
    if ((s = dupload(d, d.row_pos, p.row_pos, err))) return s;

    // neighbour codes from the finite-difference lists: the entry before (after) a block in its
    // list is its left (right) neighbour; halo entries 0 are absent, negative ones periodic images
    {
        int const NH = d.nh_max;
        std::vector<int32_t> nb(size_t(3) * 2 * NH * p.nnzb, -1);
        for (int dd = 0; dd < 3; ++dd) {
            auto const& kp = p.kinetic[dd];
            for (size_t l = 0; l + 1 < kp.start.size(); ++l) {
                int32_t const* L = &kp.list[kp.start[l]];
                int const len = int(kp.start[l + 1] - kp.start[l]);
                int const n = len - 2 * NHALO;
                for (int i = 0; i < n; ++i) {
                    uint32_t const inzb = uint32_t(L[NHALO + i] - CUBE_EXISTS);
                    for (int side = 0; side < 2; ++side) {
                        for (int k = 0; k < NH; ++k) {
                            int const pos = NHALO + i + (side ? (k + 1) : -(k + 1));
                            int32_t const e = L[pos];
                            int32_t code;
                            if (e > 0) code = e - CUBE_EXISTS;
                            else if (e == 0) code = -1;
                            else code = -2 - (-e - CUBE_EXISTS);
                            nb[((size_t(dd) * 2 + side) * NH + k) * p.nnzb + inzb] = code;
                        }
                    }
                }
            }
        }
        if ((s = dupload(d, d.nbr, nb, err))) return s;
    }

    // dyadic structures
    if ((s = dupload(d, d.add_start, p.sho_add.start, err))) return s;
    if ((s = dupload(d, d.add_index, p.sho_add.index, err))) return s;
    if ((s = dupload(d, d.image_pos, p.image_pos, err))) return s;
    if ((s = dupload(d, d.image_lmax, p.image_lmax, err))) return s;
    if ((s = dupload(d, d.image_starts, p.image_starts, err))) return s;
    if ((s = dupload(d, d.sum_start, p.sho_sum.start, err))) return s;
    if ((s = dupload(d, d.sum_index, p.sho_sum.index, err))) return s;
    if ((s = dupload(d, d.atom_lmax, p.atom_lmax, err))) return s;
    if ((s = dupload(d, d.atom_mat_off, p.atom_mat_off, err))) return s;
    if ((s = dupload(d, d.atom_starts, p.atom_starts, err))) return s;
    d.ncoeffs = p.ncoeffs();
    d.max_atom_nc = 0;
    for (uint32_t a = 0; a < p.nAtoms; ++a) d.max_atom_nc = std::max<uint32_t>(d.max_atom_nc, nsho(p.atom_lmax[a]));

    // projection work items: (image, column) pairs split into chunks of at most K blocks
    {
        uint64_t pairs = 0;
        for (auto const& c : p.sho_prj) pairs += c.index.size();
        uint32_t K = uint32_t(std::min<uint64_t>(16, std::max<uint64_t>(2, pairs / 4096)));
        std::vector<ProjChunk> chunks;
        std::vector<uint32_t> blocks, coef(1, 0), pair_chunk(size_t(p.nAtomImages) * p.nCols + 1, 0);
        for (uint32_t img = 0; img < p.nAtomImages; ++img) {
            int const nc = nsho(p.image_lmax[img]);
            for (uint32_t c = 0; c < p.nCols; ++c) {
                auto const& s_ = p.sho_prj[c];
                uint32_t const b0 = s_.start[img], b1 = s_.start[img + 1];
                for (uint32_t b = b0; b < b1; b += K) {
                    ProjChunk ch;
                    ch.image = img; ch.col = c;
                    ch.begin = uint32_t(blocks.size());
                    for (uint32_t k = b; k < std::min(b1, b + K); ++k) blocks.push_back(s_.index[k]);
                    ch.end = uint32_t(blocks.size());
                    chunks.push_back(ch);
                    coef.push_back(coef.back() + uint32_t(nc));
                }
                pair_chunk[size_t(img) * p.nCols + c + 1] = uint32_t(chunks.size());
            }
        }
        d.nchunks = uint32_t(chunks.size());
        d.total_chunk_coefs = coef.back();
        std::vector<uint32_t> bucket;
        for (int L = 0; L <= 7; ++L) {
            d.bucket_start[L] = uint32_t(bucket.size());
            for (uint32_t k = 0; k < chunks.size(); ++k)
                if (p.image_lmax[chunks[k].image] == L) bucket.push_back(k);
        }
        d.bucket_start[8] = uint32_t(bucket.size());
        if ((s = dupload(d, d.chunks, chunks, err))) return s;
        if ((s = dupload(d, d.prj_blocks, blocks, err))) return s;
        if ((s = dupload(d, d.chunk_coef, coef, err))) return s;
        if ((s = dupload(d, d.pair_chunk, pair_chunk, err))) return s;
        if ((s = dupload(d, d.bucket_chunks, bucket, err))) return s;
    }

    if ((s = dalloc(d, d.V, p.ncell() * 64, err))) return s;
    if ((s = dalloc(d, d.atom_mat, p.atom_matrix.size(), err))) return s;
    if ((s = dalloc(d, d.image_phase, p.image_phase.size(), err))) return s;
    d.uploaded = true;
    return 0;
}

int device_upload_values(Plan const& p, DevicePlan& d, std::string& err) {
    HIP_TRY(hipMemcpy(d.V, p.V_cell.data(), p.V_cell.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!p.atom_matrix.empty())
        HIP_TRY(hipMemcpy(d.atom_mat, p.atom_matrix.data(), p.atom_matrix.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!p.image_phase.empty())
        HIP_TRY(hipMemcpy(d.image_phase, p.image_phase.data(), p.image_phase.size() * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

template <typename real_t>
static Workspace<real_t>& workspace(DevicePlan& d);
template <> Workspace<double>& workspace<double>(DevicePlan& d) { return d.ws64; }
template <> Workspace<float>& workspace<float>(DevicePlan& d) { return d.ws32; }

template <typename real_t>
int device_action(DevicePlan& d, ActionScalars const& s, real_t* y, real_t const* x, hipStream_t stream,
                  std::string& err) {
    auto& ws = workspace<real_t>(d);
    if (d.nAtomImages > 0 && !ws.dcoef) {
        if (int r = dalloc(d, ws.partial, size_t(d.total_chunk_coefs) * 128, err)) return r;
        if (int r = dalloc(d, ws.dcoef, size_t(d.natomcoeffs) * d.nCols * 128, err)) return r;
    }
    if (d.nAtomImages > 0 && d.nchunks > 0) {
        for (int L = 0; L <= 7; ++L) {
            uint32_t const n = d.bucket_start[L + 1] - d.bucket_start[L];
            if (!n) continue;
            uint32_t const* bucket = d.bucket_chunks + d.bucket_start[L];
#define LAUNCH_PRJ(LL) hipLaunchKernelGGL((k_sho_project<real_t, LL>), dim3(n), dim3(128), 0, stream, ws.partial, x, \
                d.chunks, bucket, d.chunk_coef, d.prj_blocks, d.rowindx, d.row_pos, d.image_pos, \
                s.hgrid[0], s.hgrid[1], s.hgrid[2], s.hgrid[3])
            switch (L) {
                case 0: LAUNCH_PRJ(0); break; case 1: LAUNCH_PRJ(1); break;
                case 2: LAUNCH_PRJ(2); break; case 3: LAUNCH_PRJ(3); break;
                case 4: LAUNCH_PRJ(4); break; case 5: LAUNCH_PRJ(5); break;
                case 6: LAUNCH_PRJ(6); break; default: LAUNCH_PRJ(7); break;
            }
#undef LAUNCH_PRJ
            HIP_TRY(hipGetLastError());
        }
        if (d.max_atom_nc > 64 && !d.cscratch) {
            if (int r = dalloc(d, d.cscratch, size_t(d.ncoeffs) * d.nCols * 128, err)) return r;
        }
        size_t const lds = size_t(2) * std::min<uint32_t>(d.max_atom_nc, 64) * 64 * sizeof(double);
        hipLaunchKernelGGL((k_sho_atoms<real_t>), dim3(d.nAtoms, d.nCols), dim3(64), lds, stream, ws.dcoef,
                           ws.partial, d.chunk_coef, d.pair_chunk, d.sum_start, d.sum_index, d.image_starts,
                           d.atom_lmax, d.atom_mat, d.atom_mat_off, d.image_phase, d.nCols, d.cscratch,
                           d.atom_starts);
        HIP_TRY(hipGetLastError());
    }
    LocalArgs<real_t> a;
    a.y = y; a.x = x; a.dcoef = ws.dcoef; a.veff = d.veff; a.nbr = d.nbr; a.V = d.V;
    a.add_start = d.add_start; a.add_index = d.add_index; a.colindx = d.colindx; a.rowindx = d.rowindx;
    a.row_pos = d.row_pos; a.image_pos = d.image_pos; a.image_lmax = d.image_lmax; a.image_starts = d.image_starts;
    a.nnzb = d.nnzb; a.nCols = d.nCols;
    a.E_re = real_t(s.E_re); a.E_im = real_t(s.E_im);
    static int const c8i[5] = {-14350, 8064, -1008, 128, -9};
    static int const c16i[9] = {-924708642, 538137600, -94174080, 22830080, -5350800, 1053696, -156800, 15360, -735};
    for (int dd = 0; dd < 3; ++dd) {
        double const n8 = s.prefactor[dd] / 5040., n16 = s.prefactor[dd] / 302702400.;
        for (int m = 0; m < 5; ++m) a.c8[dd][m] = real_t(c8i[m] * n8);
        for (int m = 0; m < 9; ++m) a.c16[dd][m] = real_t(c16i[m] * n16);
        a.nh[dd] = (d.fd_range[dd] == 8) ? 2 : 1;
        for (int u = 0; u < 2; ++u) for (int v = 0; v < 2; ++v) a.phase[dd][u][v] = s.phase[dd][u][v];
    }
    for (int k = 0; k < 4; ++k) a.hgrid[k] = s.hgrid[k];
    // work-groups are dealt round-robin to the 8 XCDs: give each XCD a contiguous range of blocks
    uint64_t const per = (d.nnzb + 7) / 8;
    a.nblk_per_xcd = uint32_t(per);
    uint64_t const grid = per * 8;
    if (grid > 0x7fffffffull) { err = "too many blocks for one launch"; return 1; }
    if (d.nh_max == 2) hipLaunchKernelGGL((k_local<real_t, 2>), dim3(uint32_t(grid)), dim3(256), 0, stream, a);
    else hipLaunchKernelGGL((k_local<real_t, 1>), dim3(uint32_t(grid)), dim3(256), 0, stream, a);
    HIP_TRY(hipGetLastError());
    return 0;
}

template int device_action<double>(DevicePlan&, ActionScalars const&, double*, double const*, hipStream_t, std::string&);
template int device_action<float>(DevicePlan&, ActionScalars const&, float*, float const*, hipStream_t, std::string&);

}  // namespace gfa
