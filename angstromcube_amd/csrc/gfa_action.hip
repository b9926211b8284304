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
//   3. the local kernel, per (block, column slice): (V - E) x (include/green_potential.hxx:16-139) + the
//                     three finite-difference Laplacians (include/green_kinetic.hxx:24-329) + the SHO add-back
//                     of the atoms touching the block with numax <= 5 (include/green_dyadic.hxx:325-518),
//                     written once.  Two kinds, picked per plan and precision by the autotuner: k_local (LDS
//                     accumulator) and k_plane (register accumulator per z plane).  The slice is staged in
//                     LDS; neighbour blocks come through L2.  Images with numax 6, 7 are added afterwards by
//                     k_sho_add_wide.
#include "gfa_device.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdio>

namespace gfa {

namespace {

constexpr int NBLK = 2 * 64 * 64;  // elements of one block

#ifndef LOCAL_TILE
#define LOCAL_TILE 6   // xy tile edge (blocks) of the local kernel's processing order (5 to 7 measured best)
#endif

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

// The Hermite functions of every (atom image, target row) pair that the projection and the add-back use, computed
// once per plan with hermite_1d itself (so the values are those the kernels computed on the fly before):
// htab[slot][8][3][4] (FP64, numax <= 7) and xtab[slot][3][4] = xi^2 (float).
__global__ __launch_bounds__(64) void k_hermite_table(double* __restrict__ htab, float* __restrict__ xtab,
                                                      uint32_t const* __restrict__ slot_img,
                                                      uint32_t const* __restrict__ slot_row,
                                                      double const* __restrict__ image_pos,
                                                      int8_t const* __restrict__ image_lmax,
                                                      float const* __restrict__ row_pos, double hx, double hy,
                                                      double hz, double rproj) {
    uint32_t const sl = blockIdx.x;
    int const t = threadIdx.x;
    if (t >= 12) return;
    uint32_t const img = slot_img[sl], row = slot_row[sl];
    double const hgrid[4] = {hx, hy, hz, rproj};
    hermite_1d<double>(reinterpret_cast<double (*)[3][4]>(htab + size_t(sl) * 96),
                       reinterpret_cast<float (*)[4]>(xtab + size_t(sl) * 12), t, image_lmax[img],
                       image_pos + 4 * size_t(img), row_pos + 4 * size_t(row), hgrid);
}

// stage the table entry of `slot` into the LDS arrays H[L+1][3][4] (real_t) and xi2[3][4]; no barrier
template <typename real_t>
__device__ __forceinline__ void stage_hermite(real_t (*H)[3][4], float (*xi2)[4], double const* __restrict__ htab,
                                              float const* __restrict__ xtab, uint32_t slot, int L, int tid, int T) {
    double const* src = htab + size_t(slot) * 96;
    for (int e = tid; e < (L + 1) * 12; e += T) H[e / 12][(e % 12) / 4][e % 4] = real_t(src[e]);
    if (tid < 12) xi2[tid / 4][tid % 4] = xtab[size_t(slot) * 12 + tid];
}

// is point (y, x) of z plane zq inside the projection sphere (the 3D mask evaluated in float from xi^2)
__device__ __forceinline__ bool sho_inside(float const (*xi2)[4], float R2, int zq, int y, int xx) {
    float const d2z = xi2[2][zq] - R2;
    float const d2yz = xi2[1][y] + d2z;
    float const d2 = xi2[0][xx] + d2yz;
    return (d2z < 0) && (d2yz < 0) && (d2 < 0);
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
    double rproj, uint8_t const* __restrict__ col_active, double const* __restrict__ htab,
    float const* __restrict__ xtab, uint32_t const* __restrict__ prj_slot) {
    __shared__ real_t H[L + 1][3][4];
    __shared__ float xi2[3][4];
    uint32_t const ich = bucket[blockIdx.x];
    ProjChunk const ch = chunks[ich];
    if (col_active && !col_active[ch.col]) return;  // a retired column group
    int const tid = threadIdx.x, j = tid & 63, p = tid >> 6;
    (void)rowindx; (void)row_pos; (void)image_pos; (void)hx; (void)hy; (void)hz;
    constexpr int n3 = Sho<L>::n3, n2 = Sho<L>::n2;
    real_t czyx[n3];
#pragma unroll
    for (int s = 0; s < n3; ++s) czyx[s] = 0;
    float const R2 = float(rproj * rproj);
    for (uint32_t b = ch.begin; b < ch.end; ++b) {
        uint32_t const inzb = prj_blocks[b];
        __syncthreads();
        stage_hermite<real_t>(H, xi2, htab, xtab, prj_slot[b], L, tid, 128);
        __syncthreads();
        real_t const* xb = x + size_t(inzb) * NBLK + p * 4096 + j;
#pragma unroll
        for (int z = 0; z < 4; ++z) {
            float const d2z = xi2[2][z] - R2;
            if (d2z < 0) {
                // the 16 points of the plane are loaded together (unconditionally, all in the block) and the
                // sphere mask is applied as a select, so the loads are in flight at once
                real_t v[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) v[k] = xb[64 * (z * 16 + k)];
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
                            real_t const ps = (d2 < 0) ? v[y * 4 + xx] : real_t(0);
#pragma unroll
                            for (int k = 0; k <= L; ++k) ax[k] += ps * H[k][0][xx];
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
// 1b. projection on the matrix cores (the dense form): per block of the chunk, C[s][n] += sum_i P[s][i] X[i][n]
//     with P[s][i] = H_x[ix_s](x_i) H_y[iy_s](y_i) H_z[iz_s](z_i) inside the projection sphere (0 outside), i the 64
//     points, n = plane * 64 + column.  P is built in LDS per block, X comes straight from global memory as the
//     B operand.  4 waves; wave w owns the columns n = 32 w .. 32 w + 31 (two 16-column tiles) and every
//     16-coefficient tile.  Operand layout of the 16x16x4 MFMA (64 lanes): A[m][k] from lane m + 16 k, B[k][n]
//     from lane n + 16 k; accumulator r of a lane holds D[row][lane % 16] with row = 4 (lane / 16) + r in FP32 and
//     row = lane / 16 + 4 r in FP64 (measured on gfx950).
// ---------------------------------------------------------------------------------------------
template <typename real_t> struct Mfma;
template <> struct Mfma<float> {
    typedef float __attribute__((ext_vector_type(4))) acc_t;
    static __device__ __forceinline__ acc_t mac(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kl, int r) { return 4 * kl + r; }
};
template <> struct Mfma<double> {
    typedef double __attribute__((ext_vector_type(4))) acc_t;
    static __device__ __forceinline__ acc_t mac(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kl, int r) { return kl + 4 * r; }
};

template <typename real_t, int L>
__global__ __launch_bounds__(256) void k_sho_project_mfma(
    real_t* __restrict__ partial, real_t const* __restrict__ x, ProjChunk const* __restrict__ chunks,
    uint32_t const* __restrict__ bucket, uint32_t const* __restrict__ chunk_coef,
    uint32_t const* __restrict__ prj_blocks, double rproj, uint8_t const* __restrict__ col_active,
    double const* __restrict__ htab, float const* __restrict__ xtab, uint32_t const* __restrict__ prj_slot) {
    using acc_t = typename Mfma<real_t>::acc_t;
    constexpr int NC = Sho<L>::n3, MT = (NC + 15) / 16, NCP = 16 * MT;
    __shared__ real_t P[64 * NCP];           // [point][coefficient], zero beyond NC and outside the sphere
    __shared__ real_t H[L + 1][3][4];
    __shared__ float xi2[3][4];
    __shared__ uint16_t sxyz[NCP];           // coefficient s -> ix | iy << 4 | iz << 8
    uint32_t const ich = bucket[blockIdx.x];
    ProjChunk const ch = chunks[ich];
    if (col_active && !col_active[ch.col]) return;  // a retired column group
    int const tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) {
        int s = 0;
        for (int iz = 0; iz <= L; ++iz)
            for (int iy = 0; iy <= L - iz; ++iy)
                for (int ix = 0; ix <= L - iz - iy; ++ix) sxyz[s++] = uint16_t(ix | (iy << 4) | (iz << 8));
        for (; s < NCP; ++s) sxyz[s] = 0xFFFF;
    }
    float const R2 = float(rproj * rproj);
    acc_t acc[MT][2];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < 2; ++t) acc[mt][t] = acc_t{0, 0, 0, 0};
    int const p = w >> 1, jb = (w & 1) * 32 + (lane & 15), kl = lane >> 4;   // B operand: column jb (+16), row kl
    for (uint32_t b = ch.begin; b < ch.end; ++b) {
        uint32_t const inzb = prj_blocks[b];
        // the B operands of the whole block (16 steps of 4 points, 2 column tiles), issued before P is built
        real_t const* xb = x + size_t(inzb) * NBLK + p * 4096 + jb;
        real_t bv[16][2];
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) {
            bv[ks][0] = xb[(ks * 4 + kl) * 64];
            bv[ks][1] = xb[(ks * 4 + kl) * 64 + 16];
        }
        __syncthreads();   // P of the previous block is no longer read
        stage_hermite<real_t>(H, xi2, htab, xtab, prj_slot[b], L, tid, 256);
        __syncthreads();
        for (int e = tid; e < 64 * NCP; e += 256) {
            int const i = e / NCP, s = e % NCP;
            int const xx = i & 3, yy = (i >> 2) & 3, zz = i >> 4;
            uint16_t const c = sxyz[s];
            real_t v = 0;
            if (c != 0xFFFF && sho_inside(xi2, R2, zz, yy, xx))
                v = H[c & 15][0][xx] * H[(c >> 4) & 15][1][yy] * H[c >> 8][2][zz];
            P[e] = v;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                real_t const av = P[(ks * 4 + kl) * NCP + mt * 16 + (lane & 15)];
                acc[mt][0] = Mfma<real_t>::mac(av, bv[ks][0], acc[mt][0]);
                acc[mt][1] = Mfma<real_t>::mac(av, bv[ks][1], acc[mt][1]);
            }
        }
    }
    real_t* out = partial + size_t(chunk_coef[ich]) * 128 + p * 64 + jb;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int const s = mt * 16 + Mfma<real_t>::row(kl, r);
            if (s < NC) {
                out[size_t(s) * 128] = acc[mt][0][r];
                out[size_t(s) * 128 + 16] = acc[mt][1][r];
            }
        }
}

// ---------------------------------------------------------------------------------------------
// 2. per (atom, column): collect images with phase, multiply (h - E s) dVol, broadcast with
//    conjugate phase.  ATOM_WAVES waves; lane = column j, both planes per thread.  The waves share
//    the collected coefficients (LDS up to 64 coefficients, numax <= 5, else a global scratch row):
//    coefficient s is collected by wave s % ATOM_WAVES, output row i computed by wave i % ATOM_WAVES.
// ---------------------------------------------------------------------------------------------
#define ATOM_WAVES 8
template <typename real_t>
__global__ __launch_bounds__(64 * ATOM_WAVES) void k_sho_atoms(
    real_t* __restrict__ dcoef, real_t const* __restrict__ partial, uint32_t const* __restrict__ chunk_coef,
    uint32_t const* __restrict__ pair_chunk, uint32_t const* __restrict__ sum_start,
    uint32_t const* __restrict__ sum_index, uint32_t const* __restrict__ image_starts,
    int8_t const* __restrict__ atom_lmax, double const* __restrict__ atom_mat,
    uint64_t const* __restrict__ atom_mat_off, double const* __restrict__ image_phase, uint32_t nCols,
    double* __restrict__ cscratch, uint32_t const* __restrict__ atom_starts, double const* __restrict__ E_col,
    uint8_t const* __restrict__ col_active) {
    extern __shared__ double lds[];
    uint32_t const a = blockIdx.x, c = blockIdx.y;
    if (col_active && !col_active[c]) return;  // a retired column group
    // the wave index is uniform: the rows i = w, w + ATOM_WAVES, ... of M are read with scalar loads
    int const j = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int const L = atom_lmax[a];
    int const nc = ((L + 1) * (L + 2) * (L + 3)) / 6;
    bool const in_lds = nc <= 64;
    double* cre = in_lds ? lds : cscratch + (size_t(atom_starts[a]) * nCols + size_t(c) * nc) * 128;
    double* cim = cre + nc * 64;
    double const* HS = atom_mat + atom_mat_off[a];  // [2][nc][nc]: h dVol, s dVol
    double const Er = E_col[2 * c], Ei = E_col[2 * c + 1];
    uint32_t const k0 = sum_start[a], k1 = sum_start[a + 1];
    for (int s = w; s < nc; s += ATOM_WAVES) {
        double ar = 0, ai = 0;
        for (uint32_t k = k0; k < k1; ++k) {
            uint32_t const img = sum_index[k];
            double const ph_re = image_phase[2 * img], ph_im = image_phase[2 * img + 1];
            uint32_t const c0 = pair_chunk[size_t(img) * nCols + c], c1 = pair_chunk[size_t(img) * nCols + c + 1];
            double vr = 0, vi = 0;
            for (uint32_t ch = c0; ch < c1; ++ch) {
                real_t const* src = partial + (size_t(chunk_coef[ch]) + s) * 128 + j;
                vr += double(src[0]);
                vi += double(src[64]);
            }
            ar += ph_re * vr - ph_im * vi;
            ai += ph_re * vi + ph_im * vr;
        }
        if (in_lds) { lds[s * 64 + j] = ar; lds[(nc + s) * 64 + j] = ai; }   // LDS addressed as such
        else { cre[s * 64 + j] = ar; cim[s * 64 + j] = ai; }
    }
    __syncthreads();
    // d = M c with M = (h - E s) dVol of this column's energy, complex double accumulation in ascending
    // column order (include/green_dyadic.hxx:800-819, src/green_function.cxx:97-98); the rows of M are
    // wave-uniform (scalar loads), the coefficients per column come from LDS (numax <= 5) or the scratch
    auto mv = [&](int i, double const* __restrict__ cr_, double const* __restrict__ ci_, double& dr, double& di) {
#pragma unroll 4
        for (int k = 0; k < nc; ++k) {
            double const hd = HS[i * nc + k], sd = HS[nc * nc + i * nc + k];
            double const mr = hd - Er * sd, mi = -Ei * sd;
            double const cr = cr_[k * 64 + j], ci = ci_[k * 64 + j];
            dr += mr * cr - mi * ci;
            di += mr * ci + mi * cr;
        }
    };
    for (int i = w; i < nc; i += ATOM_WAVES) {
        double dr = 0, di = 0;
        if (in_lds) mv(i, lds, lds + nc * 64, dr, di);
        else mv(i, cre, cim, dr, di);
        for (uint32_t k = k0; k < k1; ++k) {
            uint32_t const img = sum_index[k];
            double const ph_re = image_phase[2 * img], ph_im = -image_phase[2 * img + 1];
            real_t* dst = dcoef + ((size_t(image_starts[img]) + i) * nCols + c) * 128 + j;
            dst[0] = real_t(ph_re * dr - ph_im * di);
            dst[64] = real_t(ph_re * di + ph_im * dr);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// 3. local part of the action: one work-group per (block, slice of JW of the 64 columns), 8 JW threads, each
//    owning points of one (column, plane).  The slice is staged in LDS; neighbour blocks are read with
//    branch-free addressing (an absent neighbour reads a zero block) so that all loads of a line are in flight
//    together.
// ---------------------------------------------------------------------------------------------
// 16 bytes of y, stored past the caches' replacement order (non-temporal): the output of the local kernels is read
// by the next kernel only, so keeping it out of L2 leaves room there for the neighbour blocks of x
__device__ __forceinline__ void st_y16(float* p, float4 const& v) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    f4v t; t[0] = v.x; t[1] = v.y; t[2] = v.z; t[3] = v.w;
    __builtin_nontemporal_store(t, reinterpret_cast<f4v*>(p));
}
__device__ __forceinline__ void st_y16(double* p, double2 const& v) {
    typedef double d2v __attribute__((ext_vector_type(2)));
    d2v t; t[0] = v.x; t[1] = v.y;
    __builtin_nontemporal_store(t, reinterpret_cast<d2v*>(p));
}

// word indices in a block header: block, column group, potential cell (-1: none), add-list range [ADD0, ADD1),
// then the neighbour codes [direction][side][k < NH] (>= 0 block, -1 absent, <= -2 periodic image -code - 2)
enum : int { IH_BLK = 0, IH_COL = 1, IH_VEFF = 2, IH_ADD0 = 3, IH_ADD1 = 4, IH_NBR = 5 };
template <int NH>
struct alignas(64) ItemHead {
    uint32_t w[NH == 1 ? 16 : 32];
    __device__ int32_t code(int d, int sd, int k) const { return int32_t(w[IH_NBR + (d * 2 + sd) * NH + k]); }
};

// one add-list entry: atom image, its numax, the Hermite slot of (image, target row), its first coefficient
struct AddEntry {
    uint32_t img, L, slot, coef0;
};

template <typename real_t>
struct LocalArgs {
    real_t* y;
    real_t const* x;
    real_t const* dcoef;
    real_t const* zero;      // one block of zeros
    uint32_t const* items;   // block headers in processing order (cache-aware), ItemHead<NH> each
    AddEntry const* add_entry;   // [add nnz]
    double const* V;
    uint32_t const* add_start;   // [nnzb + 1] add-list range of each block (the add-back kernels after the local one)
    uint16_t const* colindx;
    uint64_t nitems;         // blocks to process (all, or the active ones) listed in items
    uint32_t nCols;
    uint32_t per_xcd;        // work-items per XCD in the XCD-contiguous mapping
    double const* E_col;     // [nCols][2] energy of each column group
    double const* htab;      // Hermite functions per (image, row) slot
    float const* xtab;
    double* dot;             // if set: [nnzb][re, im][64] bilinear products sum_i x[b][i][j] y[b][i][j] per block
    int add_local;           // 1: the local kernel adds the numax <= 5 images; 0: k_sho_add_mfma adds them after it
    real_t c8[3][5];         // 8th-order weights per direction
    real_t c16[3][9];        // 16th-order weights per direction
    int nh[3];               // neighbour blocks per side used in direction d (1 or 2)
    real_t cdiag;            // sum of the centre weights of the three directions
    double phase[3][2][2];
    double hgrid[4];
};

// one z layer iz of the contraction in sho_plane: the M = L - iz triangle of coefficients (iy + ix <= M, ix
// fastest) times H_z[iz](zq), accumulated into c2[iy][ix] (triangular with rows of the full L triangle)
template <typename real_t, int L, int M>
__device__ __forceinline__ void sho_z_layer(real_t (&c2)[((L + 1) * (L + 2)) / 2], real_t const* __restrict__ src,
                                            size_t dstride, real_t hz) {
    int t = 0;
#pragma unroll
    for (int iy = 0; iy <= M; ++iy)
#pragma unroll
        for (int ix = 0; ix <= M - iy; ++ix, ++t) c2[(iy * (2 * L + 3 - iy)) / 2 + ix] += src[size_t(t) * dstride] * hz;
}

// sum over the SHO coefficients c[iz][iy][ix] (ix fastest, iz + iy + ix <= L) of one atom image at the 16
// points (y, x) of the z plane zq, v(y, x) = sum c H_x[ix](x) H_y[iy](y) H_z[iz](zq), handed to out(y, x, v).
// Contracted z first (one value H_z[iz](zq) per thread), then y and x row by row: (L+1)(L+2)(L+3)/6 +
// 4 (L+1)(L+2)/2 + 16 (L+1) FMAs.  The z layers run in a rolled loop, so at most one layer of coefficients is
// loaded ahead; ROLL keeps the row loop rolled as well (out must then accept a run-time row y).
// dsrc points to coefficient 0 of this thread's (plane, column); consecutive coefficients are dstride apart.
template <typename real_t, int L, bool ROLL, typename Out>
__device__ __forceinline__ void sho_plane(real_t const (*H)[3][4], real_t const* __restrict__ dsrc, size_t dstride,
                                          int zq, Out&& out) {
    constexpr int n2 = ((L + 1) * (L + 2)) / 2;
    real_t c2[n2];   // [iy][ix] triangular, row iy starts at iy (2L + 3 - iy) / 2
#pragma unroll
    for (int k = 0; k < n2; ++k) c2[k] = 0;
    real_t const* src = dsrc;
#pragma unroll 1
    for (int iz = 0; iz <= L; ++iz) {
        real_t const hz = H[iz][2][zq];
        switch (L - iz) {
            case 0: sho_z_layer<real_t, L, 0>(c2, src, dstride, hz); break;
            case 1: if constexpr (L >= 1) sho_z_layer<real_t, L, 1>(c2, src, dstride, hz); break;
            case 2: if constexpr (L >= 2) sho_z_layer<real_t, L, 2>(c2, src, dstride, hz); break;
            case 3: if constexpr (L >= 3) sho_z_layer<real_t, L, 3>(c2, src, dstride, hz); break;
            case 4: if constexpr (L >= 4) sho_z_layer<real_t, L, 4>(c2, src, dstride, hz); break;
            case 5: if constexpr (L >= 5) sho_z_layer<real_t, L, 5>(c2, src, dstride, hz); break;
            case 6: if constexpr (L >= 6) sho_z_layer<real_t, L, 6>(c2, src, dstride, hz); break;
            default: if constexpr (L >= 7) sho_z_layer<real_t, L, 7>(c2, src, dstride, hz); break;
        }
        src += size_t(((L - iz + 1) * (L - iz + 2)) / 2) * dstride;
    }
    auto row = [&](int y) {
        real_t by[L + 1];
#pragma unroll
        for (int ix = 0; ix <= L; ++ix) by[ix] = 0;
#pragma unroll
        for (int iy = 0; iy <= L; ++iy) {
            real_t const hy = H[iy][1][y];
#pragma unroll
            for (int ix = 0; ix <= L - iy; ++ix) by[ix] += c2[(iy * (2 * L + 3 - iy)) / 2 + ix] * hy;
        }
#pragma unroll
        for (int xx = 0; xx < 4; ++xx) {
            real_t v = 0;
#pragma unroll
            for (int ix = 0; ix <= L; ++ix) v += by[ix] * H[ix][0][xx];
            out(y, xx, v);
        }
    };
    if constexpr (ROLL) {
#pragma unroll 1
        for (int y = 0; y < 4; ++y) row(y);
    } else {
#pragma unroll
        for (int y = 0; y < 4; ++y) row(y);
    }
}

// sho_plane for a run-time numax L <= LMAX (LMAX <= 7)
template <typename real_t, int LMAX, bool ROLL, typename Out>
__device__ __forceinline__ void sho_plane_any(real_t const (*H)[3][4], real_t const* __restrict__ dsrc, size_t dstride,
                                              int zq, int L, Out&& out) {
    switch (L) {
        case 0: sho_plane<real_t, 0, ROLL>(H, dsrc, dstride, zq, out); break;
        case 1: sho_plane<real_t, 1, ROLL>(H, dsrc, dstride, zq, out); break;
        case 2: sho_plane<real_t, 2, ROLL>(H, dsrc, dstride, zq, out); break;
        case 3: sho_plane<real_t, 3, ROLL>(H, dsrc, dstride, zq, out); break;
        case 4: sho_plane<real_t, 4, ROLL>(H, dsrc, dstride, zq, out); break;
        case 5: sho_plane<real_t, 5, ROLL>(H, dsrc, dstride, zq, out); break;
        case 6: if constexpr (LMAX >= 6) sho_plane<real_t, 6, ROLL>(H, dsrc, dstride, zq, out); break;
        default: if constexpr (LMAX >= 7) sho_plane<real_t, 7, ROLL>(H, dsrc, dstride, zq, out); break;
    }
}


// SHO add-back of one atom image (numax L <= LMAX) into this thread's z plane zq of an accumulator in memory
// (LDS or global) whose points are ostride apart: ys[(zq*16 + y*4 + x) * ostride]
template <typename real_t, int LMAX>
__device__ __forceinline__ void sho_add_atom(real_t* __restrict__ ys, int ostride, real_t const (*H)[3][4],
                                             float const (*xi2)[4], float R2, real_t const* __restrict__ dsrc,
                                             size_t dstride, int zq, int L) {
    sho_plane_any<real_t, LMAX, true>(H, dsrc, dstride, zq, L, [&](int y, int xx, real_t v) {
        real_t& o = ys[(zq * 16 + y * 4 + xx) * ostride];
        o += sho_inside(xi2, R2, zq, y, xx) ? v : real_t(0);
    });
}

// value of a neighbour block at point e (a compile-time element offset from the uniform pointer b) and byte
// offset `off` (this thread's column and plane), with the Bloch phase of a periodic image applied:
// f x_p - (+/-) g x_other; off_o is the byte offset of the other plane.  Uniform base + immediate + 32-bit
// vector offset is the global load's own addressing mode: no per-load address arithmetic.
template <typename real_t, bool IMG>
__device__ __forceinline__ real_t nb_value(real_t const* __restrict__ b, int e, uint32_t off, uint32_t off_o,
                                           real_t fr, real_t fi) {
    char const* bb = reinterpret_cast<char const*>(b + e);
    real_t v = *reinterpret_cast<real_t const*>(bb + off);
    if (IMG) v = fr * v - fi * *reinterpret_cast<real_t const*>(bb + off_o);
    return v;
}

// One work-group per (block, JW-column slice), 8*JW threads.  LDS holds the slice of the block
// (xs) and the output accumulator (ys); passes: potential, kinetic x, y, z, SHO add-back, store.
// In the kinetic passes a thread owns 4 lines (4 points each) of one (column, plane); neighbour
// values come straight from global memory with branch-free addressing (absent -> zero block).
template <typename real_t, int NH, bool IMG, int JW, int LB>
__global__ __launch_bounds__(8 * JW) void k_local(LocalArgs<real_t> a) {
    constexpr int T = 8 * JW;
    constexpr int NW = 4 * (2 * NH + 1);   // points in a line window
    constexpr int NSLICE = 64 / JW;
    __shared__ real_t xs[2 * 64 * JW];
    __shared__ real_t ys[2 * 64 * JW];
    __shared__ real_t H[8][3][4];
    __shared__ float xi2[3][4];

    uint64_t lin = blockIdx.x;
    if (a.per_xcd) lin = uint64_t(blockIdx.x & 7) * a.per_xcd + (blockIdx.x >> 3);
    // slice-major: all blocks of column slice 0 in processing order, then slice 1, ...; the neighbour slices
    // a work-group reads were then read shortly before by work-groups of the same XCD
    uint64_t const item = lin % a.nitems;
    int const slice = int(lin / a.nitems);
    if (slice >= NSLICE) return;  // uniform over the work-group
    ItemHead<NH> const h = reinterpret_cast<ItemHead<NH> const*>(a.items)[item];   // one scalar load
    uint64_t const blk = h.w[IH_BLK];
    int const tid = threadIdx.x;

    // stage this column slice of the block in LDS with 16-byte loads
    using vec_t = typename std::conditional<sizeof(real_t) == 8, double2, float4>::type;
    constexpr int VE = 16 / sizeof(real_t), VPR = JW / VE;  // vectors per row
    {
        real_t const* src = a.x + blk * NBLK + slice * JW;
#pragma unroll
        for (int k = 0; k < 2 * 64 * VPR / T; ++k) {
            int const v = k * T + tid, row = v / VPR, cv = v % VPR;
            *reinterpret_cast<vec_t*>(xs + row * JW + cv * VE) = *reinterpret_cast<vec_t const*>(src + row * 64 + cv * VE);
        }
    }
    __syncthreads();

    // ---- potential: ys = (V - E) x ----
    {
        int32_t const iv = int32_t(h.w[IH_VEFF]);
        real_t const vm = (iv >= 0) ? real_t(1) : real_t(0);
        double const* Vb = a.V + uint64_t(iv >= 0 ? iv : 0) * 64;
        uint32_t const colE = h.w[IH_COL];
        real_t const E_re = real_t(a.E_col[2 * colE]), E_im = real_t(a.E_col[2 * colE + 1]);
#pragma unroll
        for (int k = 0; k < 2 * 64 * JW / T; ++k) {
            int const e = k * T + tid, jl = e % JW, row = e / JW, p = row >> 6, i = row & 63;
            real_t const Vim = E_im * real_t(1 - 2 * p);
            ys[e] = (real_t(Vb[i]) * vm + (a.cdiag - E_re)) * xs[e] + Vim * xs[((1 - p) * 64 + i) * JW + jl];
        }
    }

    // thread's (column, plane, line group) in the kinetic passes
    int const jl = tid % JW, p = (tid / JW) & 1, lg = tid / (2 * JW);   // lg in 0..3
    int const j = slice * JW + jl;
    real_t* ysp = ys + p * 64 * JW + jl;
    real_t const* xsp = xs + p * 64 * JW + jl;
    // byte offsets of this thread's column in its plane and in the other plane (neighbour loads)
    uint32_t const boff_p = uint32_t(p * 4096 + j) * sizeof(real_t), boff_o = uint32_t((1 - p) * 4096 + j) * sizeof(real_t);

#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int const stride = (d == 0) ? 1 : ((d == 1) ? 4 : 16);
        // neighbour blocks of this direction (uniform over the work-group)
        real_t const* nb[2][NH];
        real_t fr[2][NH], fi[2][NH];
#pragma unroll
        for (int sd = 0; sd < 2; ++sd)
#pragma unroll
            for (int k = 0; k < NH; ++k) {
                int32_t const code = h.code(d, sd, k);
                uint64_t const idx = (code >= 0) ? uint64_t(code) : uint64_t(code <= -2 ? -code - 2 : 0);
                nb[sd][k] = (code == -1) ? a.zero : a.x + idx * NBLK;
                bool const img = IMG && (code <= -2);
                fr[sd][k] = img ? real_t(a.phase[d][sd][0]) : real_t(1);
                fi[sd][k] = img ? real_t(a.phase[d][sd][1]) * real_t(1 - 2 * p) : real_t(0);
            }
        __syncthreads();  // the previous pass has finished its updates of ys
#pragma unroll 1
        for (int rb = 0; rb < 4; rb += LB) {
        real_t w[LB][NW];
        int il0[LB];
#pragma unroll
        for (int r = 0; r < LB; ++r) {
            int const line = lg * 4 + rb + r;  // 0..15
            int i0;
            if (d == 0) i0 = line * 4;                               // line = z*4 + y
            else if (d == 1) i0 = (line >> 2) * 16 + (line & 3);     // line = z*4 + x
            else i0 = line;                                          // line = y*4 + x
            il0[r] = i0;
            uint32_t const line_p = boff_p + uint32_t(i0) * (64 * sizeof(real_t));   // first point of the line
            uint32_t const line_o = boff_o + uint32_t(i0) * (64 * sizeof(real_t));
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int const i = i0 + q * stride;
                w[r][4 * NH + q] = xsp[i * JW];
#pragma unroll
                for (int k = 0; k < NH; ++k) {
                    w[r][4 * (NH - 1 - k) + q] = nb_value<real_t, IMG>(nb[0][k], q * stride * 64, line_p, line_o, fr[0][k], fi[0][k]);
                    w[r][4 * (NH + 1 + k) + q] = nb_value<real_t, IMG>(nb[1][k], q * stride * 64, line_p, line_o, fr[1][k], fi[1][k]);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < LB; ++r) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int const c = 4 * NH + q;
                real_t t;
                // off-centre weights only: the centre weights are applied with the potential
                if (NH == 2 && a.nh[d] == 2) {
                    t = a.c16[d][1] * (w[r][c - 1] + w[r][c + 1]);
#pragma unroll
                    for (int m = 2; m <= 8; ++m) {
                        int const lo = c - m, hi = c + m;
                        t += a.c16[d][m] * (w[r][lo < 0 ? 0 : lo] + w[r][hi >= NW ? NW - 1 : hi]);
                    }
                } else {
                    t = a.c8[d][1] * (w[r][c - 1] + w[r][c + 1]);
#pragma unroll
                    for (int m = 2; m <= 4; ++m) t += a.c8[d][m] * (w[r][c - m] + w[r][c + m]);
                }
                ysp[(il0[r] + q * stride) * JW] += t;
            }
        }
        }  // line batches
    }

    // ---- SHO add-back of the atom images (numax <= 5) touching this block; the coefficients of
    //      each image are staged in the LDS buffer of the block, which is free after the kinetic passes ----
    uint32_t const b0 = h.w[IH_ADD0], b1 = h.w[IH_ADD1];
    if (a.add_local && b1 > b0) {
        uint32_t const col = h.w[IH_COL];
        float const R2 = float(a.hgrid[3] * a.hgrid[3]);
        int const zq = lg;  // threads (jl, p, zq) own the z plane zq
        real_t* cs = xs;    // [nc][2][JW]
        for (uint32_t b = b0; b < b1; ++b) {
            AddEntry const e = a.add_entry[b];
            int const L = int(e.L);
            if (L > 5) continue;  // done by k_sho_add_wide
            int const nc = ((L + 1) * (L + 2) * (L + 3)) / 6;
            __syncthreads();
            stage_hermite<real_t>(H, xi2, a.htab, a.xtab, e.slot, L, tid, T);
            {
                real_t const* src = a.dcoef + (size_t(e.coef0) * a.nCols + col) * 128 + slice * JW;
                for (int v = tid; v < 2 * nc * VPR; v += T) {
                    int const row = v / VPR, cv = v % VPR;  // row = coefficient*2 + plane
                    *reinterpret_cast<vec_t*>(cs + row * JW + cv * VE) =
                        *reinterpret_cast<vec_t const*>(src + size_t(row >> 1) * a.nCols * 128 + (row & 1) * 64 + cv * VE);
                }
            }
            __syncthreads();
            sho_add_atom<real_t, 5>(ysp, JW, H, xi2, R2, cs + p * JW + jl, 2 * JW, zq, L);
        }
    }
    __syncthreads();

    // ---- store the slice with 16-byte stores ----
    {
        real_t* dst = a.y + blk * NBLK + slice * JW;
#pragma unroll
        for (int k = 0; k < 2 * 64 * VPR / T; ++k) {
            int const v = k * T + tid, row = v / VPR, cv = v % VPR;
            st_y16(dst + row * 64 + cv * VE, *reinterpret_cast<vec_t const*>(ys + row * JW + cv * VE));
        }
    }
    if (a.dot) {   // (x, y) = sum_i x_i y_i (no conjugation) per column, x re-read from global memory (L2)
        int const jl = tid % JW, grp = tid / JW;   // 8 groups of 8 points
        int const j = slice * JW + jl;
        real_t const* xb = a.x + blk * NBLK + j;
        double re = 0, im = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            int const i = grp * 8 + q;
            double const yr = ys[i * JW + jl], yi = ys[(64 + i) * JW + jl];
            double const xr = xb[i * 64], xi = xb[4096 + i * 64];
            re += xr * yr - xi * yi;
            im += xr * yi + xi * yr;
        }
        __syncthreads();   // xs is free: the partial sums of the 8 groups
        double* red = reinterpret_cast<double*>(xs);
        red[(grp * 2 + 0) * JW + jl] = re;
        red[(grp * 2 + 1) * JW + jl] = im;
        __syncthreads();
        if (tid < JW) {
            double sr = 0, si = 0;
            for (int g = 0; g < 8; ++g) { sr += red[(g * 2) * JW + tid]; si += red[(g * 2 + 1) * JW + tid]; }
            a.dot[(blk * 2 + 0) * 64 + slice * JW + tid] = sr;
            a.dot[(blk * 2 + 1) * 64 + slice * JW + tid] = si;
        }
    }
}


// SHO add-back of one atom image (numax <= 5) into 16 register accumulators (z plane zq of one column and plane)
template <typename real_t>
__device__ __forceinline__ void sho_add_regs(real_t (&out)[16], real_t const (*H)[3][4], float const (*xi2)[4],
                                             float R2, real_t const* __restrict__ dsrc, size_t dstride, int zq, int L) {
    sho_plane_any<real_t, 5, false>(H, dsrc, dstride, zq, L, [&](int y, int xx, real_t v) {
        out[y * 4 + xx] += sho_inside(xi2, R2, zq, y, xx) ? v : real_t(0);
    });
}

// acc += the off-centre finite-difference weights applied to a window w of 4 (2 NH + 1) values centred at c
// (the centre weights of the three directions are applied together with the potential)
template <typename real_t, int NH, int NW>
__device__ __forceinline__ void stencil(real_t& acc, LocalArgs<real_t> const& a, int d, real_t const (&w)[NW], int c) {
    if (NH == 2 && a.nh[d] == 2) {
#pragma unroll
        for (int m = 1; m <= 8; ++m) {
            int const lo = c - m, hi = c + m;
            acc += a.c16[d][m] * (w[lo < 0 ? 0 : lo] + w[hi >= NW ? NW - 1 : hi]);
        }
    } else {
#pragma unroll
        for (int m = 1; m <= 4; ++m) acc += a.c8[d][m] * (w[c - m] + w[c + m]);
    }
}

// z-direction stencil for the thread's plane ZQ: window positions z' = ZQ - 4 NH .. ZQ + 4 NH along z at every
// (y, x); inside the block from LDS (own plane from registers), outside from the neighbour blocks
template <typename real_t, int NH, bool IMG, int JW, int ZQ>
__device__ __forceinline__ void z_pass(LocalArgs<real_t> const& a, real_t (&acc)[16], real_t const (&own)[16],
                                       real_t const* __restrict__ xsp, real_t const* const (&nb)[2][NH],
                                       real_t const (&fr)[2][NH], real_t const (&fi)[2][NH], uint32_t boff_p, uint32_t boff_o) {
    constexpr int NWZ = 8 * NH + 1;   // z' = ZQ - 4 NH .. ZQ + 4 NH
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        real_t w[NWZ];
#pragma unroll
        for (int m = 0; m < NWZ; ++m) {
            int const z = ZQ - 4 * NH + m;   // z' relative to this block
            if (z == ZQ) w[m] = own[k];
            else if (z >= 0 && z < 4) w[m] = xsp[(z * 16 + k) * JW];
            else if (z < 0) {
                int const kb = (-z - 1) / 4, zz = z + 4 * (kb + 1);
                w[m] = nb_value<real_t, IMG>(nb[0][kb], (zz * 16 + k) * 64, boff_p, boff_o, fr[0][kb], fi[0][kb]);
            } else {
                int const kb = (z - 4) / 4, zz = z - 4 * (kb + 1);
                w[m] = nb_value<real_t, IMG>(nb[1][kb], (zz * 16 + k) * 64, boff_p, boff_o, fr[1][kb], fi[1][kb]);
            }
        }
        stencil<real_t, NH, NWZ>(acc[k], a, 2, w, 4 * NH);
    }
}

// neighbour blocks of blk in direction d, NH per side: base pointers (an absent neighbour reads the zero block)
// and the Bloch phase of a periodic image as (fr, fi) for plane p (fr = 1, fi = 0 otherwise)
template <typename real_t, int NH, bool IMG>
__device__ __forceinline__ void neighbours(LocalArgs<real_t> const& a, int d, ItemHead<NH> const& h, int p,
                                           real_t const* (&nb)[2][NH], real_t (&fr)[2][NH], real_t (&fi)[2][NH]) {
#pragma unroll
    for (int sd = 0; sd < 2; ++sd)
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            int32_t const code = h.code(d, sd, k);
            uint64_t const idx = (code >= 0) ? uint64_t(code) : uint64_t(code <= -2 ? -code - 2 : 0);
            nb[sd][k] = (code == -1) ? a.zero : a.x + idx * NBLK;
            bool const img = IMG && (code <= -2);
            fr[sd][k] = img ? real_t(a.phase[d][sd][0]) : real_t(1);
            fi[sd][k] = img ? real_t(a.phase[d][sd][1]) * real_t(1 - 2 * p) : real_t(0);
        }
}

// (V - E) x at the thread's 16 points (z plane zq of one column, plane p): own = x of plane p there, xso = x of the
// other plane in LDS (row stride JW), vz = the potential at the 16 points (0 where the block has no potential
// cell), E the column group's energy, whose imaginary part couples the two planes; the centre weights of the
// three finite-difference directions are applied here as well
template <typename real_t, int JW>
__device__ __forceinline__ void plane_potential(LocalArgs<real_t> const& a, double const (&vz)[16], double E_re,
                                                double E_im, int p, int zq, real_t const (&own)[16],
                                                real_t const* xso, real_t (&acc)[16]) {
    real_t const Vim = real_t(E_im) * real_t(1 - 2 * p);
    real_t const dg = a.cdiag - real_t(E_re);
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = (real_t(vz[k]) + dg) * own[k] + Vim * xso[(zq * 16 + k) * JW];
}

// kinetic x and y at the thread's 16 points: the lines of plane zq along x (the rows y) and along y (the columns
// x) through this block and NH neighbour blocks on each side, the neighbours straight from global memory.  zq is
// uniform within the wave, so the plane's offset goes into the (scalar) base pointers.
template <typename real_t, int NH, bool IMG>
__device__ __forceinline__ void plane_xy(LocalArgs<real_t> const& a, ItemHead<NH> const& h, int p, int zq,
                                         uint32_t boff_p, uint32_t boff_o, real_t const (&own)[16], real_t (&acc)[16]) {
    constexpr int NW = 4 * (2 * NH + 1);
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        real_t const* nb[2][NH];
        real_t fr[2][NH], fi[2][NH];
        neighbours<real_t, NH, IMG>(a, d, h, p, nb, fr, fi);
        int const zqu = __builtin_amdgcn_readfirstlane(zq);
        real_t const* nbz[2][NH];
#pragma unroll
        for (int sd = 0; sd < 2; ++sd)
#pragma unroll
            for (int kk = 0; kk < NH; ++kk) nbz[sd][kk] = nb[sd][kk] + zqu * 16 * 64;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            real_t w[NW];   // the line through this block and NH neighbour blocks on each side
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int const k = (d == 0) ? (r * 4 + q) : (q * 4 + r);   // point of the plane on the line
                int const e = k * 64;   // plane zq is in the base pointers
#pragma unroll
                for (int kk = 0; kk < NH; ++kk) {
                    w[4 * (NH - 1 - kk) + q] = nb_value<real_t, IMG>(nbz[0][kk], e, boff_p, boff_o, fr[0][kk], fi[0][kk]);
                    w[4 * (NH + 1 + kk) + q] = nb_value<real_t, IMG>(nbz[1][kk], e, boff_p, boff_o, fr[1][kk], fi[1][kk]);
                }
                w[4 * NH + q] = own[k];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int const k = (d == 0) ? (r * 4 + q) : (q * 4 + r);
                stencil<real_t, NH, NW>(acc[k], a, d, w, 4 * NH + q);
            }
        }
    }
}

// The end of a register-plane kernel for one block slice: the SHO add-back of the atom images (numax <= 5)
// touching the block (add-list entries b0 .. b1 - 1, column group col), the store of the thread's 16 points and,
// if a.dot is set, the bilinear product (x, y) of every column of the slice.  cs is LDS of at least 2 * 56 * JW
// elements of its own (the coefficients of one image at a time, then the partial sums of the product); xsp and
// xso point to this thread's column of x in the LDS copy of the slice (its plane and the other one), which stays
// intact, so that x is not held in registers across the add-back.  Contains work-group barriers: every thread of
// the work-group calls it.
template <typename real_t, int JW>
__device__ __forceinline__ void plane_finish(LocalArgs<real_t> const& a, uint64_t blk, uint32_t col, uint32_t b0,
                                             uint32_t b1, int slice, int p, int zq, int jl, real_t (&acc)[16],
                                             real_t const* xsp, real_t const* xso, real_t* cs,
                                             real_t (*H)[3][4], float (*xi2)[4]) {
    constexpr int T = 8 * JW;
    using vec_t = typename std::conditional<sizeof(real_t) == 8, double2, float4>::type;
    constexpr int VE = 16 / sizeof(real_t), VPR = JW / VE;
    int const tid = threadIdx.x, j = slice * JW + jl;
    if (a.add_local && b1 > b0) {
        float const R2 = float(a.hgrid[3] * a.hgrid[3]);
        for (uint32_t b = b0; b < b1; ++b) {
            AddEntry const e = a.add_entry[b];
            int const L = int(e.L);
            if (L > 5) continue;  // done by k_sho_add_wide
            int const nc = ((L + 1) * (L + 2) * (L + 3)) / 6;
            __syncthreads();
            stage_hermite<real_t>(H, xi2, a.htab, a.xtab, e.slot, L, tid, T);
            {   // [coefficient][plane][JW]
                real_t const* src = a.dcoef + (size_t(e.coef0) * a.nCols + col) * 128 + slice * JW;
                for (int v = tid; v < 2 * nc * VPR; v += T) {
                    int const row = v / VPR, cv = v % VPR;  // row = coefficient*2 + plane
                    *reinterpret_cast<vec_t*>(cs + row * JW + cv * VE) =
                        *reinterpret_cast<vec_t const*>(src + size_t(row >> 1) * a.nCols * 128 + (row & 1) * 64 + cv * VE);
                }
            }
            __syncthreads();
            sho_add_regs<real_t>(acc, H, xi2, R2, cs + p * JW + jl, 2 * JW, zq, L);
        }
    }

    real_t* dst = a.y + blk * NBLK + size_t(p * 4096 + j) + size_t(zq * 16) * 64;
#pragma unroll
    for (int k = 0; k < 16; ++k) __builtin_nontemporal_store(acc[k], &dst[size_t(k) * 64]);   // y is not re-read here
    if (a.dot) {
        // (x, y) = sum_i x_i y_i per column: plane 0 holds x_r, y_r (+ x_i from LDS), plane 1 x_i, y_i (+ x_r):
        // Re = sum x_r y_r - sum x_i y_i, Im = sum x_i y_r + sum x_r y_i
        real_t re_t = 0, im_t = 0;   // 16 products in the vector precision, then summed in FP64
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            re_t += xsp[(zq * 16 + k) * JW] * acc[k];
            im_t += xso[(zq * 16 + k) * JW] * acc[k];
        }
        double re = re_t, im = im_t;
        if (p == 1) re = -re;
        __syncthreads();   // cs is free
        double* red = reinterpret_cast<double*>(cs);   // [zq][p][re, im][JW]
        red[((zq * 2 + p) * 2 + 0) * JW + jl] = re;
        red[((zq * 2 + p) * 2 + 1) * JW + jl] = im;
        __syncthreads();
        if (tid < JW) {
            double sr = 0, si = 0;
            for (int g = 0; g < 8; ++g) { sr += red[(g * 2 + 0) * JW + tid]; si += red[(g * 2 + 1) * JW + tid]; }
            a.dot[(blk * 2 + 0) * 64 + slice * JW + tid] = sr;
            a.dot[(blk * 2 + 1) * 64 + slice * JW + tid] = si;
        }
    }
}

// Register-plane variant of the local kernel: one work-group per (block, JW-column slice), 8 JW threads; thread
// (jl, p, zq) owns the 16 points of z plane zq of column j = slice JW + jl, plane p, and accumulates potential,
// kinetic x, y, z and the SHO add-back in registers.  LDS holds the block slice (read by the z-stencil, the
// imaginary part of the potential and the bilinear product of the epilogue) and, apart, the coefficients of the
// add-back, so that x need not stay in registers across the add-back (95 VGPRs instead of 124 in FP32).  zq is
// uniform within a wave (2 JW >= 64), so the z-stencil is specialised per plane.
template <typename real_t, int NH, bool IMG, int JW>
__global__ __launch_bounds__(8 * JW) void k_plane(LocalArgs<real_t> a) {
    static_assert(2 * JW >= 64, "zq must be uniform within a wave");
    constexpr int T = 8 * JW;
    constexpr int NSLICE = 64 / JW;
    __shared__ real_t xs[2 * 64 * JW];
    __shared__ __align__(16) real_t cs[2 * 56 * JW];   // add-back coefficients (numax <= 5), then the (x, y) sums
    __shared__ real_t H[8][3][4];
    __shared__ float xi2[3][4];

    uint64_t lin = blockIdx.x;
    if (a.per_xcd) lin = uint64_t(blockIdx.x & 7) * a.per_xcd + (blockIdx.x >> 3);
    // slice-major: all blocks of column slice 0 in processing order, then slice 1, ...; the neighbour slices
    // a work-group reads were then read shortly before by work-groups of the same XCD
    uint64_t const item = lin % a.nitems;
    int const slice = int(lin / a.nitems);
    if (slice >= NSLICE) return;  // uniform over the work-group
    ItemHead<NH> const h = reinterpret_cast<ItemHead<NH> const*>(a.items)[item];   // one scalar load
    uint64_t const blk = h.w[IH_BLK];
    int const tid = threadIdx.x;

    using vec_t = typename std::conditional<sizeof(real_t) == 8, double2, float4>::type;
    constexpr int VE = 16 / sizeof(real_t), VPR = JW / VE;
    {
        real_t const* src = a.x + blk * NBLK + slice * JW;
#pragma unroll
        for (int k = 0; k < 2 * 64 * VPR / T; ++k) {
            int const v = k * T + tid, row = v / VPR, cv = v % VPR;
            *reinterpret_cast<vec_t*>(xs + row * JW + cv * VE) = *reinterpret_cast<vec_t const*>(src + row * 64 + cv * VE);
        }
    }
    int const jl = tid % JW, p = (tid / JW) & 1, zq = tid / (2 * JW);
    // the potential at the thread's plane and the energy, loaded (scalar: zq is uniform within a wave) together
    // with the slice
    double vz[16];
    {
        int32_t const iv = int32_t(h.w[IH_VEFF]);
        double const* Vb = a.V + uint64_t(iv >= 0 ? iv : 0) * 64 + __builtin_amdgcn_readfirstlane(zq) * 16;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            double const v = Vb[k];   // in bounds also without a cell (row 0)
            vz[k] = (iv >= 0) ? v : 0.0;
        }
    }
    double const E_re = a.E_col[2 * h.w[IH_COL]], E_im = a.E_col[2 * h.w[IH_COL] + 1];
    __syncthreads();

    int const j = slice * JW + jl;
    real_t const* xsp = xs + p * 64 * JW + jl;
    real_t const* xso = xs + (1 - p) * 64 * JW + jl;
    // byte offsets of this thread's column in its plane and in the other plane (neighbour loads)
    uint32_t const boff_p = uint32_t(p * 4096 + j) * sizeof(real_t), boff_o = uint32_t((1 - p) * 4096 + j) * sizeof(real_t);

    real_t own[16], acc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) own[k] = xsp[(zq * 16 + k) * JW];

    plane_potential<real_t, JW>(a, vz, E_re, E_im, p, zq, own, xso, acc);
    plane_xy<real_t, NH, IMG>(a, h, p, zq, boff_p, boff_o, own, acc);
    {
        real_t const* nb[2][NH];
        real_t fr[2][NH], fi[2][NH];
        neighbours<real_t, NH, IMG>(a, 2, h, p, nb, fr, fi);
        switch (zq) {
            case 0: z_pass<real_t, NH, IMG, JW, 0>(a, acc, own, xsp, nb, fr, fi, boff_p, boff_o); break;
            case 1: z_pass<real_t, NH, IMG, JW, 1>(a, acc, own, xsp, nb, fr, fi, boff_p, boff_o); break;
            case 2: z_pass<real_t, NH, IMG, JW, 2>(a, acc, own, xsp, nb, fr, fi, boff_p, boff_o); break;
            default: z_pass<real_t, NH, IMG, JW, 3>(a, acc, own, xsp, nb, fr, fi, boff_p, boff_o); break;
        }
    }

    plane_finish<real_t, JW>(a, blk, h.w[IH_COL], h.w[IH_ADD0], h.w[IH_ADD1], slice, p, zq, jl, acc, xsp, xso, cs,
                             H, xi2);
}

// ---------------------------------------------------------------------------------------------
// Line variant of the local kernel (kind 2).  One work-group per (block, JW-column slice).  A lane owns
// CPL = 16 / sizeof(real_t) adjacent columns of one plane, so that every global load and store moves 16 bytes per
// lane.  Thread (cq, p, l) owns line l = 0..15 of plane p in each direction: the x-line (y, z) = (l % 4, l / 4),
// the y-line (x, z) and the z-line (x, y) of the same two indices.  Every neighbour value the three stencils need
// is loaded once per work-group, straight into registers, and all of them are issued before the first stencil;
// the block's own y- and z-lines come from the LDS copy of the slice.  The y and z stencil sums meet the x-line
// accumulator (potential, x stencil, SHO add-back) in a second LDS buffer.
// ---------------------------------------------------------------------------------------------
template <typename real_t>
struct alignas(16) Lane {   // the 16 bytes of adjacent columns a lane owns at one point
    static constexpr int n = 16 / sizeof(real_t);
    real_t v[n];
};

__device__ __forceinline__ Lane<float> ld16(float const* p) {
    float4 const t = *reinterpret_cast<float4 const*>(p);
    return Lane<float>{{t.x, t.y, t.z, t.w}};
}
__device__ __forceinline__ Lane<double> ld16(double const* p) {
    double2 const t = *reinterpret_cast<double2 const*>(p);
    return Lane<double>{{t.x, t.y}};
}
__device__ __forceinline__ void st16(float* p, Lane<float> const& v) {
    *reinterpret_cast<float4*>(p) = float4{v.v[0], v.v[1], v.v[2], v.v[3]};
}
__device__ __forceinline__ void st16(double* p, Lane<double> const& v) {
    *reinterpret_cast<double2*>(p) = double2{v.v[0], v.v[1]};
}
template <typename real_t>
__device__ __forceinline__ void st16_nt(real_t* p, Lane<real_t> const& v) {   // non-temporal, for y
    if constexpr (sizeof(real_t) == 4) st_y16(p, float4{v.v[0], v.v[1], v.v[2], v.v[3]});
    else st_y16(p, double2{v.v[0], v.v[1]});
}
// 16 bytes at element e of the uniform base b plus the lane's byte offset off (the load's own addressing mode)
template <typename real_t>
__device__ __forceinline__ Lane<real_t> ldv(real_t const* __restrict__ b, int e, uint32_t off) {
    return ld16(reinterpret_cast<real_t const*>(reinterpret_cast<char const*>(b + e) + off));
}

// acc[q] += the off-centre stencil of direction d at point q of a line, for every column of the lane: the window
// is the own line (centre) and NH neighbour lines on each side (w[side][k][q], k = 0 nearest)
template <typename real_t, int NH>
__device__ __forceinline__ void line_stencil(Lane<real_t> (&acc)[4], LocalArgs<real_t> const& a, int d,
                                             Lane<real_t> const (&own)[4], Lane<real_t> const (&w)[2][NH][4]) {
    constexpr int NW = 4 * (2 * NH + 1);
#pragma unroll
    for (int c = 0; c < Lane<real_t>::n; ++c) {
        real_t win[NW];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            win[4 * NH + q] = own[q].v[c];
#pragma unroll
            for (int k = 0; k < NH; ++k) {
                win[4 * (NH - 1 - k) + q] = w[0][k][q].v[c];
                win[4 * (NH + 1 + k) + q] = w[1][k][q].v[c];
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) stencil<real_t, NH, NW>(acc[q].v[c], a, d, win, 4 * NH + q);
    }
}

// one z layer of sho_xline: the M = L - iz triangle of coefficients (iy + ix <= M, ix fastest) times
// H_z[iz](z) H_y[iy](y), accumulated into t[column][ix]
template <typename real_t, int L, int M>
__device__ __forceinline__ void xline_layer(real_t (&t)[Lane<real_t>::n][L + 1], real_t const* cs, int cstride,
                                            real_t hz, real_t const (*H)[3][4], int y) {
    int s = 0;
#pragma unroll
    for (int iy = 0; iy <= M; ++iy) {
        real_t const hzy = hz * H[iy][1][y];
#pragma unroll
        for (int ix = 0; ix <= M - iy; ++ix, ++s) {
            Lane<real_t> const cv = ld16(cs + s * cstride);
#pragma unroll
            for (int c = 0; c < Lane<real_t>::n; ++c) t[c][ix] += cv.v[c] * hzy;
        }
    }
}

// SHO add-back of one atom image (numax L) into the x-line (y, z) of a lane: the coefficients c[iz][iy][ix]
// (ix fastest) are contracted with H_z(z) H_y(y) first, t[ix], then t with H_x at the 4 points of the line.
// The z layers run in a rolled loop, so that one layer of coefficients at a time is loaded ahead.
// cs points to coefficient 0 of the lane's plane and first column; consecutive coefficients are cstride apart.
template <typename real_t, int L>
__device__ __forceinline__ void sho_xline(Lane<real_t> (&acc)[4], real_t const (*H)[3][4], float const (*xi2)[4],
                                          float R2, real_t const* cs, int cstride, int y, int z) {
    constexpr int CPL = Lane<real_t>::n;
    real_t t[CPL][L + 1];
#pragma unroll
    for (int c = 0; c < CPL; ++c)
#pragma unroll
        for (int ix = 0; ix <= L; ++ix) t[c][ix] = 0;
#pragma unroll 1
    for (int iz = 0; iz <= L; ++iz) {
        real_t const hz = H[iz][2][z];
        switch (L - iz) {
            case 0: xline_layer<real_t, L, 0>(t, cs, cstride, hz, H, y); break;
            case 1: if constexpr (L >= 1) xline_layer<real_t, L, 1>(t, cs, cstride, hz, H, y); break;
            case 2: if constexpr (L >= 2) xline_layer<real_t, L, 2>(t, cs, cstride, hz, H, y); break;
            case 3: if constexpr (L >= 3) xline_layer<real_t, L, 3>(t, cs, cstride, hz, H, y); break;
            case 4: if constexpr (L >= 4) xline_layer<real_t, L, 4>(t, cs, cstride, hz, H, y); break;
            default: if constexpr (L >= 5) xline_layer<real_t, L, 5>(t, cs, cstride, hz, H, y); break;
        }
        cs += ((L - iz + 1) * (L - iz + 2)) / 2 * cstride;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        bool const in = sho_inside(xi2, R2, z, y, q);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            real_t v = 0;
#pragma unroll
            for (int ix = 0; ix <= L; ++ix) v += t[c][ix] * H[ix][0][q];
            acc[q].v[c] += in ? v : real_t(0);
        }
    }
}

// the neighbour lines of direction d (NH blocks per side) of the lane's line starting at point i0, plane p and
// first column j0, into w; a periodic-image neighbour (uniform over the work-group) gets its Bloch phase,
// f x_p - (+/-) g x_other, which waits for its loads (image blocks are few)
template <typename real_t, int NH, bool IMG>
__device__ __forceinline__ void load_lines(LocalArgs<real_t> const& a, int d, ItemHead<NH> const& h, int p, int i0,
                                           int j0, Lane<real_t> (&w)[2][NH][4]) {
    int const stride = (d == 0) ? 1 : ((d == 1) ? 4 : 16);
    uint32_t const off = uint32_t(p * 4096 + i0 * 64 + j0) * sizeof(real_t);
#pragma unroll
    for (int sd = 0; sd < 2; ++sd)
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            int32_t const code = h.code(d, sd, k);
            uint64_t const idx = (code >= 0) ? uint64_t(code) : uint64_t(code <= -2 ? -code - 2 : 0);
            real_t const* nb = (code == -1) ? a.zero : a.x + idx * NBLK;
#pragma unroll
            for (int q = 0; q < 4; ++q) w[sd][k][q] = ldv(nb, q * stride * 64, off);
        }
    if (IMG) {
        uint32_t const offo = uint32_t((1 - p) * 4096 + i0 * 64 + j0) * sizeof(real_t);
#pragma unroll
        for (int sd = 0; sd < 2; ++sd)
#pragma unroll
            for (int k = 0; k < NH; ++k) {
                int32_t const code = h.code(d, sd, k);
                if (code <= -2) {
                    real_t const* nb = a.x + uint64_t(-code - 2) * NBLK;
                    real_t const fr = real_t(a.phase[d][sd][0]), fi = real_t(a.phase[d][sd][1]) * real_t(1 - 2 * p);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        Lane<real_t> const o = ldv(nb, q * stride * 64, offo);
#pragma unroll
                        for (int c = 0; c < Lane<real_t>::n; ++c) w[sd][k][q].v[c] = fr * w[sd][k][q].v[c] - fi * o.v[c];
                    }
                }
            }
    }
}

// PF: directions whose neighbour lines are issued before the first barrier (3: all; 2: x and y, z after the x
// stencil; 1: x, then y after the x stencil and z after the y stencil) -- loads in flight against occupancy.
// The FP64 shape with 16 columns and PF 2 needs 130 VGPRs on its own; it is held to 128 (4 waves per SIMD instead
// of 3, one 8-byte spill), which makes it 6% faster; with PF 3 the same bound spills heavily and is not applied.
template <typename real_t, int NH, bool IMG, int JW, int PF>
__global__ __launch_bounds__((JW / Lane<real_t>::n) * 32, (sizeof(real_t) == 8 && JW == 16 && PF == 2) ? 4 : 1)
void k_lines(LocalArgs<real_t> a) {
    using L16 = Lane<real_t>;
    constexpr int CPL = L16::n, LQ = JW / CPL, T = LQ * 32;
    constexpr int NSLICE = 64 / JW;
    __shared__ __align__(16) real_t xs[2 * 64 * JW];   // the slice [plane][point][column]
    __shared__ __align__(16) real_t ya[2 * 64 * JW];   // y + z stencil sums; then the coefficients of an image
    __shared__ real_t H[8][3][4];
    __shared__ float xi2[3][4];

    uint64_t lin = blockIdx.x;
    if (a.per_xcd) lin = uint64_t(blockIdx.x & 7) * a.per_xcd + (blockIdx.x >> 3);
    uint64_t const item = lin % a.nitems;   // slice-major, as in k_plane
    int const slice = int(lin / a.nitems);
    if (slice >= NSLICE) return;  // uniform over the work-group
    ItemHead<NH> const h = reinterpret_cast<ItemHead<NH> const*>(a.items)[item];   // one scalar load
    uint64_t const blk = h.w[IH_BLK];
    int const tid = threadIdx.x, cq = tid % LQ, p = (tid / LQ) & 1, l = tid / (2 * LQ);
    int const la = l & 3, lb = l >> 2;
    int const jc = cq * CPL, j0 = slice * JW + jc;
    // first points of the lines: x-line (y, z) = (la, lb), y-line (x, z) = (la, lb), z-line (x, y) = (la, lb)
    int const ix0 = lb * 16 + la * 4, iy0 = lb * 16 + la, iz0 = lb * 4 + la;

    L16 wx[2][NH][4], wy[2][NH][4], wz[2][NH][4];   // neighbour lines per direction
    load_lines<real_t, NH, IMG>(a, 0, h, p, ix0, j0, wx);
    if constexpr (PF >= 2) load_lines<real_t, NH, IMG>(a, 1, h, p, iy0, j0, wy);
    if constexpr (PF >= 3) load_lines<real_t, NH, IMG>(a, 2, h, p, iz0, j0, wz);
    // own x-line, to registers and to the LDS copy of the slice
    L16 own[4];
    {
        real_t const* src = a.x + blk * NBLK + p * 4096 + j0;
#pragma unroll
        for (int q = 0; q < 4; ++q) own[q] = ld16(src + (ix0 + q) * 64);
#pragma unroll
        for (int q = 0; q < 4; ++q) st16(xs + (p * 64 + ix0 + q) * JW + jc, own[q]);
    }
    __syncthreads();

    // ---- potential and x stencil on the x-line ----
    L16 acc[4];
    {
        L16 oth[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) oth[q] = ld16(xs + ((1 - p) * 64 + ix0 + q) * JW + jc);
        int32_t const iv = int32_t(h.w[IH_VEFF]);
        real_t const vm = (iv >= 0) ? real_t(1) : real_t(0);
        double const* Vb = a.V + uint64_t(iv >= 0 ? iv : 0) * 64 + ix0;
        uint32_t const colE = h.w[IH_COL];
        real_t const E_re = real_t(a.E_col[2 * colE]), E_im = real_t(a.E_col[2 * colE + 1]);
        real_t const Vim = E_im * real_t(1 - 2 * p);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            real_t const dg = real_t(Vb[q]) * vm + (a.cdiag - E_re);
#pragma unroll
            for (int c = 0; c < CPL; ++c) acc[q].v[c] = dg * own[q].v[c] + Vim * oth[q].v[c];
        }
        line_stencil<real_t, NH>(acc, a, 0, own, wx);
    }
    if constexpr (PF < 2) load_lines<real_t, NH, IMG>(a, 1, h, p, iy0, j0, wy);
    if constexpr (PF == 2) load_lines<real_t, NH, IMG>(a, 2, h, p, iz0, j0, wz);

    // ---- y stencil on the y-line, into ya ----
    {
        L16 o[4], s[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            o[q] = ld16(xs + (p * 64 + iy0 + 4 * q) * JW + jc);
#pragma unroll
            for (int c = 0; c < CPL; ++c) s[q].v[c] = 0;
        }
        line_stencil<real_t, NH>(s, a, 1, o, wy);
#pragma unroll
        for (int q = 0; q < 4; ++q) st16(ya + (p * 64 + iy0 + 4 * q) * JW + jc, s[q]);
    }
    if constexpr (PF < 2) load_lines<real_t, NH, IMG>(a, 2, h, p, iz0, j0, wz);
    __syncthreads();
    // ---- z stencil on the z-line, added into ya ----
    {
        L16 o[4], s[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            o[q] = ld16(xs + (p * 64 + iz0 + 16 * q) * JW + jc);
            s[q] = ld16(ya + (p * 64 + iz0 + 16 * q) * JW + jc);
        }
        line_stencil<real_t, NH>(s, a, 2, o, wz);
#pragma unroll
        for (int q = 0; q < 4; ++q) st16(ya + (p * 64 + iz0 + 16 * q) * JW + jc, s[q]);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        L16 const s = ld16(ya + (p * 64 + ix0 + q) * JW + jc);
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[q].v[c] += s.v[c];
    }

    // ---- SHO add-back of the atom images (numax <= 5) touching the block; coefficients staged in ya ----
    uint32_t const b0 = h.w[IH_ADD0], b1 = h.w[IH_ADD1];
    if (a.add_local && b1 > b0) {
        uint32_t const col = h.w[IH_COL];
        float const R2 = float(a.hgrid[3] * a.hgrid[3]);
        constexpr int VPR = JW / CPL;
        for (uint32_t b = b0; b < b1; ++b) {
            AddEntry const e = a.add_entry[b];
            int const L = int(e.L);
            if (L > 5) continue;  // done by k_sho_add_wide
            int const nc = ((L + 1) * (L + 2) * (L + 3)) / 6;
            __syncthreads();   // ya (the stencil sums, or the previous image) is no longer read
            stage_hermite<real_t>(H, xi2, a.htab, a.xtab, e.slot, L, tid, T);
            {   // [coefficient][plane][JW]
                real_t const* src = a.dcoef + (size_t(e.coef0) * a.nCols + col) * 128 + slice * JW;
                for (int v = tid; v < 2 * nc * VPR; v += T) {
                    int const row = v / VPR, cv = v % VPR;  // row = coefficient*2 + plane
                    st16(ya + row * JW + cv * CPL, ld16(src + size_t(row >> 1) * a.nCols * 128 + (row & 1) * 64 + cv * CPL));
                }
            }
            __syncthreads();
            real_t const* cs = ya + p * JW + jc;
            switch (L) {
                case 0: sho_xline<real_t, 0>(acc, H, xi2, R2, cs, 2 * JW, la, lb); break;
                case 1: sho_xline<real_t, 1>(acc, H, xi2, R2, cs, 2 * JW, la, lb); break;
                case 2: sho_xline<real_t, 2>(acc, H, xi2, R2, cs, 2 * JW, la, lb); break;
                case 3: sho_xline<real_t, 3>(acc, H, xi2, R2, cs, 2 * JW, la, lb); break;
                case 4: sho_xline<real_t, 4>(acc, H, xi2, R2, cs, 2 * JW, la, lb); break;
                default: sho_xline<real_t, 5>(acc, H, xi2, R2, cs, 2 * JW, la, lb); break;
            }
        }
    }

    // ---- store the x-line ----
    {
        real_t* dst = a.y + blk * NBLK + p * 4096 + j0;
#pragma unroll
        for (int q = 0; q < 4; ++q) st16_nt(dst + (ix0 + q) * 64, acc[q]);
    }
    if (a.dot) {
        // (x, y) = sum_i x_i y_i per column: plane 0 holds x_r, y_r (+ x_i), plane 1 x_i, y_i (+ x_r):
        // Re = sum x_r y_r - sum x_i y_i, Im = sum x_i y_r + sum x_r y_i; 4 products in the vector precision,
        // then FP64 over the 32 lines and planes of the work-group in a fixed order
        real_t re_t[CPL], im_t[CPL];   // x of both planes at the x-line from the LDS copy of the slice
#pragma unroll
        for (int c = 0; c < CPL; ++c) re_t[c] = im_t[c] = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            L16 const xp = ld16(xs + (p * 64 + ix0 + q) * JW + jc), xo = ld16(xs + ((1 - p) * 64 + ix0 + q) * JW + jc);
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                re_t[c] += xp.v[c] * acc[q].v[c];
                im_t[c] += xo.v[c] * acc[q].v[c];
            }
        }
        // the lanes of one column group within a wave (lane bits LQ and up: plane and line) are summed with
        // shuffles, then the waves' partial sums in LDS ([wave][re, im][JW], fits xs) in wave order
        constexpr int NWAVE = T / 64;
        double re[CPL], im[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            re[c] = (p == 1) ? -double(re_t[c]) : double(re_t[c]);
            im[c] = double(im_t[c]);
#pragma unroll
            for (int o = LQ; o < 64; o <<= 1) {
                re[c] += __shfl_xor(re[c], o, 64);
                im[c] += __shfl_xor(im[c], o, 64);
            }
        }
        __syncthreads();   // xs is free
        double* red = reinterpret_cast<double*>(xs);
        int const wave = tid >> 6, lane = tid & 63;
        if (lane < LQ) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                red[(wave * 2 + 0) * JW + jc + c] = re[c];
                red[(wave * 2 + 1) * JW + jc + c] = im[c];
            }
        }
        __syncthreads();
        if (tid < JW) {
            double sr = 0, si = 0;
            for (int w = 0; w < NWAVE; ++w) { sr += red[(w * 2 + 0) * JW + tid]; si += red[(w * 2 + 1) * JW + tid]; }
            a.dot[(blk * 2 + 0) * 64 + slice * JW + tid] = sr;
            a.dot[(blk * 2 + 1) * 64 + slice * JW + tid] = si;
        }
    }
}

// SHO add-back on the matrix cores (the dense form), after the local kernel: for every block touched by atom images
// with numax <= 5, y[b][n][i] += sum_img sum_s P_img[i][s] D_img[s][n] (i the 64 points, n = plane * 64 + column,
// s the coefficients padded to 64).  One work-group per block; 4 waves, wave w owns the columns n = 32 w .. 32 w + 31
// and all four 16-point tiles.  P (the A operand, zero outside the sphere and beyond the coefficients) is built in
// LDS per image; the coefficients D (the B operand) are read from global memory.  Operand layout as in
// k_sho_project_mfma.
template <typename real_t>
__global__ __launch_bounds__(256) void k_sho_add_mfma(LocalArgs<real_t> a, uint32_t const* __restrict__ blocks,
                                                      uint8_t const* __restrict__ col_active) {
    using acc_t = typename Mfma<real_t>::acc_t;
    constexpr int NCP = 64;
    __shared__ real_t P[64 * NCP];   // [point][coefficient]
    __shared__ real_t H[6][3][4];
    __shared__ float xi2[3][4];
    __shared__ uint16_t sxyz[NCP];   // coefficient s -> ix | iy << 4 | iz << 8 of the current image's numax
    uint32_t const blk = blocks[blockIdx.x];
    uint32_t const col = a.colindx[blk];
    if (col_active && !col_active[col]) return;
    int const tid = threadIdx.x, lane = tid & 63, w = tid >> 6, kl = lane >> 4;
    float const R2 = float(a.hgrid[3] * a.hgrid[3]);
    acc_t acc[4][2];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int t = 0; t < 2; ++t) acc[mt][t] = acc_t{0, 0, 0, 0};
    int const nb0 = 32 * w + (lane & 15);   // B operand column of tile 0 (tile 1: + 16)
    for (uint32_t b = a.add_start[blk]; b < a.add_start[blk + 1]; ++b) {
        AddEntry const e = a.add_entry[b];
        int const L = int(e.L);
        if (L > 5) continue;   // added by k_sho_add_wide
        int const nc = ((L + 1) * (L + 2) * (L + 3)) / 6;
        __syncthreads();   // P of the previous image is no longer read
        stage_hermite<real_t>(H, xi2, a.htab, a.xtab, e.slot, L, tid, 256);
        if (tid == 0) {
            int q = 0;
            for (int iz = 0; iz <= L; ++iz)
                for (int iy = 0; iy <= L - iz; ++iy)
                    for (int ix = 0; ix <= L - iz - iy; ++ix) sxyz[q++] = uint16_t(ix | (iy << 4) | (iz << 8));
            for (; q < NCP; ++q) sxyz[q] = 0xFFFF;
        }
        __syncthreads();
        for (int e = tid; e < 64 * NCP; e += 256) {
            int const i = e / NCP, c = e % NCP;
            int const xx = i & 3, yy = (i >> 2) & 3, zz = i >> 4;
            uint16_t const code = sxyz[c];
            real_t v = 0;
            if (code != 0xFFFF && sho_inside(xi2, R2, zz, yy, xx))
                v = H[code & 15][0][xx] * H[(code >> 4) & 15][1][yy] * H[code >> 8][2][zz];
            P[e] = v;
        }
        __syncthreads();
        real_t const* dsrc = a.dcoef + (size_t(e.coef0) * a.nCols + col) * 128 + nb0;
        size_t const dstride = size_t(a.nCols) * 128;
        int const nks = (nc + 3) / 4;
        for (int ks = 0; ks < nks; ++ks) {
            int const sk = ks * 4 + kl;
            real_t const b0v = sk < nc ? dsrc[size_t(sk) * dstride] : real_t(0);
            real_t const b1v = sk < nc ? dsrc[size_t(sk) * dstride + 16] : real_t(0);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                real_t const av = P[(mt * 16 + (lane & 15)) * NCP + ks * 4 + kl];
                acc[mt][0] = Mfma<real_t>::mac(av, b0v, acc[mt][0]);
                acc[mt][1] = Mfma<real_t>::mac(av, b1v, acc[mt][1]);
            }
        }
    }
    // y += acc: point mt * 16 + row, column n of tile t
    real_t* yb = a.y + size_t(blk) * NBLK;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int const i = mt * 16 + Mfma<real_t>::row(kl, r), n = nb0 + 16 * t;
                real_t& o = yb[(n >> 6) * 4096 + i * 64 + (n & 63)];
                o += acc[mt][t][r];
            }
}

// SHO add-back of the atom images with numax 6 or 7, y += P d, on the blocks listed in blocks[]: one work-group of
// 512 threads = (column j, plane p, z plane zq) per block, taking the block's wide images one after the other, so
// that no two work-groups update the same block.  Runs after the local kernel (which adds numax <= 5).
template <typename real_t>
__global__ __launch_bounds__(512) void k_sho_add_wide(LocalArgs<real_t> a, uint32_t const* __restrict__ blocks,
                                                      uint8_t const* __restrict__ col_active) {
    __shared__ real_t H[8][3][4];
    __shared__ float xi2[3][4];
    uint32_t const blk = blocks[blockIdx.x];
    uint32_t const col = a.colindx[blk];
    if (col_active && !col_active[col]) return;
    int const tid = threadIdx.x, j = tid & 63, p = (tid >> 6) & 1, zq = tid >> 7;
    float const R2 = float(a.hgrid[3] * a.hgrid[3]);
    real_t* yb = a.y + size_t(blk) * NBLK + p * 4096 + j;   // ys stride 64 for the direct update of y
    for (uint32_t b = a.add_start[blk]; b < a.add_start[blk + 1]; ++b) {
        AddEntry const e = a.add_entry[b];
        int const L = int(e.L);
        if (L <= 5) continue;   // added by the local kernel
        __syncthreads();        // H of the previous image is no longer read
        stage_hermite<real_t>(H, xi2, a.htab, a.xtab, e.slot, L, tid, 512);
        __syncthreads();
        real_t const* dsrc = a.dcoef + (size_t(e.coef0) * a.nCols + col) * 128 + p * 64 + j;
        sho_add_atom<real_t, 7>(yb, 64, H, xi2, R2, dsrc, size_t(a.nCols) * 128, zq, L);
    }
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

// items[i] = heads[order[i]] for i < n, one word per thread
__global__ __launch_bounds__(256) void k_gather_heads(uint32_t const* __restrict__ order, uint64_t n,
                                                      uint32_t const* __restrict__ heads, uint32_t* __restrict__ items,
                                                      int hw) {
    uint64_t const t = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (t >= n * uint64_t(hw)) return;
    uint64_t const i = t / hw, w = t % hw;
    items[t] = heads[uint64_t(order[i]) * hw + w];
}

}  // namespace

int device_gather_heads(DevicePlan const& d, uint32_t const* order, uint64_t n, uint32_t* items, hipStream_t stream,
                        std::string& err) {
    if (n == 0) return 0;
    uint64_t const words = n * uint64_t(d.head_words);
    uint64_t const grid = (words + 255) / 256;
    if (grid > 0x7fffffffull) { err = "too many blocks for one launch"; return 1; }
    hipLaunchKernelGGL(k_gather_heads, dim3(uint32_t(grid)), dim3(256), 0, stream, order, n, d.heads, items,
                       d.head_words);
    HIP_TRY(hipGetLastError());
    return 0;
}

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
    {   // processing order of the local kernel: per source column, xy tiles of TILE x TILE target blocks swept
        // along z, so that the 6 stencil neighbours of a block were read shortly before (z neighbours
        // TILE^2 blocks back); consecutive work items then share one XCD's L2
        int const TILE = LOCAL_TILE;
        std::vector<uint64_t> key(p.nnzb);
        int32_t lo[3] = {1 << 30, 1 << 30, 1 << 30};
        for (uint32_t r = 0; r < p.nRows; ++r)
            for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], p.row_coords[3 * r + k]);
        for (uint64_t b = 0; b < p.nnzb; ++b) {
            int32_t const* g = &p.row_coords[3 * size_t(p.rowindx[b])];
            uint64_t const x = uint64_t(g[0] - lo[0]), y = uint64_t(g[1] - lo[1]), z = uint64_t(g[2] - lo[2]);
            uint64_t const tx = x / TILE, ty = y / TILE;
            uint64_t const txs = (ty & 1) ? (1023 - tx) : tx;  // serpentine over tiles
            key[b] = (uint64_t(p.colindx[b]) << 48) | (ty << 38) | (txs << 28) | (z << 8) | ((y % TILE) << 4) | (x % TILE);
        }
        std::vector<uint32_t> order(p.nnzb);
        for (uint64_t b = 0; b < p.nnzb; ++b) order[b] = uint32_t(b);
        std::sort(order.begin(), order.end(), [&](uint32_t u, uint32_t v) { return key[u] < key[v]; });
        d.h_order = std::move(order);
    }

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
                            if (code <= -2) d.has_images = true;
                        }
                    }
                }
            }
        }

        // the local kernels' block headers: block, column group, potential cell, add-list range, neighbour codes
        int const HW = (NH == 1) ? 16 : 32;
        d.head_words = HW;
        std::vector<uint32_t> heads(size_t(HW) * p.nnzb, 0);
        for (uint64_t b = 0; b < p.nnzb; ++b) {
            uint32_t* h = &heads[size_t(HW) * b];
            h[IH_BLK] = uint32_t(b);
            h[IH_COL] = p.colindx[b];
            h[IH_VEFF] = uint32_t(p.veff_index[b]);
            h[IH_ADD0] = p.sho_add.start[b];
            h[IH_ADD1] = p.sho_add.start[b + 1];
            for (int c = 0; c < 6 * NH; ++c) h[IH_NBR + c] = uint32_t(nb[size_t(c) * p.nnzb + b]);
        }
        std::vector<uint32_t> items(heads.size());
        for (uint64_t i = 0; i < p.nnzb; ++i)
            std::copy_n(&heads[size_t(HW) * d.h_order[i]], HW, &items[size_t(HW) * i]);
        if ((s = dupload(d, d.heads, heads, err))) return s;
        if ((s = dupload(d, d.items, items, err))) return s;
    }

    // dyadic structures
    if ((s = dupload(d, d.add_start, p.sho_add.start, err))) return s;
    {   // blocks touched by atom images with numax > 5, whose add-back runs in a separate kernel after k_local
        std::vector<uint32_t> wide;
        for (uint64_t blk = 0; blk < p.nnzb; ++blk)
            for (uint32_t k = p.sho_add.start[blk]; k < p.sho_add.start[blk + 1]; ++k)
                if (p.image_lmax[p.sho_add.index[k]] > 5) { wide.push_back(uint32_t(blk)); break; }
        d.n_wide_blocks = uint32_t(wide.size());
        if ((s = dupload(d, d.wide_blocks, wide, err))) return s;
        // blocks touched by atom images with numax <= 5 (the add-back on the matrix cores)
        std::vector<uint32_t> narrow;
        for (uint64_t blk = 0; blk < p.nnzb; ++blk)
            for (uint32_t k = p.sho_add.start[blk]; k < p.sho_add.start[blk + 1]; ++k)
                if (p.image_lmax[p.sho_add.index[k]] <= 5) { narrow.push_back(uint32_t(blk)); break; }
        d.n_narrow_blocks = uint32_t(narrow.size());
        if ((s = dupload(d, d.narrow_blocks, narrow, err))) return s;
    }
    {   // add-list entries in one record each: image, numax, the slot of (image, target row) in the Hermite table
        // (the projection chunks below add their own slots), first coefficient of the image
        std::map<uint64_t, uint32_t> slot_of;
        std::vector<uint32_t> entry(4 * p.sho_add.index.size());
        for (uint64_t blk = 0; blk < p.nnzb; ++blk)
            for (uint32_t k = p.sho_add.start[blk]; k < p.sho_add.start[blk + 1]; ++k) {
                uint32_t const img = p.sho_add.index[k];
                uint64_t const key = (uint64_t(img) << 32) | p.rowindx[blk];
                entry[4 * size_t(k) + 0] = img;
                entry[4 * size_t(k) + 1] = uint32_t(p.image_lmax[img]);
                entry[4 * size_t(k) + 2] = slot_of.emplace(key, uint32_t(slot_of.size())).first->second;
                entry[4 * size_t(k) + 3] = p.image_starts[img];
            }
        d.slot_map = std::move(slot_of);
        if ((s = dupload(d, d.add_entry, entry, err))) return s;
    }
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
        std::vector<uint32_t> prj_slot(blocks.size());
        for (auto const& chk : chunks)
            for (uint32_t b = chk.begin; b < chk.end; ++b) {
                uint64_t const key = (uint64_t(chk.image) << 32) | p.rowindx[blocks[b]];
                auto const it = d.slot_map.emplace(key, uint32_t(d.slot_map.size())).first;
                prj_slot[b] = it->second;
            }
        if ((s = dupload(d, d.chunks, chunks, err))) return s;
        if ((s = dupload(d, d.prj_blocks, blocks, err))) return s;
        if ((s = dupload(d, d.prj_slot, prj_slot, err))) return s;
        {   // the table itself, computed on the device by hermite_1d
            size_t const ns = std::max<size_t>(1, d.slot_map.size());
            std::vector<uint32_t> simg(ns, 0), srow(ns, 0);
            for (auto const& kv : d.slot_map) { simg[kv.second] = uint32_t(kv.first >> 32); srow[kv.second] = uint32_t(kv.first & 0xffffffffu); }
            uint32_t *dimg = nullptr, *drow = nullptr;
            if ((s = dupload(d, dimg, simg, err))) return s;
            if ((s = dupload(d, drow, srow, err))) return s;
            if ((s = dalloc(d, d.htab, ns * 96, err))) return s;
            if ((s = dalloc(d, d.xtab, ns * 12, err))) return s;
            if (!d.slot_map.empty()) {
                hipLaunchKernelGGL(k_hermite_table, dim3(uint32_t(d.slot_map.size())), dim3(64), 0, 0, d.htab, d.xtab,
                                   dimg, drow, d.image_pos, d.image_lmax, d.row_pos, p.hg[0], p.hg[1], p.hg[2], p.r_proj);
                HIP_TRY(hipGetLastError());
                HIP_TRY(hipDeviceSynchronize());
            }
            d.nslots = uint32_t(d.slot_map.size());
            d.slot_map.clear();
        }
        if ((s = dupload(d, d.chunk_coef, coef, err))) return s;
        if ((s = dupload(d, d.pair_chunk, pair_chunk, err))) return s;
        if ((s = dupload(d, d.bucket_chunks, bucket, err))) return s;
    }

    if ((s = dalloc(d, d.V, p.ncell() * 64, err))) return s;
    if ((s = dalloc(d, d.atom_hs, p.atom_hs.size(), err))) return s;
    if ((s = dalloc(d, d.E_col, size_t(2) * p.nCols, err))) return s;
    if ((s = dalloc(d, d.image_phase, p.image_phase.size(), err))) return s;
    d.uploaded = true;
    return 0;
}

int device_upload_values(Plan const& p, DevicePlan& d, std::string& err) {
    HIP_TRY(hipMemcpy(d.V, p.V_cell.data(), p.V_cell.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!p.atom_hs.empty())
        HIP_TRY(hipMemcpy(d.atom_hs, p.atom_hs.data(), p.atom_hs.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.E_col, p.E_col.data(), p.E_col.size() * sizeof(std::complex<double>), hipMemcpyHostToDevice));
    if (!p.image_phase.empty())
        HIP_TRY(hipMemcpy(d.image_phase, p.image_phase.data(), p.image_phase.size() * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

// ---------------------------------------------------------------------------------------------
// 4. adjoint gradient of L = Im sum_c w_c Tr[G_c], Tr over the diagonal block of column group c.
//    With dG = -A^-1 dA G and lambda_c = A^-H P_c: dL = -Im sum_c w_c <lambda_c, dA G_c>.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// per block and point i: -Im(w_c sum_j conj(lambda[b][i][j]) G[b][i][j]), 0 outside the cell (no potential)
__global__ __launch_bounds__(256) void k_grad_v_blocks(double* __restrict__ gpart, double const* __restrict__ X,
                                                       double const* __restrict__ Lam, uint16_t const* __restrict__ colindx,
                                                       int32_t const* __restrict__ veff, double const* __restrict__ wE) {
    uint64_t const blk = blockIdx.x;
    int const j = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t const c = colindx[blk];
    double const wr = wE[4 * c], wi = wE[4 * c + 1];
    bool const inside = veff[blk] >= 0;
    for (int i = w; i < 64; i += 4) {
        size_t const er = (blk * 2 + 0) * 4096 + i * 64 + j, ei = er + 4096;
        double const gr = X[er], gi = X[ei], lr = Lam[er], li = Lam[ei];
        double const sr = wave_sum(lr * gr + li * gi), si = wave_sum(lr * gi - li * gr);
        if (j == 0) gpart[blk * 64 + i] = inside ? -(wr * si + wi * sr) : 0.0;
    }
}

// per cell block: the sum over the blocks that read its potential, in ascending block order
__global__ __launch_bounds__(64) void k_grad_v_cells(double* __restrict__ gcell, double const* __restrict__ gpart,
                                                     uint32_t const* __restrict__ cell_start,
                                                     uint32_t const* __restrict__ cell_blocks) {
    uint32_t const v = blockIdx.x;
    int const i = threadIdx.x;
    double acc = 0;
    for (uint32_t k = cell_start[v]; k < cell_start[v + 1]; ++k) acc += gpart[uint64_t(cell_blocks[k]) * 64 + i];
    gcell[uint64_t(v) * 64 + i] = acc;
}

// atom coefficients of one vector: per (atom, column group) the image partials collected with the Bloch phase,
// out[((atom_starts[a] + s) * nCols + c)][re/im][j] (the input of the atom-matrix multiply of k_sho_atoms)
__global__ __launch_bounds__(64 * ATOM_WAVES) void k_sho_collect(
    double* __restrict__ out, double const* __restrict__ partial, uint32_t const* __restrict__ chunk_coef,
    uint32_t const* __restrict__ pair_chunk, uint32_t const* __restrict__ sum_start,
    uint32_t const* __restrict__ sum_index, int8_t const* __restrict__ atom_lmax,
    double const* __restrict__ image_phase, uint32_t nCols, uint32_t const* __restrict__ atom_starts) {
    uint32_t const a = blockIdx.x, c = blockIdx.y;
    int const j = threadIdx.x & 63, w = threadIdx.x >> 6;
    int const L = atom_lmax[a];
    int const nc = ((L + 1) * (L + 2) * (L + 3)) / 6;
    for (int s = w; s < nc; s += ATOM_WAVES) {
        double ar = 0, ai = 0;
        for (uint32_t k = sum_start[a]; k < sum_start[a + 1]; ++k) {
            uint32_t const img = sum_index[k];
            double const ph_re = image_phase[2 * img], ph_im = image_phase[2 * img + 1];
            uint32_t const c0 = pair_chunk[size_t(img) * nCols + c], c1 = pair_chunk[size_t(img) * nCols + c + 1];
            double vr = 0, vi = 0;
            for (uint32_t ch = c0; ch < c1; ++ch) {
                double const* src = partial + (size_t(chunk_coef[ch]) + s) * 128 + j;
                vr += src[0];
                vi += src[64];
            }
            ar += ph_re * vr - ph_im * vi;
            ai += ph_re * vi + ph_im * vr;
        }
        double* dst = out + ((size_t(atom_starts[a]) + s) * nCols + c) * 128 + j;
        dst[0] = ar;
        dst[64] = ai;
    }
}

// per (atom a, row k): for every l, z_c = sum_j conj(cL[k][c][j]) cX[l][c][j];
// gmat[a] = [-Im sum_c w_c z_c][k][l] (dL/dM entries times 1) and [-Im sum_c w_c (-E_c) z_c][k][l]
__global__ __launch_bounds__(64) void k_grad_atoms(double* __restrict__ gmat, double const* __restrict__ cX,
                                                   double const* __restrict__ cL, int8_t const* __restrict__ atom_lmax,
                                                   uint32_t const* __restrict__ atom_starts,
                                                   uint64_t const* __restrict__ atom_mat_off,
                                                   double const* __restrict__ wE, uint32_t nCols) {
    uint32_t const a = blockIdx.x;
    int const k = blockIdx.y, j = threadIdx.x;
    int const L = atom_lmax[a];
    int const nc = ((L + 1) * (L + 2) * (L + 3)) / 6;
    if (k >= nc) return;
    size_t const base = size_t(atom_starts[a]) * nCols * 128;
    for (int l = 0; l < nc; ++l) {
        double hr = 0, sr = 0;   // -Im of the two weighted sums
        for (uint32_t c = 0; c < nCols; ++c) {
            double const* pl = cL + base + (size_t(k) * nCols + c) * 128 + j;
            double const* px = cX + base + (size_t(l) * nCols + c) * 128 + j;
            double const lr = pl[0], li = pl[64], xr = px[0], xi = px[64];
            double const zr = lr * xr + li * xi, zi = lr * xi - li * xr;   // conj(l) x
            double const w_r = wE[4 * c], w_i = wE[4 * c + 1], v_r = wE[4 * c + 2], v_i = wE[4 * c + 3];
            hr -= w_r * zi + w_i * zr;
            sr -= v_r * zi + v_i * zr;
        }
        hr = wave_sum(hr);
        sr = wave_sum(sr);
        if (j == 0) {
            gmat[atom_mat_off[a] + size_t(k) * nc + l] = hr;
            gmat[atom_mat_off[a] + size_t(nc) * nc + size_t(k) * nc + l] = sr;
        }
    }
}

// Local-kernel tile shapes.  kind 0: k_local (LDS accumulator), JW columns per work-group (16 or 32), LB
// kinetic lines per thread in flight (1, 2, 4).  kind 1: k_plane (register accumulator), JW = 32 or 64.
// kind 2: k_lines (16-byte lanes, line ownership), JW = 16 or 32, or 64 in FP32; LB = directions of neighbour
// lines issued up front (3 or 2; 16th-order stencils issue one direction at a time).
template <typename real_t, int NH, bool IMG>
static void launch_local(int kind, int jw, int lb, dim3 g, hipStream_t stream, LocalArgs<real_t> const& a) {
#define KL(JW_, LB_) hipLaunchKernelGGL((k_local<real_t, NH, IMG, JW_, LB_>), g, dim3(8 * JW_), 0, stream, a)
#define KP(JW_) hipLaunchKernelGGL((k_plane<real_t, NH, IMG, JW_>), g, dim3(8 * JW_), 0, stream, a)
#define KN(JW_, PF_) hipLaunchKernelGGL((k_lines<real_t, NH, IMG, JW_, PF_>), g, dim3((JW_ / Lane<real_t>::n) * 32), 0, stream, a)
    if (kind == 2) {   // lb: directions prefetched (3 or 2; 16th-order stencils always 1)
        if constexpr (NH == 2) {
            if (jw == 16) KN(16, 1);
            else if (jw == 64) { if constexpr (sizeof(real_t) == 4) KN(64, 1); }
            else KN(32, 1);
        } else if (lb == 2) {
            if (jw == 16) KN(16, 2);
            else if (jw == 64) { if constexpr (sizeof(real_t) == 4) KN(64, 2); }
            else KN(32, 2);
        } else {
            if (jw == 16) KN(16, 3);
            else if (jw == 64) { if constexpr (sizeof(real_t) == 4) KN(64, 3); }
            else KN(32, 3);
        }
    } else if (kind == 1) {
        if (jw == 64) KP(64); else KP(32);
    } else if (jw == 16) {
        if (lb == 1) KL(16, 1); else if (lb == 4) KL(16, 4); else KL(16, 2);
    } else {
        if (lb == 1) KL(32, 1); else if (lb == 4) KL(32, 4); else KL(32, 2);
    }
#undef KL
#undef KP
#undef KN
}

template <typename real_t>
static Workspace<real_t>& workspace(DevicePlan& d);
template <> Workspace<double>& workspace<double>(DevicePlan& d) { return d.ws64; }
template <> Workspace<float>& workspace<float>(DevicePlan& d) { return d.ws32; }

template <typename real_t>
int device_action(DevicePlan& d, ActionScalars const& s, real_t* y, real_t const* x, hipStream_t stream,
                  std::string& err, int parts, double* dot, Workspace<real_t>* ext, double* ext_cscratch) {
    auto& ws = ext ? *ext : workspace<real_t>(d);
    if (!ext) {
        // every buffer of this precision is allocated on the first call, so that later calls (also under graph
        // capture) allocate nothing; parts = 0 only allocates
        if (d.nAtomImages > 0 && !ws.dcoef) {
            if (int r = dalloc(d, ws.partial, size_t(d.total_chunk_coefs) * 128, err)) return r;
            if (int r = dalloc(d, ws.dcoef, size_t(d.natomcoeffs) * d.nCols * 128, err)) return r;
        }
        if (d.nAtomImages > 0 && d.nchunks > 0 && d.max_atom_nc > 64 && !d.cscratch) {
            if (int r = dalloc(d, d.cscratch, size_t(d.ncoeffs) * d.nCols * 128, err)) return r;
        }
        if (!ws.zero) {
            if (int r = dalloc(d, ws.zero, NBLK, err)) return r;
            HIP_TRY(hipMemsetAsync(ws.zero, 0, NBLK * sizeof(real_t), stream));
        }
    }
    // after the allocations above: the atom-multiply scratch of numax 6, 7 atoms exists from here on
    double* const cscratch = ext ? ext_cscratch : d.cscratch;
    if ((parts & ACTION_PROJECT) && d.nAtomImages > 0 && d.nchunks > 0 && d.max_atom_nc > 64 && !cscratch) {
        err = "no scratch for the atom matrices of numax 6, 7 atoms";
        return 1;
    }
    if ((parts & ACTION_PROJECT) && d.nAtomImages > 0 && d.nchunks > 0) {
        for (int L = 0; L <= 7; ++L) {
            uint32_t const n = d.bucket_start[L + 1] - d.bucket_start[L];
            if (!n) continue;
            uint32_t const* bucket = d.bucket_chunks + d.bucket_start[L];
#define LAUNCH_PRJ(LL) hipLaunchKernelGGL((k_sho_project<real_t, LL>), dim3(n), dim3(128), 0, stream, ws.partial, x, \
                d.chunks, bucket, d.chunk_coef, d.prj_blocks, d.rowindx, d.row_pos, d.image_pos, \
                s.hgrid[0], s.hgrid[1], s.hgrid[2], s.hgrid[3], d.col_active, d.htab, d.xtab, d.prj_slot)
#define LAUNCH_MFMA(LL) hipLaunchKernelGGL((k_sho_project_mfma<real_t, LL>), dim3(n), dim3(256), 0, stream, ws.partial, \
                x, d.chunks, bucket, d.chunk_coef, d.prj_blocks, s.hgrid[3], d.col_active, d.htab, d.xtab, d.prj_slot)
            if (d.prj_engine[sizeof(real_t) == 8 ? 0 : 1] == 1 && L <= 5) {   // the dense form on the matrix cores
                switch (L) {
                    case 0: LAUNCH_MFMA(0); break; case 1: LAUNCH_MFMA(1); break;
                    case 2: LAUNCH_MFMA(2); break; case 3: LAUNCH_MFMA(3); break;
                    case 4: LAUNCH_MFMA(4); break; default: LAUNCH_MFMA(5); break;
                }
            } else {
                switch (L) {
                    case 0: LAUNCH_PRJ(0); break; case 1: LAUNCH_PRJ(1); break;
                    case 2: LAUNCH_PRJ(2); break; case 3: LAUNCH_PRJ(3); break;
                    case 4: LAUNCH_PRJ(4); break; case 5: LAUNCH_PRJ(5); break;
                    case 6: LAUNCH_PRJ(6); break; default: LAUNCH_PRJ(7); break;
                }
            }
#undef LAUNCH_PRJ
#undef LAUNCH_MFMA
            HIP_TRY(hipGetLastError());
        }
        uint32_t const ncl = std::min<uint32_t>(d.max_atom_nc, 64);   // larger atoms use the global scratch
        size_t const lds = size_t(2) * ncl * 64 * sizeof(double);
        hipLaunchKernelGGL((k_sho_atoms<real_t>), dim3(d.nAtoms, d.nCols), dim3(64 * ATOM_WAVES), lds, stream, ws.dcoef,
                           ws.partial, d.chunk_coef, d.pair_chunk, d.sum_start, d.sum_index, d.image_starts,
                           d.atom_lmax, d.atom_hs, d.atom_mat_off, d.image_phase, d.nCols, cscratch,
                           d.atom_starts, d.E_col, d.col_active);
        HIP_TRY(hipGetLastError());
    }
    if (!(parts & ACTION_LOCAL)) return 0;
    LocalArgs<real_t> a;
    a.y = y; a.x = x; a.dcoef = ws.dcoef; a.V = d.V;
    a.add_start = d.add_start; a.colindx = d.colindx;
    a.nCols = d.nCols;
    a.zero = ws.zero;
    a.items = d.active_order ? d.active_items : d.items;
    a.add_entry = reinterpret_cast<AddEntry const*>(d.add_entry);
    a.nitems = d.active_order ? d.n_active : d.nnzb;
    a.E_col = d.E_col;
    a.htab = d.htab; a.xtab = d.xtab;
    a.dot = dot;
    int const engine = d.prj_engine[sizeof(real_t) == 8 ? 0 : 1];
    a.add_local = (engine == 1) ? 0 : 1;
    static int const c8i[5] = {-14350, 8064, -1008, 128, -9};
    static int const c16i[9] = {-924708642, 538137600, -94174080, 22830080, -5350800, 1053696, -156800, 15360, -735};
    double cdiag = 0;
    for (int dd = 0; dd < 3; ++dd) {
        double const n8 = s.prefactor[dd] / 5040., n16 = s.prefactor[dd] / 302702400.;
        for (int m = 0; m < 5; ++m) a.c8[dd][m] = real_t(c8i[m] * n8);
        for (int m = 0; m < 9; ++m) a.c16[dd][m] = real_t(c16i[m] * n16);
        a.nh[dd] = (d.fd_range[dd] == 8) ? 2 : 1;
        cdiag += (a.nh[dd] == 2) ? c16i[0] * n16 : c8i[0] * n8;
        for (int u = 0; u < 2; ++u) for (int v = 0; v < 2; ++v) a.phase[dd][u][v] = s.phase[dd][u][v];
    }
    a.cdiag = real_t(cdiag);
    for (int k = 0; k < 4; ++k) a.hgrid[k] = s.hgrid[k];
    // work-groups are dealt round-robin to the 8 XCDs: give each XCD a contiguous range of
    // (block, column slice) items so that neighbour blocks meet in the same L2
    int const v = sizeof(real_t) == 8 ? 0 : 1;
    int const KIND = d.tune_kind[v], JW = d.tune_jw[v], LB = d.tune_lb[v];
    if (a.nitems == 0) return 0;
    uint64_t const items = uint64_t(64 / JW) * a.nitems;
    uint64_t const per = (items + 7) / 8;
    a.per_xcd = uint32_t(per);
    uint64_t const grid = per * 8;
    if (grid > 0x7fffffffull) { err = "too many blocks for one launch"; return 1; }
    dim3 const g{uint32_t(grid), 1, 1};
    if (d.nh_max == 2) {
        if (d.has_images) launch_local<real_t, 2, true>(KIND, JW, LB, g, stream, a);
        else launch_local<real_t, 2, false>(KIND, JW, LB, g, stream, a);
    } else {
        if (d.has_images) launch_local<real_t, 1, true>(KIND, JW, LB, g, stream, a);
        else launch_local<real_t, 1, false>(KIND, JW, LB, g, stream, a);
    }
    HIP_TRY(hipGetLastError());
    if (!a.add_local && d.n_narrow_blocks > 0) {   // the numax <= 5 images on the matrix cores
        hipLaunchKernelGGL((k_sho_add_mfma<real_t>), dim3(d.n_narrow_blocks), dim3(256), 0, stream, a, d.narrow_blocks,
                           d.col_active);
        HIP_TRY(hipGetLastError());
    }
    if (d.n_wide_blocks > 0) {
        hipLaunchKernelGGL((k_sho_add_wide<real_t>), dim3(d.n_wide_blocks), dim3(512), 0, stream, a, d.wide_blocks,
                           d.col_active);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

template int device_action<double>(DevicePlan&, ActionScalars const&, double*, double const*, hipStream_t, std::string&,
                                   int, double*, Workspace<double>*, double*);
template int device_action<float>(DevicePlan&, ActionScalars const&, float*, float const*, hipStream_t, std::string&,
                                  int, double*, Workspace<float>*, double*);

// ---------------------------------------------------------------------------------------------
// The action without allocation (gfa_action_noalloc): the projection partial sums, the addition coefficients,
// the atom-multiply scratch of numax 6, 7 atoms and the zero block live in `scratch`, a fixed array of every GPU
// the library runs on (held until the process ends).  Its device address is taken once per device with
// hipGetSymbolAddress; the host-side name `scratch` is never handed to a kernel.  Calls that overlap on one GPU
// share the array and race.
// ---------------------------------------------------------------------------------------------
namespace {

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[MAX_DEVICES] = {};

// the address of `scratch` on the current device (which must be `device`), looked up on its first use there
int scratch_address(int device, unsigned char** out, std::string& err) {
    if (device < 0 || device >= MAX_DEVICES) { err = "device ordinal out of range"; return 1; }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        hipError_t const e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return 2;
        }
        g_scratch_ptr[device] = ptr;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
    return 0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill_zero(T* __restrict__ p, uint32_t n) {
    uint32_t const i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = T(0);
}

size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

size_t noalloc_scratch_bytes(DevicePlan const& d, int fp_bits) {
    size_t const s = fp_bits == 64 ? 8 : 4;
    size_t b = round256(size_t(NBLK) * s);   // zero block
    if (d.nAtomImages > 0) {
        b += round256(size_t(d.total_chunk_coefs) * 128 * s);
        b += round256(size_t(d.natomcoeffs) * d.nCols * 128 * s);
        if (d.nchunks > 0 && d.max_atom_nc > 64) b += round256(size_t(d.ncoeffs) * d.nCols * 128 * 8);
    }
    return b;
}

template <typename real_t>
int device_action_noalloc(DevicePlan& d, ActionScalars const& s, real_t* y, real_t const* x, hipStream_t stream,
                          std::string& err) {
    size_t const need = noalloc_scratch_bytes(d, sizeof(real_t) == 8 ? 64 : 32);
    if (need > NOALLOC_SCRATCH_BYTES) {
        err = "the plan needs " + std::to_string(need) + " bytes of action scratch, more than the " +
              std::to_string(NOALLOC_SCRATCH_BYTES) + " of the allocation-free action";
        return 1;
    }
    unsigned char* base = nullptr;
    if (int r = scratch_address(d.device, &base, err)) return r;
    unsigned char* cur = base;
    auto take = [&](size_t bytes) { unsigned char* p = cur; cur += round256(bytes); return p; };
    Workspace<real_t> ws;
    ws.zero = reinterpret_cast<real_t*>(take(size_t(NBLK) * sizeof(real_t)));
    double* csc = nullptr;
    if (d.nAtomImages > 0) {
        ws.partial = reinterpret_cast<real_t*>(take(size_t(d.total_chunk_coefs) * 128 * sizeof(real_t)));
        ws.dcoef = reinterpret_cast<real_t*>(take(size_t(d.natomcoeffs) * d.nCols * 128 * sizeof(real_t)));
        if (d.nchunks > 0 && d.max_atom_nc > 64)
            csc = reinterpret_cast<double*>(take(size_t(d.ncoeffs) * d.nCols * 128 * 8));
    }
    // the array keeps the previous call's contents: the zero block is cleared on every call (a kernel, so that
    // the call can be captured in a graph); the other regions are written before they are read
    hipLaunchKernelGGL((k_fill_zero<real_t>), dim3((NBLK + 255) / 256), dim3(256), 0, stream, ws.zero, uint32_t(NBLK));
    HIP_TRY(hipGetLastError());
    return device_action<real_t>(d, s, y, x, stream, err, ACTION_ALL, nullptr, &ws, csc);
}
template int device_action_noalloc<double>(DevicePlan&, ActionScalars const&, double*, double const*, hipStream_t,
                                           std::string&);
template int device_action_noalloc<float>(DevicePlan&, ActionScalars const&, float*, float const*, hipStream_t,
                                          std::string&);

// Times every k_local tile shape on the caller's vectors (the whole action, reps launches each, HIP events on
// `stream`) and keeps the fastest for this precision.  y is overwritten.
template <typename real_t>
int device_autotune(DevicePlan& d, ActionScalars const& s, real_t* y, real_t const* x, hipStream_t stream, int reps,
                    double* best_ms, std::string& err) {
    int const v = sizeof(real_t) == 8 ? 0 : 1;
    static int const SHAPES[14][3] = {{0, 16, 1}, {0, 16, 2}, {0, 16, 4}, {0, 32, 1}, {0, 32, 2}, {0, 32, 4},
                                      {1, 32, 0}, {1, 64, 0}, {2, 16, 3}, {2, 32, 3}, {2, 64, 3}, {2, 16, 2},
                                      {2, 32, 2}, {2, 64, 2}};
    double best = 1e300;
    int bkind = d.tune_kind[v], bjw = d.tune_jw[v], blb = d.tune_lb[v];
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // the shape in use before the call (or the fastest so far) is restored on every return, also on errors
    auto finish = [&](int status) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        d.tune_kind[v] = bkind; d.tune_jw[v] = bjw; d.tune_lb[v] = blb;
        return status;
    };
    auto timed = [&]() -> int {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        for (auto const& sh : SHAPES) {
            int const kind = sh[0], jw = sh[1], lb = sh[2];
            if (kind == 2 && jw == 64 && sizeof(real_t) == 8) continue;   // no FP64 kind-2 shape of 64 columns
            d.tune_kind[v] = kind; d.tune_jw[v] = jw; d.tune_lb[v] = lb;
            if (int r = device_action<real_t>(d, s, y, x, stream, err)) return r;  // warm-up
            HIP_TRY(hipEventRecord(e0, stream));
            for (int k = 0; k < reps; ++k)
                if (int r = device_action<real_t>(d, s, y, x, stream, err)) return r;
            HIP_TRY(hipEventRecord(e1, stream));
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            double const t = double(ms) / std::max(1, reps);
            if (t < best) { best = t; bkind = kind; bjw = jw; blb = lb; }
        }
        return 0;
    };
    if (int r = timed()) return finish(r);
    // the projection's engine with the fastest shape: the vector form, then the matrix cores
    if (d.nAtomImages > 0 && d.nchunks > 0) {
        d.tune_kind[v] = bkind; d.tune_jw[v] = bjw; d.tune_lb[v] = blb;
        double tbest[2] = {1e300, 1e300};
        int const keep = d.prj_engine[v];
        for (int eng = 0; eng < 2; ++eng) {
            d.prj_engine[v] = eng;
            if (int r = device_action<real_t>(d, s, y, x, stream, err)) { d.prj_engine[v] = keep; return finish(r); }
            if (hipEventRecord(e0, stream) != hipSuccess) { d.prj_engine[v] = keep; return finish(2); }
            for (int k = 0; k < reps; ++k)
                if (int r = device_action<real_t>(d, s, y, x, stream, err)) { d.prj_engine[v] = keep; return finish(r); }
            float ms = 0;
            if (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                d.prj_engine[v] = keep;
                err = "autotune: event timing failed";
                return finish(2);
            }
            tbest[eng] = double(ms) / std::max(1, reps);
        }
        d.prj_engine[v] = tbest[1] < tbest[0] ? 1 : 0;
        best = std::min(tbest[0], tbest[1]);
    }
    d.tuned[v] = true;
    if (best_ms) *best_ms = best;
    return finish(0);
}
template int device_autotune<double>(DevicePlan&, ActionScalars const&, double*, double const*, hipStream_t, int, double*, std::string&);
template int device_autotune<float>(DevicePlan&, ActionScalars const&, float*, float const*, hipStream_t, int, double*, std::string&);

// The adjoint gradient (FP64): X solves (H - E_c S) X = B, Lam solves (H - conj(E_c) S) Lam = B column group by
// column group; wE[c] = {w_c, -E_c w_c}.  Results on the host: gcell_host[ncell][64] (dL/dV per cell point) and
// gmat_host at atom_mat_off (per contributing atom [2][nc][nc], the derivative with respect to the entries of
// (h dVol-scaled, s dVol-scaled) before the SHO prefactors, see gfa_gradient).
// projection of one FP64 vector onto every atom image and collection per (atom, column group) with the Bloch
// phases: out[((atom_starts[a] + s) * nCols + c)][re/im][j]
static int project_collect(DevicePlan& d, ActionScalars const& s, double const* vec, double* out, hipStream_t stream,
                           std::string& err) {
    auto& ws = d.ws64;
    if (!ws.partial) {
        if (int r = dalloc(d, ws.partial, size_t(d.total_chunk_coefs) * 128, err)) return r;
        if (int r = dalloc(d, ws.dcoef, size_t(d.natomcoeffs) * d.nCols * 128, err)) return r;
    }
    for (int L = 0; L <= 7; ++L) {
        uint32_t const n = d.bucket_start[L + 1] - d.bucket_start[L];
        if (!n) continue;
        uint32_t const* bucket = d.bucket_chunks + d.bucket_start[L];
#define LAUNCH_PRJ(LL) hipLaunchKernelGGL((k_sho_project<double, LL>), dim3(n), dim3(128), 0, stream, ws.partial, vec, \
        d.chunks, bucket, d.chunk_coef, d.prj_blocks, d.rowindx, d.row_pos, d.image_pos, \
        s.hgrid[0], s.hgrid[1], s.hgrid[2], s.hgrid[3], (uint8_t const*)nullptr, d.htab, d.xtab, d.prj_slot)
        switch (L) {
            case 0: LAUNCH_PRJ(0); break; case 1: LAUNCH_PRJ(1); break;
            case 2: LAUNCH_PRJ(2); break; case 3: LAUNCH_PRJ(3); break;
            case 4: LAUNCH_PRJ(4); break; case 5: LAUNCH_PRJ(5); break;
            case 6: LAUNCH_PRJ(6); break; default: LAUNCH_PRJ(7); break;
        }
#undef LAUNCH_PRJ
    }
    hipLaunchKernelGGL(k_sho_collect, dim3(d.nAtoms, d.nCols), dim3(64 * ATOM_WAVES), 0, stream, out, ws.partial,
                       d.chunk_coef, d.pair_chunk, d.sum_start, d.sum_index, d.atom_lmax, d.image_phase, d.nCols,
                       d.atom_starts);
    HIP_TRY(hipGetLastError());
    return 0;
}

int device_atom_coefficients(DevicePlan& d, ActionScalars const& s, double const* X, std::vector<double>& host,
                             hipStream_t stream, std::string& err) {
    size_t const n = size_t(d.ncoeffs) * d.nCols * 128;
    host.assign(n, 0.0);
    if (d.nAtoms == 0 || d.nchunks == 0 || n == 0) return 0;
    if (!d.atom_coef)   // kept with the plan, like the other work buffers
        if (int r = dalloc(d, d.atom_coef, n, err)) return r;
    if (int r = project_collect(d, s, X, d.atom_coef, stream, err)) return r;
    HIP_TRY(hipMemcpyAsync(host.data(), d.atom_coef, n * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return 0;
}

int device_gradient(DevicePlan& d, Plan const& p, ActionScalars const& s, double const* X, double const* Lam,
                    std::vector<double> const& wE, std::vector<double>& gcell_host, std::vector<double>& gmat_host,
                    hipStream_t stream, std::string& err) {
    auto& g = d.grad;
    if (!g.ready) {
        // cell -> blocks that read its potential, ascending block order
        std::vector<uint32_t> start(p.ncell() + 1, 0), blocks;
        for (uint64_t b = 0; b < p.nnzb; ++b) if (p.veff_index[b] >= 0) ++start[p.veff_index[b] + 1];
        for (uint64_t v = 0; v < p.ncell(); ++v) start[v + 1] += start[v];
        blocks.resize(start.back());
        std::vector<uint32_t> fill(start.begin(), start.end() - 1);
        for (uint64_t b = 0; b < p.nnzb; ++b) if (p.veff_index[b] >= 0) blocks[fill[p.veff_index[b]]++] = uint32_t(b);
        int r;
        if ((r = dupload(d, g.cell_start, start, err))) return r;
        if ((r = dupload(d, g.cell_blocks, blocks, err))) return r;
        if ((r = dalloc(d, g.gpart, size_t(p.nnzb) * 64, err))) return r;
        if ((r = dalloc(d, g.gcell, size_t(p.ncell()) * 64, err))) return r;
        if ((r = dalloc(d, g.wE, size_t(4) * p.nCols, err))) return r;
        size_t const nco = std::max<size_t>(1, size_t(d.ncoeffs) * d.nCols * 128);
        if ((r = dalloc(d, g.cX, nco, err))) return r;
        if ((r = dalloc(d, g.cL, nco, err))) return r;
        if ((r = dalloc(d, g.gmat, std::max<size_t>(1, p.atom_mat_off.back()), err))) return r;
        g.ready = true;
    }
    HIP_TRY(hipMemcpyAsync(g.wE, wE.data(), wE.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_grad_v_blocks, dim3(uint32_t(p.nnzb)), dim3(256), 0, stream, g.gpart, X, Lam, d.colindx,
                       d.veff, g.wE);
    hipLaunchKernelGGL(k_grad_v_cells, dim3(uint32_t(p.ncell())), dim3(64), 0, stream, g.gcell, g.gpart, g.cell_start,
                       g.cell_blocks);
    HIP_TRY(hipGetLastError());
    if (d.nAtoms > 0 && d.nchunks > 0) {
        for (int v = 0; v < 2; ++v)
            if (int r = project_collect(d, s, v ? Lam : X, v ? g.cL : g.cX, stream, err)) return r;
        hipLaunchKernelGGL(k_grad_atoms, dim3(d.nAtoms, d.max_atom_nc), dim3(64), 0, stream, g.gmat, g.cX, g.cL,
                           d.atom_lmax, d.atom_starts, d.atom_mat_off, g.wE, d.nCols);
        HIP_TRY(hipGetLastError());
    }
    gcell_host.assign(size_t(p.ncell()) * 64, 0.0);
    gmat_host.assign(p.atom_mat_off.back(), 0.0);
    HIP_TRY(hipMemcpyAsync(gcell_host.data(), g.gcell, gcell_host.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (d.nAtoms > 0 && d.nchunks > 0 && !gmat_host.empty())
        HIP_TRY(hipMemcpyAsync(gmat_host.data(), g.gmat, gmat_host.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return 0;
}

}  // namespace gfa
