/*
 * green_oracle.c -- CPU restatement of the reference's Green-function action and of the TFQMR solve.
 * TEST INFRASTRUCTURE ONLY (see green_oracle.h).  Compile with -ffp-contract=off.
 *
 * Each function names the reference code it restates.  Loops follow the reference's order where the
 * order decides set membership (sphere masks in float); accumulation orders may differ, which only
 * changes results at the level of rounding.
 */
#include "green_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NBLK (2 * 64 * 64)

struct oracle_state {
    oracle_plan p;
    double* V;               /* [ncell][64] */
    double prefactor[3];
    double phase[3][2][2];   /* [dir][left/right][re/im] */
    double* image_phase;     /* [nai][2] */
    double** M;              /* per contributing atom [2][nc][nc] = (h - E s) dVol, rescaled */
};

static int nsho(int L) { return ((L + 1) * (L + 2) * (L + 3)) / 6; }

/* include/sho_projection.hxx:181-183 */
static double sho_1d_norm(int nu, double sigma) {
    double fact = 1;
    for (int k = 2; k <= nu; ++k) fact *= (double)k;
    return sqrt((double)(1ull << nu) / (fact * 1.7724538509055160272981674833411 * sigma));
}

/* src/kinetic_plan.cxx:214-269 (set_phase): exp(2 pi i angle), exact on multiples of 15 degrees */
static void set_phase(double ph[2][2], double angle) {
    double const arg = 2 * 3.14159265358979323846 * angle;
    double cs = cos(arg), sn = sin(arg);
    int const n24 = (int)(angle * 24);
    if (angle * 24 == n24) {
        double const sh = sqrt(0.5), s3 = sqrt(3);
        double const t[7] = {0.0, (s3 - 1) * sh * .5, .5, sh, s3 * .5, (s3 + 1) * sh * .5, 1.0};
        int const m = n24 + 120, quad = m / 6, half = m % 12;
        int const is = (half <= 6) ? half : 12 - half;
        sn = ((quad & 2) ? -1 : 1) * t[is];
        cs = (((quad + 1) & 2) ? -1 : 1) * t[6 - is];
    }
    ph[0][0] = cs; ph[0][1] = sn;
    ph[1][0] = cs; ph[1][1] = -sn;
}

static void cmul(double const a[2], double const b[2], double c[2]) {
    double const re = a[0] * b[0] - a[1] * b[1], im = a[0] * b[1] + a[1] * b[0];
    c[0] = re; c[1] = im;
}

oracle_state* oracle_create(const oracle_plan* p) {
    oracle_state* s = (oracle_state*)calloc(1, sizeof(oracle_state));
    if (!s) return NULL;
    s->p = *p;
    uint32_t const* nb = p->nb;
    size_t const ncell = (size_t)nb[0] * nb[1] * nb[2];
    s->V = (double*)calloc(ncell * 64, sizeof(double));
    /* cell blocks of the potential (src/green_function.cxx:137-159, cell-indexed) */
    size_t const ngx = 4 * (size_t)nb[0], ngy = 4 * (size_t)nb[1];
    for (uint32_t bz = 0; bz < nb[2]; ++bz)
        for (uint32_t by = 0; by < nb[1]; ++by)
            for (uint32_t bx = 0; bx < nb[0]; ++bx)
                for (int i = 0; i < 64; ++i) {
                    int const x = i & 3, y = (i >> 2) & 3, z = i >> 4;
                    s->V[(((size_t)bz * nb[1] + by) * nb[0] + bx) * 64 + i] =
                        p->V_grid[((bz * 4 + z) * ngy + (by * 4 + y)) * ngx + (bx * 4 + x)];
                }
    for (int d = 0; d < 3; ++d) {
        s->prefactor[d] = -0.5 * p->scale_T / (p->hg[d] * p->hg[d]);   /* src/kinetic_plan.cxx:313 */
        set_phase(s->phase[d], p->k[d]);
    }
    /* image phases: product over directions of phase^shift (src/green_function.cxx:236-247) */
    s->image_phase = (double*)calloc(2 * (size_t)p->nAtomImages + 2, sizeof(double));
    for (uint32_t i = 0; i < p->nAtomImages; ++i) {
        double ph[2] = {1, 0};
        for (int d = 0; d < 3; ++d) {
            int const sh = p->image_shift[4 * i + d];
            double const* base = s->phase[d][sh >= 0 ? 0 : 1];
            unsigned n = (unsigned)(sh >= 0 ? sh : -sh);
            double xb[2] = {base[0], base[1]}, xp[2] = {1, 0};
            while (n) { if (n & 1) cmul(xp, xb, xp); n >>= 1; cmul(xb, xb, xb); }
            cmul(ph, xp, ph);
        }
        s->image_phase[2 * i] = ph[0];
        s->image_phase[2 * i + 1] = ph[1];
    }
    /* atom matrices: SHO prefactor rescaling (src/energy_contour.cxx:307-323), then
     * (h - E s) dVol (src/green_function.cxx:80-101) */
    s->M = (double**)calloc(p->nAtoms + 1, sizeof(double*));
    for (uint32_t a = 0; a < p->nAtoms; ++a) {
        int const ia = p->atom_orig[a];
        int const L = (int)p->xyzZinso[ia * 8 + 5];
        double const sigma = p->xyzZinso[ia * 8 + 6];
        int const nc = nsho(L);
        double* f = (double*)malloc(nc * sizeof(double));
        int k = 0;
        for (int nz = 0; nz <= L; ++nz) {
            double const fz = sho_1d_norm(nz, sigma);
            for (int ny = 0; ny <= L - nz; ++ny) {
                double const fy = sho_1d_norm(ny, sigma);
                for (int nx = 0; nx <= L - nz - ny; ++nx) f[k++] = sho_1d_norm(nx, sigma) * fy * fz;
            }
        }
        double* M = (double*)calloc(2 * (size_t)nc * nc, sizeof(double));
        double const* hs = p->hs[ia];
        for (int i = 0; i < nc; ++i)
            for (int j = 0; j < nc; ++j) {
                double const h = f[i] * hs[i * nc + j] * f[j];
                double const o = f[i] * hs[nc * nc + i * nc + j] * f[j];
                M[i * nc + j] = (h - p->E_re * o) * p->dVol;
                M[nc * nc + i * nc + j] = (-p->E_im * o) * p->dVol;
            }
        s->M[a] = M;
        free(f);
    }
    return s;
}

void oracle_destroy(oracle_state* s) {
    if (!s) return;
    for (uint32_t a = 0; a < s->p.nAtoms; ++a) free(s->M[a]);
    free(s->M);
    free(s->V);
    free(s->image_phase);
    free(s);
}

/* ---- potential: include/green_potential.hxx:16-139 (no confinement) ---- */
static void potential(oracle_state const* s, double* y, double const* x) {
    oracle_plan const* p = &s->p;
    for (uint64_t b = 0; b < p->nnzb; ++b) {
        int32_t const iloc = p->veff_index[b];
        for (int reim = 0; reim < 2; ++reim) {
            double const Vim = p->E_im * (1 - 2 * reim);
            for (int i = 0; i < 64; ++i) {
                double const Vd = (iloc < 0) ? 0 : s->V[(size_t)iloc * 64 + i];
                double const Vt = Vd - p->E_re;
                for (int j = 0; j < 64; ++j) {
                    double v = Vt * x[((b * 2 + reim) * 64 + i) * 64 + j];
                    if (p->E_im != 0) v += Vim * x[((b * 2 + 1 - reim) * 64 + i) * 64 + j];
                    y[((b * 2 + reim) * 64 + i) * 64 + j] = v;
                }
            }
        }
    }
}

/* ---- kinetic: include/green_kinetic.hxx:24-329 and the list encoding of include/kinetic_plan.hxx:19-23.
 * One line of blocks per list: a halo entry 0 contributes zeros, a negative entry the periodic image
 * -entry-1 times the left (right) Bloch phase; the 9-point (range 4) or 17-point (range 8) stencil. ---- */
static void kinetic(oracle_state const* s, int dd, double* y, double const* x) {
    oracle_plan const* p = &s->p;
    static int const c8i[5] = {-14350, 8064, -1008, 128, -9};
    static int const c16i[9] = {-924708642, 538137600, -94174080, 22830080, -5350800, 1053696, -156800, 15360, -735};
    int const R = p->fd_range[dd];
    double c[9];
    double const norm = (R == 8) ? s->prefactor[dd] / 302702400. : s->prefactor[dd] / 5040.;
    for (int m = 0; m <= R; ++m) c[m] = (R == 8 ? c16i[m] : c8i[m]) * norm;
    int const stride = 1 << (2 * dd);
    int const nh = R / 4;   /* halo blocks per side */
    double* line = NULL;
    size_t cap = 0;
    for (uint32_t l = 0; l < p->nlists[dd]; ++l) {
        int32_t const* L = p->kin_list[dd] + p->kin_start[dd][l];
        int const len = (int)(p->kin_start[dd][l + 1] - p->kin_start[dd][l]);
        int const n = len - 8;     /* blocks in the line */
        size_t const npts = 4 * (size_t)(n + 2 * nh);
        if (npts > cap) { cap = npts; line = (double*)realloc(line, cap * sizeof(double)); }
        for (int i16 = 0; i16 < 16; ++i16) {
            /* base point of the pencil (include/green_kinetic.hxx:66-67) */
            int const i64 = (16 == stride) ? i16 : ((4 == stride) ? (16 * (i16 >> 2) + (i16 & 3)) : (4 * i16));
            for (int reim = 0; reim < 2; ++reim)
                for (int j = 0; j < 64; ++j) {
                    for (int q = 0; q < n + 2 * nh; ++q) {
                        int const pos = 4 - nh + q;   /* list position of this block */
                        int32_t const e = L[pos];
                        int const side = (q < nh) ? 0 : 1;
                        for (int i4 = 0; i4 < 4; ++i4) {
                            double v = 0;
                            int const ip = i64 + stride * i4;
                            if (e > 0) {
                                v = x[(((uint64_t)(e - 1) * 2 + reim) * 64 + ip) * 64 + j];
                            } else if (e < 0) {
                                uint64_t const jj = (uint64_t)(-e - 1);
                                double const pr = s->phase[dd][side][0];
                                double const pi = s->phase[dd][side][1] * (1. - 2 * reim);
                                v = pr * x[((jj * 2 + reim) * 64 + ip) * 64 + j]
                                  - pi * x[((jj * 2 + 1 - reim) * 64 + ip) * 64 + j];
                            }
                            line[4 * q + i4] = v;
                        }
                    }
                    for (int b = 0; b < n; ++b) {
                        uint64_t const ib = (uint64_t)(L[4 + b] - 1);
                        for (int i4 = 0; i4 < 4; ++i4) {
                            int const k = 4 * (b + nh) + i4;
                            double t = c[0] * line[k];
                            for (int m = 1; m <= R; ++m) t += c[m] * (line[k - m] + line[k + m]);
                            y[((ib * 2 + reim) * 64 + i64 + stride * i4) * 64 + j] += t;
                        }
                    }
                }
        }
    }
    free(line);
}

/* ---- Hermite-Gauss functions of one atom image on one block: include/green_dyadic.hxx:33-107 ---- */
static void hermite(double H[8][3][4], float xi2[3][4], int L, double const apos[4], float const cpos[3],
                    double const hg[3], double rproj) {
    double const R2 = rproj * rproj;
    for (int dir = 0; dir < 3; ++dir)
        for (int i4 = 0; i4 < 4; ++i4) {
            double const sig_inv = apos[3] * apos[3];
            double const xi = sig_inv * (((double)cpos[dir] * 4 + i4 + 0.5) * hg[dir] - apos[dir]);
            double const x2 = xi * xi;
            xi2[dir][i4] = (float)x2;
            double const H0 = (x2 < R2) ? exp(-0.5 * x2) : 0;
            H[0][dir][i4] = H0;
            double hn = H0, hm = 0;
            for (int nu = 0; nu < L; ++nu) {
                double const hp = xi * hn - (0.5 * nu) * hm;
                H[nu + 1][dir][i4] = hp;
                hm = hn; hn = hp;
            }
        }
}

/* ---- dyadic: SHOprj, SHOsum, SHOmul, SHOadd of include/green_dyadic.hxx:112-922 ---- */
static void dyadic(oracle_state const* s, double* y, double const* x) {
    oracle_plan const* p = &s->p;
    uint32_t const nai = p->nAtomImages, nrhs = p->nCols;
    if (nai * nrhs < 1) return;
    size_t const ncimg = p->image_starts[nai];
    size_t const ncat = p->atom_starts[p->nAtoms];
    float const R2f = (float)(p->r_proj * p->r_proj);
    double* Cpr = (double*)calloc(ncimg * nrhs * 128 + 1, sizeof(double));  /* [(a0+sho)*nrhs+irhs][2][64] */
    double H[8][3][4];
    float xi2[3][4];
    /* projection (SHOprj): per atom image and column, accumulated over the blocks of that column */
    size_t prj_off = 0;
    for (uint32_t irhs = 0; irhs < nrhs; ++irhs) {
        uint32_t const* rs = p->prj_start + (size_t)irhs * (nai + 1);
        for (uint32_t iai = 0; iai < nai; ++iai) {
            int const L = p->image_lmax[iai];
            int const nc = nsho(L);
            double* acc = (double*)calloc((size_t)nc * 128, sizeof(double));
            for (uint32_t k = rs[iai]; k < rs[iai + 1]; ++k) {
                uint64_t const b = p->prj_index[prj_off + k];
                float const* cp = p->row_pos + 4 * (size_t)p->rowindx[b];
                hermite(H, xi2, L, p->image_pos + 4 * (size_t)iai, cp, p->hg, p->r_proj);
                for (int reim = 0; reim < 2; ++reim)
                    for (int j = 0; j < 64; ++j)
                        for (int z = 0; z < 4; ++z) {
                            float const d2z = xi2[2][z] - R2f;
                            if (!(d2z < 0)) continue;
                            for (int yy = 0; yy < 4; ++yy) {
                                float const d2yz = xi2[1][yy] + d2z;
                                if (!(d2yz < 0)) continue;
                                for (int xx = 0; xx < 4; ++xx) {
                                    float const d2 = xi2[0][xx] + d2yz;
                                    if (!(d2 < 0)) continue;
                                    double const ps = x[((b * 2 + reim) * 64 + (z * 4 + yy) * 4 + xx) * 64 + j];
                                    int sho = 0;
                                    for (int iz = 0; iz <= L; ++iz)
                                        for (int iy = 0; iy <= L - iz; ++iy)
                                            for (int ix = 0; ix <= L - iz - iy; ++ix, ++sho)
                                                acc[(size_t)sho * 128 + reim * 64 + j] +=
                                                    ps * H[ix][0][xx] * H[iy][1][yy] * H[iz][2][z];
                                }
                            }
                        }
            }
            size_t const a0 = p->image_starts[iai];
            for (int sho = 0; sho < nc; ++sho)
                memcpy(Cpr + ((a0 + sho) * nrhs + irhs) * 128, acc + (size_t)sho * 128, 128 * sizeof(double));
            free(acc);
        }
        prj_off += rs[nai];
    }
    /* atom matrices: via images (SHOsum collect, SHOmul, SHOsum broadcast) or directly */
    double* Cad = (double*)calloc(ncimg * nrhs * 128 + 1, sizeof(double));
    if (nai > p->nAtoms) {
        double* aac = (double*)calloc(ncat * nrhs * 128 + 1, sizeof(double));
        for (uint32_t a = 0; a < p->nAtoms; ++a) {
            int const nc = nsho(p->atom_lmax[a]);
            size_t const a0 = p->atom_starts[a];
            for (uint32_t irhs = 0; irhs < nrhs; ++irhs) {
                for (uint32_t k = p->sum_start[a]; k < p->sum_start[a + 1]; ++k) {
                    uint32_t const iai = p->sum_index[k];
                    double const pr = s->image_phase[2 * iai], pi = s->image_phase[2 * iai + 1];
                    size_t const i0 = p->image_starts[iai];
                    for (int ai = 0; ai < nc; ++ai)
                        for (int j = 0; j < 64; ++j) {
                            double const* c = Cpr + ((i0 + ai) * nrhs + irhs) * 128;
                            double* t = aac + ((a0 + ai) * nrhs + irhs) * 128;
                            t[j] += pr * c[j] - pi * c[64 + j];
                            t[64 + j] += pr * c[64 + j] + pi * c[j];
                        }
                }
                double* d = (double*)calloc((size_t)nc * 128, sizeof(double));
                double const* M = s->M[a];
                for (int ai = 0; ai < nc; ++ai)
                    for (int j = 0; j < 64; ++j) {
                        double dr = 0, di = 0;
                        for (int aj = 0; aj < nc; ++aj) {
                            double const* c = aac + ((a0 + aj) * nrhs + irhs) * 128;
                            double const mr = M[ai * nc + aj], mi = M[nc * nc + ai * nc + aj];
                            dr += mr * c[j] - mi * c[64 + j];
                            di += mr * c[64 + j] + mi * c[j];
                        }
                        d[ai * 128 + j] = dr;
                        d[ai * 128 + 64 + j] = di;
                    }
                for (uint32_t k = p->sum_start[a]; k < p->sum_start[a + 1]; ++k) {
                    uint32_t const iai = p->sum_index[k];
                    double const pr = s->image_phase[2 * iai], pi = -s->image_phase[2 * iai + 1];
                    size_t const i0 = p->image_starts[iai];
                    for (int ai = 0; ai < nc; ++ai)
                        for (int j = 0; j < 64; ++j) {
                            double* t = Cad + ((i0 + ai) * nrhs + irhs) * 128;
                            double const dr = d[ai * 128 + j], di = d[ai * 128 + 64 + j];
                            t[j] = pr * dr - pi * di;
                            t[64 + j] = pr * di + pi * dr;
                        }
                }
                free(d);
            }
        }
        free(aac);
    } else {
        /* one image per atom: the reference multiplies image iatom's coefficients by atom iatom's matrix */
        for (uint32_t a = 0; a < p->nAtoms; ++a) {
            int const nc = nsho(p->atom_lmax[a]);
            size_t const a0 = p->atom_starts[a];
            double const* M = s->M[a];
            for (uint32_t irhs = 0; irhs < nrhs; ++irhs)
                for (int ai = 0; ai < nc; ++ai)
                    for (int j = 0; j < 64; ++j) {
                        double dr = 0, di = 0;
                        for (int aj = 0; aj < nc; ++aj) {
                            double const* c = Cpr + ((a0 + aj) * nrhs + irhs) * 128;
                            double const mr = M[ai * nc + aj], mi = M[nc * nc + ai * nc + aj];
                            dr += mr * c[j] - mi * c[64 + j];
                            di += mr * c[64 + j] + mi * c[j];
                        }
                        Cad[((a0 + ai) * nrhs + irhs) * 128 + j] = dr;
                        Cad[((a0 + ai) * nrhs + irhs) * 128 + 64 + j] = di;
                    }
        }
    }
    /* add-back (SHOadd) per block over its atom images, masked */
    for (uint64_t b = 0; b < p->nnzb; ++b) {
        uint32_t const irhs = p->colindx[b];
        float const* cp = p->row_pos + 4 * (size_t)p->rowindx[b];
        for (uint32_t k = p->add_start[b]; k < p->add_start[b + 1]; ++k) {
            uint32_t const iai = p->add_index[k];
            int const L = p->image_lmax[iai];
            size_t const a0 = p->image_starts[iai];
            hermite(H, xi2, L, p->image_pos + 4 * (size_t)iai, cp, p->hg, p->r_proj);
            for (int reim = 0; reim < 2; ++reim)
                for (int j = 0; j < 64; ++j)
                    for (int z = 0; z < 4; ++z) {
                        float const d2z = xi2[2][z] - R2f;
                        if (!(d2z < 0)) continue;
                        for (int yy = 0; yy < 4; ++yy) {
                            float const d2yz = xi2[1][yy] + d2z;
                            if (!(d2yz < 0)) continue;
                            for (int xx = 0; xx < 4; ++xx) {
                                float const d2 = xi2[0][xx] + d2yz;
                                if (!(d2 < 0)) continue;
                                double t = 0;
                                int sho = 0;
                                for (int iz = 0; iz <= L; ++iz)
                                    for (int iy = 0; iy <= L - iz; ++iy)
                                        for (int ix = 0; ix <= L - iz - iy; ++ix, ++sho)
                                            t += H[ix][0][xx] * H[iy][1][yy] * H[iz][2][z] *
                                                 Cad[((a0 + sho) * nrhs + irhs) * 128 + reim * 64 + j];
                                y[((b * 2 + reim) * 64 + (z * 4 + yy) * 4 + xx) * 64 + j] += t;
                            }
                        }
                    }
        }
    }
    free(Cpr);
    free(Cad);
}

/* the action in the reference's order (include/green_action.hxx:159-171) */
void oracle_action(oracle_state* s, int which, double* y, const double* x) {
    if (which < 0 || which == 0) potential(s, y, x);
    for (int dd = 0; dd < 3; ++dd)
        if (which < 0 || which == dd + 1) kinetic(s, dd, y, x);
    if (which < 0 || which == 4) dyadic(s, y, x);
}

/* shadow residual r0*: murmur3's 32-bit finaliser (Appleby, MurmurHash3) of a mix of seed and element
 * index; the top 24 bits as a signed fraction in [-1, 1), exact in FP32 and FP64 */
double oracle_shadow(uint64_t seed, uint64_t idx) {
    uint32_t h = (uint32_t)idx + 0x9E3779B9u * ((uint32_t)(idx >> 32) + 1u) + (uint32_t)seed * 0x27D4EB2Fu +
                 (uint32_t)(seed >> 32);
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return (double)((int32_t)h >> 8) * (1.0 / 8388608.0);
}

/* ---- TFQMR, Saad Algorithm 7.8, per column; the same scalar formulas as the device solver ----
 * The inner products sum over blocks in ascending order, or descending after oracle_set_sum_order(1): the same
 * algorithm with other rounding, which measures how far rounding alone separates two runs after m half-steps. */
static int g_reverse_sums = 0;
void oracle_set_sum_order(int reverse) { g_reverse_sums = reverse != 0; }
#define SUM_BLK(k) (g_reverse_sums ? p->nnzb - 1 - (k) : (k))
typedef struct { double re, im; } cplx;
static cplx cdiv(cplx a, cplx b) {
    double const den = b.re * b.re + b.im * b.im;
    cplx r = {(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
    return r;
}

int oracle_tfqmr(oracle_state* s, const uint32_t* subset, double tol, int maxit, uint64_t seed, double* X,
                 double* resid) {
    oracle_plan const* p = &s->p;
    uint64_t const n = p->nnzb * NBLK;
    uint32_t const ncol = 64 * p->nCols;
    double *w = calloc(n, 8), *u0 = calloc(n, 8), *u1 = calloc(n, 8), *v = calloc(n, 8), *d = calloc(n, 8);
    double *Aodd = calloc(n, 8), *Aeven = calloc(n, 8), *b = calloc(n, 8);
    cplx* rho = calloc(ncol, sizeof(cplx));
    cplx* alpha = calloc(ncol, sizeof(cplx));
    cplx* eta = calloc(ncol, sizeof(cplx));
    cplx* beta = calloc(ncol, sizeof(cplx));
    cplx* cd = calloc(ncol, sizeof(cplx));
    double* theta = calloc(ncol, 8);
    double* tau = calloc(ncol, 8);
    double* acc = calloc(3 * (size_t)ncol, 8);
    int* status = calloc(ncol, sizeof(int));
    memset(X, 0, n * 8);
    for (uint32_t c = 0; c < p->nCols; ++c)
        for (int i = 0; i < 64; ++i) b[((uint64_t)subset[c] * 2 * 64 + i) * 64 + i] = 1.0;
    memcpy(w, b, n * 8);
    memcpy(u0, b, n * 8);
#define COL(blk, j) ((uint32_t)p->colindx[blk] * 64 + (j))
#define RE(a, blk, i, j) a[(((blk) * 2 + 0) * 64 + (i)) * 64 + (j)]
#define IM(a, blk, i, j) a[(((blk) * 2 + 1) * 64 + (i)) * 64 + (j)]
#define SR(blk, i, j) oracle_shadow(seed, (((blk) * 2 + 0) * 64 + (i)) * 64 + (j))
#define SI(blk, i, j) oracle_shadow(seed, (((blk) * 2 + 1) * 64 + (i)) * 64 + (j))
    /* rho = dot(rs, b), tau = ||b|| */
    memset(acc, 0, 3 * (size_t)ncol * 8);
    for (uint64_t kb = 0; kb < p->nnzb; ++kb) {
        uint64_t const blk = SUM_BLK(kb);
        for (int i = 0; i < 64; ++i)
            for (int j = 0; j < 64; ++j) {
                uint32_t const q = COL(blk, j);
                double const br = RE(b, blk, i, j), bi = IM(b, blk, i, j), sr = SR(blk, i, j), si = SI(blk, i, j);
                acc[q] += br * br + bi * bi;
                acc[ncol + q] += sr * br + si * bi;
                acc[2 * ncol + q] += sr * bi - si * br;
            }
    }
    for (uint32_t q = 0; q < ncol; ++q) {
        tau[q] = sqrt(acc[q]);
        rho[q].re = acc[ncol + q]; rho[q].im = acc[2 * ncol + q];
        status[q] = (tau[q] <= tol * tau[q]) ? 1 : 0;
    }
    double* bn = calloc(ncol, 8);
    for (uint32_t q = 0; q < ncol; ++q) bn[q] = tau[q];
    memset(v, 0, n * 8);
    oracle_action(s, -1, v, u0);
    double* Au_even = v;
    double* u = u0;
    double* un = u1;
    int m;
    for (m = 0; m < maxit; ++m) {
        int active = 0;
        for (uint32_t q = 0; q < ncol; ++q) active += (status[q] == 0);
        if (!active) break;
        if ((m & 1) == 0) {
            memset(acc, 0, 2 * (size_t)ncol * 8);
            for (uint64_t kb = 0; kb < p->nnzb; ++kb) {
                uint64_t const blk = SUM_BLK(kb);
                for (int i = 0; i < 64; ++i)
                    for (int j = 0; j < 64; ++j) {
                        uint32_t const q = COL(blk, j);
                        double const vr = RE(v, blk, i, j), vi = IM(v, blk, i, j), sr = SR(blk, i, j), si = SI(blk, i, j);
                        acc[q] += sr * vr + si * vi;
                        acc[ncol + q] += sr * vi - si * vr;
                    }
            }
            for (uint32_t q = 0; q < ncol; ++q) {
                if (status[q]) { alpha[q].re = alpha[q].im = 0; continue; }
                cplx sg = {acc[q], acc[ncol + q]};
                if (!(sg.re * sg.re + sg.im * sg.im > 0)) { status[q] = 2; alpha[q].re = alpha[q].im = 0; continue; }
                alpha[q] = cdiv(rho[q], sg);
            }
            memset(acc, 0, (size_t)ncol * 8);
            for (uint64_t kb = 0; kb < p->nnzb; ++kb) {
                uint64_t const blk = SUM_BLK(kb);
                for (int i = 0; i < 64; ++i)
                    for (int j = 0; j < 64; ++j) {
                        cplx const a = alpha[COL(blk, j)];
                        double const ur = RE(u, blk, i, j), ui = IM(u, blk, i, j), vr = RE(v, blk, i, j), vi = IM(v, blk, i, j);
                        RE(un, blk, i, j) = ur - (a.re * vr - a.im * vi);
                        IM(un, blk, i, j) = ui - (a.re * vi + a.im * vr);
                        double const yr = RE(Au_even, blk, i, j), yi = IM(Au_even, blk, i, j);
                        double const wr = RE(w, blk, i, j) - (a.re * yr - a.im * yi);
                        double const wi = IM(w, blk, i, j) - (a.re * yi + a.im * yr);
                        RE(w, blk, i, j) = wr; IM(w, blk, i, j) = wi;
                        acc[COL(blk, j)] += wr * wr + wi * wi;
                    }
            }
        } else {
            memset(Aodd, 0, n * 8);
            oracle_action(s, -1, Aodd, u);
            memset(acc, 0, 3 * (size_t)ncol * 8);
            for (uint64_t kb = 0; kb < p->nnzb; ++kb) {
                uint64_t const blk = SUM_BLK(kb);
                for (int i = 0; i < 64; ++i)
                    for (int j = 0; j < 64; ++j) {
                        uint32_t const q = COL(blk, j);
                        cplx const a = alpha[q];
                        double const yr = RE(Aodd, blk, i, j), yi = IM(Aodd, blk, i, j);
                        double const wr = RE(w, blk, i, j) - (a.re * yr - a.im * yi);
                        double const wi = IM(w, blk, i, j) - (a.re * yi + a.im * yr);
                        RE(w, blk, i, j) = wr; IM(w, blk, i, j) = wi;
                        double const sr = SR(blk, i, j), si = SI(blk, i, j);
                        acc[q] += wr * wr + wi * wi;
                        acc[ncol + q] += sr * wr + si * wi;
                        acc[2 * ncol + q] += sr * wi - si * wr;
                    }
            }
        }
        /* theta, c, tau, eta, d coefficient, convergence; rho and beta on odd steps */
        for (uint32_t q = 0; q < ncol; ++q) {
            if (status[q]) { eta[q].re = eta[q].im = 0; cd[q].re = cd[q].im = 0; beta[q].re = beta[q].im = 0; continue; }
            cplx const a = alpha[q];
            double const a2 = a.re * a.re + a.im * a.im;
            cplx c_ = {0, 0};
            if (a2 > 0) {
                double const tr = theta[q] * theta[q] * eta[q].re, ti = theta[q] * theta[q] * eta[q].im;
                c_.re = (tr * a.re + ti * a.im) / a2;
                c_.im = (ti * a.re - tr * a.im) / a2;
            }
            double const thn = sqrt(acc[q]) / tau[q];
            double const c = 1.0 / sqrt(1.0 + thn * thn);
            double const taun = tau[q] * thn * c;
            cd[q] = c_;
            theta[q] = thn; tau[q] = taun;
            eta[q].re = c * c * a.re; eta[q].im = c * c * a.im;
            if (taun * sqrt((double)(m + 2)) <= tol * bn[q]) status[q] = 1;
            if (m & 1) {
                cplx const rn = {acc[ncol + q], acc[2 * ncol + q]};
                double const den = rho[q].re * rho[q].re + rho[q].im * rho[q].im;
                if (!(den > 0) || !(rn.re * rn.re + rn.im * rn.im > 0)) {
                    if (!status[q]) status[q] = 2;
                    beta[q].re = beta[q].im = 0;
                } else {
                    beta[q] = cdiv(rn, rho[q]);
                }
                rho[q] = rn;
            }
        }
        /* d = u + cd d; x += eta d */
        for (uint64_t blk = 0; blk < p->nnzb; ++blk)
            for (int i = 0; i < 64; ++i)
                for (int j = 0; j < 64; ++j) {
                    uint32_t const q = COL(blk, j);
                    cplx const c = cd[q], e = eta[q];
                    double const dr0 = RE(d, blk, i, j), di0 = IM(d, blk, i, j);
                    double const dr = RE(u, blk, i, j) + (c.re * dr0 - c.im * di0);
                    double const di = IM(u, blk, i, j) + (c.re * di0 + c.im * dr0);
                    RE(d, blk, i, j) = dr; IM(d, blk, i, j) = di;
                    RE(X, blk, i, j) += e.re * dr - e.im * di;
                    IM(X, blk, i, j) += e.re * di + e.im * dr;
                }
        if ((m & 1) == 0) {
            double* t = u; u = un; un = t;
        } else {
            for (uint64_t blk = 0; blk < p->nnzb; ++blk)
                for (int i = 0; i < 64; ++i)
                    for (int j = 0; j < 64; ++j) {
                        cplx const bb = beta[COL(blk, j)];
                        double const ur = RE(u, blk, i, j), ui = IM(u, blk, i, j);
                        RE(un, blk, i, j) = RE(w, blk, i, j) + (bb.re * ur - bb.im * ui);
                        IM(un, blk, i, j) = IM(w, blk, i, j) + (bb.re * ui + bb.im * ur);
                    }
            memset(Aeven, 0, n * 8);
            oracle_action(s, -1, Aeven, un);
            for (uint64_t blk = 0; blk < p->nnzb; ++blk)
                for (int i = 0; i < 64; ++i)
                    for (int j = 0; j < 64; ++j) {
                        cplx const bb = beta[COL(blk, j)];
                        double const vr = RE(v, blk, i, j), vi = IM(v, blk, i, j);
                        double const tr = RE(Aodd, blk, i, j) + (bb.re * vr - bb.im * vi);
                        double const ti = IM(Aodd, blk, i, j) + (bb.re * vi + bb.im * vr);
                        RE(v, blk, i, j) = RE(Aeven, blk, i, j) + (bb.re * tr - bb.im * ti);
                        IM(v, blk, i, j) = IM(Aeven, blk, i, j) + (bb.re * ti + bb.im * tr);
                    }
            Au_even = Aeven;
            double* t = u; u = un; un = t;
        }
    }
    /* true residual per column */
    if (resid) {
        memset(Aodd, 0, n * 8);
        oracle_action(s, -1, Aodd, X);
        memset(acc, 0, (size_t)ncol * 8);
        for (uint64_t blk = 0; blk < p->nnzb; ++blk)
            for (int i = 0; i < 64; ++i)
                for (int j = 0; j < 64; ++j) {
                    double const rr = RE(b, blk, i, j) - RE(Aodd, blk, i, j), ri = IM(b, blk, i, j) - IM(Aodd, blk, i, j);
                    acc[COL(blk, j)] += rr * rr + ri * ri;
                }
        for (uint32_t q = 0; q < ncol; ++q) resid[q] = sqrt(acc[q]) / bn[q];
    }
    free(w); free(u0); free(u1); free(v); free(d); free(Aodd); free(Aeven); free(b);
    free(rho); free(alpha); free(eta); free(beta); free(cd); free(theta); free(tau); free(acc); free(status); free(bn);
    return m;
}
