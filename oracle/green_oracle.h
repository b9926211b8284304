/*
 * green_oracle.h -- CPU restatement of the reference's Green-function action and of the TFQMR solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker of the HIP implementation: only tests/,
 * __graft_entry__.smoke() and the CPU-baseline leg of bench.py load it.  The product never calls it.
 *
 * The oracle consumes the plan's index structures in the reference's own formats (BSR, finite-difference
 * lists with the CUBE_* encoding, sparse_SHOprj/SHOadd/SHOsum) and computes everything that depends on
 * values itself: the cell blocking of the potential, the SHO prefactor rescaling and (h - E s) dVol of
 * the atom matrices, the Bloch phases.
 */
#ifndef GREEN_ORACLE_H
#define GREEN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    /* geometry */
    uint32_t nb[3];                 /* blocks per direction                          */
    double hg[3];                   /* grid spacing                                  */
    double r_proj;                  /* projection radius in units of sigma           */
    /* BSR */
    uint64_t nnzb;
    uint32_t nRows, nCols;
    const uint16_t* colindx;        /* [nnzb]   */
    const uint32_t* rowindx;        /* [nnzb]   */
    const int32_t* veff_index;      /* [nnzb]   */
    const float* row_pos;           /* [nRows][4] */
    /* finite-difference lists per direction */
    uint32_t nlists[3];
    const uint32_t* kin_start[3];
    const int32_t* kin_list[3];
    int32_t fd_range[3];
    double scale_T;
    /* atoms */
    uint32_t nAtoms, nAtomImages;
    const double* image_pos;        /* [nai][4]  */
    const int8_t* image_lmax;       /* [nai]     */
    const uint32_t* image_starts;   /* [nai+1]   */
    const int8_t* image_shift;      /* [nai][4]  */
    const uint32_t* add_start;      /* [nnzb+1]  */
    const uint32_t* add_index;
    const uint32_t* prj_start;      /* [nCols][nai+1] */
    const uint32_t* prj_index;      /* concatenated over columns */
    const uint32_t* sum_start;      /* [nAtoms+1] */
    const uint32_t* sum_index;
    const uint32_t* atom_starts;    /* [nAtoms+1] */
    const int8_t* atom_lmax;        /* [nAtoms] */
    const int32_t* atom_orig;       /* [nAtoms] original atom (row of xyzZinso) of each contributing atom */
    /* values */
    const double* V_grid;           /* [4 nb_z][4 nb_y][4 nb_x] */
    const double* const* hs;        /* per original atom: [2][nc][nc], L2-normalised basis */
    const double* xyzZinso;         /* [natoms*8] */
    double E_re, E_im, dVol;
    double k[3];                    /* k-point in units of 2 pi */
} oracle_plan;

typedef struct oracle_state oracle_state;

/* prepare the value-dependent data (potential blocks, matrices, phases); returns NULL on failure */
oracle_state* oracle_create(const oracle_plan* p);
void oracle_destroy(oracle_state* s);

/* y = (H - E S) x, FP64, x and y [nnzb][2][64][64]; which: -1 all, 0 potential, 1..3 kinetic x,y,z, 4 dyadic
 * (parts other than the potential add to y) */
void oracle_action(oracle_state* s, int which, double* y, const double* x);

/* the seeded shadow residual element used by the TFQMR */
double oracle_shadow(uint64_t seed, uint64_t idx);

/* TFQMR (same algorithm, scalars and shadow as the device solver) for B = unit columns on the
 * diagonal blocks; X [nnzb][2][64][64]; returns the number of half-steps; resid[64 nCols] receives the
 * true relative residual per column. */
void oracle_set_sum_order(int reverse);
int oracle_tfqmr(oracle_state* s, const uint32_t* subset, double tol, int maxit, uint64_t seed, double* X,
                 double* resid);

#ifdef __cplusplus
}
#endif

#endif
