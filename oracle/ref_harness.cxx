// Test infrastructure only: a C-ABI harness around the reference's own CPU implementation of the
// Green-function action, compiled from the reference sources where they lie (see oracle/Makefile).
// Nothing in the product links this file; tests and the CPU baseline load oracle/_ref/libref_green.so.
//
// The reference builds its CPU path from the same __global__ kernel bodies, run as host loops
// over emulated grids (include/green_cuda.hxx:19-45). This harness drives:
//   - green_function::construct_Green_function   (src/green_function.cxx:357-1068)
//   - green_function::update_energy_parameter    (src/green_function.cxx:64-109)
//   - green_function::update_phases              (src/green_function.cxx:218-251)
//   - green_action::action_t<real_t,2,1>::multiply (include/green_action.hxx:140-176)
//   - green_dyadic::get_projection_coefficients  (include/green_dyadic.hxx:925-1091)
//   - energy_contour::Gauss_Legendre_quadrature, ::Gauss_Fermi_Dirac_quadrature, ::get_energy_mesh
//     (src/energy_contour.cxx:139-253, :456, :2156)
// Inputs are injected directly into the plan instead of the MPI exchange (serial case):
//   - the local potential is stored cell-indexed, which is how the kernels read it
//     (include/green_potential.hxx:90-91 with veff_index from src/green_function.cxx:913-916);
//   - atom matrices are written into dyadic_plan.AtomMatrices_, the buffer that
//     green_parallel::exchange fills (src/green_function.cxx:192-208).

#include <cstdio>
#include <cstdint>
#include <cstring>
#include <cmath>
#include <vector>
#include <complex>
#include <algorithm>

#include "control.hxx"
#include "green_function.hxx"
#include "green_action.hxx"
#include "green_kinetic.hxx"
#include "green_dyadic.hxx"
#include "sho_tools.hxx"
#include "sho_projection.hxx"
#include "boundary_condition.hxx"
#include "energy_contour.hxx"

namespace energy_contour {   // defined in src/energy_contour.cxx without a header declaration
std::vector<std::complex<double>> get_energy_mesh(std::vector<std::complex<double>>& w8, double kBT, double eBot,
                                                  unsigned nBot, unsigned nPar, unsigned nFer, int nPol, int echo);
}

namespace {

struct RefPlan {
    action_plan_t p;
    uint32_t nb[3];
    int natoms;
    std::vector<double> xyzZinso;
    green_action::action_t<double,2,1,64>* act64 = nullptr;
    green_action::action_t<float ,2,1,64>* act32 = nullptr;
    uint16_t* colIndex = nullptr; // managed copy of p.colindx
};

void set_controls(double r_trunc, int const fd_range[3], int const nsrc[3], double r_proj, int echo) {
    int const e = control::echo_set_without_warning;
    control::set("green_function.truncation.radius", r_trunc, e);
    control::set("green_function.projection.radius", r_proj, e);
    control::set("green_kinetic.range.x", double(fd_range[0]), e);
    control::set("green_kinetic.range.y", double(fd_range[1]), e);
    control::set("green_kinetic.range.z", double(fd_range[2]), e);
    control::set("green_function.sources.x", double(nsrc[0]), e);
    control::set("green_function.sources.y", double(nsrc[1]), e);
    control::set("green_function.sources.z", double(nsrc[2]), e);
    control::set("green_function.potential.exchange", 0., e);
    control::set("green_function.matrices.exchange", 0., e);
    (void)echo;
}

} // namespace

extern "C" {

// Create a plan. xyzZinso[natoms*8] follows src/green_input.cxx:89-122 (x,y,z,Z,id,numax,sigma,_).
void* ref_plan_create(uint32_t const ng[3], int8_t const bc[3], double const hg[3],
                      double const* xyzZinso, int natoms, double r_trunc, int const fd_range[3],
                      int const nsrc[3], double r_proj, int echo) {
    set_controls(r_trunc, fd_range, nsrc, r_proj, echo);
    auto* rp = new RefPlan();
    rp->natoms = natoms;
    rp->xyzZinso.assign(xyzZinso, xyzZinso + 8*size_t(natoms));
    for (int d = 0; d < 3; ++d) rp->nb[d] = ng[d] >> 2;
    int8_t bcc[3] = {bc[0], bc[1], bc[2]};
    double hgc[3] = {hg[0], hg[1], hg[2]};
    uint32_t ngc[3] = {ng[0], ng[1], ng[2]};
    auto const stat = green_function::construct_Green_function(rp->p, ngc, bcc, hgc, rp->xyzZinso, echo, 1);
    if (stat) { delete rp; return nullptr; }
    rp->p.echo = 0;  // the plan's own log level defaults to 9 (include/action_plan.hxx:249)
    // replace the row-indexed potential by a cell-indexed one (read by the kernels via veff_index)
    size_t const ncell = size_t(rp->nb[0])*rp->nb[1]*rp->nb[2];
    free_memory(rp->p.Veff[0]);
    rp->p.Veff[0] = get_memory<double[64]>(ncell, 0, "Veff_cell");
    std::memset(rp->p.Veff[0], 0, ncell*64*sizeof(double));
    auto const nnzb = rp->p.colindx.size();
    rp->colIndex = get_memory<uint16_t>(std::max<size_t>(nnzb, 1), 0, "colIndex");
    for (size_t i = 0; i < nnzb; ++i) rp->colIndex[i] = rp->p.colindx[i];
    rp->act64 = new green_action::action_t<double,2,1,64>(&rp->p, 0, 0);
    rp->act32 = new green_action::action_t<float ,2,1,64>(&rp->p, 0, 0);
    return rp;
}

void ref_plan_destroy(void* h) {
    auto* rp = (RefPlan*)h;
    if (!rp) return;
    delete rp->act64;
    delete rp->act32;
    free_memory(rp->colIndex);
    delete rp;
}

// sizes: [0]nRows [1]nCols [2]nnzb [3]nAtoms [4]nAtomImages [5]natomcoeffs(images) [6]ncoeffs(atoms)
//        [7..9] nLists x,y,z [10..12] list entries x,y,z [13..15] FD range x,y,z
//        [16] SHOadd nnz [17] SHOprj nnz (sum over rhs) [18] SHOsum nnz
void ref_plan_sizes(void* h, int64_t out[24]) {
    auto& p = ((RefPlan*)h)->p;
    auto& dp = p.dyadic_plan;
    for (int i = 0; i < 24; ++i) out[i] = 0;
    out[0] = p.nRows; out[1] = p.nCols; out[2] = p.colindx.size();
    out[3] = dp.nAtoms; out[4] = dp.nAtomImages;
    out[5] = dp.AtomImageStarts ? dp.AtomImageStarts[dp.nAtomImages] : 0;
    out[6] = dp.AtomStarts ? dp.AtomStarts[dp.nAtoms] : 0;
    for (int dd = 0; dd < 3; ++dd) {
        out[7 + dd] = p.kinetic[dd].sparse_.nRows();
        out[10 + dd] = p.kinetic[dd].sparse_.nNonzeros();
        out[13 + dd] = p.kinetic[dd].FD_range_;
    }
    out[16] = dp.sparse_SHOadd.nNonzeros();
    int64_t nprj = 0;
    for (int r = 0; r < dp.nrhs; ++r) nprj += dp.sparse_SHOprj[r].nNonzeros();
    out[17] = nprj;
    out[18] = dp.sparse_SHOsum.nNonzeros();
}

// BSR structure of X
void ref_plan_bsr(void* h, uint32_t* RowStart, uint16_t* colindx, uint32_t* subset, int32_t* veff_index,
                  float* rowCubePos, float* colCubePos, int64_t* global_source_indices) {
    auto& p = ((RefPlan*)h)->p;
    auto const nnzb = p.colindx.size();
    std::memcpy(RowStart, p.RowStart, (p.nRows + 1)*sizeof(uint32_t));
    std::memcpy(colindx, p.colindx.data(), nnzb*sizeof(uint16_t));
    std::memcpy(subset, p.subset.data(), p.nCols*sizeof(uint32_t));
    std::memcpy(veff_index, p.veff_index, nnzb*sizeof(int32_t));
    std::memcpy(rowCubePos, p.rowCubePos, p.nRows*4*sizeof(float));
    std::memcpy(colCubePos, p.colCubePos, p.nCols*4*sizeof(float));
    std::memcpy(global_source_indices, p.global_source_indices.data(), p.nCols*sizeof(int64_t));
}

// finite-difference lists of one direction as CSR
void ref_plan_kinetic(void* h, int dd, uint32_t* rowStart, int32_t* lists, double* prefactor) {
    auto& kp = ((RefPlan*)h)->p.kinetic[dd];
    auto const n = kp.sparse_.nRows();
    std::memcpy(rowStart, kp.sparse_.rowStart(), (n + 1)*sizeof(uint32_t));
    std::memcpy(lists, kp.sparse_.colIndex(), kp.sparse_.nNonzeros()*sizeof(int32_t));
    *prefactor = kp.prefactor_;
}

// dyadic plan arrays
void ref_plan_dyadic(void* h, double* AtomImagePos, int8_t* AtomImageLmax, uint32_t* AtomImageStarts,
                     uint32_t* AtomImageIndex, int8_t* AtomImageShift, uint32_t* AtomStarts, int8_t* AtomLmax,
                     int64_t* global_atom_ids,
                     uint32_t* add_rowStart, uint32_t* add_colIndex,
                     uint32_t* prj_rowStart /*[nrhs][nai+1]*/, uint32_t* prj_colIndex /*concatenated*/,
                     uint32_t* sum_rowStart, uint32_t* sum_colIndex) {
    auto& p = ((RefPlan*)h)->p;
    auto& dp = p.dyadic_plan;
    auto const nai = dp.nAtomImages, na = dp.nAtoms;
    std::memcpy(AtomImagePos, dp.AtomImagePos, nai*4*sizeof(double));
    std::memcpy(AtomImageLmax, dp.AtomImageLmax, nai);
    std::memcpy(AtomImageStarts, dp.AtomImageStarts, (nai + 1)*sizeof(uint32_t));
    std::memcpy(AtomImageIndex, dp.AtomImageIndex, nai*sizeof(uint32_t));
    std::memcpy(AtomImageShift, dp.AtomImageShift, nai*4);
    std::memcpy(AtomStarts, dp.AtomStarts, (na + 1)*sizeof(uint32_t));
    std::memcpy(AtomLmax, dp.AtomLmax, na);
    for (uint32_t i = 0; i < na; ++i) global_atom_ids[i] = dp.global_atom_ids[i];
    auto const nnzb = p.colindx.size();
    std::memcpy(add_rowStart, dp.sparse_SHOadd.rowStart(), (nnzb + 1)*sizeof(uint32_t));
    std::memcpy(add_colIndex, dp.sparse_SHOadd.colIndex(), dp.sparse_SHOadd.nNonzeros()*sizeof(uint32_t));
    size_t off = 0;
    for (int r = 0; r < dp.nrhs; ++r) {
        auto const& s = dp.sparse_SHOprj[r];
        std::memcpy(prj_rowStart + size_t(r)*(nai + 1), s.rowStart(), (nai + 1)*sizeof(uint32_t));
        std::memcpy(prj_colIndex + off, s.colIndex(), s.nNonzeros()*sizeof(uint32_t));
        off += s.nNonzeros();
    }
    if (dp.sparse_SHOsum.rowStart()) {
        std::memcpy(sum_rowStart, dp.sparse_SHOsum.rowStart(), (na + 1)*sizeof(uint32_t));
        std::memcpy(sum_colIndex, dp.sparse_SHOsum.colIndex(), dp.sparse_SHOsum.nNonzeros()*sizeof(uint32_t));
    }
}

// cell potential on the full grid, V[ng_z][ng_y][ng_x], stored into cell blocks [iloc][64]
void ref_set_potential(void* h, double const* V) {
    auto* rp = (RefPlan*)h;
    auto const* nb = rp->nb;
    size_t const ngx = 4*nb[0], ngy = 4*nb[1];
    for (uint32_t bz = 0; bz < nb[2]; ++bz)
    for (uint32_t by = 0; by < nb[1]; ++by)
    for (uint32_t bx = 0; bx < nb[0]; ++bx) {
        size_t const iloc = (size_t(bz)*nb[1] + by)*nb[0] + bx;
        for (int z = 0; z < 4; ++z) for (int y = 0; y < 4; ++y) for (int x = 0; x < 4; ++x) {
            size_t const g = ((bz*4 + z)*ngy + (by*4 + y))*ngx + (bx*4 + x);
            rp->p.Veff[0][iloc][(z*4 + y)*4 + x] = V[g];
        }
    }
}

// atom matrices: hs[natoms][2][nc][nc] (h then s, zyx order, L2-normalised SHO basis), indexed by
// original atom; rescaled for unnormalized Hermite functions as src/energy_contour.cxx:307-323 does.
void ref_set_atom_matrices(void* h, double const* const* hs) {
    auto* rp = (RefPlan*)h;
    auto& dp = rp->p.dyadic_plan;
    for (uint32_t iac = 0; iac < dp.nAtoms; ++iac) {
        auto const ia = dp.global_atom_ids[iac];
        int const numax = int(rp->xyzZinso[ia*8 + 5]);
        double const sigma = rp->xyzZinso[ia*8 + 6];
        int const nc = sho_tools::nSHO(numax);
        auto const f = sho_projection::get_sho_prefactors(numax, sigma);
        double* out = dp.AtomMatrices_[iac];
        for (int m = 0; m < 2; ++m)
            for (int i = 0; i < nc; ++i)
                for (int j = 0; j < nc; ++j)
                    out[(m*nc + i)*nc + j] = f[i]*hs[ia][(m*nc + i)*nc + j]*f[j];
    }
}

void ref_set_energy(void* h, double re, double im, double dVol) {
    auto* rp = (RefPlan*)h;
    green_function::update_energy_parameter(rp->p, std::complex<double>(re, im), dVol, 0, 1);
}

void ref_set_kpoint(void* h, double const k[3]) {
    auto* rp = (RefPlan*)h;
    green_function::update_phases(rp->p, k, 0, 1);
}

// y = A x with A = H - E S, FP64 complex, x,y [nnzb][2][64][64]; returns the reference flop count
double ref_action64(void* h, double* y, double const* x) {
    auto* rp = (RefPlan*)h;
    auto const nnzb = uint32_t(rp->p.colindx.size());
    return rp->act64->multiply((double(*)[2][64][64])y, (double const(*)[2][64][64])x,
                               rp->colIndex, nnzb, rp->p.nCols);
}

double ref_action32(void* h, float* y, float const* x) {
    auto* rp = (RefPlan*)h;
    auto const nnzb = uint32_t(rp->p.colindx.size());
    return rp->act32->multiply((float(*)[2][64][64])y, (float const(*)[2][64][64])x,
                               rp->colIndex, nnzb, rp->p.nCols);
}

// single pieces of the action, for localising a mismatch: which=0 potential, 1..3 kinetic x,y,z, 4 dyadic
void ref_action_part64(void* h, int which, double* y, double const* x) {
    auto* rp = (RefPlan*)h;
    auto& p = rp->p;
    auto const nnzb = uint32_t(p.colindx.size());
    auto Y = (double(*)[2][64][64])y;
    auto X = (double const(*)[2][64][64])x;
    if (0 == which) {
        green_potential::multiply<double,2,1>(Y, X, p.Veff, p.veff_index, p.target_minus_source,
            p.grid_spacing_trunc, nnzb, p.E_param, p.V_confinement, pow2(p.r_confinement), 0);
    } else if (which <= 3) {
        green_kinetic::multiply<double,2,1>(Y, X, p.kinetic[which - 1], p.phase[which - 1], 0);
    } else {
        auto const& dp = p.dyadic_plan;
        size_t const n = size_t(dp.AtomImageStarts ? dp.AtomImageStarts[dp.nAtomImages] : 0)*p.nCols;
        auto apc = get_memory<double[2][1][64]>(std::max<size_t>(n, 1), 0, "apc");
        green_dyadic::multiply<double,2,1>(Y, apc, X, dp, p.rowindx, rp->colIndex, p.rowCubePos, nnzb, 0);
        free_memory(apc);
    }
}

// the reference's module self-tests (status 0 = pass)
int ref_selftest(int which, int echo) {
    switch (which) {
        case 0: return int(green_kinetic::all_tests(echo));
        case 1: return int(green_potential::all_tests(echo));
        case 2: return int(green_dyadic::all_tests(echo));
        case 3: return int(green_sparse::all_tests(echo));
        case 4: return int(green_function::all_tests(echo));
    }
    return -1;
}

// the reference's quadratures: number of points (or the reference's error return) and the tables
int ref_gauss_legendre(unsigned n, double* x, double* w) { return energy_contour::Gauss_Legendre_quadrature(x, w, n, 0); }
int ref_gauss_fermi_dirac(unsigned n, double* x, double* w) { return energy_contour::Gauss_Fermi_Dirac_quadrature(x, w, n, 0); }

// the reference's energy mesh: E[n][2], w[n][2] (capacity entries), returns the number of points
int ref_energy_mesh(double kBT, double eBot, unsigned nBot, unsigned nPar, unsigned nFer, int nPol, double* E,
                    double* w, int capacity) {
    std::vector<std::complex<double>> w8;
    auto const Ep = energy_contour::get_energy_mesh(w8, kBT, eBot, nBot, nPar, nFer, nPol, 0);
    int const n = int(Ep.size());
    for (int i = 0; i < n && i < capacity; ++i) {
        E[2 * i] = Ep[i].real(); E[2 * i + 1] = Ep[i].imag();
        w[2 * i] = w8[i].real(); w[2 * i + 1] = w8[i].imag();
    }
    return n;
}

// the reference's atomic density matrices of a Green function G[nnzb][2][64][64] (host memory):
// green_dyadic::get_projection_coefficients (include/green_dyadic.hxx:1040-1091); concatenated per atom into out,
// returns the number of values
int ref_density_matrices(void* h, double const* G, double* out) {
    auto& p = ((RefPlan*)h)->p;
    auto const D = green_dyadic::get_projection_coefficients<double,2,1>((double const (*)[2][64][64])G, p.dyadic_plan,
                                                                         p.rowindx, p.rowCubePos, p.colCubePos, 0);
    size_t o = 0;
    for (auto const& m : D) { std::memcpy(out + o, m.data(), m.size()*sizeof(double)); o += m.size(); }
    return int(o);
}

} // extern "C"
