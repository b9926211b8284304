// Host-only use of the C++ drop-in header: a plan of the 8x8x8 periodic cell, its sizes and tfQMRgpu members, the
// action_t contract types and the solver's action keys; no device work (runs without a GPU).
#include <complex>
#include <cstdio>
#include <vector>

#include "gfa_action.hxx"

int main() {
    uint32_t const ng[3] = {32, 32, 32};
    int8_t const bc[3] = {1, 1, 1};
    double const hg[3] = {0.25, 0.25, 0.25};
    std::vector<double> xyzZinso = {0.3, -1.2, 0.7, 1, 0, 3, 0.5, 0,   -2.0, 1.5, 0.1, 1, 1, 3, 0.5, 0};
    gfa::plan_t plan(ng, bc, hg, xyzZinso, 6.0);
    gfa::action_t<double> action(&plan);
    static_assert(gfa::action_t<double>::LM == 64, "block size");
    std::printf("nnzb %llu nRows %u nCols %u atoms %u images %u\n", (unsigned long long)plan.nnzb(),
                plan.nRows, plan.nCols, plan.sizes().nAtoms, plan.sizes().nAtomImages);
    // the members tfQMRgpu reads (include/action_plan.hxx:37-61)
    if (plan.colindx.size() != plan.nnzb() || plan.rowstart.size() != plan.nRows + 1u ||
        plan.rowstart.back() != plan.nnzb() || plan.subset.size() != plan.nCols || plan.iterations_needed != -99)
        return 3;
    // the action keys (src/green_solver.cxx:31-51): non-collinear keys are refused with a status, no exit
    std::complex<double> rho[64];
    gfa::green_solver_t s64(&plan, 0, 0, 64, 1), s32(&plan, 0, 0, 32, 1), n64(&plan, 0, 0, 64, 2), n32(&plan, 0, 0, 32, 2);
    std::printf("keys %d %d %d %d\n", s64.action_key(), s32.action_key(), n64.action_key(), n32.action_key());
    if (n64.solve(rho, plan.nCols, 100) != GFA_ERR_ARGUMENT || n32.solve(rho, plan.nCols, 100) != GFA_ERR_ARGUMENT) return 4;
    if (s64.solve(rho, plan.nCols, 0) != GFA_OK) return 5;   // zero iterations: construction check only
    std::printf("noco2 refused\n");
    try {
        int8_t const bad_bc[3] = {7, 0, 0};
        gfa::plan_t bad(ng, bad_bc, hg, xyzZinso, 6.0);
        return 1;  // must have thrown
    } catch (gfa::error const& e) {
        std::printf("rejected: %s\n", e.what());
    }
    return plan.nnzb() == 512 ? 0 : 2;
}
