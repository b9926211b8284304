// Host-only use of the C++ drop-in header: a plan of the 8x8x8 periodic cell, its sizes, and the
// action_t contract types; no device work (runs without a GPU).
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
                plan.sizes().nRows, plan.nCols(), plan.sizes().nAtoms, plan.sizes().nAtomImages);
    try {
        int8_t const bad_bc[3] = {7, 0, 0};
        gfa::plan_t bad(ng, bad_bc, hg, xyzZinso, 6.0);
        return 1;  // must have thrown
    } catch (gfa::error const& e) {
        std::printf("rejected: %s\n", e.what());
    }
    return plan.nnzb() == 512 ? 0 : 2;
}
