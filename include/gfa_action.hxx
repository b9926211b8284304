// gfa_action.hxx -- C++ drop-in for the reference's action/solver surface, header-only over the C-ABI gfa.h.
//
//   gfa::action_t<real_t, R1C2, Noco, n64>  replaces green_action::action_t (include/green_action.hxx:57-299):
//       the tfQMRgpu action contract (typedef real_t, LM, LN, multiply(...) returning flops,
//       has_preconditioner(), take_memory(), transfer(), get_plan()), for R1C2 = 2, Noco = 1.
//   gfa::green_solver_t                     replaces green_solver_t (include/green_solver.hxx:13-38):
//       solve(rho, ncubes, iterations, imag, echo) with the key from the floating-point bits.
//   gfa::plan_t                             owns a gfa_plan (action_plan_t + construct_Green_function,
//       update_potential, update_energy_parameter, update_phases of include/green_function.hxx:23-56).
//
// Errors: the reference exits on fatal errors; here every failure throws gfa::error with the
// library's message, and status-returning methods return the gfa.h status.
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "gfa.h"

namespace gfa {

struct error : std::runtime_error {
    int status;
    error(int s, std::string const& what) : std::runtime_error(what), status(s) {}
};

inline void check(int s) {
    if (s != GFA_OK) throw error(s, gfa_last_error());
}

class plan_t {
public:
    // construct_Green_function(p, ng, bc, hg, xyzZinso, echo) with the control-variable knobs as arguments
    plan_t(uint32_t const ng[3], int8_t const bc[3], double const hg[3], std::vector<double> const& xyzZinso,
           double r_trunc = 10.0, int const fd_range[3] = nullptr, int const nsources[3] = nullptr,
           double r_proj = 6.0, int echo = 0) {
        gfa_plan_desc d{};
        for (int k = 0; k < 3; ++k) {
            d.ng[k] = ng[k];
            d.bc[k] = bc[k];
            d.hg[k] = hg[k];
            d.fd_range[k] = fd_range ? fd_range[k] : 0;
            d.nsources[k] = nsources ? nsources[k] : 1;
        }
        d.xyzZinso = xyzZinso.data();
        d.natoms = int32_t(xyzZinso.size() / 8);
        d.r_trunc = r_trunc;
        d.r_proj = r_proj;
        d.echo = echo;
        d.scale_T = 1.0;
        check(gfa_plan_create(&d, &p_));
        check(gfa_plan_query(p_, &sizes_));
        dVol_ = hg[0] * hg[1] * hg[2];
    }
    ~plan_t() { gfa_plan_destroy(p_); }
    plan_t(plan_t const&) = delete;
    plan_t& operator=(plan_t const&) = delete;

    // update_potential: V on the full grid [ng_z][ng_y][ng_x]; atom matrices hs[ia] = {h, s} (nc x nc each)
    void update_potential(double const* V, std::vector<std::vector<double>> const& hs) {
        check(gfa_set_potential(p_, V));
        for (size_t ia = 0; ia < hs.size(); ++ia) {
            int const nc = int(std::sqrt(double(hs[ia].size() / 2)) + 0.5);
            check(gfa_set_atom_matrices(p_, int32_t(ia), hs[ia].data(), nc));
        }
    }
    void update_energy_parameter(std::complex<double> E) { check(gfa_set_energy(p_, E.real(), E.imag(), dVol_)); }
    void update_phases(double const k[3]) { check(gfa_set_kpoint(p_, k)); }

    gfa_plan* handle() const { return p_; }
    gfa_sizes const& sizes() const { return sizes_; }
    uint32_t nCols() const { return sizes_.nCols; }
    uint64_t nnzb() const { return sizes_.nnzb; }

private:
    gfa_plan* p_ = nullptr;
    gfa_sizes sizes_{};
    double dVol_ = 1;
};

template <typename floating_point_t = double, unsigned R1C2 = 2, unsigned Noco = 1, unsigned n64 = 64>
class action_t {
    static_assert(R1C2 == 2 && Noco == 1 && n64 == 64, "the MI355X action supports complex, collinear blocks of 64");
public:
    typedef floating_point_t real_t;
    static unsigned constexpr LM = Noco * n64, LN = LM;

    explicit action_t(plan_t* plan, int echo = 0, int check_ = 0) : p_(plan) { (void)echo; (void)check_; }

    // y = (H - E S) x on device memory, asynchronous on `stream`; returns the reference's flop count
    double multiply(real_t (*y)[R1C2][LM][LM], real_t const (*x)[R1C2][LM][LM], uint16_t const* colIndex,
                    uint32_t nnzb, uint32_t nCols = 1, unsigned l2nX = 0, void* stream = nullptr,
                    bool precondition = false) {
        double flops = 0;
        check(gfa_action_multiply(p_->handle(), 8 * int(sizeof(real_t)), (void*)y, (void const*)x, colIndex, nnzb, nCols,
                                  l2nX, stream, precondition ? 1 : 0, &flops));
        return flops;
    }
    bool has_preconditioner() const { return false; }
    void take_memory(char*& buffer) { (void)buffer; }      // the library owns its scratch
    void transfer(char* buffer, void* stream = nullptr) { (void)buffer; (void)stream; }
    plan_t* get_plan() { return p_; }

private:
    plan_t* p_;
};

// green_solver_t: action key 1000*fp + 10*R1C2 + Noco (src/green_solver.cxx:31-60) -> fp_bits of the solve
class green_solver_t {
public:
    explicit green_solver_t(plan_t* p = nullptr, int echo = 0, int check_ = 0, int fp_bits = 32)
        : p_(p), fp_bits_(fp_bits == 64 ? 64 : 32) { (void)echo; (void)check_; }

    // rho[plan.nCols][64] = G_ii / pi of the diagonal blocks (include/green_action.hxx:226-238);
    // with fp_bits 32 the FP32 tfQMR runs inside FP64 refinement
    int solve(std::complex<double> rho[], uint32_t ncubes, int iterations, int imag = 1, int echo = 0,
              gfa_solve_stats* stats = nullptr) {
        (void)imag;
        if (!p_) return GFA_ERR_STATE;
        if (ncubes != p_->nCols()) return GFA_ERR_ARGUMENT;
        gfa_solve_opts o{};
        o.fp_bits = fp_bits_;
        o.max_iterations = iterations;
        o.tolerance = 1e-9;
        o.max_refinements = 4;
        o.echo = echo;
        std::vector<double> r(size_t(ncubes) * 64 * 2);
        int const s = gfa_solve(p_->handle(), &o, r.data(), stats);
        if (s == GFA_OK)
            for (size_t i = 0; i < size_t(ncubes) * 64; ++i) rho[i] = std::complex<double>(r[2 * i], r[2 * i + 1]);
        return s;
    }

private:
    plan_t* p_;
    int fp_bits_;
};

}  // namespace gfa
