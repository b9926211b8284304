// gfa_action.hxx -- C++ drop-in for the reference's action/solver surface, header-only over the C-ABI gfa.h.
//
//   gfa::action_t<real_t, R1C2, Noco, n64>  replaces green_action::action_t (include/green_action.hxx:57-299):
//       the tfQMRgpu action contract (typedef real_t, LM, LN, multiply(...) returning flops,
//       has_preconditioner(), take_memory(), transfer(), get_plan()).  R1C2 = 2 only (tfQMRgpu has no real
//       path); Noco = 2 (non-collinear spin) has no device path here: the type exists, as the reference's
//       Act422/Act822 do, and its multiply throws gfa::error(GFA_ERR_ARGUMENT).
//   gfa::green_solver_t                     replaces green_solver_t (include/green_solver.hxx:13-38,
//       src/green_solver.cxx:31-98): the action key 1000 fp + 10 R1C2 + Noco, solve(rho, ncubes, iterations,
//       imag, echo).  Keys 32021 and 64021 solve; 32022 and 64022 return GFA_ERR_ARGUMENT; any other key too.
//   gfa::plan_t                             owns a gfa_plan (action_plan_t + construct_Green_function,
//       update_potential, update_energy_parameter, update_phases of include/green_function.hxx:23-56) and carries
//       the members tfQMRgpu reads from action_plan_t (include/action_plan.hxx:37-61).
//
// Errors: the reference exits on fatal errors; here every failure throws gfa::error with the
// library's message, and status-returning methods return the gfa.h status.
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
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
    // ---- the members tfQMRgpu reads (include/action_plan.hxx:37-61), host copies of the plan's arrays ----
    std::vector<uint16_t> colindx;       // [nnzb] column (source block) of every non-zero block of X
    std::vector<uint32_t> subset;        // [nCols] the block of X at which B = identity is non-zero (diagonal)
    uint32_t nRows = 0;                  // target block rows
    uint32_t nCols = 0;                  // source block columns, 64 right-hand sides each
    std::vector<uint32_t> rowstart;      // [nRows + 1] BSR row starts
    size_t gpu_mem = 0;                  // device bytes of the plan plus a solve's largest state (set by green_solver_t)
    float residuum_reached = 3e38f;      // stats of the last solve
    float flops_performed = 0.f;
    float flops_performed_all = 0.f;
    int iterations_needed = -99;
    int echo = 0;

    // construct_Green_function(p, ng, bc, hg, xyzZinso, echo) with the control-variable knobs as arguments
    plan_t(uint32_t const ng[3], int8_t const bc[3], double const hg[3], std::vector<double> const& xyzZinso,
           double r_trunc = 10.0, int const fd_range[3] = nullptr, int const nsources[3] = nullptr,
           double r_proj = 6.0, int echo_ = 0)
        : echo(echo_) {
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
        d.echo = echo_;
        d.scale_T = 1.0;
        check(gfa_plan_create(&d, &p_));
        check(gfa_plan_query(p_, &sizes_));
        dVol_ = hg[0] * hg[1] * hg[2];
        nRows = sizes_.nRows;
        nCols = sizes_.nCols;
        colindx = fetch<uint16_t>(GFA_EXPORT_COLINDX, sizes_.nnzb);
        subset = fetch<uint32_t>(GFA_EXPORT_SUBSET, nCols);
        rowstart = fetch<uint32_t>(GFA_EXPORT_ROW_START, size_t(nRows) + 1);
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
    uint64_t nnzb() const { return sizes_.nnzb; }

private:
    template <typename T>
    std::vector<T> fetch(int what, size_t n) const {
        std::vector<T> v(n);
        if (n) check(gfa_plan_export(p_, what, v.data(), n * sizeof(T)));
        return v;
    }
    gfa_plan* p_ = nullptr;
    gfa_sizes sizes_{};
    double dVol_ = 1;
};

template <typename floating_point_t = float, unsigned R1C2 = 2, unsigned Noco = 1, unsigned n64 = 64>
class action_t {
    static_assert(R1C2 == 2, "tfQMRgpu and this action are complex only (R1C2 = 2)");
    static_assert(n64 == 64, "blocks of 4x4x4 = 64 grid points");
public:
    typedef floating_point_t real_t;
    static unsigned constexpr LM = Noco * n64, LN = LM;

    explicit action_t(plan_t* plan, int echo = 0, int check_ = 0) : p_(plan) { (void)echo; (void)check_; }

    // y = (H - E S) x on device memory, asynchronous on `stream`; returns the reference's flop count.
    // colIndex (device memory) must hold the plan's colindx: the library checks that once per pointer, and a
    // null colIndex means the plan's own.  l2nX sizes tfQMRgpu's binary reductions and is not needed here.
    double multiply(real_t (*y)[R1C2][LM][LM], real_t const (*x)[R1C2][LM][LM], uint16_t const* colIndex,
                    uint32_t nnzb, uint32_t nCols = 1, unsigned l2nX = 0, void* stream = nullptr,
                    bool precondition = false) {
        if (Noco != 1) throw error(GFA_ERR_ARGUMENT, "the non-collinear (Noco = 2) action is not supported");
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

// green_solver_t: the action key 1000 fp + 10 R1C2 + Noco (src/green_solver.cxx:31-51) selects the precision of
// the solve; the reference reads fp and Noco from the control variables green_solver.floating.point.bits and
// green_solver.noco, here they are constructor arguments with the same mapping (fp 32 unless 64 is asked for,
// Noco 1 unless 2 is asked for).
class green_solver_t {
public:
    explicit green_solver_t(plan_t* p = nullptr, int echo = 0, int check_ = 0, int fp_bits = 32, int noco = 1)
        : p_(p) {
        (void)echo; (void)check_;
        int const fp = (fp_bits == 64) ? 64 : 32;
        action_key_ = 1000 * fp + 10 * 2 + ((noco == 2) ? 2 : 1);
        if (p_ && (action_key_ == 32021 || action_key_ == 64021)) {
            gfa_solve_opts o = options();
            uint64_t sb = 0, pb = 0;
            if (gfa_solve_memory(p_->handle(), &o, &sb, &pb) == GFA_OK) p_->gpu_mem = size_t(sb + pb);
        }
    }

    int action_key() const { return action_key_; }
    int max_refinements = 4;     // FP64 refinement steps of an FP32 solve
    int method = 0;              // inner Krylov method: 0 tfQMR (as the reference), 1 COCG
    double tolerance = 1e-9;     // tfqmrgpu::solve(action, buffer, 1e-9, ...) (include/green_action.hxx:216)
    // the FP32 policy of gfa_solve_opts (0: the library's defaults): residual reduction asked of one FP32 solve
    // (COCG: 1e-30, one FP32 solve with residual replacement), the FP32 budget and the progress check
    double inner_tolerance = 0;
    int fp32_max_iterations = 0;
    int fp32_progress_check = 0;

    // rho[plan.nCols][64] = G_ii / pi of the diagonal blocks (include/green_action.hxx:226-238).
    // iterations > 0: solve with at most that many iterations; 0: only check the construction (returns 0);
    // < 0: apply the action |iterations| times to a zero vector and report the timing, as the reference does
    // without tfQMRgpu (include/green_action.hxx:243-281).  Keys 32022/64022: GFA_ERR_ARGUMENT.
    int solve(std::complex<double> rho[], uint32_t ncubes, int iterations, int imag = 1, int echo = 0,
              gfa_solve_stats* stats = nullptr) {
        (void)imag;
        if (!p_) return GFA_ERR_STATE;
        if (action_key_ != 32021 && action_key_ != 64021) return GFA_ERR_ARGUMENT;
        if (iterations == 0) return GFA_OK;
        int const fp = action_key_ / 1000;
        if (iterations < 0) {
            double seconds = 0, flops = 0;
            int const s = gfa_action_benchmark(p_->handle(), fp, -iterations, &seconds, &flops);
            if (s == GFA_OK) {
                p_->flops_performed = float(flops);
                p_->flops_performed_all += float(flops);
                if (echo > 1) std::fprintf(stderr, "# %d calls of action.multiply: %.3e flop in %.3e s\n",
                                           -iterations, flops, seconds);
            }
            return s;
        }
        if (ncubes != p_->nCols) return GFA_ERR_ARGUMENT;
        gfa_solve_opts o = options();
        o.max_iterations = iterations;
        o.echo = echo;
        gfa_solve_stats st{};
        std::vector<double> r(size_t(ncubes) * 64 * 2);
        int const s = gfa_solve(p_->handle(), &o, r.data(), &st);
        if (stats) *stats = st;
        if (s == GFA_OK) {
            for (size_t i = 0; i < size_t(ncubes) * 64; ++i) rho[i] = std::complex<double>(r[2 * i], r[2 * i + 1]);
            p_->residuum_reached = float(st.max_residual);
            p_->iterations_needed = st.iterations;
            p_->flops_performed = float(st.flops);
            p_->flops_performed_all += float(st.flops);
        }
        return s;
    }

private:
    gfa_solve_opts options() const {
        gfa_solve_opts o{};
        o.fp_bits = action_key_ / 1000;
        o.max_iterations = 1000;
        o.tolerance = tolerance;
        o.max_refinements = max_refinements;
        o.method = method;
        o.inner_tolerance = inner_tolerance;
        o.fp32_max_iterations = fp32_max_iterations;
        o.fp32_progress_check = fp32_progress_check;
        return o;
    }
    plan_t* p_;
    int action_key_ = 0;
};

}  // namespace gfa
