// The energy contour around the solve: the quadratures and the energy mesh of src/energy_contour.cxx:139-253 and
// the density accumulation of energy_contour::Integrator::integrate (src/energy_contour.cxx:273-451), with the
// (energy, source block) column groups of a contour solved batched instead of one energy point at a time.
//
// Quadratures.  Gauss-Legendre nodes are the roots of P_n found by Newton's method from the asymptotic guess
// cos(pi (i - 1/4) / (n + 1/2)); weights 2 / ((1 - x^2) P_n'(x)^2).  The Gauss rule for the Fermi-Dirac weight
// f(x) = 1 / (1 + exp(30 x)) on [-1, infinity) (the "fermidirac" path of the contour, starting 30 kT below the
// Fermi level) comes from the
// discretised Stieltjes procedure (Gautschi, "Orthogonal Polynomials: Computation and Approximation", 2004, sec.
// 2.2.3) on a composite Gauss-Legendre grid graded towards x = 0, then the Golub-Welsch eigenproblem of the
// Jacobi matrix, solved by Sturm-sequence bisection with the weights from the Christoffel function.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "../../include/gfa.h"

namespace gfa {
void capi_set_error(std::string const& msg);
}

namespace {

constexpr double PI = 3.14159265358979323846;
typedef std::complex<double> cplx;

// number of points the reference's tables provide for a request of n (src/energy_contour.cxx:456-461): above 32
// rounded up to a multiple of 4, above 64 to a multiple of 8; 0 if n is out of that range (> 112)
int legendre_points(int n) {
    if (n <= 0) return 0;
    if (n > 64) n = ((n - 1) / 8 + 1) * 8;
    else if (n > 32) n = ((n - 1) / 4 + 1) * 4;
    return n > 112 ? -1 : n;
}

// P_n(x) and P_n'(x) by the three-term recurrence
void legendre(int n, double x, double& p, double& dp) {
    double p0 = 1, p1 = x;
    if (n == 0) { p = 1; dp = 0; return; }
    for (int k = 2; k <= n; ++k) {
        double const p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1; p1 = p2;
    }
    p = p1;
    dp = n * (x * p1 - p0) / (x * x - 1);
}

void gauss_legendre(int n, double* x, double* w) {
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(PI * (i + 0.75) / (n + 0.5));   // i-th root from the top
        double p = 0, dp = 1;
        for (int it = 0; it < 100; ++it) {
            legendre(n, z, p, dp);
            double const dz = p / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-17) break;
        }
        legendre(n, z, p, dp);
        double const wz = 2 / ((1 - z * z) * dp * dp);
        x[i] = -z; w[i] = wz;                 // ascending: negative half first
        x[n - 1 - i] = z; w[n - 1 - i] = wz;
    }
    if (n % 2) x[n / 2] = 0;
}

// Gauss rule for the weight f(x) = 1 / (1 + exp(30 x)) on [-1, infinity); the measure beyond x = 3 (below
// e^-90) is dropped
void gauss_fermi_dirac(int n, double* x, double* w) {
    // discrete measure: composite Gauss-Legendre, 32 points on each of the subintervals between the break points
    std::vector<double> br = {-1.0};
    for (double b : {-0.5, -0.25, -0.125, -0.0625, -0.03125, -0.015625, 0.0, 0.015625, 0.03125, 0.0625, 0.125,
                     0.25, 0.5, 0.75, 1.0, 1.25, 1.5, 2.0, 2.5, 3.0})
        br.push_back(b);
    double gx[32], gw[32];
    gauss_legendre(32, gx, gw);
    std::vector<double> t, lam;
    for (size_t s = 0; s + 1 < br.size(); ++s) {
        double const a = br[s], b = br[s + 1];
        for (int k = 0; k < 32; ++k) {
            double const u = 0.5 * (a + b) + 0.5 * (b - a) * gx[k];
            t.push_back(u);
            lam.push_back(0.5 * (b - a) * gw[k] / (1 + std::exp(30 * u)));
        }
    }
    size_t const N = t.size();
    // Stieltjes: monic p_{k+1} = (t - alpha_k) p_k - beta_k p_{k-1}
    std::vector<double> alpha(n), beta(n), pm(N, 0.0), pk(N, 1.0), pn(N);
    double norm_prev = 0;
    for (int k = 0; k < n; ++k) {
        double nk = 0, tk = 0;
        for (size_t j = 0; j < N; ++j) { double const q = lam[j] * pk[j] * pk[j]; nk += q; tk += q * t[j]; }
        alpha[k] = tk / nk;
        beta[k] = (k == 0) ? nk : nk / norm_prev;
        norm_prev = nk;
        for (size_t j = 0; j < N; ++j) pn[j] = (t[j] - alpha[k]) * pk[j] - (k ? beta[k] * pm[j] : 0.0);
        pm.swap(pk); pk.swap(pn);
    }
    // eigenvalues of the Jacobi matrix (diag alpha, off-diagonal sqrt(beta_1..)) by bisection on Sturm counts
    auto count_below = [&](double lam_) {   // number of eigenvalues < lam_
        int c = 0;
        double d = alpha[0] - lam_;
        if (d < 0) ++c;
        for (int k = 1; k < n; ++k) {
            double const dd = (d == 0) ? 1e-300 : d;
            d = alpha[k] - lam_ - beta[k] / dd;
            if (d < 0) ++c;
        }
        return c;
    };
    for (int i = 0; i < n; ++i) {
        double lo = -1.0, hi = 3.0;
        for (int it = 0; it < 200 && hi - lo > 1e-17; ++it) {
            double const mid = 0.5 * (lo + hi);
            if (count_below(mid) > i) hi = mid; else lo = mid;
        }
        double const z = 0.5 * (lo + hi);
        // Christoffel: w = 1 / sum_k phat_k(z)^2 with orthonormal phat_k = p_k / sqrt(beta_0 ... beta_k)
        double s = 0, q0 = 0, q1 = 1, h = beta[0];
        s += q1 * q1 / h;
        for (int k = 0; k + 1 < n; ++k) {
            double const q2 = (z - alpha[k]) * q1 - (k ? beta[k] * q0 : 0.0);
            q0 = q1; q1 = q2;
            h *= beta[k + 1];
            s += q1 * q1 / h;
        }
        x[i] = z;
        w[i] = 1 / s;
    }
}

int fd_points(int n) { return (n < 0 || n > 16) ? -1 : n; }

// src/energy_contour.cxx:139-241: bottom path (nPol != 0), parallel path (or the density-of-states line for
// nPol == 0), Fermi-Dirac path (nPol > 0), Matsubara poles (nPol >= 0) or the closing top path (nPol < 0)
int energy_mesh(double kBT, double eBot, int nBot, int nPar, int nFer, int nPol, std::vector<cplx>& E,
                std::vector<cplx>& w, std::string& err) {
    E.clear(); w.clear();
    if (!(kBT > 0)) { err = "the temperature kBT must be positive"; return 1; }
    double const E_mu = 0, pikBT = PI * kBT;
    double xi[128], wi[128];
    auto gl = [&](int n) {
        int const m = legendre_points(n);
        if (m < 0) { err = "Gauss-Legendre quadrature supports up to 112 points"; return -1; }
        if (m > 0) gauss_legendre(m, xi, wi);
        return m;
    };
    if (nPol != 0) {
        int const m = gl(nBot);
        if (m < 0) return 1;
        double const dE = std::abs(nPol) * pikBT;
        for (int i = 0; i < m; ++i) { E.emplace_back(eBot, (xi[i] + 1) * dE); w.emplace_back(0.0, wi[i] * dE); }
        int const mp = gl(nPar);
        if (mp < 0) return 1;
        double const dP = (E_mu - 30 * kBT * (nPol > 0) - eBot) * 0.5;
        for (int i = 0; i < mp; ++i) {
            E.emplace_back(eBot + (xi[i] + 1) * dP, std::abs(nPol) * 2 * pikBT);
            w.emplace_back(wi[i] * dP, 0.0);
        }
    } else {
        if (nPar <= 0) { err = "the density-of-states line needs nPar > 0"; return 1; }
        double const dE = (nPar < 2) ? 1 : ((E_mu + 60 * kBT - eBot) / (nPar - 1.));
        int const mFer = int(30 * kBT / dE);
        for (int i = mFer - nPar; i < mFer; ++i) { E.emplace_back(i * dE, pikBT); w.emplace_back(dE, 0.0); }
    }
    if (nPol > 0) {
        int const m = fd_points(nFer);
        if (m < 0) { err = "the Fermi-Dirac quadrature supports up to 16 points"; return 1; }
        if (m > 0) gauss_fermi_dirac(m, xi, wi);
        double const dE = 30 * kBT;
        for (int i = 0; i < m; ++i) { E.emplace_back(E_mu + xi[i] * dE, nPol * 2 * pikBT); w.emplace_back(wi[i] * dE, 0.0); }
    }
    if (nPol >= 0) {
        for (int i = nPol; i > 0; --i) { E.emplace_back(E_mu, (2 * i - 1) * pikBT); w.emplace_back(0.0, -2 * pikBT); }
    } else {
        int const m = gl(nFer);
        if (m < 0) return 1;
        double const dE = std::abs(nPol) * pikBT;
        for (int i = 0; i < m; ++i) { E.emplace_back(E_mu, (1 - xi[i]) * dE); w.emplace_back(0.0, -wi[i] * dE); }
    }
    return 0;
}

template <typename F>
int guarded(F&& f) {
    try {
        return f();
    } catch (std::exception const& e) {
        gfa::capi_set_error(e.what());
        return GFA_ERR_ARGUMENT;
    } catch (...) {
        gfa::capi_set_error("unknown exception");
        return GFA_ERR_ARGUMENT;
    }
}

int fail(int code, std::string const& msg) {
    gfa::capi_set_error(msg);
    return code;
}

// the source blocks of a plan description (its own box or explicit list) as block coordinates [nsrc][3]
int source_blocks(gfa_plan_desc const* desc, std::vector<int32_t>& src) {
    gfa_plan* base = nullptr;
    if (int s = gfa_plan_create(desc, &base)) return s;
    gfa_sizes sz;
    gfa_plan_query(base, &sz);
    uint32_t const nsrc = sz.nCols;
    std::vector<float> colpos(size_t(nsrc) * 4);
    int const s0 = gfa_plan_export(base, GFA_EXPORT_COL_POS, colpos.data(), colpos.size() * sizeof(float));
    gfa_plan_destroy(base);
    if (s0) return s0;
    src.assign(size_t(nsrc) * 3, 0);
    for (uint32_t c = 0; c < nsrc; ++c)   // the export is centred: block coordinate - nb / 2
        for (int d = 0; d < 3; ++d)
            src[3 * c + d] = int32_t(std::lround(double(colpos[4 * c + d]) + 0.5 * (desc->ng[d] / 4)));
    return GFA_OK;
}

// a plan of desc with the given source blocks, the potential and the raw atom matrices set
int plan_with_sources(gfa_plan_desc const* desc, std::vector<int32_t> const& srcs, double const* V, double const* hs,
                      gfa_plan** out) {
    gfa_plan_desc d = *desc;
    d.sources = srcs.data();
    d.nsources_explicit = int32_t(srcs.size() / 3);
    gfa_plan* p = nullptr;
    if (int s = gfa_plan_create(&d, &p)) return s;
    if (int s = gfa_set_potential(p, V)) { gfa_plan_destroy(p); return s; }
    size_t off = 0;
    for (int32_t ia = 0; ia < desc->natoms; ++ia) {
        int const L = int(desc->xyzZinso[8 * ia + 5]);
        int const nc = ((L + 1) * (L + 2) * (L + 3)) / 6;
        if (int s = gfa_set_atom_matrices(p, ia, hs + off, nc)) { gfa_plan_destroy(p); return s; }
        off += size_t(2) * nc * nc;
    }
    *out = p;
    return GFA_OK;
}

}  // namespace

extern "C" {

int gfa_gauss_legendre(int32_t n, double* x, double* w) {
    return guarded([&]() -> int {
        int const m = legendre_points(n);
        if (m < 0) return -fail(GFA_ERR_ARGUMENT, "Gauss-Legendre quadrature supports up to 112 points");
        if (m > 0 && (!x || !w)) return -fail(GFA_ERR_ARGUMENT, "null argument");
        if (m > 0) gauss_legendre(m, x, w);
        return m;
    });
}

int gfa_gauss_fermi_dirac(int32_t n, double* x, double* w) {
    return guarded([&]() -> int {
        int const m = fd_points(n);
        if (m < 0) return -fail(GFA_ERR_ARGUMENT, "the Fermi-Dirac quadrature supports up to 16 points");
        if (m > 0 && (!x || !w)) return -fail(GFA_ERR_ARGUMENT, "null argument");
        if (m > 0) gauss_fermi_dirac(m, x, w);
        return m;
    });
}

int gfa_energy_mesh(double kBT, double eBot, int32_t nBot, int32_t nPar, int32_t nFer, int32_t nPol, double* E,
                    double* w, int32_t capacity, int32_t* nE) {
    return guarded([&]() -> int {
        if (!nE) return fail(GFA_ERR_ARGUMENT, "null argument");
        std::vector<cplx> Ev, wv;
        std::string err;
        if (energy_mesh(kBT, eBot, nBot, nPar, nFer, nPol, Ev, wv, err)) return fail(GFA_ERR_ARGUMENT, err);
        *nE = int32_t(Ev.size());
        if (!E || !w) return GFA_OK;   // size query
        if (capacity < *nE) return fail(GFA_ERR_ARGUMENT, "energy mesh arrays too small");
        for (size_t i = 0; i < Ev.size(); ++i) {
            E[2 * i] = Ev[i].real(); E[2 * i + 1] = Ev[i].imag();
            w[2 * i] = wv[i].real(); w[2 * i + 1] = wv[i].imag();
        }
        return GFA_OK;
    });
}

int gfa_contour_density(const gfa_plan_desc* desc, const double* V, const double* hs, const gfa_contour_opts* opts,
                        double* rho, double* rho_response, gfa_contour_stats* stats) {
    return guarded([&]() -> int {
        if (!desc || !V || !opts || !rho) return fail(GFA_ERR_ARGUMENT, "null argument");
        if (opts->nE < 1 || !opts->E || !opts->wE) return fail(GFA_ERR_ARGUMENT, "no energy points");
        if (opts->nk < 1 || !opts->k) return fail(GFA_ERR_ARGUMENT, "no k-points");
        if (desc->source_radii) return fail(GFA_ERR_ARGUMENT, "the contour uses one truncation radius (source_radii must be NULL)");
        if (desc->natoms > 0 && !hs) return fail(GFA_ERR_ARGUMENT, "null atom matrices");
        auto const t0 = std::chrono::steady_clock::now();

        // the source blocks of the caller's plan description (its own box or explicit list)
        std::vector<int32_t> src;
        if (int s = source_blocks(desc, src)) return s;
        uint32_t const nsrc = uint32_t(src.size() / 3);

        int const nE = opts->nE, nk = opts->nk;
        int const per = std::max(1, int((opts->max_groups > 0 ? opts->max_groups : 64) / int(nsrc)));   // energies per solve
        double const dVol = opts->dVol > 0 ? opts->dVol : desc->hg[0] * desc->hg[1] * desc->hg[2];
        std::vector<cplx> rho_c(size_t(nsrc) * 64, 0.0), res_c(size_t(nsrc) * 64, 0.0);
        gfa_contour_stats st{};
        double max_res = 0;

        for (int e0 = 0; e0 < nE; e0 += per) {
            int const ne = std::min(per, nE - e0);
            // column groups (energy e0 + i, source c) for i < ne, c < nsrc: the same source box as the caller's
            // plan, so every group has the sparsity of its source there
            std::vector<int32_t> srcs;
            std::vector<double> E(size_t(ne) * nsrc * 2);
            for (int i = 0; i < ne; ++i)
                for (uint32_t c = 0; c < nsrc; ++c) {
                    srcs.insert(srcs.end(), &src[3 * c], &src[3 * c] + 3);
                    E[2 * (size_t(i) * nsrc + c)] = opts->E[2 * (e0 + i)] + opts->E_Fermi;
                    E[2 * (size_t(i) * nsrc + c) + 1] = opts->E[2 * (e0 + i) + 1];
                }
            gfa_plan* p = nullptr;
            if (int s = plan_with_sources(desc, srcs, V, hs, &p)) return s;
            auto done = [&](int s) { gfa_plan_destroy(p); return s; };
            if (int s = gfa_set_energies(p, E.data(), dVol)) return done(s);
            std::vector<double> rk(size_t(ne) * nsrc * 128);
            for (int ik = 0; ik < nk; ++ik) {
                double const kv[3] = {opts->k[4 * ik], opts->k[4 * ik + 1], opts->k[4 * ik + 2]};
                double const wk = opts->k[4 * ik + 3];
                if (int s = gfa_set_kpoint(p, kv)) return done(s);
                gfa_solve_stats ss;
                if (int s = gfa_solve(p, &opts->solve, rk.data(), &ss)) return done(s);
                st.solves += 1;
                st.converged_columns += ss.converged_columns;
                st.total_columns += ss.total_columns;
                max_res = std::max(max_res, ss.max_residual);
                // src/energy_contour.cxx:370-393: rho_E = sum_k w_k rho_Ek; rho += w_E rho_E; the response
                // density sums -rho_E over all but the last two energy points
                for (int i = 0; i < ne; ++i) {
                    int const e = e0 + i;
                    cplx const wE(opts->wE[2 * e], opts->wE[2 * e + 1]);
                    for (uint32_t c = 0; c < nsrc; ++c)
                        for (int q = 0; q < 64; ++q) {
                            size_t const g = (size_t(i) * nsrc + c) * 64 + q;
                            cplx const r = wk * cplx(rk[2 * g], rk[2 * g + 1]);
                            rho_c[size_t(c) * 64 + q] += wE * r;
                            if (e < nE - 2) res_c[size_t(c) * 64 + q] -= r;
                        }
                }
            }
            gfa_plan_destroy(p);
        }
        cplx res_point = 0;
        for (int e = 0; e < nE - 2; ++e) res_point -= cplx(opts->E[2 * e] + opts->E_Fermi, opts->E[2 * e + 1]);
        res_point = (res_point != cplx(0)) ? 1.0 / res_point : cplx(1);
        for (size_t g = 0; g < size_t(nsrc) * 64; ++g) {
            rho[g] = rho_c[g].real();
            if (rho_response) rho_response[g] = (res_c[g] * res_point).real();
        }
        st.max_residual = max_res;
        st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        st.sources = int32_t(nsrc);
        if (stats) *stats = st;
        return GFA_OK;
    });
}

int gfa_solve_chunked(const gfa_plan_desc* desc, const double* V, const double* hs, double E_re, double E_im,
                      const double k[3], const gfa_solve_opts* opts, uint64_t max_bytes, double* rho,
                      gfa_chunked_stats* stats) {
    return guarded([&]() -> int {
        if (!desc || !V || !opts || !rho) return fail(GFA_ERR_ARGUMENT, "null argument");
        if (desc->source_radii) return fail(GFA_ERR_ARGUMENT, "the chunked solve uses one truncation radius (source_radii must be NULL)");
        if (desc->natoms > 0 && !hs) return fail(GFA_ERR_ARGUMENT, "null atom matrices");
        auto const t0 = std::chrono::steady_clock::now();
        std::vector<int32_t> src;
        if (int s = source_blocks(desc, src)) return s;
        uint32_t const nsrc = uint32_t(src.size() / 3);
        double const dVol = desc->hg[0] * desc->hg[1] * desc->hg[2];
        double const kv[3] = {k ? k[0] : 0.0, k ? k[1] : 0.0, k ? k[2] : 0.0};
        // device bytes per source block: the plan of all sources (host side only) and its solve state
        uint64_t per_source = 0;
        {
            gfa_plan* all = nullptr;
            if (int s = plan_with_sources(desc, src, V, hs, &all)) return s;
            uint64_t sb = 0, pb = 0;
            int const s1 = gfa_solve_memory(all, opts, &sb, &pb);
            gfa_plan_destroy(all);
            if (s1) return s1;
            per_source = (sb + pb + nsrc - 1) / std::max<uint32_t>(1, nsrc);
        }
        uint32_t const chunk = max_bytes == 0 ? nsrc
                             : uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(nsrc, max_bytes / std::max<uint64_t>(1, per_source))));
        gfa_chunked_stats st{};
        st.sources = int32_t(nsrc);
        st.sources_per_chunk = int32_t(chunk);
        for (uint32_t c0 = 0; c0 < nsrc; c0 += chunk) {
            uint32_t const n = std::min(chunk, nsrc - c0);
            std::vector<int32_t> srcs(src.begin() + 3 * size_t(c0), src.begin() + 3 * size_t(c0 + n));
            gfa_plan* p = nullptr;
            if (int s = plan_with_sources(desc, srcs, V, hs, &p)) return s;
            auto done = [&](int s) { gfa_plan_destroy(p); return s; };
            if (int s = gfa_set_energy(p, E_re, E_im, dVol)) return done(s);
            if (int s = gfa_set_kpoint(p, kv)) return done(s);
            gfa_solve_stats ss;
            if (int s = gfa_solve(p, opts, rho + size_t(c0) * 128, &ss)) return done(s);
            gfa_sizes sz;
            gfa_plan_query(p, &sz);
            st.chunks += 1;
            st.converged_columns += ss.converged_columns;
            st.total_columns += ss.total_columns;
            st.max_residual = std::max(st.max_residual, ss.max_residual);
            st.iterations = std::max(st.iterations, ss.iterations);
            st.largest_chunk_bytes = std::max<uint64_t>(st.largest_chunk_bytes, sz.device_bytes + sz.solver_bytes);
            gfa_plan_destroy(p);
        }
        st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (stats) *stats = st;
        return GFA_OK;
    });
}

}  // extern "C"
