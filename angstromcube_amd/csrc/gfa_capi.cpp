// C-ABI entry points (include/gfa.h).  Every entry point catches all C++ exceptions and reports
// failures through a status code and gfa_last_error(); none terminates the process.
#include "../../include/gfa.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>
#include <new>
#include <string>

#include "gfa_device.hpp"
#include "gfa_solver.hpp"

struct gfa_plan {
    gfa::Plan host;
    gfa::DevicePlan dev;
    bool values_dirty = true;
    bool energy_set = false;
    int echo = 0;
    uint16_t const* verified_colindex = nullptr;   // a caller's device colIndex found equal to the plan's colindx
};

namespace {

thread_local std::string g_err;

int fail(int code, std::string const& msg) {
    g_err = msg;
    return code;
}

template <typename F>
int guarded(F&& f) {
    try {
        g_err.clear();
        return f();
    } catch (std::bad_alloc const&) {
        return fail(GFA_ERR_MEMORY, "host allocation failed");
    } catch (std::exception const& e) {
        return fail(GFA_ERR_ARGUMENT, e.what());
    } catch (...) {
        return fail(GFA_ERR_ARGUMENT, "unknown exception");
    }
}

int ensure_device(gfa_plan* p) {
    std::string err;
    int dev = 0;
    hipError_t const e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(GFA_ERR_DEVICE, std::string("no HIP device: ") + hipGetErrorString(e));
    if (p->dev.uploaded && p->dev.device != dev)
        return fail(GFA_ERR_STATE, "the plan was uploaded to another device");
    if (!p->dev.uploaded) {
        if (int s = gfa::device_upload_plan(p->host, p->dev, err)) return fail(s == 1 ? GFA_ERR_ARGUMENT : GFA_ERR_DEVICE, err);
    }
    if (p->values_dirty) {
        if (gfa::device_upload_values(p->host, p->dev, err)) return fail(GFA_ERR_DEVICE, err);
        p->values_dirty = false;
    }
    return GFA_OK;
}

gfa::ActionScalars scalars(gfa::Plan const& h) {
    gfa::ActionScalars s;
    s.E_re = h.E.real();
    s.E_im = h.E.imag();
    std::memcpy(s.phase, h.phase, sizeof(s.phase));
    for (int d = 0; d < 3; ++d) { s.prefactor[d] = h.kinetic[d].prefactor; s.hgrid[d] = h.hg[d]; }
    s.hgrid[3] = h.r_proj;
    return s;
}

}  // namespace

namespace gfa {
// used by the solver: run one action on a plan handle
int capi_action(gfa_plan* p, int fp_bits, void* y, void const* x, hipStream_t stream, std::string& err, int parts,
                double* dot) {
    auto const s = scalars(p->host);
    if (fp_bits == 64) return device_action<double>(p->dev, s, (double*)y, (double const*)x, stream, err, parts, dot);
    return device_action<float>(p->dev, s, (float*)y, (float const*)x, stream, err, parts, dot);
}
int capi_autotune(gfa_plan* p, int fp_bits, void* y, void const* x, hipStream_t stream, int reps, double* best_ms,
                  std::string& err) {
    auto const s = scalars(p->host);
    if (fp_bits == 64) return device_autotune<double>(p->dev, s, (double*)y, (double const*)x, stream, reps, best_ms, err);
    return device_autotune<float>(p->dev, s, (float*)y, (float const*)x, stream, reps, best_ms, err);
}
Plan const& capi_host(gfa_plan const* p) { return p->host; }
void capi_set_error(std::string const& msg) { g_err = msg; }
DevicePlan& capi_dev(gfa_plan* p) { return p->dev; }
}  // namespace gfa

extern "C" {

const char* gfa_last_error(void) { return g_err.c_str(); }

int gfa_plan_create(const gfa_plan_desc* desc, gfa_plan** plan) {
    return guarded([&]() -> int {
        if (!desc || !plan) return fail(GFA_ERR_ARGUMENT, "null argument");
        *plan = nullptr;
        if (desc->natoms < 0 || (desc->natoms > 0 && !desc->xyzZinso)) return fail(GFA_ERR_ARGUMENT, "bad atom list");
        gfa::PlanInput in;
        for (int d = 0; d < 3; ++d) {
            in.ng[d] = desc->ng[d]; in.bc[d] = desc->bc[d]; in.hg[d] = desc->hg[d];
            in.fd_range[d] = desc->fd_range[d]; in.nsources[d] = desc->nsources[d];
        }
        in.xyzZinso.assign(desc->xyzZinso, desc->xyzZinso + 8 * size_t(desc->natoms));
        in.r_trunc = desc->r_trunc;
        in.r_proj = desc->r_proj > 0 ? desc->r_proj : 6.0;
        in.scale_T = desc->scale_T != 0 ? desc->scale_T : 1.0;
        if (desc->nsources_explicit > 0) {
            if (!desc->sources) return fail(GFA_ERR_ARGUMENT, "null explicit source list");
            in.sources.assign(desc->sources, desc->sources + 3 * size_t(desc->nsources_explicit));
        }
        if (desc->source_radii) {
            if (desc->nsources_explicit <= 0) return fail(GFA_ERR_ARGUMENT, "per-source radii need an explicit source list");
            in.source_radii.assign(desc->source_radii, desc->source_radii + size_t(desc->nsources_explicit));
        }
        auto* p = new gfa_plan();
        p->echo = desc->echo;
        std::string err;
        if (p->host.build(in, err) != 0) {
            delete p;
            return fail(GFA_ERR_ARGUMENT, err);
        }
        if (p->echo > 0)
            std::printf("# gfa plan: %u rows, %u columns, %llu blocks, %u atoms, %u images\n", p->host.nRows,
                        p->host.nCols, (unsigned long long)p->host.nnzb, p->host.nAtoms, p->host.nAtomImages);
        *plan = p;
        return GFA_OK;
    });
}

int gfa_plan_destroy(gfa_plan* plan) {
    return guarded([&]() -> int {
        if (!plan) return GFA_OK;
        gfa::solver_release(plan);
        if (plan->dev.uploaded || !plan->dev.allocations.empty()) gfa::device_free(plan->dev);
        delete plan;
        return GFA_OK;
    });
}

int gfa_plan_query(const gfa_plan* plan, gfa_sizes* s) {
    return guarded([&]() -> int {
        if (!plan || !s) return fail(GFA_ERR_ARGUMENT, "null argument");
        auto const& h = plan->host;
        std::memset(s, 0, sizeof(*s));
        s->nnzb = h.nnzb; s->nRows = h.nRows; s->nCols = h.nCols;
        s->nAtoms = h.nAtoms; s->nAtomImages = h.nAtomImages;
        s->natomcoeffs = h.natomcoeffs(); s->ncoeffs = h.ncoeffs();
        for (int d = 0; d < 3; ++d) {
            s->nlists[d] = uint32_t(h.kinetic[d].start.size() - 1);
            s->fd_range[d] = h.kinetic[d].fd_range;
            s->nlist_entries[d] = h.kinetic[d].list.size();
            s->bc[d] = h.bc[d];
        }
        s->sho_add_nnz = h.sho_add.index.size();
        for (auto const& c : h.sho_prj) s->sho_prj_nnz += c.index.size();
        s->sho_sum_nnz = h.sho_sum.index.size();
        s->nchunks = plan->dev.nchunks;
        s->device_bytes = plan->dev.bytes;
        s->solver_bytes = gfa::solver_held_bytes(const_cast<gfa_plan*>(plan));
        return GFA_OK;
    });
}

int gfa_plan_export(const gfa_plan* plan, int what, void* dst, uint64_t dst_bytes) {
    return guarded([&]() -> int {
        if (!plan || !dst) return fail(GFA_ERR_ARGUMENT, "null argument");
        auto const& h = plan->host;
        std::vector<char> buf;
        auto put = [&](void const* src, size_t n) { buf.insert(buf.end(), (char const*)src, (char const*)src + n); };
        switch (what) {
            case GFA_EXPORT_ROW_START: put(h.row_start.data(), h.row_start.size() * 4); break;
            case GFA_EXPORT_COLINDX: put(h.colindx.data(), h.colindx.size() * 2); break;
            case GFA_EXPORT_SUBSET: put(h.subset.data(), h.subset.size() * 4); break;
            case GFA_EXPORT_VEFF_INDEX: put(h.veff_index.data(), h.veff_index.size() * 4); break;
            case GFA_EXPORT_ROW_POS: put(h.row_pos.data(), h.row_pos.size() * 4); break;
            case GFA_EXPORT_COL_POS: put(h.col_pos.data(), h.col_pos.size() * 4); break;
            case GFA_EXPORT_SOURCE_IDS: put(h.global_source_ids.data(), h.global_source_ids.size() * 8); break;
            case GFA_EXPORT_KIN_START_X: case GFA_EXPORT_KIN_START_Y: case GFA_EXPORT_KIN_START_Z: {
                auto const& k = h.kinetic[what - GFA_EXPORT_KIN_START_X];
                put(k.start.data(), k.start.size() * 4);
            } break;
            case GFA_EXPORT_KIN_LIST_X: case GFA_EXPORT_KIN_LIST_Y: case GFA_EXPORT_KIN_LIST_Z: {
                auto const& k = h.kinetic[what - GFA_EXPORT_KIN_LIST_X];
                put(k.list.data(), k.list.size() * 4);
            } break;
            case GFA_EXPORT_IMAGE_POS: put(h.image_pos.data(), h.image_pos.size() * 8); break;
            case GFA_EXPORT_IMAGE_LMAX: put(h.image_lmax.data(), h.image_lmax.size()); break;
            case GFA_EXPORT_IMAGE_STARTS: put(h.image_starts.data(), h.image_starts.size() * 4); break;
            case GFA_EXPORT_IMAGE_ATOM: put(h.image_atom.data(), h.image_atom.size() * 4); break;
            case GFA_EXPORT_IMAGE_SHIFT: put(h.image_shift.data(), h.image_shift.size()); break;
            case GFA_EXPORT_ATOM_STARTS: put(h.atom_starts.data(), h.atom_starts.size() * 4); break;
            case GFA_EXPORT_ATOM_LMAX: put(h.atom_lmax.data(), h.atom_lmax.size()); break;
            case GFA_EXPORT_ATOM_IDS: put(h.global_atom_ids.data(), h.global_atom_ids.size() * 8); break;
            case GFA_EXPORT_ADD_START: put(h.sho_add.start.data(), h.sho_add.start.size() * 4); break;
            case GFA_EXPORT_ADD_INDEX: put(h.sho_add.index.data(), h.sho_add.index.size() * 4); break;
            case GFA_EXPORT_PRJ_START:
                for (auto const& c : h.sho_prj) put(c.start.data(), c.start.size() * 4);
                break;
            case GFA_EXPORT_PRJ_INDEX:
                for (auto const& c : h.sho_prj) put(c.index.data(), c.index.size() * 4);
                break;
            case GFA_EXPORT_SUM_START: put(h.sho_sum.start.data(), h.sho_sum.start.size() * 4); break;
            case GFA_EXPORT_SUM_INDEX: put(h.sho_sum.index.data(), h.sho_sum.index.size() * 4); break;
            case GFA_EXPORT_ROWINDX: put(h.rowindx.data(), h.rowindx.size() * 4); break;
            case GFA_EXPORT_V_CELL: put(h.V_cell.data(), h.V_cell.size() * 8); break;
            case GFA_EXPORT_ATOM_MATRICES: put(h.atom_matrix.data(), h.atom_matrix.size() * 8); break;
            case GFA_EXPORT_IMAGE_PHASE: put(h.image_phase.data(), h.image_phase.size() * 8); break;
            case GFA_EXPORT_PHASE: put(h.phase, sizeof(h.phase)); break;
            case GFA_EXPORT_KIN_PREFACTOR:
                for (int d = 0; d < 3; ++d) put(&h.kinetic[d].prefactor, 8);
                break;
            default: return fail(GFA_ERR_ARGUMENT, "unknown export id");
        }
        if (buf.size() > dst_bytes) return fail(GFA_ERR_ARGUMENT, "destination too small: need " + std::to_string(buf.size()) + " bytes");
        if (!buf.empty()) std::memcpy(dst, buf.data(), buf.size());
        return GFA_OK;
    });
}

int gfa_set_potential(gfa_plan* plan, const double* V) {
    return guarded([&]() -> int {
        if (!plan) return fail(GFA_ERR_ARGUMENT, "null plan");
        std::string err;
        if (plan->host.set_potential(V, err)) return fail(GFA_ERR_ARGUMENT, err);
        plan->values_dirty = true;
        return GFA_OK;
    });
}

int gfa_set_atom_matrices(gfa_plan* plan, int32_t ia, const double* hs, int32_t nc) {
    return guarded([&]() -> int {
        if (!plan || !hs) return fail(GFA_ERR_ARGUMENT, "null argument");
        std::string err;
        if (plan->host.set_atom_matrices(ia, hs, nc, err)) return fail(GFA_ERR_ARGUMENT, err);
        plan->host.update_atom_matrices();
        plan->values_dirty = true;
        return GFA_OK;
    });
}

int gfa_set_energy(gfa_plan* plan, double E_re, double E_im, double dVol) {
    return guarded([&]() -> int {
        if (!plan) return fail(GFA_ERR_ARGUMENT, "null plan");
        if (!(dVol > 0)) return fail(GFA_ERR_ARGUMENT, "dVol must be positive");
        plan->host.set_energy(std::complex<double>(E_re, E_im), dVol);
        plan->energy_set = true;
        plan->values_dirty = true;
        return GFA_OK;
    });
}

int gfa_set_energies(gfa_plan* plan, const double* E, double dVol) {
    return guarded([&]() -> int {
        if (!plan || !E) return fail(GFA_ERR_ARGUMENT, "null argument");
        if (!(dVol > 0)) return fail(GFA_ERR_ARGUMENT, "dVol must be positive");
        std::vector<std::complex<double>> Es(plan->host.nCols);
        for (uint32_t c = 0; c < plan->host.nCols; ++c) Es[c] = std::complex<double>(E[2 * c], E[2 * c + 1]);
        std::string err;
        if (plan->host.set_energies(Es.data(), dVol, err)) return fail(GFA_ERR_ARGUMENT, err);
        plan->energy_set = true;
        plan->values_dirty = true;
        return GFA_OK;
    });
}

int gfa_set_kpoint(gfa_plan* plan, const double k[3]) {
    return guarded([&]() -> int {
        if (!plan || !k) return fail(GFA_ERR_ARGUMENT, "null argument");
        std::string err;
        if (plan->host.set_kpoint(k, err)) return fail(GFA_ERR_ARGUMENT, err);
        plan->values_dirty = true;
        return GFA_OK;
    });
}

int gfa_action(gfa_plan* plan, int fp_bits, void* y, const void* x, void* stream, double* nflops) {
    return guarded([&]() -> int {
        if (!plan || !y || !x) return fail(GFA_ERR_ARGUMENT, "null argument");
        if (fp_bits != 64 && fp_bits != 32) return fail(GFA_ERR_ARGUMENT, "fp_bits must be 32 or 64");
        if (y == x) return fail(GFA_ERR_ARGUMENT, "the action cannot run in place");
        if (!plan->energy_set) return fail(GFA_ERR_STATE, "set the energy before applying the action");
        if (int s = ensure_device(plan)) return s;
        std::string err;
        if (int s = gfa::capi_action(plan, fp_bits, y, x, (hipStream_t)stream, err))
            return fail(s == 1 ? GFA_ERR_ARGUMENT : GFA_ERR_DEVICE, err);
        if (nflops) *nflops = gfa::reference_flops(plan->host, fp_bits);
        return GFA_OK;
    });
}

int gfa_action_noalloc(gfa_plan* plan, int fp_bits, void* y, const void* x, void* stream, double* nflops) {
    return guarded([&]() -> int {
        if (!plan || !y || !x) return fail(GFA_ERR_ARGUMENT, "null argument");
        if (fp_bits != 64 && fp_bits != 32) return fail(GFA_ERR_ARGUMENT, "fp_bits must be 32 or 64");
        if (y == x) return fail(GFA_ERR_ARGUMENT, "the action cannot run in place");
        if (!plan->energy_set) return fail(GFA_ERR_STATE, "set the energy before applying the action");
        if (!plan->dev.uploaded)
            return fail(GFA_ERR_STATE, "the allocation-free action needs the plan on the device: call gfa_plan_prepare "
                                       "(or any other device entry point) first");
        if (int s = ensure_device(plan)) return s;   // uploaded: copies changed values only, allocates nothing
        std::string err;
        auto const sc = scalars(plan->host);
        int const r = fp_bits == 64
            ? gfa::device_action_noalloc<double>(plan->dev, sc, (double*)y, (double const*)x, (hipStream_t)stream, err)
            : gfa::device_action_noalloc<float>(plan->dev, sc, (float*)y, (float const*)x, (hipStream_t)stream, err);
        if (r) return fail(r == 1 ? GFA_ERR_ARGUMENT : GFA_ERR_DEVICE, err);
        if (nflops) *nflops = gfa::reference_flops(plan->host, fp_bits);
        return GFA_OK;
    });
}

int gfa_action_part(gfa_plan* plan, int fp_bits, int part, void* y, const void* x, void* stream) {
    return guarded([&]() -> int {
        if (!plan || !y || !x) return fail(GFA_ERR_ARGUMENT, "null argument");
        if (fp_bits != 64 && fp_bits != 32) return fail(GFA_ERR_ARGUMENT, "fp_bits must be 32 or 64");
        if (part < 1 || part > 3) return fail(GFA_ERR_ARGUMENT, "part must be 1 (projection), 2 (local) or 3 (both)");
        if (y == x) return fail(GFA_ERR_ARGUMENT, "the action cannot run in place");
        if (!plan->energy_set) return fail(GFA_ERR_STATE, "set the energy before applying the action");
        if (int s = ensure_device(plan)) return s;
        std::string err;
        if (int s = gfa::capi_action(plan, fp_bits, y, x, (hipStream_t)stream, err, part))
            return fail(s == 1 ? GFA_ERR_ARGUMENT : GFA_ERR_DEVICE, err);
        return GFA_OK;
    });
}

int gfa_atom_density_matrices(gfa_plan* plan, const void* X, double factor, double* D, void* stream) {
    return guarded([&]() -> int {
        if (!plan || !X || !D) return fail(GFA_ERR_ARGUMENT, "null argument");
        gfa::Plan const& h = plan->host;
        if (h.nAtomImages != h.nAtoms)
            return fail(GFA_ERR_ARGUMENT, "atomic density matrices need one image per atom: the reduction over periodic "
                                          "images is not defined (include/green_dyadic.hxx:1085 asserts the same)");
        if (int s = ensure_device(plan)) return s;
        std::vector<double> cx;
        std::string err;
        if (gfa::device_atom_coefficients(plan->dev, scalars(h), (double const*)X, cx, (hipStream_t)stream, err))
            return fail(GFA_ERR_DEVICE, err);
        std::vector<size_t> off(h.nAtoms + 1, 0);
        for (uint32_t a = 0; a < h.nAtoms; ++a) off[a + 1] = off[a] + size_t(gfa::nsho(h.atom_lmax[a])) * gfa::nsho(h.atom_lmax[a]);
        std::fill(D, D + off[h.nAtoms], 0.0);
        typedef std::complex<double> cplx;
        float const R2 = float(h.r_proj * h.r_proj);
        for (uint32_t iai = 0; iai < h.nAtomImages; ++iai) {   // SHOprj_right (include/green_dyadic.hxx:925-1036)
            int const L = h.image_lmax[iai];
            int const nc = gfa::nsho(L), n2 = ((L + 1) * (L + 2)) / 2;
            uint32_t const a = h.image_atom[iai], a0 = h.atom_starts[a];
            double const* apos = &h.image_pos[4 * size_t(iai)];
            double const sig_inv = apos[3] * apos[3];
            // Hermite functions of this image at the 4x4x4 points of every source block: H[c][nu][dir][i4], the
            // 1D functions of hermite_1d (include/green_dyadic.hxx:33-107) evaluated on the host
            std::vector<double> H(size_t(h.nCols) * (L + 1) * 12);
            std::vector<float> x2(size_t(h.nCols) * 12);
            for (uint32_t c = 0; c < h.nCols; ++c)
                for (int dir = 0; dir < 3; ++dir)
                    for (int i4 = 0; i4 < 4; ++i4) {
                        double const xi = sig_inv * ((double(h.col_pos[4 * c + dir]) * 4 + i4 + 0.5) * h.hg[dir] - apos[dir]);
                        double const xx = xi * xi;
                        x2[c * 12 + dir * 4 + i4] = float(xx);
                        double hn = (xx < h.r_proj * h.r_proj) ? std::exp(-0.5 * xx) : 0.0, hm = 0;
                        H[(c * (L + 1) + 0) * 12 + dir * 4 + i4] = hn;
                        for (int nu = 0; nu < L; ++nu) {
                            double const hp = xi * hn - (0.5 * nu) * hm;
                            H[(c * (L + 1) + nu + 1) * 12 + dir * 4 + i4] = hp;
                            hm = hn; hn = hp;
                        }
                    }
            std::vector<cplx> piGp(size_t(nc) * nc, 0.0), czyx(nc), byx(n2), ax(L + 1);
            for (int isho = 0; isho < nc; ++isho) {
                std::fill(czyx.begin(), czyx.end(), cplx(0));
                for (uint32_t c = 0; c < h.nCols; ++c) {
                    double const* src = &cx[((size_t(a0) + isho) * h.nCols + c) * 128];
                    auto Hc = [&](int nu, int dir, int i4) { return H[(c * (L + 1) + nu) * 12 + dir * 4 + i4]; };
                    for (int z4 = 0; z4 < 4; ++z4) {
                        float const d2z = x2[c * 12 + 8 + z4] - R2;
                        if (!(d2z < 0)) continue;
                        std::fill(byx.begin(), byx.end(), cplx(0));
                        for (int y4 = 0; y4 < 4; ++y4) {
                            float const d2yz = x2[c * 12 + 4 + y4] + d2z;
                            if (!(d2yz < 0)) continue;
                            std::fill(ax.begin(), ax.end(), cplx(0));
                            for (int x4 = 0; x4 < 4; ++x4) {
                                if (!(x2[c * 12 + x4] + d2yz < 0)) continue;
                                int const xyz = (z4 * 4 + y4) * 4 + x4;
                                cplx const pG(src[xyz], src[64 + xyz]);
                                for (int ix = 0; ix <= L; ++ix) ax[ix] += pG * Hc(ix, 0, x4);
                            }
                            for (int iy = 0; iy <= L; ++iy) {
                                int const iyx0 = (iy * (2 * L + 3 - iy)) >> 1;
                                for (int ix = 0; ix <= L - iy; ++ix) byx[iyx0 + ix] += ax[ix] * Hc(iy, 1, y4);
                            }
                        }
                        for (int jsho = 0, iz = 0; iz <= L; ++iz)
                            for (int iy = 0; iy <= L - iz; ++iy) {
                                int const iyx0 = (iy * (2 * L + 3 - iy)) >> 1;
                                for (int ix = 0; ix <= L - iz - iy; ++ix) czyx[jsho++] += byx[iyx0 + ix] * Hc(iz, 2, z4);
                            }
                    }
                }
                for (int jsho = 0; jsho < nc; ++jsho) piGp[size_t(isho) * nc + jsho] = czyx[jsho];
            }
            cplx const ph(h.image_phase[2 * iai], h.image_phase[2 * iai + 1]);
            for (size_t ij = 0; ij < piGp.size(); ++ij) D[off[a] + ij] += (piGp[ij] * ph).imag() * factor;
        }
        return GFA_OK;
    });
}

int gfa_gradient(gfa_plan* plan, const void* X, const void* Lambda, const double* E, const double* w, double* grad_V,
                 double* grad_hs, void* stream) {
    return guarded([&]() -> int {
        if (!plan || !X || !Lambda || !E || !w) return fail(GFA_ERR_ARGUMENT, "null argument");
        if (!plan->energy_set) return fail(GFA_ERR_STATE, "set the energies before the gradient");
        if (int s = ensure_device(plan)) return s;
        gfa::Plan const& h = plan->host;
        std::vector<double> wE(size_t(4) * h.nCols);
        for (uint32_t c = 0; c < h.nCols; ++c) {
            double const wr = w[2 * c], wi = w[2 * c + 1], er = E[2 * c], ei = E[2 * c + 1];
            wE[4 * c] = wr; wE[4 * c + 1] = wi;
            wE[4 * c + 2] = -(er * wr - ei * wi); wE[4 * c + 3] = -(er * wi + ei * wr);   // -E w
        }
        std::vector<double> gcell, gmat;
        std::string err;
        auto const sc = scalars(h);
        if (gfa::device_gradient(plan->dev, h, sc, (double const*)X, (double const*)Lambda, wE, gcell, gmat,
                                 (hipStream_t)stream, err))
            return fail(GFA_ERR_DEVICE, err);
        if (grad_V) {   // cell-blocked -> grid [ng_z][ng_y][ng_x] (the inverse of gfa_set_potential's blocking)
            size_t const ngx = 4 * size_t(h.nb[0]), ngy = 4 * size_t(h.nb[1]);
            for (uint32_t bz = 0; bz < h.nb[2]; ++bz)
            for (uint32_t by = 0; by < h.nb[1]; ++by)
            for (uint32_t bx = 0; bx < h.nb[0]; ++bx) {
                double const* src = &gcell[((size_t(bz) * h.nb[1] + by) * h.nb[0] + bx) * 64];
                for (int z = 0; z < 4; ++z)
                    for (int y = 0; y < 4; ++y)
                        for (int x = 0; x < 4; ++x)
                            grad_V[((bz * 4 + z) * ngy + (by * 4 + y)) * ngx + (bx * 4 + x)] = src[(z * 4 + y) * 4 + x];
            }
        }
        if (grad_hs) {  // per original atom [2][nc][nc]: dL/dh and dL/ds of the matrices given to gfa_set_atom_matrices
            size_t const natoms = h.xyzZinso.size() / 8;
            std::vector<size_t> off(natoms + 1, 0);
            for (size_t ia = 0; ia < natoms; ++ia) {
                size_t const nc = gfa::nsho(int(h.xyzZinso[ia * 8 + 5]));
                off[ia + 1] = off[ia] + 2 * nc * nc;
            }
            std::fill(grad_hs, grad_hs + off[natoms], 0.0);
            for (uint32_t a = 0; a < h.nAtoms; ++a) {
                int const ia = h.atom_orig[a];
                auto const f = h.sho_prefactors(ia);
                size_t const nc = f.size();
                double const* gm = &gmat[h.atom_mat_off[a]];
                double* out = grad_hs + off[ia];
                for (size_t k = 0; k < nc; ++k)
                    for (size_t l = 0; l < nc; ++l) {
                        double const ff = f[k] * f[l] * h.dVol;
                        out[k * nc + l] += h.scale_H * ff * gm[k * nc + l];
                        out[nc * nc + k * nc + l] += ff * gm[nc * nc + k * nc + l];
                    }
            }
        }
        return GFA_OK;
    });
}

int gfa_set_tile_shape(gfa_plan* plan, int fp_bits, int kind, int columns, int lines) {
    return guarded([&]() -> int {
        if (!plan) return fail(GFA_ERR_ARGUMENT, "null plan");
        if (fp_bits != 64 && fp_bits != 32) return fail(GFA_ERR_ARGUMENT, "fp_bits must be 32 or 64");
        if (kind == 0 && ((columns != 16 && columns != 32) || (lines != 1 && lines != 2 && lines != 4)))
            return fail(GFA_ERR_ARGUMENT, "tile shapes of kind 0: 16 or 32 columns, 1, 2 or 4 lines");
        if (kind == 1 && columns != 32 && columns != 64)
            return fail(GFA_ERR_ARGUMENT, "tile shapes of kind 1: 32 or 64 columns");
        if (kind == 2 && ((columns != 16 && columns != 32 && !(columns == 64 && fp_bits == 32)) ||
                          (lines != 2 && lines != 3)))
            return fail(GFA_ERR_ARGUMENT, "tile shapes of kind 2: 16 or 32 columns (64 in FP32), 2 or 3 directions "
                                          "prefetched");
        if (kind < 0 || kind > 2) return fail(GFA_ERR_ARGUMENT, "tile shape kind must be 0, 1 or 2");
        int const v = fp_bits == 64 ? 0 : 1;
        plan->dev.tune_kind[v] = kind;
        plan->dev.tune_jw[v] = columns;
        plan->dev.tune_lb[v] = lines;
        plan->dev.tuned[v] = true;
        gfa::solver_release(plan);  // a captured solver graph holds the previous shape
        return GFA_OK;
    });
}

int gfa_set_dyadic_engine(gfa_plan* plan, int fp_bits, int engine) {
    return guarded([&]() -> int {
        if (!plan) return fail(GFA_ERR_ARGUMENT, "null plan");
        if (fp_bits != 64 && fp_bits != 32) return fail(GFA_ERR_ARGUMENT, "fp_bits must be 32 or 64");
        if (engine != 0 && engine != 1) return fail(GFA_ERR_ARGUMENT, "dyadic engine must be 0 (vector) or 1 (matrix)");
        plan->dev.prj_engine[fp_bits == 64 ? 0 : 1] = engine;
        gfa::solver_release(plan);  // a captured solver graph holds the previous launches
        return GFA_OK;
    });
}

int gfa_get_dyadic_engine(gfa_plan* plan, int fp_bits, int* engine) {
    return guarded([&]() -> int {
        if (!plan || !engine) return fail(GFA_ERR_ARGUMENT, "null argument");
        if (fp_bits != 64 && fp_bits != 32) return fail(GFA_ERR_ARGUMENT, "fp_bits must be 32 or 64");
        *engine = plan->dev.prj_engine[fp_bits == 64 ? 0 : 1];
        return GFA_OK;
    });
}

int gfa_get_tile_shape(gfa_plan* plan, int fp_bits, int* kind, int* columns, int* lines, int* tuned) {
    return guarded([&]() -> int {
        if (!plan || !kind || !columns || !lines || !tuned) return fail(GFA_ERR_ARGUMENT, "null argument");
        if (fp_bits != 64 && fp_bits != 32) return fail(GFA_ERR_ARGUMENT, "fp_bits must be 32 or 64");
        int const v = fp_bits == 64 ? 0 : 1;
        *kind = plan->dev.tune_kind[v];
        *columns = plan->dev.tune_jw[v];
        *lines = plan->dev.tune_lb[v];
        *tuned = plan->dev.tuned[v] ? 1 : 0;
        return GFA_OK;
    });
}

int gfa_autotune(gfa_plan* plan, int fp_bits, void* y, const void* x, void* stream, int reps, double* best_ms) {
    return guarded([&]() -> int {
        if (!plan || !y || !x) return fail(GFA_ERR_ARGUMENT, "null argument");
        if (fp_bits != 64 && fp_bits != 32) return fail(GFA_ERR_ARGUMENT, "fp_bits must be 32 or 64");
        if (y == x) return fail(GFA_ERR_ARGUMENT, "the action cannot run in place");
        if (!plan->energy_set) return fail(GFA_ERR_STATE, "set the energy before tuning the action");
        if (int s = ensure_device(plan)) return s;
        std::string err;
        if (gfa::capi_autotune(plan, fp_bits, y, x, (hipStream_t)stream, reps > 0 ? reps : 3, best_ms, err))
            return fail(GFA_ERR_DEVICE, err);
        return GFA_OK;
    });
}

int gfa_action_multiply(gfa_plan* plan, int fp_bits, void* y, const void* x, const uint16_t* colIndex,
                        uint32_t nnzb, uint32_t nCols, uint32_t l2nX, void* stream, int precondition,
                        double* nflops) {
    (void)l2nX;   // sizes tfQMRgpu's binary reductions; the library's reductions need no such level count
    int const s = guarded([&]() -> int {
        if (!plan) return fail(GFA_ERR_ARGUMENT, "null plan");
        if (nnzb != plan->host.nnzb || nCols != plan->host.nCols)
            return fail(GFA_ERR_ARGUMENT, "nnzb/nCols do not match the plan");
        if (precondition) return fail(GFA_ERR_ARGUMENT, "the action has no preconditioner");
        // the blocks' columns come from the plan; a caller's colIndex (device memory) must hold the same values.
        // Checked once per pointer (a blocking copy), so that tfQMRgpu's repeated calls with its own copy cost nothing
        if (colIndex && colIndex != plan->verified_colindex) {
            if (int e = ensure_device(plan)) return e;
            if (colIndex != plan->dev.colindx) {
                std::vector<uint16_t> h(nnzb);
                hipError_t const e = hipMemcpy(h.data(), colIndex, size_t(nnzb) * sizeof(uint16_t), hipMemcpyDefault);
                if (e != hipSuccess)
                    return fail(GFA_ERR_ARGUMENT, std::string("colIndex is not readable device memory: ") +
                                                      hipGetErrorString(e));
                if (!std::equal(h.begin(), h.end(), plan->host.colindx.begin()))
                    return fail(GFA_ERR_ARGUMENT, "colIndex differs from the plan's colindx");
            }
            plan->verified_colindex = colIndex;
        }
        return GFA_OK;
    });
    if (s != GFA_OK) return s;
    return gfa_action(plan, fp_bits, y, x, stream, nflops);
}

int gfa_action_benchmark(gfa_plan* plan, int fp_bits, int32_t iterations, double* seconds, double* nflops) {
    return guarded([&]() -> int {
        if (!plan) return fail(GFA_ERR_ARGUMENT, "null plan");
        if (fp_bits != 64 && fp_bits != 32) return fail(GFA_ERR_ARGUMENT, "fp_bits must be 32 or 64");
        if (iterations < 1) return fail(GFA_ERR_ARGUMENT, "iterations must be positive");
        if (!plan->energy_set) return fail(GFA_ERR_STATE, "set the energy before applying the action");
        if (int s = ensure_device(plan)) return s;
        size_t const bytes = size_t(plan->host.nnzb) * 2 * 64 * 64 * (fp_bits / 8);
        void* v[2] = {nullptr, nullptr};
        auto release = [&](int status) {
            for (void* q : v) if (q) (void)hipFree(q);
            return status;
        };
        for (auto& q : v)
            if (hipMalloc(&q, bytes) != hipSuccess) { q = nullptr; return release(fail(GFA_ERR_MEMORY, "benchmark vectors")); }
        if (hipMemset(v[0], 0, bytes) != hipSuccess) return release(fail(GFA_ERR_DEVICE, "benchmark vectors"));
        std::string err;
        if (int s = gfa::capi_action(plan, fp_bits, v[1], v[0], nullptr, err, 0))   // workspace only
            return release(fail(s == 1 ? GFA_ERR_ARGUMENT : GFA_ERR_DEVICE, err));
        if (hipDeviceSynchronize() != hipSuccess) return release(fail(GFA_ERR_DEVICE, "benchmark set-up"));
        auto const t0 = std::chrono::steady_clock::now();
        for (int32_t it = 0; it < iterations; ++it) {   // y = A x, then swap, as the reference's loop
            if (int s = gfa::capi_action(plan, fp_bits, v[1], v[0], nullptr, err))
                return release(fail(s == 1 ? GFA_ERR_ARGUMENT : GFA_ERR_DEVICE, err));
            if (hipDeviceSynchronize() != hipSuccess) return release(fail(GFA_ERR_DEVICE, "benchmark action"));
            std::swap(v[0], v[1]);
        }
        double const dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (seconds) *seconds = dt;
        if (nflops) *nflops = gfa::reference_flops(plan->host, fp_bits) * iterations;
        return release(GFA_OK);
    });
}

int gfa_solver_release(gfa_plan* plan) {
    return guarded([&]() -> int {
        if (!plan) return fail(GFA_ERR_ARGUMENT, "null plan");
        gfa::solver_release(plan);
        return GFA_OK;
    });
}

int gfa_solve_memory(gfa_plan* plan, const gfa_solve_opts* opts, uint64_t* solver_bytes, uint64_t* plan_bytes) {
    return guarded([&]() -> int {
        if (!plan || !opts) return fail(GFA_ERR_ARGUMENT, "null argument");
        if (opts->fp_bits != 64 && opts->fp_bits != 32) return fail(GFA_ERR_ARGUMENT, "fp_bits must be 32 or 64");
        if (solver_bytes) *solver_bytes = gfa::solve_bytes(plan, opts->fp_bits, opts->method);
        if (plan_bytes) {
            if (plan->dev.uploaded) {
                *plan_bytes = plan->dev.bytes;
            } else {   // index data (about 100 bytes per block, the local kernels' block headers twice, 28 bytes
                       // per add entry and 12 per projection entry) and the potential
                gfa::Plan const& h = plan->host;
                bool wide = false;
                for (int d = 0; d < 3; ++d) wide = wide || h.kinetic[d].fd_range == 8;
                uint64_t b = h.nnzb * (120 + 2 * 4 * (wide ? 32 : 16)) + h.ncell() * 64 * 8 + h.atom_hs.size() * 8;
                b += h.sho_add.index.size() * 28;
                for (auto const& c : h.sho_prj) b += c.index.size() * 12;
                b += uint64_t(h.natomcoeffs()) * h.nCols * 128 * 8 * 2;   // coefficient workspace (both precisions)
                *plan_bytes = b;
            }
        }
        return GFA_OK;
    });
}

int gfa_plan_prepare(gfa_plan* plan, int fp_bits) {
    return guarded([&]() -> int {
        if (!plan) return fail(GFA_ERR_ARGUMENT, "null plan");
        if (fp_bits != 64 && fp_bits != 32) return fail(GFA_ERR_ARGUMENT, "fp_bits must be 32 or 64");
        if (!plan->energy_set) return fail(GFA_ERR_STATE, "set the energy before preparing the plan");
        if (int s = ensure_device(plan)) return s;
        std::string err;
        // parts = 0 allocates the action's workspace of this precision and launches nothing
        if (int s = gfa::capi_action(plan, fp_bits, nullptr, nullptr, nullptr, err, 0))
            return fail(s == 1 ? GFA_ERR_ARGUMENT : GFA_ERR_DEVICE, err);
        hipError_t const e = hipDeviceSynchronize();
        if (e != hipSuccess) return fail(GFA_ERR_DEVICE, std::string("prepare: ") + hipGetErrorString(e));
        return GFA_OK;
    });
}

int gfa_solve(gfa_plan* plan, const gfa_solve_opts* opts, double* rho, gfa_solve_stats* stats) {
    return guarded([&]() -> int {
        if (!plan || !opts) return fail(GFA_ERR_ARGUMENT, "null argument");
        if (!plan->energy_set) return fail(GFA_ERR_STATE, "set the energy before solving");
        if (int s = ensure_device(plan)) return s;
        std::string err;
        int const s = gfa::solve(plan, *opts, rho, stats, err);
        if (s) return fail(s, err);
        return GFA_OK;
    });
}

}  // extern "C"
