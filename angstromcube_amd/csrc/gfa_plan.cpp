// Host-side plan construction; see gfa_plan.hpp for the reference functions each part follows.
// Floating-point expressions that decide set membership (truncation sphere, projection spheres)
// are evaluated in the same order and precision as the reference, so the sparsity is identical.
// This file must be compiled with -ffp-contract=off.
#include "gfa_plan.hpp"

#include <atomic>
#include <cstdlib>
#include <thread>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace gfa {

namespace {
constexpr int X = 0, Y = 1, Z = 2;
inline double sq(double x) { return x * x; }

// f(i) for i in [0, n) on up to 16 host threads (GFA_PLAN_THREADS overrides), items handed out one at a time; the
// callers write per-item results and merge them afterwards in item order, so the plan does not depend on the
// thread count
template <typename F>
void parallel_for(size_t n, F&& f) {
    unsigned nt = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (char const* e = std::getenv("GFA_PLAN_THREADS")) nt = unsigned(std::max(1, std::atoi(e)));
    nt = unsigned(std::min<size_t>(nt, n));
    if (nt <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t)
        pool.emplace_back([&]() {
            for (size_t i = next.fetch_add(1); i < n; i = next.fetch_add(1)) f(i);
        });
    for (auto& t : pool) t.join();
}

template <typename T>
inline uint64_t idx3(T const n[3], int64_t x, int64_t y, int64_t z) {
    return (uint64_t(z) * uint64_t(n[Y]) + uint64_t(y)) * uint64_t(n[X]) + uint64_t(x);
}

std::string fmt(char const* f, double a = 0, double b = 0, double c = 0) {
    char s[256];
    std::snprintf(s, sizeof(s), f, a, b, c);
    return s;
}

// sho_1D_prefactor of include/sho_projection.hxx:181-183: sqrt(2^nu / (nu! sqrt(pi) sigma))
double sho_1d_norm(int nu, double sigma) {
    double fact = 1;
    for (int k = 2; k <= nu; ++k) fact *= double(k);
    double const sqrtpi = 1.7724538509055160272981674833411;  // sqrt(pi)
    return std::sqrt(double(1ull << nu) / (fact * sqrtpi * sigma));
}
}  // namespace

int64_t global_coordinate_id(int32_t x, int32_t y, int32_t z) {
    // interleave the lowest 21 bits of x, y, z as ...z1 y1 x1 z0 y0 x0
    int64_t id = 0;
    for (int b = 0; b < 21; ++b) {
        int64_t const bits = (x & 1) + 2 * (y & 1) + 4 * (z & 1);
        id |= bits << (3 * b);
        x >>= 1; y >>= 1; z >>= 1;
    }
    return id;
}

int set_phase(double phase[2][2], double angle, std::string& err) {
    if (!(std::abs(angle) < 0.503)) { err = fmt("k-point component %g outside of [-1/2, 1/2] (units of 2 pi)", angle); return 1; }
    double const arg = 2 * 3.14159265358979323846 * angle;
    double cs = std::cos(arg), sn = std::sin(arg);
    int const n24 = int(angle * 24);  // truncation, as the reference
    if (angle * 24 == n24) {          // exactly on a 15-degree grid point: use exact values
        double const sh = std::sqrt(0.5), s3 = std::sqrt(3);
        double const absval[7] = {0.0, (s3 - 1) * sh * .5, .5, sh, s3 * .5, (s3 + 1) * sh * .5, 1.0};  // sin(15 deg * i)
        int const m = n24 + 120;      // shift by 5 full turns to make it non-negative
        int const quadrant = m / 6;   // which quarter turn
        int const half = m % 12;      // position inside the half turn
        int const isin = (half <= 6) ? half : 12 - half;
        int const sgn_sin = (quadrant & 2) ? -1 : 1;
        int const sgn_cos = ((quadrant + 1) & 2) ? -1 : 1;
        sn = sgn_sin * absval[isin];
        cs = sgn_cos * absval[6 - isin];
    }
    phase[0][0] = cs; phase[0][1] = sn;    // left (forward)
    phase[1][0] = cs; phase[1][1] = -sn;   // right (backward) = conjugate
    return 0;
}

int Plan::build(PlanInput const& in, std::string& err) {
    for (int d = 0; d < 3; ++d) {
        if (in.ng[d] < 4 || (in.ng[d] & 3)) { err = fmt("grid points per direction must be a positive multiple of 4, got %g", in.ng[d]); return 1; }
        if (!(in.hg[d] > 0)) { err = "grid spacings must be positive"; return 1; }
        if (in.bc[d] < BC_ISOLATED || in.bc[d] > BC_REPEAT) { err = fmt("boundary condition %g not in {0 isolated, 1 periodic, 2 vacuum, 3 repeat}", in.bc[d]); return 1; }
        nb[d] = in.ng[d] >> 2;
        if (nb[d] > (1u << 20)) { err = "more than 2^20 blocks per direction"; return 1; }
        bc_in[d] = bc[d] = in.bc[d];
        hg[d] = in.hg[d];
    }
    if (!(in.r_trunc >= 0)) { err = "the truncation radius must be non-negative"; return 1; }
    for (double r : in.source_radii)
        if (!(r >= 0)) { err = "the truncation radius must be non-negative"; return 1; }
    if (in.xyzZinso.size() % 8) { err = "xyzZinso must hold 8 numbers per atom"; return 1; }
    xyzZinso = in.xyzZinso;
    size_t const natoms = xyzZinso.size() / 8;
    for (size_t ia = 0; ia < natoms; ++ia) {
        int const numax = int(xyzZinso[ia * 8 + 5]);
        if (numax < 0 || numax > 7) { err = fmt("atom %g has numax %g, supported are 0..7", double(ia), double(numax)); return 1; }
        if (!(xyzZinso[ia * 8 + 6] > 0)) { err = "SHO sigma must be positive"; return 1; }
    }
    r_trunc = in.r_trunc;
    for (double r : in.source_radii) r_trunc = std::max(r_trunc, r);  // the largest sphere decides wrapping
    r_proj = in.r_proj;

    std::vector<std::vector<uint8_t>> sparsity;
    std::vector<int32_t> iRow_of_coords;
    if (int s = build_bsr(in, err, sparsity, iRow_of_coords)) return s;
    for (int dd = 0; dd < 3; ++dd) {
        if (int s = build_kinetic(dd, in.fd_range[dd], sparsity, iRow_of_coords, in.scale_T, err)) return s;
    }
    if (int s = build_dyadic(err)) return s;

    V_cell.assign(ncell() * 64, 0.0);
    hs_scaled.assign(natoms, std::vector<double>());
    image_phase.assign(size_t(nAtomImages) * 2, 0.0);
    for (uint32_t i = 0; i < nAtomImages; ++i) image_phase[2 * i] = 1.0;
    atom_mat_off.assign(nAtoms + 1, 0);
    for (uint32_t a = 0; a < nAtoms; ++a) {
        int const nc = nsho(atom_lmax[a]);
        atom_mat_off[a + 1] = atom_mat_off[a] + 2ull * nc * nc;
    }
    atom_matrix.assign(atom_mat_off[nAtoms], 0.0);
    atom_hs.assign(atom_mat_off[nAtoms], 0.0);
    E_col.assign(nCols, std::complex<double>(0, 0));
    return 0;
}

// --------------------------------------------------------------------------------------------
// Sources, truncation sphere and BSR tables: src/green_function.cxx:255-354 (serial branch),
// :406-944.  Corner test (not USE_SIMPLE_RANGE_TRUNCATION), as the CMake build compiles it.
// --------------------------------------------------------------------------------------------
int Plan::build_bsr(PlanInput const& in, std::string& err,
                    std::vector<std::vector<uint8_t>>& sparsity, std::vector<int32_t>& iRow_of_coords) {
    // right-hand-side source blocks
    source_coords.clear();
    if (!in.sources.empty()) {
        if (in.sources.size() % 3) { err = "explicit sources need 3 coordinates each"; return 1; }
        for (size_t i = 0; i < in.sources.size(); i += 3) {
            for (int d = 0; d < 3; ++d) {
                int32_t const c = in.sources[i + d];
                if (c < 0 || uint32_t(c) >= nb[d]) { err = "explicit source block outside of the cell"; return 1; }
                source_coords.push_back(c);
            }
        }
    } else {
        int32_t ns[3], off[3];
        for (int d = 0; d < 3; ++d) {
            ns[d] = (in.nsources[d] < 0) ? int32_t(nb[d]) : std::min(std::max(1, in.nsources[d]), int32_t(nb[d]));
            off[d] = (int32_t(nb[d]) - ns[d]) / 2;  // centred in the cell
        }
        for (int32_t z = 0; z < ns[Z]; ++z)
            for (int32_t y = 0; y < ns[Y]; ++y)
                for (int32_t x = 0; x < ns[X]; ++x) {
                    source_coords.push_back((x + off[X]) % int32_t(nb[X]));
                    source_coords.push_back((y + off[Y]) % int32_t(nb[Y]));
                    source_coords.push_back((z + off[Z]) % int32_t(nb[Z]));
                }
    }
    nCols = uint32_t(source_coords.size() / 3);
    if (nCols < 1 || nCols > 65536) { err = "between 1 and 65536 source blocks are supported (uint16 column indices)"; return 1; }
    global_source_ids.resize(nCols);
    int32_t src_min[3] = {1 << 21, 1 << 21, 1 << 21}, src_max[3] = {-(1 << 21), -(1 << 21), -(1 << 21)};
    for (uint32_t c = 0; c < nCols; ++c) {
        int32_t const* s = &source_coords[3 * c];
        global_source_ids[c] = global_coordinate_id(s[X], s[Y], s[Z]);
        for (int d = 0; d < 3; ++d) { src_min[d] = std::min(src_min[d], s[d]); src_max[d] = std::max(src_max[d], s[d]); }
    }
    {   // largest distance of a source block centre from the centre of the source box
        double centre[3];
        for (int d = 0; d < 3; ++d) centre[d] = (((src_min[d] + src_max[d]) * 0.5) * 4 + 2) * hg[d];
        double max_d2 = 0;
        for (uint32_t c = 0; c < nCols; ++c) {
            double d2 = 0;
            for (int d = 0; d < 3; ++d) d2 += sq((source_coords[3 * c + d] * 4 + 2) * hg[d] - centre[d]);
            max_d2 = std::max(max_d2, d2);
        }
        max_distance_from_center = std::sqrt(max_d2);
    }

    // truncation: a periodic direction becomes Wrap if the sphere fits, untruncated otherwise
    double const rtrunc = std::max(0., r_trunc);
    for (int d = 0; d < 3; ++d) {
        htrunc[d] = hg[d];
        if (BC_PERIODIC == bc[d]) {
            if (2 * rtrunc > htrunc[d] * (4.0 * nb[d])) htrunc[d] = 0;  // no truncation in d
            else bc[d] = BC_WRAP;
        }
    }
    r_block_circ = 0.5 * (4 - 1) * std::sqrt(sq(hg[X]) + sq(hg[Y]) + sq(hg[Z]));
    double const rtrunc_plus = rtrunc + 2 * r_block_circ;
    int32_t itr[3], max_target[3];
    for (int d = 0; d < 3; ++d) {
        int32_t const nbox = int32_t(nb[d]) - 1;
        itr[d] = (htrunc[d] > 0) ? int32_t(std::floor(rtrunc_plus / (4 * htrunc[d]))) : nbox;
        if (BC_PERIODIC == bc[d]) itr[d] = std::min(itr[d], nbox);
        min_target[d] = src_min[d] - itr[d];
        max_target[d] = src_max[d] + itr[d];
        if (BC_ISOLATED == bc[d] || BC_PERIODIC == bc[d]) {
            min_target[d] = std::max(min_target[d], 0);
            max_target[d] = std::min(max_target[d], nbox);
        }
        num_target[d] = std::max(0, max_target[d] + 1 - min_target[d]);
    }
    uint64_t const nbox3 = uint64_t(num_target[X]) * num_target[Y] * num_target[Z];
    if (nbox3 == 0 || nbox3 > (1ull << 31)) { err = "target box empty or too large"; return 1; }

    double const r2block_circum = sq(r_block_circ * 3);
    if (!in.source_radii.empty() && in.source_radii.size() != nCols) { err = "one radius per source block needed"; return 1; }
    source_radius.assign(nCols, rtrunc);
    for (size_t c = 0; c < in.source_radii.size(); ++c) source_radius[c] = std::max(0., in.source_radii[c]);
    double const* h = htrunc;

    sparsity.assign(nCols, std::vector<uint8_t>(nbox3, 0));
    std::vector<std::vector<uint16_t>> cols_of_target(nbox3);
    std::vector<uint64_t> diag_i3(nCols, ~uint64_t(0));  // target box index of each column's own block
    // the target blocks of every source inside its sphere, in (z, y, x) order, per column in parallel
    std::vector<std::vector<uint64_t>> hits(nCols);
    std::vector<int> col_err(nCols, 0);   // 1: a block hit twice, 2: the source outside its own sphere
    parallel_for(nCols, [&](size_t cc) {
        uint32_t const c = uint32_t(cc);
        int32_t const* s = &source_coords[3 * c];
        // this source's sphere (the reference has one radius for all; src/green_function.cxx:580-581, 540)
        double const rc = source_radius[c];
        double const r2trunc = sq(rc), r2trunc_plus = sq(rc + 2 * r_block_circ);
        int32_t bfirst[3], blast[3];
        for (int d = 0; d < 3; ++d) {
            if (BC_ISOLATED == bc[d] || BC_PERIODIC == bc[d]) {
                bfirst[d] = std::max(-itr[d], min_target[d] - s[d]);
                blast[d] = std::min(itr[d], max_target[d] - s[d]);
            } else {
                bfirst[d] = -itr[d];
                blast[d] = itr[d];
            }
        }
        bool diag_hit = false;
        auto& sp = sparsity[c];
        for (int32_t bz = bfirst[Z]; bz <= blast[Z]; ++bz)
        for (int32_t by = bfirst[Y]; by <= blast[Y]; ++by)
        for (int32_t bx = bfirst[X]; bx <= blast[X]; ++bx) {
            double const d2 = sq(bx * 4 * h[X]) + sq(by * 4 * h[Y]) + sq(bz * 4 * h[Z]);
            int nci = 0;
            if (d2 < r2trunc_plus) {
                int const far = (d2 > r2block_circum);
                int const inc = 3 + 3 * far;
                for (int iz = -3; iz <= 3; iz += inc) {
                    double const d2z = sq((bz * 4 + iz) * h[Z]);
                    for (int iy = -3; iy <= 3; iy += inc) {
                        double const d2yz = sq((by * 4 + iy) * h[Y]) + d2z;
                        for (int ix = -3; ix <= 3; ix += inc) {
                            double const d2xyz = sq((bx * 4 + ix) * h[X]) + d2yz;
                            nci += (d2xyz < r2trunc);
                        }
                    }
                }
            }
            if (nci > 0) {
                int64_t const tx = s[X] + bx - min_target[X], ty = s[Y] + by - min_target[Y], tz = s[Z] + bz - min_target[Z];
                uint64_t const i3 = idx3(num_target, tx, ty, tz);
                if (sp[i3]) { col_err[c] = 1; return; }
                sp[i3] = 1;
                hits[c].push_back(i3);
                if (0 == bx && 0 == by && 0 == bz) { diag_i3[c] = i3; diag_hit = true; }
            }
        }
        if (!diag_hit) col_err[c] = 2;
    });
    for (uint32_t c = 0; c < nCols; ++c) {
        if (col_err[c] == 1) { err = "a target block was hit twice"; return 1; }
        if (col_err[c] == 2) { err = "a source block is not inside its own truncation sphere"; return 1; }
        for (uint64_t i3 : hits[c]) cols_of_target[i3].push_back(uint16_t(c));   // columns in source order
        std::vector<uint64_t>().swap(hits[c]);
    }

    // BSR rows in z, y, x order of the target box, columns in source order
    nRows = 0; nnzb = 0;
    for (uint64_t i3 = 0; i3 < nbox3; ++i3) if (!cols_of_target[i3].empty()) { ++nRows; nnzb += cols_of_target[i3].size(); }
    if (nnzb >= (1ull << 32)) { err = "more than 2^32 - 1 non-zero blocks"; return 1; }
    row_start.assign(nRows + 1, 0);
    colindx.resize(nnzb); rowindx.resize(nnzb); veff_index.assign(nnzb, -1);
    subset.assign(nCols, 0);
    row_pos.assign(size_t(nRows) * 4, 0.f); col_pos.assign(size_t(nCols) * 4, 0.f);
    row_coords.assign(size_t(nRows) * 3, 0);
    iRow_of_coords.assign(nbox3, -1);
    for (uint32_t c = 0; c < nCols; ++c)
        for (int d = 0; d < 3; ++d) col_pos[4 * c + d] = float(source_coords[3 * c + d] - nb[d] * 0.5);
    uint32_t iRow = 0;
    for (int32_t z = 0; z < num_target[Z]; ++z)
    for (int32_t y = 0; y < num_target[Y]; ++y)
    for (int32_t x = 0; x < num_target[X]; ++x) {
        uint64_t const i3 = idx3(num_target, x, y, z);
        auto const& cl = cols_of_target[i3];
        if (cl.empty()) continue;
        iRow_of_coords[i3] = int32_t(iRow);
        uint32_t const r0 = row_start[iRow];
        row_start[iRow + 1] = r0 + uint32_t(cl.size());
        int32_t const g[3] = {x + min_target[X], y + min_target[Y], z + min_target[Z]};
        for (int d = 0; d < 3; ++d) {
            row_pos[4 * iRow + d] = float(g[d] - nb[d] * 0.5);
            row_coords[3 * iRow + d] = g[d];
        }
        // cell block of the local potential, -1 outside of the cell in vacuum directions
        int32_t vidx = -1;
        {
            bool given = true;
            int32_t m[3];
            for (int d = 0; d < 3; ++d) {
                if (BC_VACUUM == bc[d] || BC_ISOLATED == bc[d]) given = given && (g[d] >= 0 && g[d] < int32_t(nb[d]));
                m[d] = g[d] % int32_t(nb[d]);
                m[d] += (m[d] < 0) * int32_t(nb[d]);
            }
            if (given) vidx = int32_t(idx3(nb, m[X], m[Y], m[Z]));
        }
        for (size_t k = 0; k < cl.size(); ++k) {
            uint32_t const inz = r0 + uint32_t(k);
            colindx[inz] = cl[k];
            rowindx[inz] = iRow;
            veff_index[inz] = vidx;
            if (diag_i3[cl[k]] == i3) subset[cl[k]] = inz;  // columns may share a source block
        }
        ++iRow;
    }
    return 0;
}

// --------------------------------------------------------------------------------------------
// Finite-difference lists: src/kinetic_plan.cxx:38-210 and src/green_function.cxx:966-989
// --------------------------------------------------------------------------------------------
int Plan::build_kinetic(int dd, int fd_request, std::vector<std::vector<uint8_t>> const& sparsity,
                        std::vector<int32_t> const& iRow_of_coords, double scale_T, std::string& err) {
    KineticPlan& kp = kinetic[dd];
    bool const periodic = (BC_PERIODIC == bc[dd]);  // wrapped directions are not periodic here
    uint32_t const periodicity = periodic ? nb[dd] : 0;
    int fd = (0 == fd_request) ? 8 : fd_request;
    if (0 == fd_request && periodic) fd = 4;        // the default range shrinks to 4 in periodic directions
    if (fd != 4 && fd != 8) { err = "finite-difference range must be 4 or 8"; return 1; }
    if (periodic && 8 == fd) { err = "the 16th-order stencil (range 8) supports no periodic direction"; return 1; }
    kp.fd_range = fd;
    kp.prefactor = -0.5 * scale_T / (hg[dd] * hg[dd]);

    int32_t num[3] = {num_target[X], num_target[Y], num_target[Z]};
    int32_t const num_dd = num[dd];
    num[dd] = 1;
    auto find_inz = [&](int32_t const idx[3], uint32_t c) -> int64_t {   // the columns of a row are ascending
        int32_t const r = iRow_of_coords[idx3(num_target, idx[X], idx[Y], idx[Z])];
        auto const b = colindx.begin() + row_start[r], e = colindx.begin() + row_start[r + 1];
        auto const it = std::lower_bound(b, e, uint16_t(c));
        return (it != e && *it == c) ? int64_t(it - colindx.begin()) : -1;
    };
    // the lines of every column in parallel, then concatenated in column order
    std::vector<std::vector<int32_t>> col_list(nCols);
    std::vector<std::vector<uint32_t>> col_len(nCols);
    std::vector<std::string> col_err(nCols);
    parallel_for(nCols, [&](size_t cc) {
        uint32_t const c = uint32_t(cc);
        std::string& err = col_err[c];
        std::vector<int32_t> line;
        auto& out = col_list[c];
        auto& lens = col_len[c];
        auto const& sp = sparsity[c];
        for (int32_t iz = 0; iz < num[Z]; ++iz)
        for (int32_t iy = 0; iy < num[Y]; ++iy)
        for (int32_t ix = 0; ix < num[X]; ++ix) {
            int32_t idx[3] = {ix, iy, iz};
            line.assign(NHALO, CUBE_IS_ZERO);
            int32_t first = -1, last = -1;
            for (int32_t id = 0; id < num_dd; ++id) {
                idx[dd] = id;
                if (sp[idx3(num_target, idx[X], idx[Y], idx[Z])]) {
                    int64_t const inz = find_inz(idx, c);
                    if (inz < 0) { err = "inconsistent sparsity pattern"; return; }
                    if (first < 0) first = id;
                    last = id;
                    line.push_back(int32_t(inz) + CUBE_EXISTS);
                }
            }
            size_t const len = line.size();  // including the leading halo
            if (len <= size_t(NHALO)) continue;
            for (int ih = 0; ih < NHALO; ++ih) line.push_back(CUBE_IS_ZERO);
            if (periodicity) {
                if (uint32_t(last + 1 - first) != periodicity) { err = "periodic line does not span the cell"; return; }
                for (int side = 0; side < 2; ++side) {
                    for (int ih = 0; ih < NHALO; ++ih) {
                        int32_t const j = side ? (last - int32_t(periodicity) + ih + 1)
                                               : (first + int32_t(periodicity) + ih - NHALO);
                        if (j < 0 || j >= num_dd) continue;
                        int32_t jdx[3] = {idx[X], idx[Y], idx[Z]};
                        jdx[dd] = j;
                        if (sp[idx3(num_target, jdx[X], jdx[Y], jdx[Z])]) {
                            int64_t const jnz = find_inz(jdx, c);
                            if (jnz >= 0) line[ih + side * len] = CUBE_NEEDS_PHASE * (int32_t(jnz) + CUBE_EXISTS);
                        }
                    }
                }
            }
            out.insert(out.end(), line.begin(), line.end());
            lens.push_back(uint32_t(line.size()));
        }
    });
    kp.start.assign(1, 0);
    kp.list.clear();
    for (uint32_t c = 0; c < nCols; ++c) {
        if (!col_err[c].empty()) { err = col_err[c]; return 1; }
        kp.list.insert(kp.list.end(), col_list[c].begin(), col_list[c].end());
        for (uint32_t n : col_len[c]) kp.start.push_back(kp.start.back() + n);
        std::vector<int32_t>().swap(col_list[c]);
    }
    return 0;
}

// --------------------------------------------------------------------------------------------
// Atom images / copies and SHO sparsity: src/dyadic_plan.cxx:159-583
// --------------------------------------------------------------------------------------------
int Plan::build_dyadic(std::string& err) {
    (void)err;
    double const cell[3] = {nb[X] * 4 * hg[X], nb[Y] * 4 * hg[Y], nb[Z] * 4 * hg[Z]};
    int32_t const natoms = int32_t(xyzZinso.size() / 8);
    double max_sigma = 0;
    for (int32_t ia = 0; ia < natoms; ++ia) max_sigma = std::max(max_sigma, xyzZinso[ia * 8 + 6]);
    double const radius = r_trunc + max_distance_from_center + 2 * max_sigma * r_proj + 2 * r_block_circ;
    int images[3] = {0, 0, 0}, copies[3] = {0, 0, 0};
    uint32_t ncopies = 1;
    for (int d = 0; d < 3; ++d) {
        int const nmx = int(std::max(0., std::ceil(radius / cell[d])));
        if (BC_PERIODIC == bc[d]) images[d] = nmx;
        else if (BC_REPEAT == bc[d]) copies[d] = nmx;
        else if (BC_WRAP == bc[d]) copies[d] = 1;
        ncopies *= uint32_t(2 * copies[d] + 1);
    }
    size_t const ncopies_all = size_t(natoms) * ncopies;

    struct Image { double pos[3]; size_t iaa; int32_t ia; int8_t shift[3]; };
    std::vector<Image> img_data;
    std::vector<std::vector<uint32_t>> cubes;  // target rows touched by each contributing image
    std::vector<uint8_t> copy_contributes(ncopies_all, 0);

    // every (periodic image shift, atom, copy) in the reference's loop order; the target rows each one touches
    // (corner test of the projection sphere) are found in parallel and the touching ones kept in that order
    struct Cand { int8_t ishift[3]; int32_t ia; size_t iaa; double pos[3]; };
    std::vector<Cand> cand;
    for (int zi = -images[Z]; zi <= images[Z]; ++zi)
    for (int yi = -images[Y]; yi <= images[Y]; ++yi)
    for (int xi = -images[X]; xi <= images[X]; ++xi) {
        int const ishift[3] = {xi, yi, zi};
        size_t iaa = 0;
        for (int32_t ia = 0; ia < natoms; ++ia)
            for (int zc = -copies[Z]; zc <= copies[Z]; ++zc)
            for (int yc = -copies[Y]; yc <= copies[Y]; ++yc)
            for (int xc = -copies[X]; xc <= copies[X]; ++xc) {
                int const cshift[3] = {xc, yc, zc};
                Cand k;
                for (int d = 0; d < 3; ++d) {
                    k.ishift[d] = int8_t(ishift[d]);
                    k.pos[d] = xyzZinso[ia * 8 + d] + (ishift[d] + cshift[d]) * cell[d];
                }
                k.ia = ia;
                k.iaa = iaa++;
                cand.push_back(k);
            }
    }
    std::vector<std::vector<uint32_t>> cand_hits(cand.size());
    parallel_for(cand.size(), [&](size_t ic) {
        Cand const& k = cand[ic];
        double const sigma = xyzZinso[k.ia * 8 + 6];
        double const r_projection = r_proj * sigma;
        double const r2projection = sq(r_projection);
        double const r2projection_plus = sq(r_projection + r_block_circ);
        double const* pos = k.pos;
        auto& hit = cand_hits[ic];
        for (uint32_t r = 0; r < nRows; ++r) {
            float const* rc = &row_pos[4 * r];
            double cd2 = 0;
            for (int d = 0; d < 3; ++d) {
                double const centre = (rc[d] * 4.f + 2.0) * hg[d];
                cd2 += sq(centre - pos[d]);
            }
            if (cd2 <= r2projection_plus) {
                double d2c[3][2];
                for (int d = 0; d < 3; ++d)
                    for (int ii = 0; ii < 2; ++ii)
                        d2c[d][ii] = sq((rc[d] * 4.0 + ii * 3 + 0.5) * hg[d] - pos[d]);
                int nci = 0;
                for (int i8 = 0; i8 < 8; ++i8) {
                    double const d2 = d2c[X][i8 & 1] + d2c[Y][(i8 >> 1) & 1] + d2c[Z][i8 >> 2];
                    nci += (d2 < r2projection);
                }
                if (nci > 0) hit.push_back(r);
            }
        }
    });
    for (size_t ic = 0; ic < cand.size(); ++ic) {
        if (cand_hits[ic].empty()) continue;
        Image im;
        for (int d = 0; d < 3; ++d) { im.pos[d] = cand[ic].pos[d]; im.shift[d] = cand[ic].ishift[d]; }
        im.iaa = cand[ic].iaa; im.ia = cand[ic].ia;
        img_data.push_back(im);
        cubes.push_back(std::move(cand_hits[ic]));
        copy_contributes[cand[ic].iaa] = 1;
    }
    nAtomImages = uint32_t(img_data.size());

    // SHO add lists (rows: inzb) and projection lists (per column: rows atom images, cols inzb)
    {
        std::vector<uint32_t> count(nnzb + 1, 0);
        for (uint32_t i = 0; i < nAtomImages; ++i)
            for (uint32_t r : cubes[i])
                for (uint32_t inz = row_start[r]; inz < row_start[r + 1]; ++inz) ++count[inz + 1];
        sho_add.start.assign(nnzb + 1, 0);
        for (uint64_t k = 0; k < nnzb; ++k) sho_add.start[k + 1] = sho_add.start[k] + count[k + 1];
        sho_add.index.assign(sho_add.start[nnzb], 0);
        std::vector<uint32_t> fill(sho_add.start.begin(), sho_add.start.end() - 1);
        std::vector<std::vector<uint32_t>> prj_count(nCols, std::vector<uint32_t>(nAtomImages + 1, 0));
        for (uint32_t i = 0; i < nAtomImages; ++i)
            for (uint32_t r : cubes[i])
                for (uint32_t inz = row_start[r]; inz < row_start[r + 1]; ++inz) {
                    sho_add.index[fill[inz]++] = i;
                    ++prj_count[colindx[inz]][i + 1];
                }
        sho_prj.assign(nCols, Csr());
        for (uint32_t c = 0; c < nCols; ++c) {
            auto& s = sho_prj[c];
            s.start.assign(nAtomImages + 1, 0);
            for (uint32_t i = 0; i < nAtomImages; ++i) s.start[i + 1] = s.start[i] + prj_count[c][i + 1];
            s.index.resize(s.start[nAtomImages]);
        }
        std::vector<std::vector<uint32_t>> pfill(nCols);
        for (uint32_t c = 0; c < nCols; ++c) pfill[c].assign(sho_prj[c].start.begin(), sho_prj[c].start.end() - 1);
        for (uint32_t i = 0; i < nAtomImages; ++i)
            for (uint32_t r : cubes[i])
                for (uint32_t inz = row_start[r]; inz < row_start[r + 1]; ++inz) {
                    uint32_t const c = colindx[inz];
                    sho_prj[c].index[pfill[c][i]++] = inz;
                }
    }

    // contributing atom copies in copy order
    std::vector<int32_t> iac_of_iaa(ncopies_all, -1);
    nAtoms = 0;
    atom_orig.clear(); global_atom_ids.clear();
    for (size_t iaa = 0; iaa < ncopies_all; ++iaa) {
        if (!copy_contributes[iaa]) continue;
        iac_of_iaa[iaa] = int32_t(nAtoms++);
        int32_t const ia = int32_t(iaa / ncopies);
        atom_orig.push_back(ia);
        global_atom_ids.push_back(int64_t(xyzZinso[ia * 8 + 4]));
    }
    image_pos.assign(size_t(nAtomImages) * 4, 0.0);
    image_lmax.assign(nAtomImages, 0);
    image_atom.assign(nAtomImages, 0);
    image_shift.assign(size_t(nAtomImages) * 4, 0);
    image_starts.assign(nAtomImages + 1, 0);
    std::vector<std::vector<uint32_t>> sum_rows(nAtoms);
    for (uint32_t i = 0; i < nAtomImages; ++i) {
        auto const& im = img_data[i];
        int const numax = int(xyzZinso[im.ia * 8 + 5]);
        double const sigma = xyzZinso[im.ia * 8 + 6];
        for (int d = 0; d < 3; ++d) { image_pos[4 * i + d] = im.pos[d]; image_shift[4 * i + d] = im.shift[d]; }
        image_pos[4 * i + 3] = 1. / std::sqrt(sigma);
        image_lmax[i] = int8_t(numax);
        image_atom[i] = uint32_t(iac_of_iaa[im.iaa]);
        image_starts[i + 1] = image_starts[i] + uint32_t(nsho(numax));
        sum_rows[image_atom[i]].push_back(i);
    }
    sho_sum.start.assign(nAtoms + 1, 0);
    sho_sum.index.clear();
    for (uint32_t a = 0; a < nAtoms; ++a) {
        sho_sum.index.insert(sho_sum.index.end(), sum_rows[a].begin(), sum_rows[a].end());
        sho_sum.start[a + 1] = uint32_t(sho_sum.index.size());
    }
    atom_lmax.assign(nAtoms, 0); atom_sigma.assign(nAtoms, 0); atom_starts.assign(nAtoms + 1, 0);
    for (uint32_t a = 0; a < nAtoms; ++a) {
        int32_t const ia = atom_orig[a];
        atom_lmax[a] = int8_t(int(xyzZinso[ia * 8 + 5]));
        atom_sigma[a] = xyzZinso[ia * 8 + 6];
        atom_starts[a + 1] = atom_starts[a] + uint32_t(nsho(atom_lmax[a]));
    }
    return 0;
}

// --------------------------------------------------------------------------------------------
// per-system and per-(E,k) data
// --------------------------------------------------------------------------------------------
int Plan::set_potential(double const* V, std::string& err) {
    if (!V) { err = "null potential"; return 1; }
    size_t const ngx = 4 * size_t(nb[X]), ngy = 4 * size_t(nb[Y]);
    for (uint32_t bz = 0; bz < nb[Z]; ++bz)
    for (uint32_t by = 0; by < nb[Y]; ++by)
    for (uint32_t bx = 0; bx < nb[X]; ++bx) {
        double* dst = &V_cell[idx3(nb, bx, by, bz) * 64];
        for (int z = 0; z < 4; ++z)
            for (int y = 0; y < 4; ++y)
                for (int x = 0; x < 4; ++x)
                    dst[(z * 4 + y) * 4 + x] = V[((bz * 4 + z) * ngy + (by * 4 + y)) * ngx + (bx * 4 + x)];
    }
    return 0;
}

// prepare for unnormalized Hermite functions: M_ij * f_i * f_j, zyx order
// (src/energy_contour.cxx:307-323 with include/sho_projection.hxx:181-209)
std::vector<double> Plan::sho_prefactors(int ia) const {
    int const numax = int(xyzZinso[ia * 8 + 5]);
    double const sigma = xyzZinso[ia * 8 + 6];
    std::vector<double> f(nsho(numax));
    int k = 0;
    for (int nz = 0; nz <= numax; ++nz) {
        double const fz = sho_1d_norm(nz, sigma);
        for (int ny = 0; ny <= numax - nz; ++ny) {
            double const fy = sho_1d_norm(ny, sigma);
            for (int nx = 0; nx <= numax - nz - ny; ++nx) f[k++] = sho_1d_norm(nx, sigma) * fy * fz;
        }
    }
    return f;
}

int Plan::set_atom_matrices(int ia, double const* hs, int nc, std::string& err) {
    int const natoms = int(xyzZinso.size() / 8);
    if (ia < 0 || ia >= natoms) { err = "atom index out of range"; return 1; }
    int const numax = int(xyzZinso[ia * 8 + 5]);
    if (nc != nsho(numax)) { err = "matrix size does not match nSHO(numax) of this atom"; return 1; }
    std::vector<double> const f = sho_prefactors(ia);
    auto& out = hs_scaled[ia];
    out.assign(2ull * nc * nc, 0.0);
    for (int m = 0; m < 2; ++m)
        for (int i = 0; i < nc; ++i)
            for (int j = 0; j < nc; ++j)
                out[(m * nc + i) * nc + j] = f[i] * hs[(m * nc + i) * nc + j] * f[j];
    return 0;
}

void Plan::set_energy(std::complex<double> E_, double dVol_) {
    E = E_;
    dVol = dVol_;
    E_col.assign(nCols, E_);
    update_atom_matrices();
}

int Plan::set_energies(std::complex<double> const* Es, double dVol_, std::string& err) {
    if (!Es) { err = "null energies"; return 1; }
    dVol = dVol_;
    E_col.assign(Es, Es + nCols);
    E = E_col[0];
    update_atom_matrices();
    return 0;
}

// (h - E s) dVol per contributing atom (src/green_function.cxx:64-109)
void Plan::update_atom_matrices() {
    for (uint32_t a = 0; a < nAtoms; ++a) {
        int const nc = nsho(atom_lmax[a]);
        double* M = &atom_matrix[atom_mat_off[a]];
        auto const& hs = hs_scaled[atom_orig[a]];
        if (hs.empty()) {
            std::fill(M, M + 2 * nc * nc, 0.0);
            std::fill(atom_hs.begin() + atom_mat_off[a], atom_hs.begin() + atom_mat_off[a + 1], 0.0);
            continue;
        }
        double const* hm = hs.data();
        double const* sm = hs.data() + nc * nc;
        double* HS = &atom_hs[atom_mat_off[a]];
        for (int i = 0; i < nc * nc; ++i) {
            M[i] = (scale_H * hm[i] - E.real() * sm[i]) * dVol;
            M[i + nc * nc] = (-E.imag() * sm[i]) * dVol;
            HS[i] = scale_H * hm[i] * dVol;        // the device forms (h - E_c s) dVol per column c
            HS[i + nc * nc] = sm[i] * dVol;
        }
    }
}

// Bloch phases of the kinetic halos and of the atom images (src/green_function.cxx:218-251)
int Plan::set_kpoint(double const k[3], std::string& err) {
    for (int d = 0; d < 3; ++d)
        if (int s = set_phase(phase[d], k[d], err)) return s;
    auto cmul = [](double const a[2], double const b[2], double c[2]) {
        double const re = a[0] * b[0] - a[1] * b[1];
        double const im = a[0] * b[1] + a[1] * b[0];
        c[0] = re; c[1] = im;
    };
    for (uint32_t i = 0; i < nAtomImages; ++i) {
        double ph[2] = {1, 0};
        for (int d = 0; d < 3; ++d) {
            int const s = image_shift[4 * i + d];
            double const* base = phase[d][s >= 0 ? 0 : 1];
            unsigned n = unsigned(s >= 0 ? s : -s);
            // exponentiation by squaring, as intpow of include/inline_math.hxx:22-32
            double xb[2] = {base[0], base[1]}, xp[2] = {1, 0};
            while (n) {
                if (n & 1) cmul(xp, xb, xp);
                n >>= 1;
                cmul(xb, xb, xb);
            }
            cmul(ph, xp, ph);
        }
        image_phase[2 * i] = ph[0];
        image_phase[2 * i + 1] = ph[1];
    }
    return 0;
}

}  // namespace gfa
