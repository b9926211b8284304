// Host-side plan of the block-sparse Green-function action (H - E S) G.
//
// The plan fixes which 4x4x4-point target blocks each source block (a column group of 64 right-hand
// sides) keeps inside its truncation sphere, the finite-difference line lists of the kinetic term and
// the atom-image sparsity of the SHO-projector term.  Construction follows the semantics of
//   green_function::construct_Green_function   src/green_function.cxx:357-1068
//   kinetic_plan::finite_difference_plan        src/kinetic_plan.cxx:38-210
//   dyadic_plan_t::dyadic_plan_t                src/dyadic_plan.cxx:159-583
// so that every index array equals the reference's (checked against the reference build by the tests).
#pragma once

#include <cstdint>
#include <complex>
#include <string>
#include <vector>

namespace gfa {

// boundary condition codes (include/boundary_condition.hxx:17-40)
enum : int8_t { BC_ISOLATED = 0, BC_PERIODIC = 1, BC_VACUUM = 2, BC_REPEAT = 3, BC_WRAP = 5 };

// finite-difference list markers (include/kinetic_plan.hxx:19-23)
constexpr int32_t CUBE_EXISTS = 1, CUBE_IS_ZERO = 0, CUBE_NEEDS_PHASE = -1;
constexpr int NHALO = 4;

inline int nsho(int numax) { return ((numax + 1) * (numax + 2) * (numax + 3)) / 6; }

struct Csr {  // compressed rows of uint32 column indices
    std::vector<uint32_t> start{0};
    std::vector<uint32_t> index;
    uint32_t rows() const { return uint32_t(start.size() - 1); }
};

struct KineticPlan {
    int fd_range = 4;            // 4: 9-point 8th order, 8: 17-point 16th order
    double prefactor = 0;        // -1/2 * scale / h^2
    std::vector<uint32_t> start; // CSR over lists
    std::vector<int32_t> list;   // entries: >0 block+1, 0 absent, <0 periodic image -(block+1)
};

struct PlanInput {
    uint32_t ng[3];
    int8_t bc[3];
    double hg[3];
    std::vector<double> xyzZinso;     // [natoms*8]: x, y, z, Z, id, numax, sigma, unused
    double r_trunc = 10.0;            // green_function.truncation.radius
    double r_proj = 6.0;              // green_function.projection.radius (units of sigma)
    int fd_range[3] = {0, 0, 0};      // 0: default (8, reduced to 4 in periodic directions)
    int nsources[3] = {1, 1, 1};      // box of source blocks at the cell centre, <0: all
    std::vector<int32_t> sources;     // optional explicit source block coordinates [n][3]
    std::vector<double> source_radii; // optional truncation radius per source block (batched instances)
    double scale_T = 1.0;             // hamiltonian.scale.kinetic
};

class Plan {
public:
    // returns 0 on success, an error text in err otherwise
    int build(PlanInput const& in, std::string& err);

    // ---- geometry ----
    uint32_t nb[3] = {0, 0, 0};
    int8_t bc_in[3] = {0, 0, 0}, bc[3] = {0, 0, 0};  // bc after wrapping
    double hg[3] = {0, 0, 0}, htrunc[3] = {0, 0, 0};
    double r_trunc = 0, r_proj = 6, r_block_circ = 0, max_distance_from_center = 0;
    int32_t min_target[3] = {0, 0, 0}, num_target[3] = {0, 0, 0};

    // ---- BSR of the Green function X ----
    uint32_t nRows = 0, nCols = 0;
    uint64_t nnzb = 0;
    std::vector<int32_t> source_coords;     // [nCols][3] global block coordinates
    std::vector<double> source_radius;      // [nCols] truncation radius of each column's sphere
    std::vector<int64_t> global_source_ids; // [nCols] bit-interleaved ids (include/global_coordinates.hxx:20-48)
    std::vector<uint32_t> row_start;        // [nRows+1]
    std::vector<uint16_t> colindx;          // [nnzb]
    std::vector<uint32_t> rowindx;          // [nnzb]
    std::vector<uint32_t> subset;           // [nCols] inzb of the diagonal block of each column
    std::vector<int32_t> veff_index;        // [nnzb] cell block of the potential, -1: none
    std::vector<float> row_pos, col_pos;    // [nRows][4], [nCols][4] block coords relative to the cell centre
    std::vector<int32_t> row_coords;        // [nRows][3] global target block coordinates

    KineticPlan kinetic[3];

    // ---- dyadic plan ----
    uint32_t nAtoms = 0, nAtomImages = 0;        // contributing atom copies, contributing images
    std::vector<double> image_pos;               // [nai][4]: x,y,z, 1/sqrt(sigma)
    std::vector<int8_t> image_lmax;              // [nai]
    std::vector<uint32_t> image_starts;          // [nai+1] coefficient offsets
    std::vector<uint32_t> image_atom;            // [nai] contributing atom index
    std::vector<int8_t> image_shift;             // [nai][4] periodic image shifts
    std::vector<uint32_t> atom_starts;           // [nAtoms+1]
    std::vector<int8_t> atom_lmax;               // [nAtoms]
    std::vector<double> atom_sigma;              // [nAtoms]
    std::vector<int32_t> atom_orig;              // [nAtoms] original atom index (row of xyzZinso)
    std::vector<int64_t> global_atom_ids;        // [nAtoms]
    Csr sho_add;                                 // rows: inzb, cols: atom images
    std::vector<Csr> sho_prj;                    // [nCols] rows: atom images, cols: inzb
    Csr sho_sum;                                 // rows: atoms, cols: atom images

    // ---- per-system and per-(E,k) data ----
    std::vector<double> xyzZinso;
    std::vector<double> V_cell;                  // [ncell][64] cell-indexed local potential
    std::vector<std::vector<double>> hs_scaled;  // per original atom: [2][nc][nc] rescaled h, s
    std::complex<double> E{0, 0};                 // energy of column 0 (the only one unless batched)
    std::vector<std::complex<double>> E_col;      // [nCols] energy of each column group
    double dVol = 0, scale_H = 1;
    std::vector<double> atom_hs;                  // per contributing atom [2][nc][nc]: h dVol, s dVol (at atom_mat_off)
    std::vector<double> atom_matrix;             // per contributing atom, [2][nc][nc] at atom_mat_off
    std::vector<uint64_t> atom_mat_off;          // [nAtoms+1]
    double phase[3][2][2] = {{{1, 0}, {1, 0}}, {{1, 0}, {1, 0}}, {{1, 0}, {1, 0}}};
    std::vector<double> image_phase;             // [nai][2] complex Bloch phase of each image

    uint64_t ncell() const { return uint64_t(nb[0]) * nb[1] * nb[2]; }
    uint64_t natomcoeffs() const { return image_starts.empty() ? 0 : image_starts.back(); }
    uint64_t ncoeffs() const { return atom_starts.empty() ? 0 : atom_starts.back(); }

    // per-(E,k) updates; return 0 or an error
    int set_potential(double const* V_grid, std::string& err);              // V[ng_z][ng_y][ng_x]
    int set_atom_matrices(int ia, double const* hs, int nc, std::string& err); // raw L2-normalised [2][nc][nc]
    void set_energy(std::complex<double> E, double dVol);
    int set_energies(std::complex<double> const* E, double dVol, std::string& err);  // one per column
    int set_kpoint(double const k[3], std::string& err);
    void update_atom_matrices();
    std::vector<double> sho_prefactors(int ia) const;  // f_i of original atom ia, zyx order
    // A^T = A: every Bloch phase the operator uses is real (kinetic halo phases only matter in periodic
    // directions; a wrapped or isolated direction has no periodic halo images)
    bool complex_symmetric() const {
        for (int d = 0; d < 3; ++d)
            for (int s = 0; s < 2; ++s)
                if (bc[d] == BC_PERIODIC && phase[d][s][1] != 0) return false;
        for (size_t i = 1; i < image_phase.size(); i += 2)
            if (image_phase[i] != 0) return false;
        return true;
    }

private:
    int build_bsr(PlanInput const& in, std::string& err,
                  std::vector<std::vector<uint8_t>>& sparsity, std::vector<int32_t>& iRow_of_coords);
    int build_kinetic(int dd, int fd_request, std::vector<std::vector<uint8_t>> const& sparsity,
                      std::vector<int32_t> const& iRow_of_coords, double scale_T, std::string& err);
    int build_dyadic(std::string& err);
};

// Bloch phase factors for one direction (src/kinetic_plan.cxx:214-269): left = exp(2 pi i angle),
// right = conj(left); values on multiples of 15 degrees are replaced by their exact forms.
int set_phase(double phase[2][2], double angle, std::string& err);

int64_t global_coordinate_id(int32_t x, int32_t y, int32_t z);

}  // namespace gfa
