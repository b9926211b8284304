// Device-side plan and action launchers for gfx950.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "gfa_plan.hpp"

namespace gfa {

// one projection work item: blocks [begin, end) of the projection list of (atom image, column)
struct ProjChunk {
    uint32_t image;
    uint32_t col;
    uint32_t begin;
    uint32_t end;
};

template <typename real_t>
struct Workspace {
    real_t* partial = nullptr;  // [total chunk coefficients][2][64] projection partial sums
    real_t* dcoef = nullptr;    // [natomcoeffs * nCols][2][64] addition coefficients per atom image
    real_t* zero = nullptr;     // one block of zeros, read in place of absent neighbours
    size_t partial_elems = 0, dcoef_elems = 0;
};

struct GradWs {  // buffers of the adjoint gradient
    bool ready = false;
    uint32_t* cell_start = nullptr;   // [ncell+1] blocks reading each cell block's potential
    uint32_t* cell_blocks = nullptr;
    double* gpart = nullptr;          // [nnzb][64]
    double* gcell = nullptr;          // [ncell][64]
    double* wE = nullptr;             // [nCols][4]: w_c, -E_c w_c
    double* cX = nullptr;             // [ncoeffs][nCols][2][64] collected atom coefficients of X
    double* cL = nullptr;             // same for lambda
    double* gmat = nullptr;           // per contributing atom [2][nc][nc]
};

struct DevicePlan {
    bool uploaded = false;
    int device = -1;
    uint64_t nnzb = 0;
    uint32_t nCols = 0, nRows = 0, nAtoms = 0, nAtomImages = 0;
    int nh_max = 1;                    // neighbour blocks per side in the widest stencil
    bool has_images = false;           // some finite-difference halo is a periodic image
    int fd_range[3] = {4, 4, 4};
    // local-kernel tile shape per precision (index 0: FP64, 1: FP32), set by the autotuner; the defaults are the
    // shapes measured fastest on a 159,256-block plan (FP64 k_lines 16 columns, 2 directions ahead; FP32 k_plane
    // 32 columns)
    int tune_kind[2] = {2, 1}, tune_jw[2] = {16, 32}, tune_lb[2] = {2, 0};
    bool tuned[2] = {false, false};
    // SHO projection and add-back per precision: 0 the sum-factorised forms on the vector ALUs (the add-back inside
    // the local kernel), 1 the dense forms on the matrix cores (the add-back in its own kernel after the local one);
    // atoms with numax > 5 always take the vector forms
    int prj_engine[2] = {0, 0};

    // index data
    uint16_t* colindx = nullptr;       // [nnzb]
    uint32_t* rowindx = nullptr;       // [nnzb]
    int32_t* veff = nullptr;           // [nnzb]
    // per-block headers of the local kernels, head_words (16, or 32 when nh_max = 2) words each: the block, its
    // column group, potential cell, add-list range and neighbour codes [direction][side][k], so that a work-group
    // reads everything it needs about its block with one scalar load
    uint32_t* heads = nullptr;         // [nnzb][head_words] in block order
    uint32_t* items = nullptr;         // [nnzb][head_words] in processing order
    int head_words = 16;
    uint32_t* add_entry = nullptr;     // [add nnz][4]: image, numax, Hermite slot, first coefficient of the image
    // Hermite functions per (atom image, target row) slot: htab[slot][8][3][4], xtab[slot][3][4] = xi^2
    double* htab = nullptr;
    float* xtab = nullptr;
    uint32_t* prj_slot = nullptr;      // [projection blocks] slot of each chunk block
    uint32_t nslots = 0;
    std::map<uint64_t, uint32_t> slot_map;  // (image << 32 | row) -> slot, during upload only
    std::vector<uint32_t> h_order;     // [nnzb] processing order of the local kernels (the headers follow it)
    float* row_pos = nullptr;          // [nRows][4]
    uint32_t* add_start = nullptr;     // [nnzb+1]
    uint32_t* wide_blocks = nullptr;   // blocks touched by an atom image with numax > 5
    uint32_t n_wide_blocks = 0;
    uint32_t* narrow_blocks = nullptr; // blocks touched by an atom image with numax <= 5
    uint32_t n_narrow_blocks = 0;
    double* image_pos = nullptr;       // [nai][4]
    int8_t* image_lmax = nullptr;      // [nai]
    uint32_t* image_starts = nullptr;  // [nai+1]
    ProjChunk* chunks = nullptr;       // [nchunks]
    uint32_t* chunk_coef = nullptr;    // [nchunks+1] coefficient offsets of the partial sums
    uint32_t* prj_blocks = nullptr;    // flat inzb lists of all chunks
    uint32_t* pair_chunk = nullptr;    // [nai*nCols+1] chunk range of each (image, column)
    uint32_t* bucket_chunks = nullptr; // chunk ids sorted by numax
    uint32_t bucket_start[9] = {0};    // ranges into bucket_chunks per numax 0..7
    uint32_t* sum_start = nullptr;     // [nAtoms+1]
    uint32_t* sum_index = nullptr;
    int8_t* atom_lmax = nullptr;       // [nAtoms]
    uint32_t* atom_starts = nullptr;   // [nAtoms+1]
    double* cscratch = nullptr;        // [ncoeffs * nCols][2][64] atom coefficients when numax > 5
    uint64_t ncoeffs = 0;
    uint32_t max_atom_nc = 0;
    uint32_t nchunks = 0;
    uint64_t total_chunk_coefs = 0;
    uint64_t natomcoeffs = 0;

    // per-system / per-(E,k) data
    double* V = nullptr;               // [ncell][64]
    double* atom_hs = nullptr;         // concatenated [2][nc][nc]: h dVol, s dVol
    double* E_col = nullptr;           // [nCols][2] energy per column group
    // retirement: when set, the action processes only the listed blocks and skips inactive columns
    uint32_t* active_order = nullptr;  // [n_active] blocks of active column groups, in processing order
    uint32_t* active_items = nullptr;  // [n_active][head_words] their headers
    uint64_t n_active = 0;
    uint8_t* col_active = nullptr;     // [nCols] 1 for an active column group
    uint64_t* atom_mat_off = nullptr;  // [nAtoms+1]
    double* image_phase = nullptr;     // [nai][2]

    Workspace<double> ws64;
    GradWs grad;
    double* atom_coef = nullptr;       // [ncoeffs][nCols][2][64] atom projections (density matrices)
    Workspace<float> ws32;
    size_t bytes = 0;
    std::vector<void*> allocations;
};

struct ActionScalars {
    double E_re, E_im;
    double phase[3][2][2];
    double prefactor[3];
    double hgrid[4];                  // grid spacings and projection radius (units of sigma)
};

// headers (heads[order[i]]) of n blocks into items, on the stream
int device_gather_heads(DevicePlan const& d, uint32_t const* order, uint64_t n, uint32_t* items, hipStream_t stream,
                        std::string& err);

// Upload the index data (once) and the per-(E,k) data; 0 or error.
int device_upload_plan(Plan const& p, DevicePlan& d, std::string& err);
int device_upload_values(Plan const& p, DevicePlan& d, std::string& err);
void device_free(DevicePlan& d);

// parts of the action: the SHO projection and atom-matrix multiply (coefficients into the workspace), and
// the local kernel (potential, kinetic, add-back of the workspace coefficients) that writes y
enum { ACTION_PROJECT = 1, ACTION_LOCAL = 2, ACTION_ALL = 3 };
// ext: a workspace the caller provides (no buffer of the plan is allocated or used); ext_cscratch its atom-multiply
// scratch for numax 6, 7 atoms
template <typename real_t>
int device_action(DevicePlan& d, ActionScalars const& s, real_t* y, real_t const* x, hipStream_t stream,
                  std::string& err, int parts = ACTION_ALL, double* dot = nullptr, Workspace<real_t>* ext = nullptr,
                  double* ext_cscratch = nullptr);

// the whole action with its workspace in the library's fixed per-device scratch array (no allocation); the plan
// must be on the device already.  Returns 1 if the plan needs more than NOALLOC_SCRATCH_BYTES.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
size_t noalloc_scratch_bytes(DevicePlan const& d, int fp_bits);
template <typename real_t>
int device_action_noalloc(DevicePlan& d, ActionScalars const& s, real_t* y, real_t const* x, hipStream_t stream,
                          std::string& err);

// flop count of one action as the reference counts it (include/green_action.hxx:154-175)
template <typename real_t>
int device_autotune(DevicePlan& d, ActionScalars const& s, real_t* y, real_t const* x, hipStream_t stream, int reps,
                    double* best_ms, std::string& err);
// the projections of an FP64 vector X onto every atom, collected over its images with the Bloch phases, on the
// host: [((atom_starts[a] + s) * nCols + c)][re/im][j]
int device_atom_coefficients(DevicePlan& d, ActionScalars const& s, double const* X, std::vector<double>& host,
                             hipStream_t stream, std::string& err);
int device_gradient(DevicePlan& d, Plan const& p, ActionScalars const& s, double const* X, double const* Lam,
                    std::vector<double> const& wE, std::vector<double>& gcell_host, std::vector<double>& gmat_host,
                    hipStream_t stream, std::string& err);
double reference_flops(Plan const& p, int fp_bits);

}  // namespace gfa
