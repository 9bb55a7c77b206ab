// bfm_ctx.hpp -- context state of the C ABI (host side of libbfm.so).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <map>
#include <string>
#include <vector>

#include "../../include/bfm.h"

namespace bfm {

// Grid parameters of one neighbor build, computed on the device.
struct Plan {
    double lo[3];
    double edge;
    double inv_edge;
    double pre;       // reference prefilter radius thr + Ldiv + 2R
    double thr;
    double fmargin;   // FP32 pre-screen margin: 1e-3 um + 1e-6 * max |coordinate|
    int dims[3];
    int reach;        // stencil half-width in bins
    int nbins;
    int ok;
};

// Grid of one replicate in a batched neighbor build: its own origin and
// dimensions, the shared edge of Plan, and the offset of its bins in the
// global bin numbering. A single colony is replicate 0 with base 0.
struct RepGrid {
    double lo[3];
    int dims[3];
    int base;
};

// Device status bits, checked at synchronisation points.
enum StatusBits {
    ST_SLOT_OVERFLOW = 1,   // a cell has more neighbors than the per-cell slot capacity
    ST_PAIR_OVERFLOW = 2,   // more pairs than max_pairs
    ST_NAN = 4,             // non-finite velocity
    ST_JKR_ROOT = 8,        // JKR quartic without admissible root
    ST_QUEUE = 16           // internal candidate queue misuse (never expected)
};

struct SurrogateWeights;

struct TimedLaunch {
    std::string name;
    hipEvent_t a, b;
};

struct Tableau {
    double A[49];
    double b[7];
    double bs[7];
    int s;
};

}  // namespace bfm

struct bfm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int max_cells = 0;
    int64_t max_pairs = 0;
    int max_bins = 0;
    int slot_cap = 32;
    double bin_scale = 1.0;
    int block_threads = 256;
    int cand_variant = 8;           // candidate kernel (8: lane per home; 6 above 2^26 cells)

    bfm_params p{};
    bool params_set = false;
    bfm2d_params p2d{};             // the 2-D model (bfm2d_set_params)
    bool p2d_set = false;
    int* d_adhere = nullptr;        // max_pairs to_adhere flags of the 2-D model
    int* d_amap = nullptr;          // 64 x 64 group adhesion map of the 2-D device list
    std::string err;

    // device status word and counters
    int* d_status = nullptr;
    unsigned long long* d_counters = nullptr;   // [0] screened candidates, [1] segment tests

    // staging for host-pointer entry points
    double* d_cells = nullptr;      // max_cells x 32 (ld = max_cells)
    double* d_cells_new = nullptr;  // max_cells x 32
    double* d_nb = nullptr;         // max_pairs x 7
    double* d_noise = nullptr;      // max_cells x 6
    double* d_out6 = nullptr;       // max_cells x 6
    double* d_err6 = nullptr;       // max_cells x 6
    int* d_assume2d = nullptr;      // max_cells
    double* d_ss = nullptr;         // unused spare

    // binning
    double* d_partials = nullptr;   // bbox partial reductions
    bfm::Plan* d_plan = nullptr;
    int* d_keys = nullptr; int* d_keys_alt = nullptr;
    int* d_vals = nullptr; int* d_vals_alt = nullptr;
    int* d_bin_count = nullptr;     // max_bins + 1
    bool bins_dirty = true;         // histogram not known to be zero (see build_neighbors)
    int* d_bin_off = nullptr;       // max_bins + 1
    int* d_tile_sum = nullptr;      // per tile of the histogram (bin-offset scan)
    unsigned* d_scan_flag = nullptr;  // tile flags of the bin-offset scan
    unsigned scan_epoch = 0;          // launch epoch of the bin-offset scan
    double* d_sorted = nullptr;     // 7 x max_cells SoA: x y z nx ny nz hl
    float4* d_sortf = nullptr;      // 2 x max_cells FP32: (x y z hl), then (nx ny nz 0), sorted order
    double* d_sgeom = nullptr;      // max_cells x 8: segment records {r, n, half_l, 0}, sorted order
    double* d_geom = nullptr;       // max_cells x 8 packed segments (distance updates)
    int* d_sid = nullptr;           // sorted -> original index
    int* d_skey = nullptr;          // sorted keys
    int* d_count = nullptr;         // per-cell neighbor counts (max_cells + 1)
    int* d_offsets = nullptr;       // exclusive scan of counts (max_cells + 1)
    int* d_slot_j = nullptr;        // slot_cap x max_cells (entry-major): partner j
    double* d_slot_rec = nullptr;   // slot_cap x max_cells records of 8 doubles: d, s, t, (partner slot), pad

    // internal neighbor list (SoA)
    int* pi = nullptr; int* pj = nullptr;
    double* pdx = nullptr; double* pdy = nullptr; double* pdz = nullptr;
    double* ps = nullptr; double* pt = nullptr;
    int* d_m = nullptr;             // device pair count
    int64_t h_m = -1;               // host copy when known
    int list_n = 0;                 // number of cells the list refers to
    bool list_sorted_i = false;     // rows sorted by i (i-role ranges contiguous)
    bool transposed = false;        // ioff/joff/perm valid for the current list

    bool list_csr = false;          // per-cell CSR (device-built lists) instead of transposes
    unsigned long long* d_cnt64 = nullptr;  // per cell: (rows as i) << 32 | (rows as j) taken by atomics
    unsigned long long* d_own64 = nullptr;  // the same counts of the rows a cell's own search kept (slots
                                            // from the top of its list; 0 for the LDS-staged variants)
    unsigned long long* d_tot64 = nullptr;  // d_cnt64 + d_own64 (the list path's scan input)
    unsigned long long* d_off64 = nullptr;  // exclusive scan of d_tot64
    int* d_jslot = nullptr;         // max_cells x slot_cap: partners with larger index
    int* d_jslotk = nullptr;        // slot_cap x max_cells: the i-role slot of each partner-list entry
    unsigned long long* d_ordj = nullptr;   // slot_cap x max_cells: j-role rows in order, i << 32 | slot
    int* d_jrow = nullptr;          // max_cells x slot_cap: row index of each jslot pair
    int* d_csr = nullptr;           // 2 x max_pairs row indices (bit 31: j role)
    double* d_rec = nullptr;        // 8 x max_pairs: per-row {d, s, t, main, hard, 0}

    // transposes of imported lists (i-role offsets live in d_offsets)
    int* d_joff = nullptr;                          // max_cells + 1
    int* d_icnt = nullptr; int* d_jcnt = nullptr;   // max_cells + 1
    int* d_perm_i = nullptr; int* d_perm_j = nullptr;  // max_pairs
    bool use_perm_i = false;

    // per-cell forces
    double* d_cc = nullptr;         // 6 x max_cells
    double* d_frec = nullptr;       // 4 x max_pairs friction rows {f, contact} (friction on)
    double* d_zst = nullptr;        // 2 x max_cells stage rz, nz for the friction rows (friction on)
    double* d_k = nullptr;          // 7 x 6 x max_cells stage velocities
    unsigned long long* d_errmax = nullptr;

    // JKR table
    double* d_tab_ov = nullptr; double* d_tab_g = nullptr; double* d_tab_r = nullptr;
    int tab_nd = 0, tab_ng = 0;

    // anisotropic JKR tables: the seven grids back to back (theta, half_l,
    // coords, Rx, phi, delta, gamma), Rx at the contact point per (theta,
    // half_l, coord), and (force, radius) per (Rx1, Rx2, phi, delta[, gamma])
    double* d_an_grid = nullptr;
    double* d_an_curv = nullptr;
    double* d_an_fr = nullptr;
    int an_n[7] = {0, 0, 0, 0, 0, 0, 0};
    double an_inv_h[7] = {0, 0, 0, 0, 0, 0, 0};
    double* d_nst = nullptr;        // 3 x max_cells stage orientations (anisotropic tables set)

    // surrogate
    bfm::SurrogateWeights* sur = nullptr;
    double sur_env[4] = {0, 0, 0, 0};
    int sur_polish = 1;

    // hipcub temporary storage
    void* d_temp = nullptr;
    size_t temp_bytes = 0;

    // timing
    bool timing = false;
    std::vector<bfm::TimedLaunch> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<double, int64_t>> ktimes;
    std::vector<std::string> korder;

    // batched replicates (0 = a single colony)
    int n_rep = 0;
    int max_reps = 0;
    int* d_rep = nullptr;                   // max_cells: replicate of each cell
    bfm::RepGrid* d_grid = nullptr;         // max(1, max_reps) replicate grids
    unsigned long long* d_rep_box = nullptr;  // max_reps x 7 ordered keys: lo[3], hi[3], max half_l

    // last build stats
    double last_edge = 0;
    int last_nbins = 0;
};
