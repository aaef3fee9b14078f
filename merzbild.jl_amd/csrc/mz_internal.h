// mz_internal.h -- host-side store handle and internal launch helpers.
#pragma once

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mzhip.h"
#include "mz_common.h"

#define MZ_MAX_SPECIES 4
#define MZ_SORT_CHUNK 1024
#define MZ_SORT_MAX_CELLS 14336   // LDS: 8 KB per chunk + 4 B per cell within 64 KB
#define MZ_MAX_BOUNCES 64
#define MZ_MAX_MOMENTS 8   // total-moment powers per compute_props_with_total_moments! call
#define MZ_NTC_MAX_CANDIDATES 100000000  // per cell and call; larger counts flag an error
#define MZ_KB 1.380649e-23  // src/constants.jl:4
#define MZ_NOALLOC_SCRATCH_WORDS (1 << 21)   // int32 words of the fixed scratch array (8 MiB per GPU)

struct MzGrid {
    double L, dx, inv_dx, min_x, max_x;
    int32_t nx;
};

struct MzWalls {
    double T[2];
    double v[2][3];
    double acc[2];
    double vsq[2][MZ_MAX_SPECIES];
};

struct MzInteraction {
    double mu1, mu2, factor, expo;
    int32_t set;
};

// All species views by value (kernel arguments, no device-side table).
struct MzViews {
    MzSpView v[MZ_MAX_SPECIES];
};

struct MzEventPair {
    hipEvent_t a, b;
    int64_t n = 1;
};

struct MzScratch {
    void *ptr = nullptr;
    size_t bytes = 0;
};

// The slab launch's shape, kept between calls (see mz_couette_slab).
struct MzSlabShapeCache {
    bool valid = false, avg = false;
    int32_t req_spc = 0, req_ppt = 0;   // the requested shape it was picked for
    int R = 0, spc = 0, n_slabs = 0, spx = 0, grid = 0;
    const void *fn = nullptr;           // nullptr: no shape fits
};

struct mz_store {
    int n_cells = 0, n_species = 0;
    std::vector<MzSpecies> sp;
    hipStream_t stream = nullptr;
    MzGrid grid{};
    bool has_grid = false;
    MzWalls walls{};
    bool has_walls = false;
    std::vector<MzInteraction> inter;
    double *d_sgwm = nullptr;    // [s1][s2][cell]
    int32_t *d_cfcnt = nullptr;  // [s1][s2][cell][5]: n1, n2, n_coll, performed, eq_w
    uint64_t seed = 0;
    double *d_replay[3] = {nullptr, nullptr, nullptr};
    int64_t replay_rows[3] = {0, 0, 0}, replay_cols[3] = {0, 0, 0};
    int32_t *d_err = nullptr;
    uint32_t *d_ticket = nullptr;  // last-block counter of k_sort_rows (left at 0 by every launch)
    double *d_np = nullptr, *d_n = nullptr, *d_v = nullptr, *d_T = nullptr;
    double *d_V = nullptr, *d_invV = nullptr;
    double *d_masses = nullptr;
    std::map<std::string, MzScratch> scratch;
    // timing
    bool timing = false;
    std::vector<std::pair<std::string, MzEventPair>> open_events;
    std::map<std::string, std::vector<MzEventPair>> pending;
    std::map<std::string, std::pair<double, int64_t>> timing_acc;
    std::vector<hipEvent_t> event_pool;   // read events, reused by later brackets
    // captured step graphs
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    std::string graph_key;
    uint64_t *d_step = nullptr;
    // slab-resident Couette kernel: requested shape (0 = automatic) and the
    // last launch's {particles per thread, cells per slab, slabs, outcome}
    int32_t slab_spc = 0, slab_ppt = 0;
    int32_t slab_info[4] = {0, 0, 0, 0};
    MzSlabShapeCache slab_shape;
    int32_t slab_min_ppt = 0;      // floor on particles per thread, raised after an aborted launch
    int32_t *h_report = nullptr;   // pinned host words the device reports outcomes into
    // host copies of what mz_couette_steps last uploaded (skips unchanged uploads)
    double masses_up[MZ_MAX_SPECIES] = {0.0, 0.0, 0.0, 0.0};
    bool masses_valid = false;
    std::vector<double> V_up;
    std::vector<double> ntc_V_up;    // of the NTC calls, at ntc_V_dev
    const void *ntc_V_dev = nullptr;
    int32_t last_path = -1;   // of mz_couette_steps: 0 eager, 1 graph, 2 slab
    int32_t n_moments = 0;    // moments of the last mz_compute_props_moments
    int32_t sort_path = 0;    // 0 automatic, 1 count matrix, 2 radix (mz_set_sort_path)
    // batched 0-D cells: item packing (0 = default) and the last launch's
    // {items, max cells per item, target particles per item, workgroups}
    int32_t b0_max_cells = 0, b0_target = 0;
    int32_t b0_info[4] = {0, 0, 0, 0};
    int32_t nnls_info[2] = {0, 0};   // last candidate NNLS merge: cells run, cells re-run in FP64
    // surface properties of the two walls: [set][species][element][MZ_SURF_Q],
    // set 0 = the last convection's values, 1 = the time average
    double *d_surf = nullptr;
    uint32_t *d_surf_ticket = nullptr;   // last-block counter of k_convect_surf (left at 0 by every launch)
    double surf_inv_area[2] = {1.0, 1.0};
    // time-averaged PhysProps: [np | n | v (x3) | T] x [species][cell], then lpa[species]
    double *d_pavg = nullptr;
    // FluxProps: [set][species][cell][9] (energy xyz, diagonal xyz, off-diagonal xy xz yz), set 1 = average
    double *d_flux = nullptr;
};

MzViews mz_views(mz_store *h);
void *mz_scratch(mz_store *h, const char *name, size_t bytes);
int mz_read_counters(mz_store *h, int species, MzCounters *c);
MzRngParams mz_rng_params(mz_store *h, int kind, uint64_t tag, int mode);
int mz_timing_begin(mz_store *h, const char *family);
int mz_timing_end(mz_store *h);
int mz_timing_end_n(mz_store *h, int64_t n);

// A timing bracket closed on every path out of its scope, early returns
// included, so an error between the two events cannot leave it open.
struct MzTimed {
    mz_store *h;
    int64_t n;
    bool open;
    MzTimed(mz_store *h_, const char *family, int64_t n_ = 1) : h(h_), n(n_) {
        open = mz_timing_begin(h, family) == 0 && h->timing;
    }
    ~MzTimed() {
        if (open) mz_timing_end_n(h, n);
    }
    MzTimed(const MzTimed &) = delete;
    MzTimed &operator=(const MzTimed &) = delete;
};

// device address of the fixed scratch array on the current device
int mz_noalloc_scratch_init(int32_t **out);
// true when another live store's stream still has work queued
bool mz_other_stores_busy(mz_store *h);
void mz_swap_buffers(MzSpecies &s);
void mz_graph_release(mz_store *h);

int mz_launch_sort(mz_store *h, int species, int compute_cells);
int mz_sort_prealloc(mz_store *h, int species);
int mz_launch_convect(mz_store *h, int species, double dt, uint64_t tag, int rng_mode, bool contiguous,
                      int32_t n_pos, const uint64_t *step_ptr = nullptr);
int mz_launch_props(mz_store *h, int variant, int ndens, double avg_inv = 0.0);
int mz_launch_convect_surf(mz_store *h, int species, double dt, uint64_t tag, int rng_mode, bool contiguous,
                           int32_t n_pos, double mass, const uint64_t *step_ptr, double avg_inv);
int mz_launch_surf_avg(mz_store *h, double inv);
int mz_launch_props_avg(mz_store *h, double inv);
int mz_surf_prealloc(mz_store *h, int species);
int mz_launch_ntc_equal(mz_store *h, int s1, int s2, double dt, const double *dV, double dw_tol,
                        int equal_weight, uint64_t tag, int rng_mode, int cell_lo, int cell_hi,
                        const uint64_t *step_ptr = nullptr);
int mz_couette_slab(mz_store *h, int32_t n_steps, uint64_t step0, double dt, const double *dV, int32_t props_every,
                    int *ran, double avg_inv = 0.0, double mass = 0.0, int retry = 0);
