// mz_merge.hip -- per-cell merging kernels (octree N:2 and NNLS).
#include "mz_internal.h"

extern "C" int mz_merge_octree(mz_store *h, int32_t species, int32_t threshold, int32_t target_np, int32_t split,
                               int32_t init_bounds, int32_t bounds_mode, int32_t max_Nbins, int32_t max_depth,
                               int32_t grid_clamp, int32_t squash, uint64_t tag, int32_t rng_mode,
                               int64_t *n_merged) {
    (void)h; (void)species; (void)threshold; (void)target_np; (void)split; (void)init_bounds;
    (void)bounds_mode; (void)max_Nbins; (void)max_depth; (void)grid_clamp; (void)squash; (void)tag;
    (void)rng_mode; (void)n_merged;
    mz_set_error("merge_octree: not available in this build");
    return -3;
}

extern "C" int mz_merge_nnls(mz_store *h, int32_t species, int32_t threshold, const int32_t *mim, int32_t n_mim,
                             double vref, int32_t scaling_vref, int32_t centered_at_mean, const double *v_multipliers,
                             int32_t n_vmult, double max_err, int32_t iteration_mult, double w_threshold,
                             int32_t squash, int32_t *status) {
    (void)h; (void)species; (void)threshold; (void)mim; (void)n_mim; (void)vref; (void)scaling_vref;
    (void)centered_at_mean; (void)v_multipliers; (void)n_vmult; (void)max_err; (void)iteration_mult;
    (void)w_threshold; (void)squash; (void)status;
    mz_set_error("merge_nnls: not available in this build");
    return -3;
}
