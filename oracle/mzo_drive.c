/*
 * mzo_drive.c -- oracle: multi-step drivers used as the CPU baseline and by
 * the long-run parity tests.
 *
 * TEST INFRASTRUCTURE ONLY (see mzo.h).  The step order follows the
 * reference's Couette test (test/test_1D_couette.jl:66-92): NTC in every
 * cell, convection, sort, compute_props_sorted!.  Random tags: step t uses
 * tag (t << 4) | 1 for NTC and (t << 4) | 2 for convection, the same layout
 * the GPU library uses.  Cells are processed in parallel with OpenMP where
 * the reference's per-cell calls are independent (equal-weight NTC creates
 * no particles; convection and props touch disjoint particles).
 */
#include "mzo.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

int mzo_omp_max_threads(void) { return omp_get_max_threads(); }

void mzo_couette_steps(const mzo_rng *proto, const mzo_grid1d *g, const mzo_walls *wl,
                       const mzo_interaction *it, mzo_cf *cfs, mzo_pv *pv, mzo_pia *pia, double dt,
                       const double *V, double mass, mzo_props *pp, int64_t n_steps, uint64_t step0,
                       int32_t equal_weight, int32_t nthreads, int32_t props_every) {
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    for (int64_t t = 0; t < n_steps; ++t) {
        uint64_t step = step0 + (uint64_t)t;
        if (equal_weight) {
#pragma omp parallel for num_threads(nt) schedule(dynamic, 8)
            for (int64_t c = 1; c <= g->n_cells; ++c) {
                mzo_rng r = *proto;
                r.tag = (step << 4) | 1u;
                mzo_ntc_single(&r, &cfs[c - 1], it, pv, pia, c, 1, dt, V[c - 1], 1e-16, 1);
            }
        } else {
            mzo_rng r = *proto;
            r.tag = (step << 4) | 1u;
            for (int64_t c = 1; c <= g->n_cells; ++c)
                mzo_ntc_single(&r, &cfs[c - 1], it, pv, pia, c, 1, dt, V[c - 1], 1e-16, 0);
        }
        {
            int64_t n_tot = pia->n_total[0];
            if (pia->contiguous[0]) {
#pragma omp parallel for num_threads(nt) schedule(static)
                for (int64_t i = 1; i <= n_tot; ++i) {
                    /* one particle at a time through the per-particle stream */
                    mzo_rng r = *proto;
                    r.tag = (step << 4) | 2u;
                    mzo_convect_one(&r, g, wl, pv, pv->index[i - 1], 1, dt);
                }
            } else {
                mzo_rng r = *proto;
                r.tag = (step << 4) | 2u;
                mzo_convect_particles(&r, g, wl, pv, pia, 1, dt);
            }
        }
        mzo_sort_particles(g, pv, pia, 1);
        if (props_every > 0 && ((t + 1) % props_every) == 0) {
            mzo_pv *pvs[1] = {pv};
#pragma omp parallel for num_threads(nt) schedule(static)
            for (int64_t c = 1; c <= g->n_cells; ++c)
                mzo_compute_props_sorted(pvs, pia, &mass, pp, NULL, c, c);
        }
    }
}
