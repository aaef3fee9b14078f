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

/* sort_particles! (src/grids/grid_sorting.jl:58-113) for the threaded
 * baseline: the same stable counting sort, split over threads.  Each thread
 * counts the cells of its own contiguous range of logical positions; cell c
 * of thread t starts after cell c of every earlier thread, so a forward
 * scatter per thread gives the reference's order (within a cell, ascending
 * previous position).  The workspace is the caller's (no allocation per
 * step); the state must be contiguous (the Couette loop keeps it so). */
typedef struct {
    int64_t *hist;     /* [nt][n_cells + 1] */
    int64_t *sorted;   /* [capacity] */
    int64_t cap;
    int nt;
} sort_ws;

static void sort_parallel(const mzo_grid1d *g, mzo_pv *pv, mzo_pia *pia, sort_ws *ws) {
    const int64_t n_tot = pia->n_total[0], nc = g->n_cells;
    const int nt = ws->nt;
#pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num();
        const int64_t lo = n_tot * t / nt, hi = n_tot * (t + 1) / nt;
        int64_t *h = ws->hist + (int64_t)t * (nc + 1);
        memset(h, 0, sizeof(int64_t) * (size_t)(nc + 1));
        for (int64_t i = lo; i < hi; ++i) {
            const int64_t id = pv->index[i];
            const int64_t c = mzo_get_cell(g, pv->x[3 * (id - 1)]);
            pv->cell[i] = c;
            h[c] += 1;
        }
#pragma omp barrier
#pragma omp single
        {
            /* exclusive offsets in (cell, thread) order; PIA group 1 per cell */
            int64_t pos = 0;
            for (int64_t c = 1; c <= nc; ++c) {
                const int64_t start = pos;
                for (int u = 0; u < nt; ++u) {
                    int64_t *hu = ws->hist + (int64_t)u * (nc + 1);
                    const int64_t k = hu[c];
                    hu[c] = pos;
                    pos += k;
                }
                mzo_ind *d = mzo_pia_at(pia, c, 1);
                const int64_t np = pos - start;
                d->start2 = 0; d->end2 = -1; d->n_group2 = 0;
                if (np > 0) { d->start1 = start + 1; d->end1 = pos; }
                else { d->start1 = 0; d->end1 = -1; }
                d->n_group1 = np;
                d->n_local = np;
            }
        }
        for (int64_t i = lo; i < hi; ++i) ws->sorted[h[pv->cell[i]]++] = pv->index[i];
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n_tot; ++i) pv->index[i] = ws->sorted[i];
    }
    pia->contiguous[0] = 1;
}

void mzo_couette_steps(const mzo_rng *proto, const mzo_grid1d *g, const mzo_walls *wl,
                       const mzo_interaction *it, mzo_cf *cfs, mzo_pv *pv, mzo_pia *pia, double dt,
                       const double *V, double mass, mzo_props *pp, int64_t n_steps, uint64_t step0,
                       int32_t equal_weight, int32_t nthreads, int32_t props_every) {
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    sort_ws ws;
    ws.nt = nt;
    ws.cap = pv->length > 0 ? pv->length : 1;
    ws.hist = (int64_t *)malloc(sizeof(int64_t) * (size_t)nt * (size_t)(g->n_cells + 1));
    ws.sorted = (int64_t *)malloc(sizeof(int64_t) * (size_t)ws.cap);
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
        if (pia->contiguous[0] && pia->n_total[0] <= ws.cap && nt > 1) sort_parallel(g, pv, pia, &ws);
        else mzo_sort_particles(g, pv, pia, 1);
        if (props_every > 0 && ((t + 1) % props_every) == 0) {
            mzo_pv *pvs[1] = {pv};
#pragma omp parallel for num_threads(nt) schedule(static)
            for (int64_t c = 1; c <= g->n_cells; ++c)
                mzo_compute_props_sorted(pvs, pia, &mass, pp, NULL, c, c);
        }
    }
    free(ws.hist);
    free(ws.sorted);
}
