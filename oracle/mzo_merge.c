/*
 * mzo_merge.c -- oracle: octree N:2 merging, Lawson-Hanson NNLS and
 * NNLS moment-preserving merging.
 *
 * TEST INFRASTRUCTURE ONLY (see mzo.h).
 */
#include "mzo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PV_W(pv, id) ((pv)->w[(id) - 1])
#define PV_V(pv, id) (&(pv)->v[3 * ((id) - 1)])
#define PV_X(pv, id) (&(pv)->x[3 * ((id) - 1)])
#define PID(pv, pos) ((pv)->index[(pos) - 1])

/* ========================================================================= */
/* Octree N:2 (src/merging/merging_octree_N2.jl).  Bins are 1-based; the
 * particle lists hold logical positions (1-based), as in the reference. */

mzo_octree *mzo_octree_new(int32_t split, int32_t init_bin_bounds, int32_t bin_bounds_compute,
                           int64_t max_Nbins, int64_t max_depth) {
    mzo_octree *oc = (mzo_octree *)calloc(1, sizeof(mzo_octree));
    oc->max_Nbins = max_Nbins;
    oc->bins = (mzo_obin *)calloc((size_t)max_Nbins + 1, sizeof(mzo_obin));
    oc->full = (mzo_ofull *)calloc((size_t)max_Nbins + 1, sizeof(mzo_ofull));
    for (int64_t i = 0; i <= max_Nbins; ++i) oc->bins[i].can_be_refined = 1;
    oc->bin_start = (int64_t *)calloc((size_t)max_Nbins + 1, sizeof(int64_t));
    oc->bin_end = (int64_t *)calloc((size_t)max_Nbins + 1, sizeof(int64_t));
    oc->buf_len = 8192;
    oc->pis = (int64_t *)calloc((size_t)oc->buf_len, sizeof(int64_t));
    oc->poct = (int64_t *)calloc((size_t)oc->buf_len, sizeof(int64_t));
    oc->psort = (int64_t *)calloc((size_t)oc->buf_len, sizeof(int64_t));
    oc->split = split;
    oc->init_bin_bounds = init_bin_bounds;
    oc->bin_bounds_compute = bin_bounds_compute;
    oc->max_depth = max_depth;
    return oc;
}

void mzo_octree_free(mzo_octree *oc) {
    if (!oc) return;
    free(oc->bins); free(oc->full); free(oc->bin_start); free(oc->bin_end);
    free(oc->pis); free(oc->poct); free(oc->psort); free(oc);
}

/* resize_octree_buffers! (:270-280) */
static void octree_resize(mzo_octree *oc, int64_t n) {
    if (oc->buf_len >= n) return;
    int64_t nl = n + MZO_DELTA_PARTICLES;
    oc->pis = (int64_t *)realloc(oc->pis, sizeof(int64_t) * nl);
    oc->poct = (int64_t *)realloc(oc->poct, sizeof(int64_t) * nl);
    oc->psort = (int64_t *)realloc(oc->psort, sizeof(int64_t) * nl);
    oc->buf_len = nl;
}

/* compute_octant (:304-316): strict > comparisons */
static int64_t compute_octant(const double *v, const double *m) {
    int64_t o = 1;
    if (v[0] > m[0]) o += 1;
    if (v[1] > m[1]) o += 2;
    if (v[2] > m[2]) o += 4;
    return o;
}

/* bin_bounds_inherit! (:341-367) */
static void bin_bounds_inherit(mzo_octree *oc, int64_t b, const double *pmin, const double *pmax,
                               const double *mid, int64_t oct) {
    mzo_obin *bn = &oc->bins[b];
    int sx = (oct - 1) & 1, sy = ((oct - 1) >> 1) & 1, sz = ((oct - 1) >> 2) & 1;
    bn->v_min[0] = sx ? mid[0] : pmin[0]; bn->v_max[0] = sx ? pmax[0] : mid[0];
    bn->v_min[1] = sy ? mid[1] : pmin[1]; bn->v_max[1] = sy ? pmax[1] : mid[1];
    bn->v_min[2] = sz ? mid[2] : pmin[2]; bn->v_max[2] = sz ? pmax[2] : mid[2];
}

/* bin_bounds_recompute! (:382-418): bounds are only written when the bin
 * holds at least one particle. */
static void bin_bounds_recompute(mzo_octree *oc, int64_t b, int64_t bs, int64_t be, mzo_pv *pv) {
    double mn[3] = {9299792458.0, 9299792458.0, 9299792458.0};
    double mx[3] = {-9299792458.0, -9299792458.0, -9299792458.0};
    for (int64_t i = bs; i <= be; ++i) {
        const double *v = PV_V(pv, PID(pv, oc->pis[i - 1]));
        for (int d = 0; d < 3; ++d) {
            if (v[d] < mn[d]) mn[d] = v[d];
            if (v[d] > mx[d]) mx[d] = v[d];
        }
        memcpy(oc->bins[b].v_min, mn, sizeof(mn));
        memcpy(oc->bins[b].v_max, mx, sizeof(mx));
    }
}

/* compute_v_mean! (:431-439).  vel_middle is not reset before the sum: the
 * previous split's middle velocity is carried into the new one, exactly as
 * the reference does. */
static void compute_v_mean(mzo_octree *oc, int64_t bs, int64_t be, mzo_pv *pv) {
    double n_tot = 0.0;
    for (int64_t i = bs; i <= be; ++i) {
        int64_t id = PID(pv, oc->pis[i - 1]);
        n_tot += PV_W(pv, id);
        for (int d = 0; d < 3; ++d) oc->vel_middle[d] = oc->vel_middle[d] + PV_W(pv, id) * PV_V(pv, id)[d];
    }
    for (int d = 0; d < 3; ++d) oc->vel_middle[d] = oc->vel_middle[d] / n_tot;
}

static int64_t post_merge_np(const mzo_octree *oc, int64_t b) {
    return oc->bins[b].np >= 2 ? 2 : oc->bins[b].np;
}

static int64_t new_bin_id(int64_t i, int64_t bin_id, int64_t Nbins) {
    return i == 1 ? bin_id : Nbins + i - 1;
}

/* split_bin! (:503-633) */
void mzo_octree_split_bin(mzo_octree *oc, int64_t bin_id, mzo_pv *pv) {
    int64_t cnt[9] = {0}, nonempty_counter[9] = {0}, nonempty_bins[9] = {0};
    double ndens[9] = {0};
    int64_t n_ne = 0;
    int64_t depth = oc->bins[bin_id].depth;
    int64_t bs = oc->bin_start[bin_id], be = oc->bin_end[bin_id];
    if (oc->bin_bounds_compute == 2) bin_bounds_recompute(oc, bin_id, bs, be, pv);
    if (oc->split == 1) {
        for (int d = 0; d < 3; ++d) oc->vel_middle[d] = 0.5 * (oc->bins[bin_id].v_min[d] + oc->bins[bin_id].v_max[d]);
    } else {
        compute_v_mean(oc, bs, be, pv);
    }
    for (int64_t i = bs; i <= be; ++i) {
        int64_t id = PID(pv, oc->pis[i - 1]);
        int64_t o = compute_octant(PV_V(pv, id), oc->vel_middle);
        cnt[o] += 1;
        oc->poct[i - bs] = o;
        ndens[o] += PV_W(pv, id);
    }
    int64_t n_eb = 0;
    if (cnt[1] > 0) { n_ne += 1; n_eb += 1; nonempty_counter[n_eb] = cnt[1]; nonempty_bins[n_eb] = 1; }
    for (int i = 2; i <= 8; ++i) {
        if (cnt[i] > 0) { n_ne += 1; n_eb += 1; nonempty_counter[n_eb] = cnt[i]; nonempty_bins[n_eb] = i; }
        cnt[i] += cnt[i - 1];
    }
    oc->bin_end[bin_id] = oc->bin_start[bin_id] + nonempty_counter[1] - 1;
    for (int64_t i = 2; i <= n_ne; ++i) {
        int64_t bi = new_bin_id(i, bin_id, oc->Nbins), bim1 = new_bin_id(i - 1, bin_id, oc->Nbins);
        oc->bin_start[bi] = oc->bin_end[bim1] + 1;
        oc->bin_end[bi] = oc->bin_start[bi] + nonempty_counter[i] - 1;
    }
    oc->total_post_merge_np -= 2;
    if (oc->bin_bounds_compute == 1) {
        memcpy(oc->v_min_parent, oc->bins[bin_id].v_min, sizeof(oc->v_min_parent));
        memcpy(oc->v_max_parent, oc->bins[bin_id].v_max, sizeof(oc->v_max_parent));
    }
    for (int64_t i = 1; i <= n_ne; ++i) {
        int64_t bi = new_bin_id(i, bin_id, oc->Nbins);
        if (oc->bin_bounds_compute == 1)
            bin_bounds_inherit(oc, bi, oc->v_min_parent, oc->v_max_parent, oc->vel_middle, nonempty_bins[i]);
        oc->bins[bi].np = nonempty_counter[i];
        oc->bins[bi].w = ndens[nonempty_bins[i]];
        oc->bins[bi].depth = depth + 1;
        oc->total_post_merge_np += post_merge_np(oc, bi);
        oc->bins[bi].can_be_refined = (oc->bins[bi].np > 2) && (oc->bins[bi].depth < oc->max_depth);
    }
    for (int64_t i = bs; i <= be; ++i) {
        int64_t pin = oc->pis[i - 1];
        int64_t j = oc->poct[i - bs];
        oc->psort[cnt[j] - 1] = pin;
        cnt[j] -= 1;
    }
    for (int64_t i = bs; i <= be; ++i) oc->pis[i - 1] = oc->psort[i - bs];
    oc->Nbins += n_ne - 1;
}

/* compute_bin_props! (:646-700) */
void mzo_octree_compute_bin_props(mzo_octree *oc, int64_t b, mzo_pv *pv) {
    int64_t bs = oc->bin_start[b], be = oc->bin_end[b];
    if (oc->bins[b].w == 0) { oc->bins[b].np = 0; return; }
    if (oc->bins[b].np == 1) {
        oc->full[b].particle_index1 = oc->pis[bs - 1];
    } else if (oc->bins[b].np >= 2) {
        oc->full[b].particle_index1 = oc->pis[bs - 1];
        oc->full[b].particle_index2 = oc->pis[bs];
    }
    if (oc->bins[b].np <= 2) return;
    double vm[3] = {0, 0, 0}, vs[3] = {0, 0, 0}, xm[3] = {0, 0, 0}, xs[3] = {0, 0, 0};
    double inv_w = 1.0 / oc->bins[b].w;
    for (int64_t i = bs; i <= be; ++i) {
        int64_t id = PID(pv, oc->pis[i - 1]);
        double w = PV_W(pv, id);
        for (int d = 0; d < 3; ++d) {
            vm[d] = vm[d] + w * PV_V(pv, id)[d];
            xm[d] = xm[d] + w * PV_X(pv, id)[d];
        }
    }
    for (int d = 0; d < 3; ++d) { vm[d] = vm[d] * inv_w; xm[d] = xm[d] * inv_w; }
    for (int64_t i = bs; i <= be; ++i) {
        int64_t id = PID(pv, oc->pis[i - 1]);
        double w = PV_W(pv, id);
        for (int d = 0; d < 3; ++d) {
            double a = PV_V(pv, id)[d] - vm[d], c = PV_X(pv, id)[d] - xm[d];
            vs[d] = vs[d] + w * (a * a);
            xs[d] = xs[d] + w * (c * c);
        }
    }
    for (int d = 0; d < 3; ++d) {
        oc->full[b].v_mean[d] = vm[d];
        oc->full[b].v_std_sq[d] = vs[d] * inv_w;
        oc->full[b].x_mean[d] = xm[d];
        oc->full[b].x_std_sq[d] = xs[d] * inv_w;
    }
}

/* init_octree! (:947-984) */
void mzo_octree_init(mzo_octree *oc, mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species) {
    mzo_ind *d = mzo_pia_at(pia, cell, species);
    oc->Nbins = 1;
    for (int64_t i = 0; i < d->n_group1; ++i) oc->pis[i] = d->start1 + i;
    if (d->n_group2 > 0)
        for (int64_t i = 0; i < d->n_group2; ++i) oc->pis[d->n_group1 + i] = d->start2 + i;
    oc->n_particles = d->n_local;
    oc->bins[1].depth = 0;
    oc->bin_start[1] = 1;
    oc->bin_end[1] = oc->n_particles;
    oc->bins[1].np = oc->n_particles;
    oc->bins[1].w = 1e50;
    oc->total_post_merge_np = post_merge_np(oc, 1);
    oc->bins[1].can_be_refined = (oc->bins[1].np > 2) && (oc->max_depth > 0);
    if (oc->init_bin_bounds == 3) {
        for (int k = 0; k < 3; ++k) { oc->bins[1].v_min[k] = -299792458.0; oc->bins[1].v_max[k] = 299792458.0; }
    } else {
        bin_bounds_recompute(oc, 1, 1, oc->n_particles, pv);
        if (oc->init_bin_bounds == 2) {
            for (int k = 0; k < 3; ++k) {
                double m = fmax(fabs(oc->bins[1].v_min[k]), fabs(oc->bins[1].v_max[k]));
                oc->bins[1].v_min[k] = -m;
                oc->bins[1].v_max[k] = m;
            }
        }
    }
}

/* compute_octree! (:998-1039): refine the heaviest refinable bin (first
 * maximum over 1..Nbins) until the post-merge count would exceed the target,
 * nothing is refinable or the bin budget runs out. */
void mzo_octree_compute(mzo_octree *oc, mzo_pv *pv, int64_t target_np) {
    for (;;) {
        int64_t refine_id = -1;
        double max_w = -1.0;
        for (int64_t b = 1; b <= oc->Nbins; ++b) {
            if ((oc->bins[b].w > max_w) && oc->bins[b].can_be_refined) {
                max_w = oc->bins[b].w;
                refine_id = b;
            }
        }
        if (refine_id == -1) break;
        else if (oc->total_post_merge_np + 14 > target_np) break;
        else mzo_octree_split_bin(oc, refine_id, pv);
        if (oc->Nbins + 7 > oc->max_Nbins) break;
    }
    if (oc->Nbins == 1) {
        oc->bins[1].w = 0.0;
        for (int64_t ii = 1; ii <= oc->n_particles; ++ii) oc->bins[1].w += PV_W(pv, PID(pv, oc->pis[ii - 1]));
    }
    for (int64_t b = 1; b <= oc->Nbins; ++b) mzo_octree_compute_bin_props(oc, b, pv);
}

/* One sign drawn from direction_signs = [-1, 1] (src/constants.jl:39).  The
 * reference draws an integer index; here one uniform u gives -1 if u < 0.5. */
static double draw_sign(mzo_rng *rng) { return mzo_rand(rng) < 0.5 ? -1.0 : 1.0; }

/* compute_new_particles! (:736-813) and its grid-clamped variant (:830-933) */
static void octree_new_particles(mzo_rng *rng, mzo_octree *oc, mzo_pv *pv, mzo_pia *pia, int64_t cell,
                                 int64_t species, const mzo_grid1d *grid) {
    for (int64_t b = 1; b <= oc->Nbins; ++b) {
        int64_t lnp = oc->bins[b].np;
        mzo_ofull *f = &oc->full[b];
        if (lnp > 2) {
            f->w1 = 0.5 * oc->bins[b].w;
            f->w2 = f->w1;
            for (int d = 0; d < 3; ++d) { f->v_std_sq[d] = sqrt(f->v_std_sq[d]); f->x_std_sq[d] = sqrt(f->x_std_sq[d]); }
            for (int d = 0; d < 3; ++d) oc->direction_vec[d] = draw_sign(rng);
            for (int d = 0; d < 3; ++d) {
                f->v1[d] = f->v_mean[d] + oc->direction_vec[d] * f->v_std_sq[d];
                f->v2[d] = f->v_mean[d] - oc->direction_vec[d] * f->v_std_sq[d];
            }
            for (int d = 0; d < 3; ++d) oc->direction_vec[d] = draw_sign(rng);
            for (int d = 0; d < 3; ++d) {
                f->x1[d] = f->x_mean[d] + oc->direction_vec[d] * f->x_std_sq[d];
                f->x2[d] = f->x_mean[d] - oc->direction_vec[d] * f->x_std_sq[d];
            }
        } else if (lnp == 2) {
            int64_t id = PID(pv, f->particle_index1);
            f->w1 = PV_W(pv, id); memcpy(f->v1, PV_V(pv, id), 24); memcpy(f->x1, PV_X(pv, id), 24);
            id = PID(pv, f->particle_index2);
            f->w2 = PV_W(pv, id); memcpy(f->v2, PV_V(pv, id), 24); memcpy(f->x2, PV_X(pv, id), 24);
        } else if (lnp == 1) {
            int64_t id = PID(pv, f->particle_index1);
            f->w1 = PV_W(pv, id); memcpy(f->v1, PV_V(pv, id), 24); memcpy(f->x1, PV_X(pv, id), 24);
        }
    }
    mzo_ind *d = mzo_pia_at(pia, cell, species);
    int64_t cur = 0;
    for (int64_t b = 1; b <= oc->Nbins; ++b) {
        int64_t lnp = oc->bins[b].np;
        mzo_ofull *f = &oc->full[b];
        if (lnp >= 2) {
            for (int half = 0; half < 2; ++half) {
                int64_t id = PID(pv, mzo_map_cont_index(d, cur));
                cur += 1;
                const double *x = half == 0 ? f->x1 : f->x2;
                PV_W(pv, id) = half == 0 ? f->w1 : f->w2;
                memcpy(PV_V(pv, id), half == 0 ? f->v1 : f->v2, 24);
                memcpy(PV_X(pv, id), x, 24);
                if (grid && lnp > 2) {
                    if (x[0] < grid->min_x) PV_X(pv, id)[0] = grid->min_x;
                    else if (x[0] > grid->max_x) PV_X(pv, id)[0] = grid->max_x;
                }
            }
        } else if (lnp == 1) {
            int64_t id = PID(pv, mzo_map_cont_index(d, cur));
            cur += 1;
            PV_W(pv, id) = f->w1;
            memcpy(PV_V(pv, id), f->v1, 24);
            memcpy(PV_X(pv, id), f->x1, 24);
        }
    }
    int64_t n_del = d->n_local - cur;
    if (!(cell == pia->n_cells) || (n_del > d->n_group2)) pia->contiguous[species - 1] = 0;
    for (int64_t i = 0; i < n_del; ++i) mzo_delete_particle_end(pv, pia, cell, species);
}

/* merge_octree_N2_based! (:1060-1094); grid may be NULL */
void mzo_merge_octree_N2_based(mzo_rng *rng, mzo_octree *oc, mzo_pv *pv, mzo_pia *pia, int64_t cell,
                               int64_t species, int64_t target_np, const mzo_grid1d *grid) {
    oc->Nbins = 0;
    octree_resize(oc, mzo_pia_at(pia, cell, species)->n_local);
    mzo_rng_set_stream(rng, (uint64_t)(cell - 1));
    mzo_octree_init(oc, pv, pia, cell, species);
    mzo_octree_compute(oc, pv, target_np);
    octree_new_particles(rng, oc, pv, pia, cell, species, grid);
}

/* ========================================================================= */
/* Lawson-Hanson NNLS (src/merging/nnls.jl:26-548, after Lawson & Hanson,
 * "Solving Least Squares Problems", 1974).  A is m x n column-major and is
 * overwritten; b is overwritten.  Returns mode (1 ok, 3 iteration cap). */

#define A_(r, c) A[((c) - 1) * m + ((r) - 1)]

static double construct_householder(double *u, int64_t len, double up) {
    if (len <= 1) return up;
    double cl = 0.0;
    for (int64_t i = 0; i < len; ++i) if (fabs(u[i]) > cl) cl = fabs(u[i]);
    double clinv = 1.0 / cl;
    double sm = 0.0;
    for (int64_t i = 0; i < len; ++i) { double t = u[i] * clinv; sm += t * t; }
    cl *= sqrt(sm);
    if (u[0] > 0) cl = -cl;
    double result = u[0] - cl;
    u[0] = cl;
    return result;
}

static void apply_householder(const double *u, int64_t len, double up, double *c) {
    if (len > 1) {
        double b = up * u[0];
        if (b >= 0) return;
        b = 1.0 / b;
        double sm = c[0] * up;
        for (int64_t i = 1; i < len; ++i) sm = sm + c[i] * u[i];
        if (sm != 0) {
            sm *= b;
            c[0] = c[0] + sm * up;
            for (int64_t i = 1; i < len; ++i) c[i] = c[i] + sm * u[i];
        }
    }
}

static void orthogonal_rotmat(double a, double b, double *c, double *s, double *sig) {
    if (fabs(a) > fabs(b)) {
        double xr = b / a, yr = sqrt(1 + xr * xr);
        *c = (1 / yr) * (a > 0 ? 1.0 : (a < 0 ? -1.0 : 0.0));
        *s = *c * xr;
        *sig = fabs(a) * yr;
    } else if (b != 0) {
        double xr = a / b, yr = sqrt(1 + xr * xr);
        *s = (1 / yr) * (b > 0 ? 1.0 : -1.0);
        *c = *s * xr;
        *sig = fabs(b) * yr;
    } else {
        *sig = 0.0; *c = 0.0; *s = 1.0;
    }
}

static int64_t solve_triangular(double *zz, const double *A, int64_t m, const int64_t *idx,
                                int64_t nsetp, int64_t jj) {
    for (int64_t l = 1; l <= nsetp; ++l) {
        int64_t ip = nsetp + 1 - l;
        if (l != 1)
            for (int64_t ii = 1; ii <= ip; ++ii) zz[ii - 1] = zz[ii - 1] - A_(ii, jj) * zz[ip];
        jj = idx[ip - 1];
        zz[ip - 1] /= A_(ip, jj);
    }
    return jj;
}

int mzo_nnls_solve(double *A, int64_t m, int64_t n, double *b, double *x, double *w, double *zz,
                   int64_t *idx, int64_t max_iter, double kkt_tol, double *rnorm, int64_t *nsetp_out) {
    const double factor = 0.01;
    int mode = 1;
    int64_t iter = 0;
    for (int64_t i = 1; i <= n; ++i) { x[i - 1] = 0.0; idx[i - 1] = i; w[i - 1] = 0.0; }
    int64_t iz2 = n, iz1 = 1, iz = 0, j = 0, jj = 0, nsetp = 0;
    double up = 0.0;
    int terminated = 0;
    for (;;) {
        if (iz1 > iz2 || nsetp >= m) { terminated = 1; break; }
        for (int64_t i = iz1; i <= iz2; ++i) {
            int64_t idxi = idx[i - 1];
            double sm = 0.0;
            for (int64_t l = nsetp + 1; l <= m; ++l) sm = sm + A_(l, idxi) * b[l - 1];
            w[idxi - 1] = sm;
        }
        for (;;) {
            double wmax = 0.0;
            int64_t izmax = 0;
            for (int64_t i = iz1; i <= iz2; ++i) {
                int64_t jx = idx[i - 1];
                if (w[jx - 1] > wmax) { wmax = w[jx - 1]; izmax = i; }
            }
            if (wmax <= kkt_tol) { terminated = 1; break; }
            iz = izmax;
            j = idx[iz - 1];
            double Asave = A_(nsetp + 1, j);
            up = construct_householder(&A_(nsetp + 1, j), m - nsetp, up);
            double unorm = 0.0;
            for (int64_t l = 1; l <= nsetp; ++l) unorm += A_(l, j) * A_(l, j);
            unorm = sqrt(unorm);
            if (((unorm + fabs(A_(nsetp + 1, j)) * factor) - unorm) > 0) {
                for (int64_t i = 0; i < m; ++i) zz[i] = b[i];
                apply_householder(&A_(nsetp + 1, j), m - nsetp, up, &zz[nsetp]);
                double ztest = zz[nsetp] / A_(nsetp + 1, j);
                if (ztest > 0) break;
            }
            A_(nsetp + 1, j) = Asave;
            w[j - 1] = 0;
        }
        if (terminated) break;
        for (int64_t i = 0; i < m; ++i) b[i] = zz[i];
        idx[iz - 1] = idx[iz1 - 1];
        idx[iz1 - 1] = j;
        iz1 += 1;
        nsetp += 1;
        if (iz1 <= iz2) {
            for (int64_t jz = iz1; jz <= iz2; ++jz) {
                jj = idx[jz - 1];
                apply_householder(&A_(nsetp, j), m - nsetp + 1, up, &A_(nsetp, jj));
            }
        }
        if (nsetp != m)
            for (int64_t l = nsetp + 1; l <= m; ++l) A_(l, j) = 0;
        w[j - 1] = 0;
        jj = solve_triangular(zz, A, m, idx, nsetp, jj);
        for (;;) {
            iter += 1;
            if (iter > max_iter) { mode = 3; terminated = 1; break; }
            double alpha = 2.0;
            for (int64_t ip = 1; ip <= nsetp; ++ip) {
                int64_t l = idx[ip - 1];
                if (zz[ip - 1] <= 0) {
                    double t = -x[l - 1] / (zz[ip - 1] - x[l - 1]);
                    if (alpha > t) { alpha = t; jj = ip; }
                }
            }
            if (alpha == 2) break;
            for (int64_t ip = 1; ip <= nsetp; ++ip) {
                int64_t l = idx[ip - 1];
                x[l - 1] = x[l - 1] + alpha * (zz[ip - 1] - x[l - 1]);
            }
            int64_t i = idx[jj - 1];
            for (;;) {
                x[i - 1] = 0;
                if (jj != nsetp) {
                    jj += 1;
                    for (int64_t jx = jj; jx <= nsetp; ++jx) {
                        int64_t ii = idx[jx - 1];
                        idx[jx - 2] = ii;
                        double cc, ss, sig;
                        orthogonal_rotmat(A_(jx - 1, ii), A_(jx, ii), &cc, &ss, &sig);
                        A_(jx - 1, ii) = sig;
                        A_(jx, ii) = 0;
                        for (int64_t l = 1; l <= n; ++l) {
                            if (l != ii) {
                                double temp = A_(jx - 1, l);
                                A_(jx - 1, l) = cc * temp + ss * A_(jx, l);
                                A_(jx, l) = -ss * temp + cc * A_(jx, l);
                            }
                        }
                        double temp = b[jx - 2];
                        b[jx - 2] = cc * temp + ss * b[jx - 1];
                        b[jx - 1] = -ss * temp + cc * b[jx - 1];
                    }
                }
                nsetp -= 1;
                iz1 -= 1;
                idx[iz1 - 1] = i;
                int allfeasible = 1;
                for (int64_t jji = 1; jji <= nsetp; ++jji) {
                    i = idx[jji - 1];
                    if (x[i - 1] <= 0) { allfeasible = 0; break; }
                }
                if (allfeasible) break;
            }
            for (int64_t q = 0; q < m; ++q) zz[q] = b[q];
            jj = solve_triangular(zz, A, m, idx, nsetp, jj);
        }
        if (terminated) break;
        for (int64_t i = 1; i <= nsetp; ++i) x[idx[i - 1] - 1] = zz[i - 1];
    }
    double sm = 0.0;
    if (nsetp < m)
        for (int64_t i = nsetp + 1; i <= m; ++i) sm += b[i - 1] * b[i - 1];
    *rnorm = sqrt(sm);
    *nsetp_out = nsetp;
    return mode;
}
#undef A_

/* scale_columns! (src/utils.jl:62-77) */
void mzo_scale_columns(double *A, int64_t m, int64_t n, double *column_norms) {
    for (int64_t i = 0; i < n; ++i) {
        double nn = 0.0;
        for (int64_t j = 0; j < m; ++j) nn = nn + A[i * m + j] * A[i * m + j];
        nn = sqrt(nn);
        nn = nn > 1e-15 ? nn : 1.0;
        column_norms[i] = 1.0 / nn;
        for (int64_t j = 0; j < m; ++j) A[i * m + j] *= column_norms[i];
    }
}

/* ========================================================================= */
/* NNLS merging (src/merging/merging_nnls.jl) */

/* NNLSMerge constructor (:143-206): base moments (mass, momentum,
 * directional energy) first, then the requested multi-indices that are not
 * already there. */
void mzo_nnls_merge_init(mzo_nnls_merge *nm, const int32_t *mim, int64_t n_mim, const int32_t *mim_pos,
                         int64_t n_mim_pos) {
    static const int32_t base[7][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2}};
    memset(nm, 0, sizeof(*nm));
    int64_t n = 0;
    for (int i = 0; i < 7; ++i) { memcpy(nm->mim[n], base[i], sizeof(base[i])); n++; }
    for (int64_t i = 0; i < n_mim; ++i) {
        int dup = 0;
        for (int b = 0; b < 7; ++b)
            if (mim[3 * i] == base[b][0] && mim[3 * i + 1] == base[b][1] && mim[3 * i + 2] == base[b][2]) dup = 1;
        if (!dup) { memcpy(nm->mim[n], &mim[3 * i], 3 * sizeof(int32_t)); n++; }
    }
    nm->n_moments_vel = n;
    for (int64_t i = 0; i < n; ++i) nm->tot_order[i] = nm->mim[i][0] + nm->mim[i][1] + nm->mim[i][2];
    nm->n_moments_pos = n_mim_pos;
    nm->pos_i_x = nm->pos_i_y = nm->pos_i_z = -1;
    for (int64_t i = 0; i < n_mim_pos; ++i) {
        memcpy(nm->mim_pos[i], &mim_pos[3 * i], 3 * sizeof(int32_t));
        const int32_t *q = &mim_pos[3 * i];
        if (q[0] == 1 && q[1] == 0 && q[2] == 0) nm->pos_i_x = i + 1;
        else if (q[0] == 0 && q[1] == 1 && q[2] == 0) nm->pos_i_y = i + 1;
        else if (q[0] == 0 && q[1] == 0 && q[2] == 1) nm->pos_i_z = i + 1;
    }
    nm->n_total_conserved = n + n_mim_pos;
    nm->vref = 1.0;
    nm->inv_vref = 1.0;
}

static double ipow(double a, int p) {
    double r = 1.0;
    for (int i = 0; i < p; ++i) r *= a;
    return r;
}

/* ccm (:464-487): (v-v0)^mim, with Julia's x^0 == 1 */
static double ccm3(double vx, double vy, double vz, double x0, double y0, double z0, const int32_t *mm) {
    return ipow(vx - x0, mm[0]) * ipow(vy - y0, mm[1]) * ipow(vz - z0, mm[2]);
}

static double vsgn_x(int o) { return (o % 2 == 1) ? -1.0 : 1.0; }
static double vsgn_y(int o) { return (o == 3 || o == 4 || o == 7 || o == 8) ? 1.0 : -1.0; }
static double vsgn_z(int o) { return o >= 5 ? 1.0 : -1.0; }

/* check_speed_bound with multiplier (:309-318) */
static double speed_bound(double s, double mn, double mx, double mult) {
    double ms = mult * s;
    if (ms > mx) return mx * mult;
    else if (ms < mn) return mn * mult;
    return ms;
}

/* merge_nnls_based! (:1457-1510) with compute_w_total_v0! (:367-449),
 * compute_lhs_and_rhs! (:509-613), compute_lhs_particles_additional!
 * (:949-1038), scale_lhs_rhs! (:1178-1276), scale_columns! and
 * compute_post_merge_particles_nnls! (:1334-1404).  Returns 1 or -1. */
int mzo_merge_nnls_based(mzo_rng *rng, mzo_nnls_merge *nm, mzo_pv *pv, mzo_pia *pia, int64_t cell,
                         int64_t species, double vref, int32_t scaling_vref, int64_t n_rand_pairs,
                         double max_err, int32_t centered_at_mean, const double *v_multipliers,
                         int64_t n_vmult, int64_t iteration_mult, double w_threshold) {
    return mzo_merge_nnls_cand(rng, nm, pv, pia, cell, species, vref, scaling_vref, n_rand_pairs, max_err,
                               centered_at_mean, v_multipliers, n_vmult, iteration_mult, w_threshold, -1);
}

/* The same merge with at most max_pcols particle columns (the first ones in
 * logical order; max_pcols < 0: all, which is merge_nnls_based! itself).
 * Every particle still enters the moments (the right-hand side). */
int mzo_merge_nnls_cand(mzo_rng *rng, mzo_nnls_merge *nm, mzo_pv *pv, mzo_pia *pia, int64_t cell,
                        int64_t species, double vref, int32_t scaling_vref, int64_t n_rand_pairs,
                        double max_err, int32_t centered_at_mean, const double *v_multipliers,
                        int64_t n_vmult, int64_t iteration_mult, double w_threshold, int64_t max_pcols) {
    mzo_ind *d = mzo_pia_at(pia, cell, species);
    int64_t n_add = (centered_at_mean ? 1 : 0) + 8 * n_vmult;
    int64_t npc = (max_pcols >= 0 && max_pcols < d->n_local) ? max_pcols : d->n_local;
    int64_t ncols = npc + n_add + n_rand_pairs;
    int64_t m = nm->n_total_conserved, nmv = nm->n_moments_vel;
    double *A = (double *)calloc((size_t)(m * ncols), sizeof(double));
    double *vp = (double *)calloc((size_t)(6 * ncols), sizeof(double));
    double *cn = (double *)malloc(sizeof(double) * ncols);
    double *x = (double *)malloc(sizeof(double) * ncols);
    double *w = (double *)malloc(sizeof(double) * ncols);
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * ncols);
    double zz[128], b[128];
    mzo_rng_set_stream(rng, (uint64_t)(cell - 1));
    nm->vref = vref;
    nm->inv_vref = 1.0 / vref;

    /* compute_w_total_v0! */
    nm->Ex = nm->Ey = nm->Ez = nm->Epx = nm->Epy = nm->Epz = 0.0;
    for (int64_t i = 0; i < m; ++i) nm->rhs[i] = 0.0;
    for (int k = 0; k < 3; ++k) nm->v0[k] = 0.0;
    if (nm->reset_x0) for (int k = 0; k < 3; ++k) nm->x0[k] = 0.0;
    nm->w_total = 0.0;
    nm->minvx = MZO_C_LIGHT; nm->maxvx = -MZO_C_LIGHT;
    nm->minvy = MZO_C_LIGHT; nm->maxvy = -MZO_C_LIGHT;
    nm->minvz = MZO_C_LIGHT; nm->maxvz = -MZO_C_LIGHT;
    for (int grp = 1; grp <= 2; ++grp) {
        if (grp == 2 && d->n_group2 <= 0) break;
        int64_t st = grp == 1 ? d->start1 : d->start2, en = grp == 1 ? d->end1 : d->end2;
        for (int64_t i = st; i <= en; ++i) {
            int64_t id = PID(pv, i);
            double pw = PV_W(pv, id), *v = PV_V(pv, id), *px = PV_X(pv, id);
            nm->w_total += pw;
            for (int k = 0; k < 3; ++k) { nm->v0[k] = nm->v0[k] + pw * v[k]; nm->x0[k] = nm->x0[k] + pw * px[k]; }
            if (v[0] > nm->maxvx) nm->maxvx = v[0];
            if (v[0] < nm->minvx) nm->minvx = v[0];
            if (v[1] > nm->maxvy) nm->maxvy = v[1];
            if (v[1] < nm->minvy) nm->minvy = v[1];
            if (v[2] > nm->maxvz) nm->maxvz = v[2];
            if (v[2] < nm->minvz) nm->minvz = v[2];
        }
    }
    for (int k = 0; k < 3; ++k) { nm->v0[k] = nm->v0[k] / nm->w_total; nm->x0[k] = nm->x0[k] / nm->w_total; }
    nm->minvx -= nm->v0[0]; nm->minvy -= nm->v0[1]; nm->minvz -= nm->v0[2];
    nm->maxvx -= nm->v0[0]; nm->maxvy -= nm->v0[1]; nm->maxvz -= nm->v0[2];

    /* compute_lhs_and_rhs! */
    int64_t col = 1, li = 0;
    for (int grp = 1; grp <= 2; ++grp) {
        if (grp == 2 && d->n_group2 <= 0) break;
        int64_t st = grp == 1 ? d->start1 : d->start2, en = grp == 1 ? d->end1 : d->end2;
        for (int64_t i = st; i <= en; ++i) {
            int64_t id = PID(pv, i);
            double pw = PV_W(pv, id), *v = PV_V(pv, id), *px = PV_X(pv, id);
            double a;
            a = v[0] - nm->v0[0]; nm->Ex = nm->Ex + pw * (a * a);
            a = v[1] - nm->v0[1]; nm->Ey = nm->Ey + pw * (a * a);
            a = v[2] - nm->v0[2]; nm->Ez = nm->Ez + pw * (a * a);
            a = px[0] - nm->x0[0]; nm->Epx = nm->Epx + pw * (a * a);
            a = px[1] - nm->x0[1]; nm->Epy = nm->Epy + pw * (a * a);
            a = px[2] - nm->x0[2]; nm->Epz = nm->Epz + pw * (a * a);
            const int has_col = li < npc;
            if (has_col) {
                double *vpc = &vp[6 * (col - 1)];
                vpc[0] = v[0]; vpc[1] = v[1]; vpc[2] = v[2]; vpc[3] = px[0]; vpc[4] = px[1]; vpc[5] = px[2];
            }
            for (int64_t q = 0; q < nmv; ++q) {
                double t = ccm3(v[0], v[1], v[2], nm->v0[0], nm->v0[1], nm->v0[2], nm->mim[q]);
                nm->rhs[q] = nm->rhs[q] + pw * t;
                if (has_col) A[(col - 1) * m + q] = t;
            }
            for (int64_t q = 0; q < nm->n_moments_pos; ++q) {
                double t = ccm3(px[0], px[1], px[2], nm->x0[0], nm->x0[1], nm->x0[2], nm->mim_pos[q]);
                nm->rhs[nmv + q] = nm->rhs[nmv + q] + pw * t;
                if (has_col) A[(col - 1) * m + nmv + q] = t;
            }
            if (has_col) col += 1;
            li += 1;
        }
    }
    for (int64_t i = 0; i < m; ++i) nm->rhs[i] /= nm->w_total;
    nm->Ex = sqrt(nm->Ex / nm->w_total); nm->Ey = sqrt(nm->Ey / nm->w_total); nm->Ez = sqrt(nm->Ez / nm->w_total);
    nm->Epx = sqrt(nm->Epx / nm->w_total); nm->Epy = sqrt(nm->Epy / nm->w_total); nm->Epz = sqrt(nm->Epz / nm->w_total);

    /* compute_lhs_particles_additional!: the centre particle column is all
     * zeros (the reference writes 0.0 even for the mass row). */
    if (centered_at_mean) {
        double *vpc = &vp[6 * (col - 1)];
        vpc[0] = nm->v0[0]; vpc[1] = nm->v0[1]; vpc[2] = nm->v0[2];
        vpc[3] = nm->x0[0]; vpc[4] = nm->x0[1]; vpc[5] = nm->x0[2];
        col += 1;
    }
    for (int64_t vm = 0; vm < n_vmult; ++vm) {
        double vmul = v_multipliers[vm];
        for (int o = 1; o <= 8; ++o) {
            double vx = speed_bound(vsgn_x(o) * nm->Ex, nm->minvx, nm->maxvx, vmul);
            double vy = speed_bound(vsgn_y(o) * nm->Ey, nm->minvy, nm->maxvy, vmul);
            double vz = speed_bound(vsgn_z(o) * nm->Ez, nm->minvz, nm->maxvz, vmul);
            for (int64_t q = 0; q < nmv; ++q) A[(col - 1) * m + q] = ccm3(vx, vy, vz, 0.0, 0.0, 0.0, nm->mim[q]);
            double *vpc = &vp[6 * (col - 1)];
            vpc[0] = nm->v0[0] + vx; vpc[1] = nm->v0[1] + vy; vpc[2] = nm->v0[2] + vz;
            vpc[3] = nm->x0[0]; vpc[4] = nm->x0[1]; vpc[5] = nm->x0[2];
            col += 1;
        }
    }
    for (int64_t r = 0; r < n_rand_pairs; ++r) {
        int64_t i = (int64_t)floor(mzo_rand(rng) * (double)d->n_local);
        int64_t k = (int64_t)floor(mzo_rand(rng) * (double)d->n_local);
        i = mzo_map_cont_index(d, i);
        k = mzo_map_cont_index(d, k);
        if (i != k) {
            int64_t ii = PID(pv, i), kk = PID(pv, k);
            double vx = 0.5 * (PV_V(pv, ii)[0] + PV_V(pv, kk)[0]);
            double vy = 0.5 * (PV_V(pv, ii)[1] + PV_V(pv, kk)[1]);
            double vz = 0.5 * (PV_V(pv, ii)[2] + PV_V(pv, kk)[2]);
            for (int64_t q = 0; q < nmv; ++q)
                A[(col - 1) * m + q] = ccm3(vx, vy, vz, nm->v0[0], nm->v0[1], nm->v0[2], nm->mim[q]);
            double px = 0.5 * (PV_X(pv, ii)[0] + PV_X(pv, kk)[0]);
            double py = 0.5 * (PV_X(pv, ii)[1] + PV_X(pv, kk)[1]);
            double pz = 0.5 * (PV_X(pv, ii)[2] + PV_X(pv, kk)[2]);
            /* the reference indexes the spatial rows with the velocity
             * multi-indices here (merging_nnls.jl:1023-1026) */
            for (int64_t q = 0; q < nm->n_moments_pos; ++q)
                A[(col - 1) * m + nmv + q] = ccm3(px, py, pz, nm->x0[0], nm->x0[1], nm->x0[2], nm->mim[q]);
            double *vpc = &vp[6 * (col - 1)];
            vpc[0] = vx; vpc[1] = vy; vpc[2] = vz; vpc[3] = px; vpc[4] = py; vpc[5] = pz;
        }
        col += 1;
    }

    /* scale_lhs_rhs! */
    if (!scaling_vref) {
        double iex = nm->Ex > 1e-6 ? 1.0 / nm->Ex : nm->inv_vref;
        double iey = nm->Ey > 1e-6 ? 1.0 / nm->Ey : nm->inv_vref;
        double iez = nm->Ez > 1e-6 ? 1.0 / nm->Ez : nm->inv_vref;
        for (int64_t q = 0; q < nmv; ++q) {
            double rv = ipow(iex, nm->mim[q][0]) * ipow(iey, nm->mim[q][1]) * ipow(iez, nm->mim[q][2]);
            for (int64_t c = 0; c < ncols; ++c) A[c * m + q] *= rv;
            nm->rhs[q] *= rv;
        }
        nm->scalevx = 1.0 / iex; nm->scalevy = 1.0 / iey; nm->scalevz = 1.0 / iez;
    } else {
        for (int64_t q = 0; q < nmv; ++q) {
            double rv = ipow(nm->inv_vref, nm->tot_order[q]);
            for (int64_t c = 0; c < ncols; ++c) A[c * m + q] *= rv;
            nm->rhs[q] *= rv;
        }
        nm->scalevx = nm->scalevy = nm->scalevz = nm->vref;
    }
    {
        double ipx = nm->Epx > 1e-6 ? 1.0 / nm->Epx : 1.0;
        double ipy = nm->Epy > 1e-6 ? 1.0 / nm->Epy : 1.0;
        double ipz = nm->Epz > 1e-6 ? 1.0 / nm->Epz : 1.0;
        for (int64_t q = 0; q < nm->n_moments_pos; ++q) {
            double rv = ipow(ipx, nm->mim_pos[q][0]) * ipow(ipy, nm->mim_pos[q][1]) * ipow(ipz, nm->mim_pos[q][2]);
            for (int64_t c = 0; c < ncols; ++c) A[c * m + nmv + q] *= rv;
            nm->rhs[nmv + q] *= rv;
        }
        nm->scalex = 1.0 / ipx; nm->scaley = 1.0 / ipy; nm->scalez = 1.0 / ipz;
    }
    mzo_scale_columns(A, m, ncols, cn);
    for (int64_t i = 0; i < m; ++i) b[i] = nm->rhs[i];
    int mode = mzo_nnls_solve(A, m, ncols, b, x, w, zz, idx, iteration_mult * ncols, 1e-14, &nm->rnorm, &nm->nsetp);
    nm->mode = mode;
    nm->lhs_ncols = ncols;

    /* compute_post_merge_particles_nnls! */
    int ret = 1;
    if (nm->rnorm > max_err) {
        ret = -1;
    } else {
        int64_t nonzero = 0;
        double ndw = 0.0;
        for (int64_t i = 0; i < ncols; ++i) {
            x[i] = x[i] * cn[i];
            if (x[i] > w_threshold) { nonzero += 1; ndw += x[i]; }
        }
        if (nonzero >= d->n_local) {
            ret = -1;
        } else {
            nm->w_total = nm->w_total / ndw;
            int64_t cur = 0;
            for (int64_t j = 0; j < ncols; ++j) {
                if (x[j] > w_threshold) {
                    int64_t id = PID(pv, mzo_map_cont_index(d, cur));
                    cur += 1;
                    PV_W(pv, id) = x[j] * nm->w_total;
                    double *vpc = &vp[6 * j];
                    PV_V(pv, id)[0] = vpc[0]; PV_V(pv, id)[1] = vpc[1]; PV_V(pv, id)[2] = vpc[2];
                    PV_X(pv, id)[0] = nm->pos_i_x > 0 ? vpc[3] : nm->x0[0];
                    PV_X(pv, id)[1] = nm->pos_i_y > 0 ? vpc[4] : nm->x0[1];
                    PV_X(pv, id)[2] = nm->pos_i_z > 0 ? vpc[5] : nm->x0[2];
                }
            }
            int64_t n_del = d->n_local - cur;
            if (!(cell == pia->n_cells) || (n_del > d->n_group2)) pia->contiguous[species - 1] = 0;
            for (int64_t i = 0; i < n_del; ++i) mzo_delete_particle_end(pv, pia, cell, species);
        }
    }
    free(A); free(vp); free(cn); free(x); free(w); free(idx);
    return ret;
}

/* ========================================================================= */
/* merge_roulette! (src/merging/merging_roulette.jl:25-184).  No @muladd in
 * that file: sums and products are plain, in the reference's order (group 1
 * then group 2). */
static void roulette_sums(const mzo_pv *pv, const mzo_ind *d, const double *vm, double *w_sum, double *v_sum,
                          double *e_sum) {
    for (int g = 0; g < 2; ++g) {
        if (g == 1 && d->n_group2 <= 0) break;
        const int64_t s = g == 0 ? d->start1 : d->start2, e = g == 0 ? d->end1 : d->end2;
        for (int64_t i = s; i <= e; ++i) {
            const int64_t id = pv->index[i - 1];
            const double w = pv->w[id - 1];
            const double *v = pv->v + 3 * (id - 1);
            if (w_sum) *w_sum += w;
            if (v_sum) for (int k = 0; k < 3; ++k) v_sum[k] += w * v[k];
            if (e_sum) {
                const double a = v[0] - vm[0], b = v[1] - vm[1], c = v[2] - vm[2];
                *e_sum += w * ((a * a + b * b) + c * c);
            }
        }
    }
}

void mzo_merge_roulette(mzo_rng *rng, mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species, int64_t target_np,
                        int conservative) {
    mzo_ind *d = mzo_pia_at(pia, cell, species);
    const int64_t n_to_delete = d->n_local - target_np;
    double w_total0 = 0.0, v0[3] = {0.0, 0.0, 0.0}, E0 = 0.0;
    if (conservative) {
        roulette_sums(pv, d, NULL, &w_total0, v0, NULL);
        for (int k = 0; k < 3; ++k) v0[k] /= w_total0;
        roulette_sums(pv, d, v0, NULL, NULL, &E0);
        E0 /= w_total0;
    }
    double w_deleted = 0.0;
    for (int64_t n = 0; n < n_to_delete; ++n) {
        int64_t i_delete = (int64_t)floor(mzo_rand(rng) * (double)d->n_local);
        i_delete = mzo_map_cont_index(d, i_delete);
        w_deleted += pv->w[pv->index[i_delete - 1] - 1];
        mzo_delete_particle(pv, pia, cell, species, i_delete);
    }
    pia->contiguous[species - 1] = 0;
    double scale_factor;
    if (conservative) {
        scale_factor = w_total0 / (w_total0 - w_deleted);
    } else {
        double w_total_new = 0.0;
        roulette_sums(pv, d, NULL, &w_total_new, NULL, NULL);
        scale_factor = (w_total_new + w_deleted) / w_total_new;
    }
    for (int g = 0; g < 2; ++g) {
        if (g == 1 && d->n_group2 <= 0) break;
        const int64_t s = g == 0 ? d->start1 : d->start2, e = g == 0 ? d->end1 : d->end2;
        for (int64_t i = s; i <= e; ++i) pv->w[pv->index[i - 1] - 1] *= scale_factor;
    }
    if (conservative) {
        double vm[3] = {0.0, 0.0, 0.0}, E_new = 0.0;
        roulette_sums(pv, d, NULL, NULL, vm, NULL);
        for (int k = 0; k < 3; ++k) vm[k] /= w_total0;
        roulette_sums(pv, d, vm, NULL, NULL, &E_new);
        E_new /= w_total0;
        const double E_scale = sqrt(E0 / E_new);
        for (int g = 0; g < 2; ++g) {
            if (g == 1 && d->n_group2 <= 0) break;
            const int64_t s = g == 0 ? d->start1 : d->start2, e = g == 0 ? d->end1 : d->end2;
            for (int64_t i = s; i <= e; ++i) {
                double *v = pv->v + 3 * (pv->index[i - 1] - 1);
                for (int k = 0; k < 3; ++k) v[k] = E_scale * (v[k] - vm[k]) + v0[k];
            }
        }
    }
}

/* ===================================================================== */
/* Grid N:2 merging (src/merging/merging_grid.jl)                          */

mzo_gridmerge *mzo_gridmerge_new(int64_t Nx, int64_t Ny, int64_t Nz, const double mult[3]) {
    mzo_gridmerge *mg = (mzo_gridmerge *)calloc(1, sizeof(mzo_gridmerge));
    mg->Nx = Nx; mg->Ny = Ny; mg->Nz = Nz;
    mg->NyNz = Ny * Nz;
    mg->Ntotal = Nx * Ny * Nz + 8;
    for (int d = 0; d < 3; ++d) mg->mult[d] = mult[d];
    mg->cells = (mzo_gcell *)calloc((size_t)mg->Ntotal, sizeof(mzo_gcell));
    return mg;
}

void mzo_gridmerge_free(mzo_gridmerge *mg) {
    if (!mg) return;
    free(mg->cells);
    free(mg);
}

/* compute_velocity_extent!(merging_grid, cell, species, species_data, phys_props) (:190-197) */
void mzo_gridmerge_extent_props(mzo_gridmerge *mg, double T, const double v[3], double mass) {
    const double vt = sqrt(2 * T * MZO_KB / mass);
    const int64_t N[3] = {mg->Nx, mg->Ny, mg->Nz};
    for (int d = 0; d < 3; ++d) {
        const double dv = mg->mult[d] * vt;
        mg->lower[d] = v[d] - dv;
        mg->upper[d] = v[d] + dv;
        mg->mid[d] = v[d];
        mg->dv[d] = 2 * dv / (double)N[d];
        mg->dv_inv[d] = 1.0 / mg->dv[d];
    }
}

/* compute_velocity_extent!(merging_grid, vx_extent, vy_extent, vz_extent) (:210-220) */
void mzo_gridmerge_extent(mzo_gridmerge *mg, const double vx[2], const double vy[2], const double vz[2]) {
    const double *e[3] = {vx, vy, vz};
    const int64_t N[3] = {mg->Nx, mg->Ny, mg->Nz};
    for (int d = 0; d < 3; ++d) {
        mg->lower[d] = e[d][0];
        mg->upper[d] = e[d][1];
        mg->mid[d] = 0.5 * (e[d][0] + e[d][1]);
        mg->dv[d] = (e[d][1] - e[d][0]) / (double)N[d];
        mg->dv_inv[d] = 1.0 / mg->dv[d];
    }
}

/* compute_grid_index (:235-270).  A velocity exactly on an upper bound gives
 * floor((upper - lower) / dv) = N, an index the reference does not bound
 * (@inbounds); indices past Ntotal are clamped here and counted.  A grid of
 * zero width (T = 0) gives NaN cell coordinates, where the reference's
 * floor(Int64, NaN) throws; here each coordinate is taken into [0, N]
 * (NaN -> 0) before the index is formed. */
static int64_t grid_coord(double f, int64_t n) {
    const double fl = floor(f);
    return fl >= 0.0 ? (fl <= (double)n ? (int64_t)fl : n) : 0;
}

int64_t mzo_gridmerge_index(mzo_gridmerge *mg, const double v[3]) {
    int outside = 0;
    for (int d = 0; d < 3 && !outside; ++d)
        if (v[d] < mg->lower[d] || v[d] > mg->upper[d]) outside = 1;
    int64_t index;
    if (!outside) {
        const double ix = (v[0] - mg->lower[0]) * mg->dv_inv[0];
        const double iy = (v[1] - mg->lower[1]) * mg->dv_inv[1];
        const double iz = (v[2] - mg->lower[2]) * mg->dv_inv[2];
        index = grid_coord(ix, mg->Nx) * mg->NyNz + grid_coord(iy, mg->Ny) * mg->Nz + grid_coord(iz, mg->Nz) + 1;
    } else {
        index = mg->Ntotal - 7;
        if (v[0] > mg->mid[0]) index += 1;
        if (v[1] > mg->mid[1]) index += 2;
        if (v[2] > mg->mid[2]) index += 4;
    }
    if (index > mg->Ntotal) {
        mg->out_of_range += 1;
        index = mg->Ntotal;
    }
    return index;
}

/* compute_grid! (:304-384): sums in particle order, group 1 then group 2 */
static void grid_accumulate(mzo_gridmerge *mg, mzo_pv *pv, int64_t pos, int pass) {
    const int64_t id = PID(pv, pos);
    const double w = PV_W(pv, id), *v = PV_V(pv, id), *x = PV_X(pv, id);
    mzo_gcell *g = &mg->cells[mzo_gridmerge_index(mg, v) - 1];
    if (pass == 0) {
        g->np += 1;
        g->w += w;
        for (int d = 0; d < 3; ++d) {
            g->v_mean[d] = g->v_mean[d] + v[d] * w;
            g->x_mean[d] = g->x_mean[d] + x[d] * w;
        }
        if (g->np == 1) g->particle_index1 = pos;
        else if (g->np == 2) g->particle_index2 = pos;
    } else {
        for (int d = 0; d < 3; ++d) {
            const double a = v[d] - g->v_mean[d], c = x[d] - g->x_mean[d];
            g->v_std_sq[d] = g->v_std_sq[d] + (a * a) * w;
            g->x_std_sq[d] = g->x_std_sq[d] + (c * c) * w;
        }
    }
}

static void grid_compute(mzo_gridmerge *mg, mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species) {
    memset(mg->cells, 0, sizeof(mzo_gcell) * (size_t)mg->Ntotal);   /* clear_merging_grid! (:278-291) */
    const mzo_ind *d = mzo_pia_at(pia, cell, species);
    for (int pass = 0; pass < 2; ++pass) {
        for (int64_t i = d->start1; i <= d->end1; ++i) grid_accumulate(mg, pv, i, pass);
        if (d->start2 > 0)
            for (int64_t i = d->start2; i <= d->end2; ++i) grid_accumulate(mg, pv, i, pass);
        for (int64_t k = 0; k < mg->Ntotal; ++k) {
            mzo_gcell *g = &mg->cells[k];
            if (pass == 0) {
                if (g->w > 0.0) {
                    for (int q = 0; q < 3; ++q) { g->v_mean[q] = g->v_mean[q] / g->w; g->x_mean[q] = g->x_mean[q] / g->w; }
                } else {
                    g->np = 0;
                }
            } else if (g->w > 0.0) {
                for (int q = 0; q < 3; ++q) { g->v_std_sq[q] = g->v_std_sq[q] / g->w; g->x_std_sq[q] = g->x_std_sq[q] / g->w; }
            }
        }
    }
}

/* compute_new_particles! (:393-482) and its grid-clamped variant (:484-595) */
void mzo_merge_grid_based(mzo_rng *rng, mzo_gridmerge *mg, mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species,
                          const mzo_grid1d *grid) {
    mzo_rng_set_stream(rng, (uint64_t)(cell - 1));
    grid_compute(mg, pv, pia, cell, species);
    const int64_t Nt = mg->Ntotal;
    double *ow = (double *)calloc((size_t)(2 * Nt), sizeof(double));
    double *ov = (double *)calloc((size_t)(6 * Nt), sizeof(double));
    double *ox = (double *)calloc((size_t)(6 * Nt), sizeof(double));
    for (int64_t k = 0; k < Nt; ++k) {
        mzo_gcell *g = &mg->cells[k];
        double *w1 = &ow[2 * k], *v1 = &ov[6 * k], *x1 = &ox[6 * k];
        if (g->np > 2) {
            w1[0] = 0.5 * g->w;
            w1[1] = w1[0];
            double sv[3], sx[3], dir[3];
            for (int d = 0; d < 3; ++d) { sv[d] = sqrt(g->v_std_sq[d]); sx[d] = sqrt(g->x_std_sq[d]); }
            for (int d = 0; d < 3; ++d) dir[d] = draw_sign(rng);
            for (int d = 0; d < 3; ++d) { v1[d] = g->v_mean[d] + dir[d] * sv[d]; v1[3 + d] = g->v_mean[d] - dir[d] * sv[d]; }
            for (int d = 0; d < 3; ++d) dir[d] = draw_sign(rng);
            for (int d = 0; d < 3; ++d) { x1[d] = g->x_mean[d] + dir[d] * sx[d]; x1[3 + d] = g->x_mean[d] - dir[d] * sx[d]; }
        } else if (g->np >= 1) {
            for (int h = 0; h < g->np; ++h) {
                const int64_t id = PID(pv, h == 0 ? g->particle_index1 : g->particle_index2);
                w1[h] = PV_W(pv, id);
                memcpy(&v1[3 * h], PV_V(pv, id), 24);
                memcpy(&x1[3 * h], PV_X(pv, id), 24);
            }
        }
    }
    mzo_ind *d = mzo_pia_at(pia, cell, species);
    int64_t cur = 0;
    for (int64_t k = 0; k < Nt; ++k) {
        const int64_t np = mg->cells[k].np;
        const int n_out = np >= 2 ? 2 : (int)np;
        for (int h = 0; h < n_out; ++h) {
            const int64_t id = PID(pv, mzo_map_cont_index(d, cur));
            cur += 1;
            PV_W(pv, id) = ow[2 * k + h];
            memcpy(PV_V(pv, id), &ov[6 * k + 3 * h], 24);
            memcpy(PV_X(pv, id), &ox[6 * k + 3 * h], 24);
            if (grid && np >= 2) {
                if (PV_X(pv, id)[0] < grid->min_x) PV_X(pv, id)[0] = grid->min_x;
                else if (PV_X(pv, id)[0] > grid->max_x) PV_X(pv, id)[0] = grid->max_x;
            }
        }
    }
    free(ow); free(ov); free(ox);
    const int64_t n_del = d->n_local - cur;
    if (!(cell == pia->n_cells) || (n_del > d->n_group2)) pia->contiguous[species - 1] = 0;
    for (int64_t i = 0; i < n_del; ++i) mzo_delete_particle_end(pv, pia, cell, species);
}
