/*
 * mzo_coll.c -- oracle: VHS interactions, NTC collisions, physical
 * properties and samplers.
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

static const double TWOPI = 6.283185307179586;   /* src/constants.jl:29 */
static const double PI_D = 3.141592653589793;

/* compute_vhs_factor (src/collisions/collision_utils.jl:98-101),
 * compute_mu_ref (:218-223) and the [i,k]/[k,i] pair built by
 * load_interaction_data (:159-202).  Gamma is C's tgamma, which agrees with
 * SpecialFunctions.gamma to about one ulp. */
void mzo_interaction_init(mzo_interaction *ik, mzo_interaction *ki, double m_i, double m_k,
                          double d, double omega, double Tref) {
    double m_r = m_i * m_k / (m_i + m_k);
    double mu1 = m_i / (m_i + m_k);
    double mu2 = m_k / (m_i + m_k);
    double mref = 0.5 * (m_i + m_k);
    double num = 30.0 * sqrt(mref * MZO_KB * Tref);
    double den = 4.0 * sqrt(PI_D) * (5.0 - 2.0 * omega) * (7.0 - 2.0 * omega) * d * d;
    double muref = num / den;
    double fac = PI_D * d * d * pow(2 * MZO_KB * Tref / m_r, omega - 0.5) / tgamma(2.5 - omega);
    mzo_interaction a = {m_r, mu1, mu2, d, omega, Tref, muref, fac};
    mzo_interaction b = {m_r, mu2, mu1, d, omega, Tref, muref, fac};
    if (ik) *ik = a;
    if (ki) *ki = b;
}

/* sigma_vhs (src/collisions/collision_cross_sections.jl:179-181) */
double mzo_sigma_vhs(const mzo_interaction *it, double g) {
    return it->vhs_factor * pow(g, 1.0 - 2 * it->vhs_o);
}

/* estimate_sigma_g_w_max (src/collisions/collision_utils.jl:418-424) */
double mzo_estimate_sigma_g_w_max(const mzo_interaction *it, double m1, double m2, double T1,
                                  double T2, double Fnum, double mult) {
    double g1 = sqrt(2 * T1 * MZO_KB / m1);
    double g2 = sqrt(2 * T2 * MZO_KB / m2);
    double g = 0.5 * (g1 + g2);
    return mult * mzo_sigma_vhs(it, g) * g * Fnum;
}

/* ------------------------------------------------------------------------- */
/* One collision of particles (pv1, id_i) and (pv2, id_k), whose relative
 * speed g has already passed g > eps().  collide_2particles_vhs!
 * (src/collisions/collision_ntc.jl:223-270) and its equal-weight variant
 * (:294-309); compute_com! (collision_utils.jl:374-376) and scatter_vhs!
 * (collision_scattering.jl:17-29).  The @muladd sites v_com = mu1 v1 + mu2 v2
 * and v_com +/- mu g' are written as fma(). */
static void collide(mzo_rng *rng, mzo_cf *cf, const mzo_interaction *it, mzo_pv *pv1, mzo_pv *pv2,
                    mzo_pia *pia, int64_t cell, int64_t s1, int64_t s2, int64_t id_i, int64_t id_k,
                    double g, double dw_tol, int equal_weight) {
    double wi = PV_W(pv1, id_i), wk = PV_W(pv2, id_k);
    double sigma = mzo_sigma_vhs(it, g);
    double sgw = sigma * g * (wi > wk ? wi : wk);
    cf->sigma_g_w_max = sgw > cf->sigma_g_w_max ? sgw : cf->sigma_g_w_max;
    if (mzo_rand(rng) < sgw / cf->sigma_g_w_max) {
        cf->n_coll_performed += 1;
        double *vi = PV_V(pv1, id_i), *vk = PV_V(pv2, id_k);
        double vcom[3];
        for (int d = 0; d < 3; ++d) vcom[d] = fma(it->mu1, vi[d], it->mu2 * vk[d]);
        if (equal_weight) {
            cf->n_eq_w_coll_performed += 1;
        } else if (fabs(wi - wk) < dw_tol) {
            cf->n_eq_w_coll_performed += 1;
        } else if (wi > wk) {
            if (pv1->length <= pia->n_total[s1 - 1]) mzo_pv_resize(pv1, pv1->length + MZO_DELTA_PARTICLES);
            vi = PV_V(pv1, id_i); vk = PV_V(pv2, id_k);   /* storage may have moved */
            mzo_update_buffer_index_new_particle(pv1, pia, cell, s1);
            double dw = wi - wk;
            PV_W(pv1, id_i) = wk;
            int64_t nid = PID(pv1, pia->n_total[s1 - 1]);
            PV_W(pv1, nid) = dw;
            memcpy(PV_V(pv1, nid), PV_V(pv1, id_i), 3 * sizeof(double));
            memcpy(PV_X(pv1, nid), PV_X(pv1, id_i), 3 * sizeof(double));
        } else {
            if (pv2->length <= pia->n_total[s2 - 1]) mzo_pv_resize(pv2, pv2->length + MZO_DELTA_PARTICLES);
            vi = PV_V(pv1, id_i); vk = PV_V(pv2, id_k);
            mzo_update_buffer_index_new_particle(pv2, pia, cell, s2);
            double dw = wk - wi;
            PV_W(pv2, id_k) = wi;
            int64_t nid = PID(pv2, pia->n_total[s2 - 1]);
            PV_W(pv2, nid) = dw;
            memcpy(PV_V(pv2, nid), PV_V(pv2, id_k), 3 * sizeof(double));
            memcpy(PV_X(pv2, nid), PV_X(pv2, id_k), 3 * sizeof(double));
        }
        vi = PV_V(pv1, id_i); vk = PV_V(pv2, id_k);
        /* scatter_vhs!: phi first, then cos(theta) */
        double phi = TWOPI * mzo_rand(rng);
        double cphi = cos(phi), sphi = sin(phi);
        double ct = 2.0 * mzo_rand(rng) - 1.0;
        double st = sqrt(1.0 - ct * ct);
        double gn[3] = {g * (st * cphi), g * (st * sphi), g * ct};
        for (int d = 0; d < 3; ++d) {
            vi[d] = fma(it->mu2, gn[d], vcom[d]);
            vk[d] = fma(-it->mu1, gn[d], vcom[d]);
        }
    }
}

static double rel_speed(const double *a, const double *b) {
    double g0 = a[0] - b[0], g1 = a[1] - b[1], g2 = a[2] - b[2];
    return sqrt(g0 * g0 + g1 * g1 + g2 * g2);
}

void mzo_collide_pair(mzo_rng *rng, mzo_cf *cf, const mzo_interaction *it, mzo_pv *pv1, mzo_pv *pv2,
                      mzo_pia *pia, int64_t cell, int64_t s1, int64_t s2, int64_t id_i, int64_t id_k,
                      double dw_tol, int equal_weight) {
    double g = rel_speed(PV_V(pv1, id_i), PV_V(pv2, id_k));
    collide(rng, cf, it, pv1, pv2, pia, cell, s1, s2, id_i, id_k, g, dw_tol, equal_weight);
}

/* ntc! single species (src/collisions/collision_ntc.jl:338-380) and
 * ntc_equal_weight! (:479-521).  The candidate count uses one draw; each
 * candidate draws i, k (k redrawn while i == k), then one acceptance draw
 * and, when accepted, two scattering draws. */
void mzo_ntc_single(mzo_rng *rng, mzo_cf *cf, const mzo_interaction *it, mzo_pv *pv, mzo_pia *pia,
                    int64_t cell, int64_t species, double dt, double V, double dw_tol, int equal_weight) {
    mzo_ind *ind = mzo_pia_at(pia, cell, species);
    mzo_rng_set_stream(rng, (uint64_t)(cell - 1));
    cf->n1 = ind->n_local;
    cf->n2 = ind->n_local;
    int64_t np = ind->n_local;
    double ncf = 0.5 * dt * (double)np * (double)(np - 1) * cf->sigma_g_w_max / V + mzo_rand(rng);
    int64_t n_coll = (int64_t)floor(ncf);
    cf->n_coll = n_coll;
    cf->n_coll_performed = 0;
    cf->n_eq_w_coll_performed = 0;
    for (int64_t c = 0; c < n_coll; ++c) {
        int64_t i = (int64_t)floor(mzo_rand(rng) * (double)ind->n_local);
        int64_t k = (int64_t)floor(mzo_rand(rng) * (double)ind->n_local);
        while (i == k && !rng->err) k = (int64_t)floor(mzo_rand(rng) * (double)ind->n_local);
        if (rng->err) return;   /* replay table exhausted */
        i = mzo_map_cont_index(ind, i);
        k = mzo_map_cont_index(ind, k);
        int64_t id_i = PID(pv, i), id_k = PID(pv, k);
        double g = rel_speed(PV_V(pv, id_i), PV_V(pv, id_k));
        if (g > 2.220446049250313e-16)
            collide(rng, cf, it, pv, pv, pia, cell, species, species, id_i, id_k, g, dw_tol, equal_weight);
        ind = mzo_pia_at(pia, cell, species);
    }
}

/* swpm! (src/collisions/collision_swpm.jl:201-288).  The collision count
 * (compute_n_coll_single_species, :171-174) has no fusable product-sum: its
 * last operation is a division, added to the draw. */
void mzo_swpm(mzo_rng *rng, mzo_cf_swpm *cf, const mzo_interaction *it, mzo_pv *pv, mzo_pia *pia, int64_t cell,
              int64_t species, double G, double dt, double V) {
    mzo_ind *ind = mzo_pia_at(pia, cell, species);
    double w_max = 0.0;
    for (int64_t i = ind->start1; i <= ind->end1; ++i) {
        double w = PV_W(pv, PID(pv, i));
        w_max = w > w_max ? w : w_max;
    }
    if (ind->n_group2 > 0) {
        for (int64_t i = ind->start2; i <= ind->end2; ++i) {
            double w = PV_W(pv, PID(pv, i));
            w_max = w > w_max ? w : w_max;
        }
    }
    const double wtf = 1.0 / (1.0 + G);
    const double inv_w_max = 1.0 / w_max;
    mzo_rng_set_stream(rng, (uint64_t)(cell - 1));
    cf->n1 = ind->n_local;
    cf->n2 = ind->n_local;
    const int64_t np = ind->n_local;
    const double ncf = 0.5 * dt * (double)np * (double)(np - 1) * cf->sigma_g_max * w_max * (G + 1) / V +
                       mzo_rand(rng);
    const int64_t n_coll = (int64_t)floor(ncf);
    cf->n_coll = n_coll;
    cf->n_coll_performed = 0;
    for (int64_t c = 0; c < n_coll; ++c) {
        int64_t i = (int64_t)floor(mzo_rand(rng) * (double)ind->n_local);
        int64_t k = (int64_t)floor(mzo_rand(rng) * (double)ind->n_local);
        while (i == k && !rng->err) k = (int64_t)floor(mzo_rand(rng) * (double)ind->n_local);
        if (rng->err) return;   /* replay table exhausted */
        i = mzo_map_cont_index(ind, i);
        k = mzo_map_cont_index(ind, k);
        const int64_t id_i = PID(pv, i), id_k = PID(pv, k);
        const double g = rel_speed(PV_V(pv, id_i), PV_V(pv, id_k));
        if (!(g > 2.220446049250313e-16)) continue;
        const double sgm = mzo_sigma_vhs(it, g) * g;
        cf->sigma_g_max = sgm > cf->sigma_g_max ? sgm : cf->sigma_g_max;
        const double wi = PV_W(pv, id_i), wk = PV_W(pv, id_k);
        if (!(mzo_rand(rng) < sgm * (wi > wk ? wi : wk) * inv_w_max / cf->sigma_g_max)) continue;
        cf->n_coll_performed += 1;
        double vcom[3];
        {
            const double *vi = PV_V(pv, id_i), *vk = PV_V(pv, id_k);
            for (int d = 0; d < 3; ++d) vcom[d] = fma(it->mu1, vi[d], it->mu2 * vk[d]);
        }
        if (pv->length <= pia->n_total[species - 1]) mzo_pv_resize(pv, pv->length + MZO_DELTA_PARTICLES);
        const double dw = (wi < wk ? wi : wk) * wtf;
        PV_W(pv, id_i) = wi - dw;
        PV_W(pv, id_k) = wk - dw;
        int64_t nid[2];
        const int64_t src[2] = {id_i, id_k};
        for (int j = 0; j < 2; ++j) {
            mzo_update_buffer_index_new_particle(pv, pia, cell, species);
            nid[j] = PID(pv, pia->n_total[species - 1]);
            PV_W(pv, nid[j]) = dw;
            memcpy(PV_V(pv, nid[j]), PV_V(pv, src[j]), 3 * sizeof(double));
            memcpy(PV_X(pv, nid[j]), PV_X(pv, src[j]), 3 * sizeof(double));
        }
        ind = mzo_pia_at(pia, cell, species);
        /* scatter_vhs! on the two new particles: phi first, then cos(theta) */
        const double phi = TWOPI * mzo_rand(rng);
        const double cphi = cos(phi), sphi = sin(phi);
        const double ct = 2.0 * mzo_rand(rng) - 1.0;
        const double st = sqrt(1.0 - ct * ct);
        const double gn[3] = {g * (st * cphi), g * (st * sphi), g * ct};
        double *v1 = PV_V(pv, nid[0]), *v2 = PV_V(pv, nid[1]);
        for (int d = 0; d < 3; ++d) {
            v1[d] = fma(it->mu2, gn[d], vcom[d]);
            v2[d] = fma(-it->mu1, gn[d], vcom[d]);
        }
    }
}

/* ntc! two species (src/collisions/collision_ntc.jl:412-453) and
 * ntc_equal_weight! (:554-595). */
void mzo_ntc_two(mzo_rng *rng, mzo_cf *cf, const mzo_interaction *it, mzo_pv *pv1, mzo_pv *pv2,
                 mzo_pia *pia, int64_t cell, int64_t s1, int64_t s2, double dt, double V,
                 double dw_tol, int equal_weight) {
    mzo_ind *i1 = mzo_pia_at(pia, cell, s1), *i2 = mzo_pia_at(pia, cell, s2);
    mzo_rng_set_stream(rng, (uint64_t)(cell - 1));
    cf->n1 = i1->n_local;
    cf->n2 = i2->n_local;
    double ncf = dt * (double)i1->n_local * (double)i2->n_local * cf->sigma_g_w_max / V + mzo_rand(rng);
    int64_t n_coll = (int64_t)floor(ncf);
    cf->n_coll = n_coll;
    cf->n_coll_performed = 0;
    cf->n_eq_w_coll_performed = 0;
    for (int64_t c = 0; c < n_coll; ++c) {
        int64_t i = (int64_t)floor(mzo_rand(rng) * (double)i1->n_local);
        int64_t k = (int64_t)floor(mzo_rand(rng) * (double)i2->n_local);
        i = mzo_map_cont_index(i1, i);
        k = mzo_map_cont_index(i2, k);
        int64_t id_i = PID(pv1, i), id_k = PID(pv2, k);
        double g = rel_speed(PV_V(pv1, id_i), PV_V(pv2, id_k));
        if (g > 2.220446049250313e-16)
            collide(rng, cf, it, pv1, pv2, pia, cell, s1, s2, id_i, id_k, g, dw_tol, equal_weight);
        i1 = mzo_pia_at(pia, cell, s1);
        i2 = mzo_pia_at(pia, cell, s2);
    }
}

/* ------------------------------------------------------------------------- */
/* Properties (src/properties/physical_props.jl) */

#define PP_IDX(pp, cell, s) (((s) - 1) * (pp)->n_cells + ((cell) - 1))

/* accumulate n, n*v over positions [s, e] */
static void acc_first(mzo_pv *pv, int64_t s, int64_t e, double *n, double *v, int64_t *np) {
    for (int64_t i = s; i <= e; ++i) {
        int64_t id = PID(pv, i);
        double w = PV_W(pv, id);
        *n += w;
        for (int d = 0; d < 3; ++d) v[d] = v[d] + PV_V(pv, id)[d] * w;
        *np += 1;
    }
}

static double acc_second(mzo_pv *pv, int64_t s, int64_t e, const double *v, double E) {
    for (int64_t i = s; i <= e; ++i) {
        int64_t id = PID(pv, i);
        double *pvv = PV_V(pv, id);
        double a = pvv[0] - v[0], b = pvv[1] - v[1], c = pvv[2] - v[2];
        E = E + PV_W(pv, id) * (a * a + b * b + c * c);
    }
    return E;
}

/* compute_props! (physical_props.jl:104-154): groups 1 and 2, two passes. */
void mzo_compute_props(mzo_pv **pvs, mzo_pia *pia, const double *masses, mzo_props *pp) {
    for (int64_t s = 1; s <= pp->n_species; ++s) {
        for (int64_t c = 1; c <= pp->n_cells; ++c) {
            mzo_ind *d = mzo_pia_at(pia, c, s);
            int64_t np = 0;
            double n = 0.0, E = 0.0, T = 0.0, v[3] = {0.0, 0.0, 0.0};
            acc_first(pvs[s - 1], d->start1, d->end1, &n, v, &np);
            if (d->n_group2 > 0) acc_first(pvs[s - 1], d->start2, d->end2, &n, v, &np);
            if (n > 0.0) {
                for (int k = 0; k < 3; ++k) v[k] /= n;
                E = acc_second(pvs[s - 1], d->start1, d->end1, v, E);
                if (d->n_group2 > 0) E = acc_second(pvs[s - 1], d->start2, d->end2, v, E);
                E *= 0.5 * masses[s - 1] / (n * MZO_KB);
                T = (2.0 / 3.0) * E;
            }
            int64_t ix = PP_IDX(pp, c, s);
            pp->lpa[s - 1] = (double)pvs[s - 1]->length;
            pp->np[ix] = (double)np;
            pp->n[ix] = n;
            for (int k = 0; k < 3; ++k) pp->v[3 * ix + k] = v[k];
            pp->T[ix] = T;
        }
    }
}

/* compute_props_sorted! (physical_props.jl:317-454): group 1 only; with a
 * grid and ndens_not_Np the number density uses inv_V.  Cells
 * [cell_lo, cell_hi]. */
void mzo_compute_props_sorted(mzo_pv **pvs, mzo_pia *pia, const double *masses, mzo_props *pp,
                              const double *inv_V, int64_t cell_lo, int64_t cell_hi) {
    for (int64_t s = 1; s <= pp->n_species; ++s) {
        for (int64_t c = cell_lo; c <= cell_hi; ++c) {
            mzo_ind *d = mzo_pia_at(pia, c, s);
            int64_t npi = 0;
            double n = 0.0, E = 0.0, T = 0.0, v[3] = {0.0, 0.0, 0.0};
            acc_first(pvs[s - 1], d->start1, d->end1, &n, v, &npi);
            double np = d->end1 >= d->start1 ? (double)(d->end1 - d->start1) + 1.0 : 0.0;
            if (n > 0.0) {
                for (int k = 0; k < 3; ++k) v[k] /= n;
                E = acc_second(pvs[s - 1], d->start1, d->end1, v, E);
                E *= 0.5 * masses[s - 1] / (n * MZO_KB);
                T = (2.0 / 3.0) * E;
            }
            int64_t ix = PP_IDX(pp, c, s);
            pp->np[ix] = np;
            pp->n[ix] = (pp->ndens_not_Np && inv_V) ? n * inv_V[c - 1] : n;
            for (int k = 0; k < 3; ++k) pp->v[3 * ix + k] = v[k];
            pp->T[ix] = T;
        }
    }
}

/* compute_props_with_total_moments! (physical_props.jl:168-245) */
void mzo_compute_props_with_total_moments(mzo_pv **pvs, mzo_pia *pia, const double *masses,
                                          mzo_props *pp) {
    if (pp->n_moments == 0) { mzo_compute_props(pvs, pia, masses, pp); return; }
    for (int64_t s = 1; s <= pp->n_species; ++s) {
        double mf = 4 * PI_D * pow(masses[s - 1] / (TWOPI * MZO_KB * pp->Tref), 1.5) * 0.5;
        double mvref = pow(masses[s - 1] / (2 * MZO_KB * pp->Tref), 0.5);
        mzo_pv *pv = pvs[s - 1];
        for (int64_t c = 1; c <= pp->n_cells; ++c) {
            mzo_ind *d = mzo_pia_at(pia, c, s);
            int64_t ix = PP_IDX(pp, c, s);
            int64_t np = 0;
            double n = 0.0, E = 0.0, T = 0.0, v[3] = {0.0, 0.0, 0.0};
            for (int64_t m = 0; m < pp->n_moments; ++m) pp->moments[m + pp->n_moments * ix] = 0.0;
            acc_first(pv, d->start1, d->end1, &n, v, &np);
            if (d->n_group2 > 0) acc_first(pv, d->start2, d->end2, &n, v, &np);
            if (n > 0.0) {
                for (int k = 0; k < 3; ++k) v[k] /= n;
                for (int grp = 1; grp <= 2; ++grp) {
                    int64_t st = grp == 1 ? d->start1 : d->start2, en = grp == 1 ? d->end1 : d->end2;
                    if (grp == 2 && d->n_group2 <= 0) break;
                    for (int64_t i = st; i <= en; ++i) {
                        int64_t id = PID(pv, i);
                        double *pvv = PV_V(pv, id);
                        double a = pvv[0] - v[0], b = pvv[1] - v[1], cc = pvv[2] - v[2];
                        double normv = sqrt(a * a + b * b + cc * cc);
                        E = E + PV_W(pv, id) * (normv * normv);
                        for (int64_t m = 0; m < pp->n_moments; ++m)
                            pp->moments[m + pp->n_moments * ix] += PV_W(pv, id) * pow(normv, (double)pp->moment_powers[m]);
                    }
                }
                E *= 0.5 * masses[s - 1] / (n * MZO_KB);
                T = (2.0 / 3.0) * E;
            }
            for (int64_t m = 0; m < pp->n_moments; ++m) {
                int mp = pp->moment_powers[m];
                double scaling = mf * pow(mvref, -(double)(3 + mp)) * tgamma((3.0 + mp) / 2.0);
                pp->moments[m + pp->n_moments * ix] /= (scaling * n);
            }
            pp->lpa[s - 1] = (double)pv->length;
            pp->np[ix] = (double)np;
            pp->n[ix] = n;
            for (int k = 0; k < 3; ++k) pp->v[3 * ix + k] = v[k];
            pp->T[ix] = T;
        }
    }
}

/* compute_mixed_moment (physical_props.jl:473-501) */
double mzo_compute_mixed_moment(mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species,
                                int px, int py, int pz, double sum_scaler, double res_scaler) {
    double r = 0.0;
    mzo_ind *d = mzo_pia_at(pia, cell, species);
    for (int grp = 1; grp <= 2; ++grp) {
        if (grp == 2 && d->n_group2 <= 0) break;
        int64_t st = grp == 1 ? d->start1 : d->start2, en = grp == 1 ? d->end1 : d->end2;
        for (int64_t i = st; i <= en; ++i) {
            int64_t id = PID(pv, i);
            double *pvv = PV_V(pv, id);
            r += PV_W(pv, id) * sum_scaler * pow(pvv[0], px) * pow(pvv[1], py) * pow(pvv[2], pz);
        }
    }
    return r * res_scaler;
}

/* clear_props! (physical_props.jl:256-262) */
void mzo_clear_props(mzo_props *pp) {
    int64_t n = pp->n_cells * pp->n_species;
    for (int64_t s = 0; s < pp->n_species; ++s) pp->lpa[s] = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        pp->np[i] = 0.0; pp->n[i] = 0.0; pp->T[i] = 0.0;
        pp->v[3 * i] = pp->v[3 * i + 1] = pp->v[3 * i + 2] = 0.0;
    }
}

/* ------------------------------------------------------------------------- */
/* Samplers (src/distributions_and_sampling.jl) */

/* sample_maxwellian! (:432-443): 4 draws per particle */
void mzo_sample_maxwellian(mzo_rng *rng, mzo_pv *pv, int64_t nparticles, int64_t offset, double m,
                           double T, const double v0[3]) {
    double vscale = sqrt(2 * MZO_KB * T / m);
    for (int64_t i = 1; i <= nparticles; ++i) {
        double vn = sqrt(-log(mzo_rand(rng)));
        double vr = sqrt(-log(mzo_rand(rng)));
        double t1 = TWOPI * mzo_rand(rng);
        double t2 = TWOPI * mzo_rand(rng);
        int64_t id = PID(pv, i + offset);
        double *v = PV_V(pv, id);
        v[0] = vscale * (vn * cos(t1)) + v0[0];
        v[1] = vscale * (vr * cos(t2)) + v0[1];
        v[2] = vscale * (vr * sin(t2)) + v0[2];
    }
}

/* sample_particles_equal_weight! (:477-509), Maxwellian branch: 3 position
 * draws per particle, then the velocities. */
void mzo_sample_particles_equal_weight(mzo_rng *rng, mzo_pv *pv, mzo_pia *pia, int64_t cell,
                                       int64_t species, int64_t nparticles, double m, double T,
                                       double Fnum, double xlo, double xhi, double ylo, double yhi,
                                       double zlo, double zhi) {
    mzo_ind *d = mzo_pia_at(pia, cell, species);
    int64_t start = pia->n_total[species - 1] + 1;
    d->n_local = nparticles;
    pia->n_total[species - 1] += nparticles;
    d->start1 = start;
    d->end1 = start - 1 + nparticles;
    d->n_group1 = nparticles;
    d->start2 = 0; d->end2 = -1; d->n_group2 = 0;
    int64_t offset = start - 1;
    for (int64_t i = 1; i <= nparticles; ++i) {
        mzo_pv_buffer_pop_to(pv, i + offset);
        int64_t id = PID(pv, i + offset);
        PV_W(pv, id) = Fnum;
        double *v = PV_V(pv, id), *x = PV_X(pv, id);
        v[0] = v[1] = v[2] = 0.0;
        double a = mzo_rand(rng), b = mzo_rand(rng), c = mzo_rand(rng);
        x[0] = xlo + a * (xhi - xlo);
        x[1] = ylo + b * (yhi - ylo);
        x[2] = zlo + c * (zhi - zlo);
        pv->cell[i + offset - 1] = cell;
    }
    double v0[3] = {0.0, 0.0, 0.0};
    mzo_sample_maxwellian(rng, pv, nparticles, offset, m, T, v0);
}

/* sample_bkw! (:195-213) at t = 0.  |v|/vscale ~ Chi(5) is drawn as the norm
 * of 5 standard normals (Box-Muller), since Distributions.jl's Chi sampler
 * is not available; theta ~ U[0, pi) and phi ~ U[0, 2 pi) as in the
 * reference. */
void mzo_sample_bkw(mzo_rng *rng, mzo_pv *pv, int64_t nparticles, int64_t offset, double m, double T,
                    const double v0[3]) {
    double vscale = sqrt(2 * MZO_KB * T / m) * sqrt(0.3);
    for (int64_t i = 1; i <= nparticles; ++i) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j) {
            double u1 = fmax(1e-300, mzo_rand(rng)), u2 = mzo_rand(rng);
            double r = sqrt(-2.0 * log(u1));
            double z1 = r * cos(TWOPI * u2), z2 = r * sin(TWOPI * u2);
            s += z1 * z1;
            if (j < 2) s += z2 * z2;
        }
        double vabs = sqrt(s);
        double th = mzo_rand(rng) * PI_D;
        double ph = mzo_rand(rng) * TWOPI;
        double st = sin(th);
        int64_t id = PID(pv, i + offset);
        double *v = PV_V(pv, id);
        v[0] = vscale * (vabs * st * cos(ph)) + v0[0];
        v[1] = vscale * (vabs * st * sin(ph)) + v0[1];
        v[2] = vscale * (vabs * cos(th)) + v0[2];
    }
}
