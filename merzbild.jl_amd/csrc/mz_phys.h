// mz_phys.h -- per-particle and per-pair physics shared by the kernels that
// launch one step phase at a time and by the slab-resident step kernel, so
// both evaluate exactly the same floating-point expressions.
//
// Reference (Merzbild.jl, src/):
//   compute_n_coll_single_species / _two_species  collisions/collision_ntc.jl:173-197
//   sigma_vhs                                      collisions/collision_cross_sections.jl:179-181
//   collide_2particles_vhs! acceptance             collisions/collision_ntc.jl:223-246
//   compute_com!                                   collisions/collision_utils.jl:374-376
//   scatter_vhs!                                   collisions/collision_scattering.jl:17-29
//   convect_single_particle!                       convection/convection_1D.jl:17-54
//   diffuse_reflection_x!, reflect_particle_x!     convection/boundary_conditions.jl:79-121
//   get_cell (uniform 1-D grid)                    grids/grid_uniform1D.jl:97-99
// The power of the relative speed, the logarithms of the wall draws and the
// sine/cosine pairs are the series forms of mz_math.h (within 2 ulp of the C
// library), which keep these one-lane chains short.
#pragma once

#include "mz_internal.h"
#include "mz_math.h"

#define MZ_G_EPS 2.220446049250313e-16  // eps(Float64): the g > eps() test of ntc!

// n_coll before the floor: 0.5 dt N (N - 1) (sigma g w)_max / V + R for one
// species, dt N1 N2 (sigma g w)_max / V + R for two.
__device__ inline double mz_ncoll_est(bool single, double dt, int32_t nla, int32_t nlb, double sgwm, double V,
                                      double R) {
    return single ? 0.5 * dt * (double)nla * (double)(nla - 1) * sgwm / V + R
                  : dt * (double)nla * (double)nlb * sgwm / V + R;
}

// |v_i - v_k| with g = v_i - v_k
__device__ inline double mz_rel_speed(double vi0, double vi1, double vi2, double vk0, double vk1, double vk2,
                                      double &g0, double &g1, double &g2) {
    g0 = vi0 - vk0;
    g1 = vi1 - vk1;
    g2 = vi2 - vk2;
    return sqrt(g0 * g0 + g1 * g1 + g2 * g2);
}

// sigma(g) g max(w_i, w_k), sigma(g) = C g^(1 - 2 omega) (expo = 1 - 2 omega);
// g^0 is 1 for every g, so pseudo-Maxwell molecules (omega = 1/2) skip pow
__device__ inline double mz_vhs_sgw(double factor, double expo, double g, double wi, double wk) {
    double sigma = factor * (expo == 0.0 ? 1.0 : mz_pow_pos(g, expo));
    return sigma * g * (wi > wk ? wi : wk);
}

// centre-of-mass velocity, mu1 v_i + mu2 v_k as one fused multiply-add
__device__ inline void mz_com(double mu1, double mu2, double vi0, double vi1, double vi2, double vk0, double vk1,
                              double vk2, double &vc0, double &vc1, double &vc2) {
    vc0 = fma(mu1, vi0, mu2 * vk0);
    vc1 = fma(mu1, vi1, mu2 * vk1);
    vc2 = fma(mu1, vi2, mu2 * vk2);
}

// scatter_vhs! from its two draws (R_phi drawn first, then R_ct): isotropic
// post-collision relative velocity of magnitude g about the centre of mass.
__device__ inline void mz_vhs_post(double mu1, double mu2, double g, double vc0, double vc1, double vc2,
                                   double R_phi, double R_ct, double &vi0, double &vi1, double &vi2, double &vk0,
                                   double &vk1, double &vk2) {
    double phi = 6.283185307179586 * R_phi;
    double sphi, cphi;
    mz_sincos(phi, &sphi, &cphi);
    double ct = 2.0 * R_ct - 1.0;
    double stt = sqrt(1.0 - ct * ct);
    double gn0 = g * (stt * cphi), gn1 = g * (stt * sphi), gn2 = g * ct;
    vi0 = fma(mu2, gn0, vc0);
    vi1 = fma(mu2, gn1, vc1);
    vi2 = fma(mu2, gn2, vc2);
    vk0 = fma(-mu1, gn0, vc0);
    vk1 = fma(-mu1, gn1, vc1);
    vk2 = fma(-mu1, gn2, vc2);
}

// scatter_vhs!: two draws, phi first, then cos(theta)
__device__ inline void mz_vhs_scatter(const MzRngParams &rp, MzStream &st, double mu1, double mu2, double g,
                                      double vc0, double vc1, double vc2, double &vi0, double &vi1, double &vi2,
                                      double &vk0, double &vk1, double &vk2) {
    double R_phi = mz_rand(rp, st);
    double R_ct = mz_rand(rp, st);
    mz_vhs_post(mu1, mu2, g, vc0, vc1, vc2, R_phi, R_ct, vi0, vi1, vi2, vk0, vk1, vk2);
}

// Diffuse re-emission from a wall: three draws (normal, tangential speed
// factors clamped at 1e-50, then the tangential angle).
__device__ inline void mz_diffuse_reflect(const MzRngParams &rp, MzStream &st, double &v0, double &v1, double &v2,
                                          double vsq, double sgn, const double *wv) {
    double R = fmax(1e-50, mz_rand(rp, st));
    double vn = sgn * sqrt(-vsq * mz_log(R));
    R = fmax(1e-50, mz_rand(rp, st));
    double vt = sqrt(-vsq * mz_log(R));
    R = 6.283185307179586 * mz_rand(rp, st);
    double s, c;
    mz_sincos(R, &s, &c);
    v0 = vn + wv[0];
    v1 = s * vt + wv[1];
    v2 = c * vt + wv[2];
}

// Free flight x + v dt (an explicit fma, as the reference's @muladd block)
// with wall interactions while the particle is outside (0, L); the result is
// clamped to [min_x, max_x].  A particle draws from stream (id - 1).
// With pre_bc >= 0, (p0, p1, p2) is the velocity a first diffuse reflection
// at wall pre_bc produces, computed ahead by mz_diffuse_reflect from the same
// stream (its draws 0-2); later bounces continue the stream at draw 3, so the
// result is the same as without it.
// Surface quantities of the 1-D grid's two walls (surface_props.jl:22-128):
// element 0 is the left wall (normal +x, boundary condition 1), element 1 the
// right wall (normal -x).  Per element MZ_SURF_Q values, in the order np,
// flux_incident, flux_reflected, force[3], normal_pressure,
// shear_pressure[3], kinetic_energy_flux.  One thread sums the contacts of
// its particle; hit() adds the incident particle (sgn = 1,
// update_surface_incident!) or subtracts the reflected one (sgn = -1,
// update_surface_reflected!), with the reference's fused forms: the normal
// projection muladd(px, n1, muladd(py, n2, pz n3)), the shear components
// muladd(-p_dot_n, n_i, p_i) and the energy 0.5 muladd(px, v1, muladd(py, v2,
// pz v3)).
#define MZ_SURF_Q 11
#define MZ_SURF_EACH(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10)
struct MzSurfAcc {
    double w;
    double q[2][MZ_SURF_Q];

    __device__ __attribute__((always_inline)) void zero() {
#define MZ_SURF_ZERO(k) q[0][k] = 0.0; q[1][k] = 0.0;
        MZ_SURF_EACH(MZ_SURF_ZERO)
#undef MZ_SURF_ZERO
    }
    __device__ __attribute__((always_inline)) void hit(int bc, double v0, double v1, double v2, double sgn) {
        const double n0 = bc == 0 ? 1.0 : -1.0, n1 = 0.0, n2 = 0.0;
        const double px = w * v0, py = w * v1, pz = w * v2;
        const double pdn = fma(px, n0, fma(py, n1, pz * n2));
        const double d0 = sgn > 0.0 ? 1.0 : 0.0;
        const double d1 = sgn > 0.0 ? w : 0.0;
        const double d2 = sgn > 0.0 ? 0.0 : -w;
        const double d3 = sgn * px;
        const double d4 = sgn * py;
        const double d5 = sgn * pz;
        const double d6 = -sgn * pdn;
        const double d7 = sgn * fma(-pdn, n0, px);
        const double d8 = sgn * fma(-pdn, n1, py);
        const double d9 = sgn * fma(-pdn, n2, pz);
        const double d10 = sgn * (0.5 * fma(px, v0, fma(py, v1, pz * v2)));
        // both elements updated through selects (adding 0.0 changes no sum),
        // every index spelled out, so q stays in registers
        const bool left = bc == 0;
#define MZ_SURF_ADD(k) q[0][k] += left ? d##k : 0.0; q[1][k] += left ? 0.0 : d##k;
        MZ_SURF_EACH(MZ_SURF_ADD)
#undef MZ_SURF_ADD
    }
};

template <class Acc>
__device__ inline void mz_convect_impl(const MzRngParams &rp, const MzGrid &g, const MzWalls &wl, int sidx,
                                       double dt, int32_t id, double &x0, double &v0, double &v1, double &v2,
                                       int32_t *err, int pre_bc, double p0, double p1, double p2, Acc &sa) {
    double t_rest = dt;
    double x_old = x0;
    double x_new = fma(v0, dt, x0);
    int bounces = 0;
    MzStream st = mz_stream((uint64_t)(id - 1));
    while ((x_new >= g.L) || (x_new <= 0.0)) {
        if (++bounces > MZ_MAX_BOUNCES) {
            atomicOr(err, MZ_ERR_BOUNCES);
            break;
        }
        int bc;
        double wn;
        if (x_new >= g.L) {
            t_rest -= fabs((g.L - x_old) / v0);
            bc = 1; wn = -1.0; x_old = g.L;
        } else {
            t_rest -= fabs(x_old / v0);
            bc = 0; wn = 1.0; x_old = 0.0;
        }
        sa.hit(bc, v0, v1, v2, 1.0);
        double acc = wl.acc[bc];
        if (acc == 0.0) {
            v0 = -v0;
        } else if (acc == 1.0) {
            if (bounces == 1 && bc == pre_bc) {
                v0 = p0; v1 = p1; v2 = p2;
                st = mz_stream((uint64_t)(id - 1), 3);
            } else {
                mz_diffuse_reflect(rp, st, v0, v1, v2, wl.vsq[bc][sidx], wn, wl.v[bc]);
            }
        } else {
            double R = mz_rand(rp, st);
            if (R < acc) mz_diffuse_reflect(rp, st, v0, v1, v2, wl.vsq[bc][sidx], wn, wl.v[bc]);
            else v0 = -v0;
        }
        sa.hit(bc, v0, v1, v2, -1.0);
        x_new = fma(v0, t_rest, x_old);
    }
    if (x_new < g.min_x) x_new = g.min_x;
    else if (x_new > g.max_x) x_new = g.max_x;
    x0 = x_new;
}

// Term k of one wall contact on wall bc: the incident particle (sgn = 1,
// update_surface_incident!) adds, the reflected one (sgn = -1,
// update_surface_reflected!) subtracts, with the reference's fused forms:
// the normal projection muladd(px, n1, muladd(py, n2, pz n3)), the shear
// components muladd(-p_dot_n, n_i, p_i) and the energy 0.5 muladd(px, v1,
// muladd(py, v2, pz v3)).
__device__ inline double mz_surf_term(int k, int bc, double w, double v0, double v1, double v2, double sgn) {
    const double n0 = bc == 0 ? 1.0 : -1.0, n1 = 0.0, n2 = 0.0;
    const double px = w * v0, py = w * v1, pz = w * v2;
    const double pdn = fma(px, n0, fma(py, n1, pz * n2));
    switch (k) {
    case 0: return sgn > 0.0 ? 1.0 : 0.0;
    case 1: return sgn > 0.0 ? w : 0.0;
    case 2: return sgn > 0.0 ? 0.0 : -w;
    case 3: return sgn * px;
    case 4: return sgn * py;
    case 5: return sgn * pz;
    case 6: return -sgn * pdn;
    case 7: return sgn * fma(-pdn, n0, px);
    case 8: return sgn * fma(-pdn, n1, py);
    case 9: return sgn * fma(-pdn, n2, pz);
    default: return sgn * (0.5 * fma(px, v0, fma(py, v1, pz * v2)));
    }
}

// Walls without surface properties.
struct MzNoSurf {
    __device__ void hit(int, double, double, double, double) {}
};

__device__ inline void mz_convect_one(const MzRngParams &rp, const MzGrid &g, const MzWalls &wl, int sidx,
                                      double dt, int32_t id, double &x0, double &v0, double &v1, double &v2,
                                      int32_t *err, int pre_bc = -1, double p0 = 0.0, double p1 = 0.0,
                                      double p2 = 0.0) {
    MzNoSurf ns;
    mz_convect_impl(rp, g, wl, sidx, dt, id, x0, v0, v1, v2, err, pre_bc, p0, p1, p2, ns);
}

// The same flight with every wall contact passed to sa.hit(bc, v0, v1, v2,
// sgn) (an MzSurfAcc, or a recorder that keeps the contacts for a later sum):
// before the reflection with sgn = 1, after it with sgn = -1.
template <class Acc>
__device__ inline void mz_convect_one_surf(Acc &sa, const MzRngParams &rp, const MzGrid &g, const MzWalls &wl,
                                           int sidx, double dt, int32_t id, double &x0, double &v0, double &v1,
                                           double &v2, int32_t *err, int pre_bc = -1, double p0 = 0.0,
                                           double p1 = 0.0, double p2 = 0.0) {
    mz_convect_impl(rp, g, wl, sidx, dt, id, x0, v0, v1, v2, err, pre_bc, p0, p1, p2, sa);
}

// get_cell: floor(x / dx) + 1 (1-based); cells outside [1, n_cells] flag a
// capacity error and are clamped so no kernel indexes out of range.
__device__ inline int32_t mz_cell_of(double x0, const MzGrid &g, int32_t n_cells, int32_t *err) {
    int32_t c = (int32_t)floor(x0 * g.inv_dx) + 1;
    if (c < 1 || c > n_cells) {
        atomicOr(err, MZ_ERR_CAPACITY);
        c = c < 1 ? 1 : n_cells;
    }
    return c;
}

// compute_props! of one cell on one wave (physical_props.jl:104-154 and
// 317-454): n = sum w, vbar = sum w v / n, and T from the centred energy,
// in one pass with shifted sums (shift = the cell's first particle), lanes
// striding over the cell's logical positions.  get(i, w, v0, v1, v2) loads
// logical particle i.  Every kernel that computes props goes through here,
// so all paths sum in the same order.
struct MzPropsOut {
    double n, v0, v1, v2, T;
};

template <class Get>
__device__ inline MzPropsOut mz_props_wave(const Get &get, int32_t tot, double mass) {
    const int lane = threadIdx.x & 63;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (tot > 0) {
        double w0;
        get(0, w0, sx, sy, sz);
    }
    double n = 0.0, a0 = 0.0, a1 = 0.0, a2 = 0.0, e = 0.0;
    for (int32_t i = lane; i < tot; i += 64) {
        double w, v0, v1, v2;
        get(i, w, v0, v1, v2);
        const double d0 = v0 - sx, d1 = v1 - sy, d2 = v2 - sz;
        n += w;
        a0 = fma(w, d0, a0);
        a1 = fma(w, d1, a1);
        a2 = fma(w, d2, a2);
        e = fma(w, d0 * d0 + d1 * d1 + d2 * d2, e);
    }
    n = mz_wave_sum(n);
    a0 = mz_wave_sum(a0);
    a1 = mz_wave_sum(a1);
    a2 = mz_wave_sum(a2);
    e = mz_wave_sum(e);
    MzPropsOut o{n, 0.0, 0.0, 0.0, 0.0};
    if (n > 0.0) {
        const double m0 = a0 / n, m1 = a1 / n, m2 = a2 / n;
        o.v0 = sx + m0;
        o.v1 = sy + m1;
        o.v2 = sz + m2;
        double E = e - n * (m0 * m0 + m1 * m1 + m2 * m2);
        if (E < 0.0) E = 0.0;
        E *= 0.5 * mass / (n * MZ_KB);
        o.T = (2.0 / 3.0) * E;
    }
    return o;
}
