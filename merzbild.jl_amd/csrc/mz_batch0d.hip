// mz_batch0d.hip -- many independent 0-D cells: every step ntc_equal_weight!
// and compute_props! in every cell (the per-cell loops of test/test_bkw.jl:
// 88-93 and test/test_2species.jl:52-65, over a batch of cells).
//
// 0-D cells never exchange particles, so a cell's steps depend on that cell
// alone.  The batched path therefore runs all steps of a call in one
// persistent launch: each wavefront pulls an item (a group of cells of
// similar size) from a queue ordered by decreasing size, and runs every step
// for it: the item's cells collide concurrently, one lane per cell, then the
// wavefront computes each cell's properties.  Items are formed on the host:
// cells sorted by occupancy are packed until an item holds `max_cells`
// cells or `target` particles, so large cells run alone and small ones share
// a wavefront (the packing is what the autotuner chooses).
//
// The plain path launches the NTC kernel (one lane per cell) and the props
// kernel (one wavefront per cell) once per step.  Both paths evaluate the
// same draws and the same expressions in the same order, so they agree bit
// for bit.
#include "mz_phys.h"

#include <algorithm>
#include <numeric>
#include <vector>

struct B0Item {
    int32_t first, ncell;
};

struct B0Args {
    MzSpView sp;
    int32_t sidx;
    double dt, mu1, mu2, factor, expo, mass;
    const double *V;
    double *sgwm;
    int32_t *cnt;
    MzRngParams rp;
    uint64_t step0;
    int32_t n_steps, props_every;
    double *out_np, *out_n, *out_v, *out_T;
    int32_t n_cells;
    const int32_t *queue;
    const B0Item *items;
    int32_t n_items;
    int32_t *next;
    double *aos;               // [slot][4]: v0, v1, v2, w of the launch's cells
};

// The persistent launch works on a packed copy of the velocities and weights,
// one 32-byte record per slot, so a collision candidate gathers two records
// instead of eight scattered words; the copy is filled when a wavefront
// takes an item and written back when it has run all steps.
__device__ inline void b0_get(const double *aos, int64_t p, double &v0, double &v1, double &v2, double &w) {
    const double4 t = reinterpret_cast<const double4 *>(aos)[p];
    v0 = t.x; v1 = t.y; v2 = t.z; w = t.w;
}
__device__ inline void b0_put_v(double *aos, int64_t p, double v0, double v1, double v2) {
    double *q = aos + 4 * p;
    q[0] = v0; q[1] = v1; q[2] = v2;
}

// ntc_equal_weight! of one cell whose logical particle i sits at slot
// i < n1 ? st1 + i : st2 + (i - n1); the candidate loop of k_ntc without
// splitting, on the cell's own draw stream.
__device__ inline void b0_ntc_cell(const B0Args &A, const MzRngParams &rp, int c, double Vc, int32_t n1,
                                   int32_t st1, int32_t st2, int32_t nla, double &sgwm, int32_t &n_coll_out,
                                   int32_t &performed, int32_t &eqw) {
    const MzSpView &sp = A.sp;
    MzStream st = mz_stream((uint64_t)c);
    double R = mz_rand(rp, st);
    int64_t n_coll = (int64_t)floor(mz_ncoll_est(true, A.dt, nla, nla, sgwm, Vc, R));
    if (!(n_coll <= MZ_NTC_MAX_CANDIDATES)) {
        atomicOr(&sp.ctr->err, MZ_ERR_SCRATCH);
        n_coll = 0;
    }
    performed = 0;
    eqw = 0;
    for (int64_t cc = 0; cc < n_coll; ++cc) {
        int32_t i = (int32_t)floor(mz_rand(rp, st) * (double)nla);
        int32_t k = (int32_t)floor(mz_rand(rp, st) * (double)nla);
        while (i == k) k = (int32_t)floor(mz_rand(rp, st) * (double)nla);
        const int64_t pi = i < n1 ? (int64_t)st1 + i : (int64_t)st2 + (i - n1);
        const int64_t pk = k < n1 ? (int64_t)st1 + k : (int64_t)st2 + (k - n1);
        double vi0, vi1, vi2, wi, vk0, vk1, vk2, wk;
        b0_get(A.aos, pi, vi0, vi1, vi2, wi);
        b0_get(A.aos, pk, vk0, vk1, vk2, wk);
        double g0, g1, g2;
        double g = mz_rel_speed(vi0, vi1, vi2, vk0, vk1, vk2, g0, g1, g2);
        if (!(g > MZ_G_EPS)) continue;
        double sgw = mz_vhs_sgw(A.factor, A.expo, g, wi, wk);
        sgwm = sgw > sgwm ? sgw : sgwm;
        if (!(mz_rand(rp, st) < sgw / sgwm)) continue;
        ++performed;
        ++eqw;
        double vc0, vc1, vc2;
        mz_com(A.mu1, A.mu2, vi0, vi1, vi2, vk0, vk1, vk2, vc0, vc1, vc2);
        double ui0, ui1, ui2, uk0, uk1, uk2;
        mz_vhs_scatter(rp, st, A.mu1, A.mu2, g, vc0, vc1, vc2, ui0, ui1, ui2, uk0, uk1, uk2);
        b0_put_v(A.aos, pi, ui0, ui1, ui2);
        b0_put_v(A.aos, pk, uk0, uk1, uk2);
    }
    n_coll_out = (int32_t)n_coll;
}

// A cell alone in its item with at least this many particles collides on
// the whole wavefront (b0_ntc_cell_wave); smaller ones use one lane.
#define B0_WAVE_MIN 256
#define B0_SPEC 2   // candidates evaluated ahead per lane and window

// The candidate starting at draw o of the cell's stream, evaluated directly
// from the current velocities (the loop body of b0_ntc_cell); returns the
// draw after it.
__device__ inline uint64_t b0_candidate(const B0Args &A, const MzRngParams &rp, int c, int32_t n1, int32_t st1,
                                        int32_t st2, int32_t nla, uint64_t o, double &sgwm, int32_t &performed,
                                        int32_t &eqw) {
    MzStream st = mz_stream((uint64_t)c, o);
    int32_t i = (int32_t)floor(mz_rand(rp, st) * (double)nla);
    int32_t k = (int32_t)floor(mz_rand(rp, st) * (double)nla);
    while (i == k) k = (int32_t)floor(mz_rand(rp, st) * (double)nla);
    const int64_t pi = i < n1 ? (int64_t)st1 + i : (int64_t)st2 + (i - n1);
    const int64_t pk = k < n1 ? (int64_t)st1 + k : (int64_t)st2 + (k - n1);
    double vi0, vi1, vi2, wi, vk0, vk1, vk2, wk;
    b0_get(A.aos, pi, vi0, vi1, vi2, wi);
    b0_get(A.aos, pk, vk0, vk1, vk2, wk);
    double g0, g1, g2;
    double g = mz_rel_speed(vi0, vi1, vi2, vk0, vk1, vk2, g0, g1, g2);
    if (!(g > MZ_G_EPS)) return st.k;
    double sgw = mz_vhs_sgw(A.factor, A.expo, g, wi, wk);
    sgwm = sgw > sgwm ? sgw : sgwm;
    if (!(mz_rand(rp, st) < sgw / sgwm)) return st.k;
    ++performed;
    ++eqw;
    double vc0, vc1, vc2;
    mz_com(A.mu1, A.mu2, vi0, vi1, vi2, vk0, vk1, vk2, vc0, vc1, vc2);
    double ui0, ui1, ui2, uk0, uk1, uk2;
    mz_vhs_scatter(rp, st, A.mu1, A.mu2, g, vc0, vc1, vc2, ui0, ui1, ui2, uk0, uk1, uk2);
    b0_put_v(A.aos, pi, ui0, ui1, ui2);
    b0_put_v(A.aos, pk, uk0, uk1, uk2);
    return st.k;
}

// ntc_equal_weight! of one large cell on a whole wavefront, with the same
// draws, order and arithmetic as b0_ntc_cell.  The candidates are taken in
// windows of 128 draws: at the start of a window at draw D, lane L evaluates
// the candidates that would start at draws D + L and D + 64 + L from the
// velocities as they are then -- its pair (draws D+L, D+L+1), relative speed, sigma g w, and the
// post-collision velocities from draws D+L+3 and D+L+4.  The candidate loop
// then runs on uniform values: each candidate starts where the previous one
// ended, takes its lane's evaluation, updates (sigma g w)_max and draws its
// acceptance number (draw o + 2) exactly as the one-lane loop does.  A
// candidate whose pair needs a redraw of k, or shares a particle with a
// collision performed earlier in the window, is evaluated directly
// (b0_candidate) and the next window starts after it.  Every lane returns
// the same sgwm and counts.
__device__ inline void b0_ntc_cell_wave(const B0Args &A, const MzRngParams &rp, int c, double Vc, int32_t n1,
                                        int32_t st1, int32_t st2, int32_t nla, double &sgwm, int32_t &n_coll_out,
                                        int32_t &performed, int32_t &eqw) {
    const MzSpView &sp = A.sp;
    const int lane = threadIdx.x & 63;
    double R;
    {
        MzStream st = mz_stream((uint64_t)c);
        R = mz_rand(rp, st);
    }
    int64_t n_coll = (int64_t)floor(mz_ncoll_est(true, A.dt, nla, nla, sgwm, Vc, R));
    if (!(n_coll <= MZ_NTC_MAX_CANDIDATES)) {
        if (lane == 0) atomicOr(&sp.ctr->err, MZ_ERR_SCRATCH);
        n_coll = 0;
    }
    performed = 0;
    eqw = 0;
    int64_t cc = 0;
    uint64_t o = 1;   // the next draw of the stream
    while (cc < n_coll) {
        // ---- window at draw D: lane L holds Philox blocks D/2 + L and
        // D/2 + 64 + L (draws 2 (D/2) .. 2 (D/2) + 255), then evaluates the
        // candidates that would start at draws D + L and D + 64 + L
        const uint64_t D = o, b0 = D >> 1;
        const double sgwm0 = sgwm;
        double h[4];
        mz_philox_pair(rp, (uint64_t)c, b0 + (uint64_t)lane, h[0], h[1]);
        mz_philox_pair(rp, (uint64_t)c, b0 + 64 + (uint64_t)lane, h[2], h[3]);
        auto draw = [&](uint64_t d) {   // any lane's draw d, 2 b0 <= d < 2 b0 + 256
            const int src = (int)((d >> 1) - b0);
            const int sl = src & 63;
            const double a0 = __shfl(h[0], sl, 64), a1 = __shfl(h[1], sl, 64);
            const double a2 = __shfl(h[2], sl, 64), a3 = __shfl(h[3], sl, 64);
            return src < 64 ? ((d & 1) ? a1 : a0) : ((d & 1) ? a3 : a2);
        };
        int32_t si[B0_SPEC], sk[B0_SPEC], sflag[B0_SPEC];
        double sgw[B0_SPEC], ui[B0_SPEC][3], uk[B0_SPEC][3];
#pragma unroll
        for (int q = 0; q < B0_SPEC; ++q) {
            const uint64_t oL = D + (uint64_t)(64 * q + lane);
            const double d0 = draw(oL), d1 = draw(oL + 1), d2 = draw(oL + 2), d3 = draw(oL + 3), d4 = draw(oL + 4);
            si[q] = (int32_t)floor(d0 * (double)nla);
            sk[q] = (int32_t)floor(d1 * (double)nla);
            const int64_t pi = si[q] < n1 ? (int64_t)st1 + si[q] : (int64_t)st2 + (si[q] - n1);
            const int64_t pk = sk[q] < n1 ? (int64_t)st1 + sk[q] : (int64_t)st2 + (sk[q] - n1);
            double vi0, vi1, vi2, wi, vk0, vk1, vk2, wk;
            b0_get(A.aos, pi, vi0, vi1, vi2, wi);
            b0_get(A.aos, pk, vk0, vk1, vk2, wk);
            double g0, g1, g2;
            const double g = mz_rel_speed(vi0, vi1, vi2, vk0, vk1, vk2, g0, g1, g2);
            sgw[q] = mz_vhs_sgw(A.factor, A.expo, g, wi, wk);
            // the acceptance test as it comes out while (sigma g w)_max stays
            // at its value at the window's start: the same operands as below
            sflag[q] = (si[q] == sk[q] ? 1 : 0) | (g > MZ_G_EPS ? 2 : 0) | (sgw[q] > sgwm0 ? 4 : 0) |
                       (d2 < sgw[q] / sgwm0 ? 8 : 0);
            double vc0, vc1, vc2;
            mz_com(A.mu1, A.mu2, vi0, vi1, vi2, vk0, vk1, vk2, vc0, vc1, vc2);
            mz_vhs_post(A.mu1, A.mu2, g, vc0, vc1, vc2, d3, d4, ui[q][0], ui[q][1], ui[q][2], uk[q][0], uk[q][1],
                        uk[q][2]);
        }
        // ---- walk the window; lane j < nd holds a particle that collided in it
        int32_t dirty = -1;
        int nd = 0;
        bool direct = false;
        while (cc < n_coll) {
            const int L = (int)(o - D);
            if (L >= 64 * B0_SPEC || nd >= 62) break;
            const int q = L >> 6, ln = L & 63;
            int32_t i, k, f;
            if (q == 0) {
                i = __builtin_amdgcn_readlane(si[0], ln);
                k = __builtin_amdgcn_readlane(sk[0], ln);
                f = __builtin_amdgcn_readlane(sflag[0], ln);
            } else {
                i = __builtin_amdgcn_readlane(si[1], ln);
                k = __builtin_amdgcn_readlane(sk[1], ln);
                f = __builtin_amdgcn_readlane(sflag[1], ln);
            }
            if ((f & 1) || __ballot(lane < nd && (dirty == i || dirty == k))) {
                direct = true;
                break;
            }
            if (!(f & 2)) {   // g <= eps: no acceptance draw
                o += 2;
                ++cc;
                continue;
            }
            bool acc;
            if (!(f & 4) && sgwm == sgwm0) {
                acc = (f & 8) != 0;   // (sigma g w)_max unchanged: the lane's test
            } else {
                const double s = mz_readlane_d(q == 0 ? sgw[0] : sgw[1], ln);
                sgwm = s > sgwm ? s : sgwm;
                const uint64_t du = o + 2;
                const int src = (int)((du >> 1) - b0);
                const int hs = (src >> 6) * 2 + (int)(du & 1);
                const double U =
                    mz_readlane_d(hs == 0 ? h[0] : (hs == 1 ? h[1] : (hs == 2 ? h[2] : h[3])), src & 63);
                acc = U < s / sgwm;
            }
            if (!acc) {
                o += 3;
                ++cc;
                continue;
            }
            ++performed;
            ++eqw;
            double a[3], e[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                a[d] = mz_readlane_d(q == 0 ? ui[0][d] : ui[1][d], ln);
                e[d] = mz_readlane_d(q == 0 ? uk[0][d] : uk[1][d], ln);
            }
            if (lane == 0) {
                const int64_t qi = i < n1 ? (int64_t)st1 + i : (int64_t)st2 + (i - n1);
                const int64_t qk = k < n1 ? (int64_t)st1 + k : (int64_t)st2 + (k - n1);
                b0_put_v(A.aos, qi, a[0], a[1], a[2]);
                b0_put_v(A.aos, qk, e[0], e[1], e[2]);
            }
            if (lane == nd) dirty = i;
            if (lane == nd + 1) dirty = k;
            nd += 2;
            o += 5;
            ++cc;
        }
        if (direct) {
            uint64_t on = 0;
            double sg = sgwm;
            int32_t pf = performed, eq = eqw;
            if (lane == 0) on = b0_candidate(A, rp, c, n1, st1, st2, nla, o, sg, pf, eq);
            o = (uint64_t)__shfl((long long)on, 0, 64);
            sgwm = mz_readlane_d(sg, 0);
            performed = __builtin_amdgcn_readlane(pf, 0);
            eqw = __builtin_amdgcn_readlane(eq, 0);
            ++cc;
        }
        // lane 0's velocity stores are complete and visible to the wave's
        // loads of the next window
        if (nd > 0 || direct) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
    n_coll_out = (int32_t)n_coll;
}

__global__ void __launch_bounds__(256, 2) k_batch0d(B0Args A) {
    const int lane = threadIdx.x & 63;
    const MzSpView &sp = A.sp;
    for (;;) {
        int item = 0;
        if (lane == 0) item = atomicAdd(A.next, 1);
        item = __shfl(item, 0, 64);
        if (item >= A.n_items) break;
        const B0Item itm = A.items[item];
        const bool own = lane < itm.ncell;
        const int c = own ? A.queue[itm.first + lane] : 0;
        int32_t n1 = 0, st1 = 0, st2 = 0, n2 = 0;
        double sgwm = 0.0, Vc = 1.0;
        if (own) {
            n1 = sp.pia.n1[c];
            st1 = sp.pia.start1[c] - 1;
            n2 = sp.pia.n2[c] > 0 ? sp.pia.n2[c] : 0;
            st2 = sp.pia.start2[c] - 1;
            sgwm = A.sgwm[c];
            Vc = A.V[c];
        }
        const int32_t nla = n1 + n2;
        int32_t n_coll = 0, perf = 0, eqw = 0;
        // pack the item's cells into the launch's records
        for (int j = 0; j < itm.ncell; ++j) {
            const int32_t jn1 = __shfl(n1, j, 64), jst1 = __shfl(st1, j, 64), jst2 = __shfl(st2, j, 64);
            const int32_t jtot = __shfl(nla, j, 64);
            for (int32_t i = lane; i < jtot; i += 64) {
                const int64_t p = i < jn1 ? (int64_t)jst1 + i : (int64_t)jst2 + (i - jn1);
                reinterpret_cast<double4 *>(A.aos)[p] = make_double4(sp.v0[p], sp.v1[p], sp.v2[p], sp.w[p]);
            }
        }
        // records written by one lane are read by others: order the wave's
        // memory operations at each hand-over
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        for (int32_t t = 0; t < A.n_steps; ++t) {
            MzRngParams rp = A.rp;
            rp.tag = ((A.step0 + (uint64_t)t) << 4) | 1u;
            if (itm.ncell == 1 && __shfl(nla, 0, 64) >= B0_WAVE_MIN) {
                const int cw = __shfl(c, 0, 64);
                const int32_t wn1 = __shfl(n1, 0, 64), wst1 = __shfl(st1, 0, 64), wst2 = __shfl(st2, 0, 64);
                double wsg = mz_readlane_d(sgwm, 0);
                b0_ntc_cell_wave(A, rp, cw, mz_readlane_d(Vc, 0), wn1, wst1, wst2, __shfl(nla, 0, 64), wsg, n_coll,
                                 perf, eqw);
                sgwm = wsg;
            } else if (own) {
                b0_ntc_cell(A, rp, c, Vc, n1, st1, st2, nla, sgwm, n_coll, perf, eqw);
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            if (A.props_every > 0 && ((t + 1) % A.props_every) == 0) {
                for (int j = 0; j < itm.ncell; ++j) {
                    const int cj = __shfl(c, j, 64);
                    const int32_t jn1 = __shfl(n1, j, 64), jst1 = __shfl(st1, j, 64), jst2 = __shfl(st2, j, 64);
                    const int32_t jtot = __shfl(nla, j, 64);
                    auto get = [&](int32_t i, double &w, double &v0, double &v1, double &v2) {
                        const int64_t p = i < jn1 ? (int64_t)jst1 + i : (int64_t)jst2 + (i - jn1);
                        b0_get(A.aos, p, v0, v1, v2, w);
                    };
                    const MzPropsOut o = mz_props_wave(get, jtot, A.mass);
                    if (lane == 0) {
                        const int64_t ix = (int64_t)A.sidx * A.n_cells + cj;
                        A.out_np[ix] = (double)jtot;
                        A.out_n[ix] = o.n;
                        A.out_v[3 * ix] = o.v0;
                        A.out_v[3 * ix + 1] = o.v1;
                        A.out_v[3 * ix + 2] = o.v2;
                        A.out_T[ix] = o.T;
                    }
                }
            }
        }
        // write the velocities back
        for (int j = 0; j < itm.ncell; ++j) {
            const int32_t jn1 = __shfl(n1, j, 64), jst1 = __shfl(st1, j, 64), jst2 = __shfl(st2, j, 64);
            const int32_t jtot = __shfl(nla, j, 64);
            for (int32_t i = lane; i < jtot; i += 64) {
                const int64_t p = i < jn1 ? (int64_t)jst1 + i : (int64_t)jst2 + (i - jn1);
                const double4 r = reinterpret_cast<const double4 *>(A.aos)[p];
                sp.v0[p] = r.x; sp.v1[p] = r.y; sp.v2[p] = r.z;
            }
        }
        if (own) {
            A.sgwm[c] = sgwm;
            int32_t *cn = A.cnt + (int64_t)c * 5;
            cn[0] = nla;
            cn[1] = nla;
            cn[2] = n_coll;
            cn[3] = perf;
            cn[4] = eqw;
        }
    }
}

// ---------------------------------------------------------------------------
// host side

extern "C" int mz_set_batch_shape(mz_store *h, int32_t max_cells, int32_t target_particles) {
    MZ_CHECK(max_cells >= 0 && max_cells <= 64, "batch shape: max_cells must be 0..64");
    MZ_CHECK(target_particles >= 0, "batch shape: negative target");
    h->b0_max_cells = max_cells;
    h->b0_target = target_particles;
    return 0;
}

extern "C" int mz_get_batch_info(mz_store *h, int32_t *info) {
    MZ_CHECK(info != nullptr, "batch info: NULL output");
    for (int k = 0; k < 4; ++k) info[k] = h->b0_info[k];
    return 0;
}

static int batch0d_launch(mz_store *h, int species, int32_t n_steps, uint64_t step0, double dt, const double *dV,
                          int32_t props_every) {
    MzSpecies &s = h->sp[species - 1];
    const int nc = h->n_cells;
    // the queue: cells by decreasing occupancy, packed into items
    std::vector<int32_t> nloc(nc);
    MZ_HIP(hipStreamSynchronize(h->stream));   // the handle's stream does not synchronise with hipMemcpy
    MZ_HIP(hipMemcpy(nloc.data(), s.pia.nloc, nc * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int32_t> order(nc);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return nloc[a] > nloc[b]; });
    const int32_t maxc = h->b0_max_cells ? h->b0_max_cells : 64;
    const int64_t target = h->b0_target ? h->b0_target : 4096;
    std::vector<B0Item> items;
    int64_t acc = 0;
    for (int32_t q = 0; q < nc; ++q) {
        const int32_t n = nloc[order[q]];
        if (items.empty() || items.back().ncell >= maxc || (acc + n > target && items.back().ncell > 0)) {
            items.push_back(B0Item{q, 0});
            acc = 0;
        }
        items.back().ncell += 1;
        acc += n;
    }
    int32_t *d_queue = (int32_t *)mz_scratch(h, "b0_queue", (size_t)nc * sizeof(int32_t));
    B0Item *d_items = (B0Item *)mz_scratch(h, "b0_items", items.size() * sizeof(B0Item));
    int32_t *d_next = (int32_t *)mz_scratch(h, "b0_next", sizeof(int32_t));
    MZ_CHECK(d_queue && d_items && d_next, "batch0d: scratch allocation failed");
    MZ_HIP(hipMemcpyAsync(d_queue, order.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    MZ_HIP(hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(B0Item), hipMemcpyHostToDevice, h->stream));
    MZ_HIP(hipMemsetAsync(d_next, 0, sizeof(int32_t), h->stream));
    B0Args A{};
    A.sp = mz_view(s);
    A.sidx = species - 1;
    A.dt = dt;
    const MzInteraction &it = h->inter[(species - 1) * h->n_species + (species - 1)];
    A.mu1 = it.mu1; A.mu2 = it.mu2; A.factor = it.factor; A.expo = it.expo;
    MZ_HIP(hipMemcpy(&A.mass, h->d_masses + (species - 1), sizeof(double), hipMemcpyDeviceToHost));
    A.V = dV;
    const size_t pair = (size_t)(species - 1) * h->n_species + (species - 1);
    A.sgwm = h->d_sgwm + pair * nc;
    A.cnt = h->d_cfcnt + pair * nc * 5;
    A.rp = mz_rng_params(h, 0, 0, 0);
    A.step0 = step0;
    A.n_steps = n_steps;
    A.props_every = props_every;
    A.out_np = h->d_np; A.out_n = h->d_n; A.out_v = h->d_v; A.out_T = h->d_T;
    A.n_cells = nc;
    A.queue = d_queue;
    A.items = d_items;
    A.n_items = (int32_t)items.size();
    A.next = d_next;
    A.aos = (double *)mz_scratch(h, "b0_aos", (size_t)std::max<int64_t>(s.cap, 1) * 4 * sizeof(double));
    MZ_CHECK(A.aos, "batch0d: record scratch allocation failed");
    int dev = 0, n_cu = 0, per_cu = 0;
    MZ_HIP(hipGetDevice(&dev));
    MZ_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    MZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_batch0d, 256, 0));
    const int64_t waves_needed = (int64_t)items.size();
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)std::max(per_cu, 1) * n_cu,
                                                                 (waves_needed + 3) / 4));
    h->b0_info[0] = (int32_t)items.size();
    h->b0_info[1] = maxc;
    h->b0_info[2] = (int32_t)target;
    h->b0_info[3] = grid;
    MzTimed tm(h, "batch0d", n_steps);
    hipLaunchKernelGGL(k_batch0d, dim3(grid), dim3(256), 0, h->stream, A);
    MZ_HIP(hipGetLastError());
    return 0;
}

extern "C" int mz_batch0d_steps(mz_store *h, int32_t species, int32_t n_steps, uint64_t step0, double dt,
                                const double *V, int32_t props_every, const double *masses, int32_t mode) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "batch0d: bad species");
    MZ_CHECK(h->inter[(species - 1) * h->n_species + (species - 1)].set, "batch0d: interaction not set");
    MZ_CHECK(masses != nullptr, "batch0d: masses is NULL");
    MZ_CHECK(n_steps >= 0, "batch0d: negative step count");
    MZ_CHECK(mode >= 0 && mode <= 2,
             "batch0d: mode must be 0 (per-step kernels), 1 (persistent) or 2 (one NTC launch per cell)");
    MZ_HIP(hipMemcpy(h->d_masses, masses, h->n_species * sizeof(double), hipMemcpyHostToDevice));
    h->masses_valid = false;   // the couette drivers upload their mass again
    const double *dV = h->d_V;
    if (V) {
        double *v = (double *)mz_scratch(h, "b0_V", h->n_cells * sizeof(double));
        MZ_CHECK(v != nullptr, "batch0d: scratch allocation failed");
        MZ_HIP(hipMemcpy(v, V, h->n_cells * sizeof(double), hipMemcpyHostToDevice));
        dV = v;
    } else {
        MZ_CHECK(h->has_grid, "batch0d: V is NULL and no grid is set");
    }
    if (n_steps == 0) return 0;
    if (mode == 1) return batch0d_launch(h, species, n_steps, step0, dt, dV, props_every);
    // mode 0: per step the NTC kernel over all cells and the props kernel;
    // mode 2: per step one NTC launch per cell (the reference's loop of
    // ntc!(..., cell, ...) calls, each a drop-in launch), then the props kernel
    MzTimed tm(h, "batch0d", n_steps);
    for (int32_t t = 0; t < n_steps; ++t) {
        const uint64_t step = step0 + (uint64_t)t;
        if (mode == 2) {
            for (int32_t c = 1; c <= h->n_cells; ++c) {
                int r = mz_launch_ntc_equal(h, species, species, dt, dV, 1e-16, 1, (step << 4) | 1u, 0, c, c);
                if (r) return r;
            }
        } else {
            int r = mz_launch_ntc_equal(h, species, species, dt, dV, 1e-16, 1, (step << 4) | 1u, 0, 1,
                                        h->n_cells);
            if (r) return r;
        }
        if (props_every > 0 && ((t + 1) % props_every) == 0) {
            int r = mz_launch_props(h, 0, 0);
            if (r) return r;
        }
    }
    return 0;
}
