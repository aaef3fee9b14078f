// mz_ntc.hip -- NTC pair selection with VHS collisions on gfx950.
//
// Reference (Merzbild.jl, src/collisions/):
//   compute_n_coll_*                collision_ntc.jl:173-197
//   collide_2particles_vhs!         collision_ntc.jl:223-270
//   collide_2particles_vhs_equal_weight!  collision_ntc.jl:294-309
//   ntc! / ntc_equal_weight!        collision_ntc.jl:338-595
//   compute_com!, compute_g!        collision_utils.jl:374-391
//   scatter_vhs!                    collision_scattering.jl:17-29
//   sigma_vhs                       collision_cross_sections.jl:179-181
//
// The candidate loop of one cell is sequential in the reference: the running
// (sigma g w)_max is updated before every acceptance test, and particles
// created by a split become candidates of the same loop.  Each cell is
// therefore one lane; cells run in parallel.  Draw order per cell stream:
// [n_coll], then per candidate [i, k (, k redraws)], [accept] if g > eps,
// [phi, cos theta] if accepted.
//
// Particle creation (weight splitting) follows the reference's append at the
// global end: created particles are first written to per-cell scratch (the
// loop sees them as group-2 members), then appended in cell order, which is
// the order a sequential loop over cells produces; ids come from the emulated
// free-id LIFO, growing `length` by DELTA_PARTICLES when it runs dry, as
// resize! does (src/particles.jl:269-298, 311-332, 426-433).
#include "mz_phys.h"

#include <algorithm>
#include <cstring>
#include <vector>

struct NtcArgs {
    MzSpView a, b;       // species s1, s2 (same view when s1 == s2)
    int32_t single;      // s1 == s2
    int32_t equal_weight;
    int32_t may_split;
    int32_t cell_lo, cell_hi;  // 0-based inclusive
    double dt, dw_tol;
    double mu1, mu2, factor, expo;
    const double *V;
    double *sgwm;        // [n_cells] of this pair
    int32_t *cnt;        // [n_cells][5]
    MzRngParams rp;
    // split path
    const int32_t *ncoll_in;   // [n_cells] from the count pass (NULL: draw here)
    const int64_t *soff;       // [n_cells] scratch offsets
    double *sc[7];             // scratch SoA: x0 x1 x2 v0 v1 v2 w, per species half
    int64_t sc_cap;            // entries per species half
    int32_t *created;          // [2][n_cells]
};

__device__ inline int64_t slot_of(const MzSpView &sp, int c, int32_t li, int32_t n1, int32_t n2) {
    // map_cont_index (src/particles.jl:364-366) for positions present before the call
    return li < n1 ? (int64_t)sp.pia.start1[c] - 1 + li : (int64_t)sp.pia.start2[c] - 1 + (li - n1);
}

// Particle handle: either a slot of the store or a scratch entry.
struct PRef {
    double *x0, *x1, *x2, *v0, *v1, *v2, *w;
    int64_t i;
};

__device__ inline PRef pref_slot(const MzSpView &sp, int64_t p) {
    return PRef{sp.x0, sp.x1, sp.x2, sp.v0, sp.v1, sp.v2, sp.w, p};
}

__device__ inline PRef pref_scratch(const NtcArgs &A, int half, int64_t e) {
    int64_t o = (int64_t)half * A.sc_cap;
    return PRef{A.sc[0] + o, A.sc[1] + o, A.sc[2] + o, A.sc[3] + o, A.sc[4] + o, A.sc[5] + o, A.sc[6] + o, e};
}

// A cell's fixed values and the candidate loop's running state.
struct NtcCell {
    int32_t n1a, n2a, n1b, n2b, nla, nlb;
    int64_t s_base, n_coll;
};
struct NtcState {
    int32_t na, nb;       // current n_local (grows with splits)
    int32_t ca, cb;       // created per species
    int32_t performed, eqw;
    double sgwm;
};

// Particle i of species half 0 (k of half 1) of cell c: a position present
// before the call or this call's scratch entry.
__device__ inline PRef ntc_ref_a(const NtcArgs &A, int c, const NtcCell &C, int32_t i) {
    return i < C.nla ? pref_slot(A.a, slot_of(A.a, c, i, C.n1a, C.n2a)) : pref_scratch(A, 0, C.s_base + (i - C.nla));
}
__device__ inline PRef ntc_ref_b(const NtcArgs &A, int c, const NtcCell &C, int32_t k) {
    return k < C.nlb ? pref_slot(A.b, slot_of(A.b, c, k, C.n1b, C.n2b)) : pref_scratch(A, 1, C.s_base + (k - C.nlb));
}

// A cell's setup: counts, (sigma g w)_max and n_coll (draw 0 of its stream,
// or the count pass's value).
template <class Draw>
__device__ __forceinline__ void ntc_setup(const NtcArgs &A, const int c, Draw &dr, NtcCell &C, NtcState &S) {
    C.n1a = A.a.pia.n1[c]; C.n2a = A.a.pia.n2[c];
    C.n1b = A.b.pia.n1[c]; C.n2b = A.b.pia.n2[c];
    C.nla = A.a.pia.nloc[c]; C.nlb = A.b.pia.nloc[c];
    S.sgwm = A.sgwm[c];
    int64_t n_coll;
    if (A.ncoll_in) {
        n_coll = A.ncoll_in[c];
        dr.seek(1);
    } else {
        double R = dr.next();
        n_coll = (int64_t)floor(mz_ncoll_est(A.single, A.dt, C.nla, C.nlb, S.sgwm, A.V[c], R));
    }
    if (!(n_coll <= MZ_NTC_MAX_CANDIDATES)) {   // also catches NaN estimates
        atomicOr(&A.a.ctr->err, MZ_ERR_SCRATCH);
        n_coll = 0;
    }
    C.n_coll = n_coll;
    S.na = C.nla; S.nb = C.nlb;
    S.ca = 0; S.cb = 0;
    S.performed = 0; S.eqw = 0;
    C.s_base = A.may_split ? A.soff[c] : 0;
}

// Whether this call's scratch has room for the split a collision of the
// pair makes (always true when it makes none).  Without room the scratch
// error is raised and the candidate ends before its scattering draws, as a
// performed collision.
__device__ inline bool ntc_split_room(const NtcArgs &A, const NtcCell &C, const NtcState &S, double wi, double wk) {
    if (A.equal_weight || fabs(wi - wk) < A.dw_tol || !A.may_split) return true;
    const bool on_a = wi > wk || A.single;
    const int32_t q = on_a ? (A.single ? S.ca + S.cb : S.ca) : S.cb;
    if (q < C.n_coll) return true;
    atomicOr(&(on_a ? A.a : A.b).ctr->err, MZ_ERR_SCRATCH);
    return false;
}

// The collision of a pair that passed the acceptance test, with room for
// its split (collide_2particles_vhs! with weight splitting,
// collision_ntc.jl:223-270): the split (the heavier particle leaves a copy of
// itself with the weight difference and takes the lighter's weight), then
// the post-collision velocities u.
__device__ __forceinline__ void ntc_collide(const NtcArgs &A, const NtcCell &C, NtcState &S, const PRef &pi,
                                            const PRef &pk, double vi0, double vi1, double vi2, double vk0,
                                            double vk1, double vk2, double wi, double wk, double ui0, double ui1,
                                            double ui2, double uk0, double uk1, double uk2) {
    ++S.performed;
    if (A.equal_weight || fabs(wi - wk) < A.dw_tol) {
        ++S.eqw;
    } else if (!A.may_split) {
        atomicOr(&A.a.ctr->err, MZ_ERR_SCRATCH);
    } else if (wi > wk) {
        PRef nw = pref_scratch(A, 0, C.s_base + (A.single ? S.ca + S.cb : S.ca));
        nw.x0[nw.i] = pi.x0[pi.i]; nw.x1[nw.i] = pi.x1[pi.i]; nw.x2[nw.i] = pi.x2[pi.i];
        nw.v0[nw.i] = vi0; nw.v1[nw.i] = vi1; nw.v2[nw.i] = vi2;
        nw.w[nw.i] = wi - wk;
        pi.w[pi.i] = wk;
        ++S.ca; ++S.na;
    } else if (A.single) {
        PRef nw = pref_scratch(A, 0, C.s_base + S.ca + S.cb);
        nw.x0[nw.i] = pk.x0[pk.i]; nw.x1[nw.i] = pk.x1[pk.i]; nw.x2[nw.i] = pk.x2[pk.i];
        nw.v0[nw.i] = vk0; nw.v1[nw.i] = vk1; nw.v2[nw.i] = vk2;
        nw.w[nw.i] = wk - wi;
        pk.w[pk.i] = wi;
        ++S.ca; ++S.na;
    } else {
        PRef nw = pref_scratch(A, 1, C.s_base + S.cb);
        nw.x0[nw.i] = pk.x0[pk.i]; nw.x1[nw.i] = pk.x1[pk.i]; nw.x2[nw.i] = pk.x2[pk.i];
        nw.v0[nw.i] = vk0; nw.v1[nw.i] = vk1; nw.v2[nw.i] = vk2;
        nw.w[nw.i] = wk - wi;
        pk.w[pk.i] = wi;
        ++S.cb; ++S.nb;
    }
    pi.v0[pi.i] = ui0; pi.v1[pi.i] = ui1; pi.v2[pi.i] = ui2;
    pk.v0[pk.i] = uk0; pk.v1[pk.i] = uk1; pk.v2[pk.i] = uk2;
}

// One candidate of cell c's loop (collision_ntc.jl:338-380); false when a
// replay row ran out.
template <class Draw>
__device__ __forceinline__ bool ntc_candidate(const NtcArgs &A, const int c, const NtcCell &C, NtcState &S,
                                              Draw &dr) {
    int32_t i = (int32_t)floor(dr.next() * (double)S.na);
    int32_t k = (int32_t)floor(dr.next() * (double)(A.single ? S.na : S.nb));
    if (A.single) {
        while (i == k) {
            k = (int32_t)floor(dr.next() * (double)S.na);
            if (dr.exhausted()) return false;   // replay row exhausted
        }
    }
    const PRef pi = ntc_ref_a(A, c, C, i);
    const PRef pk = A.single ? ntc_ref_a(A, c, C, k) : ntc_ref_b(A, c, C, k);
    const double vi0 = pi.v0[pi.i], vi1 = pi.v1[pi.i], vi2 = pi.v2[pi.i];
    const double vk0 = pk.v0[pk.i], vk1 = pk.v1[pk.i], vk2 = pk.v2[pk.i];
    double g0, g1, g2;
    const double g = mz_rel_speed(vi0, vi1, vi2, vk0, vk1, vk2, g0, g1, g2);
    if (!(g > MZ_G_EPS)) return true;
    const double wi = pi.w[pi.i], wk = pk.w[pk.i];
    const double sgw = mz_vhs_sgw(A.factor, A.expo, g, wi, wk);
    S.sgwm = sgw > S.sgwm ? sgw : S.sgwm;
    if (!(dr.next() < sgw / S.sgwm)) return true;
    if (!ntc_split_room(A, C, S, wi, wk)) {
        ++S.performed;
        return true;
    }
    double vc0, vc1, vc2;
    mz_com(A.mu1, A.mu2, vi0, vi1, vi2, vk0, vk1, vk2, vc0, vc1, vc2);
    // scatter_vhs!: phi first, then cos(theta)
    const double R_phi = dr.next();
    const double R_ct = dr.next();
    double ui0, ui1, ui2, uk0, uk1, uk2;
    mz_vhs_post(A.mu1, A.mu2, g, vc0, vc1, vc2, R_phi, R_ct, ui0, ui1, ui2, uk0, uk1, uk2);
    ntc_collide(A, C, S, pi, pk, vi0, vi1, vi2, vk0, vk1, vk2, wi, wk, ui0, ui1, ui2, uk0, uk1, uk2);
    return true;
}

// The end of a cell's loop: (sigma g w)_max, the counters, created counts.
__device__ __forceinline__ void ntc_finish_cell(const NtcArgs &A, const int c, const NtcCell &C,
                                                const NtcState &S) {
    A.sgwm[c] = S.sgwm;
    int32_t *cn = A.cnt + (int64_t)c * 5;
    cn[0] = C.nla;
    cn[1] = A.single ? C.nla : C.nlb;
    cn[2] = (int32_t)C.n_coll;
    cn[3] = S.performed;
    cn[4] = S.eqw;
    if (A.may_split) {
        A.created[c] = A.single ? S.ca + S.cb : S.ca;
        A.created[(int64_t)(A.cell_hi + 1) + c] = A.single ? 0 : S.cb;
    }
}

// The candidate loop of cell c with the draws of its stream from dr (next(),
// seek(k), exhausted()).
template <class Draw>
__device__ __forceinline__ void ntc_cell(const NtcArgs &A, const int c, Draw &dr) {
    NtcCell C;
    NtcState S;
    ntc_setup(A, c, dr, C, S);
    for (int64_t cc = 0; cc < C.n_coll; ++cc)
        if (!ntc_candidate(A, c, C, S, dr)) return;
    ntc_finish_cell(A, c, C, S);
}

// Draws of a cell's stream for one lane: mz_rand.
struct NtcLaneDraw {
    const MzRngParams &rp;
    MzStream st;
    __device__ double next() { return mz_rand(rp, st); }
    __device__ bool exhausted() const { return rp.mode == 1 && (int64_t)st.k > rp.cols; }
    __device__ void seek(uint64_t k) { st.k = k; }
};

// Draws of a cell's stream for a whole wavefront running the cell (Philox
// draws only): lane L holds Philox block base / 2 + L, i.e. draws base + 2 L
// and base + 2 L + 1, and every lane reads draw k from its lane; the block
// of 128 draws is computed again, all lanes at once, when k leaves it.
struct NtcWaveDraw {
    const MzRngParams &rp;
    uint64_t stream, k, base;
    double h0, h1;
    __device__ double next() {
        const uint64_t kk = k++;
        if (kk < base || kk >= base + 128) {
            base = kk & ~(uint64_t)1;
            mz_philox_pair(rp, stream, (base >> 1) + (uint64_t)(threadIdx.x & 63), h0, h1);
        }
        const uint64_t idx = kk - base;
        const int l = (int)(idx >> 1);
        const double a = mz_readlane_d(h0, l), b = mz_readlane_d(h1, l);
        return (idx & 1) ? b : a;
    }
    __device__ bool exhausted() const { return false; }
    __device__ void seek(uint64_t kk) { k = kk; }
    // the block holds draws D .. D + 67 afterwards
    __device__ void cover(uint64_t D) {
        if (D < base || D + 68 > base + 128) {
            base = D & ~(uint64_t)1;
            mz_philox_pair(rp, stream, (base >> 1) + (uint64_t)(threadIdx.x & 63), h0, h1);
        }
    }
    // draw s for this lane (s may differ between lanes; base <= s < base + 128)
    __device__ double at(uint64_t s) const {
        const uint64_t idx = s - base;
        const int l = (int)(idx >> 1);
        const double a = __shfl(h0, l, 64), b = __shfl(h1, l, 64);
        return (idx & 1) ? b : a;
    }
};

// ntc_cell for a single-species cell on a whole wavefront with Philox
// draws, with the candidates evaluated ahead: at a window's start at draw D,
// lane L evaluates the candidate that would start at draw D + L (pair,
// relative speed, (sigma g w), the acceptance test under the window's
// (sigma g w)_max, post-collision velocities); the wave then walks the
// candidates in the reference's order on uniform values.  Each starts where
// the previous one ended, (sigma g w)_max is updated before its acceptance
// test, and a lane's test is taken only while (sigma g w)_max is unchanged
// (the same operands, so the same result).  A candidate whose k is redrawn,
// or that shares a particle with a collision earlier in the window, is
// evaluated directly and a new window starts after it; so does a split,
// which changes n_local and with it every later pair.  The results are
// ntc_cell's.
__device__ void ntc_cell_spec(const NtcArgs &A, const int c, NtcWaveDraw &dr) {
    const int lane = threadIdx.x & 63;
    NtcCell C;
    NtcState S;
    ntc_setup(A, c, dr, C, S);
    int64_t cc = 0;
    uint64_t o = dr.k;   // the next draw of the stream
    while (cc < C.n_coll) {
        const uint64_t D = o;
        dr.cover(D);
        const int32_t na0 = S.na;
        const double sgwm0 = S.sgwm;
        const uint64_t sL = D + (uint64_t)lane;
        const double d0 = dr.at(sL), d1 = dr.at(sL + 1), d2 = dr.at(sL + 2), d3 = dr.at(sL + 3), d4 = dr.at(sL + 4);
        const int32_t si = (int32_t)floor(d0 * (double)na0), sk = (int32_t)floor(d1 * (double)na0);
        double swi, swk, sgw, vi0, vi1, vi2, vk0, vk1, vk2, ui0, ui1, ui2, uk0, uk1, uk2;
        int32_t f;
        {
            const PRef pi = ntc_ref_a(A, c, C, si), pk = ntc_ref_a(A, c, C, sk);
            vi0 = pi.v0[pi.i]; vi1 = pi.v1[pi.i]; vi2 = pi.v2[pi.i];
            vk0 = pk.v0[pk.i]; vk1 = pk.v1[pk.i]; vk2 = pk.v2[pk.i];
            swi = pi.w[pi.i];
            swk = pk.w[pk.i];
            double g0, g1, g2;
            const double g = mz_rel_speed(vi0, vi1, vi2, vk0, vk1, vk2, g0, g1, g2);
            sgw = mz_vhs_sgw(A.factor, A.expo, g, swi, swk);
            f = (si == sk ? 1 : 0) | (g > MZ_G_EPS ? 2 : 0) | (sgw > sgwm0 ? 4 : 0) | (d2 < sgw / sgwm0 ? 8 : 0);
            double vc0, vc1, vc2;
            mz_com(A.mu1, A.mu2, vi0, vi1, vi2, vk0, vk1, vk2, vc0, vc1, vc2);
            mz_vhs_post(A.mu1, A.mu2, g, vc0, vc1, vc2, d3, d4, ui0, ui1, ui2, uk0, uk1, uk2);
        }
        // walk the window; lane j < nd holds a particle that collided in it
        int32_t dirty = -1;
        int nd = 0;
        bool direct = false;
        while (cc < C.n_coll) {
            const int L = (int)(o - D);
            if (L >= 60 || nd >= 62) break;
            const int32_t i = __builtin_amdgcn_readlane(si, L), k = __builtin_amdgcn_readlane(sk, L);
            const int32_t fl = __builtin_amdgcn_readlane(f, L);
            if ((fl & 1) || __ballot(lane < nd && (dirty == i || dirty == k))) {
                direct = true;
                break;
            }
            if (!(fl & 2)) {   // g <= eps: no acceptance draw
                o += 2;
                ++cc;
                continue;
            }
            bool acc;
            if (!(fl & 4) && S.sgwm == sgwm0) {
                acc = (fl & 8) != 0;
            } else {
                const double sl = mz_readlane_d(sgw, L);
                S.sgwm = sl > S.sgwm ? sl : S.sgwm;
                acc = mz_readlane_d(d2, L) < sl / S.sgwm;
            }
            if (!acc) {
                o += 3;
                ++cc;
                continue;
            }
            const double wi = mz_readlane_d(swi, L), wk = mz_readlane_d(swk, L);
            if (!ntc_split_room(A, C, S, wi, wk)) {
                ++S.performed;
                o += 3;
                ++cc;
                continue;
            }
            const int32_t na_before = S.na;
            ntc_collide(A, C, S, ntc_ref_a(A, c, C, i), ntc_ref_a(A, c, C, k), mz_readlane_d(vi0, L),
                        mz_readlane_d(vi1, L), mz_readlane_d(vi2, L), mz_readlane_d(vk0, L), mz_readlane_d(vk1, L),
                        mz_readlane_d(vk2, L), wi, wk, mz_readlane_d(ui0, L), mz_readlane_d(ui1, L),
                        mz_readlane_d(ui2, L), mz_readlane_d(uk0, L), mz_readlane_d(uk1, L), mz_readlane_d(uk2, L));
            if (lane == nd) dirty = i;
            if (lane == nd + 1) dirty = k;
            nd += 2;
            o += 5;
            ++cc;
            if (S.na != na_before) break;   // a split changed n_local
        }
        // the stores of the window before the loads that follow
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        if (direct) {
            dr.seek(o);
            ntc_candidate(A, c, C, S, dr);   // Philox draws: never exhausted
            o = dr.k;
            ++cc;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        }
    }
    ntc_finish_cell(A, c, C, S);
}

// One lane per cell; cells run in parallel (64-lane workgroups, so the
// loop keeps its state in registers).
__global__ void __launch_bounds__(64) k_ntc(NtcArgs A) {
    const int c = A.cell_lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (c > A.cell_hi) return;
    A.rp = mz_rng_resolve(A.rp);
    NtcLaneDraw dr{A.rp, mz_stream((uint64_t)c)};
    ntc_cell(A, c, dr);
}

// One wavefront per cell (Philox draws): single-species cells take their
// candidates a window at a time (ntc_cell_spec), two-species cells run the
// loop on every lane on the same values; the draws come from registers.  For
// calls over few cells, which one lane per cell would spread over a handful
// of wavefronts.
__global__ void __launch_bounds__(64) k_ntc_wave(NtcArgs A) {
    const int c = A.cell_lo + blockIdx.x;
    if (c > A.cell_hi) return;
    A.rp = mz_rng_resolve(A.rp);
    NtcWaveDraw dr{A.rp, (uint64_t)c, 0, ~(uint64_t)0, 0.0, 0.0};
    if (A.single) ntc_cell_spec(A, c, dr);
    else ntc_cell(A, c, dr);
}

// Count pass of the split path: n_coll of every cell (draw 0 of its stream).
__global__ void k_ntc_count(NtcArgs A, int32_t *__restrict__ ncoll) {
    const int c = A.cell_lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (c > A.cell_hi) return;
    const int32_t nla = A.a.pia.nloc[c], nlb = A.b.pia.nloc[c];
    A.rp = mz_rng_resolve(A.rp);
    MzStream st = mz_stream((uint64_t)c);
    double R = mz_rand(A.rp, st);
    double sgwm = A.sgwm[c];
    double f = floor(mz_ncoll_est(A.single, A.dt, nla, nlb, sgwm, A.V[c], R));
    if (!(f <= (double)MZ_NTC_MAX_CANDIDATES)) {
        atomicOr(&A.a.ctr->err, MZ_ERR_SCRATCH);
        f = 0.0;
    }
    ncoll[c] = (int32_t)f;
}

// Exclusive scan of n int32 values into int64 offsets (one 1024-thread block);
// total goes to *total.
__global__ void __launch_bounds__(1024) k_scan_i32(const int32_t *__restrict__ in, int64_t n,
                                                   int64_t *__restrict__ out, int64_t *__restrict__ total) {
    __shared__ int64_t wsum[16];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int64_t per = (n + 1023) / 1024;
    const int64_t lo = t * per, hi = min(n, lo + per);
    int64_t s = 0;
    for (int64_t i = lo; i < hi; ++i) s += in[i];
    int64_t x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    if (t < 16) {
        int64_t a = wsum[t], b = a;
        for (int o = 1; o < 16; o <<= 1) {
            int64_t y = __shfl_up(b, o, 16);
            if ((t & 15) >= o) b += y;
        }
        wsum[t] = b - a;
    }
    __syncthreads();
    int64_t run = x - s + wsum[wv];
    for (int64_t i = lo; i < hi; ++i) {
        int64_t v = in[i];
        out[i] = run;
        run += v;
    }
    if (t == 1023) *total = run;
}

// Append pass: created particles of cell c go to positions
// n_total_old + coff[c] + j (1-based +1), ids from the LIFO in creation order.
__global__ void k_ntc_append(MzSpView sp, int32_t cell_lo, int32_t cell_hi, const int32_t *__restrict__ created,
                             const int64_t *__restrict__ coff, const int64_t *__restrict__ ctotal,
                             const int64_t *__restrict__ soff, const double *const *sc, int64_t sc_off,
                             MzCounters snap) {
    const int c = cell_lo + blockIdx.x;
    if (c > cell_hi) return;
    const int32_t k = created[c];
    if (k == 0) return;
    const int64_t base = (int64_t)snap.n_total + coff[c];  // 0-based first new slot
    for (int j = threadIdx.x; j < k; j += blockDim.x) {
        int64_t dst = base + j;
        if (dst >= sp.cap) {
            atomicOr(&sp.ctr->err, MZ_ERR_CAPACITY);
            continue;
        }
        int64_t src = sc_off + soff[c] + j;
        sp.x0[dst] = sc[0][src]; sp.x1[dst] = sc[1][src]; sp.x2[dst] = sc[2][src];
        sp.v0[dst] = sc[3][src]; sp.v1[dst] = sc[4][src]; sp.v2[dst] = sc[5][src];
        sp.w[dst] = sc[6][src];
        int64_t q = coff[c] + j;   // global creation number
        int32_t idv = q < snap.nbuffer ? sp.buffer[snap.nbuffer - 1 - q]
                                       : (int32_t)(snap.length + 1 + (q - snap.nbuffer));
        sp.id[dst] = idv;
        sp.key[dst] = c + 1;
    }
    if (threadIdx.x == 0) {
        // update_particle_indexer_new_particle! applied k times
        sp.pia.nloc[c] += k;
        sp.pia.n2[c] += k;
        if (sp.pia.start2[c] <= 0) sp.pia.start2[c] = (int32_t)(base + 1);
        sp.pia.end2[c] = (int32_t)(base + k);
    }
    (void)ctotal;
}

// Counters after the append: n_total, n_pos, and the LIFO / length as a run of
// resize! + pops leaves them.
__global__ void k_ntc_finish(MzSpView sp, const int64_t *__restrict__ ctotal, MzCounters snap) {
    const int64_t tot = *ctotal;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sp.ctr->created = (int32_t)tot;
        if (tot == 0) return;
        int64_t nt = (int64_t)snap.n_total + tot;
        if (nt > sp.cap) {
            atomicOr(&sp.ctr->err, MZ_ERR_CAPACITY);
            nt = sp.cap;
        }
        sp.ctr->n_total = (int32_t)nt;
        sp.ctr->n_pos = max(snap.n_pos, sp.ctr->n_total);
        if (tot <= snap.nbuffer) {
            sp.ctr->nbuffer = snap.nbuffer - (int32_t)tot;
        } else {
            int64_t R = tot - snap.nbuffer;
            int64_t chunks = (R + MZ_DELTA_PARTICLES - 1) / MZ_DELTA_PARTICLES;
            int64_t L2 = snap.length + chunks * MZ_DELTA_PARTICLES;
            sp.ctr->length = (int32_t)L2;
            sp.ctr->nbuffer = (int32_t)(L2 - snap.length - R);
        }
    }
    // the remaining free ids after a resize: buffer[0..nbuffer) = L', L'-1, ...
    if (tot > snap.nbuffer) {
        int64_t R = tot - snap.nbuffer;
        int64_t chunks = (R + MZ_DELTA_PARTICLES - 1) / MZ_DELTA_PARTICLES;
        int64_t L2 = snap.length + chunks * MZ_DELTA_PARTICLES;
        int64_t nbuf = L2 - snap.length - R;
        if (nbuf > sp.bufcap) {
            if (threadIdx.x == 0) atomicOr(&sp.ctr->err, MZ_ERR_BUFFER);
            return;
        }
        for (int64_t i = threadIdx.x; i < nbuf; i += blockDim.x) sp.buffer[i] = (int32_t)(L2 - i);
    }
}

static NtcArgs base_args(mz_store *h, int s1, int s2, double dt, const double *dV, double dw_tol, int equal_weight,
                         uint64_t tag, int rng_mode, int cell_lo, int cell_hi) {
    NtcArgs A{};
    A.a = mz_view(h->sp[s1 - 1]);
    A.b = mz_view(h->sp[s2 - 1]);
    A.single = s1 == s2;
    A.equal_weight = equal_weight;
    A.cell_lo = cell_lo - 1;
    A.cell_hi = cell_hi - 1;
    A.dt = dt;
    A.dw_tol = dw_tol;
    const MzInteraction &it = h->inter[(s1 - 1) * h->n_species + (s2 - 1)];
    A.mu1 = it.mu1; A.mu2 = it.mu2; A.factor = it.factor; A.expo = it.expo;
    A.V = dV;
    size_t pair = (size_t)(s1 - 1) * h->n_species + (s2 - 1);
    A.sgwm = h->d_sgwm + pair * h->n_cells;
    A.cnt = h->d_cfcnt + pair * h->n_cells * 5;
    A.rp = mz_rng_params(h, 0, tag, rng_mode);
    return A;
}

// The candidate loops over ncells cells: a wavefront per cell for calls over
// few cells (Philox draws), one lane per cell otherwise.
#define NTC_WAVE_MAX_CELLS 4096
static void launch_k_ntc(mz_store *h, const NtcArgs &A, int ncells) {
    if (ncells <= NTC_WAVE_MAX_CELLS && A.rp.mode == 0)
        hipLaunchKernelGGL(k_ntc_wave, dim3(ncells), dim3(64), 0, h->stream, A);
    else
        hipLaunchKernelGGL(k_ntc, dim3((ncells + 63) / 64), dim3(64), 0, h->stream, A);
}

int mz_launch_ntc_equal(mz_store *h, int s1, int s2, double dt, const double *dV, double dw_tol, int equal_weight,
                        uint64_t tag, int rng_mode, int cell_lo, int cell_hi, const uint64_t *step_ptr) {
    NtcArgs A = base_args(h, s1, s2, dt, dV, dw_tol, equal_weight, tag, rng_mode, cell_lo, cell_hi);
    A.may_split = 0;
    A.rp.step_ptr = step_ptr;
    A.rp.phase = 1;
    int ncells = cell_hi - cell_lo + 1;
    launch_k_ntc(h, A, ncells);
    MZ_HIP(hipGetLastError());
    return 0;
}

static int launch_ntc_split(mz_store *h, int s1, int s2, double dt, const double *dV, double dw_tol, uint64_t tag,
                            int rng_mode, int cell_lo, int cell_hi) {
    NtcArgs A = base_args(h, s1, s2, dt, dV, dw_tol, 0, tag, rng_mode, cell_lo, cell_hi);
    A.may_split = 1;
    const int nc = h->n_cells;
    const int ncells = cell_hi - cell_lo + 1;
    int32_t *ncoll = (int32_t *)mz_scratch(h, "ntc_ncoll", (size_t)nc * sizeof(int32_t));
    int64_t *soff = (int64_t *)mz_scratch(h, "ntc_soff", (size_t)(nc + 1) * sizeof(int64_t));
    int32_t *created = (int32_t *)mz_scratch(h, "ntc_created", (size_t)2 * nc * sizeof(int32_t));
    int64_t *coff = (int64_t *)mz_scratch(h, "ntc_coff", (size_t)2 * (nc + 1) * sizeof(int64_t));
    MZ_CHECK(ncoll && soff && created && coff, "ntc: scratch allocation failed");
    MZ_HIP(hipMemsetAsync(ncoll, 0, (size_t)nc * sizeof(int32_t), h->stream));
    MZ_HIP(hipMemsetAsync(created, 0, (size_t)2 * nc * sizeof(int32_t), h->stream));
    hipLaunchKernelGGL(k_ntc_count, dim3((ncells + 255) / 256), dim3(256), 0, h->stream, A, ncoll);
    hipLaunchKernelGGL(k_scan_i32, dim3(1), dim3(1024), 0, h->stream, ncoll, (int64_t)nc, soff, soff + nc);
    int64_t total = 0;
    MZ_HIP(hipMemcpyAsync(&total, soff + nc, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    MZ_HIP(hipStreamSynchronize(h->stream));
    int64_t cap = std::max<int64_t>(total, 1);
    double *sc = (double *)mz_scratch(h, "ntc_sc", (size_t)7 * 2 * cap * sizeof(double));
    MZ_CHECK(sc != nullptr, "ntc: creation scratch allocation failed");
    for (int f = 0; f < 7; ++f) A.sc[f] = sc + (size_t)f * 2 * cap;
    A.sc_cap = cap;
    A.ncoll_in = ncoll;
    A.soff = soff;
    // created is indexed [species half][cell]; the kernel writes half 1 at cell_hi + 1 + c
    A.created = created;
    launch_k_ntc(h, A, ncells);
    MZ_HIP(hipGetLastError());
    // append per species half
    double **scp = (double **)mz_scratch(h, "ntc_scp", 7 * sizeof(double *));
    MZ_CHECK(scp != nullptr, "ntc: scratch allocation failed");
    MZ_HIP(hipMemcpy(scp, A.sc, 7 * sizeof(double *), hipMemcpyHostToDevice));
    for (int half = 0; half < (A.single ? 1 : 2); ++half) {
        int sidx = half == 0 ? s1 : s2;
        const int32_t *cr = created + (half == 0 ? 0 : (int64_t)(A.cell_hi + 1));
        int64_t *co = coff + (size_t)half * (nc + 1);
        MzCounters snap;
        int r = mz_read_counters(h, sidx, &snap);
        if (r) return r;
        // created[] for half 1 is offset by cell_hi + 1 entries; scan the used range
        hipLaunchKernelGGL(k_scan_i32, dim3(1), dim3(1024), 0, h->stream, cr, (int64_t)(A.cell_hi + 1), co,
                           co + nc);
        hipLaunchKernelGGL(k_ntc_append, dim3(ncells), dim3(64), 0, h->stream, mz_view(h->sp[sidx - 1]), A.cell_lo,
                           A.cell_hi, cr, co, co + nc, soff, (const double *const *)scp, (int64_t)half * cap, snap);
        hipLaunchKernelGGL(k_ntc_finish, dim3(1), dim3(256), 0, h->stream, mz_view(h->sp[sidx - 1]), co + nc, snap);
        MZ_HIP(hipGetLastError());
    }
    return 0;
}

// The cell volumes V[n_cells] of an NTC call on the device.  They upload
// only when they differ from the last upload; an upload first waits for the
// store's stream, whose queued kernels may still read the previous values.
static int upload_ntc_volumes(mz_store *h, const double *V, const double **out) {
    double *v = (double *)mz_scratch(h, "ntc_V", h->n_cells * sizeof(double));
    MZ_CHECK(v != nullptr, "ntc: scratch allocation failed");
    if (h->ntc_V_up.size() != (size_t)h->n_cells ||
        std::memcmp(h->ntc_V_up.data(), V, h->n_cells * sizeof(double)) != 0 || h->ntc_V_dev != v) {
        MZ_HIP(hipStreamSynchronize(h->stream));
        MZ_HIP(hipMemcpy(v, V, h->n_cells * sizeof(double), hipMemcpyHostToDevice));
        h->ntc_V_up.assign(V, V + h->n_cells);
        h->ntc_V_dev = v;
    }
    *out = v;
    return 0;
}

extern "C" int mz_ntc(mz_store *h, int32_t s1, int32_t s2, double dt, const double *V, double dw_tol,
                      int32_t equal_weight, uint64_t tag, int32_t rng_mode, int32_t cell_lo, int32_t cell_hi) {
    MZ_CHECK(s1 >= 1 && s1 <= h->n_species && s2 >= 1 && s2 <= h->n_species, "ntc: bad species");
    MZ_CHECK(cell_lo >= 1 && cell_hi <= h->n_cells && cell_lo <= cell_hi, "ntc: bad cell range");
    MZ_CHECK(h->inter[(s1 - 1) * h->n_species + (s2 - 1)].set, "ntc: interaction not set");
    MZ_CHECK(rng_mode == 0 || (rng_mode == 1 && h->d_replay[0]), "ntc: replay table not set");
    const double *dV = h->d_V;
    if (V) {
        int r = upload_ntc_volumes(h, V, &dV);
        if (r) return r;
    } else {
        MZ_CHECK(h->has_grid, "ntc: V is NULL and no grid is set");
    }
    MzTimed tm(h, "ntc");
    if (equal_weight) return mz_launch_ntc_equal(h, s1, s2, dt, dV, dw_tol, 1, tag, rng_mode, cell_lo, cell_hi);
    return launch_ntc_split(h, s1, s2, dt, dV, dw_tol, tag, rng_mode, cell_lo, cell_hi);
}

// The reference's loop of per-cell ntc! calls (test/test_1D_couette.jl:67-71)
// as one call: cell cells[k] with volume V[k], in list order.  A maximal run
// of consecutive ascending cells launches as one range (its cells are
// independent and its splits append in cell order), runs launch one after
// another, so the result is the sequential loop's for any list.
extern "C" int mz_ntc_cells(mz_store *h, int32_t s1, int32_t s2, double dt, const int32_t *cells, const double *V,
                            int32_t n, double dw_tol, int32_t equal_weight, uint64_t tag, int32_t rng_mode) {
    MZ_CHECK(s1 >= 1 && s1 <= h->n_species && s2 >= 1 && s2 <= h->n_species, "ntc_cells: bad species");
    MZ_CHECK(h->inter[(s1 - 1) * h->n_species + (s2 - 1)].set, "ntc_cells: interaction not set");
    MZ_CHECK(rng_mode == 0 || (rng_mode == 1 && h->d_replay[0]), "ntc_cells: replay table not set");
    MZ_CHECK(n >= 0 && (n == 0 || (cells != nullptr && V != nullptr)), "ntc_cells: bad cell list");
    for (int32_t k = 0; k < n; ++k) MZ_CHECK(cells[k] >= 1 && cells[k] <= h->n_cells, "ntc_cells: cell out of range");
    if (n == 0) return 0;
    std::vector<double> full = h->ntc_V_up.size() == (size_t)h->n_cells ? h->ntc_V_up
                                                                           : std::vector<double>(h->n_cells, 1.0);
    MzTimed tm(h, "ntc");
    int32_t k0 = 0;
    while (k0 < n) {
        int32_t k1 = k0 + 1;
        while (k1 < n && cells[k1] == cells[k1 - 1] + 1) ++k1;
        for (int32_t k = k0; k < k1; ++k) full[cells[k] - 1] = V[k];
        const double *dv = nullptr;
        int r = upload_ntc_volumes(h, full.data(), &dv);
        if (r) return r;
        const int32_t lo = cells[k0], hi = cells[k1 - 1];
        r = equal_weight ? mz_launch_ntc_equal(h, s1, s2, dt, dv, dw_tol, 1, tag, rng_mode, lo, hi)
                         : launch_ntc_split(h, s1, s2, dt, dv, dw_tol, tag, rng_mode, lo, hi);
        if (r) return r;
        k0 = k1;
    }
    return 0;
}

// ===========================================================================
// swpm! (collision_swpm.jl:201-288): single-species collisions of
// variable-weight particles.  One lane per cell, as NTC: the loop is
// sequential (the running (sigma g)_max, and particles created by the loop
// are candidates of the same loop).  An accepted pair gives up
// min(w_i, w_k) / (1 + G) of each weight to two new particles at the pair's
// positions, which then take the pair's post-collision velocities
// (scatter_vhs!); the pair keeps its velocities.  Created particles go to
// per-cell scratch and are appended in cell order as for NTC splits.  For
// this method the pair's (sigma g w)_max array holds (sigma g)_max.

// Count pass: w_max of every cell, its candidate count (draw 0 of its
// stream, compute_n_coll_single_species, collision_swpm.jl:171-174) and the
// scratch it may need (two particles per candidate); one wave per cell.
__global__ void __launch_bounds__(256) k_swpm_count(NtcArgs A, double G, int32_t *__restrict__ ncoll,
                                                    int32_t *__restrict__ ssz, double *__restrict__ wmax) {
    const int c = A.cell_lo + blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c > A.cell_hi) return;
    const int32_t n1 = A.a.pia.n1[c], n2 = A.a.pia.n2[c] > 0 ? A.a.pia.n2[c] : 0, np = A.a.pia.nloc[c];
    double w_max = 0.0;
    for (int32_t li = lane; li < n1 + n2; li += 64) {
        const double w = A.a.w[slot_of(A.a, c, li, n1, n2)];
        w_max = w > w_max ? w : w_max;
    }
    w_max = mz_wave_max(w_max);
    if (lane != 0) return;
    A.rp = mz_rng_resolve(A.rp);
    MzStream st = mz_stream((uint64_t)c);
    const double R = mz_rand(A.rp, st);
    double f = floor(0.5 * A.dt * (double)np * (double)(np - 1) * A.sgwm[c] * w_max * (G + 1) / A.V[c] + R);
    if (!(f <= (double)MZ_NTC_MAX_CANDIDATES / 2)) {   // also catches NaN estimates
        atomicOr(&A.a.ctr->err, MZ_ERR_SCRATCH);
        f = 0.0;
    }
    ncoll[c] = (int32_t)f;
    ssz[c] = 2 * (int32_t)f;
    wmax[c] = w_max;
}

__global__ void __launch_bounds__(64) k_swpm(NtcArgs A, double G, const double *__restrict__ wmax) {
    const int c = A.cell_lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (c > A.cell_hi) return;
    A.rp = mz_rng_resolve(A.rp);
    const int32_t n1 = A.a.pia.n1[c], n2 = A.a.pia.n2[c] > 0 ? A.a.pia.n2[c] : 0;
    const int32_t nl = A.a.pia.nloc[c];
    const double wtf = 1.0 / (1.0 + G);
    const double inv_w_max = 1.0 / wmax[c];
    double sgmax = A.sgwm[c];
    MzStream st = mz_stream((uint64_t)c);
    st.k = 1;
    const int64_t n_coll = A.ncoll_in[c];
    const int64_t s_base = A.soff[c];
    int32_t na = nl, created = 0, performed = 0;
    for (int64_t cc = 0; cc < n_coll; ++cc) {
        const int32_t i = (int32_t)floor(mz_rand(A.rp, st) * (double)na);
        int32_t k = (int32_t)floor(mz_rand(A.rp, st) * (double)na);
        while (i == k) {
            k = (int32_t)floor(mz_rand(A.rp, st) * (double)na);
            if (A.rp.mode == 1 && (int64_t)st.k > A.rp.cols) return;   // replay row exhausted
        }
        PRef pi = i < nl ? pref_slot(A.a, slot_of(A.a, c, i, n1, n2)) : pref_scratch(A, 0, s_base + (i - nl));
        PRef pk = k < nl ? pref_slot(A.a, slot_of(A.a, c, k, n1, n2)) : pref_scratch(A, 0, s_base + (k - nl));
        const double vi0 = pi.v0[pi.i], vi1 = pi.v1[pi.i], vi2 = pi.v2[pi.i];
        const double vk0 = pk.v0[pk.i], vk1 = pk.v1[pk.i], vk2 = pk.v2[pk.i];
        double g0, g1, g2;
        const double g = mz_rel_speed(vi0, vi1, vi2, vk0, vk1, vk2, g0, g1, g2);
        if (!(g > MZ_G_EPS)) continue;
        const double sgm = mz_vhs_sgw(A.factor, A.expo, g, 1.0, 1.0);   // sigma g
        sgmax = sgm > sgmax ? sgm : sgmax;
        const double wi = pi.w[pi.i], wk = pk.w[pk.i];
        if (!(mz_rand(A.rp, st) < sgm * (wi > wk ? wi : wk) * inv_w_max / sgmax)) continue;
        ++performed;
        double vc0, vc1, vc2;
        mz_com(A.mu1, A.mu2, vi0, vi1, vi2, vk0, vk1, vk2, vc0, vc1, vc2);
        const double dw = (wi < wk ? wi : wk) * wtf;
        pi.w[pi.i] = wi - dw;
        pk.w[pk.i] = wk - dw;
        if (created + 1 >= 2 * n_coll) {
            atomicOr(&A.a.ctr->err, MZ_ERR_SCRATCH);
            continue;
        }
        PRef n_i = pref_scratch(A, 0, s_base + created), n_k = pref_scratch(A, 0, s_base + created + 1);
        n_i.x0[n_i.i] = pi.x0[pi.i]; n_i.x1[n_i.i] = pi.x1[pi.i]; n_i.x2[n_i.i] = pi.x2[pi.i];
        n_k.x0[n_k.i] = pk.x0[pk.i]; n_k.x1[n_k.i] = pk.x1[pk.i]; n_k.x2[n_k.i] = pk.x2[pk.i];
        n_i.w[n_i.i] = dw;
        n_k.w[n_k.i] = dw;
        created += 2;
        na += 2;
        // scatter_vhs! on the two new particles: phi first, then cos(theta)
        double ui0, ui1, ui2, uk0, uk1, uk2;
        mz_vhs_scatter(A.rp, st, A.mu1, A.mu2, g, vc0, vc1, vc2, ui0, ui1, ui2, uk0, uk1, uk2);
        n_i.v0[n_i.i] = ui0; n_i.v1[n_i.i] = ui1; n_i.v2[n_i.i] = ui2;
        n_k.v0[n_k.i] = uk0; n_k.v1[n_k.i] = uk1; n_k.v2[n_k.i] = uk2;
    }
    A.sgwm[c] = sgmax;
    int32_t *cn = A.cnt + (int64_t)c * 5;
    cn[0] = nl;
    cn[1] = nl;
    cn[2] = (int32_t)n_coll;
    cn[3] = performed;
    cn[4] = 0;
    A.created[c] = created;
}

extern "C" int mz_swpm(mz_store *h, int32_t species, double G, double dt, const double *V, uint64_t tag,
                       int32_t rng_mode, int32_t cell_lo, int32_t cell_hi) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "swpm: bad species");
    MZ_CHECK(cell_lo >= 1 && cell_hi <= h->n_cells && cell_lo <= cell_hi, "swpm: bad cell range");
    MZ_CHECK(G >= 0.0, "swpm: G must be non-negative");
    MZ_CHECK(h->inter[(species - 1) * h->n_species + (species - 1)].set, "swpm: interaction not set");
    MZ_CHECK(rng_mode == 0 || (rng_mode == 1 && h->d_replay[0]), "swpm: replay table not set");
    const double *dV = h->d_V;
    if (V) {
        double *v = (double *)mz_scratch(h, "ntc_V", h->n_cells * sizeof(double));
        MZ_CHECK(v != nullptr, "swpm: scratch allocation failed");
        MZ_HIP(hipMemcpy(v, V, h->n_cells * sizeof(double), hipMemcpyHostToDevice));
        dV = v;
    } else {
        MZ_CHECK(h->has_grid, "swpm: V is NULL and no grid is set");
    }
    NtcArgs A = base_args(h, species, species, dt, dV, 0.0, 0, tag, rng_mode, cell_lo, cell_hi);
    A.may_split = 1;
    const int nc = h->n_cells;
    const int ncells = cell_hi - cell_lo + 1;
    int32_t *ncoll = (int32_t *)mz_scratch(h, "ntc_ncoll", (size_t)nc * sizeof(int32_t));
    int32_t *ssz = (int32_t *)mz_scratch(h, "swpm_ssz", (size_t)nc * sizeof(int32_t));
    double *wmax = (double *)mz_scratch(h, "swpm_wmax", (size_t)nc * sizeof(double));
    int64_t *soff = (int64_t *)mz_scratch(h, "ntc_soff", (size_t)(nc + 1) * sizeof(int64_t));
    int32_t *created = (int32_t *)mz_scratch(h, "ntc_created", (size_t)2 * nc * sizeof(int32_t));
    int64_t *coff = (int64_t *)mz_scratch(h, "ntc_coff", (size_t)2 * (nc + 1) * sizeof(int64_t));
    MZ_CHECK(ncoll && ssz && wmax && soff && created && coff, "swpm: scratch allocation failed");
    MZ_HIP(hipMemsetAsync(ncoll, 0, (size_t)nc * sizeof(int32_t), h->stream));
    MZ_HIP(hipMemsetAsync(ssz, 0, (size_t)nc * sizeof(int32_t), h->stream));
    MZ_HIP(hipMemsetAsync(created, 0, (size_t)2 * nc * sizeof(int32_t), h->stream));
    {
        MzTimed tm(h, "swpm");
        hipLaunchKernelGGL(k_swpm_count, dim3((ncells + 3) / 4), dim3(256), 0, h->stream, A, G, ncoll, ssz, wmax);
        hipLaunchKernelGGL(k_scan_i32, dim3(1), dim3(1024), 0, h->stream, ssz, (int64_t)nc, soff, soff + nc);
        int64_t total = 0;
        MZ_HIP(hipMemcpyAsync(&total, soff + nc, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
        MZ_HIP(hipStreamSynchronize(h->stream));
        const int64_t cap = std::max<int64_t>(total, 1);
        double *sc = (double *)mz_scratch(h, "ntc_sc", (size_t)7 * cap * sizeof(double));
        MZ_CHECK(sc != nullptr, "swpm: creation scratch allocation failed");
        for (int f = 0; f < 7; ++f) A.sc[f] = sc + (size_t)f * cap;
        A.sc_cap = cap;
        A.ncoll_in = ncoll;
        A.soff = soff;
        A.created = created;
        hipLaunchKernelGGL(k_swpm, dim3((ncells + 63) / 64), dim3(64), 0, h->stream, A, G, (const double *)wmax);
        MZ_HIP(hipGetLastError());
        double **scp = (double **)mz_scratch(h, "ntc_scp", 7 * sizeof(double *));
        MZ_CHECK(scp != nullptr, "swpm: scratch allocation failed");
        MZ_HIP(hipMemcpy(scp, A.sc, 7 * sizeof(double *), hipMemcpyHostToDevice));
        MzCounters snap;
        int r = mz_read_counters(h, species, &snap);
        if (r) return r;
        hipLaunchKernelGGL(k_scan_i32, dim3(1), dim3(1024), 0, h->stream, created, (int64_t)(A.cell_hi + 1), coff,
                           coff + nc);
        hipLaunchKernelGGL(k_ntc_append, dim3(ncells), dim3(64), 0, h->stream, mz_view(h->sp[species - 1]), A.cell_lo,
                           A.cell_hi, created, coff, coff + nc, soff, (const double *const *)scp, (int64_t)0, snap);
        hipLaunchKernelGGL(k_ntc_finish, dim3(1), dim3(256), 0, h->stream, mz_view(h->sp[species - 1]), coff + nc, snap);
    }
    MZ_HIP(hipGetLastError());
    return 0;
}
