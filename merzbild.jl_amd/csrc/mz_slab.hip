// mz_slab.hip -- the 1-D Couette step with the particles resident in LDS.
//
// The reference's step (test/test_1D_couette.jl:66-97; simulations/1D/
// couette_multithreaded.jl:97-161 runs it per chunk of cells, exchanging
// particles between chunks after the sort) is
//   ntc! in every cell -> convect_particles! -> sort_particles! ->
//   compute_props_sorted!.
// Here one workgroup owns a slab of consecutive cells and keeps that slab's
// particles in LDS for all steps of a launch; only particles that leave the
// slab travel, through a mailbox in HBM, to the neighbouring slab.  The grid
// is no larger than the occupancy query allows, so on an otherwise idle
// device every slab is resident, and a slab waits only for its two
// neighbours (a flag per slab side, bounded spins, no grid barrier).  It is
// an ordinary launch: every wait ends by itself (a spin limit, then the
// abort path below), so nothing depends on cooperative-launch machinery.
//
// Bit-exact order.  After a sort, the particles of a slab occupy one
// contiguous range of logical positions in cell order.  The stable counting
// sort (grid_sorting.jl:58-113) orders the particles of a new cell by their
// previous position; a particle that moved at most into an adjacent slab
// therefore lands after the arrivals from the left slab and before the
// arrivals from the right slab, in the order it had.  Each slab rebuilds its
// cells as [left arrivals | own particles | right arrivals], each part in
// previous order, which is exactly the global stable sort restricted to the
// slab.  A move across more than one slab, a mailbox or LDS overflow, or a
// wait that exceeds its spin limit aborts the launch; the host then restores
// the state it saved and runs the same steps with the step-by-step kernels,
// which give the same results (every random stream is keyed by cell or
// particle id and step, not by launch layout).
//
// The physics is the same inline code the step-by-step kernels use
// (mz_phys.h), and props use the same lane-strided sums as k_props, so the
// two paths agree bit for bit.
#include "mz_phys.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#define SLAB_THREADS 256
#define SLAB_MAXC 32                  // cells per slab
#define SLAB_KMAX (SLAB_MAXC + 2)     // keys: local cells, left, right
#define SLAB_MB_CAP 256               // particles per mailbox (slab side, step parity)
#define SLAB_SPIN_LIMIT (1u << 22)    // polls per wait before the launch aborts
#define SLAB_LINE 16                  // 64-bit words per flag (one 128-byte line)
#define SLAB_DRAWS 64                 // NTC draws per cell computed ahead, in parallel
#define SLAB_WALL_CAP 256             // wall-cell particles whose reflection is drawn ahead
static_assert(SLAB_WALL_CAP <= SLAB_MB_CAP, "the wall hit list borrows an arrival-rank row");

struct SlabArgs {
    MzSpView a;                 // input particles, PIA, counters
    double *bx0, *bx1, *bx2, *bv0, *bv1, *bv2, *bw;   // output particles
    int32_t *bid;
    MzGrid g;
    MzWalls wl;
    double dt, mu1, mu2, factor, expo, mass;
    const double *V;            // cell volumes
    double *sgwm;               // (sigma g w)_max per cell
    int32_t *cnt;               // collision counters [cell][5]
    MzRngParams rp;
    uint64_t step0;
    int32_t n_steps, props_every;
    double *out_np, *out_n, *out_v, *out_T;
    int32_t n_cells, spc, n_slabs, spx;
    uint64_t *mb;               // mailboxes [slab][side][parity][field][SLAB_MB_CAP]
    uint64_t *flag;             // [slab][side] x SLAB_LINE: [parity] = (step + 1) << 32 | count
    uint64_t *fin;              // [slab] x SLAB_LINE: [0] = 1 when done, [1] = counts
    int32_t *abort_flag;
    // averaging phase (avg_inv > 0): wall and cell properties of every step
    // added to their time averages with weight avg_inv
    double avg_inv;
    double surf_f[2];           // mass / dt / area of the two walls
    double *surf_part;          // [slab][2][2 MZ_SURF_Q]: the slab's average increments, its last step's values
    double *pavg;               // time-averaged props, store layout
};

// Wall contacts of one step in the averaging phase: each wave appends its
// lanes' contacts (wall, weight, incident and reflected velocity) to its own
// list in lane order, and one wave sums the lists in wave order during the
// next step's collisions, so the sums are the same on every run.
#define SLAB_SURF_CAP 16   // contacts per wave and step; more abort the launch

struct SlabSurfRec {
    double (*rec)[8];          // this wave's list
    volatile int32_t *cnt;     // this wave's count
    int32_t *abort;
    double w, u0, u1, u2;      // weight, incident velocity
    __device__ void hit(int bc, double v0, double v1, double v2, double sgn) {
        if (sgn > 0.0) {
            u0 = v0; u1 = v1; u2 = v2;
            return;
        }
        const uint64_t m = __ballot(1);
        const int32_t base = *cnt;
        const int slot = base + __popcll(m & mz_lanemask_lt());
        if (slot < SLAB_SURF_CAP) {
            double *e = rec[slot];
            e[0] = (double)bc; e[1] = w; e[2] = u0; e[3] = u1; e[4] = u2; e[5] = v0; e[6] = v1; e[7] = v2;
        } else {
            *abort = 1;
        }
        if ((int)(threadIdx.x & 63) == __ffsll((unsigned long long)m) - 1) *cnt = base + __popcll(m);
    }
};

__device__ inline uint64_t *slab_mb(const SlabArgs &A, int slab, int side, int parity, int field) {
    return A.mb + ((((size_t)slab * 2 + side) * 2 + parity) * 8 + field) * SLAB_MB_CAP;
}

__device__ inline void st_agent(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t ld_agent(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every word that one slab writes and another reads (mailboxes, counts,
// flags) is accessed with agent-scope relaxed atomics, which the hardware
// keeps coherent across the XCDs' L2 caches by themselves.  Publishing
// therefore needs only the order "payload stores complete, then flag store"
// and reading only "flag load, then payload loads"; no agent-scope fence
// (which writes back or invalidates a whole L2) is issued per step.
//
// Wait until this wave's outstanding memory operations have completed
// (s_waitcnt vmcnt(0); expcnt and lgkmcnt left at their no-wait maxima).
__device__ inline void slab_wait_mem() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Poll *p until (*p >> shift) >= want; false on abort or after the spin limit.
__device__ inline bool slab_wait(const uint64_t *p, uint64_t want, int shift, const int32_t *abort_flag,
                                 uint64_t &val) {
    for (uint32_t it = 0; it < SLAB_SPIN_LIMIT; ++it) {
        uint64_t v = ld_agent(p);
        if ((v >> shift) >= want) {
            val = v;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            return true;
        }
        if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    return false;
}

__device__ inline void slab_raise_abort(const SlabArgs &A) {
    __hip_atomic_store(A.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Draw k of the cell's stream: from the block of draws the workgroup
// computed ahead, then from the stream itself.
__device__ inline double slab_draw(const MzRngParams &rp, MzStream &st, const double *pre) {
    if (st.k < SLAB_DRAWS) return pre[st.k++];
    return mz_rand(rp, st);
}

// ntc_equal_weight! of one cell over its LDS segment [base, base + nla);
// pre holds the first SLAB_DRAWS draws of the cell's stream for this step.
__device__ inline void slab_ntc_cell(const SlabArgs &A, const MzRngParams &rp, int c, double Vc, int32_t base,
                                     int32_t nla, double *sv0, double *sv1, double *sv2, const double *sw,
                                     const double *pre, double &sgwm, int32_t &n_coll_out, int32_t &performed,
                                     int32_t &eqw) {
    MzStream st = mz_stream((uint64_t)c);
    double R = slab_draw(rp, st, pre);
    int64_t n_coll = (int64_t)floor(mz_ncoll_est(true, A.dt, nla, nla, sgwm, Vc, R));
    if (!(n_coll <= MZ_NTC_MAX_CANDIDATES)) {
        atomicOr(&A.a.ctr->err, MZ_ERR_SCRATCH);
        n_coll = 0;
    }
    performed = 0;
    eqw = 0;
    for (int64_t cc = 0; cc < n_coll; ++cc) {
        int32_t i = (int32_t)floor(slab_draw(rp, st, pre) * (double)nla);
        int32_t k = (int32_t)floor(slab_draw(rp, st, pre) * (double)nla);
        while (i == k) k = (int32_t)floor(slab_draw(rp, st, pre) * (double)nla);
        const int pi = base + i, pk = base + k;
        double vi0 = sv0[pi], vi1 = sv1[pi], vi2 = sv2[pi];
        double vk0 = sv0[pk], vk1 = sv1[pk], vk2 = sv2[pk];
        double g0, g1, g2;
        double g = mz_rel_speed(vi0, vi1, vi2, vk0, vk1, vk2, g0, g1, g2);
        if (!(g > MZ_G_EPS)) continue;
        double sgw = mz_vhs_sgw(A.factor, A.expo, g, sw[pi], sw[pk]);
        sgwm = sgw > sgwm ? sgw : sgwm;
        if (!(slab_draw(rp, st, pre) < sgw / sgwm)) continue;
        ++performed;
        ++eqw;
        double vc0, vc1, vc2;
        mz_com(A.mu1, A.mu2, vi0, vi1, vi2, vk0, vk1, vk2, vc0, vc1, vc2);
        double ui0, ui1, ui2, uk0, uk1, uk2;
        const double R_phi = slab_draw(rp, st, pre);
        const double R_ct = slab_draw(rp, st, pre);
        mz_vhs_post(A.mu1, A.mu2, g, vc0, vc1, vc2, R_phi, R_ct, ui0, ui1, ui2, uk0, uk1, uk2);
        sv0[pi] = ui0; sv1[pi] = ui1; sv2[pi] = ui2;
        sv0[pk] = uk0; sv1[pk] = uk1; sv2[pk] = uk2;
    }
    n_coll_out = (int32_t)n_coll;
}

template <int R, bool AVG>
__global__ void __launch_bounds__(SLAB_THREADS, R <= 2 ? 2 : 1) k_slab_couette(SlabArgs A) {
    constexpr int P = R * SLAB_THREADS;   // LDS particle capacity
    constexpr int NCH = P / 64;           // 64-particle chunks
    // With R <= 4 the velocities the sort left are also kept in a snapshot,
    // so a step's props run next to the following NTC; the largest slabs have
    // no LDS room for it and compute props right after the scatter.
    constexpr bool SNAP = R <= 4;
    constexpr int PS = SNAP ? P : 1;
    static_assert(NCH <= 64, "chunk scan uses one wave");
    __shared__ double sx0[P], sx1[P], sx2[P], sv0[P], sv1[P], sv2[P], sw[P];
    __shared__ int32_t sid[P];
    __shared__ int32_t tab[NCH][SLAB_KMAX];   // per chunk and key: count, then exclusive prefix
    __shared__ int32_t tot[SLAB_KMAX];
    __shared__ int32_t cst[SLAB_MAXC], ccn[SLAB_MAXC], runL[SLAB_MAXC], runR[SLAB_MAXC];
    __shared__ int32_t arank[2][SLAB_MB_CAP], akey[2][SLAB_MB_CAP];
    __shared__ double sdraw[SLAB_MAXC][SLAB_DRAWS];
    __shared__ double ss0[PS], ss1[PS], ss2[PS];   // velocities as the last sort left them (props input)
    constexpr int WC = SNAP ? SLAB_WALL_CAP : 1;
    __shared__ double swl[3][WC];                  // reflected velocities drawn ahead (wall slab)
    __shared__ int32_t swf[WC];                    // 1 where swl holds the entry
    // the flight with that one bounce worked out ahead: the velocity it
    // assumed and the end position (NaN where the flight bounces again)
    constexpr int WX = (SNAP && !AVG) ? SLAB_WALL_CAP : 1;
    __shared__ double swv[WX], swx[WX];
    __shared__ int32_t s_abort, s_n, s_na[2];
    __shared__ int64_t s_red[4][4];
    constexpr int SC = AVG ? SLAB_SURF_CAP : 1;
    __shared__ double srec[AVG ? 4 : 1][SC][8];   // wall contacts of the last convection, per wave
    __shared__ int32_t scnt[4];
    __shared__ double pav[AVG ? SLAB_MAXC : 1][6];  // the slab's cells' time-averaged props

    const int b = blockIdx.x;
    const int g = (b & 7) * A.spx + (b >> 3);   // consecutive slabs on one XCD
    if (g >= A.n_slabs) return;
    // the check before the launch found a state this kernel cannot take, or
    // another slab has already aborted: touch nothing
    if (__hip_atomic_load(A.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int clo = g * A.spc;
    const int chi = min(A.n_cells, clo + A.spc) - 1;
    const int ncl = chi - clo + 1;
    const int K = ncl + 2, KL = ncl, KR = ncl + 1;
    const bool has_left = g > 0, has_right = g + 1 < A.n_slabs;
    const int left_lo = clo - A.spc;
    const int right_hi = min(A.n_cells, chi + 1 + A.spc) - 1;
    const uint64_t lt = mz_lanemask_lt();

    // ---- load the slab: cells clo..chi are one contiguous range of slots
    int32_t first = 0;
    if (wave == 0) {
        int32_t n1 = lane < ncl ? A.a.pia.n1[clo + lane] : 0;
        int32_t s1 = lane < ncl ? A.a.pia.start1[clo + lane] : 0;
        int32_t x = n1;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane < ncl) { cst[lane] = x - n1; ccn[lane] = n1; }
        uint64_t ne = __ballot(n1 > 0);
        int fl = ne ? __ffsll((unsigned long long)ne) - 1 : 0;
        int32_t s_first = __shfl(s1, fl, 64);
        if (lane == 63) { s_n = x; s_abort = x > P; }
        if (lane == 0) s_red[0][0] = ne ? s_first - 1 : 0;
    }
    __syncthreads();
    first = (int32_t)s_red[0][0];
    int32_t n_loc = s_n;
    if (s_abort) {
        if (tid == 0) slab_raise_abort(A);
        return;
    }
    for (int i = tid; i < n_loc; i += SLAB_THREADS) {
        const int64_t p = (int64_t)first + i;
        sx0[i] = A.a.x0[p]; sx1[i] = A.a.x1[p]; sx2[i] = A.a.x2[p];
        sv0[i] = A.a.v0[p]; sv1[i] = A.a.v1[p]; sv2[i] = A.a.v2[p];
        sw[i] = A.a.w[p];
        sid[i] = A.a.id[p];
        if (SNAP) { ss0[i] = sv0[i]; ss1[i] = sv1[i]; ss2[i] = sv2[i]; }
    }
    // NTC lanes: cell k runs on lane k / 4 of wave k % 4
    const int my_k = lane * 4 + wave;
    const bool ntc_lane = lane < SLAB_MAXC / 4 && my_k < ncl;
    double my_sgwm = ntc_lane ? A.sgwm[clo + my_k] : 0.0;
    const double my_V = ntc_lane ? A.V[clo + my_k] : 1.0;
    int32_t my_nla = 0, my_ncoll = 0, my_perf = 0, my_eqw = 0;

    double px0[R], px1[R], px2[R], pv0[R], pv1[R], pv2[R], pw[R];
    int32_t pid[R], pkey[R], prank[R], pcell[R];
    int32_t n_prev = n_loc;
    bool aborted = false;

    // The first SLAB_DRAWS draws of every cell's NTC stream for step stp,
    // computed by threads t0, t0 + stride, ...
    auto fill_draws = [&](uint64_t stp, int t0, int stride) {
        MzRngParams rpf = A.rp;
        rpf.tag = (stp << 4) | 1u;
        for (int j = t0; j < ncl * (SLAB_DRAWS / 2); j += stride) {
            const int k = j / (SLAB_DRAWS / 2), blk = j % (SLAB_DRAWS / 2);
            mz_philox_pair(rpf, (uint64_t)(clo + k), (uint64_t)blk, sdraw[k][2 * blk], sdraw[k][2 * blk + 1]);
        }
    };
    // compute_props_sorted! of the slab's cells from the snapshot the last
    // scatter wrote: cell k on wave (k + 1) % 4, the sums of k_props.  It runs
    // while the next step's NTC, on wave k % 4, changes the live velocities.
    auto props_cells = [&]() {
        for (int k = 0; k < ncl; ++k) {
            if (((k + 1) & 3) != wave) continue;
            const int32_t st1 = cst[k], tn = ccn[k];
            auto get = [&](int32_t i, double &w, double &v0, double &v1, double &v2) {
                const int p = st1 + i;
                if (SNAP) { w = sw[p]; v0 = ss0[p]; v1 = ss1[p]; v2 = ss2[p]; }
                else { w = sw[p]; v0 = sv0[p]; v1 = sv1[p]; v2 = sv2[p]; }
            };
            const MzPropsOut o = mz_props_wave(get, tn, A.mass);
            if (lane == 0) {
                const int c = clo + k;
                A.out_np[c] = (double)tn;
                A.out_n[c] = o.n;
                A.out_v[3 * c] = o.v0;
                A.out_v[3 * c + 1] = o.v1;
                A.out_v[3 * c + 2] = o.v2;
                A.out_T[c] = o.T;
                if (AVG) {
                    const double iv = A.avg_inv;
                    pav[k][0] = fma((double)tn, iv, pav[k][0]);
                    pav[k][1] = fma(o.n, iv, pav[k][1]);
                    pav[k][2] = fma(o.v0, iv, pav[k][2]);
                    pav[k][3] = fma(o.v1, iv, pav[k][3]);
                    pav[k][4] = fma(o.v2, iv, pav[k][4]);
                    pav[k][5] = fma(o.T, iv, pav[k][5]);
                }
            }
        }
    };
    auto props_due = [&](int32_t t) { return A.props_every > 0 && ((t + 1) % A.props_every) == 0; };
    // A slab that ends at a fully diffuse wall (one wall only) draws, while
    // NTC runs, the velocities that a first reflection gives the wall cell's
    // particles headed for that wall.  The draws depend on particle id and
    // step only, so they are the ones convection would draw; which particles
    // get one is predicted from the velocities before NTC (a particle that a
    // collision turns towards the wall draws in convection as before).  The
    // waves that carry no NTC or props cell do it: waves 2-3 of a one-cell
    // slab, wave 3 of a two-cell slab.
    int wall_side = -1;
    if (SNAP && has_left != has_right && ncl <= 2) {
        const int s = has_left ? 1 : 0;
        if (A.wl.acc[s] == 1.0) wall_side = s;
    }
    const int wall_k = wall_side == 1 ? ncl - 1 : 0;
    const int wall_w0 = ncl == 1 ? 2 : 3;   // first free wave
    // Each free wave first marks its share of the wall cell, listing the
    // predicted hits in arank[wave & 1] (the arrival ranks are not in use
    // during NTC), then draws them one per lane, so a wave spends one
    // reflection's latency rather than one per entry it marked.
    auto wall_draws = [&](uint64_t stp) {
        MzRngParams rpw = A.rp;
        rpw.tag = (stp << 4) | 2u;
        const int32_t lo = cst[wall_k], n = min(ccn[wall_k], SLAB_WALL_CAP);
        const int nthr = SLAB_THREADS - wall_w0 * 64;
        int32_t *list = arank[wave & 1];
        int nh = 0;
        for (int j0 = (wave - wall_w0) * 64; j0 < n; j0 += nthr) {
            const int j = j0 + lane;
            bool hit = false;
            if (j < n) {
                const double xn = fma(ss0[lo + j], A.dt, sx0[lo + j]);
                hit = wall_side == 0 ? (xn <= 0.0 && xn < A.g.L) : (xn >= A.g.L);
                swf[j] = hit ? 1 : 0;
            }
            const uint64_t m = __ballot(hit);
            if (hit) list[nh + __popcll(m & lt)] = j;
            nh += __popcll(m);
        }
        const double wn = wall_side == 0 ? 1.0 : -1.0;
        for (int h = lane; h < nh; h += 64) {
            const int j = list[h];
            MzStream sr = mz_stream((uint64_t)(sid[lo + j] - 1));
            double u0, u1, u2;
            mz_diffuse_reflect(rpw, sr, u0, u1, u2, A.wl.vsq[wall_side][0], wn, A.wl.v[wall_side]);
            swl[0][j] = u0; swl[1][j] = u1; swl[2][j] = u2;
            if (SNAP && !AVG) {
                // mz_convect_impl's first bounce at this wall with the velocity
                // before NTC; convection takes the result when NTC left that
                // velocity unchanged (the same operations on the same values)
                const double x0 = sx0[lo + j], v0 = ss0[lo + j];
                const double t_rest = A.dt - (wall_side == 1 ? fabs((A.g.L - x0) / v0) : fabs(x0 / v0));
                const double x_new = fma(u0, t_rest, wall_side == 1 ? A.g.L : 0.0);
                const bool again = (x_new >= A.g.L) || (x_new <= 0.0);
                swv[j] = v0;
                swx[j] = again ? __builtin_nan("") : (x_new < A.g.min_x ? A.g.min_x
                                                      : (x_new > A.g.max_x ? A.g.max_x : x_new));
            }
        }
    };
    if (A.n_steps > 0) fill_draws(A.step0, tid, SLAB_THREADS);
    const int32_t pn = A.n_cells;
    if (AVG) {
        if (tid < 4) scnt[tid] = 0;
        if (tid < ncl * 6) {
            const int k = tid / 6, f = tid % 6, c = clo + k;
            pav[k][f] = f == 0 ? A.pavg[c] : f == 1 ? A.pavg[pn + c]
                      : f < 5 ? A.pavg[2 * pn + 3 * c + (f - 2)] : A.pavg[5 * pn + c];
        }
    }
    // wave 3, lanes 0 .. 2 MZ_SURF_Q - 1: quantity (lane / MZ_SURF_Q, lane %
    // MZ_SURF_Q) of the slab's wall properties, its running time average and
    // its value in the last step
    double surf_acc = 0.0, surf_last = 0.0;
    auto surf_reduce = [&]() {
        if (lane < 2 * MZ_SURF_Q) {
            const int e = lane / MZ_SURF_Q, k = lane % MZ_SURF_Q;
            double sum = 0.0;
            for (int w = 0; w < 4; ++w) {
                const int n = min((int)scnt[w], SLAB_SURF_CAP);
                for (int j = 0; j < n; ++j) {
                    const double *r = srec[w][j];
                    if ((int)r[0] != e) continue;
                    sum += mz_surf_term(k, e, r[1], r[2], r[3], r[4], 1.0);
                    sum += mz_surf_term(k, e, r[1], r[5], r[6], r[7], -1.0);
                }
            }
            const double inst = sum * (k == 0 ? 1.0 : A.surf_f[e]);
            surf_acc = fma(inst, A.avg_inv, surf_acc);
            surf_last = inst;
        }
        if (lane < 4) scnt[lane] = 0;
    };
    __syncthreads();

    for (int32_t t = 0; t < A.n_steps; ++t) {
        const uint64_t step = A.step0 + (uint64_t)t;
        const int q = t & 1;
        n_prev = n_loc;
        // ---- 1. ntc_equal_weight! per cell (its draws were computed during
        // the previous step), next to the previous step's props; clear the
        // ranking tables.
        for (int j = tid; j < NCH * SLAB_KMAX; j += SLAB_THREADS) (&tab[0][0])[j] = 0;
        if (tid < SLAB_MAXC) { runL[tid] = 0; runR[tid] = 0; }
        if (SNAP && t > 0 && props_due(t - 1)) props_cells();
        if (ntc_lane) {
            MzRngParams rp = A.rp;
            rp.tag = (step << 4) | 1u;
            my_nla = ccn[my_k];
            slab_ntc_cell(A, rp, clo + my_k, my_V, cst[my_k], my_nla, sv0, sv1, sv2, sw, sdraw[my_k], my_sgwm,
                          my_ncoll, my_perf, my_eqw);
        }
        if (wall_side >= 0 && wave >= wall_w0) wall_draws(step);
        if (AVG && t > 0 && wave == 3) surf_reduce();
        __syncthreads();
        // ---- 2. convection, new cell, rank inside 64-particle chunks
        {
            MzRngParams rp = A.rp;
            rp.tag = (step << 4) | 2u;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = r * SLAB_THREADS + tid;
                const bool valid = i < n_loc;
                pkey[r] = -1;
                prank[r] = 0;
                if (valid) {
                    px0[r] = sx0[i]; px1[r] = sx1[i]; px2[r] = sx2[i];
                    pv0[r] = sv0[i]; pv1[r] = sv1[i]; pv2[r] = sv2[i];
                    pw[r] = sw[i];
                    pid[r] = sid[i];
                    double x_new = fma(pv0[r], A.dt, px0[r]);
                    if (x_new < A.g.L && x_new > 0.0) {
                        px0[r] = x_new < A.g.min_x ? A.g.min_x : (x_new > A.g.max_x ? A.g.max_x : x_new);
                    } else {
                        double p0 = 0.0, p1 = 0.0, p2 = 0.0;
                        int pre_bc = -1;
                        bool done = false;
                        if (wall_side >= 0) {
                            const int j = i - cst[wall_k];
                            if (j >= 0 && j < min(ccn[wall_k], SLAB_WALL_CAP) && swf[j]) {
                                p0 = swl[0][j]; p1 = swl[1][j]; p2 = swl[2][j];
                                pre_bc = wall_side;
                                if (SNAP && !AVG) {
                                    const double xe = swx[j];
                                    if (swv[j] == pv0[r] && xe == xe) {
                                        px0[r] = xe;
                                        pv0[r] = p0; pv1[r] = p1; pv2[r] = p2;
                                        done = true;
                                    }
                                }
                            }
                        }
                        if (done) {
                        } else if (AVG) {
                            SlabSurfRec sr{srec[wave], &scnt[wave], &s_abort, pw[r], 0.0, 0.0, 0.0};
                            mz_convect_one_surf(sr, rp, A.g, A.wl, 0, A.dt, pid[r], px0[r], pv0[r], pv1[r], pv2[r],
                                                &A.a.ctr->err, pre_bc, p0, p1, p2);
                        } else {
                            mz_convect_one(rp, A.g, A.wl, 0, A.dt, pid[r], px0[r], pv0[r], pv1[r], pv2[r],
                                           &A.a.ctr->err, pre_bc, p0, p1, p2);
                        }
                    }
                    const int c = mz_cell_of(px0[r], A.g, A.n_cells, &A.a.ctr->err) - 1;
                    pcell[r] = c + 1;
                    int key;
                    if (c < clo) {
                        key = KL;
                        if (c < left_lo) s_abort = 1;
                    } else if (c > chi) {
                        key = KR;
                        if (c > right_hi) s_abort = 1;
                    } else {
                        key = c - clo;
                    }
                    pkey[r] = key;
                }
                const int ch = r * 4 + wave;
                uint64_t active = __ballot(valid);
                while (active) {
                    const int leader = __ffsll((unsigned long long)active) - 1;
                    const int lk = __shfl(pkey[r], leader, 64);
                    const uint64_t m = __ballot(valid && pkey[r] == lk);
                    if (valid && pkey[r] == lk) prank[r] = __popcll(m & lt);
                    if (lane == leader) tab[ch][lk] = __popcll(m);
                    active &= ~m;
                }
            }
        }
        __syncthreads();
        // ---- 3. per key, exclusive prefix over chunks.  With at most 8
        // chunks each thread adds up the earlier chunks of its own keys in the
        // next phase and wave 0 forms the totals there, without a phase of
        // its own.
        constexpr bool SMALL_SCAN = NCH <= 8;
        if (!SMALL_SCAN) {
            for (int kk = wave; kk < K; kk += 4) {
                int32_t v = lane < NCH ? tab[lane][kk] : 0;
                int32_t x = v;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    int32_t y = __shfl_up(x, o, 64);
                    if (lane >= o) x += y;
                }
                if (lane < NCH) tab[lane][kk] = x - v;
                if (lane == 63) tot[kk] = x;
            }
            __syncthreads();
        } else if (wave == 0 && lane < K) {
            int32_t t = 0;
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) t += tab[ch][lane];
            tot[lane] = t;
        }
        // ---- 4. ranks in the slab's previous order; leavers go to the
        // mailboxes (a rank past the mailbox capacity aborts below)
        const bool bad0 = !SMALL_SCAN && (s_abort || tot[KL] > SLAB_MB_CAP || tot[KR] > SLAB_MB_CAP);
        if (!bad0) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (pkey[r] < 0) continue;
                if (SMALL_SCAN) {
                    const int ch = r * 4 + wave;
                    int32_t base = 0;
                    for (int c2 = 0; c2 < ch; ++c2) base += tab[c2][pkey[r]];
                    prank[r] += base;
                } else {
                    prank[r] += tab[r * 4 + wave][pkey[r]];
                }
                if (pkey[r] >= KL && prank[r] < SLAB_MB_CAP) {
                    const int side = pkey[r] == KL ? 0 : 1;
                    const int j = prank[r];
                    st_agent(slab_mb(A, g, side, q, 0) + j, (uint64_t)__double_as_longlong(px0[r]));
                    st_agent(slab_mb(A, g, side, q, 1) + j, (uint64_t)__double_as_longlong(px1[r]));
                    st_agent(slab_mb(A, g, side, q, 2) + j, (uint64_t)__double_as_longlong(px2[r]));
                    st_agent(slab_mb(A, g, side, q, 3) + j, (uint64_t)__double_as_longlong(pv0[r]));
                    st_agent(slab_mb(A, g, side, q, 4) + j, (uint64_t)__double_as_longlong(pv1[r]));
                    st_agent(slab_mb(A, g, side, q, 5) + j, (uint64_t)__double_as_longlong(pv2[r]));
                    st_agent(slab_mb(A, g, side, q, 6) + j, (uint64_t)__double_as_longlong(pw[r]));
                    st_agent(slab_mb(A, g, side, q, 7) + j, ((uint64_t)(uint32_t)pid[r] << 32) | (uint32_t)pcell[r]);
                }
            }
        }
        slab_wait_mem();
        __syncthreads();
        const bool bad = bad0 || s_abort || tot[KL] > SLAB_MB_CAP || tot[KR] > SLAB_MB_CAP;
        if (bad) {
            if (tid == 0) slab_raise_abort(A);
            aborted = true;
            break;
        }
        // ---- 5. publish, then wait for the neighbours' mailboxes of this step.
        // Each slab side has one word per step parity holding the step mark
        // and the count, so the poll that sees the mark also delivers the
        // count.  A neighbour may already publish the next step (the other
        // word) before this slab reads its word, but it cannot reach the step
        // after that, which rewrites this word, before this slab has
        // published the next step itself.
        const uint64_t mark = (uint64_t)(t + 1);
        if (tid == 0) {
            // (step + 1) << 32 | count in the word of this step's parity; the
            // payload stores of all waves completed before the barrier above
            if (has_left) st_agent(A.flag + ((size_t)g * 2 + 0) * SLAB_LINE + q, (mark << 32) | (uint32_t)tot[KL]);
            if (has_right) st_agent(A.flag + ((size_t)g * 2 + 1) * SLAB_LINE + q, (mark << 32) | (uint32_t)tot[KR]);
        }
        // ---- 6. as soon as its neighbour's word shows this step, wave 0 (left)
        // or 1 (right) ranks that side's arrivals by cell; the first 64 of
        // each side stay in registers until the scatter
        uint64_t aw[8];
        int a_key = 0, a_rank = 0;
        if (wave < 2) {
            int32_t na = 0;
            if (lane == 0) {
                const bool has = wave == 0 ? has_left : has_right;
                if (has) {
                    const int src = wave == 0 ? g - 1 : g + 1;
                    const int side = wave == 0 ? 1 : 0;
                    const uint64_t *f = A.flag + ((size_t)src * 2 + side) * SLAB_LINE + q;
                    uint64_t v = 0;
                    if (slab_wait(f, mark, 32, A.abort_flag, v)) {
                        na = (int32_t)(uint32_t)(v & 0xFFFFFFFFu);
                        if (na < 0 || na > SLAB_MB_CAP) {
                            s_abort = 1;
                            na = 0;
                        }
                    } else {
                        s_abort = 1;
                    }
                }
                s_na[wave] = na;
            }
            na = __shfl(na, 0, 64);
            const int src = wave == 0 ? g - 1 : g + 1;
            const int side = wave == 0 ? 1 : 0;
            volatile int32_t *run = wave == 0 ? runL : runR;
            for (int j0 = 0; j0 < na; j0 += 64) {
                const int j = j0 + lane;
                const bool valid = j < na;
                int key = 0;
                if (valid) {
                    uint64_t w7;
                    if (j0 == 0) {
#pragma unroll
                        for (int f = 0; f < 8; ++f) aw[f] = ld_agent(slab_mb(A, src, side, q, f) + j);
                        w7 = aw[7];
                    } else {
                        w7 = ld_agent(slab_mb(A, src, side, q, 7) + j);
                    }
                    key = (int32_t)(uint32_t)(w7 & 0xFFFFFFFFu) - 1 - clo;
                    if (key < 0 || key >= ncl) {
                        s_abort = 1;
                        key = key < 0 ? 0 : ncl - 1;
                    }
                }
                uint64_t active = __ballot(valid);
                while (active) {
                    const int leader = __ffsll((unsigned long long)active) - 1;
                    const int lk = __shfl(key, leader, 64);
                    const uint64_t m = __ballot(valid && key == lk);
                    const int32_t bk = run[lk];
                    if (valid && key == lk) {
                        arank[wave][j] = bk + __popcll(m & lt);
                        akey[wave][j] = lk;
                        if (j0 == 0) { a_key = lk; a_rank = bk + __popcll(m & lt); }
                    }
                    if (lane == leader) run[lk] = bk + __popcll(m);
                    active &= ~m;
                }
            }
        }
        if (wave >= 2 && t + 1 < A.n_steps) fill_draws(step + 1, tid - 128, SLAB_THREADS - 128);
        __syncthreads();
        if (s_abort) {
            if (tid == 0) slab_raise_abort(A);
            aborted = true;
            break;
        }
        // ---- 7. new cell starts: [left arrivals | own | right arrivals] per
        // cell.  A slab of at most 4 cells forms them in the scatter below,
        // each thread for its own cells.
        const bool small_bases = ncl <= 4;
        if (small_bases) {
            int32_t x = 0;
            for (int k = 0; k < ncl; ++k) {
                const int32_t cnum = runL[k] + tot[k] + runR[k];
                if (wave == 0 && lane == k) { cst[k] = x; ccn[k] = cnum; }
                x += cnum;
            }
            n_loc = x;
        } else if (wave == 0) {
            int32_t cnum = lane < ncl ? runL[lane] + tot[lane] + runR[lane] : 0;
            int32_t x = cnum;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                int32_t y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            if (lane < ncl) { cst[lane] = x - cnum; ccn[lane] = cnum; }
            if (lane == 63) {
                s_n = x;
                if (x > P) s_abort = 1;
            }
        }
        if (!small_bases) {
            __syncthreads();
            n_loc = s_n;
        }
        if (n_loc > P) s_abort = 1;   // same value on every thread
        if (s_abort) {
            if (tid == 0) slab_raise_abort(A);
            aborted = true;
            break;
        }
        // the start of cell k (its cst entry, formed here when small_bases)
        auto cstart = [&](int k) {
            if (!small_bases) return (int32_t)cst[k];
            int32_t x = 0;
            for (int k2 = 0; k2 < k; ++k2) x += runL[k2] + tot[k2] + runR[k2];
            return x;
        };
        // ---- 8. scatter into the new order
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int key = pkey[r];
            if (key < 0 || key >= ncl) continue;
            const int pos = cstart(key) + runL[key] + prank[r];
            sx0[pos] = px0[r]; sx1[pos] = px1[r]; sx2[pos] = px2[r];
            sv0[pos] = pv0[r]; sv1[pos] = pv1[r]; sv2[pos] = pv2[r];
            if (SNAP) { ss0[pos] = pv0[r]; ss1[pos] = pv1[r]; ss2[pos] = pv2[r]; }
            sw[pos] = pw[r];
            sid[pos] = pid[r];
        }
        if (wave < 2 && lane < s_na[wave]) {
            const int pos = cstart(a_key) + (wave == 0 ? 0 : runL[a_key] + tot[a_key]) + a_rank;
            sx0[pos] = __longlong_as_double((long long)aw[0]);
            sx1[pos] = __longlong_as_double((long long)aw[1]);
            sx2[pos] = __longlong_as_double((long long)aw[2]);
            sv0[pos] = __longlong_as_double((long long)aw[3]);
            sv1[pos] = __longlong_as_double((long long)aw[4]);
            sv2[pos] = __longlong_as_double((long long)aw[5]);
            if (SNAP) { ss0[pos] = sv0[pos]; ss1[pos] = sv1[pos]; ss2[pos] = sv2[pos]; }
            sw[pos] = __longlong_as_double((long long)aw[6]);
            sid[pos] = (int32_t)(aw[7] >> 32);
        }
        {
            // arrivals beyond the first 64 of a side are read again
            const int xl = max(s_na[0] - 64, 0), xr = max(s_na[1] - 64, 0);
            for (int j = tid; j < xl + xr; j += SLAB_THREADS) {
                const int s = j < xl ? 0 : 1;
                const int jj = 64 + (s == 0 ? j : j - xl);
                const int key = akey[s][jj];
                const int pos = cstart(key) + (s == 0 ? 0 : runL[key] + tot[key]) + arank[s][jj];
                const int src = s == 0 ? g - 1 : g + 1;
                const int side = s == 0 ? 1 : 0;
                sx0[pos] = __longlong_as_double((long long)ld_agent(slab_mb(A, src, side, q, 0) + jj));
                sx1[pos] = __longlong_as_double((long long)ld_agent(slab_mb(A, src, side, q, 1) + jj));
                sx2[pos] = __longlong_as_double((long long)ld_agent(slab_mb(A, src, side, q, 2) + jj));
                sv0[pos] = __longlong_as_double((long long)ld_agent(slab_mb(A, src, side, q, 3) + jj));
                sv1[pos] = __longlong_as_double((long long)ld_agent(slab_mb(A, src, side, q, 4) + jj));
                sv2[pos] = __longlong_as_double((long long)ld_agent(slab_mb(A, src, side, q, 5) + jj));
                if (SNAP) { ss0[pos] = sv0[pos]; ss1[pos] = sv1[pos]; ss2[pos] = sv2[pos]; }
                sw[pos] = __longlong_as_double((long long)ld_agent(slab_mb(A, src, side, q, 6) + jj));
                sid[pos] = (int32_t)(ld_agent(slab_mb(A, src, side, q, 7) + jj) >> 32);
            }
        }
        __syncthreads();
        if (!SNAP && props_due(t)) {
            props_cells();
            __syncthreads();
        }
    }
    if (aborted) return;
    if (SNAP && A.n_steps > 0 && props_due(A.n_steps - 1)) props_cells();
    if (AVG && A.n_steps > 0 && wave == 3) surf_reduce();

    // ---- publish the slab's counts, then place it after all earlier slabs
    if (tid == 0) {
        st_agent(A.fin + (size_t)g * SLAB_LINE + 1, ((uint64_t)(uint32_t)n_prev << 32) | (uint32_t)n_loc);
        slab_wait_mem();
        st_agent(A.fin + (size_t)g * SLAB_LINE, (uint64_t)1);
    }
    int64_t pre_prev = 0, pre_new = 0, all_new = 0;
    bool ok = true;
    for (int s = tid; s < A.n_slabs; s += SLAB_THREADS) {
        uint64_t v = 0;
        if (!slab_wait(A.fin + (size_t)s * SLAB_LINE, 1, 0, A.abort_flag, v)) {
            ok = false;
            break;
        }
        const uint64_t cn = ld_agent(A.fin + (size_t)s * SLAB_LINE + 1);
        const int64_t cp = (int64_t)(cn >> 32), cnew = (int64_t)(cn & 0xFFFFFFFFu);
        if (s < g) { pre_prev += cp; pre_new += cnew; }
        all_new += cnew;
    }
    if (!ok) s_abort = 1;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        pre_prev += __shfl_xor(pre_prev, o, 64);
        pre_new += __shfl_xor(pre_new, o, 64);
        all_new += __shfl_xor(all_new, o, 64);
    }
    if (lane == 0) { s_red[wave][0] = pre_prev; s_red[wave][1] = pre_new; s_red[wave][2] = all_new; }
    __syncthreads();
    if (s_abort) {
        if (tid == 0) slab_raise_abort(A);
        return;
    }
    const int64_t off_prev = s_red[0][0] + s_red[1][0] + s_red[2][0] + s_red[3][0];
    const int64_t off_new = s_red[0][1] + s_red[1][1] + s_red[2][1] + s_red[3][1];
    const int64_t total = s_red[0][2] + s_red[1][2] + s_red[2][2] + s_red[3][2];
    if (off_new + n_loc > A.a.cap || off_prev + n_prev > A.a.cap) {
        // counts that do not add up to the store: flag it, write nothing
        if (tid == 0) atomicOr(&A.a.ctr->err, MZ_ERR_CAPACITY);
        return;
    }
    for (int i = tid; i < n_loc; i += SLAB_THREADS) {
        const int64_t p = off_new + i;
        A.bx0[p] = sx0[i]; A.bx1[p] = sx1[i]; A.bx2[p] = sx2[i];
        A.bv0[p] = sv0[i]; A.bv1[p] = sv1[i]; A.bv2[p] = sv2[i];
        A.bw[p] = sw[i];
        A.bid[p] = sid[i];
    }
    // pv.cell keeps the cells of the last sort's input order
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = r * SLAB_THREADS + tid;
        if (i < n_prev) A.a.key[off_prev + i] = pcell[r];
    }
    if (tid < ncl) {
        const int c = clo + tid;
        const int32_t cnum = ccn[tid];
        const int32_t s1 = (int32_t)off_new + cst[tid];
        A.a.pia.start1[c] = cnum > 0 ? s1 + 1 : 0;
        A.a.pia.end1[c] = cnum > 0 ? s1 + cnum : -1;
        A.a.pia.n1[c] = cnum;
        A.a.pia.nloc[c] = cnum;
        A.a.pia.start2[c] = 0;
        A.a.pia.end2[c] = -1;
        A.a.pia.n2[c] = 0;
    }
    if (ntc_lane) {
        const int c = clo + my_k;
        A.sgwm[c] = my_sgwm;
        int32_t *cn = A.cnt + (int64_t)c * 5;
        cn[0] = my_nla;
        cn[1] = my_nla;
        cn[2] = my_ncoll;
        cn[3] = my_perf;
        cn[4] = my_eqw;
    }
    if (g == 0 && tid == 0) {
        if (total != (int64_t)A.a.ctr->n_total) atomicOr(&A.a.ctr->err, MZ_ERR_CAPACITY);
        A.a.ctr->contiguous = 1;
        A.a.ctr->n_pos = (int32_t)total;
    }
    if (AVG) {
        // the averages: cells from LDS (each written by its props lane, the
        // barrier above orders it), the wall increments per slab, lpa from
        // the vector's length, unchanged by the slab steps
        if (tid < ncl * 6) {
            const int k = tid / 6, f = tid % 6, c = clo + k;
            double *d = f == 0 ? A.pavg + c : f == 1 ? A.pavg + pn + c
                      : f < 5 ? A.pavg + 2 * pn + 3 * c + (f - 2) : A.pavg + 5 * pn + c;
            *d = pav[k][f];
        }
        if (wave == 3 && lane < 2 * MZ_SURF_Q) {
            A.surf_part[((size_t)g * 2) * 2 * MZ_SURF_Q + lane] = surf_acc;
            A.surf_part[((size_t)g * 2 + 1) * 2 * MZ_SURF_Q + lane] = surf_last;
        }
        if (g == 0 && tid == 0) {
            const double len = (double)A.a.ctr->length;
            double l = A.pavg[6 * pn];
            for (int32_t t = 0; t < A.n_steps; ++t) l = fma(len, A.avg_inv, l);
            A.pavg[6 * pn] = l;
        }
    }
}

// After an averaging-phase launch: quantity blockIdx.x of the slabs' wall
// increments added to the time average, and of their last step's values
// summed into the instantaneous set; lanes stride over the slabs, then a
// wave sum, so the order is fixed.
__global__ void __launch_bounds__(64) k_slab_surf_finalize(const double *__restrict__ part, int32_t n_slabs,
                                                           double *__restrict__ surf) {
    const int i = blockIdx.x, lane = threadIdx.x;
    double acc = 0.0, last = 0.0;
    for (int32_t g = lane; g < n_slabs; g += 64) {
        acc += part[((size_t)g * 2) * 2 * MZ_SURF_Q + i];
        last += part[((size_t)g * 2 + 1) * 2 * MZ_SURF_Q + i];
    }
    acc = mz_wave_sum(acc);
    last = mz_wave_sum(last);
    if (lane == 0) {
        surf[2 * MZ_SURF_Q + i] += acc;   // set 1 (one species)
        surf[i] = last;                   // set 0
    }
}

// Before a slab launch, on the store's stream: clears the flag words, keeps a
// copy of what the launch overwrites before it can abort (collision factors,
// PIA, collision counters, counters), and checks the launch's precondition,
// the state a sort leaves: group 1 only, every cell's range starting where
// the previous non-empty cell's ended, the ranges covering n_total, the
// store contiguous and error-free.  A state that fails it sets the abort
// word to SLAB_ABORT_STATE, and every slab then returns before it reads.
#define SLAB_ABORT_STATE 2

struct SlabPrepArgs {
    MzPia pia;
    const MzCounters *ctr;
    const double *sgwm;
    const int32_t *cnt;
    int32_t n_cells;
    char *bak;              // [nc] double | [7][nc] int32 PIA | [nc][5] int32 | MzCounters
    uint64_t *flags;        // slab flags and fin lines, cleared
    int64_t n_flag_words;
    uint64_t *abort_line;   // SLAB_LINE words; the first 32 bits are the abort word
};

__global__ void __launch_bounds__(1024) k_slab_prep(SlabPrepArgs P) {
    __shared__ int32_t wsum[16];
    __shared__ int32_t carry, bad;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int nc = P.n_cells;
    for (int64_t i = t; i < P.n_flag_words; i += 1024) P.flags[i] = 0;
    double *__restrict__ bd = (double *)P.bak;
    int32_t *__restrict__ bi = (int32_t *)(P.bak + (size_t)nc * sizeof(double));
    int32_t *__restrict__ bc = bi + (size_t)7 * nc;
    int32_t *__restrict__ bctr = bc + (size_t)5 * nc;
    const int32_t *ctr = (const int32_t *)P.ctr;
    if (t < (int)(sizeof(MzCounters) / sizeof(int32_t))) bctr[t] = ctr[t];
    if (t == 0) {
        carry = 0;
        bad = (!P.ctr->contiguous || P.ctr->err) ? 1 : 0;
    }
    __syncthreads();
    // one pass per 1024 cells: each thread loads all of its cell's words
    // before it stores any, then the pass scans n_group1
    for (int c0 = 0; c0 < nc; c0 += 1024) {
        const int c = c0 + t;
        const bool in = c < nc;
        int32_t pv[7] = {0, 0, 0, 0, 0, 0, 0}, cv[5] = {0, 0, 0, 0, 0};
        double sg = 0.0;
        if (in) {
            sg = P.sgwm[c];
            pv[0] = P.pia.start1[c]; pv[1] = P.pia.end1[c]; pv[2] = P.pia.n1[c]; pv[3] = P.pia.start2[c];
            pv[4] = P.pia.end2[c]; pv[5] = P.pia.n2[c]; pv[6] = P.pia.nloc[c];
#pragma unroll
            for (int k = 0; k < 5; ++k) cv[k] = P.cnt[(size_t)5 * c + k];
            bd[c] = sg;
#pragma unroll
            for (int k = 0; k < 7; ++k) bi[(size_t)k * nc + c] = pv[k];
#pragma unroll
            for (int k = 0; k < 5; ++k) bc[(size_t)5 * c + k] = cv[k];
        }
        const int32_t n1 = pv[2];
        bool b = in && (n1 < 0 || pv[5] != 0);
        const int32_t v = n1 > 0 ? n1 : 0;
        int32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        if (t < 16) {
            const int32_t a = wsum[t];
            int32_t s = a;
            for (int o = 1; o < 16; o <<= 1) {
                const int32_t y = __shfl_up(s, o, 16);
                if (t >= o) s += y;
            }
            wsum[t] = s - a;
        }
        __syncthreads();
        const int32_t excl = carry + wsum[wv] + x - v;
        if (in && n1 > 0 && pv[0] != excl + 1) b = true;
        if (b) bad = 1;
        __syncthreads();
        if (t == 1023) carry = excl + v;
        __syncthreads();
    }
    if (t == 0) {
        if (carry != P.ctr->n_total) bad = 1;
        int32_t *ab = (int32_t *)P.abort_line;
        ab[0] = bad ? SLAB_ABORT_STATE : 0;
        ab[1] = 0;
    }
    if (t > 0 && t < SLAB_LINE) P.abort_line[t] = 0;
}

// After a slab launch: the abort word into host memory the caller reads
// after the stream synchronises.
__global__ void k_slab_report(const int32_t *__restrict__ abort_flag, int32_t *out) {
    if (threadIdx.x == 0) *(volatile int32_t *)out = *abort_flag;
}

// ---------------------------------------------------------------------------
// host side

namespace {

typedef void (*SlabKernel)(SlabArgs);

struct SlabShape {
    int R = 0, spc = 0, n_slabs = 0, spx = 0, grid = 0;
    SlabKernel fn = nullptr;
};

SlabKernel slab_kernel(int R, bool avg = false) {
    switch (R) {
    case 1: return avg ? k_slab_couette<1, true> : k_slab_couette<1, false>;
    case 2: return avg ? k_slab_couette<2, true> : k_slab_couette<2, false>;
    case 4: return avg ? k_slab_couette<4, true> : k_slab_couette<4, false>;
    case 8: return avg ? k_slab_couette<8, true> : k_slab_couette<8, false>;
    default: return nullptr;
    }
}

// Largest per-slab occupancy for a given number of cells per slab.
int32_t max_slab_particles(const std::vector<int32_t> &n1, int spc) {
    int32_t best = 0;
    for (size_t c0 = 0; c0 < n1.size(); c0 += spc) {
        int32_t s = 0;
        for (size_t c = c0; c < std::min(n1.size(), c0 + (size_t)spc); ++c) s += n1[c];
        best = std::max(best, s);
    }
    return best;
}

}  // namespace

extern "C" int mz_set_slab_shape(mz_store *h, int32_t cells_per_slab, int32_t particles_per_thread) {
    MZ_CHECK(cells_per_slab >= 0 && cells_per_slab <= SLAB_MAXC, "slab shape: cells per slab out of range");
    MZ_CHECK(particles_per_thread == 0 || slab_kernel(particles_per_thread) != nullptr,
             "slab shape: particles per thread must be 0, 1, 2, 4 or 8");
    h->slab_spc = cells_per_slab;
    h->slab_ppt = particles_per_thread;
    return 0;
}

extern "C" int mz_slab_occupancy(int32_t particles_per_thread, int32_t *blocks_per_cu, int32_t *lds_bytes) {
    SlabKernel fn = slab_kernel(particles_per_thread);
    MZ_CHECK(fn != nullptr, "slab occupancy: particles per thread must be 1, 2, 4 or 8");
    int per_cu = 0;
    MZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, SLAB_THREADS, 0));
    hipFuncAttributes fa;
    MZ_HIP(hipFuncGetAttributes(&fa, (const void *)fn));
    if (blocks_per_cu) *blocks_per_cu = per_cu;
    if (lds_bytes) *lds_bytes = (int32_t)fa.sharedSizeBytes;
    return 0;
}

extern "C" int mz_get_slab_info(mz_store *h, int32_t *info) {
    MZ_CHECK(info != nullptr, "slab info: NULL output");
    for (int k = 0; k < 4; ++k) info[k] = h->slab_info[k];
    return 0;
}

// Shape of the slab launch: the fewest particles per thread whose LDS slab
// holds the current largest slab with room to spare, with as many slabs as
// can be resident.  Kept per store until a launch aborts or the requested
// shape changes, so a steady run reads nothing back from the device.
static int slab_pick_shape(mz_store *h, bool avg, MzSlabShapeCache &out) {
    MzSpecies &s = h->sp[0];
    const int nc = h->n_cells;
    out = MzSlabShapeCache{};
    std::vector<int32_t> n1(nc);
    MZ_HIP(hipStreamSynchronize(h->stream));
    MZ_HIP(hipMemcpy(n1.data(), s.pia.n1, nc * sizeof(int32_t), hipMemcpyDeviceToHost));
    int dev = 0, n_cu = 0;
    MZ_HIP(hipGetDevice(&dev));
    MZ_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    const int Rs[4] = {1, 2, 4, 8};
    for (int R : Rs) {
        if (h->slab_ppt && R != h->slab_ppt) continue;
        if (R < h->slab_min_ppt) continue;
        SlabKernel fn = slab_kernel(R, avg);
        int per_cu = 0;
        MZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, SLAB_THREADS, 0));
        const int max_wg = per_cu * n_cu;
        if (max_wg <= 0) continue;
        int spc = h->slab_spc ? h->slab_spc : (nc + max_wg - 1) / max_wg;
        if (spc > SLAB_MAXC) continue;
        const int n_slabs = (nc + spc - 1) / spc;
        const int spx = (n_slabs + 7) / 8;
        if (8 * spx > max_wg) continue;
        const int64_t P = (int64_t)R * SLAB_THREADS;
        // room for the slab's occupancy to fluctuate (25% plus four standard
        // deviations of a Poisson count); a slab that still overflows aborts
        // the launch and the steps run on the step graph
        const double mx = (double)max_slab_particles(n1, spc);
        if (1.25 * mx + 4.0 * std::sqrt(mx) > (double)P) continue;
        out.R = R; out.spc = spc; out.n_slabs = n_slabs; out.spx = spx; out.grid = 8 * spx;
        out.fn = (const void *)fn;
        break;
    }
    out.valid = true;
    out.avg = avg;
    out.req_spc = h->slab_spc;
    out.req_ppt = h->slab_ppt;
    return 0;
}

// Runs n_steps on the slab kernel.  *ran = 1 when it did; 0 when the state or
// the shape does not allow it, or when the launch aborted (state restored).
// Host work per call: one check-and-backup kernel, the launch, a one-word
// report kernel and one stream synchronisation.
int mz_couette_slab(mz_store *h, int32_t n_steps, uint64_t step0, double dt, const double *dV, int32_t props_every,
                    int *ran, double avg_inv, double mass, int retry) {
    const bool avg = avg_inv > 0.0;
    *ran = 0;
    MzSpecies &s = h->sp[0];
    const int nc = h->n_cells;
    // outcome codes of slab_info[3] when the slab kernel does not run
    const int32_t NOT_SORTED = -2, NO_SHAPE = -3;
    h->slab_info[0] = h->slab_info[1] = h->slab_info[2] = 0;
    h->slab_info[3] = NO_SHAPE;
    if (n_steps <= 0 || n_steps >= (1 << 28)) return 0;
    MzSlabShapeCache &sh = h->slab_shape;
    if (!sh.valid || sh.avg != avg || sh.req_spc != h->slab_spc || sh.req_ppt != h->slab_ppt) {
        int r = slab_pick_shape(h, avg, sh);
        if (r) return r;
    }
    if (!sh.fn) return 0;
    if (!h->h_report) {
        MZ_HIP(hipHostMalloc((void **)&h->h_report, 64, hipHostMallocDefault));
    }

    // scratch: mailboxes, flags, abort word, and a copy of what the kernel
    // overwrites before it can abort (collision factors, PIA, counters)
    const size_t mb_words = (size_t)sh.n_slabs * 2 * 2 * 8 * SLAB_MB_CAP;
    const size_t fl_words = (size_t)sh.n_slabs * 2 * SLAB_LINE;
    const size_t fin_words = (size_t)sh.n_slabs * SLAB_LINE;
    uint64_t *mb = (uint64_t *)mz_scratch(h, "slab_mb", mb_words * 8);
    uint64_t *fl = (uint64_t *)mz_scratch(h, "slab_flags", (fl_words + fin_words + SLAB_LINE) * 8);
    const size_t bak_bytes = (size_t)nc * (sizeof(double) + 7 * sizeof(int32_t) + 5 * sizeof(int32_t)) +
                             sizeof(MzCounters);
    char *bak = (char *)mz_scratch(h, "slab_backup", bak_bytes);
    double *surf_part = nullptr;
    if (avg) surf_part = (double *)mz_scratch(h, "slab_surf", (size_t)sh.n_slabs * 4 * MZ_SURF_Q * sizeof(double));
    MZ_CHECK(mb && fl && bak && (!avg || surf_part), "couette slab: scratch allocation failed");
    uint64_t *fin = fl + fl_words;
    uint64_t *abort_line = fin + fin_words;
    int32_t *abort_flag = (int32_t *)abort_line;

    SlabPrepArgs pa{};
    pa.pia = s.pia;
    pa.ctr = s.ctr;
    pa.sgwm = h->d_sgwm;
    pa.cnt = h->d_cfcnt;
    pa.n_cells = nc;
    pa.bak = bak;
    pa.flags = fl;
    pa.n_flag_words = (int64_t)(fl_words + fin_words);
    pa.abort_line = abort_line;
    hipLaunchKernelGGL(k_slab_prep, dim3(1), dim3(1024), 0, h->stream, pa);
    MZ_HIP(hipGetLastError());

    SlabArgs A{};
    A.a = mz_view(s);
    A.bx0 = s.xb[0]; A.bx1 = s.xb[1]; A.bx2 = s.xb[2];
    A.bv0 = s.vb[0]; A.bv1 = s.vb[1]; A.bv2 = s.vb[2];
    A.bw = s.wb; A.bid = s.idb;
    A.g = h->grid;
    A.wl = h->walls;
    A.dt = dt;
    const MzInteraction &it = h->inter[0];
    A.mu1 = it.mu1; A.mu2 = it.mu2; A.factor = it.factor; A.expo = it.expo;
    A.mass = mass;
    A.V = dV;
    A.sgwm = h->d_sgwm;
    A.cnt = h->d_cfcnt;
    A.rp = mz_rng_params(h, 0, 0, 0);
    A.step0 = step0;
    A.n_steps = n_steps;
    A.props_every = props_every;
    A.out_np = h->d_np; A.out_n = h->d_n; A.out_v = h->d_v; A.out_T = h->d_T;
    A.n_cells = nc;
    A.spc = sh.spc;
    A.n_slabs = sh.n_slabs;
    A.spx = sh.spx;
    A.mb = mb;
    A.flag = fl;
    A.fin = fin;
    A.abort_flag = abort_flag;
    A.avg_inv = avg_inv;
    A.surf_f[0] = mass / dt * h->surf_inv_area[0];
    A.surf_f[1] = mass / dt * h->surf_inv_area[1];
    A.surf_part = surf_part;
    A.pavg = h->d_pavg;
    void *args[] = {&A};
    // Every slab waits on its neighbours, so all workgroups should be
    // resident at once: the grid is at most what the occupancy query allows
    // on an otherwise idle device.  Work still running on another store's
    // stream holds CUs and would start some slabs late, so the device is
    // drained first when any other store is busy.  A slab whose neighbour
    // never arrives gives up after a bounded spin and the launch aborts; the
    // host then restores the state and runs the steps on the step graph.
    if (mz_other_stores_busy(h)) MZ_HIP(hipDeviceSynchronize());
    {
        MzTimed tm(h, "slab", n_steps);
        MZ_HIP(hipLaunchKernel(sh.fn, dim3(sh.grid), dim3(SLAB_THREADS), args, 0, h->stream));
    }
    hipLaunchKernelGGL(k_slab_report, dim3(1), dim3(64), 0, h->stream, abort_flag, h->h_report);
    MZ_HIP(hipGetLastError());
    MZ_HIP(hipStreamSynchronize(h->stream));
    const int32_t outcome = *(volatile int32_t *)h->h_report;
    h->slab_info[0] = sh.R;
    h->slab_info[1] = sh.spc;
    h->slab_info[2] = sh.n_slabs;
    if (outcome == SLAB_ABORT_STATE) {
        // precondition not met: no slab wrote anything
        h->slab_info[3] = NOT_SORTED;
        sh.valid = false;
        return 0;
    }
    h->slab_info[3] = outcome ? -1 : 1;
    if (outcome) {
        sh.valid = false;
        const double *bd = (const double *)bak;
        const int32_t *bi = (const int32_t *)(bak + (size_t)nc * sizeof(double));
        MZ_HIP(hipMemcpyAsync(h->d_sgwm, bd, nc * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        int32_t *pia_arr[7] = {s.pia.start1, s.pia.end1, s.pia.n1, s.pia.start2, s.pia.end2, s.pia.n2, s.pia.nloc};
        for (int k = 0; k < 7; ++k)
            MZ_HIP(hipMemcpyAsync(pia_arr[k], bi + (size_t)k * nc, nc * sizeof(int32_t), hipMemcpyDeviceToDevice,
                                  h->stream));
        MZ_HIP(hipMemcpyAsync(h->d_cfcnt, bi + (size_t)7 * nc, nc * 5 * sizeof(int32_t), hipMemcpyDeviceToDevice,
                              h->stream));
        MZ_HIP(hipMemcpyAsync(s.ctr, bi + (size_t)12 * nc, sizeof(MzCounters), hipMemcpyDeviceToDevice, h->stream));
        MZ_HIP(hipStreamSynchronize(h->stream));
        // a slab that outgrew its LDS is the usual cause: with an automatic
        // shape, try once more with twice the particles per thread (kept for
        // later calls on this store)
        if (!retry && h->slab_ppt == 0 && sh.R < 8) {
            h->slab_min_ppt = 2 * sh.R;
            return mz_couette_slab(h, n_steps, step0, dt, dV, props_every, ran, avg_inv, mass, 1);
        }
        return 0;
    }
    if (avg) {
        hipLaunchKernelGGL(k_slab_surf_finalize, dim3(2 * MZ_SURF_Q), dim3(64), 0, h->stream, surf_part, sh.n_slabs,
                           h->d_surf);
        MZ_HIP(hipGetLastError());
    }
    mz_swap_buffers(s);
    s.contig_known = true;
    *ran = 1;
    return 0;
}
