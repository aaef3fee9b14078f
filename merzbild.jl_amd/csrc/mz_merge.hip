// mz_merge.hip -- per-cell merging kernels for gfx950: octree N:2
// (src/merging/merging_octree_N2.jl) and NNLS moment matching
// (src/merging/merging_nnls.jl, src/merging/nnls.jl).
//
// Octree: one workgroup per merged cell.  The particle order list (the
// reference's particle_indexes_sorted) and the bins live in LDS.  Decisions
// that the reference takes sequentially -- which bin to refine, the order of
// the stable 8-way counting sort, the per-octant weight sums, the bin ids
// and the random sign draws -- are reproduced exactly:
//   * bounds come from min/max (order independent) and the split velocity
//     from 0.5 (v_min + v_max), so octants are bit-exact;
//   * the octant scatter ranks lanes with ballots and reproduces the
//     reference's placement (each child in reverse parent order);
//   * each child's weight is summed sequentially in parent order, the
//     reference's accumulation order;
//   * the refine pick is the first maximum over bins 1..Nbins.
// Bin means and variances are block reductions (they only feed the new
// particle values, compared to 1e-12).
//
// A cell of more than OCT_MAX_N particles, or one whose split loop would
// need more than OCT_MAX_BINS bins, is handed on: the LDS kernel lists it and
// a second kernel merges the listed cells with the same code on arrays in
// device memory (one slice per workgroup, one wavefront per workgroup, the
// workgroups taking the listed cells in turn).  The split loop keeps its
// state in LDS either way; the cell's arrays are then read through L2.
#include "mz_internal.h"

#include <algorithm>
#include <utility>
#include <vector>

#define OCT_MAX_N 2048
#define OCT_MAX_BINS 512
#define OCT_LANE_NP 16        // bins of at most this many particles get a lane each for their moments
#define OCT_BIG_BLOCKS 4096   // most 64-bin blocks whose maxima the device-memory path keeps in LDS

struct OctArgs {
    MzSpView sp;
    MzSpView tmp;             // second particle buffer, scratch for new particle values
    int32_t threshold, target_np, split, init_bounds, bounds_mode, max_Nbins, max_depth, grid_clamp;
    double min_x, max_x;
    MzRngParams rp;
    int32_t *old_pia;         // [n_cells][5]: start1, n1, start2, n2, n_new (n_new = -1: not merged)
    int32_t *n_merged;
    int32_t n_cells;
    const uint8_t *mask;      // cells to merge (nullptr: every cell above the threshold)
    const int32_t *thresholds, *targets;   // per cell (nullptr: threshold, target_np for every cell)
    // cells handed to the device-memory path (big_list nullptr: such a cell
    // raises MZ_ERR_MERGE_SIZE): four lists of n_cells entries, one per size
    // class (more than 16384, 8192, 4096 particles, the rest); big_info[0]
    // counts the listed cells, [1] is their largest N, [2..5] count each
    // class, [6] is the device-memory kernel's queue
    int32_t *big_list, *big_info;
    // that path's arrays, one slice of big_ncap particles and big_bcap bins per workgroup
    int32_t *big_pis, *big_psort, *big_bslot;
    double *big_stage;        // per workgroup: w, v0, v1, v2 of two tree-order buffers
    uint8_t *big_poct;
    double *big_bkey;
    void *big_bins;
    int32_t big_ncap, big_bcap;
    int32_t big_blocks;       // 1: the 64-bin blocks' maxima are kept in dynamic LDS
};

// A bin of the octree; IdxT holds particle counts and positions (int16_t in
// LDS, where a cell holds at most OCT_MAX_N = 2048 particles, int32_t in
// device memory).
template <class IdxT>
struct OctBinT {
    double w;
    double vmin[3], vmax[3];
    IdxT np, start, end;
    int8_t depth, refine;
};
using OctBin = OctBinT<int16_t>;

// The working arrays of one cell's merge: the particle order, its scatter
// target (after the splits: per-bin output slots and draw offsets), each
// particle's octant, the bins and the refine pick's keys.
template <class IdxT>
struct OctWork {
    IdxT *pis, *psort;
    uint8_t *poct;
    OctBinT<IdxT> *bins;
    double *bkey;
    int32_t *bslot, *bdraw;
    int32_t bins_cap;
    // per block of 64 bins: the largest key and its lowest bin (nullptr:
    // the pick scans the keys themselves)
    double *bmv;
    int32_t *bmi;
    // device-memory path: the particles' ids and values in tree order, in the
    // buffer of the parity of their bin's depth (a split moves its bin's
    // particles to the other buffer, so a pass reads a contiguous segment)
    int32_t *sid0;            // buffer 0's ids; buffer 1's at sid0 + sid_q
    int64_t sid_q;
    double *stg;              // w, v0, v1, v2 of buffer 0, then of buffer 1, n_stg each
    int64_t n_stg;
};


// The small per-cell state of the split loop and the output scan, in LDS.
struct OctShared {
    double red[4];
    double vel_mid[3];
    int32_t cnt8[9], start8[9];
    double ndens8[9];
    int32_t Nbins_s, total_post_s, stop_s, defer_s;
    int64_t wsc[4];
    int32_t n_new_s, ticket_s;
    double ckey[8];           // a split's children: key and bin id, by child
    int32_t cbi[8];
};

// orders LDS accesses among the lanes of one wave (a barrier for code that a
// single wave runs)
__device__ inline void oct_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int64_t oct_slot(int li, int32_t s1, int32_t n1, int32_t s2) {
    return li < n1 ? (int64_t)s1 - 1 + li : (int64_t)s2 - 1 + (li - n1);
}

// Buffer q's ids, and its values k (0: w, 1-3: v0-v2).
template <class IdxT>
__device__ __forceinline__ int32_t *oct_sid(const OctWork<IdxT> &W, int q) {
    return W.sid0 + q * W.sid_q;
}
template <class IdxT>
__device__ __forceinline__ double *oct_stg(const OctWork<IdxT> &W, int q, int k) {
    return W.stg + (int64_t)(4 * q + k) * W.n_stg;
}

// The values of the particle at tree position i of a bin of depth parity par:
// through the particle order (LDS path) or from the staged buffer.
template <class IdxT, bool BIG>
__device__ __forceinline__ void oct_ld(const OctArgs &A, const OctWork<IdxT> &W, const IdxT *pis, int32_t s1,
                                       int32_t n1, int32_t s2, int par, int i, double &w, double &a0, double &a1,
                                       double &a2) {
    if constexpr (BIG) {
        w = oct_stg(W, par, 0)[i]; a0 = oct_stg(W, par, 1)[i]; a1 = oct_stg(W, par, 2)[i]; a2 = oct_stg(W, par, 3)[i];
    } else {
        const int64_t p = oct_slot(pis[i], s1, n1, s2);
        w = A.sp.w[p]; a0 = A.sp.v0[p]; a1 = A.sp.v1[p]; a2 = A.sp.v2[p];
    }
}

// The id (place in the cell) of the particle at tree position i.
template <class IdxT, bool BIG>
__device__ __forceinline__ int oct_id(const OctWork<IdxT> &W, const IdxT *pis, int par, int i) {
    if constexpr (BIG) return oct_sid(W, par)[i];
    else return pis[i];
}

__device__ inline double blk_min(double v, double *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
    __syncthreads();
    return r;
}
__device__ inline double blk_max(double v, double *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    __syncthreads();
    return r;
}
__device__ inline double blk_sum(double v, double *red) {
    v = mz_wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return r;
}

// min/max of v over pis[bs..be] (bin_bounds_recompute!, merging_octree_N2.jl:382-418)
template <class IdxT, int NW>
__device__ void oct_bounds(const OctArgs &A, const IdxT *pis, int bs, int be, int32_t s1, int32_t n1, int32_t s2,
                           double *mn, double *mx, double *red) {
    for (int d = 0; d < 3; ++d) {
        const double *vv = d == 0 ? A.sp.v0 : (d == 1 ? A.sp.v1 : A.sp.v2);
        double a = 9299792458.0, b = -9299792458.0;
        for (int i = bs + threadIdx.x; i <= be; i += blockDim.x) {
            double x = vv[oct_slot(pis[i], s1, n1, s2)];
            a = fmin(a, x);
            b = fmax(b, x);
        }
        if (NW == 1) {
            mn[d] = mz_wave_min(a);
            mx[d] = mz_wave_max(b);
        } else {
            mn[d] = blk_min(a, red);
            mx[d] = blk_max(b, red);
        }
    }
}

// The end of an N:2 merge of cell c (compute_new_particles!,
// merging_octree_N2.jl:736-813 and merging_grid.jl:393-482): the n_new
// particles staged in tmp at the cell's first n_new logical slots replace
// the cell's data, the weights of the rest are zeroed, and the PIA loses
// N - n_new particles from the end, group 2 first (delete_particle_end!,
// particles.jl:478-566); old_pia keeps what k_merge_push_ids needs.  The
// reference clears contiguous unless the deletions stay inside the last
// cell's group 2.  Called by every thread of the block after a barrier.
__device__ void merge_finish(const MzSpView &sp, const MzSpView &tmp, int c, int N, int n_new, int32_t s1,
                             int32_t n1, int32_t s2, int32_t n2, int32_t *old_pia, int32_t *n_merged,
                             int32_t n_cells) {
    for (int li = threadIdx.x; li < N; li += blockDim.x) {
        int64_t p = oct_slot(li, s1, n1, s2);
        if (li < n_new) {
            sp.x0[p] = tmp.x0[p]; sp.x1[p] = tmp.x1[p]; sp.x2[p] = tmp.x2[p];
            sp.v0[p] = tmp.v0[p]; sp.v1[p] = tmp.v1[p]; sp.v2[p] = tmp.v2[p];
            sp.w[p] = tmp.w[p];
        } else {
            sp.w[p] = 0.0;
        }
    }
    if (threadIdx.x == 0) {
        int32_t *op = old_pia + c * 5;
        op[0] = s1; op[1] = n1; op[2] = s2; op[3] = n2; op[4] = n_new;
        int32_t n_del = N - n_new;
        int32_t d2 = min(n_del, n2), d1 = n_del - d2;
        int32_t nn2 = n2 - d2, nn1 = n1 - d1;
        sp.pia.n2[c] = nn2;
        sp.pia.end2[c] = nn2 > 0 ? s2 + nn2 - 1 : -1;
        if (nn2 == 0) sp.pia.start2[c] = 0;
        sp.pia.n1[c] = nn1;
        sp.pia.end1[c] = nn1 > 0 ? s1 + nn1 - 1 : -1;
        if (nn1 == 0) sp.pia.start1[c] = 0;
        sp.pia.nloc[c] = nn1 + nn2;
        if (!(c == n_cells - 1) || (n_del > n2)) sp.ctr->contiguous = 0;
        atomicAdd(n_merged, 1);
    }
}

// Merges cell c (N particles, target_np) with the working arrays W; NW
// waves per workgroup.  With BIG false a cell that would need more bins than
// W holds is listed for the device-memory path (or raises
// MZ_ERR_MERGE_SIZE when there is none) and left unchanged.
// Copies particle src to slot dst of the second buffer.
__device__ inline void oct_copy(const OctArgs &A, int64_t src, int64_t dst) {
    A.tmp.x0[dst] = A.sp.x0[src]; A.tmp.x1[dst] = A.sp.x1[src]; A.tmp.x2[dst] = A.sp.x2[src];
    A.tmp.v0[dst] = A.sp.v0[src]; A.tmp.v1[dst] = A.sp.v1[src]; A.tmp.v2[dst] = A.sp.v2[src];
    A.tmp.w[dst] = A.sp.w[src];
}

// compute_new_particles! for one bin (merging_octree_N2.jl:736-813): its two
// particles at slots slot, slot + 1 of the second buffer, the 3 + 3 sign
// draws at the bin's place (draw) in the cell's sequence.
__device__ inline void oct_emit(const OctArgs &A, int c, int32_t draw, int slot, double Wb, const double *vm,
                                const double *xm, const double *sdv, const double *sdx, int32_t s1, int32_t n1,
                                int32_t s2) {
    MzStream st = mz_stream((uint64_t)c);
    st.k = 6 * (uint64_t)draw;
    double dv[3], dx[3];
    for (int d = 0; d < 3; ++d) dv[d] = mz_rand(A.rp, st) < 0.5 ? -1.0 : 1.0;
    for (int d = 0; d < 3; ++d) dx[d] = mz_rand(A.rp, st) < 0.5 ? -1.0 : 1.0;
    const double w1 = 0.5 * Wb;
    for (int k = 0; k < 2; ++k) {
        const double sg = k == 0 ? 1.0 : -1.0;
        const int64_t dst = oct_slot(slot + k, s1, n1, s2);
        double x0 = xm[0] + sg * dx[0] * sdx[0];
        if (A.grid_clamp) x0 = x0 < A.min_x ? A.min_x : (x0 > A.max_x ? A.max_x : x0);
        A.tmp.x0[dst] = x0;
        A.tmp.x1[dst] = xm[1] + sg * dx[1] * sdx[1];
        A.tmp.x2[dst] = xm[2] + sg * dx[2] * sdx[2];
        A.tmp.v0[dst] = vm[0] + sg * dv[0] * sdv[0];
        A.tmp.v1[dst] = vm[1] + sg * dv[1] * sdv[1];
        A.tmp.v2[dst] = vm[2] + sg * dv[2] * sdv[2];
        A.tmp.w[dst] = w1;
    }
}

// Lists cell c (N particles) for the device-memory path, in its size class.
__device__ inline void oct_defer(const OctArgs &A, int c, int32_t N) {
    A.old_pia[c * 5 + 4] = -1;
    if (A.big_list) {
        const int cls = N > 16384 ? 0 : (N > 8192 ? 1 : (N > 4096 ? 2 : 3));
        A.big_list[(size_t)cls * A.n_cells + atomicAdd(&A.big_info[2 + cls], 1)] = c;
        atomicAdd(&A.big_info[0], 1);
        atomicMax(&A.big_info[1], N);
    } else {
        atomicOr(&A.sp.ctr->err, MZ_ERR_MERGE_SIZE);
    }
}

template <class IdxT, int NW, bool BIG>
__device__ __forceinline__ void oct_merge_cell(const OctArgs &A, const int c, const int32_t N,
                                               const int32_t target_np, const OctWork<IdxT> &W, OctShared &S) {
    IdxT *const pis = W.pis;
    IdxT *const psort = W.psort;
    uint8_t *const poct = W.poct;
    OctBinT<IdxT> *const bins = W.bins;
    // the refine pick's keys: bins[b].w when bin b may be refined, else -1
    // (contiguous, so the pick's scan reads without bank conflicts)
    double *const bkey = W.bkey;
    double *const red = S.red;
    double *const vel_mid = S.vel_mid;
    int32_t *const cnt8 = S.cnt8;
    int32_t *const start8 = S.start8;
    double *const ndens8 = S.ndens8;
    int32_t &Nbins_s = S.Nbins_s;
    int32_t &total_post_s = S.total_post_s;
    int32_t &stop_s = S.stop_s;
    const int32_t s1 = A.sp.pia.start1[c], n1 = A.sp.pia.n1[c], s2 = A.sp.pia.start2[c], n2 = A.sp.pia.n2[c];
    const int max_bins = min(A.max_Nbins, W.bins_cap);
    // init_octree! (merging_octree_N2.jl:947-984)
    if constexpr (BIG) {
        for (int i = threadIdx.x; i < N; i += blockDim.x) {
            const int64_t p = oct_slot(i, s1, n1, s2);
            W.sid0[i] = i;
            W.stg[i] = A.sp.w[p];
            W.stg[W.n_stg + i] = A.sp.v0[p];
            W.stg[2 * W.n_stg + i] = A.sp.v1[p];
            W.stg[3 * W.n_stg + i] = A.sp.v2[p];
        }
    } else {
        for (int i = threadIdx.x; i < N; i += blockDim.x) pis[i] = (IdxT)i;
    }
    double mn[3], mx[3];
    if (A.init_bounds == 3) {
        for (int d = 0; d < 3; ++d) { mn[d] = -299792458.0; mx[d] = 299792458.0; }
    } else {
        __syncthreads();
        oct_bounds<IdxT, NW>(A, pis, 0, N - 1, s1, n1, s2, mn, mx, red);
        if (A.init_bounds == 2) {
            for (int d = 0; d < 3; ++d) {
                double m = fmax(fabs(mn[d]), fabs(mx[d]));
                mn[d] = -m;
                mx[d] = m;
            }
        }
    }
    if (threadIdx.x == 0) {
        OctBinT<IdxT> &b = bins[0];
        b.w = 1e50;
        for (int d = 0; d < 3; ++d) { b.vmin[d] = mn[d]; b.vmax[d] = mx[d]; }
        b.np = N; b.start = 0; b.end = N - 1; b.depth = 0;
        b.refine = (N > 2) && (A.max_depth > 0);
        bkey[0] = b.refine ? b.w : -1.0;
        if (W.bmv) {
            W.bmv[0] = bkey[0];
            W.bmi[0] = 0;
        }
        Nbins_s = 1;
        total_post_s = N >= 2 ? 2 : N;
        for (int d = 0; d < 3; ++d) vel_mid[d] = 0.0;
        stop_s = 0;
        S.defer_s = 0;
    }
    __syncthreads();

    // compute_octree! (merging_octree_N2.jl:998-1039).  Each split depends on
    // the one before, so the loop runs on wave 0 alone, where a barrier is a
    // wave-level fence; the other waves wait at the block barrier after it.
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        for (;;) {
            // refine pick: first maximum of w over refinable bins (lowest id on ties)
            double bw = -1.0;
            int bid = -1;
            if (W.bmv) {
                // the blocks' maxima (each the lowest bin of its block's maximum)
                for (int k = lane; k < (Nbins_s + 63) >> 6; k += 64) {
                    const double kb = W.bmv[k];
                    if (kb > bw) { bw = kb; bid = W.bmi[k]; }
                }
            } else {
                for (int b = lane; b < Nbins_s; b += 64) {
                    const double kb = bkey[b];
                    if (kb > bw) { bw = kb; bid = b; }
                }
            }
            // the largest key, then the lowest bin holding it (two wave reductions)
            {
                const double mw = mz_wave_max(bw);
                bid = mw > -1.0 ? mz_wave_min_i32(bw == mw ? bid : 0x7FFFFFFF) : -1;
            }
            if (bid < 0 || total_post_s + 14 > target_np) break;
            // split_bin! (merging_octree_N2.jl:503-633)
            const int b = bid;
            // the keys of b's block, loaded with bin b: the block's maximum is
            // rebuilt from them after the split
            double k0key = -2.0;
            if (W.bmv && (b & ~63) + lane < Nbins_s) k0key = bkey[(b & ~63) + lane];
            const int bs = bins[b].start, be = bins[b].end;
            const int depth = bins[b].depth;
            const int par = depth & 1;
            if (A.bounds_mode == 2) {
                double bmn[3], bmx[3];
                for (int d = 0; d < 3; ++d) {
                    const double *vv;
                    if constexpr (BIG) vv = d == 0 ? oct_stg(W, par, 1) : (d == 1 ? oct_stg(W, par, 2) : oct_stg(W, par, 3));
                    else vv = d == 0 ? A.sp.v0 : (d == 1 ? A.sp.v1 : A.sp.v2);
                    double lo = 9299792458.0, hi = -9299792458.0;
                    for (int i = bs + lane; i <= be; i += 64) {
                        double x;
                        if constexpr (BIG) x = vv[i];
                        else x = vv[oct_slot(pis[i], s1, n1, s2)];
                        lo = fmin(lo, x);
                        hi = fmax(hi, x);
                    }
                    bmn[d] = mz_wave_min(lo);
                    bmx[d] = mz_wave_max(hi);
                }
                if (lane == 0 && be >= bs)
                    for (int d = 0; d < 3; ++d) { bins[b].vmin[d] = bmn[d]; bins[b].vmax[d] = bmx[d]; }
                oct_wave_sync();
            }
            // the parent's bounds, read once by every lane
            double pmin[3], pmax[3];
            for (int d = 0; d < 3; ++d) { pmin[d] = bins[b].vmin[d]; pmax[d] = bins[b].vmax[d]; }
            if (lane == 0) {
                if (A.split == 1) {
                    for (int d = 0; d < 3; ++d) vel_mid[d] = 0.5 * (pmin[d] + pmax[d]);
                } else {
                    // compute_v_mean! keeps the previous middle velocity in the sum
                    double nt = 0.0;
                    for (int i = bs; i <= be; ++i) {
                        double w, a0, a1, a2;
                        oct_ld<IdxT, BIG>(A, W, pis, s1, n1, s2, par, i, w, a0, a1, a2);
                        nt += w;
                        vel_mid[0] = vel_mid[0] + w * a0;
                        vel_mid[1] = vel_mid[1] + w * a1;
                        vel_mid[2] = vel_mid[2] + w * a2;
                    }
                    for (int d = 0; d < 3; ++d) vel_mid[d] = vel_mid[d] / nt;
                }
                for (int o = 0; o < 9; ++o) { cnt8[o] = 0; ndens8[o] = 0.0; }
            }
            oct_wave_sync();
            // each particle's octant (compute_octant, :304-316: strict >), the
            // octants' counts from one ballot per octant and chunk of 64, and the
            // octants' weights in the reference's accumulation order (the bin's
            // particles in order, each weight added to its octant's sum: the
            // chunk's lanes are walked in order, lane o keeping octant o's sum)
            int cq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            double sw = 0.0;
            const double mid0 = vel_mid[0], mid1 = vel_mid[1], mid2 = vel_mid[2];
            // the next chunk's particle loads are issued before this chunk's
            // ballots and sums, so their latency overlaps that work
            // A bin of at most 64 particles is one chunk: its ids and octants
            // stay in registers and the scatter writes pis directly.
            const bool one = be - bs < 64;
            double nw = 0.0, nv0 = 0.0, nv1 = 0.0, nv2 = 0.0;
            IdxT pi1 = 0;
            if (bs + lane <= be) {
                pi1 = oct_id<IdxT, BIG>(W, pis, par, bs + lane);
                oct_ld<IdxT, BIG>(A, W, pis, s1, n1, s2, par, bs + lane, nw, nv0, nv1, nv2);
            }
            // the next chunk's ids, loaded a chunk ahead of its values (the
            // staged values need no ids)
            IdxT pin = 0;
            if (!BIG && bs + 64 + lane <= be) pin = pis[bs + 64 + lane];
            int o1 = 0;
            for (int i0 = bs; i0 <= be; i0 += 64) {
                const int i = i0 + lane;
                const double wi = nw, c0 = nv0, c1 = nv1, c2 = nv2;
                if (i + 64 <= be) {
                    if constexpr (BIG) {
                        nw = oct_stg(W, par, 0)[i + 64]; nv0 = oct_stg(W, par, 1)[i + 64];
                        nv1 = oct_stg(W, par, 2)[i + 64]; nv2 = oct_stg(W, par, 3)[i + 64];
                    } else {
                        const int64_t p = oct_slot(pin, s1, n1, s2);
                        nw = A.sp.w[p]; nv0 = A.sp.v0[p]; nv1 = A.sp.v1[p]; nv2 = A.sp.v2[p];
                    }
                }
                if (!BIG && i + 128 <= be) pin = pis[i + 128];
                int o = 0;
                if (i <= be) {
                    o = 1;
                    if (c0 > mid0) o += 1;
                    if (c1 > mid1) o += 2;
                    if (c2 > mid2) o += 4;
                    if (!one) poct[i - bs] = (uint8_t)o;
                }
                o1 = o;
                uint64_t mq = 0;   // lane q: the chunk's lanes in octant q
#pragma unroll
                for (int q = 1; q <= 8; ++q) {
                    const uint64_t m = __ballot(o == q);
                    cq[q] += __popcll(m);
                    if (lane == q) mq = m;
                }
                // lane q adds its octant's weights in lane order, four at a
                // time, each fetched from its lane with ds_bpermute
                const int wlo = __double2loint(wi), whi = __double2hiint(wi);
                while (__ballot(mq != 0)) {
                    int src[4];
                    bool has[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        has[u] = mq != 0;
                        src[u] = has[u] ? (int)__builtin_ctzll(mq) : 0;
                        mq &= mq - 1;
                    }
                    double wv[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        wv[u] = __hiloint2double(__builtin_amdgcn_ds_bpermute(src[u] << 2, whi),
                                                 __builtin_amdgcn_ds_bpermute(src[u] << 2, wlo));
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (has[u]) sw += wv[u];
                }
            }
            if (lane >= 1 && lane <= 8) ndens8[lane] = sw;
            int sq[9];
            sq[0] = bs;
#pragma unroll
            for (int q = 1; q <= 8; ++q) sq[q] = q == 1 ? bs : sq[q - 1] + cq[q - 1];
#pragma unroll
            for (int q = 1; q <= 8; ++q)
                if (lane == q) { cnt8[q] = cq[q]; start8[q] = sq[q]; }
            // Scatter by octant.  The reference walks the bin forward and fills
            // each octant from its end (merging_octree_N2.jl:621-626), so every
            // child holds its particles in reverse parent order: a particle's
            // place is its octant's end less the octant's particles before it
            // (earlier chunks, then earlier lanes of its chunk).
            if (one) {
                const uint64_t lt = mz_lanemask_lt();
                int dst = 0;
#pragma unroll
                for (int q = 1; q <= 8; ++q) {
                    const uint64_t m = __ballot(o1 == q);
                    if (o1 == q) dst = sq[q] + cq[q] - 1 - __popcll(m & lt);
                }
                if (bs + lane <= be) {
                    if constexpr (BIG) {
                        // the values are still those of the bin's one chunk
                        const int q = par ^ 1;
                        oct_sid(W, q)[dst] = pi1;
                        oct_stg(W, q, 0)[dst] = nw; oct_stg(W, q, 1)[dst] = nv0; oct_stg(W, q, 2)[dst] = nv1; oct_stg(W, q, 3)[dst] = nv2;
                    } else {
                        pis[dst] = pi1;
                    }
                }
                oct_wave_sync();
            } else {
                const uint64_t lt = mz_lanemask_lt();
                int rq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                // each chunk's octants and ids are loaded a chunk ahead
                int on = bs + lane <= be ? poct[lane] : 0;
                IdxT pn = bs + lane <= be ? oct_id<IdxT, BIG>(W, pis, par, bs + lane) : (IdxT)0;
                double qw = 0.0, qa = 0.0, qb = 0.0, qc = 0.0;
                if (BIG && bs + lane <= be) oct_ld<IdxT, BIG>(A, W, pis, s1, n1, s2, par, bs + lane, qw, qa, qb, qc);
                for (int i0 = bs; i0 <= be; i0 += 64) {
                    const int i = i0 + lane;
                    const bool valid = i <= be;
                    const int o = on;
                    const IdxT pv = pn;
                    const double vw = qw, va = qa, vb = qb, vc = qc;
                    const bool vn = i + 64 <= be;
                    on = vn ? poct[i + 64 - bs] : 0;
                    pn = vn ? oct_id<IdxT, BIG>(W, pis, par, i + 64) : (IdxT)0;
                    if (BIG && vn) oct_ld<IdxT, BIG>(A, W, pis, s1, n1, s2, par, i + 64, qw, qa, qb, qc);
                    int dst = 0;
#pragma unroll
                    for (int q = 1; q <= 8; ++q) {
                        const uint64_t m = __ballot(o == q);
                        if (o == q) dst = sq[q] + cq[q] - 1 - (rq[q] + __popcll(m & lt));
                        rq[q] += __popcll(m);
                    }
                    if (valid) {
                        if constexpr (BIG) {
                            const int q = par ^ 1;
                            oct_sid(W, q)[dst] = pv;
                            oct_stg(W, q, 0)[dst] = vw; oct_stg(W, q, 1)[dst] = va; oct_stg(W, q, 2)[dst] = vb; oct_stg(W, q, 3)[dst] = vc;
                        } else {
                            psort[dst] = pv;
                        }
                    }
                }
                oct_wave_sync();
                if constexpr (!BIG) {
                    for (int i = bs + lane; i <= be; i += 256) {
                        IdxT t[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) t[u] = i + 64 * u <= be ? psort[i + 64 * u] : (IdxT)0;
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (i + 64 * u <= be) pis[i + 64 * u] = t[u];
                    }
                    oct_wave_sync();
                }
            }
            // the children (split_bin!, :588-633), lane o for octant o: the
            // k-th non-empty octant becomes bin b (k = 1) or Nb + k - 2
            // (get_new_bin_id), its particles the octant's segment
            const int Nb = Nbins_s;
            // a new bin's key and id on its octant's lane (for the blocks' maxima)
            double nkey = -2.0;
            int nbi = 0x7FFFFFFF;
            int n_ne = 0;
            {
                const int o = lane;
                const bool ne = o >= 1 && o <= 8 && cnt8[o] > 0;
                const uint64_t m = __ballot(ne);
                n_ne = __popcll(m);
                int contrib = 0;
                oct_wave_sync();   // every lane has read bin b before child 1 overwrites it
                if (ne) {
                    const int k = __popcll(m & mz_lanemask_lt()) + 1;
                    const int bi = k == 1 ? b : Nb + k - 2;
                    if (bi < W.bins_cap) {
                        OctBinT<IdxT> &nb = bins[bi];
                        nb.start = start8[o];
                        nb.end = start8[o] + cnt8[o] - 1;
                        if (A.bounds_mode == 1) {   // bin_bounds_inherit! (:341-367)
                            const int sx = (o - 1) & 1, sy = ((o - 1) >> 1) & 1, sz = ((o - 1) >> 2) & 1;
                            nb.vmin[0] = sx ? vel_mid[0] : pmin[0]; nb.vmax[0] = sx ? pmax[0] : vel_mid[0];
                            nb.vmin[1] = sy ? vel_mid[1] : pmin[1]; nb.vmax[1] = sy ? pmax[1] : vel_mid[1];
                            nb.vmin[2] = sz ? vel_mid[2] : pmin[2]; nb.vmax[2] = sz ? pmax[2] : vel_mid[2];
                        }
                        nb.np = cnt8[o];
                        nb.w = ndens8[o];
                        nb.depth = depth + 1;
                        nb.refine = (cnt8[o] > 2) && (depth + 1 < A.max_depth);
                        const double key = nb.refine ? nb.w : -1.0;
                        bkey[bi] = key;
                        S.ckey[k - 1] = key;
                        S.cbi[k - 1] = bi;
                        if (k > 1) { nkey = key; nbi = bi; }
                    }
                    contrib = cnt8[o] >= 2 ? 2 : cnt8[o];
                }
                const int added = mz_wave_sum_i(contrib);
                if (lane == 0) {
                    total_post_s += added - 2;
                    Nbins_s = Nb + n_ne - 1;
                    if (Nbins_s + 7 > A.max_Nbins) {
                        stop_s = 1;
                    } else if (Nbins_s + 7 > max_bins) {
                        // the arrays' bin budget is below the caller's max_Nbins:
                        // the cell goes to the device-memory path (or fails)
                        stop_s = 1;
                        S.defer_s = 1;
                    }
                }
            }
            oct_wave_sync();
            if (W.bmv) {
                // the keys of bin b and of the new bins Nb .. Nbins - 1 changed:
                // their blocks' maxima again (at most three blocks)
                const int nb_now = Nbins_s;
                const int k0 = b >> 6;
                // b's block from its keys as loaded, the children's keys put
                // in (b's key fell; new bins may lie in the block)
                {
                    const int i = (k0 << 6) + lane;
                    double v = i < nb_now ? k0key : -2.0;
                    for (int j = 0; j < n_ne; ++j)
                        if (S.cbi[j] == i) v = S.ckey[j];
                    const double mv = mz_wave_max(v);
                    const int vi = mz_wave_min_i32(i < nb_now && v == mv ? i : 0x7FFFFFFF);
                    v = mv;
                    if (lane == 0) {
                        W.bmv[k0] = v;
                        W.bmi[k0] = vi;
                    }
                }
                // the other blocks that gained new bins Nb .. nb_now - 1: their
                // old maximum against the new keys (an old bin wins a tie, its id
                // being lower)
                if (nb_now > Nb) {
                    for (int k = Nb >> 6; k <= (nb_now - 1) >> 6; ++k) {
                        if (k == k0) continue;
                        const bool in = (nbi >> 6) == k && nbi != 0x7FFFFFFF;
                        const double mv = mz_wave_max(in ? nkey : -2.0);
                        const int mi = mz_wave_min_i32(in && nkey == mv ? nbi : 0x7FFFFFFF);
                        const bool had = (k << 6) < Nb;
                        if (lane == 0 && (!had || mv > W.bmv[k])) {
                            W.bmv[k] = mv;
                            W.bmi[k] = mi;
                        }
                    }
                }
                oct_wave_sync();
            }
            if (stop_s) break;
        }
    }
    __syncthreads();
    if (S.defer_s) {
        if (threadIdx.x == 0) {
            if (!BIG) {
                oct_defer(A, c, N);
            } else {
                A.old_pia[c * 5 + 4] = -1;
                atomicOr(&A.sp.ctr->err, MZ_ERR_MERGE_SIZE);
            }
        }
        return;
    }
    const int Nbins = Nbins_s;
    if (Nbins == 1 && threadIdx.x == 0) {
        double s = 0.0;
        const int par0 = bins[0].depth & 1;
        for (int i = 0; i < N; ++i) s += A.sp.w[oct_slot(oct_id<IdxT, BIG>(W, pis, par0, i), s1, n1, s2)];
        bins[0].w = s;
    }
    __syncthreads();

    // compute_bin_props! (:646-700) and compute_new_particles! (:736-813,
    // 830-933).  The reference walks the bins in order; here a scan over the
    // bins gives each its output slot and its place in the cell's draw
    // sequence (six sign draws per bin of more than two particles), and the
    // waves take the bins 64 at a time, a lane per bin.  A bin of more than 64
    // particles is summed by its whole wave as a 256-wide block reduction
    // (lane l holds the block threads l, l + 64, l + 128, l + 192), so the
    // values do not depend on which wave runs the bin.
    int32_t *const bslot = W.bslot;
    int32_t *const bdraw = W.bdraw;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // each bin's output slots (2, or its particles when it has at most 2)
    // and draw offset (6 sign draws per bin of more than 2 particles): an
    // exclusive scan over the bins, two bins per thread and pass, slots in
    // the low and draws in the high 32 bits
    {
        int64_t carry = 0;
        for (int b0 = 0; b0 < Nbins; b0 += 2 * NW * 64) {
            int64_t oc[2], dr[2];
            for (int j = 0; j < 2; ++j) {
                const int b = b0 + 2 * threadIdx.x + j;
                const int32_t bnp = b < Nbins ? (bins[b].w == 0.0 ? 0 : (int32_t)bins[b].np) : 0;
                oc[j] = bnp <= 0 ? 0 : (bnp <= 2 ? bnp : 2);
                dr[j] = bnp > 2 ? 1 : 0;
            }
            const int64_t v = (oc[0] + oc[1]) | ((dr[0] + dr[1]) << 32);
            int64_t x = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int64_t y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            if (lane == 63) S.wsc[wv] = x;
            __syncthreads();
            int64_t pre = carry, tot = carry;
            for (int k = 0; k < NW; ++k) {
                if (k < wv) pre += S.wsc[k];
                tot += S.wsc[k];
            }
            const int64_t ex = pre + x - v;
            const int b = b0 + 2 * threadIdx.x;
            if (b < Nbins) {
                bslot[b] = (int32_t)(ex & 0xFFFFFFFF);
                bdraw[b] = (int32_t)(ex >> 32);
                if (b + 1 < Nbins) {
                    bslot[b + 1] = (int32_t)(ex & 0xFFFFFFFF) + (int32_t)oc[0];
                    bdraw[b + 1] = (int32_t)(ex >> 32) + (int32_t)dr[0];
                }
            }
            carry = tot;
            __syncthreads();
        }
        if (threadIdx.x == 0) S.n_new_s = (int32_t)(carry & 0xFFFFFFFF);
        __syncthreads();
    }
    // one bin of more than two particles on the whole wavefront: one particle
    // per lane up to 64, else the 256-wide block layout
    auto wave_bin = [&](const int b) {
        const int bs = bins[b].start, be = bins[b].end, par = bins[b].depth & 1;
        const double Wb = bins[b].w, inv_w = 1.0 / Wb;
        double vm[3], xm[3], sdv[3], sdx[3];
        if (be - bs < 64) {
            // one particle per lane
            const bool on = bs + lane <= be;
            double w = 0.0, u0 = 0.0, u1 = 0.0, u2 = 0.0, y0 = 0.0, y1 = 0.0, y2 = 0.0;
            if (on) {
                const int64_t p = oct_slot(oct_id<IdxT, BIG>(W, pis, par, bs + lane), s1, n1, s2);
                oct_ld<IdxT, BIG>(A, W, pis, s1, n1, s2, par, bs + lane, w, u0, u1, u2);
                y0 = A.sp.x0[p]; y1 = A.sp.x1[p]; y2 = A.sp.x2[p];
            }
            vm[0] = mz_wave_sum(on ? w * u0 : 0.0) * inv_w;
            vm[1] = mz_wave_sum(on ? w * u1 : 0.0) * inv_w;
            vm[2] = mz_wave_sum(on ? w * u2 : 0.0) * inv_w;
            xm[0] = mz_wave_sum(on ? w * y0 : 0.0) * inv_w;
            xm[1] = mz_wave_sum(on ? w * y1 : 0.0) * inv_w;
            xm[2] = mz_wave_sum(on ? w * y2 : 0.0) * inv_w;
            const double a0 = u0 - vm[0], a1 = u1 - vm[1], a2 = u2 - vm[2];
            const double e0 = y0 - xm[0], e1 = y1 - xm[1], e2 = y2 - xm[2];
            sdv[0] = sqrt(mz_wave_sum(on ? w * (a0 * a0) : 0.0) * inv_w);
            sdv[1] = sqrt(mz_wave_sum(on ? w * (a1 * a1) : 0.0) * inv_w);
            sdv[2] = sqrt(mz_wave_sum(on ? w * (a2 * a2) : 0.0) * inv_w);
            sdx[0] = sqrt(mz_wave_sum(on ? w * (e0 * e0) : 0.0) * inv_w);
            sdx[1] = sqrt(mz_wave_sum(on ? w * (e1 * e1) : 0.0) * inv_w);
            sdx[2] = sqrt(mz_wave_sum(on ? w * (e2 * e2) : 0.0) * inv_w);
            if (lane == 0) oct_emit(A, c, bdraw[b], bslot[b], Wb, vm, xm, sdv, sdx, s1, n1, s2);
            return;
        }
        // lane l stands for the block threads l + 64 q, q = 0..3
        double sv[4][3], sx[4][3];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int d = 0; d < 3; ++d) sv[q][d] = sx[q][d] = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            for (int i = bs + lane + 64 * q; i <= be; i += 256) {
                const int64_t p = oct_slot(oct_id<IdxT, BIG>(W, pis, par, i), s1, n1, s2);
                double w, u0, u1, u2;
                oct_ld<IdxT, BIG>(A, W, pis, s1, n1, s2, par, i, w, u0, u1, u2);
                sv[q][0] = fma(w, u0, sv[q][0]); sv[q][1] = fma(w, u1, sv[q][1]);
                sv[q][2] = fma(w, u2, sv[q][2]);
                sx[q][0] = fma(w, A.sp.x0[p], sx[q][0]); sx[q][1] = fma(w, A.sp.x1[p], sx[q][1]);
                sx[q][2] = fma(w, A.sp.x2[p], sx[q][2]);
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            vm[d] = ((mz_wave_sum(sv[0][d]) + mz_wave_sum(sv[1][d])) +
                     (mz_wave_sum(sv[2][d]) + mz_wave_sum(sv[3][d]))) * inv_w;
            xm[d] = ((mz_wave_sum(sx[0][d]) + mz_wave_sum(sx[1][d])) +
                     (mz_wave_sum(sx[2][d]) + mz_wave_sum(sx[3][d]))) * inv_w;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int d = 0; d < 3; ++d) sv[q][d] = sx[q][d] = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            for (int i = bs + lane + 64 * q; i <= be; i += 256) {
                const int64_t p = oct_slot(oct_id<IdxT, BIG>(W, pis, par, i), s1, n1, s2);
                double w, u0, u1, u2;
                oct_ld<IdxT, BIG>(A, W, pis, s1, n1, s2, par, i, w, u0, u1, u2);
                const double a0 = u0 - vm[0], a1 = u1 - vm[1], a2 = u2 - vm[2];
                const double e0 = A.sp.x0[p] - xm[0], e1 = A.sp.x1[p] - xm[1], e2 = A.sp.x2[p] - xm[2];
                sv[q][0] = fma(w, a0 * a0, sv[q][0]); sv[q][1] = fma(w, a1 * a1, sv[q][1]);
                sv[q][2] = fma(w, a2 * a2, sv[q][2]);
                sx[q][0] = fma(w, e0 * e0, sx[q][0]); sx[q][1] = fma(w, e1 * e1, sx[q][1]);
                sx[q][2] = fma(w, e2 * e2, sx[q][2]);
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            sdv[d] = sqrt(((mz_wave_sum(sv[0][d]) + mz_wave_sum(sv[1][d])) +
                           (mz_wave_sum(sv[2][d]) + mz_wave_sum(sv[3][d]))) * inv_w);
            sdx[d] = sqrt(((mz_wave_sum(sx[0][d]) + mz_wave_sum(sx[1][d])) +
                           (mz_wave_sum(sx[2][d]) + mz_wave_sum(sx[3][d]))) * inv_w);
        }
        if (lane == 0) oct_emit(A, c, bdraw[b], bslot[b], Wb, vm, xm, sdv, sdx, s1, n1, s2);
    };
    // compute_bin_props! and compute_new_particles! per bin.  With at least 32
    // bins per wave a bin of at most OCT_LANE_NP particles takes one lane, its
    // sums in particle order (the reference's order), and a larger bin the whole
    // wavefront; with fewer bins the waves take the bins in turn.
    if (Nbins < 32 * NW) {
        for (int b = wv; b < Nbins; b += NW) {
            const int32_t bnp = bins[b].w == 0.0 ? 0 : (int32_t)bins[b].np;
            if (bnp <= 0) continue;
            if (bnp <= 2) {
                const int bs = bins[b].start, par = bins[b].depth & 1;
                if (lane < bnp)
                    oct_copy(A, oct_slot(oct_id<IdxT, BIG>(W, pis, par, bs + lane), s1, n1, s2),
                             oct_slot(bslot[b] + lane, s1, n1, s2));
                continue;
            }
            wave_bin(b);
        }
    }
    for (int b0 = wv * 64; Nbins >= 32 * NW && b0 < Nbins; b0 += NW * 64) {
        const int bl = b0 + lane;
        int32_t bnp = 0;
        int bs = 0, par = 0;
        double Wb = 0.0;
        if (bl < Nbins) {
            Wb = bins[bl].w;
            bnp = Wb == 0.0 ? 0 : (int32_t)bins[bl].np;
            bs = bins[bl].start;
            par = bins[bl].depth & 1;
        }
        const bool lone = bnp > 0 && bnp <= OCT_LANE_NP;
        const int slot = lone ? bslot[bl] : 0;
        if (lone && bnp <= 2)
            for (int j = 0; j < bnp; ++j)
                oct_copy(A, oct_slot(oct_id<IdxT, BIG>(W, pis, par, bs + j), s1, n1, s2), oct_slot(slot + j, s1, n1, s2));
        const bool sums = lone && bnp > 2;
        const int jn = (int)mz_wave_max(sums ? (double)bnp : 0.0);
        if (jn > 0) {
            double vm[3] = {0.0, 0.0, 0.0}, xm[3] = {0.0, 0.0, 0.0};
            for (int j = 0; j < jn; ++j) {
                if (sums && j < bnp) {
                    const int64_t p = oct_slot(oct_id<IdxT, BIG>(W, pis, par, bs + j), s1, n1, s2);
                    double w, a0, a1, a2;
                    oct_ld<IdxT, BIG>(A, W, pis, s1, n1, s2, par, bs + j, w, a0, a1, a2);
                    vm[0] = vm[0] + w * a0; vm[1] = vm[1] + w * a1; vm[2] = vm[2] + w * a2;
                    xm[0] = xm[0] + w * A.sp.x0[p]; xm[1] = xm[1] + w * A.sp.x1[p]; xm[2] = xm[2] + w * A.sp.x2[p];
                }
            }
            const double inv_w = 1.0 / (sums ? Wb : 1.0);
            for (int d = 0; d < 3; ++d) { vm[d] = vm[d] * inv_w; xm[d] = xm[d] * inv_w; }
            double vs[3] = {0.0, 0.0, 0.0}, xs[3] = {0.0, 0.0, 0.0};
            for (int j = 0; j < jn; ++j) {
                if (sums && j < bnp) {
                    const int64_t p = oct_slot(oct_id<IdxT, BIG>(W, pis, par, bs + j), s1, n1, s2);
                    double w, u0, u1, u2;
                    oct_ld<IdxT, BIG>(A, W, pis, s1, n1, s2, par, bs + j, w, u0, u1, u2);
                    const double a0 = u0 - vm[0], a1 = u1 - vm[1], a2 = u2 - vm[2];
                    const double e0 = A.sp.x0[p] - xm[0], e1 = A.sp.x1[p] - xm[1], e2 = A.sp.x2[p] - xm[2];
                    vs[0] = vs[0] + w * (a0 * a0); vs[1] = vs[1] + w * (a1 * a1); vs[2] = vs[2] + w * (a2 * a2);
                    xs[0] = xs[0] + w * (e0 * e0); xs[1] = xs[1] + w * (e1 * e1); xs[2] = xs[2] + w * (e2 * e2);
                }
            }
            if (sums) {
                double sdv[3], sdx[3];
                for (int d = 0; d < 3; ++d) {
                    sdv[d] = sqrt(vs[d] * inv_w);
                    sdx[d] = sqrt(xs[d] * inv_w);
                }
                oct_emit(A, c, bdraw[bl], slot, Wb, vm, xm, sdv, sdx, s1, n1, s2);
            }
        }
        // larger bins, one at a time on the wavefront
        uint64_t bigm = __ballot(bnp > OCT_LANE_NP);
        while (bigm) {
            const int b = b0 + (int)__builtin_ctzll(bigm);
            bigm &= bigm - 1;
            wave_bin(b);
        }
    }
    __syncthreads();
    merge_finish(A.sp, A.tmp, c, N, S.n_new_s, s1, n1, s2, n2, A.old_pia, A.n_merged, A.n_cells);
}

// The cell's threshold and target (per cell when the caller gives them).
__device__ inline bool oct_pick(const OctArgs &A, int c, int32_t &N, int32_t &target) {
    N = A.sp.pia.nloc[c];
    const int32_t thr = A.thresholds ? A.thresholds[c] : A.threshold;
    target = A.targets ? A.targets[c] : A.target_np;
    return !(N <= thr || (A.mask && !A.mask[c]));
}

// One workgroup per cell, the working arrays in LDS.
__global__ void __launch_bounds__(256) k_octree_merge(OctArgs A) {
    __shared__ int16_t pis[OCT_MAX_N];
    __shared__ __align__(16) int16_t psort[OCT_MAX_N];   // after the splits: per-bin slots and draws
    __shared__ uint8_t poct[OCT_MAX_N];
    __shared__ OctBin bins[OCT_MAX_BINS];
    __shared__ double bkey[OCT_MAX_BINS];
    __shared__ OctShared S;
    const int c = blockIdx.x;
    A.rp = mz_rng_resolve(A.rp);
    int32_t N, target;
    if (!oct_pick(A, c, N, target)) {
        if (threadIdx.x == 0) A.old_pia[c * 5 + 4] = -1;
        return;
    }
    if (N > OCT_MAX_N) {
        if (threadIdx.x == 0) oct_defer(A, c, N);
        return;
    }
    OctWork<int16_t> W;
    W.pis = pis; W.psort = psort; W.poct = poct; W.bins = bins; W.bkey = bkey;
    W.bslot = reinterpret_cast<int32_t *>(psort);
    W.bdraw = W.bslot + OCT_MAX_BINS;
    W.bins_cap = OCT_MAX_BINS;
    W.bmv = nullptr;
    W.bmi = nullptr;
    oct_merge_cell<int16_t, 4, false>(A, c, N, target, W, S);
}

// The listed cells, one wavefront per workgroup, the working arrays in the
// workgroup's slice of device memory.  The workgroups take cells from a queue,
// the largest size class first, so the longest merges start first.
__global__ void __launch_bounds__(64) k_octree_merge_big(OctArgs A) {
    __shared__ OctShared S;
    // per 64-bin block: its largest key (nblk doubles), then its lowest bin
    // holding that key (nblk int32s)
    extern __shared__ double s_blk[];
    A.rp = mz_rng_resolve(A.rp);
    const int32_t n_list = A.big_info[0];
    const size_t g = blockIdx.x;
    const int nblk = (A.big_bcap + 63) >> 6;
    OctWork<int32_t> W;
    W.pis = A.big_pis + g * (size_t)A.big_ncap;
    W.psort = A.big_psort + g * (size_t)A.big_ncap;
    W.sid0 = W.pis;
    W.sid_q = W.psort - W.pis;
    W.stg = A.big_stage + g * (size_t)A.big_ncap * 8;
    W.n_stg = A.big_ncap;
    W.poct = A.big_poct + g * (size_t)A.big_ncap;
    W.bins = reinterpret_cast<OctBinT<int32_t> *>(A.big_bins) + g * (size_t)A.big_bcap;
    W.bkey = A.big_bkey + g * (size_t)A.big_bcap;
    W.bslot = A.big_bslot + g * (size_t)A.big_bcap * 2;
    W.bdraw = W.bslot + A.big_bcap;
    W.bins_cap = A.big_bcap;
    W.bmv = A.big_blocks ? s_blk : nullptr;
    W.bmi = A.big_blocks ? reinterpret_cast<int32_t *>(s_blk + nblk) : nullptr;
    for (;;) {
        if (threadIdx.x == 0) S.ticket_s = atomicAdd(&A.big_info[6], 1);
        __syncthreads();
        int32_t t = S.ticket_s;
        __syncthreads();
        if (t >= n_list) break;
        int cls = 0;
        while (cls < 3 && t >= A.big_info[2 + cls]) t -= A.big_info[2 + cls++];
        const int c = A.big_list[(size_t)cls * A.n_cells + t];
        int32_t N, target;
        oct_pick(A, c, N, target);
        if (N <= A.big_ncap) oct_merge_cell<int32_t, 1, true>(A, c, N, target, W, S);
        else if (threadIdx.x == 0) atomicOr(&A.sp.ctr->err, MZ_ERR_MERGE_SIZE);
        __syncthreads();
    }
}


// Free-id pushes in cell order: cell c pushes the ids of its deleted
// positions, last first (delete_particle_end! pushes pv.index[end]).
__global__ void k_merge_push_ids(MzSpView sp, const int32_t *__restrict__ old_pia, const int64_t *__restrict__ off,
                                 const int64_t *__restrict__ total, int32_t nbuf0, int32_t ntot0, int32_t n_cells) {
    const int c = blockIdx.x;
    if (c < n_cells) {
        const int32_t *op = old_pia + c * 5;
        const int32_t n_new = op[4];
        if (n_new >= 0) {
            const int32_t N = op[1] + op[3];
            const int32_t n_del = N - n_new;
            for (int j = threadIdx.x; j < n_del; j += blockDim.x) {
                const int li = N - 1 - j;
                const int64_t p = oct_slot(li, op[0], op[1], op[2]);
                const int64_t q = (int64_t)nbuf0 + off[c] + j;
                if (q < sp.bufcap) sp.buffer[q] = sp.id[p];
                else atomicOr(&sp.ctr->err, MZ_ERR_BUFFER);
            }
        }
    }
    if (c == 0 && threadIdx.x == 0) {
        sp.ctr->nbuffer = nbuf0 + (int32_t)*total;
        sp.ctr->n_total = ntot0 - (int32_t)*total;
        sp.ctr->deleted = (int32_t)*total;
    }
}

__global__ void k_merge_ndel(const int32_t *__restrict__ old_pia, int32_t n_cells, int32_t *__restrict__ ndel) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n_cells) {
        const int32_t *op = old_pia + c * 5;
        ndel[c] = op[4] >= 0 ? op[1] + op[3] - op[4] : 0;
    }
}

__global__ void __launch_bounds__(1024) k_scan_i32m(const int32_t *__restrict__ in, int64_t n, int64_t *__restrict__ out,
                                                    int64_t *__restrict__ total) {
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

static int octree_merge_launch(mz_store *h, int32_t species, int32_t threshold, const uint8_t *mask,
                               int32_t target_np, int32_t split, int32_t init_bounds, int32_t bounds_mode,
                               int32_t max_Nbins, int32_t max_depth, int32_t grid_clamp, uint64_t tag,
                               int32_t rng_mode, int64_t *n_merged, const int32_t *d_thresholds = nullptr,
                               const int32_t *d_targets = nullptr);

extern "C" int mz_merge_octree(mz_store *h, int32_t species, int32_t threshold, int32_t target_np, int32_t split,
                               int32_t init_bounds, int32_t bounds_mode, int32_t max_Nbins, int32_t max_depth,
                               int32_t grid_clamp, int32_t squash, uint64_t tag, int32_t rng_mode,
                               int64_t *n_merged) {
    int r = octree_merge_launch(h, species, threshold, nullptr, target_np, split, init_bounds, bounds_mode, max_Nbins,
                                max_depth, grid_clamp, tag, rng_mode, n_merged);
    if (r) return r;
    if (squash) return mz_squash(h, species);
    return 0;
}

// The caller's loop "merge cell c when its n_local exceeds thresholds[c],
// down to targets[c]" (per-cell merge_octree_N2_based!(rng, octree, pv, pia,
// cell, species, target_np) calls, merging_octree_N2.jl:1060-1094) as one
// call; thresholds and targets are host arrays of n_cells entries.
extern "C" int mz_merge_octree_targets(mz_store *h, int32_t species, const int32_t *thresholds,
                                       const int32_t *targets, int32_t split, int32_t init_bounds,
                                       int32_t bounds_mode, int32_t max_Nbins, int32_t max_depth, int32_t grid_clamp,
                                       int32_t squash, uint64_t tag, int32_t rng_mode, int64_t *n_merged) {
    MZ_CHECK(thresholds != nullptr && targets != nullptr, "merge_octree_targets: thresholds and targets needed");
    const int nc = h->n_cells;
    int32_t *d = (int32_t *)mz_scratch(h, "oct_targets", (size_t)nc * 2 * sizeof(int32_t));
    MZ_CHECK(d != nullptr, "merge_octree_targets: scratch allocation failed");
    MZ_HIP(hipMemcpyAsync(d, thresholds, (size_t)nc * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    MZ_HIP(hipMemcpyAsync(d + nc, targets, (size_t)nc * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    int32_t tmax = 2;
    for (int c = 0; c < nc; ++c) tmax = std::max(tmax, targets[c]);
    int r = octree_merge_launch(h, species, 0, nullptr, tmax, split, init_bounds, bounds_mode, max_Nbins, max_depth,
                                grid_clamp, tag, rng_mode, n_merged, d, d + nc);
    if (r) return r;
    if (squash) return mz_squash(h, species);
    return 0;
}

// The caller's loop of per-cell merge_octree_N2_based! calls
// (test/test_1D_couette_varweight.jl:78-87) as one call: cells[k] (1-based)
// in list order.  Each run of ascending cells is one launch over a cell
// mask; the runs go one after another, so freed ids reach the buffer in the
// calls' order.  n_merged adds the runs' counts.
extern "C" int mz_merge_octree_cells(mz_store *h, int32_t species, const int32_t *cells, int32_t n, int32_t target_np,
                                     int32_t split, int32_t init_bounds, int32_t bounds_mode, int32_t max_Nbins,
                                     int32_t max_depth, int32_t grid_clamp, uint64_t tag, int32_t rng_mode,
                                     int64_t *n_merged) {
    MZ_CHECK(n >= 0 && (n == 0 || cells != nullptr), "merge_octree_cells: bad cell list");
    for (int32_t k = 0; k < n; ++k)
        MZ_CHECK(cells[k] >= 1 && cells[k] <= h->n_cells, "merge_octree_cells: cell out of range");
    if (n_merged) *n_merged = 0;
    uint8_t *mask = (uint8_t *)mz_scratch(h, "oct_mask", (size_t)h->n_cells);
    MZ_CHECK(mask != nullptr, "merge_octree_cells: scratch allocation failed");
    std::vector<uint8_t> hm(h->n_cells);
    int32_t k0 = 0;
    while (k0 < n) {
        int32_t k1 = k0 + 1;
        while (k1 < n && cells[k1] > cells[k1 - 1]) ++k1;
        std::fill(hm.begin(), hm.end(), 0);
        for (int32_t k = k0; k < k1; ++k) hm[cells[k] - 1] = 1;
        MZ_HIP(hipStreamSynchronize(h->stream));
        MZ_HIP(hipMemcpy(mask, hm.data(), h->n_cells, hipMemcpyHostToDevice));
        int64_t cnt = 0;
        int r = octree_merge_launch(h, species, -1, mask, target_np, split, init_bounds, bounds_mode, max_Nbins,
                                    max_depth, grid_clamp, tag, rng_mode, n_merged ? &cnt : nullptr);
        if (r) return r;
        if (n_merged) *n_merged += cnt;
        k0 = k1;
    }
    return 0;
}

static int octree_merge_launch(mz_store *h, int32_t species, int32_t threshold, const uint8_t *mask,
                               int32_t target_np, int32_t split, int32_t init_bounds, int32_t bounds_mode,
                               int32_t max_Nbins, int32_t max_depth, int32_t grid_clamp, uint64_t tag,
                               int32_t rng_mode, int64_t *n_merged, const int32_t *d_thresholds,
                               const int32_t *d_targets) {
    if (species >= 1 && species <= h->n_species) h->sp[species - 1].contig_known = false;
    MZ_CHECK(species >= 1 && species <= h->n_species, "merge_octree: bad species");
    MZ_CHECK(split == 1 || split == 2, "merge_octree: split must be 1 (mid) or 2 (mean)");
    MZ_CHECK(init_bounds >= 1 && init_bounds <= 3, "merge_octree: bad init_bounds");
    MZ_CHECK(bounds_mode == 1 || bounds_mode == 2, "merge_octree: bad bounds_mode");
    MZ_CHECK(max_Nbins >= 8, "merge_octree: max_Nbins must be at least 8");
    MZ_CHECK(!grid_clamp || h->has_grid, "merge_octree: grid clamp needs a grid");
    MZ_CHECK(rng_mode == 0 || (rng_mode == 1 && h->d_replay[2]), "merge_octree: replay table not set");
    MzSpecies &s = h->sp[species - 1];
    const int nc = h->n_cells;
    int32_t *old_pia = (int32_t *)mz_scratch(h, "oct_oldpia", (size_t)nc * 5 * sizeof(int32_t));
    int32_t *cnt = (int32_t *)mz_scratch(h, "oct_cnt", 8 * sizeof(int32_t));
    int32_t *ndel = (int32_t *)mz_scratch(h, "oct_ndel", (size_t)nc * sizeof(int32_t));
    int64_t *off = (int64_t *)mz_scratch(h, "oct_off", (size_t)(nc + 1) * sizeof(int64_t));
    int32_t *big_list = (int32_t *)mz_scratch(h, "oct_big_list", (size_t)nc * 4 * sizeof(int32_t));
    MZ_CHECK(old_pia && cnt && ndel && off && big_list, "merge_octree: scratch allocation failed");
    MzCounters snap;
    int r = mz_read_counters(h, species, &snap);
    if (r) return r;
    // cnt[0] counts merges; cnt[1..7] are the big-cell lists' counts and queue
    MZ_HIP(hipMemsetAsync(cnt, 0, 8 * sizeof(int32_t), h->stream));
    OctArgs A{};
    A.sp = mz_view(s);
    A.tmp = A.sp;
    A.tmp.x0 = s.xb[0]; A.tmp.x1 = s.xb[1]; A.tmp.x2 = s.xb[2];
    A.tmp.v0 = s.vb[0]; A.tmp.v1 = s.vb[1]; A.tmp.v2 = s.vb[2];
    A.tmp.w = s.wb;
    A.threshold = threshold; A.target_np = target_np; A.split = split; A.init_bounds = init_bounds;
    A.bounds_mode = bounds_mode; A.max_Nbins = max_Nbins; A.max_depth = max_depth; A.grid_clamp = grid_clamp;
    A.min_x = h->grid.min_x; A.max_x = h->grid.max_x;
    A.rp = mz_rng_params(h, 2, tag, rng_mode);
    A.old_pia = old_pia;
    A.n_merged = cnt;
    A.n_cells = nc;
    A.mask = mask;
    A.thresholds = d_thresholds;
    A.targets = d_targets;
    A.big_list = big_list;
    A.big_info = cnt + 1;
    {
        MzTimed tm(h, "merge_octree");
        hipLaunchKernelGGL(k_octree_merge, dim3(nc), dim3(256), 0, h->stream, A);
        // cells above the LDS kernel's sizes: one more launch with their
        // working arrays in device memory, sized by what the list needs
        int32_t info[2] = {0, 0};
        MZ_HIP(hipMemcpyAsync(info, cnt + 1, sizeof(info), hipMemcpyDeviceToHost, h->stream));
        MZ_HIP(hipStreamSynchronize(h->stream));
        if (info[0] > 0) {
            // a cell's bins hold at least one particle each and count at least
            // one each towards the post-merge count, which stays at most the
            // target: at most min(max_Nbins, N, target) bins (target_np is the
            // largest per-cell target when there are several)
            const int64_t ncap = ((int64_t)info[1] + 255) / 256 * 256;
            const int64_t bcap = std::min<int64_t>({(int64_t)max_Nbins, ncap, (int64_t)std::max(target_np, 2)}) + 8;
            const int64_t per = ncap * (2 * sizeof(int32_t) + 8 * sizeof(double) + 1) +
                                bcap * (sizeof(OctBinT<int32_t>) + 3 * sizeof(double));
            // the blocks' maxima in LDS: 12 bytes per 64-bin block
            const int64_t nblk = (bcap + 63) / 64;
            const bool blocks = nblk <= OCT_BIG_BLOCKS;
            const size_t dyn = blocks ? (size_t)((nblk * 12 + 15) / 16 * 16) : 0;
            // workgroups: as many as can be resident (the loop is latency-bound,
            // so every resident wave helps), within a quarter of the free memory
            // (at least 2 GiB, at most 16 GiB)
            size_t mfree = 0, mtot = 0;
            MZ_HIP(hipMemGetInfo(&mfree, &mtot));
            const int64_t budget = std::min<int64_t>((int64_t)16 << 30, std::max<int64_t>((int64_t)2 << 30,
                                                                                          (int64_t)(mfree / 4)));
            int dev = 0, n_cu = 0, per_cu = 0;
            MZ_HIP(hipGetDevice(&dev));
            MZ_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
            MZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_octree_merge_big, 64, dyn));
            const int64_t G = std::max<int64_t>(1, std::min<int64_t>({(int64_t)info[0],
                                                                      (int64_t)std::max(per_cu, 1) * n_cu,
                                                                      budget / per}));
            int32_t *ib = (int32_t *)mz_scratch(h, "oct_big_idx", (size_t)(G * ncap * 2 * sizeof(int32_t)));
            uint8_t *ob = (uint8_t *)mz_scratch(h, "oct_big_oct", (size_t)(G * ncap));
            double *stg = (double *)mz_scratch(h, "oct_big_stage", (size_t)(G * ncap * 8 * sizeof(double)));
            void *bb = mz_scratch(h, "oct_big_bins", (size_t)(G * bcap * sizeof(OctBinT<int32_t>)));
            double *kb = (double *)mz_scratch(h, "oct_big_key", (size_t)(G * bcap * sizeof(double)));
            int32_t *sb = (int32_t *)mz_scratch(h, "oct_big_slot", (size_t)(G * bcap * 2 * sizeof(int32_t)));
            MZ_CHECK(ib && ob && stg && bb && kb && sb, "merge_octree: scratch allocation for large cells failed");
            A.big_pis = ib;
            A.big_psort = ib + G * ncap;
            A.big_poct = ob;
            A.big_stage = stg;
            A.big_bins = bb;
            A.big_bkey = kb;
            A.big_bslot = sb;
            A.big_ncap = (int32_t)ncap;
            A.big_bcap = (int32_t)bcap;
            A.big_blocks = blocks ? 1 : 0;
            hipLaunchKernelGGL(k_octree_merge_big, dim3((unsigned)G), dim3(64), dyn, h->stream, A);
        }
        hipLaunchKernelGGL(k_merge_ndel, dim3((nc + 255) / 256), dim3(256), 0, h->stream, old_pia, nc, ndel);
        hipLaunchKernelGGL(k_scan_i32m, dim3(1), dim3(1024), 0, h->stream, ndel, (int64_t)nc, off, off + nc);
        hipLaunchKernelGGL(k_merge_push_ids, dim3(nc), dim3(64), 0, h->stream, A.sp, old_pia, off, off + nc,
                           snap.nbuffer, snap.n_total, nc);
    }
    MZ_HIP(hipGetLastError());
    if (n_merged) {
        int32_t v = 0;
        MZ_HIP(hipMemcpyAsync(&v, cnt, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
        MZ_HIP(hipStreamSynchronize(h->stream));
        *n_merged = v;
    }
    return 0;
}

// ===========================================================================
// Grid N:2 merging (merging_grid.jl:190-703), one workgroup per cell above
// the threshold.  The reference walks the cell's particles (group 1, then
// group 2) accumulating into the velocity cells, then walks the velocity
// cells in order.  Here the particles' velocity-cell indices are counted in
// LDS, a stable ranking lists each velocity cell's particles in particle
// order, and one thread per velocity cell forms its sums in that order
// (x + v * w accumulations, then / w, unfused as the reference writes them),
// so the values equal the oracle's bit for bit.  A scan over the velocity
// cells gives each its output slots and its place in the cell's draw
// sequence (six sign draws per velocity cell of more than two particles).

#define GM_MAX_N 2048
#define GM_MAX_CELLS 4104   // Nx Ny Nz + 8

struct GridArgs {
    MzSpView sp, tmp;
    int32_t threshold, Nx, Ny, Nz, extent_mode, grid_clamp;
    double mult[3], ext_lo[3], ext_hi[3];
    const double *T, *v;      // extent_mode 0: this species' props, [cell] and [cell][3]
    double mass, min_x, max_x;
    MzRngParams rp;
    int32_t *old_pia;
    int32_t *n_merged;
    int32_t n_cells;
};

__device__ inline int64_t gm_coord(double f, int32_t n) {
    const double fl = floor(f);
    return fl >= 0.0 ? (fl <= (double)n ? (int64_t)fl : (int64_t)n) : 0;
}

// block-wide exclusive scan of a[0..n) in place (contiguous chunks per
// thread); returns the total to every thread
__device__ int32_t gm_scan(int32_t *a, int n, int32_t *wsc) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int per = (n + 255) / 256;
    const int lo = t * per, hi = min(n, lo + per);
    int32_t loc = 0;
    for (int i = lo; i < hi; ++i) loc += a[i];
    int32_t x = loc;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsc[wv] = x;
    __syncthreads();
    int32_t pre = 0;
    for (int k = 0; k < wv; ++k) pre += wsc[k];
    const int32_t total = wsc[0] + wsc[1] + wsc[2] + wsc[3];
    int32_t run = pre + x - loc;
    for (int i = lo; i < hi; ++i) {
        const int32_t v = a[i];
        a[i] = run;
        run += v;
    }
    __syncthreads();
    return total;
}

__global__ void __launch_bounds__(256) k_grid_merge(GridArgs A) {
    __shared__ uint32_t sk[GM_MAX_N];      // (velocity cell << 11) | li, sorted into velocity-cell order
    __shared__ int16_t ord[GM_MAX_N];      // logical particles grouped by velocity cell, in particle order
    __shared__ int32_t gstart[GM_MAX_CELLS + 1];
    __shared__ int32_t gaux[GM_MAX_CELLS + 1];   // counts, then running slots, then packed outputs/draws
    __shared__ int32_t wsc[4];
    __shared__ int32_t n_new_s;
    const int c = blockIdx.x;
    const int32_t N = A.sp.pia.nloc[c];
    if (N <= A.threshold) {
        if (threadIdx.x == 0) A.old_pia[c * 5 + 4] = -1;
        return;
    }
    if (N > GM_MAX_N) {
        if (threadIdx.x == 0) {
            A.old_pia[c * 5 + 4] = -1;
            atomicOr(&A.sp.ctr->err, MZ_ERR_MERGE_SIZE);
        }
        return;
    }
    A.rp = mz_rng_resolve(A.rp);
    const int32_t s1 = A.sp.pia.start1[c], n1 = A.sp.pia.n1[c], s2 = A.sp.pia.start2[c], n2 = A.sp.pia.n2[c];
    const int Nx = A.Nx, Ny = A.Ny, Nz = A.Nz, NyNz = Ny * Nz, Nt = Nx * Ny * Nz + 8;
    // compute_velocity_extent! (:190-220)
    double lo[3], up[3], mid[3], dinv[3];
    {
        const int32_t Nd[3] = {Nx, Ny, Nz};
        if (A.extent_mode == 0) {
            const double vt = sqrt(2 * A.T[c] * MZ_KB / A.mass);
            for (int d = 0; d < 3; ++d) {
                const double dv = A.mult[d] * vt, vc = A.v[3 * c + d];
                lo[d] = vc - dv;
                up[d] = vc + dv;
                mid[d] = vc;
                dinv[d] = 1.0 / (2 * dv / (double)Nd[d]);
            }
        } else {
            for (int d = 0; d < 3; ++d) {
                lo[d] = A.ext_lo[d];
                up[d] = A.ext_hi[d];
                mid[d] = 0.5 * (A.ext_lo[d] + A.ext_hi[d]);
                dinv[d] = 1.0 / ((A.ext_hi[d] - A.ext_lo[d]) / (double)Nd[d]);
            }
        }
    }
    for (int g = threadIdx.x; g <= Nt; g += 256) gaux[g] = 0;
    __syncthreads();
    // compute_grid_index (:235-270) of every particle
    for (int li = threadIdx.x; li < N; li += 256) {
        const int64_t p = oct_slot(li, s1, n1, s2);
        const double v0 = A.sp.v0[p], v1 = A.sp.v1[p], v2 = A.sp.v2[p];
        const bool outside = v0 < lo[0] || v0 > up[0] || v1 < lo[1] || v1 > up[1] || v2 < lo[2] || v2 > up[2];
        int64_t k;
        if (!outside) {
            k = gm_coord((v0 - lo[0]) * dinv[0], Nx) * NyNz + gm_coord((v1 - lo[1]) * dinv[1], Ny) * Nz +
                gm_coord((v2 - lo[2]) * dinv[2], Nz);
        } else {
            k = Nt - 8 + (v0 > mid[0] ? 1 : 0) + (v1 > mid[1] ? 2 : 0) + (v2 > mid[2] ? 4 : 0);
        }
        if (k > Nt - 1) k = Nt - 1;
        sk[li] = ((uint32_t)k << 11) | (uint32_t)li;
        atomicAdd(&gaux[k], 1);
    }
    __syncthreads();
    for (int g = threadIdx.x; g < Nt; g += 256) gstart[g] = gaux[g];
    __syncthreads();
    gm_scan(gstart, Nt, wsc);
    if (threadIdx.x == 0) gstart[Nt] = N;
    // (velocity cell, particle) keys in ascending order: a bitonic sort over
    // the next power of two, so each velocity cell's particles come out in
    // particle order
    int npad = 1;
    while (npad < N) npad <<= 1;
    for (int li = N + threadIdx.x; li < npad; li += 256) sk[li] = 0xFFFFFFFFu;
    __syncthreads();
    for (int len = 2; len <= npad; len <<= 1) {
        for (int stride = len >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < (npad >> 1); t += 256) {
                const int ia = 2 * t - (t & (stride - 1));
                const int ib = ia + stride;
                const bool asc = (ia & len) == 0;
                const uint32_t ka = sk[ia], kb = sk[ib];
                if ((ka > kb) == asc) { sk[ia] = kb; sk[ib] = ka; }
            }
            __syncthreads();
        }
    }
    for (int j = threadIdx.x; j < N; j += 256) ord[j] = (int16_t)(sk[j] & 0x7FFu);
    __syncthreads();
    // pass A: each velocity cell's weight in particle order; outputs
    // min(np, 2) (none when the weight is not positive), draws 1 if np > 2
    for (int g = threadIdx.x; g < Nt; g += 256) {
        const int b = gstart[g], e = gstart[g + 1];
        double w = 0.0;
        // four loads in flight, then the adds in particle order
        for (int j = b; j < e; j += 4) {
            double wb[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) wb[k] = j + k < e ? A.sp.w[oct_slot(ord[j + k], s1, n1, s2)] : 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (j + k < e) w += wb[k];
        }
        const int np = w > 0.0 ? e - b : 0;
        gaux[g] = (np >= 2 ? 2 : np) | ((np > 2 ? 1 : 0) << 16);
    }
    __syncthreads();
    const int32_t tot = gm_scan(gaux, Nt, wsc);
    if (threadIdx.x == 0) {
        gaux[Nt] = tot;
        n_new_s = tot & 0xFFFF;
    }
    __syncthreads();
    // pass B: compute_grid!'s means and spreads (:304-384) and
    // compute_new_particles! (:393-595), one thread per velocity cell
    for (int g = threadIdx.x; g < Nt; g += 256) {
        const int b = gstart[g], e = gstart[g + 1];
        const int slot = gaux[g] & 0xFFFF, n_out = (gaux[g + 1] & 0xFFFF) - slot;
        if (n_out == 0) continue;
        const int np = e - b;
        double ow[2], ov[2][3], ox[2][3];
        if (np > 2) {
            double w = 0.0, vm[3] = {0, 0, 0}, xm[3] = {0, 0, 0};
            // four particles' loads in flight, then the sums in particle order
            for (int j = b; j < e; j += 4) {
                double pw[4], pv[4][3], px[4][3];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int64_t p = oct_slot(ord[j + k < e ? j + k : b], s1, n1, s2);
                    pw[k] = A.sp.w[p];
                    pv[k][0] = A.sp.v0[p]; pv[k][1] = A.sp.v1[p]; pv[k][2] = A.sp.v2[p];
                    px[k][0] = A.sp.x0[p]; px[k][1] = A.sp.x1[p]; px[k][2] = A.sp.x2[p];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (j + k >= e) break;
                    w += pw[k];
                    for (int d = 0; d < 3; ++d) {
                        vm[d] = vm[d] + pv[k][d] * pw[k];
                        xm[d] = xm[d] + px[k][d] * pw[k];
                    }
                }
            }
            for (int d = 0; d < 3; ++d) { vm[d] = vm[d] / w; xm[d] = xm[d] / w; }
            double vs[3] = {0, 0, 0}, xs[3] = {0, 0, 0};
            for (int j = b; j < e; j += 4) {
                double pw[4], pv[4][3], px[4][3];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int64_t p = oct_slot(ord[j + k < e ? j + k : b], s1, n1, s2);
                    pw[k] = A.sp.w[p];
                    pv[k][0] = A.sp.v0[p]; pv[k][1] = A.sp.v1[p]; pv[k][2] = A.sp.v2[p];
                    px[k][0] = A.sp.x0[p]; px[k][1] = A.sp.x1[p]; px[k][2] = A.sp.x2[p];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (j + k >= e) break;
                    for (int d = 0; d < 3; ++d) {
                        const double a = pv[k][d] - vm[d], c = px[k][d] - xm[d];
                        vs[d] = vs[d] + (a * a) * pw[k];
                        xs[d] = xs[d] + (c * c) * pw[k];
                    }
                }
            }
            for (int d = 0; d < 3; ++d) { vs[d] = sqrt(vs[d] / w); xs[d] = sqrt(xs[d] / w); }
            MzStream st = mz_stream((uint64_t)c);
            st.k = 6 * (uint64_t)(gaux[g] >> 16);
            double dir[3];
            for (int d = 0; d < 3; ++d) dir[d] = mz_rand(A.rp, st) < 0.5 ? -1.0 : 1.0;
            for (int d = 0; d < 3; ++d) { ov[0][d] = vm[d] + dir[d] * vs[d]; ov[1][d] = vm[d] - dir[d] * vs[d]; }
            for (int d = 0; d < 3; ++d) dir[d] = mz_rand(A.rp, st) < 0.5 ? -1.0 : 1.0;
            for (int d = 0; d < 3; ++d) { ox[0][d] = xm[d] + dir[d] * xs[d]; ox[1][d] = xm[d] - dir[d] * xs[d]; }
            ow[0] = 0.5 * w;
            ow[1] = ow[0];
        } else {
            for (int h = 0; h < n_out; ++h) {
                const int64_t p = oct_slot(ord[b + h], s1, n1, s2);
                ow[h] = A.sp.w[p];
                ov[h][0] = A.sp.v0[p]; ov[h][1] = A.sp.v1[p]; ov[h][2] = A.sp.v2[p];
                ox[h][0] = A.sp.x0[p]; ox[h][1] = A.sp.x1[p]; ox[h][2] = A.sp.x2[p];
            }
        }
        for (int h = 0; h < n_out; ++h) {
            const int64_t dst = oct_slot(slot + h, s1, n1, s2);
            double x0 = ox[h][0];
            if (A.grid_clamp && n_out == 2) x0 = x0 < A.min_x ? A.min_x : (x0 > A.max_x ? A.max_x : x0);
            A.tmp.x0[dst] = x0; A.tmp.x1[dst] = ox[h][1]; A.tmp.x2[dst] = ox[h][2];
            A.tmp.v0[dst] = ov[h][0]; A.tmp.v1[dst] = ov[h][1]; A.tmp.v2[dst] = ov[h][2];
            A.tmp.w[dst] = ow[h];
        }
    }
    __syncthreads();
    merge_finish(A.sp, A.tmp, c, N, n_new_s, s1, n1, s2, n2, A.old_pia, A.n_merged, A.n_cells);
}

extern "C" int mz_merge_grid(mz_store *h, int32_t species, int32_t threshold, int32_t Nx, int32_t Ny, int32_t Nz,
                             const double *extent_multiplier, const double *extents, double mass,
                             int32_t grid_clamp, int32_t squash, uint64_t tag, int32_t rng_mode,
                             int64_t *n_merged) {
    if (species >= 1 && species <= h->n_species) h->sp[species - 1].contig_known = false;
    MZ_CHECK(species >= 1 && species <= h->n_species, "merge_grid: bad species");
    MZ_CHECK(Nx >= 1 && Ny >= 1 && Nz >= 1 && (int64_t)Nx * Ny * Nz + 8 <= GM_MAX_CELLS,
             "merge_grid: Nx Ny Nz must be positive with Nx Ny Nz + 8 <= 4104");
    MZ_CHECK(extents != nullptr || extent_multiplier != nullptr,
             "merge_grid: give the extent multipliers (extents from the props) or the explicit extents");
    MZ_CHECK(!grid_clamp || h->has_grid, "merge_grid: grid clamp needs a grid");
    MZ_CHECK(rng_mode == 0 || (rng_mode == 1 && h->d_replay[2]), "merge_grid: replay table not set");
    MzSpecies &s = h->sp[species - 1];
    const int nc = h->n_cells;
    int32_t *old_pia = (int32_t *)mz_scratch(h, "oct_oldpia", (size_t)nc * 5 * sizeof(int32_t));
    int32_t *cnt = (int32_t *)mz_scratch(h, "oct_cnt", sizeof(int32_t));
    int32_t *ndel = (int32_t *)mz_scratch(h, "oct_ndel", (size_t)nc * sizeof(int32_t));
    int64_t *off = (int64_t *)mz_scratch(h, "oct_off", (size_t)(nc + 1) * sizeof(int64_t));
    MZ_CHECK(old_pia && cnt && ndel && off, "merge_grid: scratch allocation failed");
    MzCounters snap;
    int r = mz_read_counters(h, species, &snap);
    if (r) return r;
    MZ_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t), h->stream));
    GridArgs A{};
    A.sp = mz_view(s);
    A.tmp = A.sp;
    A.tmp.x0 = s.xb[0]; A.tmp.x1 = s.xb[1]; A.tmp.x2 = s.xb[2];
    A.tmp.v0 = s.vb[0]; A.tmp.v1 = s.vb[1]; A.tmp.v2 = s.vb[2];
    A.tmp.w = s.wb;
    A.threshold = threshold; A.Nx = Nx; A.Ny = Ny; A.Nz = Nz; A.grid_clamp = grid_clamp;
    if (extents) {   // [vx_lo, vx_hi, vy_lo, vy_hi, vz_lo, vz_hi]
        A.extent_mode = 1;
        for (int d = 0; d < 3; ++d) { A.ext_lo[d] = extents[2 * d]; A.ext_hi[d] = extents[2 * d + 1]; }
    } else {
        A.extent_mode = 0;
        for (int d = 0; d < 3; ++d) A.mult[d] = extent_multiplier[d];
        A.T = h->d_T + (size_t)(species - 1) * nc;
        A.v = h->d_v + (size_t)(species - 1) * nc * 3;
        MZ_CHECK(mass > 0.0, "merge_grid: the extents from the props need the species mass");
        A.mass = mass;
    }
    A.min_x = h->grid.min_x; A.max_x = h->grid.max_x;
    A.rp = mz_rng_params(h, 2, tag, rng_mode);
    A.old_pia = old_pia;
    A.n_merged = cnt;
    A.n_cells = nc;
    {
        MzTimed tm(h, "merge_grid");
        hipLaunchKernelGGL(k_grid_merge, dim3(nc), dim3(256), 0, h->stream, A);
        hipLaunchKernelGGL(k_merge_ndel, dim3((nc + 255) / 256), dim3(256), 0, h->stream, old_pia, nc, ndel);
        hipLaunchKernelGGL(k_scan_i32m, dim3(1), dim3(1024), 0, h->stream, ndel, (int64_t)nc, off, off + nc);
        hipLaunchKernelGGL(k_merge_push_ids, dim3(nc), dim3(64), 0, h->stream, A.sp, old_pia, off, off + nc,
                           snap.nbuffer, snap.n_total, nc);
    }
    MZ_HIP(hipGetLastError());
    if (n_merged) {
        int32_t v = 0;
        MZ_HIP(hipMemcpyAsync(&v, cnt, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
        MZ_HIP(hipStreamSynchronize(h->stream));
        *n_merged = v;
    }
    if (squash) return mz_squash(h, species);
    return 0;
}

// ===========================================================================
// NNLS merging (merging_nnls.jl:367-613, 949-1038, 1178-1276, 1334-1510) with
// the Lawson-Hanson solver (nnls.jl:26-548), one workgroup per merged cell.
//
// The solver's arithmetic is +, *, / and sqrt only.  Every loop the
// reference runs over columns is spread over the threads with one thread per
// column, and each column's loop over rows keeps the reference's order; the
// sums over particles (total weight, mean, variances, right-hand side) are
// sequential; contraction into fma is off.  The pivot choices and results
// therefore equal the oracle's bit for bit.  The LHS matrix lives in global
// scratch, row-major [row][column] so that the column threads read
// consecutive addresses.

#define NNLS_MAX_ROWS 64
#define NNLS_MAX_VMULT 4

#define NNLS_MAX_POS 16

struct NnlsArgs {
    MzSpView sp, tmp;
    int32_t m;                     // rows: m_v velocity moments, then m_p spatial moments
    int32_t m_v, m_p;
    int32_t mim[NNLS_MAX_ROWS][3];
    int32_t tot_order[NNLS_MAX_ROWS];
    int32_t mim_pos[NNLS_MAX_POS][3];
    int32_t pos_ix, pos_iy, pos_iz;   // 1 when the first spatial moment in x, y, z is conserved
    int32_t n_rand;                // random-pair columns after the fictitious ones
    MzRngParams rp;                // their draws: the cell's stream, 2 per pair
    double vref, inv_vref;
    int32_t scaling_vref, centered, n_vmult, iteration_mult;
    double vmult[NNLS_MAX_VMULT];
    double max_err, w_threshold;
    const int32_t *cells;          // batch: cells to merge (0-based)
    const int64_t *aoff;           // batch: offset of each cell's scratch (doubles)
    double *scratch;
    int32_t *status;               // [n_cells]
    int32_t *old_pia;              // [n_cells][5]
    int32_t n_cells;
    int32_t max_pcols;             // particle columns: the first max_pcols (< 0: all)
    int32_t canon;                 // 1: the rows are the 35 multi-indices of total order <= 4 in graded order
};

// Particle columns of a cell with N particles.
__device__ inline int32_t nnls_pcols(const NnlsArgs &A, int32_t N) {
    return (A.max_pcols >= 0 && A.max_pcols < N) ? A.max_pcols : N;
}

__device__ inline double nnls_ipow(double a, int p) {
    double r = 1.0;
    for (int i = 0; i < p; ++i) r *= a;
    return r;
}

// A random-pair column (compute_lhs_particles_additional!,
// merging_nnls.jl:1004-1030): pair r draws i = floor(R n_local), then k the
// same way (draws 2r and 2r + 1 of the cell's stream); i == k leaves the
// column zero.  Otherwise the candidate sits at the pair's mean velocity
// and mean position.  Returns false for a zero column.
__device__ inline bool nnls_pair(const NnlsArgs &A, const MzRngParams &rp, int c, int r, int32_t N, int32_t s1,
                                 int32_t n1, int32_t s2, double vm[3], double xm[3]) {
    MzStream st = mz_stream((uint64_t)c, (uint64_t)(2 * r));
    const int64_t i = (int64_t)floor(mz_rand(rp, st) * (double)N);
    const int64_t k = (int64_t)floor(mz_rand(rp, st) * (double)N);
    if (i == k) return false;
    const int64_t pi = oct_slot((int)i, s1, n1, s2), pk = oct_slot((int)k, s1, n1, s2);
    vm[0] = 0.5 * (A.sp.v0[pi] + A.sp.v0[pk]);
    vm[1] = 0.5 * (A.sp.v1[pi] + A.sp.v1[pk]);
    vm[2] = 0.5 * (A.sp.v2[pi] + A.sp.v2[pk]);
    xm[0] = 0.5 * (A.sp.x0[pi] + A.sp.x0[pk]);
    xm[1] = 0.5 * (A.sp.x1[pi] + A.sp.x1[pk]);
    xm[2] = 0.5 * (A.sp.x2[pi] + A.sp.x2[pk]);
    return true;
}

// Row q of a candidate column before scaling: velocity rows are the central
// moments of (cx, cy, cz), the velocity relative to the mean; spatial rows
// those of (px, py, pz), the position relative to the mean, with the
// spatial multi-indices for particles and (as the reference writes it,
// merging_nnls.jl:1023-1026) the velocity multi-indices for random pairs;
// kind 0 particle, 1 zero column, 2 octant particle (zero spatial rows),
// 3 random pair.
__device__ inline double nnls_entry(const NnlsArgs &A, int kind, int q, double cx, double cy, double cz, double px,
                                    double py, double pz) {
    if (kind == 1) return 0.0;
    if (q < A.m_v) return nnls_ipow(cx, A.mim[q][0]) * nnls_ipow(cy, A.mim[q][1]) * nnls_ipow(cz, A.mim[q][2]);
    const int r = q - A.m_v;
    if (kind == 0) return nnls_ipow(px, A.mim_pos[r][0]) * nnls_ipow(py, A.mim_pos[r][1]) * nnls_ipow(pz, A.mim_pos[r][2]);
    if (kind == 3) return nnls_ipow(px, A.mim[r][0]) * nnls_ipow(py, A.mim[r][1]) * nnls_ipow(pz, A.mim[r][2]);
    return 0.0;
}

// Row scaling (scale_lhs_rhs!, merging_nnls.jl:1178-1276): velocity rows by
// the variances (or vref), spatial rows by the spatial variances.
__device__ inline double nnls_rowscale(const NnlsArgs &A, int q, double Ex, double Ey, double Ez, double Epx,
                                       double Epy, double Epz) {
    if (q < A.m_v) {
        if (!A.scaling_vref) {
            const double iex = Ex > 1e-6 ? 1.0 / Ex : A.inv_vref;
            const double iey = Ey > 1e-6 ? 1.0 / Ey : A.inv_vref;
            const double iez = Ez > 1e-6 ? 1.0 / Ez : A.inv_vref;
            return nnls_ipow(iex, A.mim[q][0]) * nnls_ipow(iey, A.mim[q][1]) * nnls_ipow(iez, A.mim[q][2]);
        }
        return nnls_ipow(A.inv_vref, A.tot_order[q]);
    }
    const int r = q - A.m_v;
    const double ipx = Epx > 1e-6 ? 1.0 / Epx : 1.0;
    const double ipy = Epy > 1e-6 ? 1.0 / Epy : 1.0;
    const double ipz = Epz > 1e-6 ? 1.0 / Epz : 1.0;
    return nnls_ipow(ipx, A.mim_pos[r][0]) * nnls_ipow(ipy, A.mim_pos[r][1]) * nnls_ipow(ipz, A.mim_pos[r][2]);
}

#pragma clang fp contract(off)

__device__ inline void nnls_argmax(double v, int i, double *rw, int *ri, double &bw, int &bi) {
    // first maximum: larger value wins, ties go to the smaller index
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double ov = __shfl_xor(v, o, 64);
        int oi = __shfl_xor(i, o, 64);
        if (oi >= 0 && (i < 0 || ov > v || (ov == v && oi < i))) { v = ov; i = oi; }
    }
    if ((threadIdx.x & 63) == 0) { rw[threadIdx.x >> 6] = v; ri[threadIdx.x >> 6] = i; }
    __syncthreads();
    double w0 = -1.0;
    int i0 = -1;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
        if (ri[k] >= 0 && (i0 < 0 || rw[k] > w0 || (rw[k] == w0 && ri[k] < i0))) { w0 = rw[k]; i0 = ri[k]; }
    }
    bw = w0;
    bi = i0;
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_nnls_merge(NnlsArgs A) {
    __shared__ double b[NNLS_MAX_ROWS], zz[NNLS_MAX_ROWS], rhs[NNLS_MAX_ROWS], rowscale[NNLS_MAX_ROWS];
    __shared__ double st_d[24];
    __shared__ int32_t st_i[16];
    __shared__ double rw[4];
    __shared__ int ri[4];
    const int c = A.cells[blockIdx.x];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int32_t s1 = A.sp.pia.start1[c], n1 = A.sp.pia.n1[c], s2 = A.sp.pia.start2[c], n2 = A.sp.pia.n2[c];
    const int32_t N = n1 + n2;
    const int32_t Np = nnls_pcols(A, N);                // particle columns
    const int m = A.m;
    const int n_add = (A.centered ? 1 : 0) + 8 * A.n_vmult;
    const int n = Np + n_add + A.n_rand;
    const MzRngParams rp = mz_rng_resolve(A.rp);
    double *Am = A.scratch + A.aoff[blockIdx.x];       // [m][n] row-major
    double *xs = Am + (int64_t)m * n;                   // x[n]
    double *ws = xs + n;                                // w[n] (dual)
    double *cn = ws + n;                                // column norms[n]
    int32_t *idx = (int32_t *)(cn + n);                 // idx[n]
#define AM(r, cc) Am[(int64_t)((r) - 1) * n + ((cc) - 1)]
    // ---- compute_w_total_v0! (sequential sums, exact min/max)
    if (tid == 0) {
        double W = 0.0, v0[3] = {0, 0, 0}, x0[3] = {0, 0, 0};
        double mn[3] = {299792458.0, 299792458.0, 299792458.0};
        double mx[3] = {-299792458.0, -299792458.0, -299792458.0};
        for (int li = 0; li < N; ++li) {
            int64_t p = oct_slot(li, s1, n1, s2);
            double w = A.sp.w[p], v[3] = {A.sp.v0[p], A.sp.v1[p], A.sp.v2[p]};
            double xx[3] = {A.sp.x0[p], A.sp.x1[p], A.sp.x2[p]};
            W += w;
            for (int d = 0; d < 3; ++d) {
                v0[d] = v0[d] + w * v[d];
                x0[d] = x0[d] + w * xx[d];
                if (v[d] > mx[d]) mx[d] = v[d];
                if (v[d] < mn[d]) mn[d] = v[d];
            }
        }
        for (int d = 0; d < 3; ++d) {
            v0[d] = v0[d] / W;
            x0[d] = x0[d] / W;
            st_d[d] = v0[d];
            st_d[3 + d] = x0[d];
            st_d[6 + d] = mn[d] - v0[d];
            st_d[9 + d] = mx[d] - v0[d];
        }
        st_d[12] = W;
    }
    __syncthreads();
    const double W = st_d[12];
    const double v0x = st_d[0], v0y = st_d[1], v0z = st_d[2];
    // ---- variances of velocity and position (compute_lhs_and_rhs!, :539-548):
    // one accumulator per thread
    if (tid < 6) {
        double E = 0.0;
        const int d = tid % 3;
        for (int li = 0; li < N; ++li) {
            int64_t p = oct_slot(li, s1, n1, s2);
            double v = tid < 3 ? (d == 0 ? A.sp.v0[p] : (d == 1 ? A.sp.v1[p] : A.sp.v2[p]))
                               : (d == 0 ? A.sp.x0[p] : (d == 1 ? A.sp.x1[p] : A.sp.x2[p]));
            double a = v - st_d[tid];
            E = E + A.sp.w[p] * (a * a);
        }
        st_d[tid < 3 ? 13 + tid : 18 + tid] = sqrt(E / W);
    }
    // ---- right-hand side: one row per thread, particles in order
    for (int q = tid; q < m; q += nt) {
        double r = 0.0;
        for (int li = 0; li < N; ++li) {
            int64_t p = oct_slot(li, s1, n1, s2);
            double t = nnls_entry(A, 0, q, A.sp.v0[p] - v0x, A.sp.v1[p] - v0y, A.sp.v2[p] - v0z,
                                  A.sp.x0[p] - st_d[3], A.sp.x1[p] - st_d[4], A.sp.x2[p] - st_d[5]);
            r = r + A.sp.w[p] * t;
        }
        rhs[q] = r / W;
    }
    __syncthreads();
    const double Ex = st_d[13], Ey = st_d[14], Ez = st_d[15];
    // ---- row scaling (scale_lhs_rhs!, :1178-1276)
    if (tid < m) {
        const double rv = nnls_rowscale(A, tid, Ex, Ey, Ez, st_d[21], st_d[22], st_d[23]);
        rowscale[tid] = rv;
        b[tid] = rhs[tid] * rv;
    }
    __syncthreads();
    // ---- columns: particles, centre, octant particles, random pairs;
    // scaled, unit norm
    for (int j = tid; j < n; j += nt) {
        double cx, cy, cz, px = 0.0, py = 0.0, pz = 0.0;
        int kind = 0;
        if (j < Np) {
            int64_t p = oct_slot(j, s1, n1, s2);
            cx = A.sp.v0[p] - v0x; cy = A.sp.v1[p] - v0y; cz = A.sp.v2[p] - v0z;
            px = A.sp.x0[p] - st_d[3]; py = A.sp.x1[p] - st_d[4]; pz = A.sp.x2[p] - st_d[5];
        } else if (j >= Np + n_add) {
            double vm[3], xm[3];
            kind = nnls_pair(A, rp, c, j - Np - n_add, N, s1, n1, s2, vm, xm) ? 3 : 1;
            cx = kind == 3 ? vm[0] - v0x : 0.0; cy = kind == 3 ? vm[1] - v0y : 0.0; cz = kind == 3 ? vm[2] - v0z : 0.0;
            px = kind == 3 ? xm[0] - st_d[3] : 0.0; py = kind == 3 ? xm[1] - st_d[4] : 0.0;
            pz = kind == 3 ? xm[2] - st_d[5] : 0.0;
        } else {
            int f = j - Np;
            if (A.centered && f == 0) {
                kind = 1;
                cx = cy = cz = 0.0;
            } else {
                kind = 2;
                f -= A.centered ? 1 : 0;
                const int vm = f / 8, o = f % 8 + 1;
                const double mult = A.vmult[vm];
                const double sx = (o % 2 == 1) ? -1.0 : 1.0;
                const double sy = (o == 3 || o == 4 || o == 7 || o == 8) ? 1.0 : -1.0;
                const double sz = o >= 5 ? 1.0 : -1.0;
                double e3[3] = {Ex, Ey, Ez}, s3[3] = {sx, sy, sz}, r3[3];
                for (int d = 0; d < 3; ++d) {
                    double ms = mult * (s3[d] * e3[d]);
                    r3[d] = ms > st_d[9 + d] ? st_d[9 + d] * mult : (ms < st_d[6 + d] ? st_d[6 + d] * mult : ms);
                }
                cx = r3[0]; cy = r3[1]; cz = r3[2];
            }
        }
        double nn = 0.0;
        for (int q = 0; q < m; ++q) {
            double t = nnls_entry(A, kind, q, cx, cy, cz, px, py, pz);
            t *= rowscale[q];
            AM(q + 1, j + 1) = t;
            nn = nn + t * t;
        }
        nn = sqrt(nn);
        nn = nn > 1e-15 ? nn : 1.0;
        const double inv = 1.0 / nn;
        cn[j] = inv;
        for (int q = 0; q < m; ++q) AM(q + 1, j + 1) *= inv;
        xs[j] = 0.0;
        ws[j] = 0.0;
        idx[j] = j + 1;
    }
    __syncthreads();

    // ---- Lawson-Hanson solve!(work, iteration_mult * n, 1e-14)
    const double kkt_tol = 1e-14, factor = 0.01;
    const int64_t max_iter = (int64_t)A.iteration_mult * n;
    // shared solver state: 0 iz1, 1 iz2, 2 nsetp, 3 j, 4 jj, 5 iz, 6 terminated, 7 mode, 8 brk, 9 i
    if (tid == 0) {
        st_i[0] = 1; st_i[1] = n; st_i[2] = 0; st_i[3] = 0; st_i[4] = 0; st_i[5] = 0;
        st_i[6] = 0; st_i[7] = 1; st_i[8] = 0; st_i[9] = 0;
        st_d[16] = 0.0;   // up
        st_d[17] = 0.0;   // iter (as double)
    }
    __syncthreads();
    for (;;) {
        if (st_i[0] > st_i[1] || st_i[2] >= m) break;
        // dual vector over the zero set
        {
            const int iz1 = st_i[0], iz2 = st_i[1], nsetp = st_i[2];
            for (int i = iz1 + tid; i <= iz2; i += nt) {
                const int idxi = idx[i - 1];
                double sm = 0.0;
                for (int l = nsetp + 1; l <= m; ++l) sm = sm + AM(l, idxi) * b[l - 1];
                ws[idxi - 1] = sm;
            }
        }
        __syncthreads();
        bool terminated = false;
        for (;;) {
            // largest positive dual (first maximum in idx order)
            const int iz1 = st_i[0], iz2 = st_i[1];
            double wmax = 0.0;
            int izmax = -1;
            for (int i = iz1 + tid; i <= iz2; i += nt) {
                const double wv = ws[idx[i - 1] - 1];
                if (wv > wmax || (izmax < 0 && wv > 0.0)) {
                    if (izmax < 0 || wv > wmax) { wmax = wv; izmax = i; }
                }
            }
            double bw;
            int bi;
            nnls_argmax(izmax >= 0 ? wmax : -1.0, izmax, rw, ri, bw, bi);
            if (bi < 0 || bw <= kkt_tol) { terminated = true; break; }
            if (tid == 0) {
                const int nsetp = st_i[2];
                const int iz = bi, j = idx[iz - 1];
                st_i[5] = iz;
                st_i[3] = j;
                const double Asave = AM(nsetp + 1, j);
                // construct_householder! on A[nsetp+1:m, j]
                double up = st_d[16];
                const int len = m - nsetp;
                if (len > 1) {
                    double cl = 0.0;
                    for (int l = nsetp + 1; l <= m; ++l) cl = fmax(cl, fabs(AM(l, j)));
                    double clinv = 1.0 / cl, sm = 0.0;
                    for (int l = nsetp + 1; l <= m; ++l) { double t = AM(l, j) * clinv; sm += t * t; }
                    cl *= sqrt(sm);
                    if (AM(nsetp + 1, j) > 0) cl = -cl;
                    up = AM(nsetp + 1, j) - cl;
                    AM(nsetp + 1, j) = cl;
                }
                st_d[16] = up;
                double unorm = 0.0;
                for (int l = 1; l <= nsetp; ++l) unorm += AM(l, j) * AM(l, j);
                unorm = sqrt(unorm);
                int accept = 0;
                if (((unorm + fabs(AM(nsetp + 1, j)) * factor) - unorm) > 0) {
                    for (int q = 0; q < m; ++q) zz[q] = b[q];
                    // apply_householder!(A[nsetp+1:m, j], up, zz[nsetp+1:m])
                    if (len > 1) {
                        double bb = up * AM(nsetp + 1, j);
                        if (bb < 0) {
                            bb = 1.0 / bb;
                            double sm = zz[nsetp] * up;
                            for (int l = 2; l <= len; ++l) sm = sm + zz[nsetp + l - 1] * AM(nsetp + l, j);
                            if (sm != 0) {
                                sm *= bb;
                                zz[nsetp] = zz[nsetp] + sm * up;
                                for (int l = 2; l <= len; ++l) zz[nsetp + l - 1] = zz[nsetp + l - 1] + sm * AM(nsetp + l, j);
                            }
                        }
                    }
                    const double ztest = zz[nsetp] / AM(nsetp + 1, j);
                    if (ztest > 0) accept = 1;
                }
                if (!accept) {
                    AM(nsetp + 1, j) = Asave;
                    ws[j - 1] = 0;
                }
                st_i[8] = accept;
            }
            __syncthreads();
            if (st_i[8]) break;
        }
        if (terminated) break;
        // move j from Z to P, transform the remaining columns
        if (tid == 0) {
            for (int q = 0; q < m; ++q) b[q] = zz[q];
            const int iz = st_i[5], iz1 = st_i[0], j = st_i[3];
            idx[iz - 1] = idx[iz1 - 1];
            idx[iz1 - 1] = j;
            st_i[0] = iz1 + 1;
            st_i[2] = st_i[2] + 1;
        }
        __syncthreads();
        {
            const int iz1 = st_i[0], iz2 = st_i[1], nsetp = st_i[2], j = st_i[3];
            const double up = st_d[16];
            const int len = m - nsetp + 1;
            for (int jz = iz1 + tid; jz <= iz2; jz += nt) {
                const int jj = idx[jz - 1];
                if (len > 1) {
                    double bb = up * AM(nsetp, j);
                    if (bb < 0) {
                        bb = 1.0 / bb;
                        double sm = AM(nsetp, jj) * up;
                        for (int l = 2; l <= len; ++l) sm = sm + AM(nsetp + l - 1, jj) * AM(nsetp + l - 1, j);
                        if (sm != 0) {
                            sm *= bb;
                            AM(nsetp, jj) = AM(nsetp, jj) + sm * up;
                            for (int l = 2; l <= len; ++l)
                                AM(nsetp + l - 1, jj) = AM(nsetp + l - 1, jj) + sm * AM(nsetp + l - 1, j);
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (tid == 0) {
            const int nsetp = st_i[2], j = st_i[3];
            if (nsetp != m)
                for (int l = nsetp + 1; l <= m; ++l) AM(l, j) = 0;
            ws[j - 1] = 0;
            // solve_triangular_system!
            int jj = st_i[4];
            for (int l = 1; l <= nsetp; ++l) {
                const int ip = nsetp + 1 - l;
                if (l != 1)
                    for (int ii = 1; ii <= ip; ++ii) zz[ii - 1] = zz[ii - 1] - AM(ii, jj) * zz[ip];
                jj = idx[ip - 1];
                zz[ip - 1] /= AM(ip, jj);
            }
            st_i[4] = jj;
        }
        __syncthreads();
        // secondary loop
        for (;;) {
            if (tid == 0) {
                st_i[8] = 0;   // 1: leave the secondary loop, 2: iteration cap
                st_d[17] += 1.0;
                if (st_d[17] > (double)max_iter) {
                    st_i[7] = 3;
                    st_i[8] = 2;
                } else {
                    const int nsetp = st_i[2];
                    double alpha = 2.0;
                    int jj = st_i[4];
                    for (int ip = 1; ip <= nsetp; ++ip) {
                        const int l = idx[ip - 1];
                        if (zz[ip - 1] <= 0) {
                            const double t = -xs[l - 1] / (zz[ip - 1] - xs[l - 1]);
                            if (alpha > t) { alpha = t; jj = ip; }
                        }
                    }
                    st_i[4] = jj;
                    if (alpha == 2) {
                        st_i[8] = 1;
                    } else {
                        for (int ip = 1; ip <= nsetp; ++ip) {
                            const int l = idx[ip - 1];
                            xs[l - 1] = xs[l - 1] + alpha * (zz[ip - 1] - xs[l - 1]);
                        }
                        st_i[9] = idx[jj - 1];   // i
                    }
                }
            }
            __syncthreads();
            if (st_i[8] == 2) { terminated = true; break; }
            if (st_i[8] == 1) break;
            // move coefficient i from P to Z (with Givens rotations), repeat
            // while some remaining coefficient is not positive
            for (;;) {
                if (tid == 0) {
                    xs[st_i[9] - 1] = 0;
                    st_i[12] = st_i[4] != st_i[2];
                    if (st_i[12]) st_i[4] += 1;   // jj += 1 persists across passes, as in nnls.jl:449-451
                }
                __syncthreads();
                if (st_i[12]) {
                    const int jj0 = st_i[4], nsetp = st_i[2];
                    for (int jx = jj0; jx <= nsetp; ++jx) {
                        if (tid == 0) {
                            const int ii = idx[jx - 1];
                            idx[jx - 2] = ii;
                            const double a = AM(jx - 1, ii), bq = AM(jx, ii);
                            double cc, ss, sig;
                            if (fabs(a) > fabs(bq)) {
                                double xr = bq / a, yr = sqrt(1 + xr * xr);
                                cc = (1 / yr) * (a > 0 ? 1.0 : (a < 0 ? -1.0 : 0.0));
                                ss = cc * xr;
                                sig = fabs(a) * yr;
                            } else if (bq != 0) {
                                double xr = a / bq, yr = sqrt(1 + xr * xr);
                                ss = (1 / yr) * (bq > 0 ? 1.0 : -1.0);
                                cc = ss * xr;
                                sig = fabs(bq) * yr;
                            } else {
                                sig = 0.0; cc = 0.0; ss = 1.0;
                            }
                            AM(jx - 1, ii) = sig;
                            AM(jx, ii) = 0;
                            st_d[18] = cc; st_d[19] = ss;
                            st_i[10] = ii;
                            double temp = b[jx - 2];
                            b[jx - 2] = cc * temp + ss * b[jx - 1];
                            b[jx - 1] = -ss * temp + cc * b[jx - 1];
                        }
                        __syncthreads();
                        const double cc = st_d[18], ss = st_d[19];
                        const int ii = st_i[10];
                        for (int l = 1 + tid; l <= n; l += nt) {
                            if (l != ii) {
                                const double temp = AM(jx - 1, l);
                                AM(jx - 1, l) = cc * temp + ss * AM(jx, l);
                                AM(jx, l) = -ss * temp + cc * AM(jx, l);
                            }
                        }
                        __syncthreads();
                    }
                }
                if (tid == 0) {
                    st_i[2] -= 1;
                    st_i[0] -= 1;
                    idx[st_i[0] - 1] = st_i[9];
                    int allfeasible = 1;
                    for (int jji = 1; jji <= st_i[2]; ++jji) {
                        const int i = idx[jji - 1];
                        st_i[9] = i;
                        if (xs[i - 1] <= 0) { allfeasible = 0; break; }
                    }
                    st_i[11] = allfeasible;
                }
                __syncthreads();
                if (st_i[11]) break;
            }
            if (tid == 0) {
                const int nsetp = st_i[2];
                for (int q = 0; q < m; ++q) zz[q] = b[q];
                int jj = st_i[4];
                for (int l = 1; l <= nsetp; ++l) {
                    const int ip = nsetp + 1 - l;
                    if (l != 1)
                        for (int ii = 1; ii <= ip; ++ii) zz[ii - 1] = zz[ii - 1] - AM(ii, jj) * zz[ip];
                    jj = idx[ip - 1];
                    zz[ip - 1] /= AM(ip, jj);
                }
                st_i[4] = jj;
            }
            __syncthreads();
        }
        if (terminated) break;
        if (tid == 0)
            for (int i = 1; i <= st_i[2]; ++i) xs[idx[i - 1] - 1] = zz[i - 1];
        __syncthreads();
    }
    // ---- compute_post_merge_particles_nnls! (:1334-1404)
    __shared__ int32_t n_new_s, ok_s;
    if (tid == 0) {
        const int nsetp = st_i[2];
        double sm = 0.0;
        if (nsetp < m)
            for (int i = nsetp + 1; i <= m; ++i) sm += b[i - 1] * b[i - 1];
        const double rnorm = sqrt(sm);
        int ok = !(rnorm > A.max_err);
        int nonzero = 0;
        double ndw = 0.0;
        if (ok) {
            for (int i = 0; i < n; ++i) {
                xs[i] = xs[i] * cn[i];
                if (xs[i] > A.w_threshold) { nonzero += 1; ndw += xs[i]; }
            }
            if (nonzero >= N) ok = 0;
        }
        ok_s = ok;
        n_new_s = ok ? nonzero : N;
        st_d[20] = W / ndw;
        A.status[c] = ok ? 1 : -1;
    }
    __syncthreads();
    const int32_t n_new = n_new_s;
    if (ok_s) {
        if (tid == 0) {
            const double wt = st_d[20];
            int cur = 0;
            for (int j = 0; j < n; ++j) {
                if (!(xs[j] > A.w_threshold)) continue;
                const int64_t dst = oct_slot(cur, s1, n1, s2);
                ++cur;
                double vx, vy, vz, xx = st_d[3], xy = st_d[4], xz = st_d[5];
                if (j < Np) {
                    int64_t p = oct_slot(j, s1, n1, s2);
                    vx = A.sp.v0[p]; vy = A.sp.v1[p]; vz = A.sp.v2[p];
                    xx = A.sp.x0[p]; xy = A.sp.x1[p]; xz = A.sp.x2[p];
                } else if (j >= Np + n_add) {
                    double vm[3] = {0, 0, 0}, xm[3] = {0, 0, 0};
                    nnls_pair(A, rp, c, j - Np - n_add, N, s1, n1, s2, vm, xm);
                    vx = vm[0]; vy = vm[1]; vz = vm[2];
                    xx = xm[0]; xy = xm[1]; xz = xm[2];
                } else {
                    int f = j - Np;
                    if (A.centered && f == 0) {
                        vx = v0x; vy = v0y; vz = v0z;
                    } else {
                        f -= A.centered ? 1 : 0;
                        const int vm = f / 8, o = f % 8 + 1;
                        const double mult = A.vmult[vm];
                        const double sx = (o % 2 == 1) ? -1.0 : 1.0;
                        const double sy = (o == 3 || o == 4 || o == 7 || o == 8) ? 1.0 : -1.0;
                        const double sz = o >= 5 ? 1.0 : -1.0;
                        double e3[3] = {Ex, Ey, Ez}, s3[3] = {sx, sy, sz}, r3[3];
                        for (int d = 0; d < 3; ++d) {
                            double ms = mult * (s3[d] * e3[d]);
                            r3[d] = ms > st_d[9 + d] ? st_d[9 + d] * mult : (ms < st_d[6 + d] ? st_d[6 + d] * mult : ms);
                        }
                        vx = v0x + r3[0]; vy = v0y + r3[1]; vz = v0z + r3[2];
                    }
                }
                A.tmp.w[dst] = xs[j] * wt;
                A.tmp.v0[dst] = vx; A.tmp.v1[dst] = vy; A.tmp.v2[dst] = vz;
                // positions: the candidate's where the first spatial moment is
                // conserved, else the mean (compute_post_merge_particles_nnls!)
                A.tmp.x0[dst] = A.pos_ix ? xx : st_d[3];
                A.tmp.x1[dst] = A.pos_iy ? xy : st_d[4];
                A.tmp.x2[dst] = A.pos_iz ? xz : st_d[5];
            }
        }
        __syncthreads();
        for (int li = tid; li < N; li += nt) {
            int64_t p = oct_slot(li, s1, n1, s2);
            if (li < n_new) {
                A.sp.x0[p] = A.tmp.x0[p]; A.sp.x1[p] = A.tmp.x1[p]; A.sp.x2[p] = A.tmp.x2[p];
                A.sp.v0[p] = A.tmp.v0[p]; A.sp.v1[p] = A.tmp.v1[p]; A.sp.v2[p] = A.tmp.v2[p];
                A.sp.w[p] = A.tmp.w[p];
            } else {
                A.sp.w[p] = 0.0;
            }
        }
    }
    if (tid == 0) {
        int32_t *op = A.old_pia + c * 5;
        op[0] = s1; op[1] = n1; op[2] = s2; op[3] = n2; op[4] = ok_s ? n_new : -1;
        if (ok_s) {
            const int32_t n_del = N - n_new;
            const int32_t d2 = min(n_del, n2), d1 = n_del - d2;
            const int32_t nn2 = n2 - d2, nn1 = n1 - d1;
            A.sp.pia.n2[c] = nn2;
            A.sp.pia.end2[c] = nn2 > 0 ? s2 + nn2 - 1 : -1;
            if (nn2 == 0) A.sp.pia.start2[c] = 0;
            A.sp.pia.n1[c] = nn1;
            A.sp.pia.end1[c] = nn1 > 0 ? s1 + nn1 - 1 : -1;
            if (nn1 == 0) A.sp.pia.start1[c] = 0;
            A.sp.pia.nloc[c] = nn1 + nn2;
            if (!(c == A.n_cells - 1) || (n_del > n2)) A.sp.ctr->contiguous = 0;
        }
    }
#undef AM
}


// ===========================================================================
// Candidate NNLS merging, one workgroup per cell with thread j owning column
// j: the same Lawson-Hanson iteration as k_nnls_merge (nnls.jl:273-548) on a
// column set of at most NF_COLS (the first particles plus the fictitious
// ones), with the scaled moment matrix resident in LDS.  The moments and the
// right-hand side are block reductions over the particles (thread-strided
// partial sums combined in a fixed order), and the steps the reference runs
// as loops over rows -- building and applying the Householder vector, the
// back substitution, the step length of the inner loop, the feasibility scan
// -- run on wave 0 with one lane per row.  Results agree with the oracle to
// rounding (sums are taken in another order), not bit for bit.
#define NF_COLS 256


// The reference's (fictitious) column f >= 0 after the particles: the
// centre particle (all-zero column) and the octant particles
// (compute_lhs_particles_additional!, merging_nnls.jl:949-1038), relative to
// the mean velocity.
__device__ inline bool nf_fict(const NnlsArgs &A, int f, const double *e3, const double *mnv, const double *mxv,
                               double &cx, double &cy, double &cz) {
    if (A.centered && f == 0) {
        cx = cy = cz = 0.0;
        return true;
    }
    f -= A.centered ? 1 : 0;
    const int vm = f / 8, o = f % 8 + 1;
    const double mult = A.vmult[vm];
    const double s3[3] = {(o % 2 == 1) ? -1.0 : 1.0, (o == 3 || o == 4 || o == 7 || o == 8) ? 1.0 : -1.0,
                          o >= 5 ? 1.0 : -1.0};
    double r3[3];
    for (int d = 0; d < 3; ++d) {
        const double ms = mult * (s3[d] * e3[d]);
        r3[d] = ms > mxv[d] ? mxv[d] * mult : (ms < mnv[d] ? mnv[d] * mult : ms);
    }
    cx = r3[0]; cy = r3[1]; cz = r3[2];
    return false;
}

// T is the storage type of the scaled moment matrix and of the solver's x
// and dual vectors; the arithmetic is FP64.  With T = float the active-set
// search runs on the rounded matrix and nf_polish re-solves the final
// passive set in FP64 (status 2 asks the host for the FP64 method).
template <int ROWS, typename T>
__global__ void __launch_bounds__(NF_COLS, sizeof(T) == 4 ? 4 : (ROWS <= 36 ? 2 : 1)) k_nnls_fast(NnlsArgs A) {
    constexpr bool MIXED = sizeof(T) == 4;
    // the matrix area also holds the moment partial sums before the matrix is
    // built, and the FP64 passive-set system after the search
    __shared__ __align__(16) unsigned char lds_mat[ROWS * NF_COLS * sizeof(T)];
    static_assert(4 * (ROWS + 8) * sizeof(double) <= sizeof(lds_mat), "moment partials fit the matrix area");
    T *Al = reinterpret_cast<T *>(lds_mat);     // scaled moment matrix [row][column]
    double *red = reinterpret_cast<double *>(lds_mat);   // [wave][ROWS + 8]
    __shared__ double bq[ROWS], zz[ROWS];
    __shared__ T xs[NF_COLS], ws[NF_COLS];
    __shared__ int16_t idx[NF_COLS];
    __shared__ double sd[24];
    __shared__ int32_t si[16];
    __shared__ double rw[4];
    __shared__ int ri[4];
    const int c = A.cells[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int32_t s1 = A.sp.pia.start1[c], n1 = A.sp.pia.n1[c], s2 = A.sp.pia.start2[c], n2 = A.sp.pia.n2[c];
    const int32_t N = n1 + n2;
    const int32_t Np = nnls_pcols(A, N);
    const int m = A.m;
    const int n_add = (A.centered ? 1 : 0) + 8 * A.n_vmult;
    const int n = Np + n_add + A.n_rand;
#define AL(r, cc) Al[((r) - 1) * NF_COLS + ((cc) - 1)]
#define RED(w, k) red[(w) * (ROWS + 8) + (k)]
    // ---- compute_w_total_v0!: W, mean velocity and position, velocity bounds
    {
        double W = 0.0, sv[3] = {0, 0, 0}, sx[3] = {0, 0, 0};
        double mn[3] = {299792458.0, 299792458.0, 299792458.0};
        double mx[3] = {-299792458.0, -299792458.0, -299792458.0};
        for (int li = tid; li < N; li += NF_COLS) {
            const int64_t p = oct_slot(li, s1, n1, s2);
            const double w = A.sp.w[p], v[3] = {A.sp.v0[p], A.sp.v1[p], A.sp.v2[p]};
            const double xx[3] = {A.sp.x0[p], A.sp.x1[p], A.sp.x2[p]};
            W += w;
            for (int d = 0; d < 3; ++d) {
                sv[d] += w * v[d];
                sx[d] += w * xx[d];
                mn[d] = fmin(mn[d], v[d]);
                mx[d] = fmax(mx[d], v[d]);
            }
        }
        double t[13] = {W, sv[0], sv[1], sv[2], sx[0], sx[1], sx[2], mn[0], mn[1], mn[2], mx[0], mx[1], mx[2]};
#pragma unroll
        for (int k = 0; k < 13; ++k) {
            const double r = k < 7 ? mz_wave_sum(t[k]) : (k < 10 ? mz_wave_min(t[k]) : mz_wave_max(t[k]));
            if (lane == 0) RED(wave, k) = r;
        }
        __syncthreads();
        if (tid < 13) {
            const int k = tid;
            double r;
            if (k < 7) r = ((RED(0, k) + RED(1, k)) + RED(2, k)) + RED(3, k);
            else if (k < 10) r = fmin(fmin(RED(0, k), RED(1, k)), fmin(RED(2, k), RED(3, k)));
            else r = fmax(fmax(RED(0, k), RED(1, k)), fmax(RED(2, k), RED(3, k)));
            sd[k] = r;
        }
        __syncthreads();
        if (tid == 0) {
            const double Wt = sd[0];
            for (int d = 0; d < 3; ++d) {
                const double v0 = sd[1 + d] / Wt, x0 = sd[4 + d] / Wt;
                const double lo = sd[7 + d] - v0, hi = sd[10 + d] - v0;
                sd[1 + d] = v0;
                sd[4 + d] = x0;
                sd[7 + d] = lo;
                sd[10 + d] = hi;
            }
        }
        __syncthreads();
    }
    const double W = sd[0];
    const double v0x = sd[1], v0y = sd[2], v0z = sd[3];
    // ---- variances and right-hand side (compute_lhs_and_rhs!, :509-613)
    {
        double acc[ROWS], e[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < ROWS; ++q) acc[q] = 0.0;
        for (int li = tid; li < N; li += NF_COLS) {
            const int64_t p = oct_slot(li, s1, n1, s2);
            const double w = A.sp.w[p];
            const double cx = A.sp.v0[p] - v0x, cy = A.sp.v1[p] - v0y, cz = A.sp.v2[p] - v0z;
            const double px = A.sp.x0[p] - sd[4], py = A.sp.x1[p] - sd[5], pz = A.sp.x2[p] - sd[6];
            e[0] += w * (cx * cx);
            e[1] += w * (cy * cy);
            e[2] += w * (cz * cz);
            e[3] += w * (px * px);
            e[4] += w * (py * py);
            e[5] += w * (pz * pz);
#pragma unroll
            for (int q = 0; q < ROWS; ++q) {
                if (q < m) acc[q] += w * nnls_entry(A, 0, q, cx, cy, cz, px, py, pz);
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double r = mz_wave_sum(e[k]);
            if (lane == 0) RED(wave, k) = r;
            const double rp_ = mz_wave_sum(e[3 + k]);
            if (lane == 0) RED(wave, ROWS + 3 + k) = rp_;
        }
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
            if (q < m) {
                const double r = mz_wave_sum(acc[q]);
                if (lane == 0) RED(wave, 3 + q) = r;
            }
        }
        __syncthreads();
        if (tid < 3) {
            const double E = ((RED(0, tid) + RED(1, tid)) + RED(2, tid)) + RED(3, tid);
            sd[13 + tid] = sqrt(E / W);
            const int kp = ROWS + 3 + tid;
            const double Ep = ((RED(0, kp) + RED(1, kp)) + RED(2, kp)) + RED(3, kp);
            sd[20 + tid] = sqrt(Ep / W);
        }
        if (tid >= 32 && tid < 32 + m) {
            const int q = tid - 32;
            zz[q] = (((RED(0, 3 + q) + RED(1, 3 + q)) + RED(2, 3 + q)) + RED(3, 3 + q)) / W;
        }
        __syncthreads();
    }
    const double Ex = sd[13], Ey = sd[14], Ez = sd[15];
    // ---- row scaling (scale_lhs_rhs_variance! / _vref!, :1178-1215); zz holds the moments
    __shared__ double rowscale[ROWS], b0[ROWS];
    if (tid < m) {
        const double rv = nnls_rowscale(A, tid, Ex, Ey, Ez, sd[20], sd[21], sd[22]);
        rowscale[tid] = rv;
        bq[tid] = zz[tid] * rv;
        b0[tid] = bq[tid];
    }
    __syncthreads();
    // ---- column j on thread j: moments of the candidate, row-scaled, unit norm
    // (scale_columns!, utils.jl:62-77); the candidate's velocity stays in
    // registers for the post-merge write
    // and position (relative to the means, and absolute for the write)
    double cvx = 0.0, cvy = 0.0, cvz = 0.0, cinv = 1.0;
    double cpx = sd[4], cpy = sd[5], cpz = sd[6];
    double cx = 0.0, cy = 0.0, cz = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    int kind = 0;   // see nnls_entry
    if (tid < n) {
        const int j = tid;
        if (j < Np) {
            const int64_t p = oct_slot(j, s1, n1, s2);
            cvx = A.sp.v0[p]; cvy = A.sp.v1[p]; cvz = A.sp.v2[p];
            cpx = A.sp.x0[p]; cpy = A.sp.x1[p]; cpz = A.sp.x2[p];
            cx = cvx - v0x; cy = cvy - v0y; cz = cvz - v0z;
            px = cpx - sd[4]; py = cpy - sd[5]; pz = cpz - sd[6];
        } else if (j >= Np + n_add) {
            double vm[3] = {0, 0, 0}, xm[3] = {0, 0, 0};
            kind = nnls_pair(A, mz_rng_resolve(A.rp), c, j - Np - n_add, N, s1, n1, s2, vm, xm) ? 3 : 1;
            cvx = vm[0]; cvy = vm[1]; cvz = vm[2];
            cpx = xm[0]; cpy = xm[1]; cpz = xm[2];
            if (kind == 3) {
                cx = cvx - v0x; cy = cvy - v0y; cz = cvz - v0z;
                px = cpx - sd[4]; py = cpy - sd[5]; pz = cpz - sd[6];
            }
        } else {
            const double e3[3] = {Ex, Ey, Ez};
            const double mnv[3] = {sd[7], sd[8], sd[9]}, mxv[3] = {sd[10], sd[11], sd[12]};
            kind = nf_fict(A, j - Np, e3, mnv, mxv, cx, cy, cz) ? 1 : 2;
            cvx = v0x + cx; cvy = v0y + cy; cvz = v0z + cz;
        }
        double nn = 0.0;
        for (int q = 0; q < m; ++q) {
            double t = nnls_entry(A, kind, q, cx, cy, cz, px, py, pz);
            t *= rowscale[q];
            AL(q + 1, j + 1) = t;
            nn = nn + t * t;
        }
        nn = sqrt(nn);
        nn = nn > 1e-15 ? nn : 1.0;
        cinv = 1.0 / nn;
        for (int q = 0; q < m; ++q) AL(q + 1, j + 1) *= cinv;
        xs[j] = 0.0;
        ws[j] = 0.0;
        idx[j] = j + 1;
    }

    // ---- Lawson-Hanson solve!(work, iteration_mult * n, 1e-14)
    const double kkt_tol = 1e-14, factor = 0.01;
    const int64_t max_iter = (int64_t)A.iteration_mult * n;
    // shared state: si 0 iz1, 1 iz2, 2 nsetp, 3 j, 4 jj, 5 iz, 7 mode, 8 flag, 9 i, 11 allfeasible, 12 rotate
    // sd 16 up, 17 iteration count, 18/19 Givens c/s
    if (tid == 0) {
        si[0] = 1; si[1] = n; si[2] = 0; si[3] = 0; si[4] = 0; si[5] = 0;
        si[7] = 1; si[8] = 0; si[9] = 0;
        si[14] = 0;   // 1: the duals of the zero set are current
        sd[16] = 0.0;
        sd[17] = 0.0;
    }
    __syncthreads();
    // back substitution (solve_triangular_system!) on wave 0, lane r = row r + 1:
    // zz <- R^-1 zz over the nsetp columns of P
    // Lane k holds the column at P position k + 1 and the reciprocal of its
    // diagonal entry; each step's column of R is loaded one step ahead.
    auto back_subst = [&]() {
        const int nsetp = si[2];
        const int colk = lane < nsetp ? idx[lane] : 1;
        const double rdk = lane < nsetp ? 1.0 / AL(lane + 1, colk) : 0.0;
        double zr = lane < m ? zz[lane] : 0.0;
        // the column of R the step at position ip subtracts (rows above ip)
        auto rcol = [&](int ip) {
            return (ip >= 1 && lane < ip) ? AL(lane + 1, __builtin_amdgcn_readlane(colk, ip)) : 0.0;
        };
        double p1 = rcol(nsetp - 1), p2 = rcol(nsetp - 2);
        for (int l = 1; l <= nsetp; ++l) {
            const int ip = nsetp + 1 - l;
            if (l != 1) {
                const double zip = mz_readlane_d(zr, ip);
                if (lane < ip) zr = zr - p1 * zip;
                p1 = p2;
                p2 = rcol(ip - 2);
            }
            if (lane == ip - 1) zr = zr * rdk;
        }
        if (lane < m) zz[lane] = zr;
        if (lane == 0 && nsetp >= 1) si[4] = __builtin_amdgcn_readlane(colk, 0);
    };
    // sum over rows lo..m of column c times bq, in four partial sums
    auto dual_of = [&](int c, int lo) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int l = lo;
        for (; l + 3 <= m; l += 4) {
            s0 = s0 + AL(l, c) * bq[l - 1];
            s1 = s1 + AL(l + 1, c) * bq[l];
            s2 = s2 + AL(l + 2, c) * bq[l + 1];
            s3 = s3 + AL(l + 3, c) * bq[l + 2];
        }
        for (; l <= m; ++l) s0 = s0 + AL(l, c) * bq[l - 1];
        return (s0 + s1) + (s2 + s3);
    };
    for (;;) {
        if (si[0] > si[1] || si[2] >= m) break;
        // dual vector over the zero set (already current after a column
        // transform that removed nothing)
        if (!si[14]) {
            const int iz1 = si[0], iz2 = si[1], nsetp = si[2];
            for (int i = iz1 + tid; i <= iz2; i += NF_COLS) {
                const int idxi = idx[i - 1];
                ws[idxi - 1] = dual_of(idxi, nsetp + 1);
            }
        }
        __syncthreads();
        bool terminated = false;
        for (;;) {
            // largest positive dual (first maximum in idx order)
            const int iz1 = si[0], iz2 = si[1];
            double wmax = 0.0;
            int izmax = -1;
            for (int i = iz1 + tid; i <= iz2; i += NF_COLS) {
                const double wv = ws[idx[i - 1] - 1];
                if (wv > wmax || (izmax < 0 && wv > 0.0)) {
                    if (izmax < 0 || wv > wmax) { wmax = wv; izmax = i; }
                }
            }
            double bw;
            int bi;
            nnls_argmax(izmax >= 0 ? wmax : -1.0, izmax, rw, ri, bw, bi);
            if (bi < 0 || bw <= kkt_tol) { terminated = true; break; }
            if (wave == 0) {
                // construct_householder! on rows nsetp+1..m of column j, the
                // column test and apply_householder! to a copy of b
                const int nsetp = si[2];
                const int j = idx[bi - 1];
                const int len = m - nsetp;
                const double a = lane < m ? AL(lane + 1, j) : 0.0;
                const bool hh = lane >= nsetp && lane < m;
                const double asave = mz_readlane_d(a, nsetp);
                double up = sd[16], top = asave;
                if (len > 1) {
                    double cl = mz_wave_max(hh ? fabs(a) : 0.0);
                    const double clinv = 1.0 / cl;
                    const double t = hh ? a * clinv : 0.0;
                    cl *= sqrt(mz_wave_sum(t * t));
                    if (asave > 0) cl = -cl;
                    up = asave - cl;
                    top = cl;
                }
                const double unorm = sqrt(mz_wave_sum(lane < nsetp ? a * a : 0.0));
                int accept = 0;
                double zr = lane < m ? bq[lane] : 0.0;
                if (((unorm + fabs(top) * factor) - unorm) > 0) {
                    if (len > 1) {
                        double bb = up * top;
                        if (bb < 0) {
                            bb = 1.0 / bb;
                            const double ur = lane == nsetp ? up : (hh ? a : 0.0);
                            double sm = mz_wave_sum(zr * ur);
                            if (sm != 0) {
                                sm *= bb;
                                if (hh) zr = zr + sm * ur;
                            }
                        }
                    }
                    const double ztest = mz_readlane_d(zr, nsetp) / top;
                    if (ztest > 0) accept = 1;
                }
                if (accept) {
                    if (lane < m) zz[lane] = zr;
                    if (lane == nsetp) AL(nsetp + 1, j) = top;
                } else if (lane == 0) {
                    ws[j - 1] = 0;
                }
                if (lane == 0) {
                    sd[16] = up;
                    si[8] = accept;
                    si[5] = bi;
                    si[3] = j;
                }
            }
            __syncthreads();
            if (si[8]) break;
        }
        if (terminated) break;
        // move j from Z to P, transform the remaining columns
        if (tid == 0) {
            const int iz = si[5], iz1 = si[0], j = si[3];
            idx[iz - 1] = idx[iz1 - 1];
            idx[iz1 - 1] = j;
            si[0] = iz1 + 1;
            si[2] = si[2] + 1;
        }
        if (tid < m) bq[tid] = zz[tid];
        __syncthreads();
        {
            const int iz1 = si[0], iz2 = si[1], nsetp = si[2], j = si[3];
            const double up = sd[16];
            const int len = m - nsetp + 1;
            for (int jz = iz1 + tid; jz <= iz2; jz += NF_COLS) {
                const int jj = idx[jz - 1];
                if (len > 1) {
                    double bb = up * AL(nsetp, j);
                    if (bb < 0) {
                        bb = 1.0 / bb;
                        double s0 = AL(nsetp, jj) * up, s1 = 0.0, s2 = 0.0, s3 = 0.0;
                        int l = 2;
                        for (; l + 2 <= len; l += 3) {
                            s1 = s1 + AL(nsetp + l - 1, jj) * AL(nsetp + l - 1, j);
                            s2 = s2 + AL(nsetp + l, jj) * AL(nsetp + l, j);
                            s3 = s3 + AL(nsetp + l + 1, jj) * AL(nsetp + l + 1, j);
                        }
                        for (; l <= len; ++l) s1 = s1 + AL(nsetp + l - 1, jj) * AL(nsetp + l - 1, j);
                        double sm = (s0 + s1) + (s2 + s3);
                        if (sm != 0) {
                            sm *= bb;
                            AL(nsetp, jj) = AL(nsetp, jj) + sm * up;
                            for (int l2 = 2; l2 <= len; ++l2)
                                AL(nsetp + l2 - 1, jj) = AL(nsetp + l2 - 1, jj) + sm * AL(nsetp + l2 - 1, j);
                        }
                    }
                }
                // the column's dual for the next pass (b is already the new one)
                ws[jj - 1] = dual_of(jj, nsetp + 1);
            }
            if (tid == 0) si[14] = 1;
        }
        __syncthreads();
        {
            const int nsetp = si[2], j = si[3];
            if (tid > nsetp && tid <= m && nsetp != m) AL(tid, j) = 0;
            if (tid == 0) ws[j - 1] = 0;
        }
        __syncthreads();
        if (wave == 0) back_subst();
        __syncthreads();
        // secondary loop
        for (;;) {
            if (wave == 0) {
                const int nsetp = si[2];
                const double iter = sd[17] + 1.0;
                int flag = 0;   // 1: leave the secondary loop, 2: iteration cap
                if (iter > (double)max_iter) {
                    flag = 2;
                } else {
                    // alpha = the first minimum of -x / (z - x) over the
                    // non-positive z in P order, if below 2
                    double t = 2.0;
                    int l = 0;
                    if (lane < nsetp) {
                        l = idx[lane];
                        const double z = zz[lane];
                        if (z <= 0) t = -xs[l - 1] / (z - xs[l - 1]);
                    }
                    const bool cand = t < 2.0;
                    double tv = cand ? t : 2.0;
                    const double alpha = mz_wave_min(tv);
                    const uint64_t hit = __ballot(cand && t == alpha);
                    if (!hit) {
                        flag = 1;
                    } else {
                        const int jmin = __ffsll((unsigned long long)hit) - 1;   // 0-based ip - 1
                        if (lane < nsetp) xs[l - 1] = xs[l - 1] + alpha * (zz[lane] - xs[l - 1]);
                        if (lane == 0) {
                            si[4] = jmin + 1;
                            si[9] = idx[jmin];
                        }
                    }
                }
                if (lane == 0) {
                    sd[17] = iter;
                    si[8] = flag;
                    if (flag == 2) si[7] = 3;
                }
            }
            __syncthreads();
            if (si[8] == 2) { terminated = true; break; }
            if (si[8] == 1) break;
            // move coefficient i from P to Z (Givens rotations), repeat while
            // some remaining coefficient is not positive
            for (;;) {
                if (tid == 0) {
                    xs[si[9] - 1] = 0;
                    si[12] = si[4] != si[2];
                    if (si[12]) si[4] += 1;   // jj += 1 persists across passes, as in nnls.jl:449-451
                    si[14] = 0;
                }
                __syncthreads();
                if (si[12]) {
                    const int jj0 = si[4], nsetp = si[2];
                    for (int jx = jj0; jx <= nsetp; ++jx) {
                        if (tid == 0) {
                            const int ii = idx[jx - 1];
                            idx[jx - 2] = ii;
                            const double a = AL(jx - 1, ii), bqv = AL(jx, ii);
                            double cc, ss, sig;
                            if (fabs(a) > fabs(bqv)) {
                                const double xr = bqv / a, yr = sqrt(1 + xr * xr);
                                cc = (1 / yr) * (a > 0 ? 1.0 : (a < 0 ? -1.0 : 0.0));
                                ss = cc * xr;
                                sig = fabs(a) * yr;
                            } else if (bqv != 0) {
                                const double xr = a / bqv, yr = sqrt(1 + xr * xr);
                                ss = (1 / yr) * (bqv > 0 ? 1.0 : -1.0);
                                cc = ss * xr;
                                sig = fabs(bqv) * yr;
                            } else {
                                sig = 0.0; cc = 0.0; ss = 1.0;
                            }
                            AL(jx - 1, ii) = sig;
                            AL(jx, ii) = 0;
                            sd[18] = cc; sd[19] = ss;
                            si[10] = ii;
                            const double temp = bq[jx - 2];
                            bq[jx - 2] = cc * temp + ss * bq[jx - 1];
                            bq[jx - 1] = -ss * temp + cc * bq[jx - 1];
                        }
                        __syncthreads();
                        const double cc = sd[18], ss = sd[19];
                        const int ii = si[10];
                        for (int l = 1 + tid; l <= n; l += NF_COLS) {
                            if (l != ii) {
                                const double temp = AL(jx - 1, l);
                                AL(jx - 1, l) = cc * temp + ss * AL(jx, l);
                                AL(jx, l) = -ss * temp + cc * AL(jx, l);
                            }
                        }
                        __syncthreads();
                    }
                }
                if (wave == 0) {
                    // nsetp -= 1, iz1 -= 1, the removed index back to Z; then
                    // the first coefficient of P that is not positive, if any
                    const int nsetp = si[2] - 1, iz1 = si[0] - 1;
                    const int32_t i0 = si[9];
                    const bool bad = lane < nsetp && xs[idx[lane] - 1] <= 0;
                    const uint64_t badm = __ballot(bad);
                    if (lane == 0) {
                        si[2] = nsetp;
                        si[0] = iz1;
                        idx[iz1 - 1] = i0;
                        si[11] = badm ? 0 : 1;
                        if (badm) si[9] = idx[__ffsll((unsigned long long)badm) - 1];
                        else if (nsetp >= 1) si[9] = idx[nsetp - 1];
                    }
                }
                __syncthreads();
                if (si[11]) break;
            }
            if (tid < m) zz[tid] = bq[tid];
            __syncthreads();
            if (wave == 0) back_subst();
            __syncthreads();
        }
        if (terminated) break;
        if (tid < si[2]) xs[idx[tid] - 1] = zz[tid];
        __syncthreads();
    }
    // ---- compute_post_merge_particles_nnls! (:1334-1404)
    __shared__ int32_t wcnt[4];
    __shared__ double wsum[4];
    double xj = 0.0;
    if constexpr (MIXED) {
        // The search's passive set P solved again in FP64 from the exact
        // columns: normal equations G y = A_P^T b by Cholesky, one refinement
        // step with the FP64 residual; the merge stands if y > 0 and the
        // residual is within max_err, otherwise the cell goes to the FP64
        // method (status 2).
        __shared__ double gy[ROWS + 1], gr[ROWS + 1];
        constexpr int NPP = ROWS + 1;
        static_assert((ROWS * NPP + NPP * NPP) * sizeof(double) <= sizeof(lds_mat), "system fits");
        double *AP = reinterpret_cast<double *>(lds_mat);   // [row][NPP]
        double *G = AP + ROWS * NPP;                          // [NPP][NPP]
        int32_t *ppos = reinterpret_cast<int32_t *>(ws);    // column -> position in P
        const int p = si[2];
        if (tid < n) ppos[tid] = -1;
        __syncthreads();
        if (tid < p) ppos[idx[tid] - 1] = tid;
        __syncthreads();
        const int pc = tid < n ? ppos[tid] : -1;
        if (pc >= 0) {
            for (int q = 0; q < m; ++q) {
                const double t = nnls_entry(A, kind, q, cx, cy, cz, px, py, pz);
                AP[q * NPP + pc] = (t * rowscale[q]) * cinv;
            }
        }
        __syncthreads();
        for (int e = tid; e < p * p; e += NF_COLS) {
            const int i = e / p, jx = e % p;
            if (jx < i) continue;
            double sm = 0.0;
            for (int q = 0; q < m; ++q) sm = sm + AP[q * NPP + i] * AP[q * NPP + jx];
            G[i * NPP + jx] = sm;
            G[jx * NPP + i] = sm;
        }
        __syncthreads();
        if (wave == 0) {
            // Cholesky, lane = row, lower triangle in place
            bool spd = true;
            for (int k = 0; k < p; ++k) {
                const double dkk = G[k * NPP + k];
                if (!(dkk > 0.0)) spd = false;
                const double d = sqrt(dkk);
                if (lane == k) G[k * NPP + k] = d;
                if (lane > k && lane < p) {
                    const double lik = G[lane * NPP + k] / d;
                    G[lane * NPP + k] = lik;
                    for (int jx = k + 1; jx <= lane; ++jx) G[lane * NPP + jx] -= lik * G[jx * NPP + k];
                }
            }
            // G y = A_P^T r by the two triangular solves; r starts as b
            auto solve = [&](double rhs) {   // lane t < p holds rhs_t, returns the solution component
                double zr = lane < p ? rhs : 0.0;
                for (int k = 0; k < p; ++k) {
                    const double zk = mz_readlane_d(zr, k) / G[k * NPP + k];
                    if (lane == k) zr = zk;
                    if (lane > k && lane < p) zr -= G[lane * NPP + k] * zk;
                }
                for (int k = p - 1; k >= 0; --k) {
                    const double yk = mz_readlane_d(zr, k) / G[k * NPP + k];
                    if (lane == k) zr = yk;
                    if (lane < k) zr -= G[k * NPP + lane] * yk;
                }
                return zr;
            };
            // residual b - A_P y (lane = row) and A_P^T of it (lane = column of P)
            auto resid = [&](double yl, double &rq) {
                if (lane < p) gy[lane] = yl;
                rq = 0.0;
                if (lane < m) {
                    double sm = b0[lane];
                    for (int t = 0; t < p; ++t) sm = sm - AP[lane * NPP + t] * gy[t];
                    rq = sm;
                    gr[lane] = sm;
                }
                double g = 0.0;
                if (lane < p)
                    for (int q = 0; q < m; ++q) g = g + AP[q * NPP + lane] * gr[q];
                return g;
            };
            double rq;
            double y = 0.0;
            if (spd) {
                y = solve(resid(0.0, rq));
                const double dy = solve(resid(y, rq));
                y = y + dy;
            }
            (void)resid(y, rq);
            const double rnorm = sqrt(mz_wave_sum(rq * rq));
            const bool pos = __ballot(lane < p && !(y > 0.0)) == 0;
            if (lane < p) gy[lane] = y;
            if (lane == 0) si[13] = (spd && pos && !(rnorm > A.max_err)) ? 1 : 2;
        }
        __syncthreads();
        if (pc >= 0) xj = gy[pc] * cinv;
    } else {
        if (wave == 0) {
            const int nsetp = si[2];
            const double r = (lane >= nsetp && lane < m) ? bq[lane] : 0.0;
            const double rnorm = sqrt(mz_wave_sum(r * r));
            if (lane == 0) si[13] = !(rnorm > A.max_err);
        }
        __syncthreads();
        xj = tid < n ? (double)xs[tid] * cinv : 0.0;
    }
    int ok = si[13] == 1;
    const bool retry = si[13] == 2;
    const bool keep = ok && tid < n && xj > A.w_threshold;
    const uint64_t km = __ballot(keep);
    {
        const double ksum = mz_wave_sum(keep ? xj : 0.0);
        if (lane == 0) { wcnt[wave] = __popcll(km); wsum[wave] = ksum; }
    }
    __syncthreads();
    const int nonzero = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
    const double ndw = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
    if (nonzero >= N) ok = 0;
    const int32_t n_new = ok ? nonzero : N;
    const bool redo = MIXED && !ok && (retry || nonzero >= N);   // the FP64 method decides
    if (ok) {
        int base = 0;
        for (int k = 0; k < wave; ++k) base += wcnt[k];
        if (keep) {
            const int cur = base + __popcll(km & mz_lanemask_lt());
            const int64_t dst = oct_slot(cur, s1, n1, s2);
            A.sp.w[dst] = xj * (W / ndw);
            A.sp.v0[dst] = cvx; A.sp.v1[dst] = cvy; A.sp.v2[dst] = cvz;
            A.sp.x0[dst] = A.pos_ix ? cpx : sd[4];
            A.sp.x1[dst] = A.pos_iy ? cpy : sd[5];
            A.sp.x2[dst] = A.pos_iz ? cpz : sd[6];
        }
        for (int li = n_new + tid; li < N; li += NF_COLS) A.sp.w[oct_slot(li, s1, n1, s2)] = 0.0;
    }
    if (tid == 0 && redo) A.status[c] = 2;
    if (tid == 0 && !redo) {
        A.status[c] = ok ? 1 : -1;
        int32_t *op = A.old_pia + c * 5;
        op[0] = s1; op[1] = n1; op[2] = s2; op[3] = n2; op[4] = ok ? n_new : -1;
        if (ok) {
            const int32_t n_del = N - n_new;
            const int32_t d2 = min(n_del, n2), d1 = n_del - d2;
            const int32_t nn2 = n2 - d2, nn1 = n1 - d1;
            A.sp.pia.n2[c] = nn2;
            A.sp.pia.end2[c] = nn2 > 0 ? s2 + nn2 - 1 : -1;
            if (nn2 == 0) A.sp.pia.start2[c] = 0;
            A.sp.pia.n1[c] = nn1;
            A.sp.pia.end1[c] = nn1 > 0 ? s1 + nn1 - 1 : -1;
            if (nn1 == 0) A.sp.pia.start1[c] = 0;
            A.sp.pia.nloc[c] = nn1 + nn2;
            if (!(c == A.n_cells - 1) || (n_del > n2)) A.sp.ctr->contiguous = 0;
        }
    }
#undef RED
#undef AL
}

// ===========================================================================
// Candidate NNLS merging with one wavefront per cell (method 3).  The
// moments, the right-hand side and the candidate columns are built in FP64
// as in k_nnls_fast.  The Lawson-Hanson search (nnls.jl:273-548) then runs
// on the scaled moment matrix rounded to FP32 and held in registers: lane l
// keeps columns l, l + 64, l + 128 and l + 192, NW_ROWS values each, so the
// loops over columns (duals, Householder updates, Givens rotations) are
// register arithmetic with the Householder vector, the right-hand side and
// the rotations broadcast from LDS.  The triangular factor R of the passive
// columns sits in LDS, one slot of NW_RS floats per passive column; the
// right-hand side, the solution and the Householder scalars are FP64 with
// one lane per row.  The final passive set P is solved again in FP64 from
// the exact columns (normal equations, Cholesky, one refinement step); a
// cell whose FP64 solution is not positive, whose residual exceeds max_err
// or which would keep all its particles goes to the FP64 method (status 2).
// Two choices differ from the reference's sequence: a coefficient found
// infeasible after a step leaves P from its own position (Lawson and
// Hanson's step 11), and ties between duals go to the smaller column.
#define NW_ROWS 35
#define NW_RS 36          // LDS stride of an R slot and length of the broadcast vectors
#define NW_K 4            // columns per lane
#define NW_RB 4           // rows per block of the register loops

struct NwCand {           // a candidate: where it sits, and its moments' arguments
    double vx, vy, vz, xx, xy, xz;   // velocity and position written for a merged particle
    double cx, cy, cz, px, py, pz;   // relative to the cell's mean velocity and position
    int kind;                        // as nnls_entry
};

// Candidate column j of a cell (particle, fictitious particle or random
// pair), as the column set-up of k_nnls_fast; st holds W, the mean
// velocity, the mean position, the velocity bounds and the deviations.
__device__ inline NwCand nw_candidate(const NnlsArgs &A, int j, int c, int32_t N, int32_t Np, int n_add, int32_t s1,
                                      int32_t n1, int32_t s2, const double *st) {
    NwCand r;
    r.vx = r.vy = r.vz = 0.0;
    r.xx = st[4]; r.xy = st[5]; r.xz = st[6];
    r.cx = r.cy = r.cz = r.px = r.py = r.pz = 0.0;
    r.kind = 0;
    if (j < Np) {
        const int64_t p = oct_slot(j, s1, n1, s2);
        r.vx = A.sp.v0[p]; r.vy = A.sp.v1[p]; r.vz = A.sp.v2[p];
        r.xx = A.sp.x0[p]; r.xy = A.sp.x1[p]; r.xz = A.sp.x2[p];
        r.cx = r.vx - st[1]; r.cy = r.vy - st[2]; r.cz = r.vz - st[3];
        r.px = r.xx - st[4]; r.py = r.xy - st[5]; r.pz = r.xz - st[6];
    } else if (j >= Np + n_add) {
        double vm[3] = {0, 0, 0}, xm[3] = {0, 0, 0};
        r.kind = nnls_pair(A, mz_rng_resolve(A.rp), c, j - Np - n_add, N, s1, n1, s2, vm, xm) ? 3 : 1;
        r.vx = vm[0]; r.vy = vm[1]; r.vz = vm[2];
        r.xx = xm[0]; r.xy = xm[1]; r.xz = xm[2];
        if (r.kind == 3) {
            r.cx = r.vx - st[1]; r.cy = r.vy - st[2]; r.cz = r.vz - st[3];
            r.px = r.xx - st[4]; r.py = r.xy - st[5]; r.pz = r.xz - st[6];
        }
    } else {
        const double e3[3] = {st[13], st[14], st[15]};
        const double mnv[3] = {st[7], st[8], st[9]}, mxv[3] = {st[10], st[11], st[12]};
        r.kind = nf_fict(A, j - Np, e3, mnv, mxv, r.cx, r.cy, r.cz) ? 1 : 2;
        r.vx = st[1] + r.cx; r.vy = st[2] + r.cy; r.vz = st[3] + r.cz;
    }
    return r;
}

// Multi-index q of the 35 of total order <= 4 in graded order: order 0, then
// for each order i the pairs (a, b) with a, then b, ascending, c = i - a - b.
__host__ __device__ constexpr int nw_canon(int q, int d) {
    if (q == 0) return 0;
    int k = 1;
    for (int i = 1; i <= 4; ++i)
        for (int a = 0; a <= i; ++a)
            for (int b = 0; b <= i - a; ++b) {
                if (k == q) return d == 0 ? a : (d == 1 ? b : i - a - b);
                ++k;
            }
    return -1;
}

template <typename T, int... Q>
__device__ __forceinline__ void nw_mono4_rows(std::integer_sequence<int, Q...>, const T (&px)[5], const T (&py)[5],
                                              const T (&pz)[5], T (&out)[NW_ROWS]) {
    ((out[Q] = (px[std::integral_constant<int, nw_canon(Q, 0)>::value] *
                py[std::integral_constant<int, nw_canon(Q, 1)>::value]) *
               pz[std::integral_constant<int, nw_canon(Q, 2)>::value]),
     ...);
}

// The 35 monomials x^a y^b z^c in graded order with constant exponents, each
// formed as nnls_entry forms it: powers by repeated multiplication, then
// (x^a y^b) z^c.
__device__ __forceinline__ void nw_mono4(double x, double y, double z, double (&out)[NW_ROWS]) {
    static_assert(NW_ROWS == 35, "35 multi-indices of total order <= 4");
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double x3 = x2 * x, y3 = y2 * y, z3 = z2 * z;
    const double px[5] = {1.0, x, x2, x3, x3 * x}, py[5] = {1.0, y, y2, y3, y3 * y}, pz[5] = {1.0, z, z2, z3, z3 * z};
    nw_mono4_rows(std::make_integer_sequence<int, NW_ROWS>{}, px, py, pz, out);
}
template <int... Q>
__device__ __forceinline__ void nw_mono4_acc_rows(std::integer_sequence<int, Q...>, const double (&px)[5],
                                                  const double (&py)[5], const double (&pz)[5], double w,
                                                  double (&acc)[NW_ROWS]) {
    ((acc[Q] += w * ((px[std::integral_constant<int, nw_canon(Q, 0)>::value] *
                      py[std::integral_constant<int, nw_canon(Q, 1)>::value]) *
                     pz[std::integral_constant<int, nw_canon(Q, 2)>::value])),
     ...);
}
// acc[q] += w x^a y^b z^c over the 35 multi-indices, each monomial used as it
// is formed
__device__ __forceinline__ void nw_mono4_acc(double x, double y, double z, double w, double (&acc)[NW_ROWS]) {
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double x3 = x2 * x, y3 = y2 * y, z3 = z2 * z;
    const double px[5] = {1.0, x, x2, x3, x3 * x}, py[5] = {1.0, y, y2, y3, y3 * y}, pz[5] = {1.0, z, z2, z3, z3 * z};
    nw_mono4_acc_rows(std::make_integer_sequence<int, NW_ROWS>{}, px, py, pz, w, acc);
}
template <int... Q>
__device__ __forceinline__ void nw_mono4_scaled_rows(std::integer_sequence<int, Q...>, const double (&px)[5],
                                                     const double (&py)[5], const double (&pz)[5], const double *rs,
                                                     double (&out)[NW_ROWS]) {
    ((out[Q] = ((px[std::integral_constant<int, nw_canon(Q, 0)>::value] *
                 py[std::integral_constant<int, nw_canon(Q, 1)>::value]) *
                pz[std::integral_constant<int, nw_canon(Q, 2)>::value]) *
               rs[Q]),
     ...);
}
// out[q] = x^a y^b z^c rs[q] over the 35 multi-indices
__device__ __forceinline__ void nw_mono4_scaled(double x, double y, double z, const double *rs,
                                                double (&out)[NW_ROWS]) {
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double x3 = x2 * x, y3 = y2 * y, z3 = z2 * z;
    const double px[5] = {1.0, x, x2, x3, x3 * x}, py[5] = {1.0, y, y2, y3, y3 * y}, pz[5] = {1.0, z, z2, z3, z3 * z};
    nw_mono4_scaled_rows(std::make_integer_sequence<int, NW_ROWS>{}, px, py, pz, rs, out);
}
template <int... Q>
__device__ __forceinline__ void nw_mono4_dual_rows(std::integer_sequence<int, Q...>, const double (&px)[5],
                                                   const double (&py)[5], const double (&pz)[5], const double *rs,
                                                   const double *r, double &nn, double &wr) {
    (([&] {
         const double t = ((px[std::integral_constant<int, nw_canon(Q, 0)>::value] *
                            py[std::integral_constant<int, nw_canon(Q, 1)>::value]) *
                           pz[std::integral_constant<int, nw_canon(Q, 2)>::value]) *
                          rs[Q];
         nn = fma(t, t, nn);
         wr = fma(t, r[Q], wr);
     }()),
     ...);
}
// sum of squares and of products with r of the row-scaled monomials, each
// used as it is formed
__device__ __forceinline__ void nw_mono4_dual(double x, double y, double z, const double *rs, const double *r,
                                              double &nn, double &wr) {
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double x3 = x2 * x, y3 = y2 * y, z3 = z2 * z;
    const double px[5] = {1.0, x, x2, x3, x3 * x}, py[5] = {1.0, y, y2, y3, y3 * y}, pz[5] = {1.0, z, z2, z3, z3 * z};
    nw_mono4_dual_rows(std::make_integer_sequence<int, NW_ROWS>{}, px, py, pz, rs, r, nn, wr);
}
// the same in FP32 (the search's copy of the matrix)
__device__ __forceinline__ void nw_mono4(double xd, double yd, double zd, float (&out)[NW_ROWS]) {
    const float x = (float)xd, y = (float)yd, z = (float)zd;
    const float x2 = x * x, y2 = y * y, z2 = z * z;
    const float x3 = x2 * x, y3 = y2 * y, z3 = z2 * z;
    const float px[5] = {1.0f, x, x2, x3, x3 * x}, py[5] = {1.0f, y, y2, y3, y3 * y},
                pz[5] = {1.0f, z, z2, z3, z3 * z};
    nw_mono4_rows(std::make_integer_sequence<int, NW_ROWS>{}, px, py, pz, out);
}

__device__ inline float nw_readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Row exponents of the search matrix are read from LDS, packed a | b << 8 |
// c << 16 per row: ev for the velocity rows, ep for the spatial rows of
// particles, er for the spatial rows of random pairs (the velocity
// multi-indices, as nnls_entry); indexing the argument block by a row number
// inside the unrolled row loops would make the compiler copy the whole block
// to scratch.
__device__ inline int32_t nw_pack3(const int32_t *e) { return e[0] | (e[1] << 8) | (e[2] << 16); }
__device__ inline double nw_pow3(int32_t e, double x, double y, double z) {
    e = __builtin_amdgcn_readfirstlane(e);
    return nnls_ipow(x, e & 255) * nnls_ipow(y, (e >> 8) & 255) * nnls_ipow(z, e >> 16);
}
// nnls_entry with the exponents from the LDS tables
__device__ inline double nw_entry(const int32_t *ev, const int32_t *ep, const int32_t *er, int m_v, int kind, int q,
                                  double cx, double cy, double cz, double px, double py, double pz) {
    if (kind == 1) return 0.0;
    if (q < m_v) return nw_pow3(ev[q], cx, cy, cz);
    if (kind == 0) return nw_pow3(ep[q], px, py, pz);
    if (kind == 3) return nw_pow3(er[q], px, py, pz);
    return 0.0;
}

// The exact FP64 column of a candidate, row-scaled and of unit norm (the
// column the FP32 search holds rounded); returns 1 / its norm before
// normalisation.
__device__ __forceinline__ double nw_exact_column(const NnlsArgs &A, const NwCand &cd, const double *rs,
                                                  const int32_t *ev, const int32_t *ep, const int32_t *er, int m,
                                                  int m_v, double (&col)[NW_ROWS]) {
    double nn = 0.0;
    if (A.canon) {
        if (cd.kind == 1) {
#pragma unroll
            for (int q = 0; q < NW_ROWS; ++q) col[q] = 0.0;
        } else {
            nw_mono4_scaled(cd.cx, cd.cy, cd.cz, rs, col);
#pragma unroll
            for (int q = 0; q < NW_ROWS; ++q) nn = nn + col[q] * col[q];
        }
    } else {
#pragma unroll
        for (int q = 0; q < NW_ROWS; ++q) {
            col[q] = q < m ? nw_entry(ev, ep, er, m_v, cd.kind, q, cd.cx, cd.cy, cd.cz, cd.px, cd.py, cd.pz) * rs[q]
                           : 0.0;
            nn = nn + col[q] * col[q];
        }
    }
    nn = sqrt(nn);
    nn = nn > 1e-15 ? nn : 1.0;
    const double ci = 1.0 / nn;
#pragma unroll
    for (int q = 0; q < NW_ROWS; ++q) col[q] *= ci;
    return ci;
}

// Column j of the search matrix: the candidate's moments, row-scaled by rs,
// unit norm (scale_columns!, utils.jl:62-77), rounded to FP32; zero for
// j >= n.  Clears the column's x.
__device__ __forceinline__ void nw_build_column(const NnlsArgs &A, float (&col)[NW_ROWS], int j, int n, int c,
                                                int32_t N, int32_t Np, int n_add, int32_t s1, int32_t n1,
                                                int32_t s2, const double *st, const double *rs, double *xs,
                                                const int32_t *ev, const int32_t *ep, const int32_t *er, int m,
                                                int m_v) {
    double nn = 0.0;
    if (j < n) {
        const NwCand cd = nw_candidate(A, j, c, N, Np, n_add, s1, n1, s2, st);
        if (A.canon) {
            // the search's copy in FP32 throughout (the FP64 solve rebuilds its columns)
            float mo[NW_ROWS];
            nw_mono4(cd.cx, cd.cy, cd.cz, mo);
            float nf = 0.0f;
#pragma unroll
            for (int q = 0; q < NW_ROWS; ++q) {
                const float t = (cd.kind == 1 ? 0.0f : mo[q]) * (float)rs[q];
                col[q] = t;
                nf = __builtin_fmaf(t, t, nf);
            }
            nn = nf;
        } else {
#pragma unroll
            for (int q = 0; q < NW_ROWS; ++q) {
                double t = 0.0;
                if (q < m) t = nw_entry(ev, ep, er, m_v, cd.kind, q, cd.cx, cd.cy, cd.cz, cd.px, cd.py, cd.pz) * rs[q];
                col[q] = (float)t;
                nn = nn + t * t;
            }
        }
        xs[j] = 0.0;
    } else {
#pragma unroll
        for (int q = 0; q < NW_ROWS; ++q) col[q] = 0.0f;
    }
    nn = sqrt(nn);
    nn = nn > 1e-15 ? nn : 1.0;
    const float ci = (float)(1.0 / nn);
#pragma unroll
    for (int q = 0; q < NW_ROWS; ++q) col[q] *= ci;
}

// Dual r . a_j of candidate column j when in_z: the largest goes to wmax
// (its column to jmax), the largest norm before normalisation (zero columns
// aside) to nmax.
__device__ __forceinline__ void nw_kkt_dual(const NnlsArgs &A, bool in_z, int j, int c, int32_t N, int32_t Np,
                                            int n_add, int32_t s1, int32_t n1, int32_t s2, const double *st,
                                            const double *rs, const int32_t *ev, const int32_t *ep,
                                            const int32_t *er, int m, int m_v, const double *r, double &wmax,
                                            int &jmax, double &nmax) {
    if (!in_z) return;
    const NwCand cd = nw_candidate(A, j, c, N, Np, n_add, s1, n1, s2, st);
    if (cd.kind == 1) return;   // a zero column: dual 0
    double nn = 0.0, wr = 0.0;
    if (A.canon) {
        nw_mono4_dual(cd.cx, cd.cy, cd.cz, rs, r, nn, wr);
    } else {
        for (int q = 0; q < m; ++q) {
            const double t = nw_entry(ev, ep, er, m_v, cd.kind, q, cd.cx, cd.cy, cd.cz, cd.px, cd.py, cd.pz) * rs[q];
            nn = fma(t, t, nn);
            wr = fma(t, r[q], wr);
        }
    }
    nn = sqrt(nn);
    nn = nn > 1e-15 ? nn : 1.0;
    const double w = wr / nn;
    if (w > wmax) {
        wmax = w;
        jmax = j;
    }
    nmax = fmax(nmax, nn);
}

__global__ void __launch_bounds__(64, 2) k_nnls_wave(NnlsArgs A) {
    constexpr int R = NW_ROWS, RS = NW_RS;
    // the FP64 passive-set system (A_P [row][position], two vectors) aliases
    // the search's x by column (FP64) and R by slot (FP32)
    constexpr int POL = R * R + 2 * RS;
    constexpr int SRCH = NF_COLS + (R * RS + 1) / 2;
    __shared__ __align__(16) double s_area[POL > SRCH ? POL : SRCH];
    __shared__ __align__(16) float s_u[RS], s_bz[RS], s_col[RS];
    __shared__ float2 s_cs[RS];
    __shared__ int32_t s_pcol[RS], s_pslot[RS];
    __shared__ double s_rs[RS], s_gy[RS], s_gr[RS], s_st[16];
    __shared__ int32_t s_ev[RS], s_ep[RS], s_er[RS];
    double *s_xs = s_area;
    float *s_R = reinterpret_cast<float *>(s_area + NF_COLS);   // R by slot: [slot][row]
    const int c = A.cells[blockIdx.x];
    const int lane = threadIdx.x;
    const int32_t s1 = A.sp.pia.start1[c], n1 = A.sp.pia.n1[c], s2 = A.sp.pia.start2[c], n2 = A.sp.pia.n2[c];
    const int32_t N = n1 + n2;
    const int32_t Np = nnls_pcols(A, N);
    const int m = A.m;
    const int n_add = (A.centered ? 1 : 0) + 8 * A.n_vmult;
    const int n = Np + n_add + A.n_rand;
    const int m_v = A.m_v;
    if (lane < m) {
        if (lane < m_v) {
            s_ev[lane] = nw_pack3(A.mim[lane]);
        } else {
            s_ep[lane] = nw_pack3(A.mim_pos[lane - m_v]);
            s_er[lane] = nw_pack3(A.mim[lane - m_v]);
        }
    }
    // ---- compute_w_total_v0!: W, mean velocity and position, velocity bounds
    {
        double Ws = 0.0, sv[3] = {0, 0, 0}, sx[3] = {0, 0, 0};
        double mn[3] = {299792458.0, 299792458.0, 299792458.0};
        double mx[3] = {-299792458.0, -299792458.0, -299792458.0};
        for (int li = lane; li < N; li += 64) {
            const int64_t p = oct_slot(li, s1, n1, s2);
            const double w = A.sp.w[p], v[3] = {A.sp.v0[p], A.sp.v1[p], A.sp.v2[p]};
            const double xx[3] = {A.sp.x0[p], A.sp.x1[p], A.sp.x2[p]};
            Ws += w;
            for (int d = 0; d < 3; ++d) {
                sv[d] += w * v[d];
                sx[d] += w * xx[d];
                mn[d] = fmin(mn[d], v[d]);
                mx[d] = fmax(mx[d], v[d]);
            }
        }
        const double Wt = mz_wave_sum(Ws);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double v0 = mz_wave_sum(sv[d]) / Wt, x0 = mz_wave_sum(sx[d]) / Wt;
            const double lo = mz_wave_min(mn[d]) - v0, hi = mz_wave_max(mx[d]) - v0;
            if (lane == 0) { s_st[1 + d] = v0; s_st[4 + d] = x0; s_st[7 + d] = lo; s_st[10 + d] = hi; }
        }
        if (lane == 0) s_st[0] = Wt;
    }
    oct_wave_sync();
    const double W = s_st[0];
    // ---- variances and right-hand side (compute_lhs_and_rhs!, :509-613), row scaling (:1178-1215)
    double bq = 0.0;   // lane q < m: row q of the scaled right-hand side, then of Q^T b
    {
        const double v0x = s_st[1], v0y = s_st[2], v0z = s_st[3], x0x = s_st[4], x0y = s_st[5], x0z = s_st[6];
        double acc[R], e[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < R; ++q) acc[q] = 0.0;
#pragma unroll 1
        for (int li = lane; li < N; li += 64) {
            const int64_t p = oct_slot(li, s1, n1, s2);
            const double w = A.sp.w[p];
            const double cx = A.sp.v0[p] - v0x, cy = A.sp.v1[p] - v0y, cz = A.sp.v2[p] - v0z;
            const double px = A.sp.x0[p] - x0x, py = A.sp.x1[p] - x0y, pz = A.sp.x2[p] - x0z;
            e[0] += w * (cx * cx);
            e[1] += w * (cy * cy);
            e[2] += w * (cz * cz);
            e[3] += w * (px * px);
            e[4] += w * (py * py);
            e[5] += w * (pz * pz);
            if (A.canon) {
                nw_mono4_acc(cx, cy, cz, w, acc);
            } else {
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    if (q < m) acc[q] += w * nw_entry(s_ev, s_ep, s_er, m_v, 0, q, cx, cy, cz, px, py, pz);
                }
            }
        }
        double Es[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) Es[k] = sqrt(mz_wave_sum(e[k]) / W);
        double mom = 0.0;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            if (q < m) {
                const double t = mz_wave_sum(acc[q]);
                if (lane == q) mom = t / W;
            }
        }
        if (lane == 0) { s_st[13] = Es[0]; s_st[14] = Es[1]; s_st[15] = Es[2]; }
        if (lane < m) {
            const double rv = nnls_rowscale(A, lane, Es[0], Es[1], Es[2], Es[3], Es[4], Es[5]);
            s_rs[lane] = rv;
            bq = mom * rv;
        }
    }
    const double b0 = bq;
    oct_wave_sync();
    // ---- the candidate columns: row-scaled, unit norm (scale_columns!, utils.jl:62-77), rounded to FP32
    float a[NW_K][R];
    float wd[NW_K];     // dual of each column
    int ppos[NW_K];     // position in P; -1 in Z; -2 no column
#pragma unroll
    for (int k = 0; k < NW_K; ++k) {
        wd[k] = 0.0f;
        ppos[k] = k * 64 + lane < n ? -1 : -2;
    }
    // one call per register slot (a loop over the slots is too long for the
    // compiler to unroll, and a[k] must keep constant indices)
    static_assert(NW_K == 4, "one call per register slot");
    nw_build_column(A, a[0], lane, n, c, N, Np, n_add, s1, n1, s2, s_st, s_rs, s_xs, s_ev, s_ep, s_er, m,
                    m_v);
    nw_build_column(A, a[1], 64 + lane, n, c, N, Np, n_add, s1, n1, s2, s_st, s_rs, s_xs, s_ev, s_ep, s_er, m,
                    m_v);
    nw_build_column(A, a[2], 128 + lane, n, c, N, Np, n_add, s1, n1, s2, s_st, s_rs, s_xs, s_ev, s_ep, s_er, m,
                    m_v);
    nw_build_column(A, a[3], 192 + lane, n, c, N, Np, n_add, s1, n1, s2, s_st, s_rs, s_xs, s_ev, s_ep, s_er, m,
                    m_v);
    if (lane < RS) {
        s_pslot[lane] = lane;
        s_pcol[lane] = 0;
        s_u[lane] = 0.0f;
        s_bz[lane] = 0.0f;
        s_col[lane] = 0.0f;
    }
    oct_wave_sync();

    // ---- Lawson-Hanson solve!(work, iteration_mult * n, 1e-14)
    const double kkt_tol = 1e-14, factor = 0.01;
    const int64_t max_iter = (int64_t)A.iteration_mult * n;
    int nsetp = 0;
    int64_t iter = 0;
    double up = 0.0;
    double z = 0.0;   // lane t < nsetp: solution of R z = (Q^T b)[0, nsetp) at position t
    bool duals_ok = false;
    // R by rows in registers: lane r holds row r of R over the positions of P
    // (Rrow[t] = R[r][t]), so that the back substitution reads no LDS and, with
    // its steps unrolled, indexes the row by constants.  LDS keeps R by slot
    // for the Givens rotations of a removal, after which the rows are read
    // again.
    float Rrow[R];
#pragma unroll
    for (int t = 0; t < R; ++t) Rrow[t] = 0.0f;
    auto back_subst = [&]() {
        const int ns = nsetp;
        float dg = 1.0f;   // this lane's diagonal entry R[lane][lane]
        const uint64_t lbit = 1ull << lane;
#pragma unroll
        for (int t = 0; t < R; ++t)
            if ((lbit >> t) & 1ull) dg = Rrow[t];
        const double rdk = lane < ns ? 1.0 / (double)dg : 0.0;
        double y = lane < ns ? bq : 0.0;
#pragma unroll
        for (int ip = R - 1; ip >= 0; --ip) {
            if (ip < ns) {
                if (lane == ip) y = y * rdk;
                const double zip = mz_readlane_d(y, ip);
                if (lane < ip) y = y - (double)Rrow[ip] * zip;
            }
        }
        z = y;
    };
    auto reload_rows = [&]() {
        const int ns = nsetp;
        const int sl = lane < ns ? s_pslot[lane] : 0;
        const int rl = lane < R ? lane : 0;
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const float v = s_R[__builtin_amdgcn_readlane(sl, t) * RS + rl];
            Rrow[t] = (t < ns && lane < R) ? v : 0.0f;
        }
    };
    // duals of the zero set over rows >= lo: w = A_Z^T (Q^T b) with rows < lo zero
    auto duals = [&](int lo) {
        float d[NW_K] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int qb = 0; qb < RS; qb += NW_RB) {
            if (qb + NW_RB > lo) {
                const float4 bz = *reinterpret_cast<const float4 *>(s_bz + qb);
                const float bv[NW_RB] = {bz.x, bz.y, bz.z, bz.w};
#pragma unroll
                for (int k = 0; k < NW_K; ++k)
#pragma unroll
                    for (int r = 0; r < NW_RB; ++r)
                        if (qb + r < R) d[k] = __builtin_fmaf(a[k][qb + r], bv[r], d[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < NW_K; ++k)
            if (ppos[k] == -1) wd[k] = d[k];
    };
    for (;;) {
        if (nsetp >= n || nsetp >= m) break;
        if (!duals_ok) {
            if (lane < RS) s_bz[lane] = (lane >= nsetp && lane < m) ? (float)bq : 0.0f;
            oct_wave_sync();
            duals(nsetp);
        }
        // the largest positive dual; construct_householder! on rows nsetp..m-1 of
        // its column, the column test and apply_householder! to a copy of b
        int bj = -1;
        double top = 0.0, av = 0.0;
        bool stop = false;
        for (;;) {
            // the lane's best (first maximum over its slots), the wave's maximum,
            // the lowest lane holding it
            float bwl = 0.0f;
            int bkl = -1;
#pragma unroll
            for (int k = 0; k < NW_K; ++k) {
                if (ppos[k] == -1 && wd[k] > bwl) {
                    bwl = wd[k];
                    bkl = k;
                }
            }
            const float bw = __ockl_wfred_max_f32(bwl);
            if (!((double)bw > kkt_tol)) {
                stop = true;
                break;
            }
            const uint64_t hitm = __ballot(bkl >= 0 && bwl == bw);
            const int bl0 = __ffsll((unsigned long long)hitm) - 1;
            bj = __builtin_amdgcn_readlane(bkl, bl0) * 64 + bl0;
            const int bk = bj >> 6, bl = bj & 63;
            // the column's register slot as a mask: a test k == bk would let the
            // compiler index the register array by bk
            const uint32_t bsel = 1u << bk;
#pragma unroll
            for (int k = 0; k < NW_K; ++k) {
                if (((bsel >> k) & 1u) && lane == bl) {
#pragma unroll
                    for (int q = 0; q < R; ++q) s_col[q] = a[k][q];
                }
            }
            oct_wave_sync();
            av = lane < m ? (double)s_col[lane] : 0.0;
            const int len = m - nsetp;
            const bool hh = lane >= nsetp && lane < m;
            const double asave = mz_readlane_d(av, nsetp);
            double upn = up;
            top = asave;
            // the column's norm below and above row nsetp (two independent sums)
            const double sq_lo = mz_wave_sum(hh ? av * av : 0.0);
            const double unorm = sqrt(mz_wave_sum(lane < nsetp ? av * av : 0.0));
            if (len > 1) {
                double cl = sqrt(sq_lo);
                if (asave > 0) cl = -cl;
                upn = asave - cl;
                top = cl;
            }
            bool accept = false;
            double zr = lane < m ? bq : 0.0;
            if (((unorm + fabs(top) * factor) - unorm) > 0) {
                if (len > 1) {
                    double bb = upn * top;
                    if (bb < 0) {
                        bb = 1.0 / bb;
                        const double ur = lane == nsetp ? upn : (hh ? av : 0.0);
                        double sm = mz_wave_sum(zr * ur);
                        if (sm != 0) {
                            sm *= bb;
                            if (hh) zr = zr + sm * ur;
                        }
                    }
                }
                const double ztest = mz_readlane_d(zr, nsetp) / top;
                accept = ztest > 0;
            }
            up = upn;
            if (accept) {
                bq = zr;
                break;
            }
#pragma unroll
            for (int k = 0; k < NW_K; ++k)
                if (((bsel >> k) & 1u) && lane == bl) wd[k] = 0.0f;
        }
        if (stop) break;
        // move column bj from Z to P at position p, transform the remaining columns
        const int p = nsetp;
        const int bl = bj & 63;
        const uint32_t bsel = 1u << (bj >> 6);
        nsetp = p + 1;
#pragma unroll
        for (int k = 0; k < NW_K; ++k)
            if (((bsel >> k) & 1u) && lane == bl) ppos[k] = p;
        if (lane < RS) {
            s_u[lane] = lane == p ? (float)up : ((lane > p && lane < m) ? (float)av : 0.0f);
            s_bz[lane] = (lane >= nsetp && lane < m) ? (float)bq : 0.0f;
        }
        if (lane == 0) s_pcol[p] = bj;
        oct_wave_sync();
        {
            const double bbd = up * top;
            const bool apply = m - p > 1 && bbd < 0;
            const float bbf = (float)(1.0 / bbd);
            float sm[NW_K] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (apply) {
#pragma unroll
                for (int qb = 0; qb < RS; qb += NW_RB) {
                    if (qb + NW_RB > p) {
                        const float4 u4 = *reinterpret_cast<const float4 *>(s_u + qb);
                        const float uv[NW_RB] = {u4.x, u4.y, u4.z, u4.w};
#pragma unroll
                        for (int k = 0; k < NW_K; ++k)
#pragma unroll
                            for (int r = 0; r < NW_RB; ++r)
                                if (qb + r < R) sm[k] = __builtin_fmaf(a[k][qb + r], uv[r], sm[k]);
                    }
                }
            }
            float f[NW_K], d[NW_K];
#pragma unroll
            for (int k = 0; k < NW_K; ++k) {
                f[k] = (apply && ppos[k] == -1 && sm[k] != 0.0f) ? sm[k] * bbf : 0.0f;
                d[k] = 0.0f;
            }
#pragma unroll
            for (int qb = 0; qb < RS; qb += NW_RB) {
                if (qb + NW_RB > p) {
                    const float4 u4 = *reinterpret_cast<const float4 *>(s_u + qb);
                    const float4 b4 = *reinterpret_cast<const float4 *>(s_bz + qb);
                    const float uv[NW_RB] = {u4.x, u4.y, u4.z, u4.w};
                    const float bv[NW_RB] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
                    for (int k = 0; k < NW_K; ++k)
#pragma unroll
                        for (int r = 0; r < NW_RB; ++r)
                            if (qb + r < R) {
                                a[k][qb + r] = __builtin_fmaf(f[k], uv[r], a[k][qb + r]);
                                d[k] = __builtin_fmaf(a[k][qb + r], bv[r], d[k]);
                            }
                }
            }
#pragma unroll
            for (int k = 0; k < NW_K; ++k)
                if (ppos[k] == -1) wd[k] = d[k];
        }
        duals_ok = true;
        // column bj: top at row p, zero below; its R column into a free slot
        const int slot = s_pslot[p];
#pragma unroll
        for (int k = 0; k < NW_K; ++k) {
            if (((bsel >> k) & 1u) && lane == bl) {
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const float v = q < p ? a[k][q] : (q == p ? (float)top : 0.0f);
                    a[k][q] = v;
                    s_R[slot * RS + q] = v;
                }
                wd[k] = 0.0f;
            }
        }
        oct_wave_sync();
        {
            const float v = s_R[slot * RS + (lane < R ? lane : 0)];
            const uint64_t pbit = 1ull << p;
#pragma unroll
            for (int t = 0; t < R; ++t)
                if ((pbit >> t) & 1ull) Rrow[t] = lane < R ? v : 0.0f;
        }
        back_subst();
        // secondary loop
        bool capped = false;
        for (;;) {
            ++iter;
            if (iter > max_iter) {
                capped = true;
                break;
            }
            // alpha = the first minimum of -x / (z - x) over the non-positive z
            // in P order, if below 2
            double t = 2.0;
            int l = 0;
            if (lane < nsetp) {
                l = s_pcol[lane];
                if (z <= 0) t = -s_xs[l] / (z - s_xs[l]);
            }
            const bool cand = t < 2.0;
            const double alpha = mz_wave_min(cand ? t : 2.0);
            const uint64_t hit = __ballot(cand && t == alpha);
            if (!hit) break;
            if (lane < nsetp) s_xs[l] = s_xs[l] + alpha * (z - s_xs[l]);
            int rp = __ffsll((unsigned long long)hit) - 1;
            // move the coefficient at position rp from P to Z (Givens rotations
            // close the gap), then any remaining one that is not positive
            for (;;) {
                oct_wave_sync();
                const int i = s_pcol[rp];
                const int last = nsetp - 1;
                if (lane == 0) s_xs[i] = 0.0;
                const int sl = lane < nsetp ? s_pslot[lane] : 0;
                const int pc = lane < nsetp ? s_pcol[lane] : 0;
                if (rp < last) {
                    // rotation pp acts on rows pp - 1 and pp and takes the column at
                    // position pp to (sig, 0) there; lane t > pp rotates its column's
                    // rows, carrying the lower one to the next step
                    float carry = (lane > rp && lane <= last) ? s_R[sl * RS + rp] : 0.0f;
                    for (int pp = rp + 1; pp <= last; ++pp) {
                        const float r1 = (lane >= pp && lane <= last) ? s_R[sl * RS + pp] : 0.0f;
                        const double ra = (double)nw_readlane_f(carry, pp), rb = (double)nw_readlane_f(r1, pp);
                        double cc, ss, sig;
                        if (fabs(ra) > fabs(rb)) {
                            const double xr = rb / ra, yr = sqrt(1 + xr * xr);
                            cc = (1 / yr) * (ra > 0 ? 1.0 : (ra < 0 ? -1.0 : 0.0));
                            ss = cc * xr;
                            sig = fabs(ra) * yr;
                        } else if (rb != 0) {
                            const double xr = ra / rb, yr = sqrt(1 + xr * xr);
                            ss = (1 / yr) * (rb > 0 ? 1.0 : -1.0);
                            cc = ss * xr;
                            sig = fabs(rb) * yr;
                        } else {
                            sig = 0.0; cc = 0.0; ss = 1.0;
                        }
                        if (lane > pp && lane <= last) {
                            const double c0 = carry, c1 = r1;
                            s_R[sl * RS + pp - 1] = (float)(cc * c0 + ss * c1);
                            carry = (float)(-ss * c0 + cc * c1);
                        }
                        if (lane == pp) s_R[sl * RS + pp - 1] = (float)sig;
                        if (lane == 0) s_cs[pp - 1] = make_float2((float)cc, (float)ss);
                        const double t0 = mz_readlane_d(bq, pp - 1), t1 = mz_readlane_d(bq, pp);
                        if (lane == pp - 1) bq = cc * t0 + ss * t1;
                        if (lane == pp) bq = -ss * t0 + cc * t1;
                    }
                    oct_wave_sync();
                    // the same rotations on every column held in registers
#pragma unroll
                    for (int q = 0; q + 1 < R; ++q) {
                        if (q >= rp && q < last) {
                            const float2 cs = s_cs[q];
#pragma unroll
                            for (int k = 0; k < NW_K; ++k) {
                                const float t0 = a[k][q], t1 = a[k][q + 1];
                                a[k][q] = cs.x * t0 + cs.y * t1;
                                a[k][q + 1] = -cs.y * t0 + cs.x * t1;
                            }
                        }
                    }
                }
                // positions after rp move down by one; rp's slot goes to the end
                const int pc_next = __shfl(pc, lane < 63 ? lane + 1 : 63, 64);
                const int sl_next = __shfl(sl, lane < 63 ? lane + 1 : 63, 64);
                const int sl_free = __builtin_amdgcn_readlane(sl, rp);
                oct_wave_sync();
                if (lane >= rp && lane < last) {
                    s_pcol[lane] = pc_next;
                    s_pslot[lane] = sl_next;
                }
                if (lane == last) s_pslot[lane] = sl_free;
#pragma unroll
                for (int k = 0; k < NW_K; ++k) {
                    if (ppos[k] == rp) ppos[k] = -1;
                    else if (ppos[k] > rp) ppos[k] -= 1;
                }
                nsetp = last;
                duals_ok = false;
                oct_wave_sync();
                const bool bad = lane < nsetp && s_xs[s_pcol[lane]] <= 0.0;
                const uint64_t badm = __ballot(bad);
                if (!badm) break;
                rp = __ffsll((unsigned long long)badm) - 1;
            }
            reload_rows();
            back_subst();
        }
        if (capped) break;
        if (lane < nsetp) s_xs[s_pcol[lane]] = z;
        oct_wave_sync();
    }

    // ---- the passive set P solved again in FP64 from the exact columns:
    // Householder QR with one lane per column of P (lane 63 carries b), the
    // back substitution with one lane per row; the residual norm is the
    // solver's, that of the rows of Q^T b below P.
    // compute_post_merge_particles_nnls! (:1334-1404)
    oct_wave_sync();
    const int p = nsetp;
    double *AP = s_area;              // [row][position]: the exact columns of P
    double *s_v = s_area + R * R;     // a column of R, broadcast to the rows
    double *s_uh = s_v + RS;          // the Householder vector
    int jt = NF_COLS;
    double cinv = 0.0;
    NwCand cd{};
    double col[R];                    // lane t < p: column t; lane 63: b
    if (lane < m) s_gr[lane] = b0;
    if (lane < p) {
        jt = s_pcol[lane];
        cd = nw_candidate(A, jt, c, N, Np, n_add, s1, n1, s2, s_st);
        cinv = nw_exact_column(A, cd, s_rs, s_ev, s_ep, s_er, m, m_v, col);
#pragma unroll
        for (int q = 0; q < R; ++q) AP[q * R + lane] = col[q];
    }
    oct_wave_sync();
    if (lane == 63) {
#pragma unroll
        for (int q = 0; q < R; ++q) col[q] = q < m ? s_gr[q] : 0.0;
    } else if (lane >= p) {
#pragma unroll
        for (int q = 0; q < R; ++q) col[q] = 0.0;
    }
    bool rank_ok = true;
    for (int k = 0; k < p; ++k) {
        if (lane == k) {
#pragma unroll
            for (int q = 0; q < R; ++q) s_v[q] = col[q];
        }
        oct_wave_sync();
        const double vk = (lane >= k && lane < m) ? s_v[lane] : 0.0;
        const double nrm = sqrt(mz_wave_sum(vk * vk));
        const double akk = s_v[k];
        const double alpha = akk > 0 ? -nrm : nrm;
        const double uk = akk - alpha;
        if (!(nrm > 0.0)) rank_ok = false;
        if (lane < RS) s_uh[lane] = lane == k ? uk : ((lane > k && lane < m) ? vk : 0.0);
        oct_wave_sync();
        const double beta = rank_ok ? -1.0 / (alpha * uk) : 0.0;
        if ((lane > k && lane < p) || lane == 63) {
            // the Householder vector read eight entries at a time (bounds the registers)
            double d = 0.0;
#pragma unroll
            for (int qb = 0; qb < R; qb += 8) {
                double u8[8];
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    if (qb + i < R) u8[i] = s_uh[qb + i];
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    if (qb + i < R) d = fma(u8[i], col[qb + i], d);
                __builtin_amdgcn_sched_barrier(0);
            }
            d *= beta;
#pragma unroll
            for (int qb = 0; qb < R; qb += 8) {
                double u8[8];
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    if (qb + i < R) u8[i] = s_uh[qb + i];
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    if (qb + i < R) col[qb + i] = fma(-d, u8[i], col[qb + i]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (lane == k) {
            // row k becomes alpha (a mask, not q == k: the row stays a constant index)
            const uint64_t kbit = 1ull << k;
#pragma unroll
            for (int q = 0; q < R; ++q)
                if ((kbit >> q) & 1ull) col[q] = alpha;
        }
    }
    double rr = 0.0;
    if (lane == 63) {
        const uint64_t inP = p >= 64 ? ~0ull : ((1ull << p) - 1);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            if (!((inP >> q) & 1ull) && q < m) rr = fma(col[q], col[q], rr);
            s_gr[q] = col[q];
        }
    }
    oct_wave_sync();
    const double rnorm = sqrt(mz_readlane_d(rr, 63));
    // R y = (Q^T b)[0, p): one lane per row, column ip of R broadcast per step
    double cr = lane < p ? s_gr[lane] : 0.0;
    double y = 0.0;
    for (int ip = p - 1; ip >= 0; --ip) {
        if (lane == ip) {
#pragma unroll
            for (int q = 0; q < R; ++q) s_v[q] = col[q];
        }
        oct_wave_sync();
        const double yi = mz_readlane_d(cr, ip) / s_v[ip];
        if (lane == ip) y = yi;
        if (lane < ip) cr = fma(-s_v[lane], yi, cr);
        oct_wave_sync();
    }
    const bool pos = __ballot(lane < p && !(y > 0.0 && y < 1e300)) == 0;
    bool ok = rank_ok && pos && !(rnorm > A.max_err);
    // A merge that fails on the residual is decided here when the FP64
    // solution certifies that no non-negative solution reaches max_err
    // (otherwise the FP64 method decides).  With r = b - A_P y and the duals
    // w = A^T r, convexity gives for every x* >= 0
    //   |A x* - b|^2 >= |r|^2 - 2 (max(0, max_{j outside P} w_j) |x*|_1 + |w_P . y|),
    // and the mass row bounds |x*|_1 of a minimiser by (b_0 + |r|) max_j |a_j| / rs_0
    // (its entry in column j is rs_0 / |a_j| before normalisation).
    bool certified_fail = false;
    if (!ok && rank_ok && p > 0) {
        if (lane < p) s_gy[lane] = y;
        oct_wave_sync();
        double r = 0.0;
        if (lane < m) {
            r = b0;
            for (int t = 0; t < p; ++t) r = fma(-AP[lane * R + t], s_gy[t], r);
        }
        oct_wave_sync();
        if (lane < m) s_gr[lane] = r;
        oct_wave_sync();
        const double r2 = mz_wave_sum(r * r);
        double wP = 0.0;
        if (lane < p) {
            double w = 0.0;
            for (int q = 0; q < m; ++q) w = fma(AP[q * R + lane], s_gr[q], w);
            wP = w * y;
        }
        double wZ = 0.0, nmax = lane < p ? 1.0 / cinv : 0.0;
        int jz = -1;
        static_assert(NW_K == 4, "one call per register slot");
        nw_kkt_dual(A, ppos[0] == -1, lane, c, N, Np, n_add, s1, n1, s2, s_st, s_rs, s_ev, s_ep, s_er, m, m_v, s_gr,
                    wZ, jz, nmax);
        nw_kkt_dual(A, ppos[1] == -1, 64 + lane, c, N, Np, n_add, s1, n1, s2, s_st, s_rs, s_ev, s_ep, s_er, m, m_v,
                    s_gr, wZ, jz, nmax);
        nw_kkt_dual(A, ppos[2] == -1, 128 + lane, c, N, Np, n_add, s1, n1, s2, s_st, s_rs, s_ev, s_ep, s_er, m, m_v,
                    s_gr, wZ, jz, nmax);
        nw_kkt_dual(A, ppos[3] == -1, 192 + lane, c, N, Np, n_add, s1, n1, s2, s_st, s_rs, s_ev, s_ep, s_er, m, m_v,
                    s_gr, wZ, jz, nmax);
        const double wZmax = mz_wave_max(wZ);
        const double amax = mz_wave_max(nmax);
        const double wPy = fabs(mz_wave_sum(wP));
        const double b00 = mz_readlane_d(b0, 0);
        const double x1 = (b00 + sqrt(r2)) * amax / s_rs[0];
        const double gap = wZmax * x1 + wPy;
        certified_fail = r2 - 2.0 * gap > A.max_err * A.max_err;
    }
    const double xj = lane < p ? y * cinv : 0.0;
    const bool keep = ok && lane < p && xj > A.w_threshold;
    const uint64_t km = __ballot(keep);
    const int nonzero = __popcll(km);
    const double ndw = mz_wave_sum(keep ? xj : 0.0);
    if (nonzero >= N) ok = false;
    const int32_t n_new = ok ? nonzero : N;
    if (ok) {
        // merged particles in column order
        int rank = 0;
        for (int u = 0; u < p; ++u) {
            const int ju = __builtin_amdgcn_readlane(jt, u);
            if (((km >> u) & 1ull) && ju < jt) ++rank;
        }
        if (keep) {
            const int64_t dst = oct_slot(rank, s1, n1, s2);
            A.sp.w[dst] = xj * (W / ndw);
            A.sp.v0[dst] = cd.vx; A.sp.v1[dst] = cd.vy; A.sp.v2[dst] = cd.vz;
            A.sp.x0[dst] = A.pos_ix ? cd.xx : s_st[4];
            A.sp.x1[dst] = A.pos_iy ? cd.xy : s_st[5];
            A.sp.x2[dst] = A.pos_iz ? cd.xz : s_st[6];
        }
        for (int li = n_new + lane; li < N; li += 64) A.sp.w[oct_slot(li, s1, n1, s2)] = 0.0;
    }
    if (lane == 0) {
        if (!ok && certified_fail) {
            A.status[c] = -1;
            int32_t *op = A.old_pia + c * 5;
            op[0] = s1; op[1] = n1; op[2] = s2; op[3] = n2; op[4] = -1;
        } else if (!ok) {
            A.status[c] = 2;   // the FP64 method decides
        } else {
            A.status[c] = 1;
            int32_t *op = A.old_pia + c * 5;
            op[0] = s1; op[1] = n1; op[2] = s2; op[3] = n2; op[4] = n_new;
            const int32_t n_del = N - n_new;
            const int32_t d2 = min(n_del, n2), d1 = n_del - d2;
            const int32_t nn2 = n2 - d2, nn1 = n1 - d1;
            A.sp.pia.n2[c] = nn2;
            A.sp.pia.end2[c] = nn2 > 0 ? s2 + nn2 - 1 : -1;
            if (nn2 == 0) A.sp.pia.start2[c] = 0;
            A.sp.pia.n1[c] = nn1;
            A.sp.pia.end1[c] = nn1 > 0 ? s1 + nn1 - 1 : -1;
            if (nn1 == 0) A.sp.pia.start1[c] = 0;
            A.sp.pia.nloc[c] = nn1 + nn2;
            if (!(c == A.n_cells - 1) || (n_del > n2)) A.sp.ctr->contiguous = 0;
        }
    }
}

#pragma clang fp contract(on)

// The wave method evaluates its rows with constant exponents when they are
// exactly the 35 multi-indices of total order <= 4 and there are no spatial
// rows: the rows are then listed in graded order (the order of a system's
// rows changes neither its solutions nor the merged particles' moments).
static void nnls_wave_rows(NnlsArgs &A) {
    A.canon = 0;
    if (A.m != NW_ROWS || A.m_p != 0) return;
    bool seen[NW_ROWS] = {};
    for (int q = 0; q < A.m; ++q) {
        int hit = -1;
        for (int k = 0; k < NW_ROWS; ++k)
            if (A.mim[q][0] == nw_canon(k, 0) && A.mim[q][1] == nw_canon(k, 1) && A.mim[q][2] == nw_canon(k, 2))
                hit = k;
        if (hit < 0 || seen[hit]) return;
        seen[hit] = true;
    }
    for (int q = 0; q < NW_ROWS; ++q) {
        for (int d = 0; d < 3; ++d) A.mim[q][d] = nw_canon(q, d);
        A.tot_order[q] = nw_canon(q, 0) + nw_canon(q, 1) + nw_canon(q, 2);
    }
    A.canon = 1;
}

// cell_list (0-based, ascending) replaces the threshold's selection when given.
static int nnls_merge_impl(mz_store *h, int32_t species, int32_t threshold, const int32_t *mim, int32_t n_mim,
                           double vref, int32_t scaling_vref, int32_t centered_at_mean, const double *v_multipliers,
                           int32_t n_vmult, double max_err, int32_t iteration_mult, double w_threshold,
                           int32_t squash, int32_t max_pcols, int32_t method, int32_t *status,
                           const int32_t *cell_list = nullptr, int32_t n_list = 0, const int32_t *mim_pos = nullptr,
                           int32_t n_mim_pos = 0, int32_t n_rand_pairs = 0, uint64_t tag = 0, int32_t rng_mode = 0) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "merge_nnls: bad species");
    MZ_CHECK(n_mim_pos >= 0 && n_mim_pos <= NNLS_MAX_POS && (n_mim_pos == 0 || mim_pos != nullptr),
             "merge_nnls: at most 16 spatial multi-indices");
    MZ_CHECK(n_rand_pairs >= 0, "merge_nnls: negative n_rand_pairs");
    MZ_CHECK(rng_mode == 0 || (rng_mode == 1 && h->d_replay[2]), "merge_nnls: replay table not set");
    MZ_CHECK(method >= 0 && method <= 3,
             "merge_nnls: method must be 0 (reference layout), 1 (LDS-resident), 2 (FP32 search in LDS, FP64 "
             "solve) or 3 (FP32 search in registers, one wavefront per cell, FP64 solve)");
    MZ_CHECK(n_vmult >= 0 && n_vmult <= NNLS_MAX_VMULT, "merge_nnls: too many v_multipliers");
    MZ_CHECK(n_mim >= 0, "merge_nnls: bad multi-index count");
    MzSpecies &s = h->sp[species - 1];
    const int nc = h->n_cells;
    NnlsArgs A{};
    // NNLSMerge constructor (merging_nnls.jl:143-206): base moments first,
    // then the requested ones not already present
    static const int32_t base[7][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2}};
    int m = 0;
    for (int i = 0; i < 7; ++i) { for (int d = 0; d < 3; ++d) A.mim[m][d] = base[i][d]; ++m; }
    for (int i = 0; i < n_mim; ++i) {
        bool dup = false;
        for (int k = 0; k < 7; ++k)
            if (mim[3 * i] == base[k][0] && mim[3 * i + 1] == base[k][1] && mim[3 * i + 2] == base[k][2]) dup = true;
        if (dup) continue;
        MZ_CHECK(m < NNLS_MAX_ROWS, "merge_nnls: too many moments");
        for (int d = 0; d < 3; ++d) {
            MZ_CHECK(mim[3 * i + d] >= 0 && mim[3 * i + d] <= 16, "merge_nnls: bad multi-index");
            A.mim[m][d] = mim[3 * i + d];
        }
        ++m;
    }
    for (int q = 0; q < m; ++q) A.tot_order[q] = A.mim[q][0] + A.mim[q][1] + A.mim[q][2];
    // spatial moments follow the velocity rows (merging_nnls.jl:128-141); the
    // first spatial moments present decide where merged particles sit
    A.m_v = m;
    A.m_p = n_mim_pos;
    A.pos_ix = A.pos_iy = A.pos_iz = 0;
    for (int i = 0; i < n_mim_pos; ++i) {
        for (int d = 0; d < 3; ++d) {
            MZ_CHECK(mim_pos[3 * i + d] >= 0 && mim_pos[3 * i + d] <= 16, "merge_nnls: bad spatial multi-index");
            A.mim_pos[i][d] = mim_pos[3 * i + d];
        }
        const int32_t *q = &mim_pos[3 * i];
        if (q[0] == 1 && q[1] == 0 && q[2] == 0) A.pos_ix = 1;
        else if (q[0] == 0 && q[1] == 1 && q[2] == 0) A.pos_iy = 1;
        else if (q[0] == 0 && q[1] == 0 && q[2] == 1) A.pos_iz = 1;
    }
    // a random pair's spatial rows take the velocity multi-indices (see nnls_entry)
    MZ_CHECK(n_rand_pairs == 0 || n_mim_pos <= m, "merge_nnls: more spatial than velocity moments with random pairs");
    m += n_mim_pos;
    MZ_CHECK(m <= NNLS_MAX_ROWS, "merge_nnls: too many moments");
    A.m = m;
    A.n_rand = n_rand_pairs;
    A.rp = mz_rng_params(h, 2, tag, rng_mode);
    A.vref = vref;
    A.inv_vref = 1.0 / vref;
    A.scaling_vref = scaling_vref;
    A.centered = centered_at_mean ? 1 : 0;
    A.n_vmult = n_vmult;
    for (int i = 0; i < n_vmult; ++i) A.vmult[i] = v_multipliers[i];
    A.max_err = max_err;
    A.iteration_mult = iteration_mult;
    A.w_threshold = w_threshold;
    A.sp = mz_view(s);
    A.tmp = A.sp;
    A.tmp.x0 = s.xb[0]; A.tmp.x1 = s.xb[1]; A.tmp.x2 = s.xb[2];
    A.tmp.v0 = s.vb[0]; A.tmp.v1 = s.vb[1]; A.tmp.v2 = s.vb[2];
    A.tmp.w = s.wb;
    A.n_cells = nc;
    A.max_pcols = max_pcols < 0 ? -1 : max_pcols;
    if (method >= 1) {
        MZ_CHECK(max_pcols >= 0 && max_pcols + A.centered + 8 * n_vmult + n_rand_pairs <= NF_COLS,
                 "merge_nnls: the LDS-resident methods need max_pcols + fictitious columns + random pairs <= 256");
    }
    h->nnls_info[0] = 0;
    h->nnls_info[1] = 0;
    // cells to merge and their scratch sizes (host side)
    MzCounters snap;
    int r = mz_read_counters(h, species, &snap);
    if (r) return r;
    std::vector<int32_t> nloc(nc);
    MZ_HIP(hipMemcpy(nloc.data(), s.pia.nloc, nc * sizeof(int32_t), hipMemcpyDeviceToHost));
    const int n_add = A.centered + 8 * n_vmult + n_rand_pairs;
    std::vector<int32_t> cells;
    std::vector<int64_t> offs;
    int64_t tot = 0;
    for (int k = 0; k < (cell_list ? n_list : nc); ++k) {
        const int c = cell_list ? cell_list[k] : k;
        if (cell_list || nloc[c] > threshold) {
            const int64_t n = (A.max_pcols >= 0 ? std::min<int64_t>(nloc[c], A.max_pcols) : nloc[c]) + n_add;
            cells.push_back(c);
            offs.push_back(tot);
            tot += ((int64_t)m * n + 3 * n + (n + 1) / 2 + 32) & ~(int64_t)31;
        }
    }
    int32_t *old_pia = (int32_t *)mz_scratch(h, "oct_oldpia", (size_t)nc * 5 * sizeof(int32_t));
    int32_t *d_status = (int32_t *)mz_scratch(h, "nnls_status", (size_t)nc * sizeof(int32_t));
    int32_t *ndel = (int32_t *)mz_scratch(h, "oct_ndel", (size_t)nc * sizeof(int32_t));
    int64_t *off = (int64_t *)mz_scratch(h, "oct_off", (size_t)(nc + 1) * sizeof(int64_t));
    MZ_CHECK(old_pia && d_status && ndel && off, "merge_nnls: scratch allocation failed");
    MZ_HIP(hipMemsetAsync(d_status, 0, nc * sizeof(int32_t), h->stream));
    // old_pia[c][4] = -1 marks "not merged" for the id pushes
    {
        std::vector<int32_t> init((size_t)nc * 5, 0);
        for (int c = 0; c < nc; ++c) init[(size_t)c * 5 + 4] = -1;
        MZ_HIP(hipMemcpy(old_pia, init.data(), init.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    // the LDS-resident methods' cell lists (allocated and uploaded before the
    // timed region: staging of the input, not merging)
    int32_t *dcells = nullptr, *dredo = nullptr;
    if (!cells.empty() && method >= 1) {
        dcells = (int32_t *)mz_scratch(h, "nnls_cells", cells.size() * sizeof(int32_t));
        dredo = (int32_t *)mz_scratch(h, "nnls_cells_fp64", cells.size() * sizeof(int32_t));
        MZ_CHECK(dcells && dredo, "merge_nnls: scratch allocation failed");
        MZ_HIP(hipMemcpy(dcells, cells.data(), cells.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    {
        MzTimed tm(h, "merge_nnls");
        if (!cells.empty() && method >= 1) {
            A.status = d_status;
            A.old_pia = old_pia;
            // kind 0: FP64 in LDS, 1: FP32 search in LDS, 2: FP32 search in registers
            auto launch = [&](int kind, const int32_t *list, size_t count) {
                A.cells = list;
                const dim3 g((unsigned)count), b(NF_COLS);
                if (m <= 35) {
                    if (kind == 2) {
                        NnlsArgs Aw = A;
                        nnls_wave_rows(Aw);
                        hipLaunchKernelGGL(k_nnls_wave, g, dim3(64), 0, h->stream, Aw);
                    }
                    else if (kind == 1) hipLaunchKernelGGL((k_nnls_fast<35, float>), g, b, 0, h->stream, A);
                    else hipLaunchKernelGGL((k_nnls_fast<35, double>), g, b, 0, h->stream, A);
                } else {   // more than 35 moments: FP64 only (the FP32 matrix area cannot hold the FP64 system)
                    hipLaunchKernelGGL((k_nnls_fast<NNLS_MAX_ROWS, double>), g, b, 0, h->stream, A);
                }
                return hipGetLastError();
            };
            const bool mixed = method >= 2 && m <= 35;
            MZ_HIP(launch(mixed ? method - 1 : 0, dcells, cells.size()));
            h->nnls_info[0] = (int32_t)cells.size();
            if (mixed) {
                // cells the FP64 re-solve did not confirm run the FP64 method
                std::vector<int32_t> st(nc);
                MZ_HIP(hipStreamSynchronize(h->stream));
                MZ_HIP(hipMemcpy(st.data(), d_status, nc * sizeof(int32_t), hipMemcpyDeviceToHost));
                std::vector<int32_t> redo;
                for (int32_t cc : cells)
                    if (st[cc] == 2) redo.push_back(cc);
                h->nnls_info[1] = (int32_t)redo.size();
                if (!redo.empty()) {
                    MZ_HIP(hipMemcpy(dredo, redo.data(), redo.size() * sizeof(int32_t), hipMemcpyHostToDevice));
                    MZ_HIP(launch(0, dredo, redo.size()));
                }
            }
        } else if (!cells.empty()) {
            // batches bounded to ~1 GiB of matrix scratch
            const int64_t budget = (int64_t)1 << 27;   // doubles
            size_t i0 = 0;
            while (i0 < cells.size()) {
                size_t i1 = i0;
                const int64_t base_off = offs[i0];
                while (i1 < cells.size() && (i1 == i0 || offs[i1] - base_off + 1 < budget)) ++i1;
                const int64_t span = (i1 < cells.size() ? offs[i1] : tot) - base_off;
                double *scr = (double *)mz_scratch(h, "nnls_A", (size_t)span * sizeof(double));
                int32_t *dcells = (int32_t *)mz_scratch(h, "nnls_cells", (i1 - i0) * sizeof(int32_t));
                int64_t *doffs = (int64_t *)mz_scratch(h, "nnls_offs", (i1 - i0) * sizeof(int64_t));
                MZ_CHECK(scr && dcells && doffs, "merge_nnls: scratch allocation failed");
                std::vector<int64_t> lo(offs.begin() + i0, offs.begin() + i1);
                for (auto &v : lo) v -= base_off;
                MZ_HIP(hipMemcpy(dcells, cells.data() + i0, (i1 - i0) * sizeof(int32_t), hipMemcpyHostToDevice));
                MZ_HIP(hipMemcpy(doffs, lo.data(), (i1 - i0) * sizeof(int64_t), hipMemcpyHostToDevice));
                A.cells = dcells;
                A.aoff = doffs;
                A.scratch = scr;
                A.status = d_status;
                A.old_pia = old_pia;
                hipLaunchKernelGGL(k_nnls_merge, dim3((unsigned)(i1 - i0)), dim3(256), 0, h->stream, A);
                MZ_HIP(hipGetLastError());
                MZ_HIP(hipStreamSynchronize(h->stream));
                i0 = i1;
            }
        }
        hipLaunchKernelGGL(k_merge_ndel, dim3((nc + 255) / 256), dim3(256), 0, h->stream, old_pia, nc, ndel);
        hipLaunchKernelGGL(k_scan_i32m, dim3(1), dim3(1024), 0, h->stream, ndel, (int64_t)nc, off, off + nc);
        hipLaunchKernelGGL(k_merge_push_ids, dim3(nc), dim3(64), 0, h->stream, A.sp, old_pia, off, off + nc,
                           snap.nbuffer, snap.n_total, nc);
    }
    MZ_HIP(hipGetLastError());
    // the handle's stream does not synchronise with hipMemcpy: wait for it first
    MZ_HIP(hipStreamSynchronize(h->stream));
    if (status) MZ_HIP(hipMemcpy(status, d_status, nc * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (squash) return mz_squash(h, species);
    return 0;
}

extern "C" int mz_merge_nnls(mz_store *h, int32_t species, int32_t threshold, const int32_t *mim, int32_t n_mim,
                             const int32_t *mim_pos, int32_t n_mim_pos, double vref, int32_t scaling_vref,
                             int32_t n_rand_pairs, int32_t centered_at_mean, const double *v_multipliers,
                             int32_t n_vmult, double max_err, int32_t iteration_mult, double w_threshold,
                             int32_t squash, uint64_t tag, int32_t rng_mode, int32_t *status) {
    if (species >= 1 && species <= h->n_species) h->sp[species - 1].contig_known = false;
    return nnls_merge_impl(h, species, threshold, mim, n_mim, vref, scaling_vref, centered_at_mean, v_multipliers,
                           n_vmult, max_err, iteration_mult, w_threshold, squash, -1, 0, status, nullptr, 0, mim_pos,
                           n_mim_pos, n_rand_pairs, tag, rng_mode);
}

extern "C" int mz_get_nnls_info(mz_store *h, int32_t *info) {
    MZ_CHECK(info != nullptr, "nnls info: NULL output");
    info[0] = h->nnls_info[0];
    info[1] = h->nnls_info[1];
    return 0;
}

extern "C" int mz_merge_nnls_cand(mz_store *h, int32_t species, int32_t threshold, const int32_t *mim,
                                  int32_t n_mim, const int32_t *mim_pos, int32_t n_mim_pos, double vref,
                                  int32_t scaling_vref, int32_t n_rand_pairs, int32_t centered_at_mean,
                                  const double *v_multipliers, int32_t n_vmult, double max_err,
                                  int32_t iteration_mult, double w_threshold, int32_t squash, int32_t max_pcols,
                                  int32_t method, uint64_t tag, int32_t rng_mode, int32_t *status) {
    if (species >= 1 && species <= h->n_species) h->sp[species - 1].contig_known = false;
    return nnls_merge_impl(h, species, threshold, mim, n_mim, vref, scaling_vref, centered_at_mean, v_multipliers,
                           n_vmult, max_err, iteration_mult, w_threshold, squash, max_pcols, method, status, nullptr,
                           0, mim_pos, n_mim_pos, n_rand_pairs, tag, rng_mode);
}

// The caller's loop of per-cell merge_nnls_based! calls as one call: cells[k]
// (1-based) in list order, status_out[k] = 1 (merged) or -1 (failed), the
// reference's return value (merging_nnls.jl:1457-1510).  Each run of
// ascending cells is one launch, runs in order; method as
// mz_merge_nnls_cand (max_pcols < 0: every particle a column, method 0).
extern "C" int mz_merge_nnls_cells(mz_store *h, int32_t species, const int32_t *cells, int32_t n,
                                   const int32_t *mim, int32_t n_mim, const int32_t *mim_pos, int32_t n_mim_pos,
                                   double vref, int32_t scaling_vref, int32_t n_rand_pairs, int32_t centered_at_mean,
                                   const double *v_multipliers, int32_t n_vmult, double max_err,
                                   int32_t iteration_mult, double w_threshold, int32_t max_pcols, int32_t method,
                                   uint64_t tag, int32_t rng_mode, int32_t *status_out) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "merge_nnls_cells: bad species");
    MZ_CHECK(n >= 0 && (n == 0 || (cells != nullptr && status_out != nullptr)), "merge_nnls_cells: bad cell list");
    for (int32_t k = 0; k < n; ++k)
        MZ_CHECK(cells[k] >= 1 && cells[k] <= h->n_cells, "merge_nnls_cells: cell out of range");
    h->sp[species - 1].contig_known = false;
    std::vector<int32_t> st(h->n_cells), run;
    int32_t k0 = 0;
    while (k0 < n) {
        int32_t k1 = k0 + 1;
        while (k1 < n && cells[k1] > cells[k1 - 1]) ++k1;
        run.clear();
        for (int32_t k = k0; k < k1; ++k) run.push_back(cells[k] - 1);
        int r = nnls_merge_impl(h, species, 0, mim, n_mim, vref, scaling_vref, centered_at_mean, v_multipliers,
                                n_vmult, max_err, iteration_mult, w_threshold, 0, max_pcols < 0 ? -1 : max_pcols,
                                max_pcols < 0 ? 0 : method, st.data(), run.data(), (int32_t)run.size(), mim_pos,
                                n_mim_pos, n_rand_pairs, tag, rng_mode);
        if (r) return r;
        for (int32_t k = k0; k < k1; ++k) status_out[k] = st[cells[k] - 1];
        k0 = k1;
    }
    return 0;
}

// ===========================================================================
// merge_roulette!(rng, pv, pia, cell, species, target_np; conservative)
// (merging_roulette.jl:25-184) on every cell holding more than `threshold`
// particles, one workgroup per cell.  Thread 0 replays the reference's
// deletion loop on the cell's positions in LDS: draw k of the cell's stream
// picks i = floor(R n_local), the particle at i trades places with the last of
// its group, and that last one is deleted.  The outcome is a permutation of
// the cell's slots (survivors first in each group, the deleted particles in
// the group tails, as the reference's index array holds them), applied to the
// data in one gather; then the re-weighting and, when conservative, the
// velocity correction.  Sums are the reference's sequential ones.
#define MZ_ROUL_MAX 8192   // particles per merged cell (the LDS position lists)

struct RoulArgs {
    MzSpView sp;
    MzSpView tmp;              // second particle buffer: the cell's data before the merge
    int32_t threshold, target, conservative, n_cells;
    MzRngParams rp;
    const int64_t *off;        // deleted particles before each cell (buffer order)
    int32_t nbuf0;
    int32_t *n_merged;
};

__global__ void __launch_bounds__(256) k_roulette_merge(RoulArgs A) {
    __shared__ int32_t lst[MZ_ROUL_MAX];   // position k of the cell holds the particle first at lst[k]
    __shared__ int32_t dl[MZ_ROUL_MAX];    // deleted particles in draw order
    __shared__ int32_t s_m1, s_m2;
    __shared__ double s_sc[5];              // scale factor; conservative: E_scale, v0 (3)
    __shared__ double s_vm[3];
    const int c = blockIdx.x, tid = threadIdx.x;
    const int32_t s1 = A.sp.pia.start1[c], n1 = A.sp.pia.n1[c];
    const int32_t s2 = A.sp.pia.start2[c], n2 = A.sp.pia.n2[c] > 0 ? A.sp.pia.n2[c] : 0;
    const int32_t N = n1 + n2;
    if (N <= A.threshold) return;
    if (N > MZ_ROUL_MAX) {
        if (tid == 0) atomicOr(&A.sp.ctr->err, MZ_ERR_CAPACITY);
        return;
    }
    const int32_t n_del = N - A.target > 0 ? N - A.target : 0;
    auto slot = [&](int32_t li) { return oct_slot(li, s1, n1, s2); };
    for (int k = tid; k < N; k += blockDim.x) {
        lst[k] = k;
        const int64_t p = slot(k);
        A.tmp.x0[p] = A.sp.x0[p]; A.tmp.x1[p] = A.sp.x1[p]; A.tmp.x2[p] = A.sp.x2[p];
        A.tmp.v0[p] = A.sp.v0[p]; A.tmp.v1[p] = A.sp.v1[p]; A.tmp.v2[p] = A.sp.v2[p];
        A.tmp.w[p] = A.sp.w[p];
        A.tmp.id[p] = A.sp.id[p];
    }
    __syncthreads();
    if (tid == 0) {
        const MzRngParams rp = mz_rng_resolve(A.rp);
        MzStream st = mz_stream((uint64_t)c);
        int32_t m1 = n1, m2 = n2;
        for (int32_t n = 0; n < n_del; ++n) {
            const int32_t i = (int32_t)floor(mz_rand(rp, st) * (double)(m1 + m2));
            int32_t at, last;
            if (m2 > 0 && i >= m1) {
                at = n1 + (i - m1);
                last = n1 + m2 - 1;
                --m2;
            } else {
                at = i;
                last = m1 - 1;
                --m1;
            }
            const int32_t chosen = lst[at];
            lst[at] = lst[last];
            lst[last] = chosen;
            dl[n] = chosen;
        }
        s_m1 = m1;
        s_m2 = m2;
        // the reference's sums, in its order (group 1, then group 2)
        double w_deleted = 0.0;
        for (int32_t n = 0; n < n_del; ++n) w_deleted += A.tmp.w[slot(dl[n])];
        double sc;
        if (A.conservative) {
            double wt = 0.0, v0[3] = {0.0, 0.0, 0.0}, E0 = 0.0;
            for (int32_t k = 0; k < N; ++k) {
                const int64_t p = slot(k);
                const double w = A.tmp.w[p];
                wt += w;
                v0[0] += w * A.tmp.v0[p]; v0[1] += w * A.tmp.v1[p]; v0[2] += w * A.tmp.v2[p];
            }
            for (int d = 0; d < 3; ++d) v0[d] /= wt;
            for (int32_t k = 0; k < N; ++k) {
                const int64_t p = slot(k);
                const double a = A.tmp.v0[p] - v0[0], b = A.tmp.v1[p] - v0[1], e = A.tmp.v2[p] - v0[2];
                E0 += A.tmp.w[p] * ((a * a + b * b) + e * e);
            }
            E0 /= wt;
            sc = wt / (wt - w_deleted);
            // survivors after re-weighting: new mean and energy
            double vm[3] = {0.0, 0.0, 0.0}, En = 0.0;
            for (int g = 0; g < 2; ++g) {
                const int32_t k0 = g == 0 ? 0 : n1, k1 = g == 0 ? m1 : n1 + m2;
                for (int32_t k = k0; k < k1; ++k) {
                    const int64_t p = slot(lst[k]);
                    const double w = A.tmp.w[p] * sc;
                    vm[0] += w * A.tmp.v0[p]; vm[1] += w * A.tmp.v1[p]; vm[2] += w * A.tmp.v2[p];
                }
            }
            for (int d = 0; d < 3; ++d) vm[d] /= wt;
            for (int g = 0; g < 2; ++g) {
                const int32_t k0 = g == 0 ? 0 : n1, k1 = g == 0 ? m1 : n1 + m2;
                for (int32_t k = k0; k < k1; ++k) {
                    const int64_t p = slot(lst[k]);
                    const double a = A.tmp.v0[p] - vm[0], b = A.tmp.v1[p] - vm[1], e = A.tmp.v2[p] - vm[2];
                    En += A.tmp.w[p] * sc * ((a * a + b * b) + e * e);
                }
            }
            En /= wt;
            s_sc[1] = sqrt(E0 / En);
            s_sc[2] = v0[0]; s_sc[3] = v0[1]; s_sc[4] = v0[2];
            s_vm[0] = vm[0]; s_vm[1] = vm[1]; s_vm[2] = vm[2];
        } else {
            double wn = 0.0;
            for (int g = 0; g < 2; ++g) {
                const int32_t k0 = g == 0 ? 0 : n1, k1 = g == 0 ? m1 : n1 + m2;
                for (int32_t k = k0; k < k1; ++k) wn += A.tmp.w[slot(lst[k])];
            }
            sc = (wn + w_deleted) / wn;
        }
        s_sc[0] = sc;
    }
    __syncthreads();
    const int32_t m1 = s_m1, m2 = s_m2;
    const double sc = s_sc[0];
    // the gather: position k takes particle lst[k]; deleted ones weigh 0
    for (int k = tid; k < N; k += blockDim.x) {
        const int64_t q = slot(k), p = slot(lst[k]);
        const bool alive = k < n1 ? k < m1 : k < n1 + m2;
        double v0 = A.tmp.v0[p], v1 = A.tmp.v1[p], v2 = A.tmp.v2[p];
        if (alive && A.conservative) {
            const double es = s_sc[1];
            v0 = es * (v0 - s_vm[0]) + s_sc[2];
            v1 = es * (v1 - s_vm[1]) + s_sc[3];
            v2 = es * (v2 - s_vm[2]) + s_sc[4];
        }
        A.sp.x0[q] = A.tmp.x0[p]; A.sp.x1[q] = A.tmp.x1[p]; A.sp.x2[q] = A.tmp.x2[p];
        A.sp.v0[q] = v0; A.sp.v1[q] = v1; A.sp.v2[q] = v2;
        A.sp.w[q] = alive ? A.tmp.w[p] * sc : 0.0;
        A.sp.id[q] = A.tmp.id[p];
    }
    // free ids in draw order, after all earlier cells' (the reference's cell loop)
    for (int n = tid; n < n_del; n += blockDim.x) {
        const int64_t q = (int64_t)A.nbuf0 + A.off[c] + n;
        if (q < A.sp.bufcap) A.sp.buffer[q] = A.tmp.id[slot(dl[n])];
        else atomicOr(&A.sp.ctr->err, MZ_ERR_BUFFER);
    }
    if (tid == 0) {
        // delete_particle_end_group1!/2! leave (0, -1) for an emptied group
        A.sp.pia.n1[c] = m1;
        A.sp.pia.end1[c] = m1 > 0 ? s1 + m1 - 1 : -1;
        if (m1 == 0) A.sp.pia.start1[c] = 0;
        A.sp.pia.n2[c] = m2;
        A.sp.pia.end2[c] = m2 > 0 ? s2 + m2 - 1 : -1;
        if (m2 == 0) A.sp.pia.start2[c] = 0;
        A.sp.pia.nloc[c] = m1 + m2;
        A.sp.ctr->contiguous = 0;
        atomicAdd(A.n_merged, 1);
    }
}

__global__ void k_roulette_ndel(MzSpView sp, int32_t n_cells, int32_t threshold, int32_t target,
                                int32_t *__restrict__ ndel) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n_cells) {
        const int32_t N = sp.pia.n1[c] + (sp.pia.n2[c] > 0 ? sp.pia.n2[c] : 0);
        ndel[c] = (N > threshold && N <= MZ_ROUL_MAX && N > target) ? N - target : 0;
    }
}

__global__ void k_roulette_counters(MzSpView sp, const int64_t *__restrict__ total, int32_t nbuf0, int32_t ntot0) {
    sp.ctr->nbuffer = nbuf0 + (int32_t)*total;
    sp.ctr->n_total = ntot0 - (int32_t)*total;
    sp.ctr->deleted = (int32_t)*total;
}

extern "C" int mz_merge_roulette(mz_store *h, int32_t species, int32_t threshold, int32_t target_np,
                                 int32_t conservative, uint64_t tag, int32_t rng_mode, int64_t *n_merged) {
    if (species >= 1 && species <= h->n_species) h->sp[species - 1].contig_known = false;
    MZ_CHECK(species >= 1 && species <= h->n_species, "merge_roulette: bad species");
    MZ_CHECK(target_np >= 0 && threshold >= 0, "merge_roulette: negative threshold or target");
    MZ_CHECK(rng_mode == 0 || (rng_mode == 1 && h->d_replay[2]), "merge_roulette: replay table not set");
    MzSpecies &s = h->sp[species - 1];
    const int nc = h->n_cells;
    int32_t *cnt = (int32_t *)mz_scratch(h, "oct_cnt", sizeof(int32_t));
    int32_t *ndel = (int32_t *)mz_scratch(h, "oct_ndel", (size_t)nc * sizeof(int32_t));
    int64_t *off = (int64_t *)mz_scratch(h, "oct_off", (size_t)(nc + 1) * sizeof(int64_t));
    MZ_CHECK(cnt && ndel && off, "merge_roulette: scratch allocation failed");
    MzCounters snap;
    int r = mz_read_counters(h, species, &snap);
    if (r) return r;
    MZ_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t), h->stream));
    RoulArgs A{};
    A.sp = mz_view(s);
    A.tmp = A.sp;
    A.tmp.x0 = s.xb[0]; A.tmp.x1 = s.xb[1]; A.tmp.x2 = s.xb[2];
    A.tmp.v0 = s.vb[0]; A.tmp.v1 = s.vb[1]; A.tmp.v2 = s.vb[2];
    A.tmp.w = s.wb;
    A.tmp.id = s.idb;
    A.threshold = threshold;
    A.target = target_np;
    A.conservative = conservative ? 1 : 0;
    A.n_cells = nc;
    A.rp = mz_rng_params(h, 2, tag, rng_mode);
    A.off = off;
    A.nbuf0 = snap.nbuffer;
    A.n_merged = cnt;
    {
        MzTimed tm(h, "merge_roulette");
        hipLaunchKernelGGL(k_roulette_ndel, dim3((nc + 255) / 256), dim3(256), 0, h->stream, A.sp, nc, threshold,
                           target_np, ndel);
        hipLaunchKernelGGL(k_scan_i32m, dim3(1), dim3(1024), 0, h->stream, ndel, (int64_t)nc, off, off + nc);
        hipLaunchKernelGGL(k_roulette_merge, dim3(nc), dim3(256), 0, h->stream, A);
        hipLaunchKernelGGL(k_roulette_counters, dim3(1), dim3(1), 0, h->stream, A.sp, off + nc, snap.nbuffer,
                           snap.n_total);
    }
    MZ_HIP(hipGetLastError());
    if (n_merged) {
        int32_t v = 0;
        MZ_HIP(hipMemcpyAsync(&v, cnt, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
        MZ_HIP(hipStreamSynchronize(h->stream));
        *n_merged = v;
    }
    return 0;
}
