// mz_common.h -- shared definitions of the gfx950 particle-path library.
//
// Device store layout (one per species):
//   slot p (0-based) holds the particle at logical position p+1 of the
//   reference's index array (src/particles.jl:194-212).  Particle data is
//   kept physically in logical order as SoA FP64 arrays, and the reference's
//   physical particle id travels with it in `id`, so `id[p]` is exactly the
//   reference's `index[p+1]`.  Sorting therefore permutes the data itself;
//   every later kernel reads its cell's particles as one contiguous,
//   coalesced segment instead of chasing the index indirection.
//   ParticleIndexerArray fields live on the device as int32 arrays per cell.
//   The reference's LIFO of free physical ids, its length and n_total are
//   emulated on the device so index and PIA stay bit-exact through particle
//   creation (NTC splits) and deletion (merges).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#define MZ_DELTA_PARTICLES 256  // src/Merzbild.jl:23

// ---------------------------------------------------------------------------
// error reporting (host)
void mz_set_error(const std::string &msg);
#define MZ_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            mz_set_error(std::string(#call) + ": " + hipGetErrorString(e_));              \
            return -2;                                                                    \
        }                                                                                 \
    } while (0)
#define MZ_CHECK(cond, msg)                                                               \
    do {                                                                                  \
        if (!(cond)) {                                                                    \
            mz_set_error(msg);                                                            \
            return -1;                                                                    \
        }                                                                                 \
    } while (0)

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Draw k of stream s for
// a call tagged `tag` under key `seed`: counter (k/2, s, tag_lo, tag_hi);
// half k%2 of the output gives 64 bits whose top 53 make a double in [0,1).
__host__ __device__ inline void mz_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                          uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Random draws: Philox (mode 0) or a replay table [rows][cols] (mode 1).
struct MzRngParams {
    uint64_t seed;
    uint64_t tag;
    int32_t mode;
    int32_t pad;
    const double *table;
    int64_t rows, cols;
    int32_t *err;  // device flag, set when a replay row runs out
    // When step_ptr is set (captured step graphs), the tag is
    // (*step_ptr << 4) | phase, read once per thread.
    const uint64_t *step_ptr;
    uint32_t phase;
};

__device__ inline MzRngParams mz_rng_resolve(MzRngParams rp) {
    if (rp.step_ptr) rp.tag = (*rp.step_ptr << 4) | rp.phase;
    return rp;
}

struct MzStream {
    uint64_t stream;
    uint64_t k;
    // cached Philox block: two draws per block
    double spare;
    int has_spare;
};

// Draws 2 blk and 2 blk + 1 of a Philox stream.
__device__ inline void mz_philox_pair(const MzRngParams &rp, uint64_t stream, uint64_t blk, double &d0,
                                      double &d1) {
    uint32_t o[4];
    mz_philox((uint32_t)blk, (uint32_t)stream, (uint32_t)rp.tag, (uint32_t)(rp.tag >> 32), (uint32_t)rp.seed,
              (uint32_t)(rp.seed >> 32), o);
    uint64_t b0 = ((uint64_t)o[1] << 32) | o[0];
    uint64_t b1 = ((uint64_t)o[3] << 32) | o[2];
    d0 = (double)(b0 >> 11) * 0x1.0p-53;
    d1 = (double)(b1 >> 11) * 0x1.0p-53;
}

__device__ inline double mz_rand(const MzRngParams &rp, MzStream &st) {
    uint64_t k = st.k++;
    if (rp.mode == 1) {
        if ((int64_t)st.stream >= rp.rows || (int64_t)k >= rp.cols) {
            if (rp.err) atomicOr(rp.err, 1);
            return 0.5;
        }
        return rp.table[(int64_t)st.stream * rp.cols + (int64_t)k];
    }
    if (st.has_spare && (k & 1u)) {
        st.has_spare = 0;
        return st.spare;
    }
    double d0, d1;
    mz_philox_pair(rp, st.stream, k >> 1, d0, d1);
    if (k & 1u) return d1;
    st.spare = d1;
    st.has_spare = 1;
    return d0;
}

__device__ inline MzStream mz_stream(uint64_t s, uint64_t k0 = 0) {
    MzStream st;
    st.stream = s;
    st.k = k0;
    st.spare = 0.0;
    st.has_spare = 0;
    return st;
}

// ---------------------------------------------------------------------------
// per-species device store
struct MzPia {
    int32_t *start1, *end1, *n1, *start2, *end2, *n2, *nloc;  // [n_cells], 1-based positions as the reference
};

// scalar state of a species, kept on the device
struct MzCounters {
    int32_t n_total;     // pia.n_total[species]
    int32_t contiguous;  // pia.contiguous[species]
    int32_t nbuffer;     // pv.nbuffer
    int32_t length;      // length(pv)
    int32_t n_pos;       // extent of used logical positions (>= n_total when holes exist)
    int32_t err;         // error bits, see MZ_ERR_*
    int32_t created;     // particles created by the last NTC call
    int32_t deleted;     // particles deleted by the last merge call
};

enum {
    MZ_ERR_CAPACITY = 1,       // slot capacity exceeded
    MZ_ERR_REPLAY = 2,         // replay table exhausted
    MZ_ERR_BOUNCES = 4,        // convection wall-bounce cap reached
    MZ_ERR_SCRATCH = 8,        // NTC creation scratch exhausted
    MZ_ERR_BUFFER = 16,        // id buffer capacity exceeded
    MZ_ERR_MERGE_SIZE = 32,    // a merged cell is larger than the merge kernel supports
};

struct MzSpecies {
    int64_t cap;  // slots
    // particle data in logical order (A) and the sort's second buffer (B)
    double *x[3], *v[3], *w;
    int32_t *id;
    double *xb[3], *vb[3], *wb;
    int32_t *idb;
    int32_t *key;       // cell key per slot before the last sort (reference pv.cell)
    int32_t *keyb;      // second key buffer (squash moves keys with the data)
    int32_t *buffer;    // free-id LIFO (reference pv.buffer), capacity bufcap
    int64_t bufcap;
    MzPia pia;
    MzCounters *ctr;    // device
    // host-side knowledge: true when ctr->contiguous is known to be 1 (set by
    // the calls that leave the species sorted or squashed, cleared by every
    // call that may clear it), so those calls need not read it back
    bool contig_known = false;
};

// Plain-old-data view of a species passed to kernels by value.
struct MzSpView {
    double *x0, *x1, *x2, *v0, *v1, *v2, *w;
    int32_t *id;
    int32_t *key;
    int32_t *buffer;
    int64_t cap, bufcap;
    MzPia pia;
    MzCounters *ctr;
};

// n_total clamped to the slot capacity, so no kernel indexes past the arrays
// even after a capacity error has been flagged.
__device__ inline int32_t mz_ntotal(const MzSpView &sp) {
    int32_t n = sp.ctr->n_total;
    return n < 0 ? 0 : (n > (int32_t)sp.cap ? (int32_t)sp.cap : n);
}
__device__ inline int32_t mz_npos(const MzSpView &sp) {
    int32_t n = sp.ctr->n_pos;
    return n < 0 ? 0 : (n > (int32_t)sp.cap ? (int32_t)sp.cap : n);
}

inline MzSpView mz_view(const MzSpecies &s) {
    MzSpView v;
    v.x0 = s.x[0]; v.x1 = s.x[1]; v.x2 = s.x[2];
    v.v0 = s.v[0]; v.v1 = s.v[1]; v.v2 = s.v[2];
    v.w = s.w; v.id = s.id; v.key = s.key; v.buffer = s.buffer;
    v.cap = s.cap; v.bufcap = s.bufcap;
    v.pia = s.pia; v.ctr = s.ctr;
    return v;
}

// ---------------------------------------------------------------------------
// wave64 helpers
// Wavefront reductions over the active lanes, every lane receiving the
// result: the ROCm device library's DPP row operations (a few VALU cycles
// per step) rather than LDS permutes.  Callers keep all lanes active.
extern "C" __device__ double __ockl_wfred_add_f64(double);
extern "C" __device__ double __ockl_wfred_max_f64(double);
extern "C" __device__ double __ockl_wfred_min_f64(double);
extern "C" __device__ unsigned long long __ockl_wfred_max_u64(unsigned long long);
extern "C" __device__ float __ockl_wfred_max_f32(float);
extern "C" __device__ int __ockl_wfred_min_i32(int);
// Last-block reductions publish their partials with agent-scope relaxed
// atomic stores, which the hardware keeps coherent across the XCDs' L2
// caches; the writer then only has to wait for its stores to complete
// (s_waitcnt vmcnt(0); expcnt and lgkmcnt left at their no-wait maxima)
// before it takes a ticket, instead of an agent-scope fence, which writes back
// a whole L2.  The last block reads the partials with agent-scope atomic loads.
__device__ inline void mz_wait_stores() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ inline double mz_wave_sum(double v) { return __ockl_wfred_add_f64(v); }
__device__ inline double mz_wave_max(double v) { return __ockl_wfred_max_f64(v); }
__device__ inline double mz_wave_min(double v) { return __ockl_wfred_min_f64(v); }
__device__ inline uint64_t mz_wave_max_u64(uint64_t v) { return __ockl_wfred_max_u64(v); }
__device__ inline int32_t mz_wave_min_i32(int32_t v) { return __ockl_wfred_min_i32(v); }
__device__ inline int mz_wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// value of v in lane l (l uniform across the wave)
__device__ inline double mz_readlane_d(double v, int l) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ inline uint64_t mz_lanemask_lt() {
    uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}
