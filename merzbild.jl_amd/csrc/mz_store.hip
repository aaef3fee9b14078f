// mz_store.hip -- device particle store, grid sorting, squash, 1-D
// convection and physical properties for gfx950.
//
// Reference behaviour restated here (file:line in Merzbild.jl):
//   sort_particles!        src/grids/grid_sorting.jl:58-182
//   squash_pia!            src/particles.jl:622-701
//   convect_particles!     src/convection/convection_1D.jl:17-157
//   reflect/diffuse walls  src/convection/boundary_conditions.jl:79-121
//   compute_props!         src/properties/physical_props.jl:104-154
//   compute_props_sorted!  src/properties/physical_props.jl:317-454
#include "mz_phys.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

// ===========================================================================
// error string
static thread_local std::string g_err;
void mz_set_error(const std::string &msg) { g_err = msg; }
extern "C" const char *mz_last_error(void) { return g_err.c_str(); }

extern "C" int mz_device_info(char *name, int name_len, int *n_cu, double *hbm_gbps_peak) {
    int dev = 0;
    MZ_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    MZ_HIP(hipGetDeviceProperties(&p, dev));
    if (name && name_len > 0) {
        std::strncpy(name, p.gcnArchName, name_len - 1);
        name[name_len - 1] = 0;
    }
    if (n_cu) *n_cu = p.multiProcessorCount;
    // MI355X published HBM3E peak
    if (hbm_gbps_peak) *hbm_gbps_peak = 8000.0;
    return 0;
}

// ===========================================================================
// timing
// an event from the store's pool (created when the pool is empty)
static int timing_event(mz_store *h, hipEvent_t *e) {
    if (!h->event_pool.empty()) {
        *e = h->event_pool.back();
        h->event_pool.pop_back();
        return 0;
    }
    MZ_HIP(hipEventCreate(e));
    return 0;
}

int mz_timing_begin(mz_store *h, const char *family) {
    if (!h->timing) return 0;
    MzEventPair ev;
    if (timing_event(h, &ev.a)) return -1;
    if (timing_event(h, &ev.b)) {
        h->event_pool.push_back(ev.a);
        return -1;
    }
    if (hipEventRecord(ev.a, h->stream) != hipSuccess) {
        h->event_pool.push_back(ev.a);
        h->event_pool.push_back(ev.b);
        mz_set_error("timing: hipEventRecord failed");
        return -1;
    }
    h->open_events.push_back({std::string(family), ev});
    return 0;
}

int mz_timing_end(mz_store *h) { return mz_timing_end_n(h, 1); }

// closes the last open bracket; it counts as n launches of its family
int mz_timing_end_n(mz_store *h, int64_t n) {
    if (!h->timing || h->open_events.empty()) return 0;
    auto fe = h->open_events.back();
    h->open_events.pop_back();
    MZ_HIP(hipEventRecord(fe.second.b, h->stream));
    fe.second.n = n;
    h->pending[fe.first].push_back(fe.second);
    return 0;
}

static int resolve_timing(mz_store *h) {
    if (h->pending.empty()) return 0;
    MZ_HIP(hipStreamSynchronize(h->stream));
    for (auto &kv : h->pending) {
        auto &acc = h->timing_acc[kv.first];
        for (auto &ev : kv.second) {
            float ms = 0.f;
            MZ_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
            acc.first += ms;
            acc.second += ev.n;
            h->event_pool.push_back(ev.a);
            h->event_pool.push_back(ev.b);
        }
    }
    h->pending.clear();
    return 0;
}

extern "C" int mz_timing_enable(mz_store *h, int32_t on) {
    h->timing = on != 0;
    return 0;
}

extern "C" int mz_timing_get(mz_store *h, const char *family, double *mean_ms, int64_t *count) {
    int r = resolve_timing(h);
    if (r) return r;
    auto it = h->timing_acc.find(family);
    if (it == h->timing_acc.end() || it->second.second == 0) {
        if (mean_ms) *mean_ms = 0.0;
        if (count) *count = 0;
        return 0;
    }
    if (mean_ms) *mean_ms = it->second.first / (double)it->second.second;
    if (count) *count = it->second.second;
    return 0;
}

extern "C" int mz_timing_reset(mz_store *h) {
    int r = resolve_timing(h);
    if (r) return r;
    h->timing_acc.clear();
    return 0;
}

// ===========================================================================
// live stores of the process (mz_other_stores_busy)
static std::mutex g_stores_mu;
static std::vector<mz_store *> g_stores;

bool mz_other_stores_busy(mz_store *h) {
    std::lock_guard<std::mutex> lk(g_stores_mu);
    for (mz_store *o : g_stores)
        if (o != h && hipStreamQuery(o->stream) == hipErrorNotReady) return true;
    return false;
}

// ===========================================================================
// create / destroy
static int alloc_species(MzSpecies &s, int64_t cap, int n_cells) {
    s.cap = cap;
    s.bufcap = cap + 4 * MZ_DELTA_PARTICLES;
    for (int d = 0; d < 3; ++d) {
        MZ_HIP(hipMalloc(&s.x[d], cap * sizeof(double)));
        MZ_HIP(hipMalloc(&s.v[d], cap * sizeof(double)));
        MZ_HIP(hipMalloc(&s.xb[d], cap * sizeof(double)));
        MZ_HIP(hipMalloc(&s.vb[d], cap * sizeof(double)));
    }
    MZ_HIP(hipMalloc(&s.w, cap * sizeof(double)));
    MZ_HIP(hipMalloc(&s.wb, cap * sizeof(double)));
    MZ_HIP(hipMalloc(&s.id, cap * sizeof(int32_t)));
    MZ_HIP(hipMalloc(&s.idb, cap * sizeof(int32_t)));
    MZ_HIP(hipMalloc(&s.key, cap * sizeof(int32_t)));
    MZ_HIP(hipMalloc(&s.keyb, cap * sizeof(int32_t)));
    MZ_HIP(hipMalloc(&s.buffer, s.bufcap * sizeof(int32_t)));
    int32_t **f[7] = {&s.pia.start1, &s.pia.end1, &s.pia.n1, &s.pia.start2, &s.pia.end2, &s.pia.n2, &s.pia.nloc};
    for (auto p : f) MZ_HIP(hipMalloc(p, n_cells * sizeof(int32_t)));
    MZ_HIP(hipMalloc(&s.ctr, sizeof(MzCounters)));
    MZ_HIP(hipMemset(s.ctr, 0, sizeof(MzCounters)));
    return 0;
}

extern "C" int mz_create(int32_t n_cells, int32_t n_species, const int64_t *capacity, mz_store **out) {
    MZ_CHECK(n_cells > 0 && n_species > 0 && n_species <= MZ_MAX_SPECIES, "mz_create: bad sizes");
    MZ_CHECK(capacity != nullptr, "mz_create: capacity is NULL");
    mz_store *h = new mz_store();
    h->n_cells = n_cells;
    h->n_species = n_species;
    h->sp.resize(n_species);
    MZ_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    for (int s = 0; s < n_species; ++s) {
        MZ_CHECK(capacity[s] > 0 && capacity[s] < (int64_t)INT32_MAX / 2, "mz_create: capacity out of range");
        int r = alloc_species(h->sp[s], capacity[s], n_cells);
        if (r) return r;
        // empty PIA: ParticleIndexer() = (0, 0, -1, 0, 0, -1, 0)
        std::vector<int64_t> pia((size_t)n_cells * 7, 0);
        for (int c = 0; c < n_cells; ++c) { pia[c * 7 + 2] = -1; pia[c * 7 + 5] = -1; }
        r = mz_upload_species(h, s + 1, 0, nullptr, nullptr, nullptr, nullptr, pia.data(), 0, 1, nullptr, 0, 0);
        if (r) return r;
    }
    int nss = n_species * n_species;
    MZ_HIP(hipMalloc(&h->d_sgwm, (size_t)nss * n_cells * sizeof(double)));
    MZ_HIP(hipMemset(h->d_sgwm, 0, (size_t)nss * n_cells * sizeof(double)));
    MZ_HIP(hipMalloc(&h->d_cfcnt, (size_t)nss * n_cells * 5 * sizeof(int32_t)));
    MZ_HIP(hipMemset(h->d_cfcnt, 0, (size_t)nss * n_cells * 5 * sizeof(int32_t)));
    MZ_HIP(hipMalloc(&h->d_err, sizeof(int32_t)));
    MZ_HIP(hipMemset(h->d_err, 0, sizeof(int32_t)));
    MZ_HIP(hipMalloc(&h->d_ticket, sizeof(uint32_t)));
    MZ_HIP(hipMemset(h->d_ticket, 0, sizeof(uint32_t)));
    size_t pn = (size_t)n_species * n_cells;
    MZ_HIP(hipMalloc(&h->d_np, pn * sizeof(double)));
    MZ_HIP(hipMalloc(&h->d_n, pn * sizeof(double)));
    MZ_HIP(hipMalloc(&h->d_T, pn * sizeof(double)));
    MZ_HIP(hipMalloc(&h->d_v, pn * 3 * sizeof(double)));
    MZ_HIP(hipMemset(h->d_np, 0, pn * sizeof(double)));
    MZ_HIP(hipMemset(h->d_n, 0, pn * sizeof(double)));
    MZ_HIP(hipMemset(h->d_T, 0, pn * sizeof(double)));
    MZ_HIP(hipMemset(h->d_v, 0, pn * 3 * sizeof(double)));
    MZ_HIP(hipMalloc(&h->d_V, n_cells * sizeof(double)));
    MZ_HIP(hipMalloc(&h->d_invV, n_cells * sizeof(double)));
    MZ_HIP(hipMalloc(&h->d_masses, MZ_MAX_SPECIES * sizeof(double)));
    MZ_HIP(hipMemset(h->d_masses, 0, MZ_MAX_SPECIES * sizeof(double)));
    MZ_HIP(hipMalloc(&h->d_flux, (size_t)2 * 9 * pn * sizeof(double)));
    MZ_HIP(hipMemset(h->d_flux, 0, (size_t)2 * 9 * pn * sizeof(double)));
    const size_t surf_bytes = (size_t)2 * n_species * 2 * MZ_SURF_Q * sizeof(double);
    MZ_HIP(hipMalloc(&h->d_surf, surf_bytes));
    MZ_HIP(hipMemset(h->d_surf, 0, surf_bytes));
    MZ_HIP(hipMalloc(&h->d_surf_ticket, 2 * sizeof(uint32_t)));
    MZ_HIP(hipMemset(h->d_surf_ticket, 0, 2 * sizeof(uint32_t)));
    const size_t pavg_bytes = (pn * 6 + n_species) * sizeof(double);
    MZ_HIP(hipMalloc(&h->d_pavg, pavg_bytes));
    MZ_HIP(hipMemset(h->d_pavg, 0, pavg_bytes));
    h->inter.resize(nss);
    MZ_HIP(hipStreamSynchronize(h->stream));
    // the fixed scratch array of mz_sort_noalloc: its address on this device
    int rs = mz_noalloc_scratch_init(nullptr);
    if (rs) return rs;
    {
        std::lock_guard<std::mutex> lk(g_stores_mu);
        g_stores.push_back(h);
    }
    *out = h;
    return 0;
}

static void free_species(MzSpecies &s) {
    for (int d = 0; d < 3; ++d) {
        hipFree(s.x[d]); hipFree(s.v[d]); hipFree(s.xb[d]); hipFree(s.vb[d]);
    }
    hipFree(s.w); hipFree(s.wb); hipFree(s.id); hipFree(s.idb); hipFree(s.key); hipFree(s.keyb);
    hipFree(s.buffer);
    hipFree(s.pia.start1); hipFree(s.pia.end1); hipFree(s.pia.n1); hipFree(s.pia.start2);
    hipFree(s.pia.end2); hipFree(s.pia.n2); hipFree(s.pia.nloc);
    hipFree(s.ctr);
}

extern "C" int mz_destroy(mz_store *h) {
    if (!h) return 0;
    {
        std::lock_guard<std::mutex> lk(g_stores_mu);
        g_stores.erase(std::remove(g_stores.begin(), g_stores.end(), h), g_stores.end());
    }
    hipStreamSynchronize(h->stream);
    // timing events still open or not yet read
    for (auto &fe : h->open_events) {
        hipEventDestroy(fe.second.a);
        hipEventDestroy(fe.second.b);
    }
    for (auto &kv : h->pending)
        for (auto &ev : kv.second) {
            hipEventDestroy(ev.a);
            hipEventDestroy(ev.b);
        }
    for (hipEvent_t e : h->event_pool) hipEventDestroy(e);
    h->event_pool.clear();
    h->open_events.clear();
    h->pending.clear();
    if (h->h_report) hipHostFree(h->h_report);
    mz_graph_release(h);
    for (auto &s : h->sp) free_species(s);
    hipFree(h->d_sgwm); hipFree(h->d_cfcnt); hipFree(h->d_err); hipFree(h->d_ticket);
    hipFree(h->d_np); hipFree(h->d_n); hipFree(h->d_T); hipFree(h->d_v); hipFree(h->d_V);
    hipFree(h->d_invV);
    hipFree(h->d_masses);
    hipFree(h->d_flux);
    hipFree(h->d_surf);
    hipFree(h->d_surf_ticket);
    hipFree(h->d_pavg);
    for (int k = 0; k < 3; ++k) if (h->d_replay[k]) hipFree(h->d_replay[k]);
    for (auto &sc : h->scratch) if (sc.second.ptr) hipFree(sc.second.ptr);
    hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

extern "C" int mz_sync(mz_store *h) {
    MZ_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

void *mz_scratch(mz_store *h, const char *name, size_t bytes) {
    auto &sc = h->scratch[name];
    if (sc.bytes < bytes) {
        if (sc.ptr) {
            hipStreamSynchronize(h->stream);
            hipFree(sc.ptr);
        }
        sc.ptr = nullptr;
        size_t nb = std::max(bytes, (size_t)256);
        if (hipMalloc(&sc.ptr, nb) != hipSuccess) {
            sc.bytes = 0;
            return nullptr;
        }
        sc.bytes = nb;
    }
    return sc.ptr;
}

// ===========================================================================
// upload / download
extern "C" int mz_upload_species(mz_store *h, int32_t species, int64_t n_pos, const double *x,
                                 const double *v, const double *w, const int64_t *ids,
                                 const int64_t *pia, int64_t n_total, int32_t contiguous,
                                 const int64_t *buffer, int64_t nbuffer, int64_t length) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "upload: bad species");
    MzSpecies &s = h->sp[species - 1];
    s.contig_known = false;
    h->slab_shape.valid = false;
    h->slab_min_ppt = 0;
    MZ_CHECK(n_pos >= 0 && n_pos <= s.cap, "upload: n_pos exceeds capacity");
    MZ_CHECK(length >= 0 && length < (int64_t)INT32_MAX / 2, "upload: ParticleVector length out of range");
    MZ_CHECK(nbuffer >= 0 && nbuffer <= length, "upload: bad nbuffer");
    MZ_CHECK(pia != nullptr, "upload: pia is NULL");
    // the free-id LIFO holds at most `length` ids; creations can grow length up
    // to the slot capacity plus one resize! chunk
    int64_t need = std::max<int64_t>(length, s.cap) + 2 * MZ_DELTA_PARTICLES;
    if (need > s.bufcap) {
        MZ_HIP(hipStreamSynchronize(h->stream));
        hipFree(s.buffer);
        s.bufcap = need;
        MZ_HIP(hipMalloc(&s.buffer, s.bufcap * sizeof(int32_t)));
    }
    std::vector<double> tmp((size_t)std::max<int64_t>(n_pos, 1));
    for (int d = 0; d < 3 && n_pos > 0; ++d) {
        for (int64_t i = 0; i < n_pos; ++i) tmp[i] = x[3 * i + d];
        MZ_HIP(hipMemcpy(s.x[d], tmp.data(), n_pos * sizeof(double), hipMemcpyHostToDevice));
        for (int64_t i = 0; i < n_pos; ++i) tmp[i] = v[3 * i + d];
        MZ_HIP(hipMemcpy(s.v[d], tmp.data(), n_pos * sizeof(double), hipMemcpyHostToDevice));
    }
    if (n_pos > 0) {
        MZ_HIP(hipMemcpy(s.w, w, n_pos * sizeof(double), hipMemcpyHostToDevice));
        std::vector<int32_t> iv(n_pos);
        for (int64_t i = 0; i < n_pos; ++i) {
            MZ_CHECK(ids[i] >= 1 && ids[i] <= INT32_MAX, "upload: id out of range");
            iv[i] = (int32_t)ids[i];
        }
        MZ_HIP(hipMemcpy(s.id, iv.data(), n_pos * sizeof(int32_t), hipMemcpyHostToDevice));
        std::fill(iv.begin(), iv.end(), 0);
        MZ_HIP(hipMemcpy(s.key, iv.data(), n_pos * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    int nc = h->n_cells;
    std::vector<int32_t> col((size_t)nc);
    int32_t *dst[7] = {s.pia.nloc, s.pia.start1, s.pia.end1, s.pia.n1, s.pia.start2, s.pia.end2, s.pia.n2};
    int64_t max_end = 0;
    for (int f = 0; f < 7; ++f) {
        for (int c = 0; c < nc; ++c) {
            int64_t val = pia[(int64_t)c * 7 + f];
            MZ_CHECK(val >= -1 && val <= s.cap, "upload: PIA value out of range");
            col[c] = (int32_t)val;
            if (f == 2 || f == 5) max_end = std::max(max_end, val);
        }
        MZ_HIP(hipMemcpy(dst[f], col.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    MZ_CHECK(max_end <= n_pos, "upload: PIA references positions beyond n_pos");
    // the consistency that check_pia_is_correct demands (src/particles.jl:863-907),
    // plus group ranges inside [1, n_pos], so no kernel can index past the store
    int64_t sum_local = 0;
    for (int c = 0; c < nc; ++c) {
        const int64_t *row = pia + (int64_t)c * 7;
        int64_t nl = row[0], s1 = row[1], e1 = row[2], g1 = row[3], s2 = row[4], e2 = row[5], g2 = row[6];
        MZ_CHECK(g1 >= 0 && g2 >= 0 && nl == g1 + g2, "upload: PIA n_local != n_group1 + n_group2");
        MZ_CHECK(g1 == 0 || (s1 >= 1 && e1 == s1 + g1 - 1 && e1 <= n_pos), "upload: PIA group 1 out of range");
        MZ_CHECK(g2 == 0 || (s2 >= 1 && e2 == s2 + g2 - 1 && e2 <= n_pos), "upload: PIA group 2 out of range");
        sum_local += nl;
    }
    MZ_CHECK(sum_local == n_total, "upload: n_total differs from the sum of n_local");
    MZ_CHECK(n_total >= 0 && n_total <= s.cap && (!contiguous || n_total <= n_pos),
             "upload: n_total out of range");
    if (nbuffer > 0) {
        std::vector<int32_t> b(nbuffer);
        for (int64_t i = 0; i < nbuffer; ++i) b[i] = (int32_t)buffer[i];
        MZ_HIP(hipMemcpy(s.buffer, b.data(), nbuffer * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    MzCounters c{};
    c.n_total = (int32_t)n_total;
    c.contiguous = contiguous ? 1 : 0;
    s.contig_known = contiguous != 0;
    c.nbuffer = (int32_t)nbuffer;
    c.length = (int32_t)length;
    c.n_pos = (int32_t)std::max<int64_t>(max_end, contiguous ? n_total : max_end);
    MZ_HIP(hipMemcpy(s.ctr, &c, sizeof(c), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int mz_upload_cell_keys(mz_store *h, int32_t species, int64_t n, const int64_t *cells) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "upload_cell_keys: bad species");
    MzSpecies &s = h->sp[species - 1];
    MZ_CHECK(n >= 0 && n <= s.cap, "upload_cell_keys: n exceeds capacity");
    std::vector<int32_t> k((size_t)std::max<int64_t>(n, 1));
    for (int64_t i = 0; i < n; ++i) {
        MZ_CHECK(cells[i] >= 1 && cells[i] <= h->n_cells, "upload_cell_keys: cell out of range");
        k[i] = (int32_t)cells[i];
    }
    if (n > 0) MZ_HIP(hipMemcpy(s.key, k.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    return 0;
}

int mz_read_counters(mz_store *h, int species, MzCounters *c) {
    MZ_HIP(hipStreamSynchronize(h->stream));
    MZ_HIP(hipMemcpy(c, h->sp[species - 1].ctr, sizeof(MzCounters), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int mz_get_scalars(mz_store *h, int32_t species, int64_t *out) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "get_scalars: bad species");
    MzCounters c;
    int r = mz_read_counters(h, species, &c);
    if (r) return r;
    int32_t rerr = 0;
    MZ_HIP(hipMemcpy(&rerr, h->d_err, sizeof(int32_t), hipMemcpyDeviceToHost));
    out[0] = c.n_total; out[1] = c.contiguous; out[2] = c.nbuffer; out[3] = c.length;
    out[4] = c.n_pos; out[5] = c.err | (rerr ? MZ_ERR_REPLAY : 0);
    return 0;
}

extern "C" int mz_download_species(mz_store *h, int32_t species, int64_t n_pos_max, double *x, double *v,
                                   double *w, int64_t *ids, int64_t *cell_keys, int64_t *pia,
                                   int64_t *buffer, int64_t buffer_max) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "download: bad species");
    MzSpecies &s = h->sp[species - 1];
    MzCounters c;
    int r = mz_read_counters(h, species, &c);
    if (r) return r;
    int64_t n = std::min<int64_t>(c.n_pos, n_pos_max);
    std::vector<double> tmp((size_t)std::max<int64_t>(n, 1));
    if (n > 0) {
        for (int d = 0; d < 3; ++d) {
            if (x) {
                MZ_HIP(hipMemcpy(tmp.data(), s.x[d], n * sizeof(double), hipMemcpyDeviceToHost));
                for (int64_t i = 0; i < n; ++i) x[3 * i + d] = tmp[i];
            }
            if (v) {
                MZ_HIP(hipMemcpy(tmp.data(), s.v[d], n * sizeof(double), hipMemcpyDeviceToHost));
                for (int64_t i = 0; i < n; ++i) v[3 * i + d] = tmp[i];
            }
        }
        if (w) MZ_HIP(hipMemcpy(w, s.w, n * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<int32_t> iv(n);
        if (ids) {
            MZ_HIP(hipMemcpy(iv.data(), s.id, n * sizeof(int32_t), hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) ids[i] = iv[i];
        }
        if (cell_keys) {
            MZ_HIP(hipMemcpy(iv.data(), s.key, n * sizeof(int32_t), hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) cell_keys[i] = iv[i];
        }
    }
    if (pia) {
        int nc = h->n_cells;
        std::vector<int32_t> col(nc);
        int32_t *src[7] = {s.pia.nloc, s.pia.start1, s.pia.end1, s.pia.n1, s.pia.start2, s.pia.end2, s.pia.n2};
        for (int f = 0; f < 7; ++f) {
            MZ_HIP(hipMemcpy(col.data(), src[f], nc * sizeof(int32_t), hipMemcpyDeviceToHost));
            for (int cc = 0; cc < nc; ++cc) pia[(int64_t)cc * 7 + f] = col[cc];
        }
    }
    if (buffer && c.nbuffer > 0) {
        int64_t nb = std::min<int64_t>(c.nbuffer, buffer_max);
        std::vector<int32_t> b(nb);
        MZ_HIP(hipMemcpy(b.data(), s.buffer, nb * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < nb; ++i) buffer[i] = b[i];
    }
    return 0;
}

// ===========================================================================
// parameters
extern "C" int mz_set_grid_1d(mz_store *h, double L, int64_t nx, double wall_offset) {
    MZ_CHECK(nx == h->n_cells, "set_grid_1d: nx must equal the store's cell count");
    MZ_CHECK(L > 0.0, "set_grid_1d: L must be positive");
    double dx = L / (double)nx;
    h->grid.L = L;
    h->grid.nx = (int32_t)nx;
    h->grid.dx = dx;
    h->grid.inv_dx = 1.0 / dx;
    h->grid.min_x = dx * wall_offset;
    h->grid.max_x = L - dx * wall_offset;
    h->has_grid = true;
    // Cell1D(xlo, xhi, V): V = dx, inv_V = 1/V (grid_uniform1D.jl:20-31, 76-82)
    std::vector<double> V(nx, dx), iV(nx, 1.0 / dx);
    MZ_HIP(hipMemcpy(h->d_V, V.data(), nx * sizeof(double), hipMemcpyHostToDevice));
    MZ_HIP(hipMemcpy(h->d_invV, iV.data(), nx * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int mz_set_walls_1d(mz_store *h, const double *T, const double *v, const double *acc,
                               const double *vsq) {
    for (int b = 0; b < 2; ++b) {
        h->walls.T[b] = T[b];
        for (int d = 0; d < 3; ++d) h->walls.v[b][d] = v[3 * b + d];
        h->walls.acc[b] = acc[b];
        for (int s = 0; s < h->n_species; ++s) h->walls.vsq[b][s] = vsq[b * h->n_species + s];
    }
    h->has_walls = true;
    return 0;
}

extern "C" int mz_set_interaction(mz_store *h, int32_t s1, int32_t s2, double mu1, double mu2,
                                  double vhs_factor, double vhs_o) {
    MZ_CHECK(s1 >= 1 && s1 <= h->n_species && s2 >= 1 && s2 <= h->n_species, "set_interaction: bad species");
    MzInteraction &it = h->inter[(s1 - 1) * h->n_species + (s2 - 1)];
    it.mu1 = mu1;
    it.mu2 = mu2;
    it.factor = vhs_factor;
    it.expo = 1.0 - 2 * vhs_o;   // sigma_vhs = factor * g^(1 - 2 omega)
    it.set = 1;
    return 0;
}

extern "C" int mz_set_sigma_g_w_max(mz_store *h, int32_t s1, int32_t s2, const double *sgwm) {
    MZ_CHECK(s1 >= 1 && s1 <= h->n_species && s2 >= 1 && s2 <= h->n_species, "set_sgwm: bad species");
    size_t off = ((size_t)(s1 - 1) * h->n_species + (s2 - 1)) * h->n_cells;
    MZ_HIP(hipMemcpy(h->d_sgwm + off, sgwm, h->n_cells * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int mz_get_collision_factors(mz_store *h, int32_t s1, int32_t s2, double *sgwm, int64_t *n1,
                                        int64_t *n2, int64_t *n_coll, int64_t *n_performed,
                                        int64_t *n_eq_w) {
    MZ_CHECK(s1 >= 1 && s1 <= h->n_species && s2 >= 1 && s2 <= h->n_species, "get_cf: bad species");
    MZ_HIP(hipStreamSynchronize(h->stream));
    size_t pair = (size_t)(s1 - 1) * h->n_species + (s2 - 1);
    int nc = h->n_cells;
    if (sgwm) MZ_HIP(hipMemcpy(sgwm, h->d_sgwm + pair * nc, nc * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int32_t> cnt((size_t)nc * 5);
    MZ_HIP(hipMemcpy(cnt.data(), h->d_cfcnt + pair * nc * 5, cnt.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    int64_t *outs[5] = {n1, n2, n_coll, n_performed, n_eq_w};
    for (int f = 0; f < 5; ++f)
        if (outs[f])
            for (int c = 0; c < nc; ++c) outs[f][c] = cnt[(size_t)c * 5 + f];
    return 0;
}

extern "C" int mz_set_seed(mz_store *h, uint64_t seed) {
    h->seed = seed;
    return 0;
}

extern "C" int mz_set_replay(mz_store *h, int32_t kind, const double *table, int64_t rows, int64_t cols) {
    MZ_CHECK(kind >= 0 && kind < 3, "set_replay: bad kind");
    if (h->d_replay[kind]) {
        MZ_HIP(hipStreamSynchronize(h->stream));
        hipFree(h->d_replay[kind]);
        h->d_replay[kind] = nullptr;
    }
    h->replay_rows[kind] = rows;
    h->replay_cols[kind] = cols;
    if (rows > 0 && cols > 0) {
        MZ_HIP(hipMalloc(&h->d_replay[kind], rows * cols * sizeof(double)));
        MZ_HIP(hipMemcpy(h->d_replay[kind], table, rows * cols * sizeof(double), hipMemcpyHostToDevice));
    }
    MZ_HIP(hipMemset(h->d_err, 0, sizeof(int32_t)));
    return 0;
}

MzRngParams mz_rng_params(mz_store *h, int kind, uint64_t tag, int mode) {
    MzRngParams rp;
    rp.seed = h->seed;
    rp.tag = tag;
    rp.mode = mode;
    rp.pad = 0;
    rp.table = h->d_replay[kind];
    rp.rows = h->replay_rows[kind];
    rp.cols = h->replay_cols[kind];
    rp.err = h->d_err;
    rp.step_ptr = nullptr;
    rp.phase = 0;
    return rp;
}

// ===========================================================================
// Sorting: stable counting sort by cell (grid_sorting.jl:58-182).
//   1. k_sort_hist: per chunk of MZ_SORT_CHUNK positions, cell keys (from x,
//      or the stored keys) and a per-chunk histogram in LDS; the histogram is
//      written cell-major [cell][chunk].
//   2. k_sort_rows: per cell, an exclusive scan over chunks gives each
//      chunk's offset inside the cell; the cell totals, scanned by the block
//      that finishes last, give the cell starts and the new PIA.
//   3. k_sort_scatter: per chunk, its slot of particle i is cell start +
//      chunk offset + rank of i among the chunk's equal keys in position
//      order, so the sort is stable like the reference's reverse scatter.
// The data moves with the sort (x, v, w, id); the key array stays in the
// pre-sort order, like the reference's pv.cell.

__global__ void k_sort_hist(MzSpView sp, MzGrid g, int32_t compute_cells, int32_t n_cells, int32_t nchunks,
                            int32_t *__restrict__ hist) {
    extern __shared__ int32_t sh[];
    const int32_t n = mz_ntotal(sp);
    for (int c = threadIdx.x; c < n_cells; c += blockDim.x) sh[c] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * MZ_SORT_CHUNK;
    for (int i = threadIdx.x; i < MZ_SORT_CHUNK; i += blockDim.x) {
        int64_t p = base + i;
        if (p < n) {
            int32_t c;
            if (compute_cells) {
                c = mz_cell_of(sp.x0[p], g, n_cells, &sp.ctr->err);
                sp.key[p] = c;
            } else {
                c = sp.key[p];
                if (c < 1 || c > n_cells) {
                    atomicOr(&sp.ctr->err, MZ_ERR_CAPACITY);
                    c = c < 1 ? 1 : n_cells;
                }
            }
            atomicAdd(&sh[c - 1], 1);
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < n_cells; c += blockDim.x) hist[(int64_t)c * nchunks + blockIdx.x] = sh[c];
}

// One wave per cell scans that cell's row of the count matrix over chunks
// (exclusive, in place) and leaves the cell total; the block that finishes
// last turns the totals into cell starts and writes the PIA (group 1 only,
// group 2 cleared, as grid_sorting.jl:76-98 writes it).
__global__ void __launch_bounds__(256) k_sort_rows(MzSpView sp, int32_t n_cells, int32_t nchunks,
                                                   int32_t *__restrict__ hist, int32_t *__restrict__ cstart,
                                                   uint32_t *__restrict__ ticket) {
    __shared__ int32_t is_last;
    __shared__ int32_t wsum[4];
    __shared__ int32_t carry_s;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = blockIdx.x * 4 + wv;
    if (c < n_cells) {
        int32_t *row = hist + (int64_t)c * nchunks;
        int32_t carry = 0;
        for (int b0 = 0; b0 < nchunks; b0 += 64) {
            int b = b0 + lane;
            int32_t v = b < nchunks ? row[b] : 0;
            int32_t x = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                int32_t y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            if (b < nchunks) row[b] = carry + x - v;
            carry += __shfl(x, 63, 64);
        }
        if (lane == 0) __hip_atomic_store(&cstart[c], carry, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    mz_wait_stores();
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = atomicAdd(ticket, 1u);
        is_last = (t == gridDim.x - 1);
        carry_s = 0;
    }
    __syncthreads();
    if (!is_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const int32_t total = mz_ntotal(sp);
    for (int c0 = 0; c0 < n_cells; c0 += 256) {
        const int cc = c0 + threadIdx.x;
        int32_t v = cc < n_cells ? __hip_atomic_load(&cstart[cc], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
        int32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        int32_t pre = carry_s;
        for (int k = 0; k < wv; ++k) pre += wsum[k];
        int32_t excl = pre + x - v;
        if (cc < n_cells) {
            cstart[cc] = excl;
            sp.pia.start2[cc] = 0;
            sp.pia.end2[cc] = -1;
            sp.pia.n2[cc] = 0;
            sp.pia.start1[cc] = v > 0 ? excl + 1 : 0;
            sp.pia.end1[cc] = v > 0 ? excl + v : -1;
            sp.pia.n1[cc] = v;
            sp.pia.nloc[cc] = v;
        }
        __syncthreads();
        if (threadIdx.x == 255) carry_s = excl + v;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *ticket = 0;
        sp.ctr->contiguous = 1;
        sp.ctr->n_pos = total;
    }
}

// One block per chunk: the chunk's keys go to LDS, wave 0 ranks them in
// position order (64 at a time; lanes with equal keys are ranked with a
// ballot, so equal keys keep their order: stable, as the reference's reverse
// scatter), then all threads move the particles to their slots.
__global__ void __launch_bounds__(256) k_sort_scatter(MzSpView a, MzSpView b, int32_t n_cells, int32_t nchunks,
                                                     const int32_t *__restrict__ rows,
                                                     const int32_t *__restrict__ cstart) {
    extern __shared__ int32_t run[];   // [n_cells]: next slot of each cell for this chunk
    __shared__ int32_t keys[MZ_SORT_CHUNK];
    __shared__ int32_t dst[MZ_SORT_CHUNK];
    const int32_t n = mz_ntotal(a);
    const int64_t p0 = (int64_t)blockIdx.x * MZ_SORT_CHUNK;
    if (p0 >= n) return;
    const int cnt = (int)min((int64_t)MZ_SORT_CHUNK, (int64_t)n - p0);
    for (int k = threadIdx.x; k < n_cells; k += 256) run[k] = cstart[k] + rows[(int64_t)k * nchunks + blockIdx.x];
    for (int i = threadIdx.x; i < cnt; i += 256) {
        int32_t k = a.key[p0 + i];
        keys[i] = k < 1 ? 1 : (k > n_cells ? n_cells : k);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        volatile int32_t *vrun = run;
        const int lane = threadIdx.x;
        const uint64_t lt = mz_lanemask_lt();
        for (int bt = 0; bt < cnt; bt += 64) {
            const int i = bt + lane;
            const bool valid = i < cnt;
            const int32_t key = valid ? keys[i] : 0;
            uint64_t active = __ballot(valid);
            while (active) {
                const int leader = __ffsll((unsigned long long)active) - 1;
                const int32_t k = __shfl(key, leader, 64);
                const uint64_t mask = __ballot(valid && key == k);
                const int32_t bk = vrun[k - 1];
                if (valid && key == k) dst[i] = bk + __popcll(mask & lt);
                vrun[k - 1] = bk + __popcll(mask);
                active &= ~mask;
            }
        }
    }
    __syncthreads();
#pragma unroll 4
    for (int i = threadIdx.x; i < cnt; i += 256) {
        const int64_t p = p0 + i, d = dst[i];
        b.x0[d] = a.x0[p]; b.x1[d] = a.x1[p]; b.x2[d] = a.x2[p];
        b.v0[d] = a.v0[p]; b.v1[d] = a.v1[p]; b.v2[d] = a.v2[p];
        b.w[d] = a.w[p];
        b.id[d] = a.id[p];
    }
}

// Sorting with many cells.  The count matrix of the path above holds
// n_cells x chunks entries, so its traffic grows with the cell count (12,288
// cells over 12.8M particles: 1.2 GB per sort).  Past a bound the sort runs
// as a least-significant-digit radix sort of the (key, position) pairs on
// 8-bit digits of key - 1, whose matrices hold 256 x chunks entries:
//   k_rs_keys     keys (as k_sort_hist computes them), the clamped copy the
//                 passes read, and the cell totals (wave-aggregated atomics);
//   per digit:    k_rs_hist (per-chunk digit histogram), k_rs_rows (one block
//                 per digit scans its row over the chunks), k_rs_scatter
//                 (stable rank inside the chunk, as k_sort_scatter ranks);
//   k_rs_gather   moves the particle data by the final permutation;
//   k_rs_pia      scans the cell totals into the PIA.
// Every pass is stable, so the result is the stable sort by cell, the same
// permutation as the count-matrix path and the reference's reverse scatter.

#define MZ_RS_CHUNK 4096

__global__ void k_rs_zero(int32_t *__restrict__ a, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = 0;
}

__global__ void __launch_bounds__(256) k_rs_keys(MzSpView sp, MzGrid g, int32_t compute_cells, int32_t n_cells,
                                                 int32_t *__restrict__ kout, int32_t *__restrict__ counts) {
    const int32_t n = mz_ntotal(sp);
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p0 = (int64_t)blockIdx.x * blockDim.x; p0 < n; p0 += stride) {
        const int64_t p = p0 + threadIdx.x;
        const bool valid = p < n;
        int32_t c = 0;
        if (valid) {
            if (compute_cells) {
                c = mz_cell_of(sp.x0[p], g, n_cells, &sp.ctr->err);
                sp.key[p] = c;
            } else {
                c = sp.key[p];
                if (c < 1 || c > n_cells) {
                    atomicOr(&sp.ctr->err, MZ_ERR_CAPACITY);
                    c = c < 1 ? 1 : n_cells;
                }
            }
            kout[p] = c;
        }
        uint64_t active = __ballot(valid);
        while (active) {
            const int leader = __ffsll((unsigned long long)active) - 1;
            const int32_t k = __shfl(c, leader, 64);
            const uint64_t m = __ballot(valid && c == k);
            if (lane == leader) atomicAdd(&counts[k - 1], (int32_t)__popcll(m));
            active &= ~m;
        }
    }
}

__global__ void __launch_bounds__(256) k_rs_hist(MzSpView sp, const int32_t *__restrict__ ksrc, int32_t shift,
                                                 int32_t nchunks, int32_t *__restrict__ mat) {
    __shared__ int32_t h[256];
    const int32_t n = mz_ntotal(sp);
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t p0 = (int64_t)blockIdx.x * MZ_RS_CHUNK;
    for (int i = threadIdx.x; i < MZ_RS_CHUNK; i += 256) {
        const int64_t p = p0 + i;
        if (p < n) atomicAdd(&h[((ksrc[p] - 1) >> shift) & 255], 1);
    }
    __syncthreads();
    mat[(int64_t)threadIdx.x * nchunks + blockIdx.x] = h[threadIdx.x];
}

// One block per digit: exclusive scan of its row over the chunks, in place;
// the row total goes to tot[digit].
__global__ void __launch_bounds__(256) k_rs_rows(int32_t nchunks, int32_t *__restrict__ mat,
                                                 int32_t *__restrict__ tot) {
    __shared__ int32_t wsum[4];
    __shared__ int32_t carry_s;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int32_t *row = mat + (int64_t)blockIdx.x * nchunks;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nchunks; b0 += 256) {
        const int b = b0 + threadIdx.x;
        const int32_t v = b < nchunks ? row[b] : 0;
        int32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        int32_t pre = carry_s;
        for (int k = 0; k < wv; ++k) pre += wsum[k];
        if (b < nchunks) row[b] = pre + x - v;
        __syncthreads();
        if (threadIdx.x == 255) carry_s = pre + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry_s;
}

// One block per chunk.  psrc == nullptr means the identity permutation.
// kdst == nullptr skips the keys (the last pass needs only the permutation).
__global__ void __launch_bounds__(256) k_rs_scatter(MzSpView sp, const int32_t *__restrict__ ksrc,
                                                    const int32_t *__restrict__ psrc, int32_t *__restrict__ kdst,
                                                    int32_t *__restrict__ pdst, int32_t shift, int32_t nchunks,
                                                    const int32_t *__restrict__ mat,
                                                    const int32_t *__restrict__ tot) {
    __shared__ int32_t run[256];
    __shared__ int32_t dig[MZ_RS_CHUNK];
    __shared__ int32_t dst[MZ_RS_CHUNK];
    __shared__ int32_t wsum[4];
    const int32_t n = mz_ntotal(sp);
    const int64_t p0 = (int64_t)blockIdx.x * MZ_RS_CHUNK;
    if (p0 >= n) return;
    const int cnt = (int)min((int64_t)MZ_RS_CHUNK, (int64_t)n - p0);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    {   // digit starts: exclusive scan of the 256 row totals, plus this chunk's offset
        const int32_t v = tot[threadIdx.x];
        int32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        int32_t pre = 0;
        for (int k = 0; k < wv; ++k) pre += wsum[k];
        run[threadIdx.x] = pre + x - v + mat[(int64_t)threadIdx.x * nchunks + blockIdx.x];
    }
    for (int i = threadIdx.x; i < cnt; i += 256) dig[i] = ((ksrc[p0 + i] - 1) >> shift) & 255;
    __syncthreads();
    if (threadIdx.x < 64) {
        volatile int32_t *vrun = run;
        const uint64_t lt = mz_lanemask_lt();
        for (int bt = 0; bt < cnt; bt += 64) {
            const int i = bt + lane;
            const bool valid = i < cnt;
            const int32_t d = valid ? dig[i] : -1;
            uint64_t active = __ballot(valid);
            while (active) {
                const int leader = __ffsll((unsigned long long)active) - 1;
                const int32_t k = __shfl(d, leader, 64);
                const uint64_t m = __ballot(valid && d == k);
                const int32_t bk = vrun[k];
                if (valid && d == k) dst[i] = bk + __popcll(m & lt);
                vrun[k] = bk + __popcll(m);
                active &= ~m;
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += 256) {
        const int64_t p = p0 + i;
        const int32_t d = dst[i];
        if (kdst) kdst[d] = ksrc[p];
        pdst[d] = psrc ? psrc[p] : (int32_t)p;
    }
}

__global__ void __launch_bounds__(256) k_rs_gather(MzSpView a, MzSpView b, const int32_t *__restrict__ perm) {
    const int32_t n = mz_ntotal(a);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; d < n; d += stride) {
        const int64_t p = perm[d];
        b.x0[d] = a.x0[p]; b.x1[d] = a.x1[p]; b.x2[d] = a.x2[p];
        b.v0[d] = a.v0[p]; b.v1[d] = a.v1[p]; b.v2[d] = a.v2[p];
        b.w[d] = a.w[p];
        b.id[d] = a.id[p];
    }
}

// One block: the cell totals scanned into the PIA (group 1 only, group 2
// cleared, as grid_sorting.jl:76-98 writes it).
__global__ void __launch_bounds__(1024) k_rs_pia(MzSpView sp, int32_t n_cells, const int32_t *__restrict__ counts) {
    __shared__ int32_t wsum[16];
    __shared__ int32_t carry_s;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    if (t == 0) carry_s = 0;
    __syncthreads();
    for (int c0 = 0; c0 < n_cells; c0 += 1024) {
        const int c = c0 + t;
        const int32_t v = c < n_cells ? counts[c] : 0;
        int32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        int32_t pre = carry_s;
        for (int k = 0; k < wv; ++k) pre += wsum[k];
        const int32_t excl = pre + x - v;
        if (c < n_cells) {
            sp.pia.start2[c] = 0;
            sp.pia.end2[c] = -1;
            sp.pia.n2[c] = 0;
            sp.pia.start1[c] = v > 0 ? excl + 1 : 0;
            sp.pia.end1[c] = v > 0 ? excl + v : -1;
            sp.pia.n1[c] = v;
            sp.pia.nloc[c] = v;
        }
        __syncthreads();
        if (t == 1023) carry_s = excl + v;
        __syncthreads();
    }
    if (t == 0) {
        sp.ctr->contiguous = 1;
        sp.ctr->n_pos = mz_ntotal(sp);
    }
}

// the count matrix costs 4 B x cells x chunks; past 32 B per particle slot
// (or past the LDS histogram's cell bound) the radix path moves less
static bool sort_uses_radix(const mz_store *h, const MzSpecies &s) {
    const int64_t nchunks = (s.cap + MZ_SORT_CHUNK - 1) / MZ_SORT_CHUNK;
    return h->sort_path == 2 ||
           (h->sort_path == 0 && (h->n_cells > MZ_SORT_MAX_CELLS || (int64_t)h->n_cells * nchunks > 8 * s.cap));
}

struct RsScratch {
    int32_t *kA, *kB, *pA, *pB, *mat, *tot, *counts;
};

static bool rs_scratch(mz_store *h, const MzSpecies &s, RsScratch &r) {
    const int64_t cap = s.cap;
    const int64_t nchunks = (cap + MZ_RS_CHUNK - 1) / MZ_RS_CHUNK;
    r.kA = (int32_t *)mz_scratch(h, "rs_kA", (size_t)cap * sizeof(int32_t));
    r.kB = (int32_t *)mz_scratch(h, "rs_kB", (size_t)cap * sizeof(int32_t));
    r.pA = (int32_t *)mz_scratch(h, "rs_pA", (size_t)cap * sizeof(int32_t));
    r.pB = (int32_t *)mz_scratch(h, "rs_pB", (size_t)cap * sizeof(int32_t));
    r.mat = (int32_t *)mz_scratch(h, "rs_mat", (size_t)256 * nchunks * sizeof(int32_t));
    r.tot = (int32_t *)mz_scratch(h, "rs_tot", 256 * sizeof(int32_t));
    r.counts = (int32_t *)mz_scratch(h, "rs_counts", (size_t)h->n_cells * sizeof(int32_t));
    return r.kA && r.kB && r.pA && r.pB && r.mat && r.tot && r.counts;
}

// allocates the scratch the next sort of this species uses (before a graph
// capture, which must not allocate)
int mz_sort_prealloc(mz_store *h, int species) {
    const MzSpecies &s = h->sp[species - 1];
    if (sort_uses_radix(h, s)) {
        RsScratch r;
        MZ_CHECK(rs_scratch(h, s, r), "sort: scratch allocation failed");
        return 0;
    }
    const int64_t nchunks = (s.cap + MZ_SORT_CHUNK - 1) / MZ_SORT_CHUNK;
    MZ_CHECK(mz_scratch(h, "sort_hist", (size_t)h->n_cells * nchunks * sizeof(int32_t)) != nullptr &&
                 mz_scratch(h, "sort_cstart", (size_t)h->n_cells * sizeof(int32_t)) != nullptr,
             "sort: scratch allocation failed");
    return 0;
}

static int launch_radix_sort(mz_store *h, MzSpecies &s, MzSpView a, MzSpView b, int compute_cells) {
    const int nc = h->n_cells;
    const int64_t cap = s.cap;
    const int32_t nchunks = (int32_t)((cap + MZ_RS_CHUNK - 1) / MZ_RS_CHUNK);
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < (int64_t)nc) ++bits;
    const int passes = (bits + 7) / 8;
    RsScratch rs;
    MZ_CHECK(rs_scratch(h, s, rs), "sort: scratch allocation failed");
    int32_t *kA = rs.kA, *kB = rs.kB, *pA = rs.pA, *pB = rs.pB, *mat = rs.mat, *tot = rs.tot, *counts = rs.counts;
    const int gblocks = (int)std::min<int64_t>((cap + 255) / 256, 8192);
    hipLaunchKernelGGL(k_rs_zero, dim3((nc + 255) / 256), dim3(256), 0, h->stream, counts, (int64_t)nc);
    hipLaunchKernelGGL(k_rs_keys, dim3(gblocks), dim3(256), 0, h->stream, a, h->grid, compute_cells, nc, kA, counts);
    int32_t *ks = kA, *kd = kB, *ps = nullptr, *pd = pA;
    for (int q = 0; q < passes; ++q) {
        const int shift = 8 * q;
        const bool last = q == passes - 1;
        hipLaunchKernelGGL(k_rs_hist, dim3(nchunks), dim3(256), 0, h->stream, a, ks, shift, nchunks, mat);
        hipLaunchKernelGGL(k_rs_rows, dim3(256), dim3(256), 0, h->stream, nchunks, mat, tot);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nchunks), dim3(256), 0, h->stream, a, ks, ps, last ? nullptr : kd, pd,
                           shift, nchunks, mat, tot);
        ps = pd;
        pd = pd == pA ? pB : pA;
        std::swap(ks, kd);
    }
    hipLaunchKernelGGL(k_rs_gather, dim3(gblocks), dim3(256), 0, h->stream, a, b, ps);
    hipLaunchKernelGGL(k_rs_pia, dim3(1), dim3(1024), 0, h->stream, a, nc, counts);
    MZ_HIP(hipGetLastError());
    return 0;
}

MzViews mz_views(mz_store *h) {
    MzViews v{};
    for (int s = 0; s < h->n_species; ++s) v.v[s] = mz_view(h->sp[s]);
    return v;
}

void mz_swap_buffers(MzSpecies &s) {
    for (int d = 0; d < 3; ++d) {
        std::swap(s.x[d], s.xb[d]);
        std::swap(s.v[d], s.vb[d]);
    }
    std::swap(s.w, s.wb);
    std::swap(s.id, s.idb);
}

static MzSpView view_b(const MzSpecies &s) {
    MzSpView v = mz_view(s);
    v.x0 = s.xb[0]; v.x1 = s.xb[1]; v.x2 = s.xb[2];
    v.v0 = s.vb[0]; v.v1 = s.vb[1]; v.v2 = s.vb[2];
    v.w = s.wb; v.id = s.idb;
    return v;
}

// launches the three sort kernels on the current A buffers; the caller swaps
// The count-matrix sort's three kernels over the count matrix hist
// [n_cells][chunks] and the cell starts cstart [n_cells]; every launch
// overwrites both in full, so neither needs clearing.
static int launch_count_sort(mz_store *h, MzSpecies &s, int compute_cells, int32_t *hist, int32_t *cstart) {
    const int nc = h->n_cells;
    const int32_t nchunks = (int32_t)((s.cap + MZ_SORT_CHUNK - 1) / MZ_SORT_CHUNK);
    MzSpView a = mz_view(s), b = view_b(s);
    size_t lds = (size_t)nc * sizeof(int32_t);
    hipLaunchKernelGGL(k_sort_hist, dim3(nchunks), dim3(256), lds, h->stream, a, h->grid, compute_cells, nc,
                       nchunks, hist);
    hipLaunchKernelGGL(k_sort_rows, dim3((nc + 3) / 4), dim3(256), 0, h->stream, a, nc, nchunks, hist, cstart,
                       h->d_ticket);
    hipLaunchKernelGGL(k_sort_scatter, dim3(nchunks), dim3(256), lds, h->stream, a, b, nc, nchunks, hist, cstart);
    MZ_HIP(hipGetLastError());
    return 0;
}

int mz_launch_sort(mz_store *h, int species, int compute_cells) {
    MzSpecies &s = h->sp[species - 1];
    int nc = h->n_cells;
    MZ_CHECK(!compute_cells || h->has_grid, "sort: no grid set");
    int32_t nchunks = (int32_t)((s.cap + MZ_SORT_CHUNK - 1) / MZ_SORT_CHUNK);
    const bool radix = sort_uses_radix(h, s);
    MZ_CHECK(radix || nc <= MZ_SORT_MAX_CELLS, "sort: too many cells for the LDS histogram");
    if (radix) return launch_radix_sort(h, s, mz_view(s), view_b(s), compute_cells);
    int32_t *hist = (int32_t *)mz_scratch(h, "sort_hist", (size_t)nc * nchunks * sizeof(int32_t));
    int32_t *cstart = (int32_t *)mz_scratch(h, "sort_cstart", (size_t)nc * sizeof(int32_t));
    MZ_CHECK(hist != nullptr && cstart != nullptr, "sort: scratch allocation failed");
    return launch_count_sort(h, s, compute_cells, hist, cstart);
}

extern "C" int mz_set_sort_path(mz_store *h, int32_t path) {
    MZ_CHECK(path >= 0 && path <= 2, "set_sort_path: path must be 0 (automatic), 1 (count matrix) or 2 (radix)");
    h->sort_path = path;
    mz_graph_release(h);
    return 0;
}

extern "C" int mz_sort(mz_store *h, int32_t species, int32_t compute_cells) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "sort: bad species");
    int r = 0;
    if (!h->sp[species - 1].contig_known) {
        MzCounters c;
        r = mz_read_counters(h, species, &c);
        if (r) return r;
        if (!c.contiguous) {
            r = mz_squash(h, species);
            if (r) return r;
        }
    }
    {
        MzTimed tm(h, "sort");
        r = mz_launch_sort(h, species, compute_cells);
    }
    if (r) return r;
    mz_swap_buffers(h->sp[species - 1]);
    h->sp[species - 1].contig_known = true;
    return 0;
}

// ===========================================================================
// The sort without allocating: the count matrix, the cell starts and the
// squash's group starts live in `scratch`, a fixed device array that each GPU
// holds from the library's load to the process's end.  Host code never
// passes `scratch` itself to a kernel (in host code the name is a stand-in,
// not a device address): the device address comes from hipGetSymbolAddress,
// once per device, and is kept in g_scratch_dev.  Every call on a GPU shares
// that GPU's array, so calls that overlap on one GPU race.
static __device__ __align__(256) int32_t scratch[MZ_NOALLOC_SCRATCH_WORDS];
static int launch_squash(mz_store *h, MzSpecies &s, int32_t *news);
#define MZ_MAX_DEVICES 64
static int32_t *g_scratch_dev[MZ_MAX_DEVICES];
static std::mutex g_scratch_mu;

// Looks up (first call on the current device) or returns the device address
// of `scratch` on the current device.
int mz_noalloc_scratch_init(int32_t **out) {
    int dev = 0;
    MZ_HIP(hipGetDevice(&dev));
    MZ_CHECK(dev >= 0 && dev < MZ_MAX_DEVICES, "no-alloc scratch: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            mz_set_error(std::string("no-alloc scratch: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
            return -1;
        }
        g_scratch_dev[dev] = (int32_t *)p;
    }
    if (out) *out = g_scratch_dev[dev];
    return 0;
}

extern "C" int mz_sort_noalloc(mz_store *h, int32_t species, int32_t compute_cells) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "sort_noalloc: bad species");
    MzSpecies &s = h->sp[species - 1];
    MZ_CHECK(!compute_cells || h->has_grid, "sort_noalloc: no grid set");
    MZ_CHECK(!sort_uses_radix(h, s),
             "sort_noalloc: only the count-matrix sort runs without allocating (at most 14336 cells, n_cells x "
             "chunks <= 8 x capacity, sort path not radix)");
    const int64_t nc = h->n_cells;
    const int64_t nchunks = (s.cap + MZ_SORT_CHUNK - 1) / MZ_SORT_CHUNK;
    const int64_t need = std::max<int64_t>(nc * nchunks + nc, 2 * nc);
    MZ_CHECK(need <= MZ_NOALLOC_SCRATCH_WORDS,
             "sort_noalloc: n_cells x (chunks + 1) exceeds the fixed scratch of 2^21 int32 words (8 MiB); "
             "chunks = ceil(capacity / 1024)");
    int32_t *scr = nullptr;
    int r = mz_noalloc_scratch_init(&scr);
    if (r) return r;
    if (!s.contig_known) {
        MzCounters c;
        r = mz_read_counters(h, species, &c);
        if (r) return r;
        if (!c.contiguous) {
            r = launch_squash(h, s, scr);
            if (r) return r;
        }
        s.contig_known = true;
    }
    // no timing bracket: its events would be allocations
    r = launch_count_sort(h, s, compute_cells, scr, scr + nc * nchunks);
    if (r) return r;
    mz_swap_buffers(s);
    return 0;
}

// ===========================================================================
// squash_pia! (particles.jl:622-701).  Group-1 segments are packed to the
// front in cell order, then group-2 segments follow in cell order; each
// segment keeps its internal order.  k_squash_plan computes the new starts
// with a scan; k_squash_move copies each cell's segments into the B buffers.

__global__ void __launch_bounds__(1024) k_squash_plan(MzSpView sp, int32_t n_cells, int32_t *__restrict__ news) {
    // news[c] = new start1 (0-based), news[n_cells + c] = new start2 (0-based)
    __shared__ int32_t wsum[16];
    __shared__ int32_t carry;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
        const int32_t *cnt = pass == 0 ? sp.pia.n1 : sp.pia.n2;
        for (int c0 = 0; c0 < n_cells; c0 += 1024) {
            int c = c0 + t;
            int32_t v = c < n_cells ? cnt[c] : 0;
            int32_t x = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                int32_t y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            if (lane == 63) wsum[wv] = x;
            __syncthreads();
            if (t < 16) {
                int32_t a = wsum[t], b = a;
                for (int o = 1; o < 16; o <<= 1) {
                    int32_t y = __shfl_up(b, o, 16);
                    if ((t & 15) >= o) b += y;
                }
                wsum[t] = b - a;
            }
            __syncthreads();
            int32_t excl = carry + wsum[wv] + x - v;
            if (c < n_cells) news[pass * n_cells + c] = excl;
            __syncthreads();
            if (t == 1023) carry = excl + v;
            __syncthreads();
        }
    }
}

__global__ void k_squash_move(MzSpView a, MzSpView b, int32_t n_cells, const int32_t *__restrict__ news) {
    const int c = blockIdx.x;
    for (int g = 0; g < 2; ++g) {
        int32_t cnt = g == 0 ? a.pia.n1[c] : a.pia.n2[c];
        int32_t src = (g == 0 ? a.pia.start1[c] : a.pia.start2[c]) - 1;
        int32_t dst = news[g * n_cells + c];
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
            int64_t p = src + i, q = dst + i;
            b.x0[q] = a.x0[p]; b.x1[q] = a.x1[p]; b.x2[q] = a.x2[p];
            b.v0[q] = a.v0[p]; b.v1[q] = a.v1[p]; b.v2[q] = a.v2[p];
            b.w[q] = a.w[p]; b.id[q] = a.id[p]; b.key[q] = a.key[p];
        }
    }
}

__global__ void k_squash_pia(MzSpView sp, int32_t n_cells, const int32_t *__restrict__ news) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n_cells) {
        int32_t n1 = sp.pia.n1[c], n2 = sp.pia.n2[c];
        if (n1 > 0) { sp.pia.start1[c] = news[c] + 1; sp.pia.end1[c] = news[c] + n1; }
        if (n2 > 0) { sp.pia.start2[c] = news[n_cells + c] + 1; sp.pia.end2[c] = news[n_cells + c] + n2; }
    }
    if (c == 0) {
        sp.ctr->contiguous = 1;
        sp.ctr->n_pos = sp.ctr->n_total;
    }
}

// The squash kernels over the new group starts news [2][n_cells] (written
// in full by k_squash_plan); the species' buffers swap, and it is contiguous.
static int launch_squash(mz_store *h, MzSpecies &s, int32_t *news) {
    const int nc = h->n_cells;
    MzSpView a = mz_view(s), b = view_b(s);
    b.key = s.keyb;
    hipLaunchKernelGGL(k_squash_plan, dim3(1), dim3(1024), 0, h->stream, a, nc, news);
    hipLaunchKernelGGL(k_squash_move, dim3(nc), dim3(256), 0, h->stream, a, b, nc, news);
    hipLaunchKernelGGL(k_squash_pia, dim3((nc + 255) / 256), dim3(256), 0, h->stream, a, nc, news);
    MZ_HIP(hipGetLastError());
    mz_swap_buffers(s);
    std::swap(s.key, s.keyb);
    s.contig_known = true;
    return 0;
}

extern "C" int mz_squash(mz_store *h, int32_t species) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "squash: bad species");
    if (h->sp[species - 1].contig_known) return 0;
    MzCounters c;
    int r = mz_read_counters(h, species, &c);
    if (r) return r;
    if (c.contiguous) {
        h->sp[species - 1].contig_known = true;
        return 0;
    }
    int32_t *news = (int32_t *)mz_scratch(h, "squash", (size_t)2 * h->n_cells * sizeof(int32_t));
    MZ_CHECK(news != nullptr, "squash: scratch allocation failed");
    return launch_squash(h, h->sp[species - 1], news);
}

// ===========================================================================
// restore_particle_ordering! (particles.jl:1086-1137).  Afterwards index[i]
// == i for every position and the free-id buffer reads length, length - 1,
// ...; every logical position keeps its particle.  The store already holds
// each particle's data at its logical position, so the reference's swaps
// of particle data have nothing to move here: only the identities change.
__global__ void k_restore_ordering(MzSpView sp) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t len = sp.ctr->length, nb = sp.ctr->nbuffer;
    if (p < sp.cap && p < len) sp.id[p] = (int32_t)(p + 1);
    if (p < sp.bufcap && p < nb) sp.buffer[p] = (int32_t)(len - p);
}

extern "C" int mz_restore_particle_ordering(mz_store *h, int32_t species) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "restore_particle_ordering: bad species");
    MzSpecies &s = h->sp[species - 1];
    const int64_t n = std::max<int64_t>(s.cap, s.bufcap);
    hipLaunchKernelGGL(k_restore_ordering, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, mz_view(s));
    MZ_HIP(hipGetLastError());
    return 0;
}

// ===========================================================================
// Convection (convection_1D.jl:17-54, boundary_conditions.jl:79-121).
// x + v t is an explicit fma, matching the reference's @muladd block, so the
// cell keys that follow are reproducible bit for bit.  One thread per
// particle; a particle draws from stream (id - 1).

__global__ void k_convect(MzSpView sp, MzGrid g, MzWalls wl, int sidx, double dt, MzRngParams rp,
                          const uint8_t *__restrict__ mask) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    rp = mz_rng_resolve(rp);
    const int32_t lim = mask ? mz_npos(sp) : mz_ntotal(sp);
    if (p >= lim) return;
    if (mask && !mask[p]) return;
    double x0 = sp.x0[p], v0 = sp.v0[p];
    double x_new = fma(v0, dt, x0);
    if (x_new < g.L && x_new > 0.0) {
        // free flight (the common case): only x changes
        sp.x0[p] = x_new < g.min_x ? g.min_x : (x_new > g.max_x ? g.max_x : x_new);
        return;
    }
    double v1 = sp.v1[p], v2 = sp.v2[p];
    mz_convect_one(rp, g, wl, sidx, dt, sp.id[p], x0, v0, v1, v2, &sp.ctr->err);
    sp.x0[p] = x0;
    sp.v0[p] = v0;
    sp.v1[p] = v1;
    sp.v2[p] = v2;
}

__global__ void k_mark_valid(MzSpView sp, int32_t n_cells, uint8_t *__restrict__ mask) {
    const int c = blockIdx.x;
    for (int g = 0; g < 2; ++g) {
        int32_t cnt = g == 0 ? sp.pia.n1[c] : sp.pia.n2[c];
        int32_t st = (g == 0 ? sp.pia.start1[c] : sp.pia.start2[c]) - 1;
        for (int i = threadIdx.x; i < cnt; i += blockDim.x)
            if (st + i >= 0 && st + i < sp.cap) mask[st + i] = 1;
    }
}

int mz_launch_convect(mz_store *h, int species, double dt, uint64_t tag, int rng_mode, bool contiguous,
                      int32_t n_pos, const uint64_t *step_ptr) {
    MzSpecies &s = h->sp[species - 1];
    MZ_CHECK(h->has_grid && h->has_walls, "convect: grid or walls not set");
    uint8_t *mask = nullptr;
    if (!contiguous) {
        mask = (uint8_t *)mz_scratch(h, "convect_mask", (size_t)s.cap);
        MZ_CHECK(mask != nullptr, "convect: scratch allocation failed");
        MZ_HIP(hipMemsetAsync(mask, 0, (size_t)s.cap, h->stream));
        hipLaunchKernelGGL(k_mark_valid, dim3(h->n_cells), dim3(256), 0, h->stream, mz_view(s), h->n_cells, mask);
    }
    MzRngParams rp = mz_rng_params(h, 1, tag, rng_mode);
    rp.step_ptr = step_ptr;
    rp.phase = 2;
    int64_t n = contiguous ? s.cap : std::max<int64_t>(n_pos, 1);
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_convect, dim3(blocks), dim3(256), 0, h->stream, mz_view(s), h->grid, h->walls,
                       species - 1, dt, rp, mask);
    MZ_HIP(hipGetLastError());
    return 0;
}

extern "C" int mz_convect(mz_store *h, int32_t species, double dt, uint64_t tag, int32_t rng_mode) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "convect: bad species");
    MZ_CHECK(rng_mode == 0 || (rng_mode == 1 && h->d_replay[1]), "convect: replay table not set");
    // a species known to be contiguous uses positions [0, n_total), which the
    // kernel reads on the device; otherwise the extent comes from the counters
    MzCounters c{};
    c.contiguous = 1;
    if (!h->sp[species - 1].contig_known) {
        int r = mz_read_counters(h, species, &c);
        if (r) return r;
    }
    MzTimed tm(h, "convect");
    return mz_launch_convect(h, species, dt, tag, rng_mode, c.contiguous != 0, c.n_pos);
}

// ===========================================================================
// Convection with surface properties (convection_1D.jl:72-112, 176-207;
// surface_props.jl:77-189).  Only particles whose flight ends outside
// (0, L) touch a wall; each such thread sums its contacts (MzSurfAcc), a
// block with any contact reduces them (wave sums, then its waves in order)
// into its row of partials and lists itself, and k_surf_finalize adds the
// listed rows in block order.  (A last-block pass inside the convection
// kernel costs one same-address atomic per block; at 2,000 blocks that takes
// longer than a second launch.)  The sums are
// thus the same on every run; they differ from the reference's particle-order
// sum only by rounding.

#define MZ_SURF_BLOCK 256

__global__ void k_fill_u8(uint8_t *__restrict__ a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

// Wave sums of the 2 x MZ_SURF_Q accumulators, written by lane 0.  The
// recursion spells out every index at compile time, so the accumulators stay
// in registers (a counted loop around the reduction calls keeps them in
// scratch memory).
template <int I>
__device__ inline void surf_wave_sums(const MzSurfAcc &sa, double *dst, int lane) {
    if constexpr (I < 2 * MZ_SURF_Q) {
        const double t = mz_wave_sum(sa.q[I / MZ_SURF_Q][I % MZ_SURF_Q]);
        if (lane == 0) dst[I] = t;
        surf_wave_sums<I + 1>(sa, dst, lane);
    }
}

// Where k_surf_finalize puts the wall properties.
struct MzSurfOut {
    double *surf;              // the store's sets
    uint32_t *ticket;          // [1]: blocks with contacts listed by the last convection
    int32_t n_species;
    double f0, f1;             // (mass / dt) / area of the two walls
    double avg_inv;
};

__global__ void __launch_bounds__(MZ_SURF_BLOCK) k_convect_surf(MzSpView sp, MzGrid g, MzWalls wl, int sidx,
                                                                double dt, MzRngParams rp,
                                                                const uint8_t *__restrict__ mask,
                                                                double *__restrict__ part,
                                                                int32_t *__restrict__ pflag,
                                                                int32_t *__restrict__ plist, MzSurfOut so) {
    constexpr int NW = MZ_SURF_BLOCK / 64, Q2 = 2 * MZ_SURF_Q;
    __shared__ double red[NW][Q2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + tid;
    rp = mz_rng_resolve(rp);
    const int32_t lim = mask ? mz_npos(sp) : mz_ntotal(sp);
    const bool active = p < lim && (!mask || mask[p]);
    MzSurfAcc sa;
    sa.zero();
    int hit = 0;
    if (active) {
        double x0 = sp.x0[p], v0 = sp.v0[p];
        const double x_new = fma(v0, dt, x0);
        if (x_new < g.L && x_new > 0.0) {
            sp.x0[p] = x_new < g.min_x ? g.min_x : (x_new > g.max_x ? g.max_x : x_new);
        } else {
            hit = 1;
            double v1 = sp.v1[p], v2 = sp.v2[p];
            sa.w = sp.w[p];
            mz_convect_one_surf(sa, rp, g, wl, sidx, dt, sp.id[p], x0, v0, v1, v2, &sp.ctr->err);
            sp.x0[p] = x0;
            sp.v0[p] = v0;
            sp.v1[p] = v1;
            sp.v2[p] = v2;
        }
    }
    const int any = __syncthreads_or(hit);
    if (any) {
        surf_wave_sums<0>(sa, red[wave], lane);
        __syncthreads();
        if (tid < Q2) {
            double t = red[0][tid];
            for (int w = 1; w < NW; ++w) t += red[w][tid];
            part[(size_t)blockIdx.x * Q2 + tid] = t;
        }
        if (tid == 0) plist[atomicAdd(so.ticket + 1, 1u)] = (int32_t)blockIdx.x;
    }
    if (tid == 0) pflag[blockIdx.x] = any;
}

// One block: the rows of the listed blocks in block order (the list is
// sorted here; from the per-block flags if more than MZ_SURF_BLOCK blocks
// had contacts), scaled (surface_props_scale!: every quantity but np times
// (mass / dt) / area), written as the species' instantaneous set
// (clear_props! leaves the other species at zero) and, with avg_inv > 0,
// added to the time average (avg_props!).  The list count is left at zero.
__global__ void __launch_bounds__(MZ_SURF_BLOCK) k_surf_finalize(const double *__restrict__ part,
                                                                 const int32_t *__restrict__ pflag,
                                                                 const int32_t *__restrict__ plist, int32_t nblk,
                                                                 int32_t sidx, MzSurfOut so) {
    constexpr int Q2 = 2 * MZ_SURF_Q, NJ = MZ_SURF_BLOCK / Q2;
    __shared__ double red[NJ][Q2];
    __shared__ int32_t order[MZ_SURF_BLOCK];
    const int tid = threadIdx.x;
    const int32_t nl = (int32_t)so.ticket[1];
    const bool listed = nl <= MZ_SURF_BLOCK;
    if (listed) {
        // ids are distinct: rank by count
        const int32_t id = tid < nl ? plist[tid] : 0;
        order[tid] = id;
        __syncthreads();
        int rank = 0;
        for (int j = 0; j < nl; ++j) rank += order[j] < id;
        __syncthreads();
        if (tid < nl) order[rank] = id;
        __syncthreads();
    }
    // thread (j, q): quantity q over rows j, j + NJ, ... ; then the NJ parts in order
    if (tid < NJ * Q2) {
        const int q = tid % Q2, j = tid / Q2;
        double sum = 0.0;
        if (listed) {
            for (int i = j; i < nl; i += NJ) sum += part[(size_t)order[i] * Q2 + q];
        } else {
            for (int32_t b = j; b < nblk; b += NJ)
                if (pflag[b]) sum += part[(size_t)b * Q2 + q];
        }
        red[j][q] = sum;
    }
    __syncthreads();
    if (tid < Q2) {
        const int e = tid / MZ_SURF_Q, k = tid % MZ_SURF_Q;
        double v = red[0][tid];
        for (int j = 1; j < NJ; ++j) v += red[j][tid];
        const double inst = v * (k == 0 ? 1.0 : (e == 0 ? so.f0 : so.f1));
        const size_t nset = (size_t)so.n_species * Q2;
        for (int s = 0; s < so.n_species; ++s) {
            const size_t ix = ((size_t)s * 2 + e) * MZ_SURF_Q + k;
            so.surf[ix] = s == sidx ? inst : 0.0;
            if (s == sidx && so.avg_inv > 0.0) so.surf[nset + ix] = fma(inst, so.avg_inv, so.surf[nset + ix]);
        }
    }
    if (tid == 0) so.ticket[1] = 0;
}

// avg_props!(avg, inst, n) for every species: avg = muladd(inst, 1 / n, avg)
__global__ void k_surf_avg(double *__restrict__ avg, const double *__restrict__ inst, int32_t n, double inv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) avg[i] = fma(inst[i], inv, avg[i]);
}

__global__ void k_surf_fill(double *__restrict__ a, int32_t n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// Partial sums for every block a launch over the species' slots can have
// (allocated before a step graph is captured, which cannot allocate).
int mz_surf_prealloc(mz_store *h, int species) {
    const int64_t blocks = (h->sp[species - 1].cap + MZ_SURF_BLOCK - 1) / MZ_SURF_BLOCK;
    MZ_CHECK(mz_scratch(h, "surf_part", (size_t)blocks * 2 * MZ_SURF_Q * sizeof(double)) != nullptr &&
                 mz_scratch(h, "surf_flag", (size_t)blocks * sizeof(int32_t)) != nullptr &&
                 mz_scratch(h, "surf_list", (size_t)blocks * sizeof(int32_t)) != nullptr,
             "convect: scratch allocation failed");
    return 0;
}

int mz_launch_convect_surf(mz_store *h, int species, double dt, uint64_t tag, int rng_mode, bool contiguous,
                           int32_t n_pos, double mass, const uint64_t *step_ptr, double avg_inv) {
    MzSpecies &s = h->sp[species - 1];
    MZ_CHECK(h->has_grid && h->has_walls, "convect: grid or walls not set");
    MZ_CHECK(mass > 0.0 && dt > 0.0, "convect: surface properties need mass > 0 and dt > 0");
    uint8_t *mask = nullptr;
    if (!contiguous) {
        mask = (uint8_t *)mz_scratch(h, "convect_mask", (size_t)s.cap);
        MZ_CHECK(mask != nullptr, "convect: scratch allocation failed");
        hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((s.cap + 255) / 256)), dim3(256), 0, h->stream, mask,
                           (int64_t)s.cap);
        hipLaunchKernelGGL(k_mark_valid, dim3(h->n_cells), dim3(256), 0, h->stream, mz_view(s), h->n_cells, mask);
    }
    MzRngParams rp = mz_rng_params(h, 1, tag, rng_mode);
    rp.step_ptr = step_ptr;
    rp.phase = 2;
    const int64_t n = contiguous ? s.cap : std::max<int64_t>(n_pos, 1);
    const int blocks = (int)((n + MZ_SURF_BLOCK - 1) / MZ_SURF_BLOCK);
    int r = mz_surf_prealloc(h, species);
    if (r) return r;
    double *part = (double *)mz_scratch(h, "surf_part", 0);
    int32_t *pflag = (int32_t *)mz_scratch(h, "surf_flag", 0);
    const double fb = mass / dt;
    MzSurfOut so{h->d_surf, h->d_surf_ticket, h->n_species, fb * h->surf_inv_area[0], fb * h->surf_inv_area[1],
                 avg_inv};
    int32_t *plist = (int32_t *)mz_scratch(h, "surf_list", 0);
    hipLaunchKernelGGL(k_convect_surf, dim3(blocks), dim3(MZ_SURF_BLOCK), 0, h->stream, mz_view(s), h->grid,
                       h->walls, species - 1, dt, rp, mask, part, pflag, plist, so);
    hipLaunchKernelGGL(k_surf_finalize, dim3(1), dim3(MZ_SURF_BLOCK), 0, h->stream, part, pflag, plist, blocks,
                       species - 1, so);
    MZ_HIP(hipGetLastError());
    return 0;
}

extern "C" int mz_convect_surf(mz_store *h, int32_t species, double dt, uint64_t tag, int32_t rng_mode,
                               double mass) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "convect: bad species");
    MZ_CHECK(rng_mode == 0 || (rng_mode == 1 && h->d_replay[1]), "convect: replay table not set");
    // a species known to be contiguous uses positions [0, n_total), which the
    // kernel reads on the device; otherwise the extent comes from the counters
    MzCounters c{};
    c.contiguous = 1;
    if (!h->sp[species - 1].contig_known) {
        int r = mz_read_counters(h, species, &c);
        if (r) return r;
    }
    MzTimed tm(h, "convect");
    return mz_launch_convect_surf(h, species, dt, tag, rng_mode, c.contiguous != 0, c.n_pos, mass, nullptr, 0.0);
}

extern "C" int mz_set_surf_areas(mz_store *h, const double *areas) {
    MZ_CHECK(areas != nullptr, "set_surf_areas: areas is NULL");
    MZ_CHECK(areas[0] > 0.0 && areas[1] > 0.0, "set_surf_areas: areas must be positive");
    h->surf_inv_area[0] = 1.0 / areas[0];
    h->surf_inv_area[1] = 1.0 / areas[1];
    return 0;
}

int mz_launch_surf_avg(mz_store *h, double inv) {
    const int32_t n = h->n_species * 2 * MZ_SURF_Q;
    hipLaunchKernelGGL(k_surf_avg, dim3(1), dim3(256), 0, h->stream, h->d_surf + n, h->d_surf, n, inv);
    MZ_HIP(hipGetLastError());
    return 0;
}

extern "C" int mz_surf_avg(mz_store *h, double n_avg_timesteps) {
    MZ_CHECK(n_avg_timesteps != 0.0, "surf_avg: zero averaging count");
    return mz_launch_surf_avg(h, 1.0 / n_avg_timesteps);
}

// avg_props!(phys_props_avg, phys_props, n) (physical_props.jl:281-299):
// avg = muladd(props, 1 / n, avg) for np, n, v and T of every cell and
// species; lpa is averaged from the particle vector's length.
__global__ void k_props_avg(double *__restrict__ avg, const double *__restrict__ np, const double *__restrict__ nd,
                            const double *__restrict__ v, const double *__restrict__ T, int32_t pn, double inv,
                            MzViews views, int32_t n_species) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < pn) {
        avg[i] = fma(np[i], inv, avg[i]);
        avg[pn + i] = fma(nd[i], inv, avg[pn + i]);
        for (int d = 0; d < 3; ++d) avg[2 * pn + 3 * i + d] = fma(v[3 * i + d], inv, avg[2 * pn + 3 * i + d]);
        avg[5 * pn + i] = fma(T[i], inv, avg[5 * pn + i]);
    }
    if (i < n_species) avg[6 * pn + i] = fma((double)views.v[i].ctr->length, inv, avg[6 * pn + i]);
}

int mz_launch_props_avg(mz_store *h, double inv) {
    const int32_t pn = h->n_species * h->n_cells;
    hipLaunchKernelGGL(k_props_avg, dim3((pn + 255) / 256), dim3(256), 0, h->stream, h->d_pavg, h->d_np, h->d_n,
                       h->d_v, h->d_T, pn, inv, mz_views(h), h->n_species);
    MZ_HIP(hipGetLastError());
    return 0;
}

extern "C" int mz_props_avg(mz_store *h, double n_avg_timesteps) {
    MZ_CHECK(n_avg_timesteps != 0.0, "props_avg: zero averaging count");
    return mz_launch_props_avg(h, 1.0 / n_avg_timesteps);
}

extern "C" int mz_props_avg_clear(mz_store *h) {
    const int32_t n = h->n_species * h->n_cells * 6 + h->n_species;
    hipLaunchKernelGGL(k_surf_fill, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->d_pavg, n);
    MZ_HIP(hipGetLastError());
    return 0;
}

extern "C" int mz_get_props_avg(mz_store *h, double *lpa, double *np, double *n, double *v, double *T) {
    MZ_HIP(hipStreamSynchronize(h->stream));
    const size_t pn = (size_t)h->n_species * h->n_cells;
    if (np) MZ_HIP(hipMemcpy(np, h->d_pavg, pn * sizeof(double), hipMemcpyDeviceToHost));
    if (n) MZ_HIP(hipMemcpy(n, h->d_pavg + pn, pn * sizeof(double), hipMemcpyDeviceToHost));
    if (v) MZ_HIP(hipMemcpy(v, h->d_pavg + 2 * pn, 3 * pn * sizeof(double), hipMemcpyDeviceToHost));
    if (T) MZ_HIP(hipMemcpy(T, h->d_pavg + 5 * pn, pn * sizeof(double), hipMemcpyDeviceToHost));
    if (lpa) MZ_HIP(hipMemcpy(lpa, h->d_pavg + 6 * pn, h->n_species * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int mz_surf_clear(mz_store *h, int32_t which) {
    MZ_CHECK(which == 0 || which == 1, "surf_clear: which is 0 (instantaneous) or 1 (average)");
    const int32_t n = h->n_species * 2 * MZ_SURF_Q;
    hipLaunchKernelGGL(k_surf_fill, dim3(1), dim3(256), 0, h->stream, h->d_surf + (size_t)which * n, n);
    MZ_HIP(hipGetLastError());
    return 0;
}

extern "C" int mz_get_surf(mz_store *h, int32_t which, double *out) {
    MZ_CHECK(which == 0 || which == 1, "get_surf: which is 0 (instantaneous) or 1 (average)");
    MZ_CHECK(out != nullptr, "get_surf: out is NULL");
    const size_t n = (size_t)h->n_species * 2 * MZ_SURF_Q;
    MZ_HIP(hipStreamSynchronize(h->stream));
    MZ_HIP(hipMemcpy(out, h->d_surf + which * n, n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int mz_set_surf(mz_store *h, int32_t which, const double *in) {
    MZ_CHECK(which == 0 || which == 1, "set_surf: which is 0 (instantaneous) or 1 (average)");
    MZ_CHECK(in != nullptr, "set_surf: in is NULL");
    const size_t n = (size_t)h->n_species * 2 * MZ_SURF_Q;
    MZ_HIP(hipStreamSynchronize(h->stream));
    MZ_HIP(hipMemcpy(h->d_surf + which * n, in, n * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

// ===========================================================================
// Physical properties: one wave per (cell, species), one pass over the cell's
// particles.  The temperature uses shifted sums, E = sum w |v - s|^2 -
// n |vbar - s|^2 with s the cell's first particle velocity, which keeps the
// cancellation of the one-pass form at the level of the reference's
// two-pass sum while reading each particle once.

__global__ void __launch_bounds__(256) k_props(MzViews views, int32_t n_cells, int32_t variant,
                                               int32_t ndens, const double *__restrict__ inv_V_cells,
                                               const double *__restrict__ masses, double *__restrict__ out_np,
                                               double *__restrict__ out_n, double *__restrict__ out_v,
                                               double *__restrict__ out_T, double *__restrict__ avg,
                                               double avg_inv) {
    const int sidx = blockIdx.y;
    const MzSpView &sp = views.v[sidx];
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // avg_props! of the results when avg is set (avg = muladd(props, avg_inv,
    // avg), physical_props.jl:281-299; lpa from the vector's length)
    if (avg && blockIdx.x == 0 && threadIdx.x == 0) {
        const size_t pn = (size_t)gridDim.y * n_cells;
        avg[6 * pn + sidx] = fma((double)sp.ctr->length, avg_inv, avg[6 * pn + sidx]);
    }
    if (c >= n_cells) return;
    int32_t st1 = sp.pia.start1[c] - 1, n1 = sp.pia.n1[c];
    int32_t st2 = sp.pia.start2[c] - 1, n2 = variant == 0 ? sp.pia.n2[c] : 0;
    if (variant == 1) {
        int32_t s1 = sp.pia.start1[c], e1 = sp.pia.end1[c];
        n1 = e1 >= s1 ? e1 - s1 + 1 : 0;
    }
    const int32_t tot = n1 + (n2 > 0 ? n2 : 0);
    auto get = [&](int32_t i, double &w, double &v0, double &v1, double &v2) {
        const int64_t p = i < n1 ? (int64_t)st1 + i : (int64_t)st2 + (i - n1);
        w = sp.w[p]; v0 = sp.v0[p]; v1 = sp.v1[p]; v2 = sp.v2[p];
    };
    const MzPropsOut o = mz_props_wave(get, tot, masses[sidx]);
    if (lane == 0) {
        const int64_t ix = (int64_t)sidx * n_cells + c;
        out_np[ix] = (double)tot;
        out_n[ix] = (variant == 1 && ndens) ? o.n * inv_V_cells[c] : o.n;
        out_v[3 * ix] = o.v0;
        out_v[3 * ix + 1] = o.v1;
        out_v[3 * ix + 2] = o.v2;
        out_T[ix] = o.T;
        if (avg) {
            const size_t pn = (size_t)gridDim.y * n_cells;
            avg[ix] = fma((double)tot, avg_inv, avg[ix]);
            avg[pn + ix] = fma(out_n[ix], avg_inv, avg[pn + ix]);
            avg[2 * pn + 3 * ix] = fma(o.v0, avg_inv, avg[2 * pn + 3 * ix]);
            avg[2 * pn + 3 * ix + 1] = fma(o.v1, avg_inv, avg[2 * pn + 3 * ix + 1]);
            avg[2 * pn + 3 * ix + 2] = fma(o.v2, avg_inv, avg[2 * pn + 3 * ix + 2]);
            avg[5 * pn + ix] = fma(o.T, avg_inv, avg[5 * pn + ix]);
        }
    }
}

int mz_launch_props(mz_store *h, int variant, int ndens, double avg_inv) {
    dim3 grid((h->n_cells + 3) / 4, h->n_species);
    hipLaunchKernelGGL(k_props, grid, dim3(256), 0, h->stream, mz_views(h), h->n_cells, variant, ndens,
                       h->d_invV, h->d_masses, h->d_np, h->d_n, h->d_v, h->d_T,
                       avg_inv > 0.0 ? h->d_pavg : nullptr, avg_inv);
    MZ_HIP(hipGetLastError());
    return 0;
}

extern "C" int mz_compute_props(mz_store *h, int32_t variant, int32_t ndens_not_Np, const double *masses) {
    MZ_CHECK(variant == 0 || variant == 1, "props: bad variant");
    MZ_CHECK(masses != nullptr, "props: masses is NULL");
    MZ_HIP(hipMemcpyAsync(h->d_masses, masses, h->n_species * sizeof(double), hipMemcpyHostToDevice, h->stream));
    h->masses_valid = false;   // the couette drivers upload their mass again
    int r;
    {
        MzTimed tm(h, "props");
        r = mz_launch_props(h, variant, ndens_not_Np);
    }
    if (r) return r;
    MZ_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

// compute_flux_props! / compute_flux_props_sorted! (flux_props.jl:77-131,
// 199-249): per (cell, species), with C = v - vbar from the store's last
// props, sum w C |C|^2, w (Cx^2, Cy^2, Cz^2) and w (Cx Cy, Cx Cz, Cy Cz)
// over groups 1 and 2, scaled by m / V (and 1/2 for the energy flux).  One
// wave per (cell, species) of cells [lo, hi]; other cells are left as they
// are, as the reference's cell_chunk form leaves them.
__global__ void __launch_bounds__(256) k_flux(MzViews views, int32_t n_cells, int32_t lo, int32_t hi,
                                              const double *__restrict__ masses, const double *__restrict__ inv_V,
                                              const double *__restrict__ vbar, double *__restrict__ out) {
    const int sidx = blockIdx.y;
    const MzSpView &sp = views.v[sidx];
    const int c = lo + blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c > hi) return;
    const int32_t st1 = sp.pia.start1[c] - 1, n1 = sp.pia.n1[c];
    const int32_t st2 = sp.pia.start2[c] - 1, n2 = sp.pia.n2[c] > 0 ? sp.pia.n2[c] : 0;
    const int32_t tot = n1 + n2;
    const int64_t ix = (int64_t)sidx * n_cells + c;
    const double u0 = vbar[3 * ix], u1 = vbar[3 * ix + 1], u2 = vbar[3 * ix + 2];
    double q[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) q[j] = 0.0;
    for (int32_t i = lane; i < tot; i += 64) {
        const int64_t p = i < n1 ? (int64_t)st1 + i : (int64_t)st2 + (i - n1);
        const double w = sp.w[p];
        const double c0 = sp.v0[p] - u0, c1 = sp.v1[p] - u1, c2 = sp.v2[p] - u2;
        const double cx = c0 * c0, cy = c1 * c1, cz = c2 * c2;
        const double csq = cx + cy + cz;
        q[0] = q[0] + w * c0 * csq; q[1] = q[1] + w * c1 * csq; q[2] = q[2] + w * c2 * csq;
        q[3] = q[3] + w * cx; q[4] = q[4] + w * cy; q[5] = q[5] + w * cz;
        q[6] = q[6] + w * (c0 * c1); q[7] = q[7] + w * (c0 * c2); q[8] = q[8] + w * (c1 * c2);
    }
#pragma unroll
    for (int j = 0; j < 9; ++j) q[j] = mz_wave_sum(q[j]);
    if (lane == 0) {
        const double m = masses[sidx], iv = inv_V[c];
        double *o = out + 9 * ix;   // [species][cell][energy xyz | diagonal xyz | off-diagonal xy xz yz]
        for (int j = 0; j < 3; ++j) o[j] = 0.5 * q[j] * m * iv;
        for (int j = 3; j < 9; ++j) o[j] = q[j] * m * iv;
    }
}

// avg_props! of FluxProps (flux_props.jl:152-186): avg = avg + inst / n
__global__ void k_flux_avg(const double *__restrict__ inst, double *__restrict__ avg, int64_t n, double inv_n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) avg[i] = avg[i] + inst[i] * inv_n;
}

extern "C" int mz_compute_flux_props(mz_store *h, const double *masses, int32_t cell_lo, int32_t cell_hi) {
    MZ_CHECK(masses != nullptr, "flux_props: masses is NULL");
    MZ_CHECK(h->has_grid, "flux_props: no grid set (the cell volumes come from it)");
    MZ_CHECK(cell_lo >= 1 && cell_hi <= h->n_cells && cell_lo <= cell_hi, "flux_props: bad cell range");
    MZ_HIP(hipMemcpyAsync(h->d_masses, masses, h->n_species * sizeof(double), hipMemcpyHostToDevice, h->stream));
    h->masses_valid = false;   // the couette drivers upload their mass again
    const int ncell = cell_hi - cell_lo + 1;
    {
        MzTimed tm(h, "flux");
        hipLaunchKernelGGL(k_flux, dim3((ncell + 3) / 4, h->n_species), dim3(256), 0, h->stream, mz_views(h),
                           h->n_cells, cell_lo - 1, cell_hi - 1, h->d_masses, h->d_invV, h->d_v, h->d_flux);
    }
    MZ_HIP(hipGetLastError());
    MZ_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int mz_flux_avg(mz_store *h, double n_avg_timesteps) {
    MZ_CHECK(n_avg_timesteps > 0, "flux_avg: n_avg_timesteps must be positive");
    const int64_t n = (int64_t)9 * h->n_species * h->n_cells;
    hipLaunchKernelGGL(k_flux_avg, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->d_flux,
                       h->d_flux + n, n, 1.0 / n_avg_timesteps);
    MZ_HIP(hipGetLastError());
    return 0;
}

extern "C" int mz_flux_clear(mz_store *h, int32_t which) {
    MZ_CHECK(which == 0 || which == 1, "flux_clear: which must be 0 (instantaneous) or 1 (average)");
    const int64_t n = (int64_t)9 * h->n_species * h->n_cells;
    MZ_HIP(hipMemsetAsync(h->d_flux + which * n, 0, n * sizeof(double), h->stream));
    return 0;
}

extern "C" int mz_get_flux(mz_store *h, int32_t which, double *kinetic_energy_flux, double *diagonal_momentum_flux,
                           double *off_diagonal_momentum_flux) {
    MZ_CHECK(which == 0 || which == 1, "get_flux: which must be 0 (instantaneous) or 1 (average)");
    const int64_t n = (int64_t)9 * h->n_species * h->n_cells;
    std::vector<double> buf(n);
    MZ_HIP(hipStreamSynchronize(h->stream));
    MZ_HIP(hipMemcpy(buf.data(), h->d_flux + which * n, n * sizeof(double), hipMemcpyDeviceToHost));
    const int64_t m = (int64_t)h->n_species * h->n_cells;
    for (int64_t k = 0; k < m; ++k)
        for (int j = 0; j < 3; ++j) {
            if (kinetic_energy_flux) kinetic_energy_flux[3 * k + j] = buf[9 * k + j];
            if (diagonal_momentum_flux) diagonal_momentum_flux[3 * k + j] = buf[9 * k + 3 + j];
            if (off_diagonal_momentum_flux) off_diagonal_momentum_flux[3 * k + j] = buf[9 * k + 6 + j];
        }
    return 0;
}

// compute_props_with_total_moments! (physical_props.jl:168-245): n, v and T
// over groups 1 and 2 as compute_props!, plus the total moments
// sum w |v - vbar|^p / (scaling_p n) for the requested powers p.  One wave per
// (cell, species); the mean is reduced first, then the centred sums.
__global__ void __launch_bounds__(256) k_props_moments(MzViews views, int32_t n_cells, int32_t n_mom,
                                                       const int32_t *__restrict__ powers,
                                                       const double *__restrict__ scale,   // [species][n_mom]
                                                       const double *__restrict__ masses, double *__restrict__ out_np,
                                                       double *__restrict__ out_n, double *__restrict__ out_v,
                                                       double *__restrict__ out_T, double *__restrict__ out_m) {
    const int sidx = blockIdx.y;
    const MzSpView &sp = views.v[sidx];
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c >= n_cells) return;
    const int32_t st1 = sp.pia.start1[c] - 1, n1 = sp.pia.n1[c];
    const int32_t st2 = sp.pia.start2[c] - 1, n2 = sp.pia.n2[c] > 0 ? sp.pia.n2[c] : 0;
    const int32_t tot = n1 + n2;
    double n = 0.0, a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int32_t i = lane; i < tot; i += 64) {
        const int64_t p = i < n1 ? (int64_t)st1 + i : (int64_t)st2 + (i - n1);
        const double w = sp.w[p];
        n += w;
        a0 = a0 + sp.v0[p] * w;
        a1 = a1 + sp.v1[p] * w;
        a2 = a2 + sp.v2[p] * w;
    }
    n = mz_wave_sum(n);
    a0 = mz_wave_sum(a0);
    a1 = mz_wave_sum(a1);
    a2 = mz_wave_sum(a2);
    double m0 = 0.0, m1 = 0.0, m2 = 0.0, E = 0.0;
    double mom[MZ_MAX_MOMENTS];
#pragma unroll
    for (int j = 0; j < MZ_MAX_MOMENTS; ++j) mom[j] = 0.0;
    if (n > 0.0) {
        m0 = a0 / n; m1 = a1 / n; m2 = a2 / n;
        for (int32_t i = lane; i < tot; i += 64) {
            const int64_t p = i < n1 ? (int64_t)st1 + i : (int64_t)st2 + (i - n1);
            const double w = sp.w[p];
            const double d0 = sp.v0[p] - m0, d1 = sp.v1[p] - m1, d2 = sp.v2[p] - m2;
            const double normv = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            E = E + w * (normv * normv);
#pragma unroll
            for (int j = 0; j < MZ_MAX_MOMENTS; ++j)
                if (j < n_mom) mom[j] += w * pow(normv, (double)powers[j]);
        }
        E = mz_wave_sum(E);
#pragma unroll
        for (int j = 0; j < MZ_MAX_MOMENTS; ++j)
            if (j < n_mom) mom[j] = mz_wave_sum(mom[j]);
    }
    if (lane == 0) {
        double T = 0.0;
        if (n > 0.0) T = (2.0 / 3.0) * (E * (0.5 * masses[sidx] / (n * MZ_KB)));
        const int64_t ix = (int64_t)sidx * n_cells + c;
        out_np[ix] = (double)tot;
        out_n[ix] = n;
        out_v[3 * ix] = m0;
        out_v[3 * ix + 1] = m1;
        out_v[3 * ix + 2] = m2;
        out_T[ix] = T;
        for (int j = 0; j < n_mom; ++j) out_m[ix * n_mom + j] = mom[j] / (scale[(int64_t)sidx * n_mom + j] * n);
    }
}

extern "C" int mz_compute_props_moments(mz_store *h, const int32_t *powers, int32_t n_moments, double Tref,
                                        const double *masses) {
    MZ_CHECK(n_moments >= 1 && n_moments <= MZ_MAX_MOMENTS, "props moments: 1..8 moments");
    MZ_CHECK(powers != nullptr && masses != nullptr, "props moments: NULL argument");
    MZ_CHECK(Tref > 0.0, "props moments: Tref must be positive");
    const int ns = h->n_species;
    // moment_factor * moment_vref^-(3 + p) * gamma((3 + p) / 2) per species and power
    std::vector<double> sc((size_t)ns * n_moments);
    for (int s = 0; s < ns; ++s) {
        const double mf = 4 * M_PI * std::pow(masses[s] / (2 * M_PI * MZ_KB * Tref), 1.5) * 0.5;
        const double mvref = std::pow(masses[s] / (2 * MZ_KB * Tref), 0.5);
        for (int j = 0; j < n_moments; ++j)
            sc[(size_t)s * n_moments + j] =
                mf * std::pow(mvref, -(double)(3 + powers[j])) * std::tgamma((3.0 + powers[j]) / 2.0);
    }
    int32_t *d_pow = (int32_t *)mz_scratch(h, "props_pow", MZ_MAX_MOMENTS * sizeof(int32_t));
    double *d_sc = (double *)mz_scratch(h, "props_scale", (size_t)ns * MZ_MAX_MOMENTS * sizeof(double));
    double *d_m = (double *)mz_scratch(h, "props_moments", (size_t)ns * h->n_cells * MZ_MAX_MOMENTS * sizeof(double));
    MZ_CHECK(d_pow && d_sc && d_m, "props moments: scratch allocation failed");
    MZ_HIP(hipMemcpyAsync(d_pow, powers, n_moments * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    MZ_HIP(hipMemcpyAsync(d_sc, sc.data(), sc.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    MZ_HIP(hipMemcpyAsync(h->d_masses, masses, ns * sizeof(double), hipMemcpyHostToDevice, h->stream));
    h->masses_valid = false;   // the couette drivers upload their mass again
    {
        MzTimed tm(h, "props");
        dim3 grid((h->n_cells + 3) / 4, ns);
        hipLaunchKernelGGL(k_props_moments, grid, dim3(256), 0, h->stream, mz_views(h), h->n_cells, n_moments,
                           d_pow, d_sc, h->d_masses, h->d_np, h->d_n, h->d_v, h->d_T, d_m);
        MZ_HIP(hipGetLastError());
    }
    MZ_HIP(hipStreamSynchronize(h->stream));
    h->n_moments = n_moments;
    return 0;
}

extern "C" int mz_get_props_moments(mz_store *h, double *moments) {
    MZ_CHECK(h->n_moments > 0, "props moments: none computed yet");
    MZ_CHECK(moments != nullptr, "props moments: NULL output");
    MZ_HIP(hipStreamSynchronize(h->stream));
    double *d_m = (double *)mz_scratch(h, "props_moments", 0);
    MZ_HIP(hipMemcpy(moments, d_m, (size_t)h->n_species * h->n_cells * h->n_moments * sizeof(double),
                     hipMemcpyDeviceToHost));
    return 0;
}

// compute_mixed_moment (physical_props.jl:473-501) over a range of cells:
// sum over groups 1 and 2 of w * sum_scaler * vx^px * vy^py * vz^pz, times
// res_scaler.  One wave per cell; the sum is a wave reduction, not the
// reference's left-to-right loop, so the last bits can differ.
__global__ void __launch_bounds__(256) k_mixed_moment(MzViews views, int32_t sidx, int32_t c0, int32_t n_out,
                                                      int32_t px, int32_t py, int32_t pz, double sum_scaler,
                                                      double res_scaler, double *__restrict__ out) {
    const MzSpView &sp = views.v[sidx];
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= n_out) return;
    const int c = c0 + k;
    const int32_t st1 = sp.pia.start1[c] - 1, n1 = sp.pia.n1[c] > 0 ? sp.pia.n1[c] : 0;
    const int32_t st2 = sp.pia.start2[c] - 1, n2 = sp.pia.n2[c] > 0 ? sp.pia.n2[c] : 0;
    const int32_t tot = n1 + n2;
    const double ex = (double)px, ey = (double)py, ez = (double)pz;
    double r = 0.0;
    for (int32_t i = lane; i < tot; i += 64) {
        const int64_t p = i < n1 ? (int64_t)st1 + i : (int64_t)st2 + (i - n1);
        r += sp.w[p] * sum_scaler * pow(sp.v0[p], ex) * pow(sp.v1[p], ey) * pow(sp.v2[p], ez);
    }
    r = mz_wave_sum(r);
    if (lane == 0) out[k] = r * res_scaler;
}

extern "C" int mz_compute_mixed_moment(mz_store *h, int32_t species, int32_t cell_lo, int32_t cell_hi, int32_t px,
                                       int32_t py, int32_t pz, double sum_scaler, double res_scaler,
                                       double *out) {
    MZ_CHECK(species >= 1 && species <= h->n_species, "mixed moment: species out of range");
    MZ_CHECK(cell_lo >= 1 && cell_hi <= h->n_cells && cell_lo <= cell_hi, "mixed moment: cells out of range");
    MZ_CHECK(out != nullptr, "mixed moment: NULL output");
    const int32_t n_out = cell_hi - cell_lo + 1;
    double *d_out = (double *)mz_scratch(h, "mixed_moment", (size_t)h->n_cells * sizeof(double));
    MZ_CHECK(d_out, "mixed moment: scratch allocation failed");
    {
        MzTimed tm(h, "mixed_moment");
        hipLaunchKernelGGL(k_mixed_moment, dim3((n_out + 3) / 4), dim3(256), 0, h->stream, mz_views(h),
                           species - 1, cell_lo - 1, n_out, px, py, pz, sum_scaler, res_scaler, d_out);
        MZ_HIP(hipGetLastError());
    }
    MZ_HIP(hipMemcpyAsync(out, d_out, (size_t)n_out * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    MZ_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int mz_get_props(mz_store *h, double *lpa, double *np, double *n, double *v, double *T) {
    MZ_HIP(hipStreamSynchronize(h->stream));
    size_t pn = (size_t)h->n_species * h->n_cells;
    if (np) MZ_HIP(hipMemcpy(np, h->d_np, pn * sizeof(double), hipMemcpyDeviceToHost));
    if (n) MZ_HIP(hipMemcpy(n, h->d_n, pn * sizeof(double), hipMemcpyDeviceToHost));
    if (v) MZ_HIP(hipMemcpy(v, h->d_v, pn * 3 * sizeof(double), hipMemcpyDeviceToHost));
    if (T) MZ_HIP(hipMemcpy(T, h->d_T, pn * sizeof(double), hipMemcpyDeviceToHost));
    if (lpa) {
        for (int s = 1; s <= h->n_species; ++s) {
            MzCounters c;
            int r = mz_read_counters(h, s, &c);
            if (r) return r;
            lpa[s - 1] = (double)c.length;
        }
    }
    return 0;
}
