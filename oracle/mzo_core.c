/*
 * mzo_core.c -- oracle: random draws, ParticleVector / ParticleIndexerArray
 * bookkeeping, grid sorting and 1-D convection.
 *
 * TEST INFRASTRUCTURE ONLY (see mzo.h).  Every function cites the reference
 * file:line it restates.
 */
#include "mzo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Round: multiply words
 * 0 and 2 by the two Philox constants, swap/xor with the key, bump the key by
 * the Weyl constants between rounds.  Pinned against rocRAND's independent
 * implementation and the published known-answer vectors in tests/. */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void mzo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Draw k of stream s: block k/2 of Philox with counter (k/2, s, tag_lo,
 * tag_hi); half k%2 of the 128-bit output gives 64 bits, whose top 53 bits
 * make a double in [0,1) like Julia's rand(rng, Float64). */
double mzo_rand(mzo_rng *r) {
    uint64_t k = r->k++;
    if (r->mode == 1) {
        if ((int64_t)r->stream >= r->rows || (int64_t)k >= r->cols) {
            r->err = 1;
            return 0.5;
        }
        return r->table[(int64_t)r->stream * r->cols + (int64_t)k];
    }
    uint32_t ctr[4] = {(uint32_t)(k >> 1), (uint32_t)r->stream, (uint32_t)r->tag,
                       (uint32_t)(r->tag >> 32)};
    uint32_t key[2] = {(uint32_t)r->seed, (uint32_t)(r->seed >> 32)};
    uint32_t out[4];
    mzo_philox4x32_10(ctr, key, out);
    int h = (int)(k & 1u) * 2;
    uint64_t bits = ((uint64_t)out[h + 1] << 32) | (uint64_t)out[h];
    return (double)(bits >> 11) * 0x1.0p-53;
}

void mzo_rng_set_stream(mzo_rng *r, uint64_t stream) {
    r->stream = stream;
    r->k = 0;
}

/* ------------------------------------------------------------------------- */
/* ParticleVector (src/particles.jl:194-298) */

#define PV_W(pv, id) ((pv)->w[(id) - 1])
#define PV_V(pv, id) (&(pv)->v[3 * ((id) - 1)])
#define PV_X(pv, id) (&(pv)->x[3 * ((id) - 1)])
#define PID(pv, pos) ((pv)->index[(pos) - 1])

static void pv_reserve(mzo_pv *pv, int64_t n) {
    if (n <= pv->cap) return;
    int64_t nc = pv->cap > 0 ? pv->cap : 16;
    while (nc < n) nc *= 2;
    pv->w = (double *)realloc(pv->w, sizeof(double) * nc);
    pv->v = (double *)realloc(pv->v, sizeof(double) * 3 * nc);
    pv->x = (double *)realloc(pv->x, sizeof(double) * 3 * nc);
    pv->index = (int64_t *)realloc(pv->index, sizeof(int64_t) * nc);
    pv->cell = (int64_t *)realloc(pv->cell, sizeof(int64_t) * nc);
    pv->buffer = (int64_t *)realloc(pv->buffer, sizeof(int64_t) * nc);
    pv->cap = nc;
}

/* ParticleVector(np): particles zeroed, index 1:np, cell 0, buffer np:-1:1
 * (src/particles.jl:209-212) */
mzo_pv *mzo_pv_new(int64_t np) {
    mzo_pv *pv = (mzo_pv *)calloc(1, sizeof(mzo_pv));
    pv_reserve(pv, np > 0 ? np : 1);
    pv->length = np;
    for (int64_t i = 0; i < np; ++i) {
        pv->w[i] = 0.0;
        pv->v[3 * i] = pv->v[3 * i + 1] = pv->v[3 * i + 2] = 0.0;
        pv->x[3 * i] = pv->x[3 * i + 1] = pv->x[3 * i + 2] = 0.0;
        pv->index[i] = i + 1;
        pv->cell[i] = 0;
        pv->buffer[i] = np - i;
    }
    pv->nbuffer = np;
    return pv;
}

void mzo_pv_free(mzo_pv *pv) {
    if (!pv) return;
    free(pv->w); free(pv->v); free(pv->x);
    free(pv->index); free(pv->cell); free(pv->buffer);
    free(pv);
}

/* Base.resize!(pv, n) (src/particles.jl:269-298): new slots hold zero
 * particles, index[i] = i for them, old buffer content moves to the tail and
 * the new ids are written to the head in descending order so the lowest new
 * id is popped first. */
void mzo_pv_resize(mzo_pv *pv, int64_t n) {
    int64_t old_len = pv->length;
    pv_reserve(pv, n);
    int64_t n_diff = n - old_len;
    for (int64_t i = old_len; i < n; ++i) {
        pv->w[i] = 0.0;
        pv->v[3 * i] = pv->v[3 * i + 1] = pv->v[3 * i + 2] = 0.0;
        pv->x[3 * i] = pv->x[3 * i + 1] = pv->x[3 * i + 2] = 0.0;
        pv->index[i] = i + 1;
        pv->cell[i] = 0;
    }
    /* buffer[n_diff+1:n] .= buffer[1:old_len] (copy first: ranges overlap) */
    if (old_len > 0) {
        int64_t *tmp = (int64_t *)malloc(sizeof(int64_t) * old_len);
        memcpy(tmp, pv->buffer, sizeof(int64_t) * old_len);
        memcpy(pv->buffer + n_diff, tmp, sizeof(int64_t) * old_len);
        free(tmp);
    }
    for (int64_t i = 1; i <= n_diff; ++i) pv->buffer[i - 1] = n + 1 - i;
    pv->nbuffer += n_diff;
    pv->length = n;
}

/* update_particle_buffer_new_particle!(pv, position) (src/particles.jl:311-315) */
void mzo_pv_buffer_pop_to(mzo_pv *pv, int64_t position) {
    pv->index[position - 1] = pv->buffer[pv->nbuffer - 1];
    pv->nbuffer -= 1;
}

/* ------------------------------------------------------------------------- */
/* ParticleIndexerArray (src/particles.jl:104-174) */

mzo_pia *mzo_pia_new(int64_t n_cells, int64_t n_species) {
    mzo_pia *p = (mzo_pia *)calloc(1, sizeof(mzo_pia));
    p->n_cells = n_cells;
    p->n_species = n_species;
    p->ind = (mzo_ind *)calloc((size_t)(n_cells * n_species), sizeof(mzo_ind));
    for (int64_t i = 0; i < n_cells * n_species; ++i) {
        /* ParticleIndexer() = (0, 0, -1, 0, 0, -1, 0)  (src/particles.jl:84) */
        p->ind[i].n_local = 0; p->ind[i].start1 = 0; p->ind[i].end1 = -1; p->ind[i].n_group1 = 0;
        p->ind[i].start2 = 0; p->ind[i].end2 = -1; p->ind[i].n_group2 = 0;
    }
    p->n_total = (int64_t *)calloc((size_t)n_species, sizeof(int64_t));
    p->contiguous = (int8_t *)calloc((size_t)n_species, 1);
    for (int64_t s = 0; s < n_species; ++s) p->contiguous[s] = 1;
    return p;
}

void mzo_pia_free(mzo_pia *p) {
    if (!p) return;
    free(p->ind); free(p->n_total); free(p->contiguous); free(p);
}

mzo_ind *mzo_pia_at(mzo_pia *pia, int64_t cell, int64_t species) {
    return &pia->ind[(species - 1) * pia->n_cells + (cell - 1)];
}

/* map_cont_index (src/particles.jl:364-366) */
int64_t mzo_map_cont_index(const mzo_ind *ind, int64_t i) {
    return i < ind->n_group1 ? i + ind->start1 : (i - ind->n_group1) + ind->start2;
}

/* update_particle_indexer_new_particle! (src/particles.jl:426-433) */
void mzo_update_indexer_new_particle(mzo_pia *pia, int64_t cell, int64_t species) {
    mzo_ind *ind = mzo_pia_at(pia, cell, species);
    pia->n_total[species - 1] += 1;
    ind->n_local += 1;
    ind->n_group2 += 1;
    ind->start2 = ind->start2 > 0 ? ind->start2 : pia->n_total[species - 1];
    ind->end2 = pia->n_total[species - 1];
}

/* update_particle_indexer_new_lower_count! (src/particles.jl:395-413) */
void mzo_update_indexer_new_lower_count(mzo_pia *pia, int64_t cell, int64_t species, int64_t new_count) {
    mzo_ind *ind = mzo_pia_at(pia, cell, species);
    int64_t diff = ind->n_local - new_count;
    ind->n_local = new_count;
    pia->n_total[species - 1] -= diff;
    if (new_count > ind->n_group1) {
        ind->end2 -= diff;
        ind->n_group2 -= diff;
    } else {
        diff -= ind->n_group2;
        ind->start2 = 0; ind->end2 = -1; ind->n_group2 = 0;
        ind->end1 -= diff;
        ind->n_group1 -= diff;
    }
}

/* update_buffer_index_new_particle! (src/particles.jl:718-721) */
void mzo_update_buffer_index_new_particle(mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species) {
    mzo_update_indexer_new_particle(pia, cell, species);
    mzo_pv_buffer_pop_to(pv, pia->n_total[species - 1]);
}

/* delete_particle_end_group1! / _group2! / delete_particle_end!
 * (src/particles.jl:478-566) */
static void delete_end_group(mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species, int grp) {
    mzo_ind *ind = mzo_pia_at(pia, cell, species);
    int64_t *n_g = grp == 1 ? &ind->n_group1 : &ind->n_group2;
    int64_t *s_g = grp == 1 ? &ind->start1 : &ind->start2;
    int64_t *e_g = grp == 1 ? &ind->end1 : &ind->end2;
    if (*n_g == 0) return;
    int64_t pos = *e_g;
    PV_W(pv, PID(pv, pos)) = 0.0;
    ind->n_local -= 1;
    *e_g -= 1;
    *n_g -= 1;
    pia->n_total[species - 1] -= 1;
    if (*e_g < *s_g) { *s_g = 0; *e_g = -1; }
    pv->nbuffer += 1;
    pv->buffer[pv->nbuffer - 1] = PID(pv, pos);
}

/* delete_particle!(pv, pia, cell, species, i) (src/particles.jl:448-461): the
 * particle at position i trades places with the last one of its group, which
 * is then deleted. */
void mzo_delete_particle(mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species, int64_t i) {
    mzo_ind *ind = mzo_pia_at(pia, cell, species);
    if (ind->n_group2 > 0 && i >= ind->start2 && i <= ind->end2) {
        int64_t last = pv->index[ind->end2 - 1];
        pv->index[ind->end2 - 1] = pv->index[i - 1];
        pv->index[i - 1] = last;
        delete_end_group(pv, pia, cell, species, 2);
    } else {
        int64_t last = pv->index[ind->end1 - 1];
        pv->index[ind->end1 - 1] = pv->index[i - 1];
        pv->index[i - 1] = last;
        delete_end_group(pv, pia, cell, species, 1);
    }
}

void mzo_delete_particle_end(mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species) {
    mzo_ind *ind = mzo_pia_at(pia, cell, species);
    if (ind->n_group2 > 0) delete_end_group(pv, pia, cell, species, 2);
    else delete_end_group(pv, pia, cell, species, 1);
}

/* squash_pia! (src/particles.jl:622-701) */
void mzo_squash_pia(mzo_pv *pv, mzo_pia *pia, int64_t species) {
    if (pia->contiguous[species - 1]) return;
    int64_t nc = pia->n_cells;
    if (nc == 1) {
        mzo_ind *ind = mzo_pia_at(pia, 1, species);
        if (ind->n_group2 > 0) {
            int64_t e1 = ind->end1 > 0 ? ind->end1 : 0;
            int64_t offset = ind->start2 - (e1 + 1);
            if (offset > 0) {
                ind->start2 -= offset;
                ind->end2 -= offset;
                for (int64_t j = ind->start2; j <= ind->end2; ++j) {
                    pv->index[j - 1] = pv->index[j + offset - 1];
                    pv->cell[j - 1] = pv->cell[j + offset - 1];
                }
            }
        }
    } else {
        mzo_ind *i1 = mzo_pia_at(pia, 1, species);
        int64_t last_end = i1->end1 > 0 ? i1->end1 : 0;
        for (int64_t i = 1; i <= nc - 1; ++i) {
            mzo_ind *in = mzo_pia_at(pia, i + 1, species);
            int64_t offset = in->start1 - (last_end + 1);
            if (offset > 0) {
                in->start1 -= offset;
                in->end1 -= offset;
                for (int64_t j = in->start1; j <= in->end1; ++j) {
                    pv->index[j - 1] = pv->index[j + offset - 1];
                    pv->cell[j - 1] = pv->cell[j + offset - 1];
                }
            }
            last_end = in->end1 > 0 ? in->end1 : last_end;
        }
        for (int64_t i = 1; i <= nc; ++i) {
            mzo_ind *in = mzo_pia_at(pia, i, species);
            if (in->n_group2 > 0) {
                int64_t offset = in->start2 - (last_end + 1);
                if (offset > 0) {
                    in->start2 -= offset;
                    in->end2 -= offset;
                    for (int64_t j = in->start2; j <= in->end2; ++j) {
                        pv->index[j - 1] = pv->index[j + offset - 1];
                        pv->cell[j - 1] = pv->cell[j + offset - 1];
                    }
                }
                last_end = in->end2 > 0 ? in->end2 : last_end;
            }
        }
    }
    pia->contiguous[species - 1] = 1;
}

/* check_pia_is_correct (src/particles.jl:863-907); returns 1 if correct. */
int mzo_check_pia_is_correct(mzo_pia *pia, int64_t species, int64_t *where) {
    int64_t n_tot = 0;
    *where = 0;
    for (int64_t i = 1; i <= pia->n_cells; ++i) {
        mzo_ind *d = mzo_pia_at(pia, i, species);
        if (d->n_local != d->n_group1 + d->n_group2) { *where = i; return 0; }
        n_tot += d->n_local;
        if (d->n_group1 > 0) {
            if (d->n_group1 != d->end1 - d->start1 + 1) { *where = i; return 0; }
        } else if (d->start1 != 0 || d->end1 != -1) { *where = i; return 0; }
        if (d->n_group2 > 0) {
            if (d->n_group2 != d->end2 - d->start2 + 1) { *where = i; return 0; }
        } else if (d->start2 != 0 || d->end2 != -1) { *where = i; return 0; }
    }
    if (n_tot != pia->n_total[species - 1]) { *where = 0; return 0; }
    return 1;
}

/* check_unique_index (src/particles.jl:942-986) */
int mzo_check_unique_index(mzo_pv *pv, mzo_pia *pia, int64_t species, int64_t *code) {
    int64_t *cnt = (int64_t *)calloc((size_t)pv->length + 1, sizeof(int64_t));
    int64_t retcode = 1;
    for (int64_t c = 1; c <= pia->n_cells; ++c) {
        mzo_ind *d = mzo_pia_at(pia, c, species);
        for (int64_t i = d->start1; i <= d->end1; ++i) cnt[PID(pv, i)] += 1;
        for (int64_t i = d->start2; i <= d->end2; ++i) cnt[PID(pv, i)] += 1;
    }
    int64_t n_index = 0;
    for (int64_t i = 1; i <= pv->length; ++i) if (cnt[i] > n_index) n_index = cnt[i];
    for (int64_t i = 0; i < pv->nbuffer; ++i)
        if (cnt[pv->buffer[i]] > 0) { retcode = -1; break; }
    free(cnt);
    if (n_index > 1) { *code = n_index * retcode; return 0; }
    if (retcode == -1) { *code = -1; return 0; }
    *code = 0;
    return 1;
}

/* check_unique_buffer (src/particles.jl:1009-1025) */
int mzo_check_unique_buffer(mzo_pv *pv, int64_t *where) {
    char *seen = (char *)calloc((size_t)pv->length + 2, 1);
    *where = 0;
    for (int64_t i = 1; i <= pv->nbuffer; ++i) {
        int64_t pid = pv->buffer[i - 1];
        if (pid < 0 || pid > pv->length) { *where = i; free(seen); return 0; }
        if (seen[pid]) { *where = i; free(seen); return 0; }
        seen[pid] = 1;
    }
    free(seen);
    return 1;
}

/* count_disordered_particles (src/particles.jl:805-837) */
int64_t mzo_count_disordered_particles(mzo_pv *pv, mzo_pia *pia, int64_t species, int use_offset) {
    int64_t count = 0;
    if (use_offset) {
        for (int64_t c = 1; c <= pia->n_cells; ++c) {
            mzo_ind *d = mzo_pia_at(pia, c, species);
            if (d->n_group1 > 0) {
                int64_t offset = PID(pv, d->start1) - d->start1;
                for (int64_t i = d->start1 + 1; i <= d->end1; ++i) {
                    if (PID(pv, i) != i + offset) {
                        count += 1;
                        offset = PID(pv, i) - i;
                    }
                }
            }
        }
    } else {
        for (int64_t i = 1; i <= pia->n_total[species - 1]; ++i)
            if (PID(pv, i) != i) count += 1;
    }
    return count;
}

/* restore_particle_ordering! (src/particles.jl:1086-1137) */
void mzo_restore_particle_ordering(mzo_pv *pv) {
    int64_t n_total = pv->length;
    int64_t n_used = n_total - pv->nbuffer;
    int64_t *inv = (int64_t *)calloc((size_t)n_total + 1, sizeof(int64_t));
    for (int64_t i = 1; i <= n_used; ++i) inv[PID(pv, i)] = i;
    for (int64_t i = 1; i <= n_used; ++i) {
        int64_t cur = PID(pv, i);
        if (cur != i) {
            int64_t displaced = inv[i];
            /* swap particle data of physical slots i and cur */
            double tw = PV_W(pv, i); PV_W(pv, i) = PV_W(pv, cur); PV_W(pv, cur) = tw;
            for (int d = 0; d < 3; ++d) {
                double t = PV_V(pv, i)[d]; PV_V(pv, i)[d] = PV_V(pv, cur)[d]; PV_V(pv, cur)[d] = t;
                t = PV_X(pv, i)[d]; PV_X(pv, i)[d] = PV_X(pv, cur)[d]; PV_X(pv, cur)[d] = t;
            }
            pv->index[i - 1] = i;
            inv[i] = i;
            if (displaced > 0) {
                pv->index[displaced - 1] = cur;
                inv[cur] = displaced;
            } else {
                inv[cur] = 0;
            }
        }
    }
    for (int64_t i = n_used + 1; i <= n_total; ++i) pv->index[i - 1] = i;
    for (int64_t i = 1; i <= pv->nbuffer; ++i) pv->buffer[i - 1] = n_total - i + 1;
    free(inv);
}

/* ------------------------------------------------------------------------- */
/* Grid1DUniform (src/grids/grid_uniform1D.jl:49-99) */
void mzo_grid1d_init(mzo_grid1d *g, double L, int64_t nx, double wall_offset) {
    double dx = L / (double)nx;
    g->L = L;
    g->n_cells = nx;
    g->dx = dx;
    g->inv_dx = 1.0 / dx;
    g->min_x = dx * wall_offset;
    g->max_x = L - dx * wall_offset;
}

int64_t mzo_get_cell(const mzo_grid1d *g, double x0) {
    return (int64_t)floor(x0 * g->inv_dx) + 1;
}

/* Counting sort of the index array (src/grids/grid_sorting.jl:58-113 and
 * 128-182): histogram by cell, inclusive scan writing group 1 of every cell
 * and clearing group 2, reverse-order scatter (stable), contiguous = true.
 * pv->cell is written per logical position before the scatter and is not
 * permuted, as in the reference. */
static void sort_common(int64_t n_cells, mzo_pv *pv, mzo_pia *pia, int64_t species, int64_t n_tot,
                        int64_t *cnt) {
    for (int64_t cell = 1; cell <= n_cells; ++cell) {
        cnt[cell] = cnt[cell] + cnt[cell - 1];
        int64_t cs = cnt[cell - 1] + 1, np = cnt[cell] - cnt[cell - 1], ce = cnt[cell];
        mzo_ind *d = mzo_pia_at(pia, cell, species);
        d->start2 = 0; d->end2 = -1; d->n_group2 = 0;
        if (np > 0) { d->start1 = cs; d->end1 = ce; }
        else { d->start1 = 0; d->end1 = -1; }
        d->n_group1 = np;
        d->n_local = np;
    }
    int64_t *sorted = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_tot > 0 ? n_tot : 1));
    for (int64_t i = n_tot; i >= 1; --i) {
        int64_t c = pv->cell[i - 1];
        sorted[cnt[c] - 1] = pv->index[i - 1];
        cnt[c] -= 1;
    }
    for (int64_t i = 1; i <= n_tot; ++i) pv->index[i - 1] = sorted[i - 1];
    free(sorted);
    pia->contiguous[species - 1] = 1;
}

void mzo_sort_particles(const mzo_grid1d *g, mzo_pv *pv, mzo_pia *pia, int64_t species) {
    int64_t n_tot = pia->n_total[species - 1];
    int64_t *cnt = (int64_t *)calloc((size_t)g->n_cells + 1, sizeof(int64_t));
    if (!pia->contiguous[species - 1]) mzo_squash_pia(pv, pia, species);
    for (int64_t i = 1; i <= n_tot; ++i) {
        int64_t c = mzo_get_cell(g, PV_X(pv, PID(pv, i))[0]);
        pv->cell[i - 1] = c;
        cnt[c] += 1;
    }
    sort_common(g->n_cells, pv, pia, species, n_tot, cnt);
    free(cnt);
}

void mzo_sort_particles_precomputed(mzo_pv *pv, mzo_pia *pia, int64_t species) {
    int64_t n_tot = pia->n_total[species - 1];
    int64_t *cnt = (int64_t *)calloc((size_t)pia->n_cells + 1, sizeof(int64_t));
    if (!pia->contiguous[species - 1]) mzo_squash_pia(pv, pia, species);
    for (int64_t i = 1; i <= n_tot; ++i) cnt[pv->cell[i - 1]] += 1;
    sort_common(pia->n_cells, pv, pia, species, n_tot, cnt);
    free(cnt);
}

/* ------------------------------------------------------------------------- */
/* MaxwellWalls1D (src/convection/boundary_conditions.jl:49-53) */
void mzo_walls_init(mzo_walls *wl, const double *masses, int64_t n_species, double T_l, double T_r,
                    double vy_l, double vy_r, double acc_l, double acc_r) {
    memset(wl, 0, sizeof(*wl));
    wl->T[0] = T_l; wl->T[1] = T_r;
    wl->v[0][1] = vy_l; wl->v[1][1] = vy_r;
    wl->acc[0] = acc_l; wl->acc[1] = acc_r;
    wl->n_species = n_species;
    for (int64_t s = 0; s < n_species; ++s) {
        wl->vsq[0 * n_species + s] = 2 * MZO_KB * T_l / masses[s];
        wl->vsq[1 * n_species + s] = 2 * MZO_KB * T_r / masses[s];
    }
}

/* diffuse_reflection_x! (boundary_conditions.jl:79-93): 3 draws */
static void diffuse_reflection_x(mzo_rng *rng, double *v, double vsq, double sgn, const double *wv) {
    double R = fmax(1e-50, mzo_rand(rng));
    double v_normal = sgn * sqrt(-vsq * log(R));
    R = fmax(1e-50, mzo_rand(rng));
    double v_tang = sqrt(-vsq * log(R));
    R = 6.283185307179586 * mzo_rand(rng);
    double v_t1 = sin(R) * v_tang;
    double v_t2 = cos(R) * v_tang;
    v[0] = v_normal + wv[0];
    v[1] = v_t1 + wv[1];
    v[2] = v_t2 + wv[2];
}

/* reflect_particle_x! (boundary_conditions.jl:108-121) */
static void reflect_particle_x(mzo_rng *rng, double *v, double vsq, double sgn, const double *wv,
                               double acc) {
    if (acc == 0.0) {
        v[0] = -v[0];
    } else if (acc == 1.0) {
        diffuse_reflection_x(rng, v, vsq, sgn, wv);
    } else {
        double R = mzo_rand(rng);
        if (R < acc) diffuse_reflection_x(rng, v, vsq, sgn, wv);
        else v[0] = -v[0];
    }
}

/* convect_single_particle! (src/convection/convection_1D.jl:17-54).  The
 * @muladd block makes x + v*t a fused multiply-add; it is written as fma()
 * so the cell index downstream is reproducible bit for bit. */
static void convect_single(mzo_rng *rng, const mzo_grid1d *g, const mzo_walls *wl, double *x,
                           double *v, int64_t species, double dt) {
    double t_rest = dt;
    double x_old = x[0];
    double x_new = fma(v[0], dt, x[0]);
    while ((x_new >= g->L) || (x_new <= 0.0)) {
        int bc;
        double wn;
        if (x_new >= g->L) {
            t_rest -= fabs((g->L - x_old) / v[0]);
            bc = 1; wn = -1.0; x_old = g->L;
        } else {
            t_rest -= fabs(x_old / v[0]);
            bc = 0; wn = 1.0; x_old = 0.0;
        }
        reflect_particle_x(rng, v, wl->vsq[bc * wl->n_species + species - 1], wn, wl->v[bc],
                           wl->acc[bc]);
        x_new = fma(v[0], t_rest, x_old);
    }
    /* clamp(x_new, min_x, max_x) */
    if (x_new < g->min_x) x_new = g->min_x;
    else if (x_new > g->max_x) x_new = g->max_x;
    x[0] = x_new;
}

/* convect_particles! (convection_1D.jl:130-157).  Each particle draws from
 * its own stream (physical id - 1), restarted for every call. */
void mzo_convect_particles(mzo_rng *rng, const mzo_grid1d *g, const mzo_walls *wl, mzo_pv *pv,
                           mzo_pia *pia, int64_t species, double dt) {
    if (pia->contiguous[species - 1]) {
        int64_t n_tot = pia->n_total[species - 1];
        for (int64_t i = 1; i <= n_tot; ++i) {
            int64_t id = PID(pv, i);
            mzo_rng_set_stream(rng, (uint64_t)(id - 1));
            convect_single(rng, g, wl, PV_X(pv, id), PV_V(pv, id), species, dt);
        }
    } else {
        for (int64_t c = 1; c <= g->n_cells; ++c) {
            mzo_ind *d = mzo_pia_at(pia, c, species);
            for (int64_t i = d->start1; i <= d->end1; ++i) {
                int64_t id = PID(pv, i);
                mzo_rng_set_stream(rng, (uint64_t)(id - 1));
                convect_single(rng, g, wl, PV_X(pv, id), PV_V(pv, id), species, dt);
            }
            if (d->n_group2 > 0) {
                for (int64_t i = d->start2; i <= d->end2; ++i) {
                    int64_t id = PID(pv, i);
                    mzo_rng_set_stream(rng, (uint64_t)(id - 1));
                    convect_single(rng, g, wl, PV_X(pv, id), PV_V(pv, id), species, dt);
                }
            }
        }
    }
}

/* ---------------- surface properties ----------------
 * update_surface_incident! / update_surface_reflected!
 * (surface_props.jl:77-97, 111-128).  The reference's @muladd block turns the
 * normal projection px n1 + py n2 + pz n3 into muladd(px, n1, muladd(py, n2,
 * pz n3)), px - p_dot_n n1 into muladd(-p_dot_n, n1, px) and the energy term
 * likewise; 0.5 times the energy term is exact, so fusing its accumulation
 * changes nothing. */
enum { SQ_NP, SQ_FI, SQ_FR, SQ_F0, SQ_F1, SQ_F2, SQ_PN, SQ_S0, SQ_S1, SQ_S2, SQ_KE };

static double *surf_at(mzo_surf *sp, int64_t species, int64_t element) {
    return sp->q + ((species - 1) * sp->n_elements + element - 1) * MZO_SURF_Q;
}

static void surf_update(mzo_surf *sp, double w, const double *v, int64_t species, int64_t element, double sgn) {
    const double *nrm = sp->normals + 3 * (element - 1);
    double *q = surf_at(sp, species, element);
    const double px = w * v[0], py = w * v[1], pz = w * v[2];
    const double p_dot_n = fma(px, nrm[0], fma(py, nrm[1], pz * nrm[2]));
    const double ke = 0.5 * fma(px, v[0], fma(py, v[1], pz * v[2]));
    if (sgn > 0.0) {
        q[SQ_NP] += 1.0;
        q[SQ_FI] += w;
    } else {
        q[SQ_FR] -= w;
    }
    q[SQ_F0] += sgn * px;
    q[SQ_F1] += sgn * py;
    q[SQ_F2] += sgn * pz;
    q[SQ_PN] -= sgn * p_dot_n;
    q[SQ_S0] += sgn * fma(-p_dot_n, nrm[0], px);
    q[SQ_S1] += sgn * fma(-p_dot_n, nrm[1], py);
    q[SQ_S2] += sgn * fma(-p_dot_n, nrm[2], pz);
    q[SQ_KE] += sgn * ke;
}

void mzo_surf_update_incident(mzo_surf *sp, double w, const double *v, int64_t species, int64_t element) {
    surf_update(sp, w, v, species, element, 1.0);
}

void mzo_surf_update_reflected(mzo_surf *sp, double w, const double *v, int64_t species, int64_t element) {
    surf_update(sp, w, v, species, element, -1.0);
}

/* surface_props_scale! (surface_props.jl:144-160): every quantity but np
 * times (mass / dt) / area. */
void mzo_surf_scale(mzo_surf *sp, int64_t species, double mass, double dt) {
    const double factor_base = mass / dt;
    for (int64_t e = 1; e <= sp->n_elements; ++e) {
        const double factor = factor_base * sp->inv_areas[e - 1];
        double *q = surf_at(sp, species, e);
        for (int k = SQ_FI; k < MZO_SURF_Q; ++k) q[k] *= factor;
    }
}

/* clear_props!(surf_props) (surface_props.jl:171-189) */
void mzo_surf_clear(mzo_surf *sp) {
    const int64_t n = sp->n_species * sp->n_elements * MZO_SURF_Q;
    for (int64_t i = 0; i < n; ++i) sp->q[i] = 0.0;
}

/* avg_props!(surf_props_avg, surf_props, n) (surface_props.jl:202-222):
 * avg = muladd(inst, 1 / n, avg) for every quantity. */
void mzo_surf_avg(mzo_surf *avg, const mzo_surf *sp, double n_avg_timesteps) {
    const double inv = 1.0 / n_avg_timesteps;
    const int64_t n = sp->n_species * sp->n_elements * MZO_SURF_Q;
    for (int64_t i = 0; i < n; ++i) avg->q[i] = fma(sp->q[i], inv, avg->q[i]);
}

/* convect_single_particle! with surface properties (convection_1D.jl:72-112):
 * each wall hit adds the incident particle, reflects it, then subtracts the
 * reflected one, on element bc + 1. */
static void convect_single_surf(mzo_rng *rng, const mzo_grid1d *g, const mzo_walls *wl, double w, double *x,
                                double *v, int64_t species, double dt, mzo_surf *sp) {
    double t_rest = dt;
    double x_old = x[0];
    double x_new = fma(v[0], dt, x[0]);
    while ((x_new >= g->L) || (x_new <= 0.0)) {
        int bc;
        double wn;
        if (x_new >= g->L) {
            t_rest -= fabs((g->L - x_old) / v[0]);
            bc = 1; wn = -1.0; x_old = g->L;
        } else {
            t_rest -= fabs(x_old / v[0]);
            bc = 0; wn = 1.0; x_old = 0.0;
        }
        mzo_surf_update_incident(sp, w, v, species, bc + 1);
        reflect_particle_x(rng, v, wl->vsq[bc * wl->n_species + species - 1], wn, wl->v[bc], wl->acc[bc]);
        mzo_surf_update_reflected(sp, w, v, species, bc + 1);
        x_new = fma(v[0], t_rest, x_old);
    }
    if (x_new < g->min_x) x_new = g->min_x;
    else if (x_new > g->max_x) x_new = g->max_x;
    x[0] = x_new;
}

/* convect_particles!(rng, grid, walls, pv, pia, species, species_data,
 * surf_props, dt) (convection_1D.jl:176-207): clear_props! of every species,
 * the particles in the reference's order, then surface_props_scale!. */
void mzo_convect_particles_surf(mzo_rng *rng, const mzo_grid1d *g, const mzo_walls *wl, mzo_pv *pv,
                                mzo_pia *pia, int64_t species, double dt, mzo_surf *sp, double mass) {
    mzo_surf_clear(sp);
    if (pia->contiguous[species - 1]) {
        int64_t n_tot = pia->n_total[species - 1];
        for (int64_t i = 1; i <= n_tot; ++i) {
            int64_t id = PID(pv, i);
            mzo_rng_set_stream(rng, (uint64_t)(id - 1));
            convect_single_surf(rng, g, wl, PV_W(pv, id), PV_X(pv, id), PV_V(pv, id), species, dt, sp);
        }
    } else {
        for (int64_t c = 1; c <= g->n_cells; ++c) {
            mzo_ind *d = mzo_pia_at(pia, c, species);
            for (int grp = 0; grp < 2; ++grp) {
                if (grp == 1 && d->n_group2 <= 0) break;
                const int64_t s = grp == 0 ? d->start1 : d->start2, e = grp == 0 ? d->end1 : d->end2;
                for (int64_t i = s; i <= e; ++i) {
                    int64_t id = PID(pv, i);
                    mzo_rng_set_stream(rng, (uint64_t)(id - 1));
                    convect_single_surf(rng, g, wl, PV_W(pv, id), PV_X(pv, id), PV_V(pv, id), species, dt, sp);
                }
            }
        }
    }
    mzo_surf_scale(sp, species, mass, dt);
}

void mzo_props_avg(mzo_props *avg, const mzo_props *p, double n_avg_timesteps) {
    const double inv = 1.0 / n_avg_timesteps;
    const int64_t pn = p->n_cells * p->n_species;
    for (int64_t s = 0; s < p->n_species; ++s) avg->lpa[s] = fma(p->lpa[s], inv, avg->lpa[s]);
    for (int64_t i = 0; i < pn; ++i) {
        avg->np[i] = fma(p->np[i], inv, avg->np[i]);
        avg->n[i] = fma(p->n[i], inv, avg->n[i]);
        for (int d = 0; d < 3; ++d) avg->v[3 * i + d] = fma(p->v[3 * i + d], inv, avg->v[3 * i + d]);
        avg->T[i] = fma(p->T[i], inv, avg->T[i]);
    }
}

/* One particle (physical id) through its own stream; used by the parallel
 * drivers. */
void mzo_convect_one(mzo_rng *rng, const mzo_grid1d *g, const mzo_walls *wl, mzo_pv *pv, int64_t id,
                     int64_t species, double dt) {
    mzo_rng_set_stream(rng, (uint64_t)(id - 1));
    convect_single(rng, g, wl, PV_X(pv, id), PV_V(pv, id), species, dt);
}

/* compute_flux_props! (flux_props.jl:77-131), the reference's expressions
 * left to right: (w C) |C|^2, w C_k^2, w (C_k C_l), then 0.5 q m / V and q m / V */
void mzo_compute_flux_props(mzo_pv **pvs, const mzo_pia *pia, const double *masses, const mzo_props *pp,
                            const double *inv_V, double *flux, int64_t cell_lo, int64_t cell_hi) {
    const int64_t nc = pia->n_cells;
    for (int64_t s = 1; s <= pia->n_species; ++s) {
        mzo_pv *pv = pvs[s - 1];
        for (int64_t c = cell_lo; c <= cell_hi; ++c) {
            const mzo_ind *d = mzo_pia_at((mzo_pia *)pia, c, s);
            const double *u = &pp->v[3 * ((s - 1) * nc + (c - 1))];
            double q[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int g = 0; g < 2; ++g) {
                int64_t a = g == 0 ? d->start1 : d->start2, b = g == 0 ? d->end1 : d->end2;
                if (g == 1 && d->n_group2 <= 0) continue;
                for (int64_t i = a; i <= b; ++i) {
                    const int64_t id = PID(pv, i);
                    const double w = PV_W(pv, id), *v = PV_V(pv, id);
                    const double c0 = v[0] - u[0], c1 = v[1] - u[1], c2 = v[2] - u[2];
                    const double cx = c0 * c0, cy = c1 * c1, cz = c2 * c2;
                    const double csq = cx + cy + cz;
                    q[0] = q[0] + w * c0 * csq; q[1] = q[1] + w * c1 * csq; q[2] = q[2] + w * c2 * csq;
                    q[3] = q[3] + w * cx; q[4] = q[4] + w * cy; q[5] = q[5] + w * cz;
                    q[6] = q[6] + w * (c0 * c1); q[7] = q[7] + w * (c0 * c2); q[8] = q[8] + w * (c1 * c2);
                }
            }
            double *o = &flux[9 * ((s - 1) * nc + (c - 1))];
            const double m = masses[s - 1], iv = inv_V[c - 1];
            for (int j = 0; j < 3; ++j) o[j] = 0.5 * q[j] * m * iv;
            for (int j = 3; j < 9; ++j) o[j] = q[j] * m * iv;
        }
    }
}
