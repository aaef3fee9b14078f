/*
 * mzo.h -- CPU oracle for the Merzbild.jl per-timestep particle path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference method (Merzbild.jl v0.7.10, src/).  It is used by tests/, by the
 * smoke check and by bench.py's cpu_baseline leg, always as the checker and
 * never as the product path.  The product is the HIP library under
 * merzbild.jl_amd/csrc, which never links or calls this code.
 *
 * Layout mirrors the reference so that the integer outputs (index arrays,
 * ParticleIndexerArray fields, octree bins) can be compared bit for bit:
 *   - a ParticleVector stores particle data by physical id (1-based) and an
 *     `index` array mapping logical position -> physical id
 *     (src/particles.jl:194-212);
 *   - cells and species are 1-based in every argument, as in the reference.
 *
 * Randomness: every rand(rng, Float64) of the reference is replaced by
 * mzo_rand(), which draws either from Philox4x32-10 (Salmon et al., SC'11,
 * "Parallel random numbers: as easy as 1, 2, 3") keyed by (seed, tag, stream,
 * draw#) or from a caller-supplied replay table [stream][draw#].  The GPU
 * library uses the same counter layout, so both produce identical streams.
 * Stream conventions: NTC, octree and NNLS merges use stream = cell-1;
 * convection uses stream = physical id - 1.
 *
 * Parity status of each function is recorded in its comment.
 */
#ifndef MZO_H
#define MZO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MZO_DELTA_PARTICLES 256   /* src/Merzbild.jl:23 */
#define MZO_KB 1.380649e-23       /* src/constants.jl:4 */
#define MZO_C_LIGHT 299792458.0   /* src/constants.jl:9 */

/* ---------------- random draws ---------------- */
typedef struct {
    int32_t mode;       /* 0 = Philox4x32-10, 1 = replay table */
    int32_t err;        /* set to 1 if a replay table row runs out */
    uint64_t seed;      /* Philox key */
    uint64_t tag;       /* per-call tag (counter words 2,3) */
    uint64_t stream;    /* counter word 1 */
    uint64_t k;         /* draw number within the stream */
    const double *table;
    int64_t rows, cols;
} mzo_rng;

void mzo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double mzo_rand(mzo_rng *r);
void mzo_rng_set_stream(mzo_rng *r, uint64_t stream);

/* ---------------- particles / indexing (src/particles.jl) ---------------- */
typedef struct {
    int64_t n_local, start1, end1, n_group1, start2, end2, n_group2;
} mzo_ind;

typedef struct {
    int64_t n_cells, n_species;
    mzo_ind *ind;        /* [(species-1)*n_cells + (cell-1)] */
    int64_t *n_total;    /* [n_species] */
    int8_t *contiguous;  /* [n_species] */
} mzo_pia;

typedef struct {
    int64_t length, cap;
    double *w;           /* [cap]  by physical id-1 */
    double *v;           /* [3*cap] */
    double *x;           /* [3*cap] */
    int64_t *index;      /* [cap]  logical position-1 -> physical id (1-based) */
    int64_t *cell;       /* [cap] */
    int64_t *buffer;     /* [cap]  LIFO of free physical ids */
    int64_t nbuffer;
} mzo_pv;

mzo_pv *mzo_pv_new(int64_t np);
void mzo_pv_free(mzo_pv *pv);
void mzo_pv_resize(mzo_pv *pv, int64_t n);
void mzo_pv_buffer_pop_to(mzo_pv *pv, int64_t position);
mzo_pia *mzo_pia_new(int64_t n_cells, int64_t n_species);
void mzo_pia_free(mzo_pia *pia);
mzo_ind *mzo_pia_at(mzo_pia *pia, int64_t cell, int64_t species);
int64_t mzo_map_cont_index(const mzo_ind *ind, int64_t i);
void mzo_update_indexer_new_particle(mzo_pia *pia, int64_t cell, int64_t species);
void mzo_update_indexer_new_lower_count(mzo_pia *pia, int64_t cell, int64_t species, int64_t new_count);
void mzo_update_buffer_index_new_particle(mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species);
void mzo_delete_particle_end(mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species);
void mzo_delete_particle(mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species, int64_t i);
void mzo_squash_pia(mzo_pv *pv, mzo_pia *pia, int64_t species);
int mzo_check_pia_is_correct(mzo_pia *pia, int64_t species, int64_t *where);
int mzo_check_unique_index(mzo_pv *pv, mzo_pia *pia, int64_t species, int64_t *code);
int mzo_check_unique_buffer(mzo_pv *pv, int64_t *where);
int64_t mzo_count_disordered_particles(mzo_pv *pv, mzo_pia *pia, int64_t species, int use_offset);
void mzo_restore_particle_ordering(mzo_pv *pv);

/* ---------------- grid, sorting, convection ---------------- */
typedef struct {
    double L;
    int64_t n_cells;
    double dx, inv_dx, min_x, max_x;
} mzo_grid1d;

void mzo_grid1d_init(mzo_grid1d *g, double L, int64_t nx, double wall_offset);
int64_t mzo_get_cell(const mzo_grid1d *g, double x0);
void mzo_sort_particles(const mzo_grid1d *g, mzo_pv *pv, mzo_pia *pia, int64_t species);
void mzo_sort_particles_precomputed(mzo_pv *pv, mzo_pia *pia, int64_t species);

typedef struct {
    double T[2];
    double v[2][3];
    double acc[2];
    int64_t n_species;
    double vsq[2 * 16];  /* [bc*n_species + species-1] = 2 k_B T_bc / m_species */
} mzo_walls;

void mzo_walls_init(mzo_walls *wl, const double *masses, int64_t n_species, double T_l, double T_r,
                    double vy_l, double vy_r, double acc_l, double acc_r);
void mzo_convect_particles(mzo_rng *rng, const mzo_grid1d *g, const mzo_walls *wl, mzo_pv *pv,
                           mzo_pia *pia, int64_t species, double dt);
void mzo_convect_one(mzo_rng *rng, const mzo_grid1d *g, const mzo_walls *wl, mzo_pv *pv, int64_t id,
                     int64_t species, double dt);

/* ---------------- surface properties (src/properties/surface_props.jl) ----
 * SurfProps of n_elements surface elements and n_species species.  Element e
 * (1-based) of species s keeps its MZO_SURF_Q quantities at
 * q[((s - 1) * n_elements + e - 1) * MZO_SURF_Q + k], k in the order
 * np, flux_incident, flux_reflected, force[3], normal_pressure,
 * shear_pressure[3], kinetic_energy_flux.  normals[3 (e - 1) + i] is
 * component i of element e's normal.  For the 1-D grid element 1 is the left
 * wall (normal +x) and element 2 the right wall (normal -x), the elements of
 * boundary condition 1 and 2 (surface_props.jl:63-64).
 * Parity: pinned by the known answers of test/test_surface_props_1D_uniform.jl
 * (tests/test_oracle_surf.py) and by the invariants of specular walls. */
#define MZO_SURF_Q 11
typedef struct {
    int64_t n_elements, n_species;
    double *inv_areas;   /* [n_elements] */
    double *normals;     /* [3 * n_elements] */
    double *q;           /* [n_species * n_elements * MZO_SURF_Q] */
} mzo_surf;

void mzo_surf_update_incident(mzo_surf *sp, double w, const double *v, int64_t species, int64_t element);
void mzo_surf_update_reflected(mzo_surf *sp, double w, const double *v, int64_t species, int64_t element);
void mzo_surf_scale(mzo_surf *sp, int64_t species, double mass, double dt);
void mzo_surf_clear(mzo_surf *sp);
void mzo_surf_avg(mzo_surf *avg, const mzo_surf *sp, double n_avg_timesteps);
void mzo_convect_particles_surf(mzo_rng *rng, const mzo_grid1d *g, const mzo_walls *wl, mzo_pv *pv,
                                mzo_pia *pia, int64_t species, double dt, mzo_surf *sp, double mass);

/* ---------------- collisions ---------------- */
typedef struct {
    double m_r, mu1, mu2, vhs_d, vhs_o, vhs_Tref, vhs_muref, vhs_factor;
} mzo_interaction;

typedef struct {
    int64_t n1, n2;
    double sigma_g_w_max;
    int64_t n_coll, n_coll_performed, n_eq_w_coll_performed;
} mzo_cf;

void mzo_interaction_init(mzo_interaction *ik, mzo_interaction *ki, double m_i, double m_k,
                          double d, double omega, double Tref);
double mzo_sigma_vhs(const mzo_interaction *it, double g);
double mzo_estimate_sigma_g_w_max(const mzo_interaction *it, double m1, double m2, double T1,
                                  double T2, double Fnum, double mult);
void mzo_ntc_single(mzo_rng *rng, mzo_cf *cf, const mzo_interaction *it, mzo_pv *pv, mzo_pia *pia,
                    int64_t cell, int64_t species, double dt, double V, double dw_tol, int equal_weight);
void mzo_ntc_two(mzo_rng *rng, mzo_cf *cf, const mzo_interaction *it, mzo_pv *pv1, mzo_pv *pv2,
                 mzo_pia *pia, int64_t cell, int64_t s1, int64_t s2, double dt, double V,
                 double dw_tol, int equal_weight);
/* one collision of two given particles (ids), as collide_2particles_vhs! */
void mzo_collide_pair(mzo_rng *rng, mzo_cf *cf, const mzo_interaction *it, mzo_pv *pv1, mzo_pv *pv2,
                      mzo_pia *pia, int64_t cell, int64_t s1, int64_t s2, int64_t id_i, int64_t id_k,
                      double dw_tol, int equal_weight);

/* CollisionFactorsSWPM (src/collisions/collision_swpm.jl:16-23) */
typedef struct {
    int64_t n1, n2;
    double sigma_g_max;
    int64_t n_coll, n_coll_performed;
} mzo_cf_swpm;

/* swpm!(rng, cf, collision_data, interaction, particles, pia, cell, species, G,
 * dt, V) (src/collisions/collision_swpm.jl:201-288) with
 * compute_n_coll_single_species (:171-174): single-species collisions of
 * variable-weight particles; an accepted pair gives up min(w_i, w_k) / (1 + G)
 * of each weight to two new particles, which carry the pair's post-collision
 * velocities.  Draw order per cell stream: [n_coll], per candidate [i, k (, k
 * redraws)], [accept] if g > eps, [phi, cos theta] if accepted.  Parity:
 * pinned by closed forms on replayed draws and the conservation laws
 * (tests/test_oracle_swpm.py). */
void mzo_swpm(mzo_rng *rng, mzo_cf_swpm *cf, const mzo_interaction *it, mzo_pv *pv, mzo_pia *pia, int64_t cell,
              int64_t species, double G, double dt, double V);

/* ---------------- properties (src/properties/physical_props.jl) ---------------- */
typedef struct {
    int64_t n_cells, n_species, n_moments;
    int32_t ndens_not_Np;
    double Tref;
    double *lpa;      /* [n_species] */
    double *np;       /* [n_species*n_cells], cell fastest */
    double *n;
    double *v;        /* [3*n_cells*n_species] */
    double *T;
    double *moments;  /* [n_moments*n_cells*n_species] */
    const int32_t *moment_powers;
} mzo_props;

/* avg_props!(phys_props_avg, phys_props, n_avg_timesteps)
 * (physical_props.jl:281-299): avg = muladd(props, 1 / n, avg) for lpa, np,
 * n, v and T.  Parity: pinned by closed forms (averaging constant props gives
 * them back; tests/test_oracle_surf.py). */
void mzo_props_avg(mzo_props *avg, const mzo_props *p, double n_avg_timesteps);

void mzo_compute_props(mzo_pv **pvs, mzo_pia *pia, const double *masses, mzo_props *pp);
void mzo_compute_props_sorted(mzo_pv **pvs, mzo_pia *pia, const double *masses, mzo_props *pp,
                              const double *inv_V, int64_t cell_lo, int64_t cell_hi);
void mzo_compute_props_with_total_moments(mzo_pv **pvs, mzo_pia *pia, const double *masses,
                                          mzo_props *pp);
double mzo_compute_mixed_moment(mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species,
                                int px, int py, int pz, double sum_scaler, double res_scaler);
void mzo_clear_props(mzo_props *pp);
/* compute_flux_props! (src/properties/flux_props.jl:77-131) over cells
 * cell_lo..cell_hi (1-based), groups 1 and 2, in particle order, with the
 * mean velocities in pp->v; flux is [species][cell][9]: energy flux xyz,
 * diagonal momentum flux xyz, off-diagonal xy xz yz.  Parity: pinned by
 * test/test_flux_computes.jl's values and by Maxwellian closed forms
 * (tests/test_oracle_flux.py). */
void mzo_compute_flux_props(mzo_pv **pvs, const mzo_pia *pia, const double *masses, const mzo_props *pp,
                            const double *inv_V, double *flux, int64_t cell_lo, int64_t cell_hi);

/* ---------------- octree N:2 merging (src/merging/merging_octree_N2.jl) ---------------- */
typedef struct {
    int64_t np;
    double w;
    double v_min[3], v_max[3];
    int64_t depth;
    int32_t can_be_refined;
} mzo_obin;

typedef struct {
    double v_mean[3], v_std_sq[3], x_mean[3], x_std_sq[3];
    int64_t particle_index1, particle_index2;
    double w1, w2, v1[3], v2[3], x1[3], x2[3];
} mzo_ofull;

typedef struct {
    int64_t max_Nbins, Nbins, n_particles;
    mzo_obin *bins;
    mzo_ofull *full;
    int64_t *bin_start, *bin_end;
    int64_t *pis, *poct, *psort;   /* particle_indexes_sorted, particle_octants, particles_sort_output */
    int64_t buf_len;
    int32_t bin_bounds_compute;    /* 1 inherit, 2 recompute */
    int32_t split;                 /* 1 mid, 2 mean */
    int32_t init_bin_bounds;       /* 1 minmax, 2 minmax-symmetric, 3 +-c */
    int64_t max_depth;
    int64_t total_post_merge_np;
    double vel_middle[3], v_min_parent[3], v_max_parent[3], direction_vec[3];
} mzo_octree;

mzo_octree *mzo_octree_new(int32_t split, int32_t init_bin_bounds, int32_t bin_bounds_compute,
                           int64_t max_Nbins, int64_t max_depth);
void mzo_octree_free(mzo_octree *oc);
void mzo_octree_init(mzo_octree *oc, mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species);
void mzo_octree_split_bin(mzo_octree *oc, int64_t bin_id, mzo_pv *pv);
void mzo_octree_compute_bin_props(mzo_octree *oc, int64_t bin_id, mzo_pv *pv);
void mzo_octree_compute(mzo_octree *oc, mzo_pv *pv, int64_t target_np);
void mzo_merge_octree_N2_based(mzo_rng *rng, mzo_octree *oc, mzo_pv *pv, mzo_pia *pia, int64_t cell,
                               int64_t species, int64_t target_np, const mzo_grid1d *grid);

/* ---------------- grid N:2 merging (src/merging/merging_grid.jl) ----------------
 * GridN2Merge (:72-115): Nx x Ny x Nz velocity cells plus the 8 octants
 * outside the grid, indexed 1..Ntotal as compute_grid_index (:235-270) does.
 * Parity: pinned by test/test_merging_grid_merging.jl's answers (extents,
 * conservation, zero-weight octants, two index groups) and closed forms
 * (tests/test_oracle_grid_merge.py). */
typedef struct {
    int64_t np;
    double w, v_mean[3], v_std_sq[3], x_mean[3], x_std_sq[3];
    int64_t particle_index1, particle_index2;
} mzo_gcell;

typedef struct {
    int64_t Nx, Ny, Nz, NyNz, Ntotal;
    double mult[3], lower[3], upper[3], mid[3], dv[3], dv_inv[3];
    mzo_gcell *cells;   /* [Ntotal], velocity cell index k at cells[k - 1] */
    int64_t out_of_range;   /* indices past Ntotal met (a velocity on an upper bound), clamped */
} mzo_gridmerge;

mzo_gridmerge *mzo_gridmerge_new(int64_t Nx, int64_t Ny, int64_t Nz, const double mult[3]);
void mzo_gridmerge_free(mzo_gridmerge *mg);
/* compute_velocity_extent! from a cell's T and mean velocity (:190-197) or from explicit bounds (:210-220) */
void mzo_gridmerge_extent_props(mzo_gridmerge *mg, double T, const double v[3], double mass);
void mzo_gridmerge_extent(mzo_gridmerge *mg, const double vx[2], const double vy[2], const double vz[2]);
int64_t mzo_gridmerge_index(mzo_gridmerge *mg, const double v[3]);
/* compute_grid! + compute_new_particles! (:304-595) on the extents set
 * above; grid may be NULL (no clamping).  Draws from stream cell - 1. */
void mzo_merge_grid_based(mzo_rng *rng, mzo_gridmerge *mg, mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species,
                          const mzo_grid1d *grid);

/* merge_roulette!(rng, pv, pia, cell, species, target_np; conservative)
 * (src/merging/merging_roulette.jl:25-184): delete uniformly drawn particles
 * (one draw each, from the cell's stream) until target_np remain, then
 * re-weight the rest to keep the cell's total weight; conservative also
 * restores the mean velocity and the temperature.  Always clears
 * pia.contiguous.  Parity: pinned by test/test_roulette_merging.jl's answers
 * and the conservation laws (tests/test_oracle_roulette.py). */
void mzo_merge_roulette(mzo_rng *rng, mzo_pv *pv, mzo_pia *pia, int64_t cell, int64_t species, int64_t target_np,
                        int conservative);

/* ---------------- NNLS merging (src/merging/merging_nnls.jl, nnls.jl) ---------------- */
int mzo_nnls_solve(double *A, int64_t m, int64_t n, double *b, double *x, double *w, double *zz,
                   int64_t *idx, int64_t max_iter, double kkt_tol, double *rnorm, int64_t *nsetp);

typedef struct {
    double v0[3], x0[3];
    double vref, inv_vref;
    double Ex, Ey, Ez, Epx, Epy, Epz, w_total;
    double minvx, maxvx, minvy, maxvy, minvz, maxvz;
    double scalevx, scalevy, scalevz, scalex, scaley, scalez;
    int64_t n_total_conserved, n_moments_vel, n_moments_pos;
    int32_t mim[64][3];
    int32_t tot_order[64];
    int32_t mim_pos[64][3];
    int64_t pos_i_x, pos_i_y, pos_i_z;
    double rhs[128];
    /* diagnostics of the last solve */
    double rnorm;
    int64_t nsetp, lhs_ncols;
    int32_t mode;
    int32_t reset_x0;   /* 0: reproduce the reference (x0 carried between calls) */
} mzo_nnls_merge;

void mzo_nnls_merge_init(mzo_nnls_merge *nm, const int32_t *mim, int64_t n_mim, const int32_t *mim_pos,
                         int64_t n_mim_pos);
int mzo_merge_nnls_based(mzo_rng *rng, mzo_nnls_merge *nm, mzo_pv *pv, mzo_pia *pia, int64_t cell,
                         int64_t species, double vref, int32_t scaling_vref, int64_t n_rand_pairs,
                         double max_err, int32_t centered_at_mean, const double *v_multipliers,
                         int64_t n_vmult, int64_t iteration_mult, double w_threshold);
int mzo_merge_nnls_cand(mzo_rng *rng, mzo_nnls_merge *nm, mzo_pv *pv, mzo_pia *pia, int64_t cell,
                        int64_t species, double vref, int32_t scaling_vref, int64_t n_rand_pairs,
                        double max_err, int32_t centered_at_mean, const double *v_multipliers,
                        int64_t n_vmult, int64_t iteration_mult, double w_threshold, int64_t max_pcols);
void mzo_scale_columns(double *A, int64_t m, int64_t n, double *column_norms);

/* ---------------- multi-step drivers (CPU baseline) ---------------- */
int mzo_omp_max_threads(void);
void mzo_couette_steps(const mzo_rng *proto, const mzo_grid1d *g, const mzo_walls *wl,
                       const mzo_interaction *it, mzo_cf *cfs, mzo_pv *pv, mzo_pia *pia, double dt,
                       const double *V, double mass, mzo_props *pp, int64_t n_steps, uint64_t step0,
                       int32_t equal_weight, int32_t nthreads, int32_t props_every);

/* ---------------- samplers (src/distributions_and_sampling.jl) ---------------- */
void mzo_sample_maxwellian(mzo_rng *rng, mzo_pv *pv, int64_t nparticles, int64_t offset, double m,
                           double T, const double v0[3]);
void mzo_sample_particles_equal_weight(mzo_rng *rng, mzo_pv *pv, mzo_pia *pia, int64_t cell,
                                       int64_t species, int64_t nparticles, double m, double T,
                                       double Fnum, double xlo, double xhi, double ylo, double yhi,
                                       double zlo, double zhi);
void mzo_sample_bkw(mzo_rng *rng, mzo_pv *pv, int64_t nparticles, int64_t offset, double m, double T,
                    const double v0[3]);

#ifdef __cplusplus
}
#endif
#endif
