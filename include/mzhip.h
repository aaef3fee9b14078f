/*
 * mzhip.h -- C ABI of libmzhip.so, the gfx950 (MI355X) implementation of
 * Merzbild.jl's per-timestep particle path.
 *
 * The reference path is a set of Julia functions over in-place-mutated
 * structs (src/Merzbild.jl:25-71).  A Julia shim maps those signatures onto
 * these entry points with ccall; Python ctypes drives them in the tests.
 *
 * Conventions
 *   - every function returns 0 on success, <0 on error; mz_last_error()
 *     returns a message for the calling thread;
 *   - cells and species are 1-based and PIA values are the reference's
 *     1-based positions, so tables round-trip unchanged;
 *   - host buffers are owned by the caller and copied; device state is owned
 *     by the store handle; one HIP stream per handle; calls on one handle are
 *     not re-entrant;
 *   - rng_mode 0 = Philox4x32-10 keyed by (seed, tag, stream, draw#),
 *     rng_mode 1 = replay of a caller-supplied table [stream][draw#];
 *     streams are cell-1 for NTC and merges, physical id-1 for convection.
 */
#ifndef MZHIP_H
#define MZHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mz_store mz_store;

/* ------------------------------------------------------------------ core */
const char *mz_last_error(void);
int mz_device_info(char *name, int name_len, int *n_cu, double *hbm_gbps_peak);

/* A store of n_species ParticleVectors + one ParticleIndexerArray over
 * n_cells cells: ParticleVector(np) / ParticleIndexerArray(n_cells,
 * n_species) (src/particles.jl:138-174, 209-212).  capacity[s] bounds the
 * logical positions of species s+1. */
int mz_create(int32_t n_cells, int32_t n_species, const int64_t *capacity, mz_store **out);
int mz_destroy(mz_store *h);
int mz_sync(mz_store *h);

/* Upload one species in the reference layout: x, v ([n_pos][3]) and w of
 * the particles at logical positions 1..n_pos, their physical ids
 * (pv.index[1:n_pos]), the PIA table pia[n_cells][7] =
 * (n_local, start1, end1, n_group1, start2, end2, n_group2), n_total,
 * contiguous, the free-id buffer pv.buffer[1:nbuffer] and length(pv).
 * Replaces building ParticleVector / ParticleIndexerArray on the host
 * (src/particles.jl:56-212). */
int mz_upload_species(mz_store *h, int32_t species, int64_t n_pos, const double *x, const double *v,
                      const double *w, const int64_t *ids, const int64_t *pia, int64_t n_total,
                      int32_t contiguous, const int64_t *buffer, int64_t nbuffer, int64_t length);

/* Set pv.cell[1:n] (the cell of each logical position), as a caller does
 * before sort_particles!(gridsort, pv, pia, species)
 * (src/grids/grid_sorting.jl:128-182). */
int mz_upload_cell_keys(mz_store *h, int32_t species, int64_t n, const int64_t *cells);

/* Scalars of a species: out[0..5] = n_total, contiguous, nbuffer, length,
 * n_pos, error bits. */
int mz_get_scalars(mz_store *h, int32_t species, int64_t *out);

/* Download one species (arrays sized as in mz_upload_species; any pointer
 * may be NULL).  cell_keys receives pv.cell as the last sort left it. */
int mz_download_species(mz_store *h, int32_t species, int64_t n_pos_max, double *x, double *v,
                        double *w, int64_t *ids, int64_t *cell_keys, int64_t *pia,
                        int64_t *buffer, int64_t buffer_max);

/* ------------------------------------------------------------ parameters */
/* Grid1DUniform(L, nx; wall_offset) (src/grids/grid_uniform1D.jl:49-86) */
int mz_set_grid_1d(mz_store *h, double L, int64_t nx, double wall_offset);
/* MaxwellWalls1D (src/convection/boundary_conditions.jl:29-53): T[2],
 * v[2][3], accommodation[2]; vsq[2][n_species] = 2 k_B T_wall / m. */
int mz_set_walls_1d(mz_store *h, const double *T, const double *v, const double *acc, const double *vsq);
/* Interaction[s1,s2] (src/collisions/collision_utils.jl:73-82): mu1, mu2,
 * vhs_factor, vhs_o. */
int mz_set_interaction(mz_store *h, int32_t s1, int32_t s2, double mu1, double mu2, double vhs_factor,
                       double vhs_o);
/* CollisionFactors[s1,s2,cell].sigma_g_w_max round trip
 * (src/collisions/collision_ntc.jl:18-32). */
int mz_set_sigma_g_w_max(mz_store *h, int32_t s1, int32_t s2, const double *sgwm);
/* sgwm and the counters n1, n2, n_coll, n_coll_performed,
 * n_eq_w_coll_performed of the last ntc call (any pointer may be NULL). */
int mz_get_collision_factors(mz_store *h, int32_t s1, int32_t s2, double *sgwm, int64_t *n1, int64_t *n2,
                             int64_t *n_coll, int64_t *n_performed, int64_t *n_eq_w);
int mz_set_seed(mz_store *h, uint64_t seed);
/* compute_flux_props! (src/properties/flux_props.jl:77-131) over cells
 * cell_lo..cell_hi (1-based; all cells is the whole-grid form, a range the
 * cell_chunk form of compute_flux_props_sorted!, :199-249), groups 1 and 2,
 * all species, with the mean velocities of the store's last mz_compute_props
 * and the grid's cell volumes.  The store keeps an instantaneous and an
 * averaged set (FluxProps, :27-58). */
int mz_compute_flux_props(mz_store *h, const double *masses, int32_t cell_lo, int32_t cell_hi);
/* avg_props!(flux_props_avg, flux_props, n_avg_timesteps) (:152-186) on the store's two sets */
int mz_flux_avg(mz_store *h, double n_avg_timesteps);
/* clear_props!(flux_props) (:133-137) of set which (0 instantaneous, 1 average) */
int mz_flux_clear(mz_store *h, int32_t which);
/* Copy set which out; each array is [species][cell][3] (the reference's
 * (3, n_cells, n_species) arrays in column-major order).  NULL skips one. */
int mz_get_flux(mz_store *h, int32_t which, double *kinetic_energy_flux, double *diagonal_momentum_flux,
                double *off_diagonal_momentum_flux);

/* Replay tables (device copies): kind 0 = NTC, 1 = convection, 2 = merges. */
int mz_set_replay(mz_store *h, int32_t kind, const double *table, int64_t rows, int64_t cols);

/* ------------------------------------------------------------ operations */
/* sort_particles!(gridsort, grid, pv, pia, species) when compute_cells = 1
 * and sort_particles!(gridsort, pv, pia, species) with the stored cell keys
 * when 0 (src/grids/grid_sorting.jl:58-113, 128-182).  Squashes first if the
 * PIA is not contiguous. */
int mz_sort(mz_store *h, int32_t species, int32_t compute_cells);
/* mz_sort without allocating: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The count matrix, the
 * cell starts and (when the species is not contiguous) the squash's group
 * starts use a fixed device array of 2^21 int32 words (8 MiB) that each GPU
 * holds until the process ends; no hipMalloc, hipFree or allocating library
 * call is made.  It rejects with an error the radix path (more than 14336
 * cells, or a count matrix over 8 x capacity words) and shapes whose
 * n_cells x (ceil(capacity / 1024) + 1) exceeds 2^21.  Results equal
 * mz_sort's.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (several stores' streams at once) race. */
int mz_sort_noalloc(mz_store *h, int32_t species, int32_t compute_cells);
/* Which sort runs (a tuning knob; the reference has one sort): 0 chooses by
 * size, 1 the per-chunk count matrix over cells (at most 14,336 cells), 2 the
 * 8-bit-digit radix sort, whose traffic does not grow with the cell count.
 * Both give the same stable order. */
int mz_set_sort_path(mz_store *h, int32_t path);

/* squash_pia!(pv, pia, species) (src/particles.jl:622-701) */
int mz_squash(mz_store *h, int32_t species);
/* restore_particle_ordering!(pv, inv_map) (src/particles.jl:1086-1137):
 * afterwards the index is the identity and the free-id buffer holds
 * length, length - 1, ... ; the particles at the logical positions and the
 * PIA are unchanged (the store already keeps data in logical order, so no
 * data moves). */
int mz_restore_particle_ordering(mz_store *h, int32_t species);

/* convect_particles!(rng, grid, walls, pv, pia, species, species_data, dt)
 * (src/convection/convection_1D.jl:130-157). */
int mz_convect(mz_store *h, int32_t species, double dt, uint64_t tag, int32_t rng_mode);

/* Surface properties of the 1-D grid's two walls (SurfProps(pia, grid),
 * src/properties/surface_props.jl:22-64): element 1 = left wall (normal +x),
 * element 2 = right wall (normal -x).  The store keeps two sets, which = 0
 * (the last convection's values) and which = 1 (the time average), each
 * double[n_species][2][11] with the quantities of one element in the order
 * np, flux_incident, flux_reflected, force[3], normal_pressure,
 * shear_pressure[3], kinetic_energy_flux.
 *
 * convect_particles!(rng, grid, walls, pv, pia, species, species_data,
 * surf_props, dt) (src/convection/convection_1D.jl:176-207): convection as in
 * mz_convect, plus clear_props! of set 0, the contacts of every particle that
 * reaches a wall (update_surface_incident!/update_surface_reflected!,
 * surface_props.jl:77-128) and surface_props_scale! with the given mass
 * (surface_props.jl:144-160).  The sums are reproducible from run to run;
 * they differ from the reference's particle-order sums by rounding only. */
int mz_convect_surf(mz_store *h, int32_t species, double dt, uint64_t tag, int32_t rng_mode, double mass);
/* SurfProps areas (inv_areas = 1 / areas) of the two elements; default 1. */
int mz_set_surf_areas(mz_store *h, const double *areas);
/* avg_props!(surf_props_avg, surf_props, n_avg_timesteps)
 * (surface_props.jl:202-222): set 1 += set 0 / n_avg_timesteps, on device. */
int mz_surf_avg(mz_store *h, double n_avg_timesteps);
/* clear_props!(surf_props) (surface_props.jl:171-189) of set `which`. */
int mz_surf_clear(mz_store *h, int32_t which);
/* copy set `which` to / from host memory (double[n_species][2][11]). */
int mz_get_surf(mz_store *h, int32_t which, double *out);
int mz_set_surf(mz_store *h, int32_t which, const double *in);

/* ntc! / ntc_equal_weight! for cells cell_lo..cell_hi, single species when
 * s1 == s2, two species otherwise (src/collisions/collision_ntc.jl:338-595).
 * V[n_cells] are the cell volumes.  equal_weight selects ntc_equal_weight!.
 * Cells are processed in parallel; particles created by splits are appended
 * to group 2 in cell order, as a sequential loop over cells would. */
int mz_ntc(mz_store *h, int32_t s1, int32_t s2, double dt, const double *V, double dw_tol,
           int32_t equal_weight, uint64_t tag, int32_t rng_mode, int32_t cell_lo, int32_t cell_hi);
/* The caller's loop of per-cell ntc! / ntc_equal_weight! calls as one call
 * (the loop of test/test_1D_couette.jl:67-71; collision_ntc.jl:338,479 per
 * cell): cell cells[k] (1-based) with volume V[k], in list order, all with
 * the pair (s1, s2), dt, dw_tol and tag.  Runs of consecutive ascending
 * cells launch together; the result equals the calls one by one. */
int mz_ntc_cells(mz_store *h, int32_t s1, int32_t s2, double dt, const int32_t *cells, const double *V,
                 int32_t n, double dw_tol, int32_t equal_weight, uint64_t tag, int32_t rng_mode);

/* swpm!(rng, collision_factors_swpm, collision_data, interaction, pv, pia,
 * cell, species, G, dt, V) (src/collisions/collision_swpm.jl:201-288) for
 * cells cell_lo..cell_hi of one species, cells in parallel, each on its own
 * stream (cell - 1) under `tag`.  For this method the pair's array of
 * mz_set_sigma_g_w_max / mz_get_collision_factors holds (sigma g)_max (the
 * reference's CollisionFactorsSWPM.sigma_g_max), and the counters n1, n2,
 * n_coll, n_coll_performed (n_eq_w = 0).  Created particles (two per
 * performed collision) are appended to group 2 in cell order, as a sequential
 * loop over cells would append them. */
int mz_swpm(mz_store *h, int32_t species, double G, double dt, const double *V, uint64_t tag, int32_t rng_mode,
            int32_t cell_lo, int32_t cell_hi);

/* compute_props! (variant 0, groups 1 and 2) or compute_props_sorted!
 * (variant 1, group 1; ndens_not_Np with inv_V from the grid)
 * (src/properties/physical_props.jl:104-154, 317-454).  Results stay on the
 * device until mz_get_props. masses[n_species]. */
int mz_compute_props(mz_store *h, int32_t variant, int32_t ndens_not_Np, const double *masses);
/* compute_props_with_total_moments! (src/properties/physical_props.jl:168-245):
 * n, v, T as compute_props! plus the total moments for powers[n_moments]
 * (at most 8), scaled by the Maxwellian values at Tref. */
int mz_compute_props_moments(mz_store *h, const int32_t *powers, int32_t n_moments, double Tref,
                             const double *masses);
/* moments[n_species][n_cells][n_moments] of the last mz_compute_props_moments. */
int mz_get_props_moments(mz_store *h, double *moments);
/* compute_mixed_moment(particles, pia, cell, species, powers; sum_scaler,
 * res_scaler) (src/properties/physical_props.jl:473-501) for every cell in
 * cell_lo..cell_hi (1-based, inclusive): out[cell - cell_lo] =
 * res_scaler * sum over groups 1 and 2 of w sum_scaler vx^px vy^py vz^pz. */
int mz_compute_mixed_moment(mz_store *h, int32_t species, int32_t cell_lo, int32_t cell_hi, int32_t px,
                            int32_t py, int32_t pz, double sum_scaler, double res_scaler, double *out);
/* PhysProps arrays [n_species][n_cells] (v: [n_species][n_cells][3]); any may be NULL. */
int mz_get_props(mz_store *h, double *lpa, double *np, double *n, double *v, double *T);

/* avg_props!(phys_props_avg, phys_props, n_avg_timesteps)
 * (src/properties/physical_props.jl:281-299) on the device: the store's
 * time-averaged props += the last computed props / n_avg_timesteps (np, n,
 * v, T; lpa from the particle vector's length).  mz_props_avg_clear zeroes
 * the average, mz_get_props_avg copies it out in mz_get_props's layout. */
int mz_props_avg(mz_store *h, double n_avg_timesteps);
int mz_props_avg_clear(mz_store *h);
int mz_get_props_avg(mz_store *h, double *lpa, double *np, double *n, double *v, double *T);

/* merge_octree_N2_based!(rng, octree, pv, pia, cell, species, target_np[, grid])
 * for every cell whose n_local exceeds `threshold`
 * (src/merging/merging_octree_N2.jl:1060-1094), then squash_pia! if
 * squash != 0.  split: 1 mid, 2 mean; init_bounds: 1 min/max, 2 symmetric,
 * 3 +-c; bounds_mode: 1 inherit, 2 recompute. n_merged may be NULL. */
int mz_merge_octree(mz_store *h, int32_t species, int32_t threshold, int32_t target_np, int32_t split,
                    int32_t init_bounds, int32_t bounds_mode, int32_t max_Nbins, int32_t max_depth,
                    int32_t grid_clamp, int32_t squash, uint64_t tag, int32_t rng_mode, int64_t *n_merged);
/* The caller's loop "if n_local of cell c exceeds thresholds[c], call
 * merge_octree_N2_based!(rng, octree, pv, pia, c, species, targets[c])"
 * (merging_octree_N2.jl:1060-1094; a per-cell target such as each cell's
 * initial count) as one call: thresholds and targets are host arrays of
 * n_cells entries, the other arguments as mz_merge_octree.  Cells of more
 * than 2048 particles or more than 512 bins run on a device-memory path
 * with the same results. */
int mz_merge_octree_targets(mz_store *h, int32_t species, const int32_t *thresholds, const int32_t *targets,
                            int32_t split, int32_t init_bounds, int32_t bounds_mode, int32_t max_Nbins,
                            int32_t max_depth, int32_t grid_clamp, int32_t squash, uint64_t tag, int32_t rng_mode,
                            int64_t *n_merged);
/* The caller's loop of per-cell merge_octree_N2_based!(rng, octree, pv, pia,
 * cell, species, target_np[, grid]) calls as one call
 * (merging_octree_N2.jl:1060,1088; the loop of
 * test/test_1D_couette_varweight.jl:78-87): cells[k] (1-based) in list
 * order, whatever their occupancy; no squash (the caller's squash_pia!
 * follows, as in the reference).  Runs of ascending cells launch together;
 * the result equals the calls one by one. */
int mz_merge_octree_cells(mz_store *h, int32_t species, const int32_t *cells, int32_t n, int32_t target_np,
                          int32_t split, int32_t init_bounds, int32_t bounds_mode, int32_t max_Nbins,
                          int32_t max_depth, int32_t grid_clamp, uint64_t tag, int32_t rng_mode, int64_t *n_merged);

/* merge_nnls_based! for every cell whose n_local exceeds `threshold`
 * (src/merging/merging_nnls.jl:1457-1510).  mim[n_mim][3] are the requested
 * multi-indices (base moments are added as the reference does);
 * mim_pos[n_mim_pos][3] the spatial ones (NNLSMerge's
 * multi_index_moments_pos, :128; at most 16).  n_rand_pairs random-pair
 * columns follow the fictitious ones (:1004-1030), drawn from the cell's
 * stream of the merge replay table or Philox (seed, tag).  status[n_cells]
 * receives 1, -1, or 0 for cells not merged; NULL allowed. */
int mz_merge_nnls(mz_store *h, int32_t species, int32_t threshold, const int32_t *mim, int32_t n_mim,
                  const int32_t *mim_pos, int32_t n_mim_pos, double vref, int32_t scaling_vref, int32_t n_rand_pairs,
                  int32_t centered_at_mean, const double *v_multipliers, int32_t n_vmult, double max_err,
                  int32_t iteration_mult, double w_threshold, int32_t squash, uint64_t tag, int32_t rng_mode,
                  int32_t *status);

/* merge_roulette!(rng, pv, pia, cell, species, target_np; conservative)
 * (src/merging/merging_roulette.jl:25-184) on every cell holding more than
 * `threshold` particles (at most 8192 per cell; larger cells flag a capacity
 * error and are left alone), each drawing from its own stream (cell - 1)
 * under `tag`.  Deleted ids are pushed to the free-id buffer in cell order,
 * in draw order within a cell, as the reference's cell loop pushes them; the
 * PIA is left not contiguous.  *n_merged (may be NULL) receives the number
 * of cells merged. */
int mz_merge_roulette(mz_store *h, int32_t species, int32_t threshold, int32_t target_np, int32_t conservative,
                      uint64_t tag, int32_t rng_mode, int64_t *n_merged);

/* merge_grid_based!(rng, merging_grid, pv, pia, cell, species, species_data,
 * phys_props[, grid]) and (..., vx_extent, vy_extent, vz_extent[, grid])
 * (src/merging/merging_grid.jl:597-703) on every cell with n_local >
 * threshold, for GridN2Merge(Nx, Ny, Nz, extent_multiplier) (:72-115,
 * Nx Ny Nz + 8 <= 4104, at most 2,048 particles per merged cell).  With
 * extents == NULL the velocity grid spans v -/+ extent_multiplier[d] *
 * sqrt(2 k T / mass) from the cell's T and v of the last mz_compute_props of
 * this species (compute_velocity_extent!, :190-197); otherwise extents =
 * {vx_lo, vx_hi, vy_lo, vy_hi, vz_lo, vz_hi} (:210-220).  grid_clamp = 1
 * puts x of the merged pairs back into the grid (:484-595).  Six sign draws
 * per velocity cell of more than two particles, from stream cell - 1 of the
 * merge family.  Deleted ids are pushed as by merge_octree; squash = 1 then
 * runs squash_pia!.  *n_merged (may be NULL) receives the cells merged. */
int mz_merge_grid(mz_store *h, int32_t species, int32_t threshold, int32_t Nx, int32_t Ny, int32_t Nz,
                  const double *extent_multiplier, const double *extents, double mass, int32_t grid_clamp,
                  int32_t squash, uint64_t tag, int32_t rng_mode, int64_t *n_merged);

/* merge_nnls_based! with a bounded candidate set: the columns are the first
 * max_pcols particles of the cell (logical order) plus the reference's
 * fictitious particles, while the moments to match are still those of all
 * the cell's particles (src/merging/merging_nnls.jl:1457-1510 with the column
 * set of compute_lhs_and_rhs!, :509-613, cut at max_pcols; max_pcols < 0 is
 * mz_merge_nnls).  method 0 runs the reference's column-by-column layout
 * (matrix in device memory, results equal to the CPU restatement bit for
 * bit); method 1 keeps the matrix in LDS with one thread per column and
 * block-parallel moment sums (needs max_pcols + fictitious columns <= 256,
 * results equal to rounding); method 2 runs the active-set search on the
 * matrix stored in FP32 and solves the final passive set again in FP64
 * (normal equations, one refinement step), re-running in method 1 every
 * cell whose FP64 solution is not positive or misses max_err (at most 35
 * moments; more run method 1).  Spatial moments and random pairs as
 * mz_merge_nnls. */
int mz_merge_nnls_cand(mz_store *h, int32_t species, int32_t threshold, const int32_t *mim, int32_t n_mim,
                       const int32_t *mim_pos, int32_t n_mim_pos, double vref, int32_t scaling_vref,
                       int32_t n_rand_pairs, int32_t centered_at_mean, const double *v_multipliers, int32_t n_vmult,
                       double max_err, int32_t iteration_mult, double w_threshold, int32_t squash, int32_t max_pcols,
                       int32_t method, uint64_t tag, int32_t rng_mode, int32_t *status);
/* The caller's loop of per-cell merge_nnls_based!(rng, nnls_merging, pv, pia,
 * cell, species; ...) calls as one call (merging_nnls.jl:1457-1510):
 * cells[k] (1-based) in list order; status_out[k] is the reference's return
 * value, 1 (merged) or -1 (failed).  No squash.  max_pcols < 0 takes every
 * particle as a column (the reference's candidate set, method 0);
 * otherwise max_pcols and method as in mz_merge_nnls_cand. */
int mz_merge_nnls_cells(mz_store *h, int32_t species, const int32_t *cells, int32_t n, const int32_t *mim,
                        int32_t n_mim, const int32_t *mim_pos, int32_t n_mim_pos, double vref, int32_t scaling_vref,
                        int32_t n_rand_pairs, int32_t centered_at_mean, const double *v_multipliers,
                        int32_t n_vmult, double max_err, int32_t iteration_mult, double w_threshold,
                        int32_t max_pcols, int32_t method, uint64_t tag, int32_t rng_mode, int32_t *status_out);
/* info[0] = cells the last mz_merge_nnls_cand call merged or tried,
 * info[1] = cells its method 2 handed to the FP64 method. */
int mz_get_nnls_info(mz_store *h, int32_t *info);

/* ----------------------------------------------------------- time stepping */
/* One 1-D Couette step repeated n_steps times: ntc! in every cell,
 * convection, sort, compute_props_sorted! every props_every steps
 * (test/test_1D_couette.jl:66-97).  Step t uses tags (t << 4) | 1 for NTC
 * and (t << 4) | 2 for convection.  mode 0 launches kernels one by one,
 * mode 1 replays a captured hipGraph of the step, mode 2 runs all steps in
 * one launch with each slab of cells resident in LDS (equal weights; it
 * falls back to mode 1 when the state is not the sorted state a sort leaves,
 * or when a particle crosses more than one slab in a step).  Mode 2 needs
 * every slab resident at once: the caller keeps other kernels of this
 * process off the device meanwhile (the library itself drains the device
 * first when another store's stream is busy).  Per call the host enqueues a
 * check-and-backup kernel, the launch and a report kernel, and waits once.
 * The chunked loop this replaces (collide, convect, sort per chunk, exchange
 * between chunks, re-sort and props) is simulations/1D/
 * couette_multithreaded.jl:97-161. */
int mz_couette_steps(mz_store *h, int32_t n_steps, uint64_t step0, double dt, const double *V,
                     int32_t equal_weight, int32_t props_every, const double *masses, int32_t mode);
/* The averaging phase of simulations/1D/couette_with_surface_quantities.jl:
 * 73-115: steps as mz_couette_steps with equal weights and props every step,
 * where convection also computes the wall properties (mz_convect_surf with
 * masses[0]) and each step adds the wall and cell properties / n_avg_timesteps
 * to the store's time averages (mz_surf_avg, mz_props_avg).  mode 0 launches
 * kernels one by one; modes 1 and 2 replay a captured hipGraph of the step. */
int mz_couette_avg_steps(mz_store *h, int32_t n_steps, uint64_t step0, double dt, const double *V,
                         const double *masses, int32_t mode, double n_avg_timesteps);
/* Which path the last mz_couette_steps call took: 0 eager, 1 graph, 2 slab. */
int mz_couette_last_path(mz_store *h);
/* Slab shape of mode 2 (the reference's n_chunks, couette_multithreaded.jl:
 * 13-14, 31): cells per slab (1..32) and particles per thread (1, 2, 4, 8) of
 * the LDS capacity 256 x ppt per slab; 0 lets the library choose. */
int mz_set_slab_shape(mz_store *h, int32_t cells_per_slab, int32_t particles_per_thread);
/* Workgroups per CU the slab kernel with the given particles per thread can
 * keep resident, and its LDS bytes per workgroup. */
int mz_slab_occupancy(int32_t particles_per_thread, int32_t *blocks_per_cu, int32_t *lds_bytes);
/* info[0..3] = particles per thread, cells per slab, slabs, outcome (1 ran,
 * -1 aborted and replayed on mode 1, -2 the state was not the sorted state,
 * -3 no slab shape fits the cells in LDS) of the last mode-2 call. */
int mz_get_slab_info(mz_store *h, int32_t *info);

/* Many independent 0-D cells: n_steps of ntc_equal_weight! in every cell of
 * `species` and compute_props! every props_every steps (the per-cell loops of
 * test/test_bkw.jl:88-93, test/test_2species.jl:52-65 over a batch of
 * cells).  mode 0 launches the NTC kernel over all cells and the props
 * kernel once per step; mode 2 launches NTC once per cell per step (the
 * reference's loop of ntc! calls, each a drop-in launch), then props; mode 1
 * runs all steps in one persistent launch whose wavefronts drain a queue of
 * cell groups ordered by size, large cells colliding on a whole wavefront.
 * All give identical results.  Step t uses tag (t << 4) | 1.  V[n_cells] as
 * for mz_ntc (NULL: the grid's cells). */
int mz_batch0d_steps(mz_store *h, int32_t species, int32_t n_steps, uint64_t step0, double dt, const double *V,
                     int32_t props_every, const double *masses, int32_t mode);
/* Packing of mode 1: at most max_cells (1..64) cells and about
 * target_particles particles per wavefront item; 0 keeps the default. */
int mz_set_batch_shape(mz_store *h, int32_t max_cells, int32_t target_particles);
/* info[0..3] = items, max cells per item, target particles, workgroups of the last mode-1 launch. */
int mz_get_batch_info(mz_store *h, int32_t *info);

/* ------------------------------------------------------------------ timing */
/* When enabled, the library brackets each launch of the named kernel family
 * with hipEvents on its stream.  Families: "sort", "convect", "ntc", "props",
 * "step" (a whole step).  Returns the mean duration in ms and the count. */
int mz_timing_enable(mz_store *h, int32_t on);
int mz_timing_get(mz_store *h, const char *family, double *mean_ms, int64_t *count);
int mz_timing_reset(mz_store *h);

#ifdef __cplusplus
}
#endif
#endif
