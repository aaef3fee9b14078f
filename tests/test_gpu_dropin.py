"""The drop-in interface (merzbild.jl_amd.dropin) driven by the reference's
own time loops, call line for call line, against the oracle running the
same sequence: the index array and the PIA bit for bit, velocities,
positions and temperatures to 1e-12.

* test/test_1D_couette.jl:66-97: ntc! per cell, convect_particles!,
  sort_particles!, compute_props_sorted! every output_freq steps.  The
  per-cell ntc! calls batch: one NTC launch per step.
* test/test_1D_couette_varweight.jl:78-105: ntc! per cell, and a merge and
  squash of that cell when it holds more than merge_threshold particles;
  the n_local read flushes per cell, so this loop runs cell by cell."""
import ctypes as C

import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def _setup(hip, mzo, case):
    D = hip.dropin
    st = helpers.store_from_case(hip, case)
    particles = [D.DeviceParticleVector(st, 1)]
    pia = D.DevicePIA(st)
    grid = hip.Grid1DUniform(case["L"], case["n_cells"])
    species_data = case["species"]
    boundaries = hip.MaxwellWalls1D(species_data, case["T"], case["T"], -case["v_wall"], case["v_wall"], 1.0, 1.0)
    return D, st, particles, pia, grid, species_data, boundaries


def test_dropin_couette_loop_vs_oracle(hip, mzo):
    case = helpers.couette_case(mzo, n_cells=30, ppc=100, seed=3, dt=2e-10, sgwm_mult=50.0)
    D, st, particles, pia, grid, species_data, boundaries = _setup(hip, mzo, case)
    rng = D.DeviceRNG(case["seed"])
    collision_factors = D.CollisionFactorsArray(st)   # initial values uploaded by store_from_case
    collision_data = D.CollisionData()
    interaction_data = [[case["it"]]]
    gridsorter = D.GridSortInPlace(grid, case["n_cells"] * case["ppc"])
    phys_props = D.PhysProps(pia)
    n_timesteps, output_freq = 24, 8
    Δt = case["dt"]
    st.timing(True)
    for t in range(1, n_timesteps + 1):
        for cell in range(1, grid.n_cells + 1):
            D.ntc(rng, collision_factors[1, 1, cell],
                  collision_data, interaction_data, particles[0], pia, cell, 1, Δt, grid.cells[cell].V)
        D.convect_particles(rng, grid, boundaries, particles[0], pia, 1, species_data, Δt)
        D.sort_particles(gridsorter, grid, particles[0], pia, 1)
        if t % output_freq == 0:
            D.compute_props_sorted(particles, pia, species_data, phys_props)
    # one NTC launch per step
    assert st.timing_get("ntc")[1] == n_timesteps
    st.timing(False)
    for step in range(1, n_timesteps + 1):
        helpers.oracle_couette_step(mzo, case, step)
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, case["pv"], case["pia"], rtol=1e-12)
    assert np.allclose(phys_props.T[0], helpers.oracle_props_sorted(mzo, case), rtol=1e-12, atol=0)
    # the reads of the stand-ins agree with the downloaded state
    assert pia.n_total[1] == got["n_total"] and pia.contiguous[1]
    for cell in (1, 17, 30):
        ind = pia.indexer[cell, 1]
        assert (ind.n_local, ind.start1, ind.end1) == tuple(got["pia"][cell - 1][:3])
    p5 = particles[0][5]
    assert np.array_equal(p5.v, got["v"][4]) and p5.w == got["w"][4]
    cf = st.collision_factors(1, 1)
    ref_sg = np.array([case["cfs"][c].sigma_g_w_max for c in range(case["n_cells"])])
    assert np.allclose(cf["sigma_g_w_max"], ref_sg, rtol=1e-12, atol=0)
    assert collision_factors[1, 1, 7].n_coll_performed == cf["n_coll_performed"][6]
    st.close()


def test_dropin_varweight_loop_vs_oracle(hip, mzo):
    merge_threshold, merge_target = 130, 40
    case = helpers.couette_case(mzo, n_cells=6, ppc=125, seed=12, dt=2e-10, sgwm_mult=40.0,
                                weights=lambda rng, n: rng.lognormal(0.0, 1.0, n) / np.exp(0.5))
    D, st, particles, pia, grid, species_data, boundaries = _setup(hip, mzo, case)
    rng = D.DeviceRNG(case["seed"])
    collision_factors = D.CollisionFactorsArray(st)
    collision_data = D.CollisionData()
    interaction_data = [[case["it"]]]
    gridsorter = D.GridSortInPlace(grid, 6 * 125)
    oc = D.OctreeN2Merge("OctreeBinMidSplit", init_bin_bounds="OctreeInitBinMinMaxVel", max_Nbins=6000)
    Δt = case["dt"]
    n_timesteps = 6
    merges = 0
    for t in range(1, n_timesteps + 1):
        for cell in range(1, grid.n_cells + 1):
            D.ntc(rng, collision_factors[1, 1, cell],
                  collision_data, interaction_data, particles[0], pia, cell, 1, Δt, grid.cells[cell].V)
            if pia.indexer[cell, 1].n_local > merge_threshold:
                D.merge_octree_N2_based(rng, oc, particles[0], pia, cell, 1, merge_target, grid)
                D.squash_pia(particles, pia)
                merges += 1
        D.convect_particles(rng, grid, boundaries, particles[0], pia, 1, species_data, Δt)
        D.sort_particles(gridsorter, grid, particles[0], pia, 1)
    assert merges > 0
    # the oracle, same sequence and tags
    pv, opia, g = case["pv"], case["pia"], case["grid"]
    og = mzo.grid1d(case["L"], case["n_cells"])
    octree = mzo.Octree(split=1, init_bin_bounds=1, bin_bounds_compute=1, max_Nbins=6000, max_depth=10)
    o_merges = 0
    for step in range(1, n_timesteps + 1):
        r_ntc = mzo.RNG(seed=case["seed"], tag=(step << 4) | 1)
        r_merge = mzo.RNG(seed=case["seed"], tag=(step << 4) | 3)
        for c in range(1, case["n_cells"] + 1):
            mzo.lib().mzo_ntc_single(r_ntc.ref, C.byref(case["cfs"][c - 1]), C.byref(case["it"]), pv.p, opia.p, c,
                                     1, Δt, g.dx, 1e-16, 0)
            if opia.ind(c).n_local > merge_threshold:
                mzo.lib().mzo_merge_octree_N2_based(r_merge.ref, octree.p, pv.p, opia.p, c, 1, merge_target,
                                                    C.byref(og))
                mzo.lib().mzo_squash_pia(pv.p, opia.p, 1)
                o_merges += 1
        r = mzo.RNG(seed=case["seed"], tag=(step << 4) | 2)
        mzo.lib().mzo_convect_particles(r.ref, C.byref(g), C.byref(case["walls"]), pv.p, opia.p, 1, Δt)
        mzo.lib().mzo_sort_particles(C.byref(g), pv.p, opia.p, 1)
    assert o_merges == merges
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, pv, opia, rtol=1e-12)
    st.close()


def test_dropin_per_cell_targets_vs_oracle(hip, mzo):
    """A loop that merges every cell above its own threshold down to its own
    target_np (merge_octree_N2_based!(rng, octree, pv, pia, cell, 1,
    targets[cell], grid) per cell, then one squash_pia!): the queued calls
    with different targets go to the device as one launch, and the state
    matches the oracle's per-cell calls."""
    nc = 6
    thresholds = [0] + [118 + 3 * c for c in range(1, nc + 1)]
    targets = [0] + [28 + 6 * c for c in range(1, nc + 1)]
    case = helpers.couette_case(mzo, n_cells=nc, ppc=125, seed=14, dt=2e-10, sgwm_mult=40.0,
                                weights=lambda rng, n: rng.lognormal(0.0, 1.0, n) / np.exp(0.5))
    D, st, particles, pia, grid, species_data, boundaries = _setup(hip, mzo, case)
    rng = D.DeviceRNG(case["seed"])
    collision_factors = D.CollisionFactorsArray(st)
    collision_data = D.CollisionData()
    interaction_data = [[case["it"]]]
    gridsorter = D.GridSortInPlace(grid, nc * 125)
    oc = D.OctreeN2Merge("OctreeBinMidSplit", init_bin_bounds="OctreeInitBinMinMaxVel", max_Nbins=6000)
    Δt = case["dt"]
    n_timesteps = 6
    merges = 0
    for t in range(1, n_timesteps + 1):
        for cell in range(1, grid.n_cells + 1):
            D.ntc(rng, collision_factors[1, 1, cell],
                  collision_data, interaction_data, particles[0], pia, cell, 1, Δt, grid.cells[cell].V)
        to_merge = [cell for cell in range(1, grid.n_cells + 1) if pia.indexer[cell, 1].n_local > thresholds[cell]]
        for cell in to_merge:
            D.merge_octree_N2_based(rng, oc, particles[0], pia, cell, 1, targets[cell], grid)
        merges += len(to_merge)
        D.squash_pia(particles, pia)
        D.convect_particles(rng, grid, boundaries, particles[0], pia, 1, species_data, Δt)
        D.sort_particles(gridsorter, grid, particles[0], pia, 1)
    assert merges > 2
    pv, opia, g = case["pv"], case["pia"], case["grid"]
    og = mzo.grid1d(case["L"], case["n_cells"])
    o_merges = 0
    for step in range(1, n_timesteps + 1):
        r_ntc = mzo.RNG(seed=case["seed"], tag=(step << 4) | 1)
        r_merge = mzo.RNG(seed=case["seed"], tag=(step << 4) | 3)
        for c in range(1, nc + 1):
            mzo.lib().mzo_ntc_single(r_ntc.ref, C.byref(case["cfs"][c - 1]), C.byref(case["it"]), pv.p, opia.p, c,
                                     1, Δt, g.dx, 1e-16, 0)
        for c in [c for c in range(1, nc + 1) if opia.ind(c).n_local > thresholds[c]]:
            octree = mzo.Octree(split=1, init_bin_bounds=1, bin_bounds_compute=1, max_Nbins=6000, max_depth=10)
            mzo.lib().mzo_merge_octree_N2_based(r_merge.ref, octree.p, pv.p, opia.p, c, 1, targets[c], C.byref(og))
            o_merges += 1
        mzo.lib().mzo_squash_pia(pv.p, opia.p, 1)
        r = mzo.RNG(seed=case["seed"], tag=(step << 4) | 2)
        mzo.lib().mzo_convect_particles(r.ref, C.byref(g), C.byref(case["walls"]), pv.p, opia.p, 1, Δt)
        mzo.lib().mzo_sort_particles(C.byref(g), pv.p, opia.p, 1)
    assert o_merges == merges
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, pv, opia, rtol=1e-12)
    st.close()
