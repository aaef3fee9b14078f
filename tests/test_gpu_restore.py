"""restore_particle_ordering! on the device against the oracle: the
reference's known answers (test/test_particle_index_sorting.jl), a state
after weight splits, octree merges and squash, and the Couette loop with a
restore every 10 steps (simulations/1D/couette_benchmarking.jl:57-85)."""
import ctypes as C

import numpy as np
import pytest

import helpers
from test_oracle_restore import CASES, reference_case_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", CASES)
def test_restore_reference_cases_on_device(hip, mzo, case):
    pv = reference_case_state(mzo, case)
    n_used = case[1]
    pia = mzo.PIA(1, 1)
    if n_used:
        pia.set(1, 1, n_used, 1, n_used, n_used)
    pia.set_n_total(1, n_used)
    st = hip.DeviceStore(1, 1, 16)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    hip.restore_particle_ordering(st, 1)
    mzo.lib().mzo_restore_particle_ordering(pv.p)
    got = st.download_species(1)
    assert got["nbuffer"] == case[3] and got["length"] == 10
    assert list(got["buffer"][:case[3]]) == case[4]
    assert list(got["index"][:n_used]) == list(range(1, n_used + 1))
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=0.0)
    st.close()


def test_restore_after_splits_and_merges(hip, mzo):
    """Lognormal weights: NTC splits grow the cells, octree merges above 150
    particles free ids in cell order, squash, sort; then restore on both."""
    case = helpers.couette_case(mzo, n_cells=10, ppc=140, seed=44, dt=2e-10, sgwm_mult=40.0,
                                weights=lambda rng, n: rng.lognormal(0.0, 1.0, n) / np.exp(0.5))
    st = helpers.store_from_case(hip, case)
    pv, pia, g = case["pv"], case["pia"], case["grid"]
    og = mzo.grid1d(case["L"], case["n_cells"])
    octree = mzo.Octree(split=1, init_bin_bounds=1, bin_bounds_compute=1, max_Nbins=4096, max_depth=8)
    for step in range(1, 5):
        hip.ntc(st, 1, 1, case["dt"], V=g.dx, tag=(step << 4) | 1)
        hip.merge_octree_N2_based(st, 1, threshold=150, target_np=60, max_depth=8, grid_clamp=True, squash=True,
                                  tag=(step << 4) | 3)
        hip.sort_particles(st, 1)
        r = mzo.RNG(seed=case["seed"], tag=(step << 4) | 1)
        for c in range(1, case["n_cells"] + 1):
            mzo.lib().mzo_ntc_single(r.ref, C.byref(case["cfs"][c - 1]), C.byref(case["it"]), pv.p, pia.p, c, 1,
                                     case["dt"], g.dx, 1e-16, 0)
        r = mzo.RNG(seed=case["seed"], tag=(step << 4) | 3)
        for c in range(1, case["n_cells"] + 1):
            if pia.ind(c).n_local > 150:
                mzo.lib().mzo_merge_octree_N2_based(r.ref, octree.p, pv.p, pia.p, c, 1, 60, C.byref(og))
        mzo.lib().mzo_squash_pia(pv.p, pia.p, 1)
        mzo.lib().mzo_sort_particles(C.byref(g), pv.p, pia.p, 1)
    before = st.download_species(1)
    helpers.assert_state_matches(mzo, before, pv, pia, rtol=1e-12)
    n = before["n_total"]
    assert not np.array_equal(before["index"][:n], np.arange(1, n + 1)), "ids already in order: nothing to test"
    assert before["nbuffer"] > 0
    hip.restore_particle_ordering(st, 1)
    mzo.lib().mzo_restore_particle_ordering(pv.p)
    got = st.download_species(1)
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=1e-12)
    assert np.array_equal(got["index"][:n], np.arange(1, n + 1))
    for k in ("x", "v", "w"):
        assert np.array_equal(got[k][:n], before[k][:n]), k
    st.close()


def test_restore_cadence_couette_loop(hip, mzo):
    """The benchmarking loop through the drop-in interface: ntc_equal_weight!
    per cell, convection, sort, restore_particle_ordering! every 10 steps;
    the identities feed the convection's random streams, so any difference
    in the restore shows in the particles."""
    D = hip.dropin
    case = helpers.couette_case(mzo, n_cells=20, ppc=60, seed=9, dt=2e-10, sgwm_mult=50.0)
    st = helpers.store_from_case(hip, case)
    particles = [D.DeviceParticleVector(st, 1)]
    pia = D.DevicePIA(st)
    grid = hip.Grid1DUniform(case["L"], case["n_cells"])
    boundaries = hip.MaxwellWalls1D(case["species"], case["T"], case["T"], -case["v_wall"], case["v_wall"], 1.0, 1.0)
    rng = D.DeviceRNG(case["seed"])
    cf, cd, gridsorter = D.CollisionFactorsArray(st), D.CollisionData(), D.GridSortInPlace(grid)
    index_inv_map = np.zeros(case["n_cells"] * case["ppc"], dtype=np.int64)
    n_timesteps = 30
    for t in range(1, n_timesteps + 1):
        for cell in range(1, grid.n_cells + 1):
            D.ntc_equal_weight(rng, cf[1, 1, cell], cd, None, particles[0], pia, cell, 1, case["dt"],
                               grid.cells[cell].V)
        D.convect_particles(rng, grid, boundaries, particles[0], pia, 1, case["species"], case["dt"])
        D.sort_particles(gridsorter, grid, particles[0], pia, 1)
        if t % 10 == 0:
            D.restore_particle_ordering(particles[0], index_inv_map)
    for step in range(1, n_timesteps + 1):
        helpers.oracle_couette_step(mzo, case, step, equal_weight=True)
        if step % 10 == 0:
            mzo.lib().mzo_restore_particle_ordering(case["pv"].p)
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, case["pv"], case["pia"], rtol=1e-12)
    n = got["n_total"]
    assert np.array_equal(got["index"][:n], np.arange(1, n + 1))
    st.close()
