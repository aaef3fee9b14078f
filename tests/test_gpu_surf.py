"""Convection with surface properties on the device (mz_convect_surf and the
SurfProps set operations) against the CPU oracle.

Particle state after the step is held to the usual bar (the following sort
bit for bit, x and v within 1e-12).  Surface sums are taken in a different
order than the reference's particle loop, so they are held to 1e-12 of the
sum of the magnitudes they are made of; np is an exact count.
"""
import ctypes as C

import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _walls(mzo, hip, case, acc):
    T, vw = case["T"], case["v_wall"]
    case["walls"] = mzo.walls([helpers.MASS_AR], T, T + 50.0, -vw, vw, acc, acc)
    return hip.MaxwellWalls1D(case["species"], T, T + 50.0, -vw, vw, acc, acc)


def _scales(v_before, v_after, w, f):
    """Per-quantity magnitude sums over every particle (a bound on the sum of
    the magnitudes of the wall terms)."""
    nb, na = np.abs(v_before).sum(axis=1), np.abs(v_after).sum(axis=1)
    mom = f * np.sum(w * (nb + na))
    ke = f * 0.5 * np.sum(w * (np.sum(v_before ** 2, axis=1) + np.sum(v_after ** 2, axis=1)))
    flux = f * np.sum(w)
    return np.array([0.0, flux, flux] + [mom] * 7 + [ke])


def _compare(got, ref, scale):
    assert np.array_equal(got[:, :, 0], ref[:, :, 0]), (got[:, :, 0], ref[:, :, 0])
    err = np.abs(got - ref)
    bound = RTOL * scale[None, None, :]
    assert np.all(err[:, :, 1:] <= bound[:, :, 1:]), (err / np.maximum(scale, 1e-300)).max()


@pytest.mark.parametrize("acc", [1.0, 0.5, 0.0])
def test_convect_surf_vs_oracle(hip, mzo, acc):
    """Couette walls (diffuse, half-diffuse, specular), 3,200 particles, a
    step long enough that a few hundred reach the walls."""
    case = helpers.couette_case(mzo, n_cells=16, ppc=200, seed=11, dt=2e-9)
    st = helpers.store_from_case(hip, case)
    st.set_walls(_walls(mzo, hip, case, acc))
    n = case["pia"].n_total()
    _, v0, w0, _ = case["pv"].logical_arrays(n)
    tag = (5 << 4) | 2
    sp = hip.SurfProps.grid1d(1)
    hip.convect_particles(st, 1, case["dt"], tag=tag, species_data=case["species"], surf_props=sp)
    ref = mzo.Surf.grid1d(1)
    mzo.convect_particles_surf(mzo.RNG(seed=case["seed"], tag=tag), case["grid"], case["walls"], case["pv"],
                               case["pia"], 1, case["dt"], ref, helpers.MASS_AR)
    _, v1, _, _ = case["pv"].logical_arrays(n)
    assert ref.np.sum() > 40
    _compare(sp._to_device(), ref.q, _scales(v0, v1, w0, helpers.MASS_AR / case["dt"]))
    hip.sort_particles(st, 1)
    mzo.lib().mzo_sort_particles(C.byref(case["grid"]), case["pv"].p, case["pia"].p, 1)
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, case["pv"], case["pia"], rtol=RTOL)
    # the same inputs give the same sums on every run
    st2 = helpers.store_from_case(hip, helpers.couette_case(mzo, n_cells=16, ppc=200, seed=11, dt=2e-9))
    st2.set_walls(_walls(mzo, hip, case, acc))
    sp2 = hip.SurfProps.grid1d(1)
    hip.convect_particles(st2, 1, case["dt"], tag=tag, species_data=case["species"], surf_props=sp2)
    assert np.array_equal(sp2._to_device(), sp._to_device())
    st.close()
    st2.close()


def test_convect_surf_replay_and_areas(hip, mzo):
    """Replayed draws, and element areas 2 and 4 (inv_areas 0.5, 0.25)."""
    case = helpers.couette_case(mzo, n_cells=8, ppc=300, seed=3, dt=3e-9)
    st = helpers.store_from_case(hip, case)
    n = case["pia"].n_total()
    _, v0, w0, _ = case["pv"].logical_arrays(n)
    table = np.random.default_rng(2).uniform(size=(case["pv"].length, 12))
    st.set_replay("convect", table)
    sp = hip.SurfProps(2, 1, [2.0, 4.0], np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]]).T)
    hip.convect_particles(st, 1, case["dt"], tag=7, rng_mode=1, species_data=case["species"], surf_props=sp)
    ref = mzo.Surf(2, 1, [2.0, 4.0], [[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])
    mzo.convect_particles_surf(mzo.RNG(table=table), case["grid"], case["walls"], case["pv"], case["pia"], 1,
                               case["dt"], ref, helpers.MASS_AR)
    _, v1, _, _ = case["pv"].logical_arrays(n)
    assert ref.np.sum() > 20
    _compare(sp._to_device(), ref.q, _scales(v0, v1, w0, helpers.MASS_AR / case["dt"]))
    st.close()


def _kat_store(hip, mzo, deleted):
    """test/test_surface_props_1D_uniform.jl:312-471 on the device: L = 4, 8
    cells, specular walls, 4 particles at x = 0.1 moving at -1 and 4 at 3.9
    of which one moves at +1; optionally two of cell 1 deleted (PIA not
    contiguous)."""
    from test_oracle_surf import _two_cell_case
    pv, pia = _two_cell_case(mzo)
    for _ in range(deleted):
        mzo.lib().mzo_delete_particle_end(pv.p, pia.p, 1, 1)
    if deleted:
        pia.set_contiguous(1, False)
    st = hip.DeviceStore(8, 1, 64)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    sp = [hip.Species("Ar", helpers.MASS_AR)]
    st.set_species_data(sp)
    st.set_grid(hip.Grid1DUniform(4.0, 8))
    st.set_walls(hip.MaxwellWalls1D(sp, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0))
    st.set_seed(1234)
    return st, sp


@pytest.mark.parametrize("deleted", [0, 2])
def test_convect_surf_reference_kat(hip, mzo, deleted):
    st, sp = _kat_store(hip, mzo, deleted)
    surf = hip.SurfProps.grid1d(1)
    hip.convect_particles(st, 1, 0.2, species_data=sp, surf_props=surf)
    f = helpers.MASS_AR / 0.2
    eps = np.finfo(float).eps
    n_left = 4 - deleted
    assert list(surf.np[:, 0]) == [float(n_left), 1.0]
    assert abs(surf.flux_incident[0, 0] / f - n_left) < 3 * eps
    assert abs(surf.flux_incident[1, 0] / f - 3.0) < 3 * eps
    assert abs(surf.flux_reflected[0, 0] / f + n_left) < 3 * eps
    assert abs(surf.flux_reflected[1, 0] / f + 3.0) < 3 * eps
    assert np.all(surf.kinetic_energy_flux == 0.0)
    st.close()


def test_surf_sets_on_device(hip, mzo):
    """avg_props! and clear_props! on the store's sets:
    test/test_surface_props_1D_uniform.jl:200-300's averaging answers, and
    convect_particles! clearing every species of the instantaneous set."""
    st = hip.DeviceStore(4, 2, 64)
    a = hip.SurfProps.grid1d(2)
    a.np[:, 0] = [20, 10]
    a.flux_incident[0, 0] = 6
    a.force[:, 0, 0] = [1, 3, -8]
    a.kinetic_energy_flux[:, 1] = [3, 7]
    a.push(st, 0)
    st.surf_clear(1)
    st.surf_avg(2)
    b = hip.SurfProps.grid1d(2)
    b.np[:, 0] = [10, 20]
    b.flux_incident[0, 0] = 5
    b.force[:, 0, 0] = [2, 4, 8]
    b.kinetic_energy_flux[:, 1] = [2, 6]
    b.push(st, 0)
    st.surf_avg(2)
    avg = hip.SurfProps.grid1d(2).fetch(st, 1)
    assert list(avg.np[:, 0]) == [15.0, 15.0] and avg.flux_incident[0, 0] == 5.5
    assert list(avg.force[:, 0, 0]) == [1.5, 3.5, 0.0] and list(avg.kinetic_energy_flux[:, 1]) == [2.5, 6.5]
    st.surf_clear(1)
    assert not hip.SurfProps.grid1d(2).fetch(st, 1)._to_device().any()
    # convection of species 2 clears species 1's instantaneous values
    rng = np.random.default_rng(4)
    n = 64
    x = np.column_stack([rng.uniform(0, 1e-5, n), np.zeros(n), np.zeros(n)])
    v = rng.normal(0, 500, (n, 3))
    pia = np.zeros((4, 7), dtype=np.int64)
    pia[:, 2] = pia[:, 5] = -1
    pia[0] = (n, 1, n, n, 0, -1, 0)
    st.upload_species(2, x, v, np.ones(n), np.arange(1, n + 1), pia, n, True)
    sp = [hip.Species("Ar", helpers.MASS_AR), hip.Species("He", 6.65e-27)]
    st.set_species_data(sp)
    st.set_grid(hip.Grid1DUniform(1e-5, 4))
    st.set_walls(hip.MaxwellWalls1D(sp, 300.0, 300.0, 0.0, 0.0, 1.0, 1.0))
    a.push(st, 0)
    got = hip.convect_particles(st, 2, 1e-7, species_data=sp, surf_props=hip.SurfProps.grid1d(2))
    assert not got.np[:, 0].any() and not got.force[:, :, 0].any()
    assert got.np[:, 1].sum() > 0
    st.close()


def test_equilibrium_wall_pressure_device(hip):
    """200,000 particles, 500 steps of device convection between diffuse
    walls at the gas temperature, averaged on the device (avg_props!): the
    wall pressure, incident mass flux and exchanges match kinetic theory
    (about 500,000 contacts per wall; 1 % is about ten standard deviations)."""
    from test_oracle_surf import _equilibrium, check_equilibrium, equilibrium_expectations
    n, L, T, dt, steps = 200000, 1e-3, 300.0, 5e-8, 500
    w = 1e23 * L / n
    x, v = _equilibrium(n, L, T, 33)
    st = hip.DeviceStore(10, 1, n + 1024)
    pia = np.zeros((10, 7), dtype=np.int64)
    pia[:, 2] = pia[:, 5] = -1
    pia[0] = (n, 1, n, n, 0, -1, 0)
    xs = np.column_stack([x, np.zeros(n), np.zeros(n)])
    st.upload_species(1, xs, v, np.full(n, w), np.arange(1, n + 1), pia, n, True)
    sp = [hip.Species("Ar", helpers.MASS_AR)]
    st.set_species_data(sp)
    st.set_grid(hip.Grid1DUniform(L, 10))
    st.set_walls(hip.MaxwellWalls1D(sp, T, T, 0.0, 0.0, 1.0, 1.0))
    st.set_seed(17)
    surf = hip.SurfProps.grid1d(1)
    st.surf_clear(1)
    for t in range(steps):
        hip.convect_particles(st, 1, dt, tag=(t << 4) | 2, species_data=sp, surf_props=surf, fetch=False)
        st.surf_avg(steps)
    avg = hip.SurfProps.grid1d(1).fetch(st, 1)._to_device()[0]
    assert st.scalars(1)["err"] == 0
    check_equilibrium(avg, equilibrium_expectations(n, L, T, w), 0.01)
    st.close()


def _oracle_avg_run(mzo, case, steps, n_avg, step0=1):
    """The averaging phase on the oracle: returns the averaged wall
    properties, the averaged props and the per-quantity magnitude scales."""
    pv, pia, g = case["pv"], case["pia"], case["grid"]
    inst, avg = mzo.Surf.grid1d(1), mzo.Surf.grid1d(1)
    pp, pavg = mzo.Props(case["n_cells"], 1), mzo.Props(case["n_cells"], 1)
    n = pia.n_total()
    scale = np.zeros(11)
    f = helpers.MASS_AR / case["dt"]
    for step in range(step0, step0 + steps):
        r = mzo.RNG(seed=case["seed"], tag=(step << 4) | 1)
        for c in range(1, case["n_cells"] + 1):
            mzo.lib().mzo_ntc_single(r.ref, C.byref(case["cfs"][c - 1]), C.byref(case["it"]), pv.p, pia.p, c, 1,
                                     case["dt"], g.dx, 1e-16, 1)
        _, vb, wb, _ = pv.logical_arrays(n)
        mzo.convect_particles_surf(mzo.RNG(seed=case["seed"], tag=(step << 4) | 2), g, case["walls"], pv, pia, 1,
                                   case["dt"], inst, helpers.MASS_AR)
        _, va, _, _ = pv.logical_arrays(n)
        scale += _scales(vb, va, wb, f) / n_avg
        avg.avg_from(inst, n_avg)
        mzo.lib().mzo_sort_particles(C.byref(g), pv.p, pia.p, 1)
        mzo.compute_props_sorted([pv], pia, [helpers.MASS_AR], pp)
        pp.lpa[0] = pv.length
        mzo.lib().mzo_props_avg(pavg.ref, pp.ref, float(n_avg))
    return avg, pavg, scale


@pytest.mark.parametrize("mode,shape", [("slab", (0, 0)), ("slab", (1, 1)), ("slab", (3, 2))])
def test_slab_avg_steps_vs_oracle(hip, mzo, mode, shape):
    """Thirty steps of the averaging phase on the slab kernel (CFL 0.17)
    against the oracle, held to the same bars as the step-graph path."""
    steps = 30
    case = helpers.couette_case(mzo, n_cells=20, ppc=100, seed=3, dt=2e-10, sgwm_mult=50.0)
    st = helpers.store_from_case(hip, case)
    st.set_slab_shape(*shape)
    st.surf_clear(1)
    st.props_avg_clear()
    path = hip.couette_avg_steps(st, steps, 1, case["dt"], steps, species_data=case["species"], mode=mode)
    assert path == "slab", st.slab_info()
    avg, pavg, scale = _oracle_avg_run(mzo, case, steps, steps)
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, case["pv"], case["pia"], rtol=1e-11)
    gp = hip.PhysProps(case["n_cells"], 1).fetch_avg(st)
    assert np.array_equal(gp.np, pavg.np) and gp.lpa[0] == pavg.lpa[0]
    for name in ("n", "v", "T"):
        assert np.allclose(getattr(gp, name), getattr(pavg, name), rtol=1e-11, atol=0), name
    assert avg.np.sum() * steps > 50   # contacts over the run (np is averaged per step)
    _compare(hip.SurfProps.grid1d(1).fetch(st, 1)._to_device(), avg.q, scale)
    st.close()


def test_slab_avg_equals_graph(hip, mzo):
    """500 cells x 200 particles, 1,000 steps of the averaging phase: slab
    kernel and step graph leave the same particles and the same averaged
    props bit for bit, and wall averages equal to rounding."""
    steps = 1000
    case = helpers.couette_case(mzo, n_cells=500, ppc=200, seed=21, dt=2e-10)
    a = helpers.store_from_case(hip, case)
    b = helpers.store_from_case(hip, case)
    for st in (a, b):
        st.surf_clear(1)
        st.props_avg_clear()
    assert hip.couette_avg_steps(a, steps, 7, case["dt"], steps, species_data=case["species"], mode="graph") == "graph"
    path = hip.couette_avg_steps(b, steps, 7, case["dt"], steps, species_data=case["species"], mode="slab")
    assert path == "slab", b.slab_info()
    da, db = a.download_species(1), b.download_species(1)
    n = da["n_total"]
    assert da["err"] == 0 and db["err"] == 0
    assert np.array_equal(da["pia"], db["pia"]) and np.array_equal(da["index"][:n], db["index"][:n])
    for k in ("x", "v", "w"):
        assert np.array_equal(da[k][:n], db[k][:n]), k
    pa = hip.PhysProps(500, 1).fetch_avg(a)
    pb = hip.PhysProps(500, 1).fetch_avg(b)
    for k in ("lpa", "np", "n", "v", "T"):
        assert np.array_equal(getattr(pa, k), getattr(pb, k)), k
    sa = hip.SurfProps.grid1d(1).fetch(a, 1)._to_device()[0]
    sb = hip.SurfProps.grid1d(1).fetch(b, 1)._to_device()[0]
    assert np.array_equal(sa[:, 0], sb[:, 0]) and sa[:, 0].min() * steps > 1000
    for e in range(2):
        mom = np.abs(sa[e, 3:10]).sum()
        bound = np.array([0.0, sa[e, 1], sa[e, 1]] + [mom] * 7 + [np.abs(sa[e, 10]) + sa[e, 6] * 1000.0])
        assert np.all(np.abs(sa[e] - sb[e]) <= 1e-10 * bound), (e, sa[e], sb[e])
    ia = hip.SurfProps.grid1d(1).fetch(a, 0)._to_device()[0]
    ib = hip.SurfProps.grid1d(1).fetch(b, 0)._to_device()[0]
    assert np.array_equal(ia[:, 0], ib[:, 0])
    a.close()
    b.close()


@pytest.mark.parametrize("mode", ["eager", "graph"])
def test_couette_avg_steps_vs_oracle(hip, mzo, mode):
    """Twelve steps of the averaging phase of
    simulations/1D/couette_with_surface_quantities.jl (ntc!, convection with
    wall properties, their time average, sort, compute_props_sorted!, the
    props' time average), kernel by kernel or as a replayed hipGraph, against
    the oracle: particle state as in the plain Couette parity test, averaged
    np bit for bit, averaged n, v, T to 1e-11, averaged wall properties to
    1e-12 of their magnitude sums."""
    steps, n_avg = 12, 12
    case = helpers.couette_case(mzo, n_cells=20, ppc=100, seed=3, dt=2e-9, sgwm_mult=50.0)
    st = helpers.store_from_case(hip, case)
    st.surf_clear(1)
    st.props_avg_clear()
    path = hip.couette_avg_steps(st, steps, 1, case["dt"], n_avg, species_data=case["species"], mode=mode)
    assert path == mode
    pv, pia, g = case["pv"], case["pia"], case["grid"]
    inst, avg = mzo.Surf.grid1d(1), mzo.Surf.grid1d(1)
    pp, pavg = mzo.Props(case["n_cells"], 1), mzo.Props(case["n_cells"], 1)
    n = pia.n_total()
    mom, ke, flux = 0.0, 0.0, 0.0
    f = helpers.MASS_AR / case["dt"]
    for step in range(1, steps + 1):
        r = mzo.RNG(seed=case["seed"], tag=(step << 4) | 1)
        for c in range(1, case["n_cells"] + 1):
            mzo.lib().mzo_ntc_single(r.ref, C.byref(case["cfs"][c - 1]), C.byref(case["it"]), pv.p, pia.p, c, 1,
                                     case["dt"], g.dx, 1e-16, 1)
        _, vb, wb, _ = pv.logical_arrays(n)
        mzo.convect_particles_surf(mzo.RNG(seed=case["seed"], tag=(step << 4) | 2), g, case["walls"], pv, pia, 1,
                                   case["dt"], inst, helpers.MASS_AR)
        _, va, _, _ = pv.logical_arrays(n)
        s = _scales(vb, va, wb, f) / n_avg
        flux, mom, ke = flux + s[1], mom + s[3], ke + s[10]
        avg.avg_from(inst, n_avg)
        mzo.lib().mzo_sort_particles(C.byref(g), pv.p, pia.p, 1)
        mzo.compute_props_sorted([pv], pia, [helpers.MASS_AR], pp)
        pp.lpa[0] = pv.length
        mzo.lib().mzo_props_avg(pavg.ref, pp.ref, float(n_avg))
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=1e-11)
    gp = hip.PhysProps(case["n_cells"], 1).fetch_avg(st)
    assert np.array_equal(gp.np, pavg.np)
    assert gp.lpa[0] == pavg.lpa[0]
    for name in ("n", "v", "T"):
        a, b = getattr(gp, name), getattr(pavg, name)
        assert np.allclose(a, b, rtol=1e-11, atol=0), name
    sa = hip.SurfProps.grid1d(1).fetch(st, 1)._to_device()
    assert avg.np.sum() > 0
    _compare(sa, avg.q, np.array([0.0, flux, flux] + [mom] * 7 + [ke]))
    st.close()
