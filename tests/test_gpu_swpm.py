"""SWPM collisions on the device (mz_swpm) against the oracle's swpm!: the
candidate counts, acceptances, appended group-2 positions, ids (free-id LIFO
with resize! growth), PIA and counters bit for bit; weights and velocities
to 1e-12; weight, momentum and energy conserved."""
import ctypes as C
import math

import numpy as np
import pytest

import helpers
from test_gpu_parity import _zero_d_case, _ntc_setup

pytestmark = pytest.mark.gpu

RTOL = 1e-12


@pytest.mark.parametrize("rng_mode,G", [(0, 1.0), (0, 0.0), (1, 2.5)])
def test_swpm_vs_oracle(hip, mzo, rng_mode, G):
    """24 cells x 200 particles of weights U(0.5, 2), about 60 candidates
    per cell; then a second call on the last cell (whose group 2 ends at the
    global end), over particles the first call created."""
    nc, npc = 24, 200
    w = np.random.default_rng(4).uniform(0.5, 2.0, nc * npc)
    pv, pia = _zero_d_case(mzo, nc, npc, seed=8, weights=w)
    it = mzo.interactions(["Ar"])[0][0]
    vs = np.sqrt(helpers.KB * 300.0 / helpers.MASS_AR)
    sgm = it.vhs_factor * (2 * vs) ** (1.0 - 2 * it.vhs_o) * 2 * vs
    st, it, _ = _ntc_setup(hip, mzo, pv, pia, nc, sgm, cap=nc * npc * 4)
    w_max = 2e12
    dt = 1.0
    V = 0.5 * dt * npc * (npc - 1) * sgm * w_max * (G + 1) / 60.0
    tag = (9 << 4) | 1
    if rng_mode == 1:
        table = np.random.default_rng(5).uniform(size=(nc, 8192))
        st.set_replay("ntc", table)
        r = mzo.RNG(table=table)
    else:
        r = mzo.RNG(seed=77, tag=tag)
    _, v0, w0, _ = pv.logical_arrays(nc * npc)
    hip.swpm(st, 1, G, dt, V=np.full(nc, V), tag=tag, rng_mode=rng_mode)
    cfs = (mzo.CFSwpm * nc)()
    for c in range(1, nc + 1):
        cfs[c - 1].sigma_g_max = sgm
        mzo.lib().mzo_swpm(r.ref, C.byref(cfs[c - 1]), C.byref(it), pv.p, pia.p, c, 1, G, dt, V)
    got = st.download_species(1)
    assert got["err"] == 0
    assert got["n_total"] > nc * npc + 100
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=RTOL)
    cf = st.collision_factors(1, 1)
    assert list(cf["n_coll"]) == [cfs[c].n_coll for c in range(nc)]
    assert list(cf["n_coll_performed"]) == [cfs[c].n_coll_performed for c in range(nc)]
    assert np.allclose(cf["sigma_g_w_max"], [cfs[c].sigma_g_max for c in range(nc)], rtol=1e-12, atol=0)
    n = got["n_total"]
    w1, v1 = got["w"][:n], got["v"][:n]
    assert abs(w1.sum() / w0.sum() - 1) < 1e-13
    p0, p1 = (w0[:, None] * v0).sum(0), (w1[:, None] * v1).sum(0)
    assert np.all(np.abs(p1 - p0) <= 1e-12 * (w0[:, None] * np.abs(v0)).sum(0))
    assert abs((w1 * (v1 ** 2).sum(1)).sum() / (w0 * (v0 ** 2).sum(1)).sum() - 1) < 1e-12
    tag2 = (10 << 4) | 1
    r2 = mzo.RNG(seed=77, tag=tag2) if rng_mode == 0 else mzo.RNG(table=table)
    hip.swpm(st, 1, G, dt, V=np.full(nc, V), tag=tag2, rng_mode=rng_mode, cells=(nc, nc))
    mzo.lib().mzo_swpm(r2.ref, C.byref(cfs[nc - 1]), C.byref(it), pv.p, pia.p, nc, 1, G, dt, V)
    helpers.assert_state_matches(mzo, st.download_species(1), pv, pia, rtol=RTOL)
    st.close()


def _bkw_cells(hip, mzo, nc, ppc, seed):
    """nc independent BKW boxes (V = 1, n = 1e23) of ppc particles each, one
    per cell of a 1-D grid, every particle at its cell's centre so that
    octree merging (which places new particles around the bin's mean
    position) and sorting keep each particle in its cell.  Returns the
    oracle state, the device store holding the same particles, and the
    pieces the steps need."""
    import analytic_cases as ac
    T0 = ac.BKW["T0"]
    its = mzo.interactions(["Ar"], table=mzo.PSEUDO_MAXWELL)
    it = its[0][0]
    pv, pia = mzo.ParticleVector(nc * ppc), mzo.PIA(nc, 1)
    r = mzo.RNG(seed=seed, tag=77)
    v0 = (C.c_double * 3)(0.0, 0.0, 0.0)
    for c in range(1, nc + 1):
        r.s.stream, r.s.k = 2 * c, 0
        mzo.lib().mzo_sample_particles_equal_weight(r.ref, pv.p, pia.p, c, 1, ppc, ac.M_AR, T0,
                                                    ac.BKW["ndens"] / ppc, c - 0.5, c - 0.5, 0.5, 0.5, 0.5, 0.5)
        r.s.stream, r.s.k = 2 * c + 1, 0
        mzo.lib().mzo_sample_bkw(r.ref, pv.p, ppc, (c - 1) * ppc, ac.M_AR, T0, v0)
    pia.set_contiguous(1, True)
    sgm = mzo.lib().mzo_estimate_sigma_g_w_max(C.byref(it), ac.M_AR, ac.M_AR, T0, T0, 1.0, 1.0)
    cfs = [mzo.CFSwpm() for _ in range(nc)]
    for cf in cfs:
        cf.sigma_g_max = sgm
    st = hip.DeviceStore(nc, 1, 3 * nc * ppc)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    st.set_grid(hip.Grid1DUniform(float(nc), nc))
    from test_gpu_analytic import _set_interactions
    _set_interactions(hip, st, its)
    st.set_sigma_g_w_max(1, 1, [sgm] * nc)
    st.set_seed(seed)
    # dt = 0.025 of the BKW reference time (test/test_bkw_varweight_octree_swpm.jl:55-90)
    d = mzo.PSEUDO_MAXWELL[("Ar", "Ar")][0]
    tref = 1.0 / (ac.BKW["ndens"] * np.pi * d * d) / np.sqrt(2 * helpers.KB * T0 / ac.M_AR)
    return dict(pv=pv, pia=pia, it=it, cfs=cfs, st=st, grid=mzo.grid1d(float(nc), nc),
                dt=ac.BKW["dt_scaled"] * tref, nc=nc, seed=seed)


def _device_step(hip, b, step, thr, target):
    """swpm! in every cell, octree N:2 merging of the cells above thr, then
    sorting, which puts the new particles back in their cells' runs."""
    hip.swpm(b["st"], 1, 1.0, b["dt"], V=[1.0] * b["nc"], tag=(step << 4) | 1)
    m = hip.merge_octree_N2_based(b["st"], 1, thr, target, max_Nbins=512, squash=True, tag=(step << 4) | 3)
    hip.sort_particles(b["st"], 1)
    return m


def _oracle_step(mzo, b, step, thr, target):
    pv, pia = b["pv"], b["pia"]
    r = mzo.RNG(seed=b["seed"], tag=(step << 4) | 1)
    for c in range(1, b["nc"] + 1):
        r.s.stream, r.s.k = c - 1, 0
        mzo.lib().mzo_swpm(r.ref, C.byref(b["cfs"][c - 1]), C.byref(b["it"]), pv.p, pia.p, c, 1, 1.0, b["dt"], 1.0)
    r = mzo.RNG(seed=b["seed"], tag=(step << 4) | 3)
    oc = mzo.Octree(split=1, init_bin_bounds=1, bin_bounds_compute=1, max_Nbins=512, max_depth=10)
    for c in range(1, b["nc"] + 1):
        if pia.ind(c).n_local > thr:
            r.s.stream, r.s.k = c - 1, 0
            mzo.lib().mzo_merge_octree_N2_based(r.ref, oc.p, pv.p, pia.p, c, 1, target, None)
    mzo.lib().mzo_squash_pia(pv.p, pia.p, 1)
    mzo.lib().mzo_sort_particles(C.byref(b["grid"]), pv.p, pia.p, 1)


def test_swpm_octree_sort_trajectory_vs_oracle(hip, mzo):
    """30 steps of the loop of test/test_bkw_varweight_octree_swpm.jl (swpm!
    with G = 1, octree N:2 merging above a threshold) over 16 BKW boxes of
    1,000 particles, merged above 1,250 to 1,000 (a bin budget of 512, the
    device octree's), with a sort after each step: the device state against
    the oracle's, run on the same draws, every 10 steps."""
    b = _bkw_cells(hip, mzo, 16, 1000, seed=21)
    merged = 0
    for step in range(1, 31):
        merged += _device_step(hip, b, step, 1250, 1000)
        _oracle_step(mzo, b, step, 1250, 1000)
        if step % 10 == 0:
            got = b["st"].download_species(1)
            assert got["err"] == 0
            helpers.assert_state_matches(mzo, got, b["pv"], b["pia"], rtol=1e-11)
    assert merged > 16
    b["st"].close()


def test_bkw_swpm_octree_on_device(hip, mzo):
    """The BKW relaxation with SWPM collisions (G = 1) and octree N:2
    merging (test/test_bkw_varweight_octree_swpm.jl), 500 steps on the
    device over 16 BKW boxes of 1,000 particles merged above 1,250 to 1,000
    (the reference: one box merged above 10,000 to 8,000; the device octree
    takes cells of at most 2,048 particles).  Kept exactly: every box's
    density and temperature (the reference asserts 1e-11 and 5e-4 K), and
    the particle count stays bounded.  The moments average over the boxes;
    octree merging of 1,000-particle cells perturbs the high moments far
    more than the reference's 8,000-particle cell (the oracle, run over 8
    seeds, strays up to 9%, 16% and 26% from the analytic 4th, 6th and 8th
    moments), so the bounds here are 15%, 30% and 45%: they catch a method
    that drifts away (a merge that does not keep the distribution sends the
    8th moment past 100%), not the fine shape."""
    import analytic_cases as ac
    b = _bkw_cells(hip, mzo, 16, 1000, seed=1234)
    st = b["st"]
    sp = [hip.Species("Ar", ac.M_AR)]
    pp = hip.PhysProps(16, 1, ac.BKW["moments"], Tref=ac.BKW["T0"])
    hip.compute_props_with_total_moments(st, sp, pp)
    T0, n0 = pp.T[0].copy(), pp.n[0].copy()
    mom = [pp.moments[0, :, :3].mean(0)]
    merged = 0
    for step in range(1, ac.BKW["steps"] + 1):
        merged += _device_step(hip, b, step, 1250, 1000)
        hip.compute_props_with_total_moments(st, sp, pp)
        mom.append(pp.moments[0, :, :3].mean(0))
        if step % 100 == 0:
            print(f"step {step}: np {pp.np[0].min():.0f}..{pp.np[0].max():.0f}", flush=True)
    assert st.scalars(1)["err"] == 0 and merged > 1000
    assert np.max(np.abs(pp.T[0] - T0)) < 1e-9 and np.max(np.abs(pp.n[0] / n0 - 1)) < 1e-12
    assert np.all(pp.np[0] <= 1500)
    mom = np.array(mom)
    t = np.arange(ac.BKW["steps"] + 1) * ac.BKW["dt_scaled"]
    d = mzo.PSEUDO_MAXWELL[("Ar", "Ar")][0]
    it = b["it"]
    kappa = np.pi * d * d * (it.m_r / (2 * helpers.KB * ac.BKW["T0"])) ** (-0.5) / math.gamma(1.5)
    tref = 1.0 / (ac.BKW["ndens"] * np.pi * d * d) / np.sqrt(2 * helpers.KB * ac.BKW["T0"] / ac.M_AR)
    magic = tref * (4 * np.pi * ac.BKW["ndens"] * kappa) / (4 * np.pi)
    for j, (p, tol) in enumerate(((4, 0.15), (6, 0.30), (8, 0.45))):
        a = ac.bkw_analytic(t, magic, p)
        assert np.max(np.abs(a - mom[:, j]) / a) < tol, (p, np.max(np.abs(a - mom[:, j]) / a))
    st.close()
