"""GPU parity: every operation of the HIP path against the CPU oracle on the
same inputs (and, where the reference's tests give them, its expected values).

Integer state (index arrays, PIA, LIFO buffer, counters) must match bit for
bit; velocities and positions within 1e-12 relative (sin/cos/log/pow may
differ by an ulp between the device and host math libraries)."""
import ctypes as C

import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def _pv_from(mzo, parts):
    pv = mzo.ParticleVector(len(parts))
    for i, (w, v, x) in enumerate(parts, start=1):
        pv.pop_to(i)
        pv.set(i, w, v, x)
    return pv


def _store(hip, mzo, pv, pia, n_cells, cap=None):
    st = hip.DeviceStore(n_cells, 1, cap or max(64, 2 * pv.length))
    helpers.upload_from_oracle(st, mzo, pv, pia)
    return st


# ------------------------------------------------------------------ sorting
def test_sort_reference_kats(hip, mzo):
    """test/test_grid_sorting.jl:26-67, 98-125, 145-178 through the device."""
    pv = mzo.ParticleVector(8)
    pia = mzo.PIA(2, 1)
    for i in range(1, 9):
        pv.pop_to(i)
        pv.set(i, 1.0, [0, 0, 0], [0.99 * 8.0 * (9.0 - i) / 8, 0.0, 0.0])
    pia.set(1, 1, 4, 1, 4, 4)
    pia.set(2, 1, 4, 5, 8, 4)
    pia.set_n_total(1, 8)
    st = _store(hip, mzo, pv, pia, 2)
    st.set_grid(hip.Grid1DUniform(8.0, 2))
    hip.sort_particles(st, 1)
    got = st.download_species(1)
    assert list(got["index"]) == [5, 6, 7, 8, 1, 2, 3, 4]
    assert got["pia"].tolist() == [[4, 1, 4, 4, 0, -1, 0], [4, 5, 8, 4, 0, -1, 0]]
    assert np.array_equal(got["x"][:4, 0], [0.99 * 8.0 * (9.0 - i - 4) / 8 for i in range(1, 5)])

    # 4 cells, uneven [3, 1, 0, 4]
    pv2 = mzo.ParticleVector(8)
    pia2 = mzo.PIA(4, 1)
    xs = [6.75] * 4 + [2.5] + [0.5] * 3
    for i in range(1, 9):
        pv2.pop_to(i)
        pv2.set(i, 1.0, [0, 0, 0], [xs[i - 1], 0.0, 0.0])
    pia2.set(1, 1, 8, 1, 8, 8)
    pia2.set_n_total(1, 8)
    st2 = _store(hip, mzo, pv2, pia2, 4)
    st2.set_grid(hip.Grid1DUniform(8.0, 4))
    hip.sort_particles(st2, 1)
    got = st2.download_species(1)
    assert list(got["index"]) == [6, 7, 8, 5, 1, 2, 3, 4]
    assert got["pia"].tolist() == [[3, 1, 3, 3, 0, -1, 0], [1, 4, 4, 1, 0, -1, 0],
                                   [0, 0, -1, 0, 0, -1, 0], [4, 5, 8, 4, 0, -1, 0]]
    assert list(got["cell"]) == [4, 4, 4, 4, 2, 1, 1, 1]     # pv.cell is not permuted
    # precomputed-cell variant (test_grid_sorting.jl:198-226): positions play
    # no role, the caller's pv.cell decides
    up = helpers.oracle_upload_arrays(mzo, pv2, pia2)
    up["x"][:, 0] = [0.05] * 4 + [0.1] + [0.3] * 3
    st2.upload_species(1, up["x"], up["v"], up["w"], np.arange(1, 9), [[8, 1, 8, 8, 0, -1, 0]] + [[0, 0, -1, 0, 0, -1, 0]] * 3,
                       8, True, up["buffer"], up["length"])
    st2.upload_cell_keys(1, [4, 4, 4, 4, 2, 1, 1, 1])
    hip.sort_particles(st2, 1, compute_cells=False)
    got = st2.download_species(1)
    assert list(got["index"]) == [6, 7, 8, 5, 1, 2, 3, 4]
    assert got["pia"].tolist() == [[3, 1, 3, 3, 0, -1, 0], [1, 4, 4, 1, 0, -1, 0],
                                   [0, 0, -1, 0, 0, -1, 0], [4, 5, 8, 4, 0, -1, 0]]
    st.close(); st2.close()


@pytest.mark.parametrize("n,nc,seed", [(5000, 37, 1), (100000, 500, 2), (1, 3, 3), (3000, 1, 4)])
def test_sort_random_vs_oracle(hip, mzo, n, nc, seed):
    """Random positions in random logical order: index, PIA and keys bit-exact
    against the oracle's counting sort (and so against numpy's stable argsort)."""
    rng = np.random.default_rng(seed)
    pv = mzo.ParticleVector(n)
    pia = mzo.PIA(nc, 1)
    for i in range(1, n + 1):
        pv.pop_to(i)
    pv.index[:n] = rng.permutation(n) + 1
    pv.x[:n, 0] = rng.uniform(0, 1.0, n)
    pv.x[:n, 1:] = rng.normal(size=(n, 2))
    pv.v[:n] = rng.normal(size=(n, 3))
    pv.w[:n] = rng.uniform(0.5, 2, n)
    pia.set(1, 1, n, 1, n, n)
    pia.set_n_total(1, n)
    grid = mzo.grid1d(1.0, nc)
    st = _store(hip, mzo, pv, pia, nc, cap=n + 2048)
    st.set_grid(hip.Grid1DUniform(1.0, nc))
    hip.sort_particles(st, 1)
    mzo.lib().mzo_sort_particles(C.byref(grid), pv.p, pia.p, 1)
    got = st.download_species(1)
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=0.0)
    assert np.array_equal(got["cell"], pv.cell[:n])
    st.close()


def test_squash_then_sort_vs_oracle(hip, mzo):
    """A non-contiguous PIA with holes (after deletions) is squashed before
    the sort, as sort_particles! does (grid_sorting.jl:66-68)."""
    rng = np.random.default_rng(7)
    nc, n = 6, 600
    pv = mzo.ParticleVector(n)
    pia = mzo.PIA(nc, 1)
    for i in range(1, n + 1):
        pv.pop_to(i)
        pv.set(i, 1.0 + i, rng.normal(size=3), [rng.uniform(0, 6.0), 0.0, 0.0])
    for c in range(1, nc + 1):
        pia.set(c, 1, 100, 100 * (c - 1) + 1, 100 * c, 100)
    pia.set_n_total(1, n)
    # delete from the end of some cells (leaves holes), as merges do
    for c, k in ((2, 10), (4, 37), (6, 5)):
        for _ in range(k):
            mzo.lib().mzo_delete_particle_end(pv.p, pia.p, c, 1)
    pia.set_contiguous(1, False)
    st = _store(hip, mzo, pv, pia, nc)
    hip.squash_pia(st, 1)
    mzo.lib().mzo_squash_pia(pv.p, pia.p, 1)
    helpers.assert_state_matches(mzo, st.download_species(1), pv, pia, rtol=0.0)
    st.set_grid(hip.Grid1DUniform(6.0, nc))
    hip.sort_particles(st, 1)
    mzo.lib().mzo_sort_particles(C.byref(mzo.grid1d(6.0, nc)), pv.p, pia.p, 1)
    helpers.assert_state_matches(mzo, st.download_species(1), pv, pia, rtol=0.0)
    st.close()


# --------------------------------------------------------------- convection
def test_convection_reference_kat(hip, mzo):
    """test/test_convection_1D.jl:21-52 (specular walls) through the device."""
    parts = [(1.0, [-1.25, -1.5, 4.0], [23.0, -8.0, 7.5]), (2.0, [11.0, -3.0, 1.0], [49.0, 6.0, -3.0]),
             (3.0, [-20.0, 0.0, 2.0], [17.0, 1.0, 3.0]), (4.0, [-49.0, -20.0, 13.0], [1.55, -1.0, 9.0])]
    pv = _pv_from(mzo, parts)
    pia = mzo.PIA(100, 1)
    pia.set(1, 1, 4, 1, 4, 4)
    pia.set_n_total(1, 4)
    st = _store(hip, mzo, pv, pia, 100)
    sp = [hip.Species("Ar", helpers.MASS_AR)]
    st.set_grid(hip.Grid1DUniform(50.0, 100))
    st.set_walls(hip.MaxwellWalls1D(sp, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0))
    hip.convect_particles(st, 1, 2.0)
    hip.sort_particles(st, 1)
    got = st.download_species(1)
    eps = np.finfo(float).eps
    assert list(got["index"]) == [4, 1, 3, 2]
    assert np.max(np.abs(got["x"][0] - [3.55, -1.0, 9.0])) < 3.65e-15
    assert list(got["v"][0]) == [-49.0, -20.0, 13.0] and got["w"][0] == 4.0
    assert np.max(np.abs(got["x"][1] - [20.5, -8.0, 7.5])) < 2 * eps
    assert np.max(np.abs(got["x"][2] - [23.0, 1.0, 3.0])) < 2 * eps and list(got["v"][2]) == [20.0, 0.0, 2.0]
    assert np.max(np.abs(got["x"][3] - [29.0, 6.0, -3.0])) < 2 * eps and list(got["v"][3]) == [-11.0, -3.0, 1.0]
    st.close()


@pytest.mark.parametrize("rng_mode", [0, 1])
def test_convect_diffuse_vs_oracle(hip, mzo, rng_mode):
    """Diffuse Couette walls, Philox or replayed draws: the following sort is
    bit-exact and post-reflection velocities agree to 1e-12."""
    case = helpers.couette_case(mzo, n_cells=16, ppc=200, seed=11, dt=2e-9)
    st = helpers.store_from_case(hip, case)
    tag = (3 << 4) | 2
    if rng_mode == 1:
        n_ids = case["pv"].length
        table = np.random.default_rng(1).uniform(size=(n_ids, 12))
        st.set_replay("convect", table)
        r = mzo.RNG(table=table)
    else:
        r = mzo.RNG(seed=case["seed"], tag=tag)
    hip.convect_particles(st, 1, case["dt"], tag=tag, rng_mode=rng_mode)
    mzo.lib().mzo_convect_particles(r.ref, C.byref(case["grid"]), C.byref(case["walls"]), case["pv"].p,
                                    case["pia"].p, 1, case["dt"])
    hip.sort_particles(st, 1)
    mzo.lib().mzo_sort_particles(C.byref(case["grid"]), case["pv"].p, case["pia"].p, 1)
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, case["pv"], case["pia"], rtol=RTOL)
    st.close()


# --------------------------------------------------------------- properties
def test_props_vs_oracle(hip, mzo):
    """compute_props! over groups 1 + 2 and compute_props_sorted! (with and
    without number density) against the oracle's two-pass sums."""
    rng = np.random.default_rng(5)
    nc, n = 7, 2100
    parts = [(rng.uniform(0.5, 2.0) * 1e10, rng.normal([100.0, -500.0, 3.0], 300.0), [0, 0, 0]) for _ in range(n)]
    pv = _pv_from(mzo, parts)
    pia = mzo.PIA(nc, 1)
    pos = 1
    g2 = n - 700
    for c in range(1, nc + 1):
        k1 = 200
        pia.set(c, 1, k1 + 100, pos, pos + k1 - 1, k1, g2 + 1, g2 + 100, 100)
        pos += k1
        g2 += 100
    pia.set(3, 1, 0, 0, -1, 0)   # an empty cell
    pia.set_n_total(1, n - 300)
    pia.set_contiguous(1, False)
    st = _store(hip, mzo, pv, pia, nc, cap=4096)
    sp = [hip.Species("Ar", helpers.MASS_AR)]
    pp = hip.compute_props(st, sp)
    ref = mzo.Props(nc, 1)
    mzo.compute_props([pv], pia, [helpers.MASS_AR], ref)
    assert np.array_equal(pp.np, ref.np)
    assert np.allclose(pp.n, ref.n, rtol=1e-14, atol=0)
    assert np.allclose(pp.v, ref.v, rtol=RTOL, atol=1e-9)
    assert np.allclose(pp.T, ref.T, rtol=RTOL, atol=0)
    st.set_grid(hip.Grid1DUniform(7e-6, nc))
    pp2 = hip.compute_props_sorted(st, sp, ndens_not_Np=True)
    ref2 = mzo.Props(nc, 1, ndens_not_Np=True)
    mzo.compute_props_sorted([pv], pia, [helpers.MASS_AR], ref2, inv_V=np.full(nc, 1.0 / 1e-6))
    assert np.array_equal(pp2.np, ref2.np)
    assert np.allclose(pp2.n, ref2.n, rtol=1e-14, atol=0)
    assert np.allclose(pp2.T, ref2.T, rtol=RTOL, atol=0)
    # identical particles give T = 0 exactly (test_computes.jl:33)
    pv3 = _pv_from(mzo, [(5e21, [-1.0, 2.0, -4.0], [10.0, 20.0, 30.0])] * 2000)
    pia3 = mzo.PIA(1, 1)
    pia3.set(1, 1, 2000, 1, 2000, 2000)
    pia3.set_n_total(1, 2000)
    st3 = _store(hip, mzo, pv3, pia3, 1)
    pp3 = hip.compute_props(st3, sp)
    assert pp3.T[0, 0] == 0.0 and pp3.np[0, 0] == 2000 and np.allclose(pp3.v[0, 0], [-1.0, 2.0, -4.0], atol=1e-13)
    st.close(); st3.close()


# ---------------------------------------------------------------------- NTC
def _zero_d_case(mzo, n_cells, n_per, seed, weights=None, T=300.0):
    rng = np.random.default_rng(seed)
    n = n_cells * n_per
    pv = mzo.ParticleVector(n)
    pia = mzo.PIA(n_cells, 1)
    vs = np.sqrt(KB_T(T))
    for i in range(1, n + 1):
        pv.pop_to(i)
        w = 1e12 if weights is None else 1e12 * weights[i - 1]
        pv.set(i, w, rng.normal(0, vs, 3), [rng.uniform(), rng.uniform(), rng.uniform()])
    for c in range(1, n_cells + 1):
        pia.set(c, 1, n_per, (c - 1) * n_per + 1, c * n_per, n_per)
    pia.set_n_total(1, n)
    return pv, pia


def KB_T(T):
    return helpers.KB * T / helpers.MASS_AR


def _ntc_setup(hip, mzo, pv, pia, n_cells, sgwm, cap=None):
    st = _store(hip, mzo, pv, pia, n_cells, cap=cap)
    it = mzo.interactions(["Ar"])[0][0]
    st.set_interactions([[type("I", (), dict(mu1=it.mu1, mu2=it.mu2, vhs_factor=it.vhs_factor, vhs_o=it.vhs_o))()]])
    st.set_sigma_g_w_max(1, 1, sgwm)
    st.set_seed(77)
    cfs = (mzo.CF * n_cells)()
    for c in range(n_cells):
        cfs[c].sigma_g_w_max = sgwm
    return st, it, cfs


@pytest.mark.parametrize("rng_mode,equal_weight", [(0, True), (1, True), (0, False)])
def test_ntc_single_species_vs_oracle(hip, mzo, rng_mode, equal_weight):
    """ntc! / ntc_equal_weight! on 48 cells x 300 equal-weight particles with
    ~100 candidates per cell: candidate counts and acceptances exact,
    velocities to 1e-12."""
    nc, npc = 48, 300
    pv, pia = _zero_d_case(mzo, nc, npc, seed=21)
    # sigma g w ~ 2e-4 for these particles; 0.5 dt N (N-1) sgwm / V ~ 100
    sgwm = 2e-4
    st, it, cfs = _ntc_setup(hip, mzo, pv, pia, nc, sgwm)
    dt, V = 11.0, 1.0
    tag = (5 << 4) | 1
    if rng_mode == 1:
        table = np.random.default_rng(3).uniform(size=(nc, 4096))
        st.set_replay("ntc", table)
        r = mzo.RNG(table=table)
    else:
        r = mzo.RNG(seed=77, tag=tag)
    hip.ntc(st, 1, 1, dt, V=np.full(nc, V), equal_weight=equal_weight, tag=tag, rng_mode=rng_mode)
    for c in range(1, nc + 1):
        mzo.lib().mzo_ntc_single(r.ref, C.byref(cfs[c - 1]), C.byref(it), pv.p, pia.p, c, 1, dt, V, 1e-16,
                                 1 if equal_weight else 0)
    assert r.s.err == 0
    cf = st.collision_factors(1, 1)
    assert np.array_equal(cf["n_coll"], [cfs[c].n_coll for c in range(nc)])
    assert np.array_equal(cf["n_coll_performed"], [cfs[c].n_coll_performed for c in range(nc)])
    assert np.array_equal(cf["n_eq_w_coll_performed"], [cfs[c].n_eq_w_coll_performed for c in range(nc)])
    assert np.allclose(cf["sigma_g_w_max"], [cfs[c].sigma_g_w_max for c in range(nc)], rtol=RTOL, atol=0)
    assert min(cf["n_coll"]) > 20 and sum(cf["n_coll_performed"]) > 100
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=RTOL)
    st.close()


@pytest.mark.parametrize("rng_mode", [0, 1])
def test_ntc_weight_splitting_vs_oracle(hip, mzo, rng_mode):
    """ntc! with variable weights: every accepted collision splits the heavier
    particle; the appended group-2 positions, ids (free-id LIFO with resize!
    growth), PIA and counters match the oracle's sequential loop over cells
    bit for bit; velocities to 1e-12; mass and momentum are conserved."""
    nc, npc = 24, 200
    w = np.random.default_rng(4).uniform(0.5, 2.0, nc * npc)
    pv, pia = _zero_d_case(mzo, nc, npc, seed=8, weights=w)
    st, it, cfs = _ntc_setup(hip, mzo, pv, pia, nc, 3e-4, cap=nc * npc * 2)
    dt = 8.0
    tag = (9 << 4) | 1
    if rng_mode == 1:
        table = np.random.default_rng(5).uniform(size=(nc, 8192))
        st.set_replay("ntc", table)
        r = mzo.RNG(table=table)
    else:
        r = mzo.RNG(seed=77, tag=tag)
    m0 = sum(pv.w[:nc * npc])
    hip.ntc(st, 1, 1, dt, V=np.ones(nc), equal_weight=False, tag=tag, rng_mode=rng_mode)
    for c in range(1, nc + 1):
        mzo.lib().mzo_ntc_single(r.ref, C.byref(cfs[c - 1]), C.byref(it), pv.p, pia.p, c, 1, dt, 1.0, 1e-16, 0)
    got = st.download_species(1)
    assert got["err"] == 0
    assert got["n_total"] > nc * npc + 50            # splits happened and length grew
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=RTOL)
    assert abs(got["w"][:got["n_total"]].sum() / m0 - 1) < 1e-14
    # A second call on the last cell extends its group 2, which ends at the
    # global end.  (On any other cell the reference's end2 = n_total would make
    # group 2 span the particles other cells appended after it, an inconsistent
    # PIA; that input is outside parity.)
    tag2 = (10 << 4) | 1
    r2 = mzo.RNG(seed=77, tag=tag2) if rng_mode == 0 else mzo.RNG(table=table)
    hip.ntc(st, 1, 1, dt, V=np.ones(nc), equal_weight=False, tag=tag2, rng_mode=rng_mode, cells=(nc, nc))
    mzo.lib().mzo_ntc_single(r2.ref, C.byref(cfs[nc - 1]), C.byref(it), pv.p, pia.p, nc, 1, dt, 1.0, 1e-16, 0)
    helpers.assert_state_matches(mzo, st.download_species(1), pv, pia, rtol=RTOL)
    st.close()


def test_ntc_two_species_vs_oracle(hip, mzo):
    """The 0-D two-gas step (test/test_2species.jl:52-62): Ar-Ar, He-Ar,
    He-He on one cell, ten steps, against the oracle."""
    mAr, mHe = 66.3e-27, 6.65e-27
    Fnum = 5e12
    nAr, nHe = 400, 4000
    pvs, pia = [mzo.ParticleVector(nAr), mzo.ParticleVector(nHe)], mzo.PIA(1, 2)
    r = mzo.RNG(seed=1234, tag=99)
    for s, (n, m, T) in enumerate(((nAr, mAr, 3000.0), (nHe, mHe, 360.0)), start=1):
        r.s.stream, r.s.k = s, 0
        mzo.lib().mzo_sample_particles_equal_weight(r.ref, pvs[s - 1].p, pia.p, 1, s, n, m, T, Fnum,
                                                    0, 1, 0, 1, 0, 1)
    its = mzo.interactions(["Ar", "He"])
    st = hip.DeviceStore(1, 2, [2 * nAr, 2 * nHe])
    for s in (1, 2):
        helpers.upload_from_oracle(st, mzo, pvs[s - 1], pia, species=s)
    st.set_interactions([[type("I", (), dict(mu1=its[i][k].mu1, mu2=its[i][k].mu2, vhs_factor=its[i][k].vhs_factor,
                                             vhs_o=its[i][k].vhs_o))() for k in range(2)] for i in range(2)])
    masses = [mAr, mHe]
    cfs = {}
    for (i, k) in ((1, 1), (2, 1), (2, 2), (1, 2)):
        sg = mzo.lib().mzo_estimate_sigma_g_w_max(C.byref(its[i - 1][k - 1]), masses[i - 1], masses[k - 1],
                                                  [3000.0, 360.0][i - 1], [3000.0, 360.0][k - 1], Fnum, 1.0)
        cf = mzo.CF()
        cf.sigma_g_w_max = sg
        cfs[(i, k)] = cf
        st.set_sigma_g_w_max(i, k, [sg])
    st.set_seed(1234)
    dt = 2.5e-3
    for step in range(1, 11):
        for (s1, s2) in ((1, 1), (2, 1), (2, 2)):
            tag = (step << 4) | (s1 * 4 + s2)
            hip.ntc(st, s1, s2, dt, V=[1.0], equal_weight=False, tag=tag)
            rr = mzo.RNG(seed=1234, tag=tag)
            if s1 == s2:
                mzo.lib().mzo_ntc_single(rr.ref, C.byref(cfs[(s1, s1)]), C.byref(its[s1 - 1][s1 - 1]),
                                         pvs[s1 - 1].p, pia.p, 1, s1, dt, 1.0, 1e-16, 0)
            else:
                mzo.lib().mzo_ntc_two(rr.ref, C.byref(cfs[(s1, s2)]), C.byref(its[s1 - 1][s2 - 1]), pvs[s1 - 1].p,
                                      pvs[s2 - 1].p, pia.p, 1, s1, s2, dt, 1.0, 1e-16, 0)
    for s in (1, 2):
        helpers.assert_state_matches(mzo, st.download_species(s), pvs[s - 1], pia, species=s, rtol=1e-11)
    pp = hip.compute_props(st, [hip.Species("Ar", mAr), hip.Species("He", mHe)])
    ref = mzo.Props(1, 2)
    mzo.compute_props(pvs, pia, masses, ref)
    assert np.allclose(pp.T, ref.T, rtol=1e-11)
    st.close()


# ------------------------------------------------------------ whole steps
@pytest.mark.parametrize("graph", [False, True])
def test_couette_steps_vs_oracle(hip, mzo, graph):
    """Twelve full Couette steps (ntc!, convection, sort,
    compute_props_sorted!), eager or as a replayed hipGraph, against the
    oracle: index and PIA bit-exact, velocities and T to 1e-11."""
    case = helpers.couette_case(mzo, n_cells=20, ppc=100, seed=3, dt=2e-9, sgwm_mult=50.0)
    st = helpers.store_from_case(hip, case)
    hip.couette_steps(st, 12, 1, case["dt"], equal_weight=True, props_every=1, graph=graph)
    for step in range(1, 13):
        helpers.oracle_couette_step(mzo, case, step, equal_weight=True)
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, case["pv"], case["pia"], rtol=1e-11)
    pp = hip.PhysProps(case["n_cells"], 1).fetch(st)
    assert np.allclose(pp.T[0], helpers.oracle_props_sorted(mzo, case), rtol=1e-11)
    cf = st.collision_factors(1, 1)
    assert sum(cf["n_coll_performed"]) > 0
    st.close()
