"""SWPM collisions (swpm!): the oracle pinned by closed forms on replayed
draws (the collision count, the weight transfer min(w_i, w_k) / (1 + G), the
two new particles with the pair's centre-of-mass velocity plus and minus the
scattered relative velocity) and by the conservation laws of the method
(mass, momentum and energy are kept exactly, up to rounding).  CPU only.
"""
import ctypes as C

import numpy as np

from test_oracle_kat import make_pv

M_AR = 66.3e-27


def _it(mzo):
    return mzo.interactions(["Ar"])[0][0]


def test_swpm_single_collision_closed_form(mzo):
    """Two particles (w 3 and 1) and replayed draws: one candidate (count
    0.5 + 0.6), i = 0, k = 1, accepted, phi = pi / 2, cos(theta) = 0.5."""
    it = _it(mzo)
    vi, vk = np.array([300.0, -100.0, 50.0]), np.array([-200.0, 40.0, 10.0])
    pv = make_pv(mzo, [(3.0, vi, [0.1, 0.0, 0.0]), (1.0, vk, [0.2, 0.0, 0.0])])
    pia = mzo.PIA(1, 1)
    pia.set(1, 1, 2, 1, 2, 2)
    pia.set_n_total(1, 2)
    G, dt, V = 1.0, 1e-6, 1.0
    g = np.linalg.norm(vi - vk)
    cf = mzo.CFSwpm()
    # 0.5 dt np (np - 1) sgm w_max (G + 1) / V = 0.5
    cf.sigma_g_max = 0.5 / (0.5 * dt * 2 * 1 * 3.0 * (G + 1) / V)
    table = np.array([[0.6, 0.1, 0.7, 0.0, 0.25, 0.75, 0.5, 0.5]])
    r = mzo.RNG(table=table)
    mzo.lib().mzo_swpm(r.ref, C.byref(cf), C.byref(it), pv.p, pia.p, 1, 1, G, dt, V)
    assert cf.n_coll == 1 and cf.n_coll_performed == 1
    sgm = it.vhs_factor * g ** (1.0 - 2 * it.vhs_o) * g
    assert abs(cf.sigma_g_max / max(sgm, 0.5 / (dt * 6.0)) - 1) < 1e-14
    assert pia.n_total() == 4 and pia.ind(1).astuple() == (4, 1, 2, 2, 3, 4, 2)
    dw = 1.0 / (1.0 + G)
    w = [pv.get(p)[0] for p in range(1, 5)]
    assert w == [3.0 - dw, 1.0 - dw, dw, dw]
    # the originals keep their velocities and positions; the new ones sit on them
    assert np.array_equal(pv.get(1)[1], vi) and np.array_equal(pv.get(2)[1], vk)
    assert np.array_equal(pv.get(3)[2], pv.get(1)[2]) and np.array_equal(pv.get(4)[2], pv.get(2)[2])
    vcom = 0.5 * (vi + vk)
    phi, ct = 2 * np.pi * 0.25, 2 * 0.75 - 1.0
    st = np.sqrt(1 - ct * ct)
    gn = g * np.array([st * np.cos(phi), st * np.sin(phi), ct])
    assert np.allclose(pv.get(3)[1], vcom + 0.5 * gn, rtol=0, atol=1e-12 * g)
    assert np.allclose(pv.get(4)[1], vcom - 0.5 * gn, rtol=0, atol=1e-12 * g)


def test_swpm_rejected_and_slow_pairs(mzo):
    """A pair with g below eps draws no acceptance; a rejected pair changes
    nothing but the running (sigma g)_max, which is max(its start, sigma g)."""
    it = _it(mzo)
    v = np.array([100.0, 0.0, 0.0])
    parts = [(1.0, v, [0.1, 0, 0]), (1.0, v, [0.2, 0, 0]), (1.0, -v, [0.3, 0, 0]), (4.0, 0 * v, [0.4, 0, 0])]
    pv = make_pv(mzo, parts)
    pia = mzo.PIA(1, 1)
    pia.set(1, 1, 4, 1, 4, 4)
    pia.set_n_total(1, 4)
    g = 200.0
    sgm = it.vhs_factor * g ** (1.0 - 2 * it.vhs_o) * g
    G, dt, V = 0.0, 1.0, 1.0
    for start in (0.5 * sgm, 3.0 * sgm):
        cf = mzo.CFSwpm()
        cf.sigma_g_max = start
        # count floor(base + 0.6) = 2 with base = 1.5 (np = 4, w_max = 4)
        dt = 1.5 / (0.5 * 4 * 3 * start * 4.0 * (G + 1) / V)
        # candidate 1: positions 0, 1 (g = 0); candidate 2: 0, 2, acceptance
        # ratio sgm * 1 / 4 / max(start, sgm) <= 1/4 against the draw 0.5
        table = np.array([[0.6, 0.1, 0.3, 0.1, 0.6, 0.5, 0.5, 0.5]])
        pv = make_pv(mzo, parts)
        pia = mzo.PIA(1, 1)
        pia.set(1, 1, 4, 1, 4, 4)
        pia.set_n_total(1, 4)
        mzo.lib().mzo_swpm(mzo.RNG(table=table).ref, C.byref(cf), C.byref(it), pv.p, pia.p, 1, 1, G, dt, V)
        assert cf.n_coll == 2 and cf.n_coll_performed == 0
        assert pia.n_total() == 4 and [pv.get(q)[0] for q in range(1, 5)] == [1.0, 1.0, 1.0, 4.0]
        assert abs(cf.sigma_g_max / max(start, sgm) - 1) < 1e-14


def test_swpm_conservation_and_count(mzo):
    """300 variable-weight particles, Philox draws: the collision count is
    floor(0.5 dt N (N - 1) (sigma g)_max w_max (G + 1) / V + R0); total
    weight, momentum and energy are kept; every performed collision adds two
    particles."""
    it = _it(mzo)
    rng = np.random.default_rng(3)
    n = 300
    vs = np.sqrt(1.380649e-23 * 300.0 / M_AR)
    parts = [(float(rng.lognormal(0, 1)), rng.normal(0, vs, 3), [0.5, 0, 0]) for _ in range(n)]
    pv = make_pv(mzo, parts)
    pia = mzo.PIA(1, 1)
    pia.set(1, 1, n, 1, n, n)
    pia.set_n_total(1, n)
    cf = mzo.CFSwpm()
    cf.sigma_g_max = it.vhs_factor * (2 * vs) ** (1.0 - 2 * it.vhs_o) * 2 * vs
    G, dt = 1.0, 2e-5
    w_max = max(p[0] for p in parts)
    V = 0.5 * dt * n * (n - 1) * cf.sigma_g_max * w_max * (G + 1) / 150.0   # about 150 candidates
    expect_n = 0.5 * dt * n * (n - 1) * cf.sigma_g_max * w_max * (G + 1) / V
    r = mzo.RNG(seed=7, tag=11)
    R0 = r.draws(0, 1)[0]
    _, v0, w0, _ = pv.logical_arrays(n)
    r = mzo.RNG(seed=7, tag=11)
    mzo.lib().mzo_swpm(r.ref, C.byref(cf), C.byref(it), pv.p, pia.p, 1, 1, G, dt, V)
    assert cf.n_coll == int(np.floor(expect_n + R0)) and cf.n_coll > 50
    m = pia.n_total()
    assert m == n + 2 * cf.n_coll_performed and cf.n_coll_performed > 10
    d = pia.ind(1)
    pos = list(range(d.start1, d.end1 + 1)) + list(range(d.start2, d.end2 + 1))
    w1 = np.array([pv.get(p)[0] for p in pos])
    v1 = np.array([pv.get(p)[1] for p in pos])
    assert abs(w1.sum() / w0.sum() - 1) < 1e-13
    p0, p1 = (w0[:, None] * v0).sum(0), (w1[:, None] * v1).sum(0)
    scale = (w0[:, None] * np.abs(v0)).sum(0)
    assert np.all(np.abs(p1 - p0) <= 1e-12 * scale)
    e0, e1 = (w0 * (v0 ** 2).sum(1)).sum(), (w1 * (v1 ** 2).sum(1)).sum()
    assert abs(e1 / e0 - 1) < 1e-12
    # every weight stays positive (new particles may themselves collide later in the loop)
    assert np.all(w1 > 0)
