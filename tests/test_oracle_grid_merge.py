"""Grid N:2 merging (merge_grid_based!, src/merging/merging_grid.jl): the
oracle pinned to the answers of test/test_merging_grid_merging.jl (extents,
conservation of n, v and T, empty octants, two index groups) and to closed
forms computed here with numpy from the particles themselves: each velocity
cell's pair carries half its weight at its mean plus and minus its standard
deviation, the pairs come out in cell order, and cells of one or two
particles are copied.  CPU only.
"""
import ctypes as C

import numpy as np
import pytest

from test_oracle_kat import make_pv

KB = 1.380649e-23
M_AR = 66.3e-27


def _octant_particles(mzo, weights):
    """create_24_3particles_in_octant (test_merging_grid_merging.jl:4-44):
    three particles per octant at speeds 9 - o -/+ 0.5 and 9 - o, weight
    o * weights[o]."""
    parts = []
    for o in range(1, 9):
        for dv in (-0.5, 0.5, 0.0):
            val = 9.0 - o + dv
            vx = -val if o % 2 == 1 else val
            vy = val if o in (3, 4, 7, 8) else -val
            vz = val if o >= 5 else -val
            parts.append((o * weights[o - 1], [vx, vy, vz], [1.0, -10.0, 3.0]))
    return make_pv(mzo, parts)


def _pia(mzo, n, groups=None):
    pia = mzo.PIA(1, 1)
    if groups is None:
        pia.set(1, 1, n, 1, n, n)
    else:
        (s1, e1), (s2, e2) = groups
        pia.set(1, 1, n, s1, e1, e1 - s1 + 1, s2, e2, e2 - s2 + 1)
    pia.set_n_total(1, n)
    return pia


def _props(mzo, pv, pia):
    pp = mzo.Props(1, 1)
    mzo.compute_props([pv], pia, [M_AR], pp)
    return pp


def _merge(mzo, mg, pv, pia, seed=1234, grid=None):
    r = mzo.RNG(seed=seed, tag=5)
    mzo.lib().mzo_merge_grid_based(r.ref, mg.p, pv.p, pia.p, 1, 1, C.byref(grid) if grid is not None else None)
    return r


def _live(pv, pia):
    d = pia.ind(1)
    pos = list(range(d.start1, d.end1 + 1)) + (list(range(d.start2, d.end2 + 1)) if d.n_group2 > 0 else [])
    return [pv.get(p) for p in pos]


def test_grid_constructors_and_extents(mzo):
    """test_merging_grid_merging.jl:222-248: Ntotal, the explicit extents
    (lower, upper, mid, dv); and the extents from a cell's T and v
    (:190-197): v -/+ mult sqrt(2 k T / m)."""
    mg = mzo.GridMerge(4, 8, 2, [3.0, 2.0, 1.0])
    assert mg.s.Ntotal == 4 * 8 * 2 + 8 and mg.s.NyNz == 16
    mg.extent([-3.0, 2.0], [-2.5, 2.5], [-2.0, 4.5])
    assert list(mg.s.lower) == [-3.0, -2.5, -2.0] and list(mg.s.upper) == [2.0, 2.5, 4.5]
    assert list(mg.s.mid) == [-0.5, 0.0, 1.25]
    assert np.max(np.abs(np.array(mg.s.dv) - [1.25, 0.625, 3.25])) < 2 * np.finfo(float).eps
    mg = mzo.GridMerge(6, 6, 6, 1.5)
    assert list(mg.s.mult) == [1.5, 1.5, 1.5]
    mg.extent_props(300.0, [100.0, -50.0, 7.0], M_AR)
    vt = np.sqrt(2 * 300.0 * KB / M_AR)
    assert np.allclose(mg.s.lower, np.array([100.0, -50.0, 7.0]) - 1.5 * vt, rtol=1e-15, atol=0)
    assert np.allclose(mg.s.upper, np.array([100.0, -50.0, 7.0]) + 1.5 * vt, rtol=1e-15, atol=0)
    assert np.allclose(mg.s.dv, 2 * 1.5 * vt / 6, rtol=1e-15)


def test_grid_index_is_the_cell_holding_the_velocity(mzo):
    """For velocities inside the grid the index decodes to the velocity cell
    whose bounds hold v; outside, to the octant of v about the grid's middle."""
    rng = np.random.default_rng(3)
    mg = mzo.GridMerge(3, 4, 5, 1.0)
    mg.extent([-1.0, 2.0], [-4.0, 4.0], [0.0, 5.0])
    lo, dv = np.array(mg.s.lower), np.array(mg.s.dv)
    for _ in range(2000):
        v = rng.uniform([-2.0, -6.0, -1.0], [3.0, 6.0, 6.0])
        k = mg.index(list(v))
        inside = np.all(v >= [-1.0, -4.0, 0.0]) and np.all(v <= [2.0, 4.0, 5.0])
        if inside:
            q = k - 1
            ix, iy, iz = q // 20, (q // 5) % 4, q % 5
            cell_lo = lo + dv * np.array([ix, iy, iz])
            assert np.all(cell_lo <= v + 1e-12) and np.all(v < cell_lo + dv + 1e-12), (v, k)
        else:
            o = (v[0] > 0.5) + 2 * (v[1] > 0.0) + 4 * (v[2] > 2.5)
            assert k == 3 * 4 * 5 + 1 + o


def test_grid_merge_maxwellian_conserves(mzo):
    """test_merging_grid_merging.jl:48-91: 5,000 drifting Maxwellian
    particles, a 2 x 2 x 2 grid at 3.5 thermal speeds: fewer particles,
    n kept to eps, T to 3e-12 K and v to 6.5e-12 m/s."""
    n = 5000
    pv, pia = mzo.ParticleVector(n), mzo.PIA(1, 1)
    r = mzo.RNG(seed=1234, tag=1)
    mzo.lib().mzo_sample_particles_equal_weight(r.ref, pv.p, pia.p, 1, 1, n, M_AR, 300.0, 1e8, 0, 1, 0, 1, 0, 1)
    ids = pv.index[:n]
    pv.v[ids - 1] += [2000.0, 500.0, -400.0]
    p0 = _props(mzo, pv, pia)
    mg = mzo.GridMerge(2, 2, 2, 3.5)
    mg.extent_props(p0.T[0, 0], p0.v[0, 0], M_AR)
    _merge(mzo, mg, pv, pia)
    assert pia.n_total() < n and pia.n_total() == pia.ind(1).n_local
    p1 = _props(mzo, pv, pia)
    assert p1.np[0, 0] == pia.n_total()
    assert abs(p1.n[0, 0] - p0.n[0, 0]) <= np.finfo(float).eps * p0.n[0, 0]
    assert abs(p1.T[0, 0] - p0.T[0, 0]) < 3e-12
    assert np.all(np.abs(p1.v[0, 0] - p0.v[0, 0]) < 6.5e-12)


@pytest.mark.parametrize("weights,groups", [([1, 1, 1, 1, 1, 1, 1, 0], None),
                                            ([1, 1, 0, 1, 0, 1, 1, 0], ((1, 7), (8, 24)))])
def test_grid_merge_octants_with_empty_ones(mzo, weights, groups):
    """test_merging_grid_merging.jl:93-206: three particles per octant, some
    octants of zero weight, one or two index groups, a 2 x 2 x 2 grid at 0.5
    thermal speeds: velocity cells and octants of zero total weight vanish,
    the others give min(count, 2) particles; n, v, T kept."""
    pv = _octant_particles(mzo, [float(w) for w in weights])
    pia = _pia(mzo, 24, groups)
    p0 = _props(mzo, pv, pia)
    mg = mzo.GridMerge(2, 2, 2, 0.5)
    mg.extent_props(p0.T[0, 0], p0.v[0, 0], M_AR)
    parts = _live(pv, pia)
    idx = np.array([mg.index(list(v)) for w, v, x in parts])
    ws = np.array([w for w, v, x in parts])
    expect = sum(min(int((idx == g).sum()), 2) for g in set(idx.tolist()) if ws[idx == g].sum() > 0)
    _merge(mzo, mg, pv, pia)
    assert pia.n_total() == expect < 24
    p1 = _props(mzo, pv, pia)
    assert abs(p1.n[0, 0] - p0.n[0, 0]) <= np.finfo(float).eps * p0.n[0, 0]
    assert abs(p1.T[0, 0] - p0.T[0, 0]) < 1e-14 * p0.T[0, 0]
    assert np.all(np.abs(p1.v[0, 0] - p0.v[0, 0]) < 1e-14 * np.abs(p0.v[0, 0]).max() + 1e-300)
    d = pia.ind(1)
    if groups is None:
        assert d.end1 == pia.n_total() and (d.start2, d.end2) == (0, -1)
    else:
        assert d.end1 - d.start1 + 1 + d.end2 - d.start2 + 1 == pia.n_total() and d.start2 > 0
    for w, v, x in _live(pv, pia):
        assert w > 0 and not np.isnan(x[0])


def test_grid_merge_pairs_closed_form(mzo):
    """400 variable-weight particles on a 3 x 3 x 3 grid: the merged
    particles are, in velocity-cell order, for each cell of three or more
    particles a pair of half its weight at its weighted mean -/+ its weighted
    standard deviation (one sign draw per axis), and the originals for cells
    of one or two; the draws are six per cell of three or more; with a grid
    given, x is clamped to it."""
    rng = np.random.default_rng(11)
    n = 400
    vs = np.sqrt(KB * 300.0 / M_AR)
    parts = [(float(rng.lognormal(0, 0.7)), rng.normal(0, vs, 3), [rng.uniform(-0.2, 1.2), 0.1, 0.2])
             for _ in range(n)]
    pv = make_pv(mzo, parts)
    pia = _pia(mzo, n)
    W = np.array([p[0] for p in parts])
    V = np.array([p[1] for p in parts])
    X = np.array([p[2] for p in parts])
    mg = mzo.GridMerge(3, 3, 3, 1.2)
    T = 300.0
    mg.extent_props(T, [0.0, 0.0, 0.0], M_AR)
    idx = np.array([mg.index(list(v)) for v in V])
    grid = mzo.grid1d(1.0, 4)
    r = _merge(mzo, mg, pv, pia, grid=grid)
    out = _live(pv, pia)
    k = 0
    n_big = 0
    for g in range(1, mg.s.Ntotal + 1):
        sel = np.nonzero(idx == g)[0]
        if len(sel) == 0:
            continue
        if len(sel) <= 2:
            for j in sel:
                w, v, x = out[k]
                assert w == W[j] and np.array_equal(v, V[j])
                assert x[0] == min(max(X[j, 0], 0.0), 1.0) if len(sel) == 2 else x[0] == X[j, 0]
                k += 1
            continue
        n_big += 1
        wt = W[sel].sum()
        vm = (W[sel, None] * V[sel]).sum(0) / wt
        sd = np.sqrt((W[sel, None] * (V[sel] - vm) ** 2).sum(0) / wt)
        (w1, v1, x1), (w2, v2, x2) = out[k], out[k + 1]
        assert w1 == w2 and abs(w1 / (0.5 * wt) - 1) < 1e-14
        assert np.allclose(0.5 * (v1 + v2), vm, rtol=0, atol=1e-12 * vs)
        assert np.allclose(np.abs(v1 - v2) / 2, sd, rtol=1e-10, atol=1e-12 * vs)
        assert 0.0 <= x1[0] <= 1.0 and 0.0 <= x2[0] <= 1.0
        k += 2
    assert k == len(out) == pia.n_total() and n_big > 3
    assert r.s.k == 6 * n_big


def test_grid_merge_one_cell_grid(mzo):
    """A 1 x 1 x 1 grid wide enough for every particle: one pair with the
    cell's total weight, mean velocity and (per axis) spread; energy kept."""
    rng = np.random.default_rng(2)
    parts = [(float(rng.uniform(0.5, 2)), rng.normal(0, 300.0, 3), [0.5, 0.0, 0.0]) for _ in range(50)]
    pv = make_pv(mzo, parts)
    pia = _pia(mzo, 50)
    W = np.array([p[0] for p in parts])
    V = np.array([p[1] for p in parts])
    mg = mzo.GridMerge(1, 1, 1, 100.0)
    mg.extent([-1e4, 1e4], [-1e4, 1e4], [-1e4, 1e4])
    _merge(mzo, mg, pv, pia)
    assert pia.n_total() == 2
    out = _live(pv, pia)
    e0 = (W * (V ** 2).sum(1)).sum()
    e1 = sum(w * (v ** 2).sum() for w, v, x in out)
    assert abs(e1 / e0 - 1) < 1e-13
    assert abs(sum(w for w, v, x in out) / W.sum() - 1) < 1e-15
