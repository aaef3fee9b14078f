"""Roulette merging (merge_roulette!): the oracle pinned to the answers of
test/test_roulette_merging.jl and to the conservation laws.  The reference's
cases draw from StableRNG; where its asserts depend on which particles the
draws pick, the RNG-independent part is asserted (stated at each case).
CPU only.
"""
import numpy as np
import pytest

from test_oracle_kat import make_pv

X2 = [0.55, 0.95, 0.8, 0.85, 0.99]


def _pia3(mzo, cells):
    """cells: [(n1, s1, e1, n2, s2, e2)] for cells 1..3"""
    pia = mzo.PIA(3, 1)
    for c, (n1, s1, e1, n2, s2, e2) in enumerate(cells, start=1):
        pia.set(c, 1, n1 + n2, s1 if n1 else 0, e1 if n1 else -1, n1, s2 if n2 else 0, e2 if n2 else -1, n2)
    pia.set_n_total(1, 12)
    pia.set_contiguous(1, True)
    return pia


def _case_contiguous(mzo):
    """test_roulette_merging.jl:5-57"""
    parts = [(float(i), [0.5, 3.0, 4.0], [[0.05, 0.05, 0.05, 0.45][i - 1], 0.0, 1.0]) for i in range(1, 5)]
    parts += [(3.0, [-0.5, -3.0, -4.0], [X2[i - 5], 0.0, 2.0]) for i in range(5, 10)]
    parts += [(3.0, [20.0, 20.0, 20.0], [[10.0, 11.0, 12.0][i - 10], 0.0, 3.0]) for i in range(10, 13)]
    return make_pv(mzo, parts), _pia3(mzo, [(4, 1, 4, 0, 0, -1), (5, 5, 9, 0, 0, -1), (3, 10, 12, 0, 0, -1)])


def _case_groups(mzo, varied=False, conservative_weights=False):
    """test_roulette_merging.jl:99-146 (uniform cell-2 data) and 188-247 /
    290-349 (varied velocities, and weights for the conservative case)."""
    parts = []
    for i in range(1, 5):
        v = [0.5 - 0.1 * i, 3.0 + 2.0 * i ** 2, 4.0 - 0.3 * i] if varied else [0.5, 3.0, 4.0]
        parts.append((float(i), v, [[0.05, 0.05, 0.05, 0.45][i - 1], 0.0, 1.0]))
    v5 = [-0.5 + 0.2 * 25, -3.0 + (5 * 0.2) ** 3, -4.0 + 0.3 * 5] if varied else [-0.5, -3.0, -4.0]
    parts.append((2.0, v5, [X2[0], 0.0, 2.0]))
    for i in range(6, 9):
        v = [20.0 - i, 20.0 + 0.5 * i, 20.0 - 0.2 * i] if varied else [20.0, 20.0, 20.0]
        w = 3.0 + 0.5 * i if conservative_weights else 3.0
        parts.append((w, v, [[10.0, 11.0, 12.0][i - 6], 0.0, 3.0]))
    for i in range(9, 13):
        v = [-0.5 - i, -3.0 - 0.05 * i ** 2, -4.0 + 0.5 * i] if varied else [-0.5, -3.0, -4.0]
        w = 2.0 + i if conservative_weights else 2.0
        parts.append((w, v, [X2[i - 8], 0.0, 2.0]))   # x_cell2[i-7], 1-based
    return make_pv(mzo, parts), _pia3(mzo, [(4, 1, 4, 0, 0, -1), (1, 5, 5, 4, 9, 12), (3, 6, 8, 0, 0, -1)])


def _cell_particles(pv, pia, cell):
    d = pia.ind(cell)
    pos = list(range(d.start1, d.end1 + 1)) + (list(range(d.start2, d.end2 + 1)) if d.n_group2 > 0 else [])
    return [pv.get(p) for p in pos]


def _moments(parts):
    w = np.array([p[0] for p in parts])
    v = np.array([p[1] for p in parts])
    wt = w.sum()
    v0 = (w[:, None] * v).sum(axis=0) / wt
    E = (w * ((v - v0) ** 2).sum(axis=1)).sum() / wt
    return wt, v0, E


def _roulette(mzo, pv, pia, cell, target, conservative=False, seed=1234):
    r = mzo.RNG(seed=seed, tag=7)
    r.s.stream, r.s.k = cell - 1, 0
    mzo.lib().mzo_merge_roulette(r.ref, pv.p, pia.p, cell, 1, target, 1 if conservative else 0)


def test_roulette_contiguous_kat(mzo):
    """test_roulette_merging.jl:59-95: cell 2 (five particles of weight 3)
    merged to three; the neighbours untouched, the three left weigh 5 each."""
    pv, pia = _case_contiguous(mzo)
    before = {c: _cell_particles(pv, pia, c) for c in (1, 3)}
    _roulette(mzo, pv, pia, 2, 3)
    assert not pia.contiguous(1)
    assert pia.ind(1).end1 == 4 and pia.ind(3).start1 == 10
    assert pia.ind(2).astuple() == (3, 5, 7, 3, 0, -1, 0)
    for c in (1, 3):
        after = _cell_particles(pv, pia, c)
        assert all(a[0] == b[0] and np.array_equal(a[1], b[1]) for a, b in zip(after, before[c]))
    for w, v, x in _cell_particles(pv, pia, 2):
        assert abs(w - 5.0) < 1e-15 and v[0] == -0.5 and x[0] in X2 and x[2] == 2.0
    # the two deleted particles: weight 0, their ids on the buffer, n_total down by 2
    assert pia.n_total() == 10 and pv.nbuffer == 2
    for pid in pv.buffer[:2]:
        assert pv.w[pid - 1] == 0.0 and 5 <= pid <= 9


def test_roulette_groups_kat(mzo):
    """test_roulette_merging.jl:148-184: cell 2 with one particle in group 1
    and four in group 2, merged to one.  The reference's StableRNG keeps the
    group-1 particle; with other draws the survivor may sit in group 2, so the
    assert is: one particle of weight 10 (= 2 + 4 x 2) with cell 2's data,
    and consistent group bounds."""
    for seed in (1234, 1, 2, 3):
        pv, pia = _case_groups(mzo)
        _roulette(mzo, pv, pia, 2, 1, seed=seed)
        n_local, s1, e1, n1, s2, e2, n2 = pia.ind(2).astuple()
        assert n_local == 1 and n1 + n2 == 1
        if n1:
            assert (s1, e1, s2, e2) == (5, 5, 0, -1)
        else:
            assert (s1, e1) == (0, -1) and s2 == e2 and 9 <= s2 <= 12
        (w, v, x), = _cell_particles(pv, pia, 2)
        assert abs(w - 10.0) < 1e-15 and v[0] == -0.5 and x[0] in X2
        assert pia.ind(3).start1 == 6 and pia.ind(1).end1 == 4


@pytest.mark.parametrize("seed", [1234, 5, 6])
def test_roulette_weight_conserved_momentum_not(mzo, seed):
    """test_roulette_merging.jl:186-288: the non-conservative merge keeps
    every cell's total weight; cell 2's mean velocity and temperature move
    (other cells keep theirs exactly)."""
    pv, pia = _case_groups(mzo, varied=True)
    before = {c: _moments(_cell_particles(pv, pia, c)) for c in (1, 2, 3)}
    _roulette(mzo, pv, pia, 2, 3, seed=seed)
    assert pia.ind(2).n_local == 3
    for c in (1, 2, 3):
        wt, v0, E = _moments(_cell_particles(pv, pia, c))
        assert abs(wt - before[c][0]) <= 2e-15 * before[c][0]
        if c == 2:
            assert np.max(np.abs(v0 - before[c][1])) > 1e-10 or abs(E - before[c][2]) > 1e-10
        else:
            assert np.array_equal(v0, before[c][1]) and E == before[c][2]


@pytest.mark.parametrize("seed", [1234, 5, 6])
def test_roulette_conservative(mzo, seed):
    """test_roulette_merging.jl:290-392: the conservative merge keeps the
    total weight, the mean velocity and the temperature of every cell."""
    pv, pia = _case_groups(mzo, varied=True, conservative_weights=True)
    before = {c: _moments(_cell_particles(pv, pia, c)) for c in (1, 2, 3)}
    _roulette(mzo, pv, pia, 2, 3, conservative=True, seed=seed)
    assert pia.ind(2).n_local == 3
    for c in (1, 2, 3):
        wt, v0, E = _moments(_cell_particles(pv, pia, c))
        assert abs(wt - before[c][0]) <= 1e-14 * before[c][0]
        assert np.max(np.abs(v0 - before[c][1])) <= 1e-13 * np.max(np.abs(before[c][1]))
        assert abs(E - before[c][2]) <= 1e-13 * before[c][2]


def test_roulette_draws_are_uniform(mzo):
    """Which particles survive: over many cells merged 10 -> 4, every
    position survives about 40% of the time (each of the 10 particles is
    equally likely to be among the survivors)."""
    hits = np.zeros(10)
    n_cells = 4000
    parts = [(1.0, [float(j), 0.0, 0.0], [0.5, 0.0, 0.0]) for _ in range(n_cells) for j in range(10)]
    pv = make_pv(mzo, parts)
    pia = mzo.PIA(n_cells, 1)
    for c in range(1, n_cells + 1):
        pia.set(c, 1, 10, 10 * (c - 1) + 1, 10 * c, 10)
    pia.set_n_total(1, 10 * n_cells)
    for c in range(1, n_cells + 1):
        _roulette(mzo, pv, pia, c, 4, seed=99)
        for w, v, x in _cell_particles(pv, pia, c):
            hits[int(v[0])] += 1
    frac = hits / n_cells
    assert np.all(np.abs(frac - 0.4) < 5 * np.sqrt(0.24 / n_cells)), frac
