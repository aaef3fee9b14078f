"""Roulette merging on the device (mz_merge_roulette) against the oracle's
merge_roulette!: the index array, PIA and free-id buffer bit for bit, weights
and velocities to 1e-15 (the device sums in the reference's order), and the
conservation laws at full size."""

import numpy as np
import pytest

import helpers
from test_oracle_roulette import _case_contiguous, _case_groups

pytestmark = pytest.mark.gpu

RTOL = 1e-15


def _grouped_cells(mzo, counts, seed):
    """Cells whose first two thirds sit in group 1 (contiguous, in cell
    order) and the rest in group 2 at the end of the array (in cell order),
    the layout weight splitting leaves; lognormal weights, drifting
    Maxwellians."""
    rng = np.random.default_rng(seed)
    n = int(sum(counts))
    nc = len(counts)
    pv = mzo.ParticleVector(n)
    pia = mzo.PIA(nc, 1)
    vs = np.sqrt(helpers.KB * 300.0 / helpers.MASS_AR)
    k1s = [2 * k // 3 for k in counts]
    pos1, pos2 = 1, sum(k1s) + 1
    for c, (k, k1) in enumerate(zip(counts, k1s), start=1):
        k2 = k - k1
        for j in range(k):
            p = pos1 + j if j < k1 else pos2 + (j - k1)
            pv.pop_to(p)
            w = 1e14 * rng.lognormal(0.0, 1.0)
            v = rng.normal(0, vs, 3) + np.array([0.0, 300.0 * (2 * c / nc - 1), 0.0])
            pv.set(p, w, v, [(c - 1 + rng.uniform()) * 1e-6, 0.0, 0.0])
        pia.set(c, 1, k, pos1 if k1 else 0, pos1 + k1 - 1 if k1 else -1, k1,
                pos2 if k2 else 0, pos2 + k2 - 1 if k2 else -1, k2)
        pos1 += k1
        pos2 += k2
    pia.set_n_total(1, n)
    pia.set_contiguous(1, False)
    return pv, pia


def _store(hip, mzo, pv, pia, nc, cap):
    st = hip.DeviceStore(nc, 1, cap)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    st.set_seed(1234)
    return st


def _oracle_roulette(mzo, pv, pia, threshold, target, conservative, tag, seed=1234):
    r = mzo.RNG(seed=seed, tag=tag)
    for c in range(1, pia.n_cells + 1):
        if pia.ind(c).n_local > threshold:
            r.s.stream, r.s.k = c - 1, 0
            mzo.lib().mzo_merge_roulette(r.ref, pv.p, pia.p, c, 1, target, 1 if conservative else 0)


@pytest.mark.parametrize("case,target,conservative", [("contiguous", 3, False), ("groups", 1, False),
                                                      ("groups_varied", 3, False), ("groups_cons", 3, True)])
def test_roulette_reference_cases(hip, mzo, case, target, conservative):
    """The cases of test/test_roulette_merging.jl, cell 2 merged on the
    device and by the oracle with the same draws."""
    make = {"contiguous": lambda: _case_contiguous(mzo), "groups": lambda: _case_groups(mzo),
            "groups_varied": lambda: _case_groups(mzo, varied=True),
            "groups_cons": lambda: _case_groups(mzo, varied=True, conservative_weights=True)}[case]
    pv, pia = make()
    st = _store(hip, mzo, pv, pia, 3, 64)
    threshold = 4   # only cell 2 (five particles) is above it
    tag = (3 << 4) | 3
    assert hip.merge_roulette(st, 1, threshold, target, conservative=conservative, tag=tag) == 1
    _oracle_roulette(mzo, pv, pia, threshold, target, conservative, tag)
    got = st.download_species(1)
    assert got["err"] == 0 and not got["contiguous"]
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=RTOL)
    st.close()


@pytest.mark.parametrize("conservative", [False, True])
def test_roulette_many_cells_vs_oracle(hip, mzo, conservative):
    """60 cells of 0..2,000 particles in groups 1 and 2, those above 300
    merged to 120: state against the oracle, and the conservation laws."""
    rng = np.random.default_rng(8)
    counts = [int(k) for k in rng.integers(0, 2000, 60)]
    counts[5], counts[6] = 0, 301
    pv, pia = _grouped_cells(mzo, counts, seed=4)
    st = _store(hip, mzo, pv, pia, len(counts), sum(counts) + 512)
    before = st.download_species(1)
    tag = (11 << 4) | 3
    n_merged = hip.merge_roulette(st, 1, 300, 120, conservative=conservative, tag=tag)
    assert n_merged == sum(1 for k in counts if k > 300)
    _oracle_roulette(mzo, pv, pia, 300, 120, conservative, tag)
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=RTOL)
    assert got["n_total"] == sum(min(k, 120) if k > 300 else k for k in counts)
    # per cell: total weight kept; conservative also keeps momentum and energy
    for c in range(len(counts)):
        rb, ra = before["pia"][c], got["pia"][c]

        def sel(row):
            return np.concatenate([np.arange(row[1] - 1, row[2]) if row[3] > 0 else np.zeros(0, int),
                                   np.arange(row[4] - 1, row[5]) if row[6] > 0 else np.zeros(0, int)]).astype(int)
        ib, ia = sel(rb), sel(ra)
        wb, wa = before["w"][ib], got["w"][ia]
        if len(ib) == 0:
            continue
        assert abs(wa.sum() / wb.sum() - 1) < 1e-13   # the re-weighting rounds once per particle
        if conservative and counts[c] > 300:
            pb = (wb[:, None] * before["v"][ib]).sum(0)
            pa = (wa[:, None] * got["v"][ia]).sum(0)
            assert np.all(np.abs(pa - pb) <= 1e-12 * np.abs(wb[:, None] * before["v"][ib]).sum(0))
            eb = (wb * (before["v"][ib] ** 2).sum(1)).sum()
            ea = (wa * (got["v"][ia] ** 2).sum(1)).sum()
            assert abs(ea / eb - 1) < 1e-12
    st.close()
