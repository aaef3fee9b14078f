"""compute_mixed_moment (physical_props.jl:473-501) on the device against the
oracle's restatement, over two-group cells (group 2 after every cell's
group 1, as weight splitting leaves it) and an empty cell, through the
batched entry point and the drop-in's per-cell signature.  The device sums
each cell by a wave reduction and the reference sums left to right, so the
bar is 1e-13 of the sum of the terms' magnitudes."""
import numpy as np
import pytest

import helpers
from test_gpu_roulette import _grouped_cells

pytestmark = pytest.mark.gpu

TOL = 1e-13
POWERS = [(0, 0, 0), (1, 0, 0), (0, 2, 0), (2, 1, 0), (3, 0, 2), (1, 1, 1), (0, 0, 4)]


def _magnitude(pv, pia, cell, powers, sum_scaler):
    """sum of |w sum_scaler vx^px vy^py vz^pz| over the cell's particles"""
    tab = pia.table(1)[cell - 1]
    pos = list(range(tab[1], tab[2] + 1)) + (list(range(tab[4], tab[5] + 1)) if tab[6] > 0 else [])
    ids = pv.index[np.array(pos, dtype=np.int64) - 1] - 1 if pos else np.zeros(0, dtype=np.int64)
    v = pv.v[ids]
    t = pv.w[ids] * sum_scaler * v[:, 0] ** powers[0] * v[:, 1] ** powers[1] * v[:, 2] ** powers[2]
    return float(np.abs(t).sum())


def test_mixed_moment_vs_oracle(hip, mzo):
    counts = [40, 0, 301, 5, 1200, 64, 2]
    pv, pia = _grouped_cells(mzo, counts, seed=17)
    nc = len(counts)
    st = hip.DeviceStore(nc, 1, sum(counts) + 64)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    ndens = 1e-17
    for powers in POWERS:
        for ss, rs in ((1.0, 1.0), (ndens, 1.0 / ndens)):
            got = hip.compute_mixed_moment(st, 1, powers, sum_scaler=ss, res_scaler=rs)
            for c in range(1, nc + 1):
                ref = mzo.lib().mzo_compute_mixed_moment(pv.p, pia.p, c, 1, *powers, ss, rs)
                bound = TOL * _magnitude(pv, pia, c, powers, ss) * rs
                assert abs(got[c - 1] - ref) <= bound, (powers, ss, c, got[c - 1], ref)
            if counts[1] == 0:
                assert got[1] == 0.0
    # a sub-range and a single cell give the same values as the full call
    full = hip.compute_mixed_moment(st, 1, (2, 1, 0))
    assert np.array_equal(hip.compute_mixed_moment(st, 1, (2, 1, 0), cells=(3, 6)), full[2:6])
    assert hip.compute_mixed_moment(st, 1, (2, 1, 0), cells=5) == full[4]
    st.close()


def test_mixed_moment_dropin(hip, mzo):
    """the reference's signature: compute_mixed_moment(particles, pia, cell,
    species, powers; sum_scaler, res_scaler)"""
    D = hip.dropin
    counts = [30, 70, 11]
    pv, pia = _grouped_cells(mzo, counts, seed=4)
    st = hip.DeviceStore(len(counts), 1, sum(counts) + 16)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    particles = [D.DeviceParticleVector(st, 1)]
    dpia = D.DevicePIA(st)
    for c in (1, 2, 3):
        got = D.compute_mixed_moment(particles, dpia, c, 1, [1, 0, 2], sum_scaler=2.0, res_scaler=0.5)
        ref = mzo.lib().mzo_compute_mixed_moment(pv.p, pia.p, c, 1, 1, 0, 2, 2.0, 0.5)
        assert abs(got - ref) <= TOL * _magnitude(pv, pia, c, (1, 0, 2), 2.0) * 0.5
    st.close()
