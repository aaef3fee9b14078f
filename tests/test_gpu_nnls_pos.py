"""NNLS merging with spatial moments (NNLSMerge's multi_index_moments_pos)
and random-pair columns on the device against the oracle.

* the reference's spatial-moment case (test/test_nnls_merging.jl:331-408)
  through mz_merge_nnls: bit for bit with the oracle, and the reference's
  known answers on the merged particles;
* many cells with spatial moments and random pairs (Philox draws from each
  cell's stream): the reference-layout kernel bit for bit with the oracle,
  the LDS-resident kernels (FP64 and mixed) conserving every matched moment
  to 1e-12 relative where they merge.  A random pair's spatial rows take the
  velocity multi-indices (merging_nnls.jl:1023-1026), so the mean position
  is conserved only without random pairs; with them the velocity moments
  are checked."""
import numpy as np
import pytest

import helpers
from test_gpu_merge import _bimaxwell_cells, _mim
from test_oracle_kat import nnls_particles2

pytestmark = pytest.mark.gpu

POS1 = [[0, 0, 1], [1, 0, 0], [0, 1, 0]]


def test_nnls_spatial_moments_reference_case(hip, mzo):
    parts = nnls_particles2(mult=2.0, mult2=2.0)
    pv = mzo.ParticleVector(16)
    for i, (w, v, x) in enumerate(parts, start=1):
        pv.pop_to(i)
        pv.set(i, w, v, x)
    pia = mzo.PIA(1, 1)
    pia.set(1, 1, 16, 1, 16, 16)
    pia.set_n_total(1, 16)
    st = hip.DeviceStore(1, 1, 64)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    status = hip.merge_nnls_based(st, 1, threshold=1, multi_index_moments=[], multi_index_moments_pos=POS1,
                                  squash=False)
    nm = mzo.NNLSMerge([], mim_pos=POS1)
    res = nm.merge(mzo.RNG(seed=1234), pv, pia, 1, 1)
    assert status[0] == res == 1
    got = st.download_species(1)
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=0.0)
    n = got["n_total"]
    assert n < 16
    w0 = np.array([p[0] for p in parts])
    v0 = np.array([p[1] for p in parts])
    x0 = np.array([p[2] for p in parts])
    tw = w0.sum()
    w, v, x = got["w"][:n], got["v"][:n], got["x"][:n]
    assert abs(w.sum() - tw) < 3.75e-15
    assert np.max(np.abs((w[:, None] * v).sum(0) / w.sum() - (w0[:, None] * v0).sum(0) / tw)) < 1e-15
    assert np.max(np.abs((w[:, None] * x).sum(0) / w.sum() - (w0[:, None] * x0).sum(0) / tw)) < 1e-15
    st.close()


def _spatial_moments(w, x, rows, x_mean):
    d = x - x_mean
    t = np.array([d[:, 0] ** a * d[:, 1] ** b * d[:, 2] ** c for (a, b, c) in rows])
    return (t * w).sum(1), (np.abs(t) * w).sum(1)


@pytest.mark.parametrize("n_pairs", [0, 6])
def test_nnls_spatial_and_pairs_batch_vs_oracle(hip, mzo, n_pairs):
    rng = np.random.default_rng(17 + n_pairs)
    counts = [int(k) for k in rng.integers(40, 160, size=16)]
    nc, n = len(counts), sum(counts)
    pos = POS1 + [[1, 1, 0], [2, 0, 0]]
    mim = _mim(3)
    tag = (5 << 4) | 3
    pv, pia = _bimaxwell_cells(mzo, counts, rng)
    st = hip.DeviceStore(nc, 1, n + 256)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    st.set_seed(777)
    status = hip.merge_nnls_based(st, 1, threshold=30, multi_index_moments=mim, multi_index_moments_pos=pos,
                                  n_rand_pairs=n_pairs, tag=tag, squash=False)
    ref = []
    for c in range(1, nc + 1):
        nm = mzo.NNLSMerge(mim, mim_pos=pos)
        ref.append(nm.merge(mzo.RNG(seed=777, tag=tag), pv, pia, c, 1, n_rand_pairs=n_pairs))
    assert list(status) == ref
    assert sum(s == 1 for s in ref) >= nc // 2
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=0.0)
    st.close()


@pytest.mark.parametrize("method,n_pairs", [("lds", 0), ("mixed", 0), ("lds", 8), ("mixed", 8)])
def test_nnls_spatial_and_pairs_lds_conserve(hip, mzo, method, n_pairs):
    rng = np.random.default_rng(23)
    counts = [int(k) for k in rng.integers(60, 300, size=12)]
    nc, n = len(counts), sum(counts)
    pos = POS1
    mim = _mim(3)
    pv, pia = _bimaxwell_cells(mzo, counts, rng)
    st = hip.DeviceStore(nc, 1, n + 256)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    st.set_seed(5)
    before = st.download_species(1)
    status = hip.merge_nnls_candidates(st, 1, 30, mim, 200, method=method, multi_index_moments_pos=pos,
                                       n_rand_pairs=n_pairs, tag=(9 << 4) | 3, squash=False)
    assert sum(s == 1 for s in status) >= nc // 2
    got = st.download_species(1)
    assert got["err"] == 0
    rows = [(0, 0, 0)] + [tuple(t) for t in mim]
    for c in range(nc):
        if status[c] != 1:
            continue
        b, a = before["pia"][c], got["pia"][c]
        wb, vb, xb = (before[k][b[1] - 1:b[2]] for k in ("w", "v", "x"))
        wa, va, xa = (got[k][a[1] - 1:a[2]] for k in ("w", "v", "x"))
        assert len(wa) < len(wb)
        mb, sb = helpers_moments(wb, vb, rows)
        ma, _ = helpers_moments(wa, va, rows)
        assert np.max(np.abs(ma - mb) / sb) < 1e-12
        if n_pairs:
            continue
        xm = (wb[:, None] * xb).sum(0) / wb.sum()
        assert np.max(np.abs((wa[:, None] * xa).sum(0) / wa.sum() - xm)) < 1e-12 * max(1.0, np.abs(xm).max())
    st.close()


def helpers_moments(w, v, rows):
    t = np.array([v[:, 0] ** a * v[:, 1] ** b * v[:, 2] ** c for (a, b, c) in rows])
    return (t * w).sum(1), (np.abs(t) * w).sum(1)
