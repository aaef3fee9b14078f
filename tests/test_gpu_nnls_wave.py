"""Candidate NNLS merging with one wavefront per cell (method "wave": the
active-set search on the FP32 matrix held in registers, the final passive
set solved again in FP64, cells it cannot confirm re-run by the FP64 LDS
method) against the oracle, and the 64-column candidate set (47 particles +
17 fictitious ones) for the reference-layout, mixed and wave methods.

A merged cell must keep all 35 moments (mass, momentum and energy among
them) to 1e-12 of their scale; a cell's status must equal the oracle's
(merge_nnls_based! on the same column set); failed and skipped cells are
left untouched.  The merged particles themselves may differ from the
oracle's (another non-negative solution of the same moment system)."""
import numpy as np
import pytest

import helpers
from test_gpu_merge import _bimaxwell_cells, _cell_moments, _mim, _nnls_rows

pytestmark = pytest.mark.gpu


def _check_fp32_search(hip, mzo, method, scaling, max_pcols, seed, min_merged):
    rng = np.random.default_rng(seed)
    counts = [int(k) for k in rng.integers(40, 800, size=24)]
    counts[4] = 25
    nc, n = len(counts), sum(counts)
    pv, pia = _bimaxwell_cells(mzo, counts, rng)
    st = hip.DeviceStore(nc, 1, n + 256)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    pre = st.download_species(1)
    vref = np.sqrt(2 * helpers.KB * 300.0 / helpers.MASS_AR)
    mim = _mim(4)
    status = hip.merge_nnls_candidates(st, 1, 30, mim, max_pcols, method=method, vref=vref, scaling=scaling,
                                       squash=False)
    tried, redone = st.nnls_info()
    assert tried == sum(k > 30 for k in counts) and 0 <= redone <= tried
    nm = mzo.NNLSMerge(mim)
    nm.s.reset_x0 = 1
    ref_status = [nm.merge(mzo.RNG(seed=1), pv, pia, c, 1, vref=vref, scaling=scaling, max_pcols=max_pcols)
                  if pia.ind(c).n_local > 30 else 0 for c in range(1, nc + 1)]
    assert list(status) == ref_status
    assert sum(s == 1 for s in ref_status) >= min_merged
    got = st.download_species(1)
    assert got["err"] == 0
    rows = _nnls_rows()
    for c in range(nc):
        r0, r1 = pre["pia"][c], got["pia"][c]
        a0, b0 = r0[1] - 1, r0[2]
        w_pre, v_pre = pre["w"][a0:b0], pre["v"][a0:b0]
        a1, b1 = r1[1] - 1, r1[2]
        w, v = got["w"][a1:b1], got["v"][a1:b1]
        if status[c] != 1:
            assert list(r1) == list(r0)
            assert np.array_equal(w, w_pre) and np.array_equal(v, v_pre)
            continue
        assert r1[0] < r0[0] and np.all(w > 0)
        m0, scale = _cell_moments(w_pre, v_pre, rows)
        m1, _ = _cell_moments(w, v, rows)
        assert np.max(np.abs(m1 - m0) / scale) < 1e-12, c
    st.close()
    return redone


@pytest.mark.parametrize("scaling", ["variance", "vref"])
def test_nnls_candidates_wave_vs_oracle(hip, mzo, scaling):
    redone = _check_fp32_search(hip, mzo, "wave", scaling, 239, 24, 18)
    print("cells re-run in FP64:", redone)


@pytest.mark.parametrize("method", ["mixed", "wave"])
def test_nnls_64_columns_fp32_search(hip, mzo, method):
    """47 particle columns + 17 fictitious ones: the 35-moment system is
    mostly infeasible, and every status still equals the oracle's."""
    _check_fp32_search(hip, mzo, method, "variance", 47, 25, 0)


def test_nnls_64_columns_reference_layout(hip, mzo):
    rng = np.random.default_rng(26)
    counts = [int(k) for k in rng.integers(40, 600, size=16)]
    nc, n = len(counts), sum(counts)
    pv, pia = _bimaxwell_cells(mzo, counts, rng)
    st = hip.DeviceStore(nc, 1, n + 256)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    mim = _mim(4)
    status = hip.merge_nnls_candidates(st, 1, 30, mim, 47, method="reference", squash=False)
    nm = mzo.NNLSMerge(mim)
    nm.s.reset_x0 = 1
    ref_status = [nm.merge(mzo.RNG(seed=1), pv, pia, c, 1, max_pcols=47) if pia.ind(c).n_local > 30 else 0
                  for c in range(1, nc + 1)]
    assert list(status) == ref_status
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=1e-13)
    st.close()


@pytest.mark.parametrize("n_pairs", [0, 8])
def test_nnls_wave_spatial_and_pairs_conserve(hip, mzo, n_pairs):
    """Spatial moments (the mean position) and random-pair columns through
    the wave method: velocity moments to 1e-12; the mean position to 1e-12
    without random pairs (a pair's spatial rows take the velocity
    multi-indices, merging_nnls.jl:1023-1026)."""
    rng = np.random.default_rng(23)
    counts = [int(k) for k in rng.integers(60, 300, size=12)]
    nc, n = len(counts), sum(counts)
    pos = [[0, 0, 1], [1, 0, 0], [0, 1, 0]]
    mim = _mim(3)
    pv, pia = _bimaxwell_cells(mzo, counts, rng)
    st = hip.DeviceStore(nc, 1, n + 256)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    st.set_seed(5)
    before = st.download_species(1)
    status = hip.merge_nnls_candidates(st, 1, 30, mim, 200, method="wave", multi_index_moments_pos=pos,
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
        mb, sb = _cell_moments(wb, vb, rows)
        ma, _ = _cell_moments(wa, va, rows)
        assert np.max(np.abs(ma - mb) / sb) < 1e-12
        if n_pairs:
            continue
        xm = (wb[:, None] * xb).sum(0) / wb.sum()
        assert np.max(np.abs((wa[:, None] * xa).sum(0) / wa.sum() - xm)) < 1e-12 * max(1.0, np.abs(xm).max())
    st.close()
