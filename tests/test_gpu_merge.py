"""GPU parity of the merging kernels against the CPU oracle.

Octree N:2: bin assignment follows from exact arithmetic, so the merged
particle count, the positions the new particles occupy, the PIA, the ids and
the free-id LIFO must match bit for bit; values to 1e-12; mass, momentum and
energy are conserved to 1e-12 relative."""
import ctypes as C

import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def _var_weight_cells(mzo, counts, seed, L_cell=1e-6):
    rng = np.random.default_rng(seed)
    n = int(sum(counts))
    nc = len(counts)
    pv = mzo.ParticleVector(n)
    pia = mzo.PIA(nc, 1)
    vs = np.sqrt(helpers.KB * 300.0 / helpers.MASS_AR)
    pos = 1
    for c, k in enumerate(counts, start=1):
        for j in range(k):
            pv.pop_to(pos + j)
            w = 1e14 * rng.lognormal(0.0, 1.0) / np.exp(0.5)
            v = rng.normal(0, vs, 3) + np.array([0.0, 500.0 * (2 * c / nc - 1), 0.0])
            x = [(c - 1 + rng.uniform()) * L_cell, rng.uniform(), rng.uniform()]
            pv.set(pos + j, w, v, x)
        if k:
            pia.set(c, 1, k, pos, pos + k - 1, k)
        pos += k
    pia.set_n_total(1, n)
    return pv, pia


def _moments(w, v):
    m = w.sum()
    p = (w[:, None] * v).sum(0)
    e = (w * (v ** 2).sum(1)).sum()
    return m, p, e


@pytest.mark.parametrize("split,init,bounds,clamp", [(1, 1, 1, True), (1, 2, 1, False), (1, 3, 2, False),
                                                     (2, 1, 1, False)])
def test_octree_merge_vs_oracle(hip, mzo, split, init, bounds, clamp):
    counts = [120, 450, 401, 30, 800, 1500, 402, 0, 2000, 399]
    nc = len(counts)
    pv, pia = _var_weight_cells(mzo, counts, seed=31 + split + init)
    st = hip.DeviceStore(nc, 1, sum(counts) + 512)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    grid = hip.Grid1DUniform(1e-6 * nc, nc)
    st.set_grid(grid)
    st.set_seed(4242)
    tag = (17 << 4) | 3
    before = st.download_species(1)
    n = before["n_total"]
    m0, p0, e0 = _moments(before["w"][:n], before["v"][:n])
    n_merged = hip.merge_octree_N2_based(st, 1, threshold=400, target_np=64, split=split, init_bin_bounds=init,
                                         bin_bounds_compute=bounds, max_depth=8, grid_clamp=clamp, squash=False,
                                         tag=tag)
    assert n_merged == sum(1 for k in counts if k > 400)
    og = mzo.grid1d(1e-6 * nc, nc)
    oc = mzo.Octree(split=split, init_bin_bounds=init, bin_bounds_compute=bounds, max_Nbins=4096, max_depth=8)
    r = mzo.RNG(seed=4242, tag=tag)
    for c in range(1, nc + 1):
        if pia.ind(c).n_local > 400:
            if split == 2:
                # the mean split carries vel_middle between calls on one
                # OctreeN2Merge (merging_octree_N2.jl:431-439); the device
                # merges cells independently, i.e. with a fresh octree per cell
                oc = mzo.Octree(split=split, init_bin_bounds=init, bin_bounds_compute=bounds, max_depth=8)
            mzo.lib().mzo_merge_octree_N2_based(r.ref, oc.p, pv.p, pia.p, c, 1, 64, C.byref(og) if clamp else None)
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=1e-12)
    # conservation over the used positions
    tab = got["pia"]
    sel = np.concatenate([np.arange(r_[1] - 1, r_[2]) for r_ in tab if r_[3] > 0] +
                         [np.arange(r_[4] - 1, r_[5]) for r_ in tab if r_[6] > 0])
    m1, p1, e1 = _moments(got["w"][sel], got["v"][sel])
    assert abs(m1 / m0 - 1) < 1e-12
    assert np.all(np.abs(p1 - p0) <= 1e-12 * np.abs(p0).max())
    assert abs(e1 / e0 - 1) < 1e-12
    # squash afterwards, as the reference's driver does
    hip.squash_pia(st, 1)
    mzo.lib().mzo_squash_pia(pv.p, pia.p, 1)
    helpers.assert_state_matches(mzo, st.download_species(1), pv, pia, rtol=1e-12)
    st.close()


def _octree_vs_oracle(hip, mzo, counts, thresholds, targets, seed, split=1, init=1, bounds=1, per_cell=False):
    nc = len(counts)
    pv, pia = _var_weight_cells(mzo, counts, seed=seed)
    st = hip.DeviceStore(nc, 1, sum(counts) + 512)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    st.set_grid(hip.Grid1DUniform(1e-6 * nc, nc))
    st.set_seed(977)
    tag = (5 << 4) | 3
    before = st.download_species(1)
    n = before["n_total"]
    m0, p0, e0 = _moments(before["w"][:n], before["v"][:n])
    kw = dict(split=split, init_bin_bounds=init, bin_bounds_compute=bounds, max_Nbins=4096, max_depth=10,
              squash=False, tag=tag)
    if per_cell:
        n_merged = hip.merge_octree_targets(st, 1, thresholds, targets, **kw)
    else:
        n_merged = hip.merge_octree_N2_based(st, 1, threshold=thresholds[0], target_np=targets[0], **kw)
    assert n_merged == sum(1 for k, t in zip(counts, thresholds) if k > t)
    r = mzo.RNG(seed=977, tag=tag)
    for c in range(1, nc + 1):
        if pia.ind(c).n_local > thresholds[c - 1]:
            oc = mzo.Octree(split=split, init_bin_bounds=init, bin_bounds_compute=bounds, max_Nbins=4096,
                            max_depth=10)
            mzo.lib().mzo_merge_octree_N2_based(r.ref, oc.p, pv.p, pia.p, c, 1, int(targets[c - 1]), None)
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=1e-12)
    tab = got["pia"]
    sel = np.concatenate([np.arange(r_[1] - 1, r_[2]) for r_ in tab if r_[3] > 0] +
                         [np.arange(r_[4] - 1, r_[5]) for r_ in tab if r_[6] > 0])
    m1, p1, e1 = _moments(got["w"][sel], got["v"][sel])
    assert abs(m1 / m0 - 1) < 1e-12
    assert np.all(np.abs(p1 - p0) <= 1e-12 * np.abs(p0).max())
    assert abs(e1 / e0 - 1) < 1e-12
    st.close()


@pytest.mark.parametrize("target,counts,split,bounds", [
    (64, [3000, 5000, 2049, 900, 2100, 401], 1, 1),
    (1500, [1800, 2048, 4000, 600, 12000], 1, 1),
    (700, [2500, 1900, 30000], 2, 2)])
def test_octree_merge_large_cells_vs_oracle(hip, mzo, target, counts, split, bounds):
    """Cells above 2048 particles, and cells whose split loop needs more than
    512 bins (targets of 700 and 1500 with max_Nbins 4096), go to the
    device-memory path: the same results as the oracle's per-cell merges."""
    nc = len(counts)
    _octree_vs_oracle(hip, mzo, counts, [400] * nc, [target] * nc, seed=61 + target, split=split, bounds=bounds)


def test_octree_merge_per_cell_targets_vs_oracle(hip, mzo):
    """The batched form of "merge cell c above thresholds[c] down to
    targets[c]" against the oracle's per-cell calls with each cell's target."""
    counts = [500, 2500, 300, 1200, 3100, 90, 6000]
    thresholds = [400, 2000, 350, 1000, 3000, 60, 4500]
    targets = [100, 1500, 64, 700, 2000, 30, 3000]
    _octree_vs_oracle(hip, mzo, counts, thresholds, targets, seed=73, per_cell=True)


def test_octree_merge_reference_kat(hip, mzo):
    """test/test_octree_merging.jl:76-163 on the device: 8 octants x 3
    particles merged to 16 (one split, depth 1), then to 2 (no split)."""
    parts = []
    for o in range(1, 9):
        for dv in (-0.5, 0.5, 0.0):
            val = 9.0 - o + dv
            vz = val if o >= 5 else -val
            vx = -val if o % 2 == 1 else val
            vy = val if o in (3, 4, 7, 8) else -val
            parts.append((float(o), [vx, vy, vz], [1.0, -10.0, 3.0]))
    pv = mzo.ParticleVector(24)
    for i, (w, v, x) in enumerate(parts, start=1):
        pv.pop_to(i)
        pv.set(i, w, v, x)
    pia = mzo.PIA(1, 1)
    pia.set(1, 1, 24, 1, 24, 24)
    pia.set_n_total(1, 24)
    st = hip.DeviceStore(1, 1, 64)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    st.set_seed(1234)
    sp = [hip.Species("Ar", helpers.MASS_AR)]
    p0 = hip.compute_props(st, sp)
    n0, T0, v0 = p0.n[0, 0], p0.T[0, 0], p0.v[0, 0].copy()
    hip.merge_octree_N2_based(st, 1, threshold=2, target_np=16, init_bin_bounds=hip.OCTREE_INIT_C, squash=False)
    got = st.download_species(1)
    assert got["n_total"] == 16 and got["pia"][0][0] == 16
    p1 = hip.compute_props(st, sp)
    assert abs(p1.n[0, 0] - n0) < np.finfo(float).eps * n0 and abs(p1.T[0, 0] - T0) < 1e-14
    assert np.all(np.abs(p1.v[0, 0] - v0) < 1e-14)
    hip.merge_octree_N2_based(st, 1, threshold=2, target_np=2, init_bin_bounds=hip.OCTREE_INIT_C, squash=False)
    got = st.download_species(1)
    assert got["n_total"] == 2
    assert got["w"][0] == 0.5 * sum(3 * i for i in range(1, 9)) == got["w"][1]
    st.close()


def _mim(order):
    return [t for i in range(1, order + 1) for t in (
        [(a, b, i - a - b) for a in range(i + 1) for b in range(i - a + 1)])]


def test_nnls_merge_reference_case(hip, mzo):
    """test/test_nnls_merging.jl:104-147 on the device: 50 particles, moments
    to order 4, :vref scaling; the merged particles equal the oracle's (same
    pivots) and conserve mass, momentum and energy."""
    parts = [(1.0, [i, 5 - 2.0 * i, -100.0 + 0.25 * i * i], [-0.5 * i, 27 + i, 400.0 - 0.5 * i * i])
             for i in range(1, 51)]
    pv = mzo.ParticleVector(50)
    for i, (w, v, x) in enumerate(parts, start=1):
        pv.pop_to(i)
        pv.set(i, w, v, x)
    pia = mzo.PIA(1, 1)
    pia.set(1, 1, 50, 1, 50, 50)
    pia.set_n_total(1, 50)
    st = hip.DeviceStore(1, 1, 128)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    vref = np.sqrt(2 * helpers.KB * 300.0 / helpers.MASS_AR)
    mim = _mim(4)
    status = hip.merge_nnls_based(st, 1, threshold=1, multi_index_moments=mim, vref=vref, scaling="vref",
                                  squash=False)
    nm = mzo.NNLSMerge(mim)
    nm.s.reset_x0 = 1
    res = nm.merge(mzo.RNG(seed=1), pv, pia, 1, 1, vref=vref, scaling="vref")
    assert status[0] == res == 1
    got = st.download_species(1)
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=1e-13)
    n = got["n_total"]
    assert n < 50
    w, v = got["w"][:n], got["v"][:n]
    ws, vs = np.array([p[0] for p in parts]), np.array([p[1] for p in parts])
    assert abs(w.sum() - ws.sum()) < 1.5e-14 * ws.sum()
    assert np.all(np.abs((w[:, None] * v).sum(0) - (ws[:, None] * vs).sum(0)) < 1e-12 * np.abs(vs).max() * ws.sum())
    assert abs((w * (v ** 2).sum(1)).sum() / (ws * (vs ** 2).sum(1)).sum() - 1) < 1e-12
    st.close()


def _bimaxwell_cells(mzo, counts, rng):
    nc = len(counts)
    n = sum(counts)
    pv = mzo.ParticleVector(n)
    pia = mzo.PIA(nc, 1)
    pos = 1
    for c, k in enumerate(counts, start=1):
        ratio = rng.uniform(0.5, 2.0)
        sig = np.sqrt(helpers.KB * 300.0 / helpers.MASS_AR) * np.array([np.sqrt(ratio), 1.0, 1.0])
        drift = rng.normal(size=3)
        drift *= rng.uniform(0, 500.0) / np.linalg.norm(drift)
        for j in range(k):
            pv.pop_to(pos + j)
            pv.set(pos + j, 1e12 * rng.lognormal(0, 0.5), drift + sig * rng.normal(size=3),
                   [rng.uniform(), rng.uniform(), rng.uniform()])
        pia.set(c, 1, k, pos, pos + k - 1, k)
        pos += k
    pia.set_n_total(1, n)
    return pv, pia


@pytest.mark.parametrize("scaling", ["variance", "vref"])
def test_nnls_merge_batch_vs_oracle(hip, mzo, scaling):
    """Many cells of bi-Maxwellian particles with lognormal weights merged in
    one launch; every cell's status and merged particles equal the oracle's
    sequential merges; ids pushed to the free-id LIFO in cell order."""
    rng = np.random.default_rng(9)
    counts = [int(k) for k in rng.integers(40, 400, size=24)]
    counts[3] = 20                      # below the threshold: untouched
    nc = len(counts)
    n = sum(counts)
    pv, pia = _bimaxwell_cells(mzo, counts, rng)
    st = hip.DeviceStore(nc, 1, n + 256)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    vref = np.sqrt(2 * helpers.KB * 300.0 / helpers.MASS_AR)
    mim = _mim(4)
    status = hip.merge_nnls_based(st, 1, threshold=30, multi_index_moments=mim, vref=vref, scaling=scaling,
                                  squash=False)
    nm = mzo.NNLSMerge(mim)
    nm.s.reset_x0 = 1
    ref_status = []
    for c in range(1, nc + 1):
        if pia.ind(c).n_local > 30:
            ref_status.append(nm.merge(mzo.RNG(seed=1), pv, pia, c, 1, vref=vref, scaling=scaling))
        else:
            ref_status.append(0)
    assert list(status) == ref_status
    assert sum(s == 1 for s in ref_status) > 15
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=1e-13)
    st.close()


def _cell_moments(w, v, rows):
    """Weighted moments sum w vx^a vy^b vz^c and their scale sum w |.|."""
    t = np.array([v[:, 0] ** a * v[:, 1] ** b * v[:, 2] ** c for (a, b, c) in rows])
    return (t * w).sum(1), (np.abs(t) * w).sum(1)


def _nnls_rows():
    return [(0, 0, 0)] + _mim(4)


@pytest.mark.parametrize("max_pcols", [239, 100, 20])
def test_nnls_candidates_reference_layout_vs_oracle(hip, mzo, max_pcols):
    """Columns cut to the first max_pcols particles (plus the 17 fictitious
    ones), the moments of all particles matched: the reference-layout kernel
    equals the oracle's merge_nnls_based! on the same column set bit for bit,
    for merged and failed cells alike (20 columns are mostly infeasible)."""
    rng = np.random.default_rng(21)
    counts = [int(k) for k in rng.integers(40, 600, size=20)]
    counts[5] = 25
    nc, n = len(counts), sum(counts)
    pv, pia = _bimaxwell_cells(mzo, counts, rng)
    st = hip.DeviceStore(nc, 1, n + 256)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    mim = _mim(4)
    status = hip.merge_nnls_candidates(st, 1, 30, mim, max_pcols, method="reference", squash=False)
    nm = mzo.NNLSMerge(mim)
    nm.s.reset_x0 = 1
    ref_status = [nm.merge(mzo.RNG(seed=1), pv, pia, c, 1, max_pcols=max_pcols) if pia.ind(c).n_local > 30 else 0
                  for c in range(1, nc + 1)]
    assert list(status) == ref_status
    if max_pcols == 239:
        assert sum(s == 1 for s in ref_status) >= 15
    if max_pcols == 20:
        assert sum(s == -1 for s in ref_status) >= 10
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=1e-13)
    st.close()


@pytest.mark.parametrize("scaling", ["variance", "vref"])
def test_nnls_candidates_lds_vs_oracle(hip, mzo, scaling):
    """The LDS-resident method on the same column set: every cell's status
    equals the oracle's, a merged cell keeps the oracle's particle count, its
    new particles match the oracle's to 1e-9, and all 35 moments of the cell
    (mass, momentum and energy among them) are preserved to 1e-12 of their
    scale; failed and skipped cells are untouched."""
    rng = np.random.default_rng(22)
    counts = [int(k) for k in rng.integers(40, 800, size=24)]
    counts[2] = 25
    nc, n = len(counts), sum(counts)
    pv, pia = _bimaxwell_cells(mzo, counts, rng)
    st = hip.DeviceStore(nc, 1, n + 256)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    pre = st.download_species(1)
    vref = np.sqrt(2 * helpers.KB * 300.0 / helpers.MASS_AR)
    mim = _mim(4)
    status = hip.merge_nnls_candidates(st, 1, 30, mim, 239, method="lds", vref=vref, scaling=scaling,
                                       squash=False)
    nm = mzo.NNLSMerge(mim)
    nm.s.reset_x0 = 1
    ref_status = [nm.merge(mzo.RNG(seed=1), pv, pia, c, 1, vref=vref, scaling=scaling, max_pcols=239)
                  if pia.ind(c).n_local > 30 else 0 for c in range(1, nc + 1)]
    assert list(status) == ref_status
    assert sum(s == 1 for s in ref_status) >= 18
    got = st.download_species(1)
    assert got["err"] == 0
    ref = helpers.oracle_upload_arrays(mzo, pv, pia)
    rows = _nnls_rows()
    for c in range(nc):
        r0, r1, r2 = pre["pia"][c], got["pia"][c], ref["pia"][c]
        assert list(r1) == list(r2), c
        a0, b0 = r0[1] - 1, r0[2]
        w_pre, v_pre = pre["w"][a0:b0], pre["v"][a0:b0]
        a1, b1 = r1[1] - 1, r1[2]
        w, v = got["w"][a1:b1], got["v"][a1:b1]
        if status[c] != 1:
            assert np.array_equal(w, w_pre) and np.array_equal(v, v_pre)
            continue
        assert np.allclose(w, ref["w"][a1:b1], rtol=1e-9, atol=0)
        assert np.allclose(v, ref["v"][a1:b1], rtol=1e-9, atol=1e-9 * np.abs(v_pre).max())
        m0, scale = _cell_moments(w_pre, v_pre, rows)
        m1, _ = _cell_moments(w, v, rows)
        assert np.max(np.abs(m1 - m0) / scale) < 1e-12, c
    st.close()


@pytest.mark.parametrize("scaling", ["variance", "vref"])
def test_nnls_candidates_mixed_vs_oracle(hip, mzo, scaling):
    """The mixed method (active-set search on the FP32-stored matrix, final
    passive set re-solved in FP64, unconfirmed cells re-run in FP64): every
    cell's status equals the oracle's, and every merged cell preserves all 35
    moments to 1e-12 of their scale; failed and skipped cells are untouched.
    The merged particle set may differ from the oracle's (another non-negative
    solution of the same moment system)."""
    rng = np.random.default_rng(24)
    counts = [int(k) for k in rng.integers(40, 800, size=24)]
    counts[4] = 25
    nc, n = len(counts), sum(counts)
    pv, pia = _bimaxwell_cells(mzo, counts, rng)
    st = hip.DeviceStore(nc, 1, n + 256)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    pre = st.download_species(1)
    vref = np.sqrt(2 * helpers.KB * 300.0 / helpers.MASS_AR)
    mim = _mim(4)
    status = hip.merge_nnls_candidates(st, 1, 30, mim, 239, method="mixed", vref=vref, scaling=scaling,
                                       squash=False)
    tried, redone = st.nnls_info()
    assert tried == sum(k > 30 for k in counts) and 0 <= redone <= tried
    nm = mzo.NNLSMerge(mim)
    nm.s.reset_x0 = 1
    ref_status = [nm.merge(mzo.RNG(seed=1), pv, pia, c, 1, vref=vref, scaling=scaling, max_pcols=239)
                  if pia.ind(c).n_local > 30 else 0 for c in range(1, nc + 1)]
    assert list(status) == ref_status
    assert sum(s == 1 for s in ref_status) >= 18
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
