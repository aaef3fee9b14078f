"""Grid N:2 merging on the device (mz_merge_grid) against the oracle's
merge_grid_based!: index, PIA and free-id buffer bit for bit, weights,
velocities and positions to 1e-15 (the device forms every velocity cell's
sums in particle order, as the oracle does), with extents from the props
and explicit ones, with and without clamping to the grid, over one- and
two-group cells; and the conservation laws at full size."""
import ctypes as C

import numpy as np
import pytest

import helpers
from test_gpu_roulette import _grouped_cells

pytestmark = pytest.mark.gpu

RTOL = 1e-15


def _oracle(mzo, pv, pia, threshold, mg, ext=None, props=None, grid=None, seed=1234, tag=0):
    r = mzo.RNG(seed=seed, tag=tag)
    for c in range(1, pia.n_cells + 1):
        if pia.ind(c).n_local > threshold:
            if ext is not None:
                mg.extent(*ext)
            else:
                mg.extent_props(props.T[0, c - 1], props.v[0, c - 1], helpers.MASS_AR)
            mzo.lib().mzo_merge_grid_based(r.ref, mg.p, pv.p, pia.p, c, 1, C.byref(grid) if grid is not None else None)
    mzo.lib().mzo_squash_pia(pv.p, pia.p, 1)


@pytest.mark.parametrize("N,mult,ext,clamp", [((2, 2, 2), 3.5, None, False), ((4, 4, 4), 2.0, None, True),
                                              ((4, 8, 2), 1.0, ((-800.0, 900.0), (-700.0, 700.0), (-500.0, 600.0)),
                                               False),
                                              ((1, 1, 1), 5.0, None, False)])
def test_grid_merge_vs_oracle(hip, mzo, N, mult, ext, clamp):
    counts = [0, 5, 40, 301, 900, 2000, 1500, 12, 777]
    nc = len(counts)
    pv, pia = _grouped_cells(mzo, counts, seed=21)
    st = hip.DeviceStore(nc, 1, sum(counts) + 512)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    st.set_grid(hip.Grid1DUniform(1e-6 * nc, nc))
    st.set_seed(1234)
    sp = [hip.Species("Ar", helpers.MASS_AR)]
    pp = hip.compute_props(st, sp)
    ref = mzo.Props(nc, 1)
    mzo.compute_props([pv], pia, [helpers.MASS_AR], ref)
    assert np.allclose(pp.T, ref.T, rtol=1e-12)
    tag = (9 << 4) | 3
    mg = hip.GridN2Merge(*N, extent_multiplier=mult)
    merged = hip.merge_grid_based(st, 1, 30, mg, species_data=sp, extents=ext, grid_clamp=clamp, tag=tag)
    assert merged == sum(1 for k in counts if k > 30)
    omg = mzo.GridMerge(*N, mult)
    # the oracle takes the extents from the device's props, so both grids are identical
    _oracle(mzo, pv, pia, 30, omg, ext=ext, props=pp, grid=mzo.grid1d(1e-6 * nc, nc) if clamp else None, tag=tag)
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=RTOL)
    st.close()


def test_grid_merge_conserves_at_size(hip, mzo):
    """300 cells of 300..2,000 two-group particles merged on a 4 x 4 x 4
    grid at 3.5 thermal speeds: every cell keeps its weight, momentum and
    energy to 1e-12, and holds at most 2 (4^3 + 8) particles."""
    rng = np.random.default_rng(6)
    counts = [int(k) for k in rng.integers(300, 2000, 300)]
    pv, pia = _grouped_cells(mzo, counts, seed=8)
    nc = len(counts)
    st = hip.DeviceStore(nc, 1, sum(counts) + 512)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    st.set_seed(3)
    sp = [hip.Species("Ar", helpers.MASS_AR)]
    hip.compute_props(st, sp)
    before = st.download_species(1)
    merged = hip.merge_grid_based(st, 1, 200, hip.GridN2Merge(4, extent_multiplier=3.5), species_data=sp, tag=77)
    assert merged == nc
    got = st.download_species(1)
    assert got["err"] == 0 and got["n_total"] < sum(counts)

    def sel(row):
        return np.concatenate([np.arange(row[1] - 1, row[2]) if row[3] > 0 else np.zeros(0, int),
                               np.arange(row[4] - 1, row[5]) if row[6] > 0 else np.zeros(0, int)]).astype(int)
    for c in range(nc):
        ib, ia = sel(before["pia"][c]), sel(got["pia"][c])
        assert len(ia) <= 2 * (64 + 8)
        wb, wa = before["w"][ib], got["w"][ia]
        vb, va = before["v"][ib], got["v"][ia]
        assert abs(wa.sum() / wb.sum() - 1) < 1e-13
        pb, pa = (wb[:, None] * vb).sum(0), (wa[:, None] * va).sum(0)
        assert np.all(np.abs(pa - pb) <= 1e-12 * (wb[:, None] * np.abs(vb)).sum(0))
        eb, ea = (wb * (vb ** 2).sum(1)).sum(), (wa * (va ** 2).sum(1)).sum()
        assert abs(ea / eb - 1) < 1e-12
    st.close()


def _grid_case(hip, mzo, counts, N, seed):
    pv, pia = _grouped_cells(mzo, counts, seed=seed)
    nc = len(counts)
    st = hip.DeviceStore(nc, 1, sum(counts) + 512)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    st.set_seed(1234)
    sp = [hip.Species("Ar", helpers.MASS_AR)]
    pp = hip.compute_props(st, sp)
    return pv, pia, st, sp, pp


def test_grid_merge_at_kernel_limits(hip, mzo):
    """The kernel's largest case: a 16 x 16 x 16 velocity grid (4,104 bins
    with the 8 outer ones, its bin limit) and a cell of exactly 2,048
    particles (its particle limit, the bitonic sort over 2,048 keys), next
    to smaller cells: bit for bit against the oracle."""
    counts = [2048, 0, 700, 2047, 33]
    pv, pia, st, sp, pp = _grid_case(hip, mzo, counts, (16, 16, 16), seed=31)
    tag = (5 << 4) | 2
    merged = hip.merge_grid_based(st, 1, 30, hip.GridN2Merge(16, 16, 16, extent_multiplier=3.0),
                                  species_data=sp, tag=tag)
    assert merged == 4
    _oracle(mzo, pv, pia, 30, mzo.GridMerge(16, 16, 16, 3.0), props=pp, tag=tag)
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=RTOL)
    st.close()


def test_grid_merge_rejects_oversize_cell(hip, mzo):
    """A cell of 2,049 particles is beyond the kernel: the call flags the
    merge-size error and leaves that cell as it was, while the other cells
    merge exactly as the oracle merges them."""
    counts = [2049, 400, 90]
    pv, pia, st, sp, pp = _grid_case(hip, mzo, counts, (4, 4, 4), seed=32)
    tag = 11
    before = st.download_species(1)
    hip.merge_grid_based(st, 1, 30, hip.GridN2Merge(4, extent_multiplier=3.5), species_data=sp, squash=False,
                         tag=tag)
    got = st.download_species(1)
    assert got["err"] & 32
    assert np.array_equal(got["pia"][0], before["pia"][0])
    s1, e1 = int(before["pia"][0][1]), int(before["pia"][0][2])
    assert np.array_equal(got["w"][s1 - 1:e1], before["w"][s1 - 1:e1])
    assert np.array_equal(got["v"][s1 - 1:e1], before["v"][s1 - 1:e1])
    r = mzo.RNG(seed=1234, tag=tag)
    omg = mzo.GridMerge(4, 4, 4, 3.5)
    for c in (2, 3):
        omg.extent_props(pp.T[0, c - 1], pp.v[0, c - 1], helpers.MASS_AR)
        mzo.lib().mzo_merge_grid_based(r.ref, omg.p, pv.p, pia.p, c, 1, None)
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=RTOL)
    st.close()
