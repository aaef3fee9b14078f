"""Two checks of edge cases the other GPU tests leave open.

* The particle dump of a device store (write_netcdf(nc_filename, pv, pia,
  species, species_data), io.jl:842-892) on a layout whose cells keep a
  group 2 after all other cells' particles: w, v, x and cell of the file
  against the oracle's ParticleVector walked cell by cell, group 1 then
  group 2, as io.jl:864-890 walks it; the dimension's name and size.
* Grid merging at the kernel's limits: a 16 x 16 x 16 velocity grid with a
  cell of exactly 2,048 particles (the largest the LDS sort takes) against
  the oracle, and a cell of 2,049 particles refused with the size error,
  left untouched."""
import ctypes as C

import numpy as np
import pytest

import helpers
from test_gpu_roulette import _grouped_cells

pytestmark = pytest.mark.gpu


def test_particle_dump_two_group_layout(hip, mzo, tmp_path):
    from scipy.io import netcdf_file
    counts = [3, 0, 7, 5, 1, 9]
    pv, pia = _grouped_cells(mzo, counts, seed=4)
    nc = len(counts)
    st = hip.DeviceStore(nc, 1, 64)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    sp = [hip.Species("Ar", helpers.MASS_AR)]
    path = str(tmp_path / "dump2.nc")
    hip.write_netcdf(path, st, sp, species=1)
    w_ref, v_ref, x_ref, cell_ref = [], [], [], []
    for c in range(1, nc + 1):
        ind = pia.ind(c)
        pos = list(range(ind.start1, ind.end1 + 1)) if ind.n_group1 > 0 else []
        if ind.n_group2 > 0:
            pos += list(range(ind.start2, ind.end2 + 1))
        for i in pos:
            w, v, x = pv.get(i)
            w_ref.append(w)
            v_ref.append(v)
            x_ref.append(x)
            cell_ref.append(c)
    n = pia.n_total()
    assert len(w_ref) == n == sum(counts)
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.dimensions['nparticles_["Ar"]'] == n
        assert np.array_equal(f.variables["w"][:], np.array(w_ref))
        assert np.array_equal(f.variables["v"][:], np.array(v_ref).T)
        assert np.array_equal(f.variables["x"][:], np.array(x_ref).T)
        assert list(f.variables["cell"][:]) == cell_ref
    st.close()


def test_grid_merge_16cubed_at_the_size_limit(hip, mzo):
    from test_gpu_grid_merge import _oracle
    counts = [2048, 300, 35]
    nc = len(counts)
    pv, pia = _grouped_cells(mzo, counts, seed=12)
    st = hip.DeviceStore(nc, 1, sum(counts) + 512)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    st.set_grid(hip.Grid1DUniform(1e-6 * nc, nc))
    st.set_seed(99)
    sp = [hip.Species("Ar", helpers.MASS_AR)]
    pp = hip.compute_props(st, sp)
    tag = (3 << 4) | 3
    mg = hip.GridN2Merge(16, 16, 16, extent_multiplier=3.0)
    merged = hip.merge_grid_based(st, 1, 30, mg, species_data=sp, tag=tag)
    assert merged == nc
    _oracle(mzo, pv, pia, 30, mzo.GridMerge(16, 16, 16, 3.0), props=pp, tag=tag, seed=99)
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=1e-15)
    st.close()


def test_grid_merge_refuses_2049(hip, mzo):
    counts = [2049, 40]
    nc = len(counts)
    pv, pia = _grouped_cells(mzo, counts, seed=13)
    st = hip.DeviceStore(nc, 1, sum(counts) + 512)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    st.set_grid(hip.Grid1DUniform(1e-6 * nc, nc))
    sp = [hip.Species("Ar", helpers.MASS_AR)]
    hip.compute_props(st, sp)
    before = st.download_species(1)
    hip.merge_grid_based(st, 1, 100, hip.GridN2Merge(4, extent_multiplier=3.0), species_data=sp, tag=5,
                         squash=False)
    got = st.download_species(1)
    assert got["err"] & 32   # MZ_ERR_MERGE_SIZE: the cell is larger than the kernel takes
    assert np.array_equal(got["pia"][0], before["pia"][0])
    r = before["pia"][0]
    seg = np.concatenate([np.arange(r[1] - 1, r[2]), np.arange(r[4] - 1, r[5])])
    assert np.array_equal(got["w"][seg], before["w"][seg])
    st.close()
