"""Flux properties on the device (mz_compute_flux_props) against the
oracle's compute_flux_props!: two species over two-group cells, the whole
grid and a cell range, to 1e-12 (the device sums each cell over a wave, the
oracle in particle order); the store's averaged set against avg_props! on
host copies; clearing."""
import numpy as np
import pytest

import helpers
from test_oracle_flux import _kat, M_AR

pytestmark = pytest.mark.gpu


def test_flux_kat_on_device(hip, mzo):
    pv, pia = _kat(mzo)
    st = hip.DeviceStore(8, 1, 64)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    st.set_grid(hip.Grid1DUniform(4.0, 8))
    sp = [hip.Species("Ar", M_AR)]
    hip.compute_props(st, sp)
    fp = hip.compute_flux_props(st, sp)
    pp = mzo.Props(8, 1)
    mzo.compute_props([pv], pia, [M_AR], pp)
    ref = mzo.compute_flux_props([pv], pia, [M_AR], pp, np.full(8, 2.0))
    assert np.allclose(fp.kinetic_energy_flux[0], ref[0, :, 0:3], rtol=1e-14, atol=0)
    assert np.allclose(fp.diagonal_momentum_flux[0], ref[0, :, 3:6], rtol=1e-14, atol=0)
    assert np.allclose(fp.off_diagonal_momentum_flux[0], ref[0, :, 6:9], rtol=1e-14, atol=1e-40)
    st.close()


@pytest.mark.parametrize("cells", [None, (3, 17)])
def test_flux_vs_oracle_two_species(hip, mzo, cells):
    rng = np.random.default_rng(5)
    nc, masses = 24, [M_AR, 6.64e-27]
    pvs, pia = [], mzo.PIA(nc, 2)
    L = 1e-4
    for s, m in enumerate(masses, start=1):
        counts = rng.integers(0, 120, nc)
        counts[4] = 0
        n = int(counts.sum())
        pv = mzo.ParticleVector(n)
        pos = 1
        k1s = counts * 2 // 3
        p2 = int(k1s.sum()) + 1
        vs = np.sqrt(helpers.KB * 400.0 / m)
        for c in range(nc):
            k, k1 = int(counts[c]), int(k1s[c])
            for j in range(k):
                q = pos + j if j < k1 else p2 + (j - k1)
                pv.pop_to(q)
                pv.set(q, float(rng.lognormal(0, 0.5)), rng.normal(0, vs, 3) + [50.0 * s, 0.0, -20.0],
                       [(c + rng.uniform()) * L / nc, 0.0, 0.0])
            k2 = k - k1
            pia.set(c + 1, s, k, pos if k1 else 0, pos + k1 - 1 if k1 else -1, k1, p2 if k2 else 0,
                    p2 + k2 - 1 if k2 else -1, k2)
            pos += k1
            p2 += k2
        pia.set_n_total(s, n)
        pia.set_contiguous(s, False)
        pvs.append(pv)
    st = hip.DeviceStore(nc, 2, 4096)
    for s in (1, 2):
        helpers.upload_from_oracle(st, mzo, pvs[s - 1], pia, species=s)
    st.set_grid(hip.Grid1DUniform(L, nc))
    sp = [hip.Species("Ar", masses[0]), hip.Species("He", masses[1])]
    hip.compute_props(st, sp)
    st.flux_clear(0)
    fp = hip.compute_flux_props(st, sp, cells=cells)
    pp = mzo.Props(nc, 2)
    mzo.compute_props(pvs, pia, masses, pp)
    ref = mzo.compute_flux_props(pvs, pia, masses, pp, np.full(nc, nc / L), cells=cells)
    for got, sl in ((fp.kinetic_energy_flux, slice(0, 3)), (fp.diagonal_momentum_flux, slice(3, 6)),
                    (fp.off_diagonal_momentum_flux, slice(6, 9))):
        r = ref[:, :, sl]
        assert np.allclose(got, r, rtol=1e-12, atol=1e-11 * np.abs(r).max())
    if cells is not None:
        assert not fp.kinetic_energy_flux[:, : cells[0] - 1].any() and not fp.diagonal_momentum_flux[:, cells[1]:].any()
    # the averaged set: avg + inst / n over two calls, against host avg_props!
    st.flux_clear(1)
    st.flux_avg(4)
    st.flux_avg(4)
    avg = hip.FluxProps(nc, 2).fetch(st, 1)
    host = hip.FluxProps(nc, 2)
    hip.avg_props(host, fp, 4)
    hip.avg_props(host, fp, 4)
    assert np.array_equal(avg.kinetic_energy_flux, host.kinetic_energy_flux)
    assert np.array_equal(avg.diagonal_momentum_flux, host.diagonal_momentum_flux)
    st.flux_clear(1)
    assert not hip.FluxProps(nc, 2).fetch(st, 1).diagonal_momentum_flux.any()
    st.close()


def test_particle_dump_from_store(hip, mzo, tmp_path):
    """write_netcdf(nc_filename, pv, pia, species, species_data) (io.jl:842-892)
    from a device store.  The layout is the one particle creation leaves
    (io.jl:864-890 walks it): cell 1's group 2 sits after cell 6's particles
    and the index array is a permutation, so the file order differs from
    both the position order and the id order.  The expected file is the
    oracle's own walk over its ParticleVector and PIA."""
    from scipy.io import netcdf_file
    pv, _ = _kat(mzo)
    pv.index[:6] = [3, 1, 6, 5, 2, 4]          # position p holds particle id index[p]
    for i in range(6):
        pv.x[i] = [0.25 + 0.5 * i, 1.0 + i, -2.0 * i]
    pia = mzo.PIA(8, 1)
    pia.set(1, 1, 4, 1, 2, 2, 5, 6, 2)        # cell 1: group 1 at 1-2, group 2 at 5-6
    pia.set(6, 1, 2, 3, 4, 2)                 # cell 6 between them, at 3-4
    pia.set_n_total(1, 6)
    pia.set_contiguous(1, False)
    st = hip.DeviceStore(8, 1, 64)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    sp = [hip.Species("Ar", M_AR)]
    path = str(tmp_path / "dump.nc")
    hip.write_netcdf(path, st, sp, species=1)
    walk = [(1, p) for p in (1, 2, 5, 6)] + [(6, p) for p in (3, 4)]
    ids = [int(pv.index[p - 1]) - 1 for _, p in walk]
    with netcdf_file(path, "r", mmap=False) as f:
        assert list(f.variables["cell"][:]) == [c for c, _ in walk]
        assert np.array_equal(f.variables["w"][:], pv.w[ids])
        assert np.array_equal(f.variables["v"][:], pv.v[ids].T)
        assert np.array_equal(f.variables["x"][:], pv.x[ids].T)
    st.close()
