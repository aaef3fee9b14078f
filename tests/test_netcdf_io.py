"""NCDataHolder / write_netcdf / close_netcdf (src/io.jl:116-219): the
reference's own output arrays (tests/golden/reference_fixtures.npz, taken
from test/data/*.nc) written through the writer one output time at a time
come back unchanged, with the reference's variable names, dimension order
and ndens comment."""
import os
import sys

import numpy as np
import pytest
from scipy.io import netcdf_file

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = np.load(os.path.join(ROOT, "tests", "golden", "reference_fixtures.npz"))


def _pkg():
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    return ge.load_package()


def _io():
    return _pkg()   # NCDataHolder, write_netcdf, close_netcdf are package attributes


class _Sp:
    def __init__(self, name):
        self.name = name


def test_two_species_fixture_round_trip(tmp_path):
    pkg, io = _pkg(), _io()
    T, nd, npart = FIX["2species_T"], FIX["2species_ndens"], FIX["2species_np"]
    nt, ns, nc = T.shape
    pp = pkg.PhysProps(nc, ns)
    path = str(tmp_path / "two_gas.nc")
    ds = io.NCDataHolder(path, [_Sp("Ar"), _Sp("He")], pp, global_attributes={"seed": 1234})
    for t in range(nt):
        pp.T[:], pp.n[:], pp.np[:] = T[t], nd[t], npart[t]
        pp.lpa[:] = npart[t].sum(axis=1)
        io.write_netcdf(ds, pp, t, sync_freq=100)
    io.close_netcdf(ds)
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.variables["T"].dimensions == ("timestep", "n_species", "n_cells")
        assert np.array_equal(f.variables["T"][:], T)
        assert np.array_equal(f.variables["ndens"][:], nd)
        assert np.array_equal(f.variables["np"][:], npart)
        assert np.array_equal(f.variables["timestep"][:], np.arange(nt, dtype=float))
        names = [b"".join(r).rstrip(b"\0").decode() for r in f.variables["species_names"][:]]
        assert names == ["Ar", "He"]
        assert "physical particles" in getattr(f, "COMMENT ndens").decode()
        assert f.seed == 1234
        assert "moments" not in f.variables


def test_couette_and_moments_layout(tmp_path):
    pkg, io = _pkg(), _io()
    v, T = FIX["couette_v"], FIX["couette_T"]
    nt, ns, nc, _ = v.shape
    pp = pkg.PhysProps(nc, ns, ndens_not_Np=True)
    path = str(tmp_path / "couette.nc")
    ds = io.NCDataHolder(path, [_Sp("Ar")], pp, names_skip_list=["length_particle_array"])
    for t in range(nt):
        pp.v[:], pp.T[:], pp.n[:] = v[t], T[t], FIX["couette_ndens"][t]
        io.write_netcdf(ds, pp, 1000 * t)
    io.close_netcdf(ds)
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.variables["v"].dimensions == ("timestep", "n_species", "n_cells", "vel_components")
        assert np.array_equal(f.variables["v"][:], v)
        assert "length_particle_array" not in f.variables
        assert "number density" in getattr(f, "COMMENT ndens").decode()
    mom = FIX["bkw_moments"]
    nt, ns, nc, nm = mom.shape
    pp = pkg.PhysProps(nc, ns, [4, 6, 8, 10])
    path = str(tmp_path / "bkw.nc")
    ds = io.NCDataHolder(path, [_Sp("Ar")], pp)
    for t in range(nt):
        pp.moments[:] = mom[t]
        io.write_netcdf(ds, pp, t)
    io.close_netcdf(ds)
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.variables["moments"].dimensions == ("timestep", "n_species", "n_cells", "n_moments")
        assert np.array_equal(f.variables["moments"][:], mom)
        assert list(f.variables["moment_powers"][:]) == [4, 6, 8, 10]


def test_rejects_unknown_skip_name(tmp_path):
    pkg, io = _pkg(), _io()
    with pytest.raises(ValueError):
        io.NCDataHolder(str(tmp_path / "x.nc"), [_Sp("Ar")], pkg.PhysProps(2, 1), names_skip_list=["bogus"])


def test_surf_fixture_round_trip(tmp_path):
    """NCDataHolderSurf / write_netcdf for SurfProps (src/io.jl:400-610): the
    reference's wall-property series (test/data/couette_..._vw200to150_surf.nc,
    six outputs) written one output at a time come back unchanged, in the
    reference's variable names and dimension order."""
    pkg, io = _pkg(), _io()
    names = ("np", "flux_incident", "flux_reflected", "force", "normal_pressure", "shear_pressure",
             "kinetic_energy_flux")
    fx = {n: FIX["couette_surf_" + n] for n in names}
    ts = FIX["couette_surf_timestep"]
    nt, ns, ne = fx["np"].shape
    sp = pkg.SurfProps.grid1d(ns)
    assert ne == sp.n_elements
    path = str(tmp_path / "surf.nc")
    ds = io.NCDataHolderSurf(path, [_Sp("Ar")], sp, global_attributes={"seed": 1234})
    for t in range(nt):
        for n in names:
            a = fx[n][t]
            getattr(sp, n)[...] = a.T if a.ndim == 2 else np.transpose(a, (2, 1, 0))
        io.write_netcdf(ds, sp, ts[t], sync_freq=2)
    io.close_netcdf(ds)
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.variables["np"].dimensions == ("timestep", "n_species", "n_elements")
        assert f.variables["force"].dimensions == ("timestep", "n_species", "n_elements", "vector_components")
        for n in names:
            assert np.array_equal(f.variables[n][:], fx[n]), n
        assert np.array_equal(f.variables["timestep"][:], ts)
        assert [b"".join(r).rstrip(b"\0").decode() for r in f.variables["species_names"][:]] == ["Ar"]


def test_surf_skip_list(tmp_path):
    """IOSkipListSurf (src/io.jl:334-398): "np"/"nparticles", "fluxes" (both
    fluxes), "force", "normal_pressure", "shear_pressure", "kinetic_energy_flux"."""
    pkg, io = _pkg(), _io()
    sp = pkg.SurfProps.grid1d(1)
    path = str(tmp_path / "skip.nc")
    ds = io.NCDataHolderSurf(path, [_Sp("Ar")], sp, names_skip_list=["nparticles", "force", "shear_pressure"])
    io.write_netcdf(ds, sp, 0)
    io.close_netcdf(ds)
    with netcdf_file(path, "r", mmap=False) as f:
        for n in ("np", "force", "shear_pressure"):
            assert n not in f.variables
        for n in ("flux_incident", "flux_reflected", "normal_pressure", "kinetic_energy_flux"):
            assert n in f.variables
    path = str(tmp_path / "skip2.nc")
    ds = io.NCDataHolderSurf(path, [_Sp("Ar")], sp, names_skip_list=["fluxes", "normal_pressure",
                                                                      "kinetic_energy_flux"])
    io.close_netcdf(ds)
    with netcdf_file(path, "r", mmap=False) as f:
        assert set(f.variables) == {"timestep", "species_names", "np", "force", "shear_pressure"}
    with pytest.raises(ValueError):
        io.NCDataHolderSurf(str(tmp_path / "bad.nc"), [_Sp("Ar")], sp, names_skip_list=["T"])


def test_flux_round_trip_and_skip_list(tmp_path):
    """NCDataHolderFlux and write_netcdf of FluxProps (src/io.jl:612-780),
    as test/test_io_fluxes.jl exercises them: three outputs read back in the
    reference's layout ([3, n_cells, n_species, timestep] column-major), and
    the skip list dropping one or two of the three variables."""
    pkg, io = _pkg(), _io()
    rng = np.random.default_rng(2)
    fp = pkg.FluxProps(5, 2)
    path = str(tmp_path / "flux.nc")
    ds = io.NCDataHolderFlux(path, [_Sp("Ar"), _Sp("He")], fp)
    series = []
    for t in (0, 10, 20):
        for name in ("kinetic_energy_flux", "diagonal_momentum_flux", "off_diagonal_momentum_flux"):
            getattr(fp, name)[...] = rng.normal(size=(2, 5, 3))
        series.append({k: getattr(fp, k).copy() for k in ("kinetic_energy_flux", "diagonal_momentum_flux",
                                                          "off_diagonal_momentum_flux")})
        io.write_netcdf(ds, fp, t)
    io.close_netcdf(ds)
    with netcdf_file(path, "r", mmap=False) as f:
        assert list(f.variables["timestep"][:]) == [0.0, 10.0, 20.0]
        assert f.variables["kinetic_energy_flux"].dimensions == ("timestep", "n_species", "n_cells",
                                                                 "vector_components")
        for k, ser in enumerate(series):
            for name, a in ser.items():
                assert np.array_equal(f.variables[name][k], a)
        names = [b"".join(r).rstrip(b"\0").decode() for r in f.variables["species_names"][:]]
        assert names == ["Ar", "He"]
    path = str(tmp_path / "skip.nc")
    ds = io.NCDataHolderFlux(path, [_Sp("Ar"), _Sp("He")], fp, names_skip_list=["kinetic_energy_flux",
                                                                               "off_diagonal_momentum_flux"])
    io.write_netcdf(ds, fp, 0)
    io.close_netcdf(ds)
    with netcdf_file(path, "r", mmap=False) as f:
        assert set(f.variables) == {"timestep", "species_names", "diagonal_momentum_flux"}
    with pytest.raises(ValueError):
        io.NCDataHolderFlux(str(tmp_path / "bad.nc"), [_Sp("Ar"), _Sp("He")], fp, names_skip_list=["force"])


def test_particle_dump(tmp_path):
    """write_netcdf(nc_filename, pv, pia, species, species_data) and the
    all-species form (src/io.jl:842-984): particles cell by cell, group 1
    then group 2, with their cells; the reference's dimension and variable
    names."""
    pkg, io = _pkg(), _io()
    rng = np.random.default_rng(9)
    n = 10
    x, v, w = rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.uniform(1, 2, n)
    # cell 1: positions 1-2 and 9-10; cell 2: none; cell 3: positions 3-8
    pia = np.array([[4, 1, 2, 2, 9, 10, 2], [0, 0, -1, 0, 0, -1, 0], [6, 3, 8, 6, 0, -1, 0]])
    sp = [{"x": x, "v": v, "w": w, "pia": pia}, {"x": x[:3], "v": v[:3], "w": w[:3],
                                                  "pia": np.array([[0, 0, -1, 0, 0, -1, 0], [3, 1, 3, 3, 0, -1, 0],
                                                                   [0, 0, -1, 0, 0, -1, 0]])}]
    order = [0, 1, 8, 9, 2, 3, 4, 5, 6, 7]
    path = str(tmp_path / "p1.nc")
    io.write_netcdf(path, sp, [_Sp("Ar"), _Sp("He")], species=1)
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.dimensions['nparticles_["Ar"]'] == n
        assert np.array_equal(f.variables["w"][:], w[order])
        assert np.array_equal(f.variables["v"][:], v[order].T) and np.array_equal(f.variables["x"][:], x[order].T)
        assert list(f.variables["cell"][:]) == [1, 1, 1, 1, 3, 3, 3, 3, 3, 3]
    path = str(tmp_path / "p2.nc")
    io.write_particles_netcdf(path, sp, [_Sp("Ar"), _Sp("He")])
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.dimensions["nparticles_Ar"] == n and f.dimensions["nparticles_He"] == 3
        assert np.array_equal(f.variables["w_Ar"][:], w[order]) and list(f.variables["cell_He"][:]) == [2, 2, 2]
        names = [b"".join(r).rstrip(b"\0").decode() for r in f.variables["species_names"][:]]
        assert names == ["Ar", "He"]
