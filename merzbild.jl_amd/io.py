"""NetCDF output of PhysProps time series in the layout of the reference's
NCDataHolder (src/io.jl:116-318), so device runs can be set beside the
reference's own output files (test/data/*.nc) and read by its
post-processing scripts.

Variables (netCDF order: the unlimited timestep dimension first):
    timestep               [timestep]
    species_names          [n_species, name_strlen]   (chars)
    length_particle_array  [timestep, n_species]
    moment_powers          [n_moments]                 (int32)
    moments                [timestep, n_species, n_cells, n_moments]
    np, ndens, T           [timestep, n_species, n_cells]
    v                      [timestep, n_species, n_cells, vel_components]
and the global attribute "COMMENT ndens" (io.jl:160-165).  `ndens` holds
PhysProps.n: the number density when ndens_not_Np, otherwise the number of
physical particles in the cell.

The reference writes NetCDF-4; this writer produces NetCDF classic with
64-bit offsets (scipy.io.netcdf_file), which the same readers (libnetcdf,
NetCDF.jl, NCDatasets.jl, netCDF4-python, xarray) read with the same
variable names and shapes.  Classic files allow one unlimited dimension, so
the moment variables are written only when the PhysProps tracks moments
(the reference writes them with a zero-length dimension otherwise).
"""
import numpy as np
from scipy.io import netcdf_file

_SKIPPABLE = ("length_particle_array", "moments", "np", "ndens", "v", "T")


class NCDataHolder:
    """NCDataHolder(nc_filename, species_data, phys_props[, names_skip_list];
    global_attributes) (io.jl:116-150).  names_skip_list names variables
    not to write, as IOSkipList does (io.jl:1-40)."""

    def __init__(self, nc_filename, species_data, phys_props, names_skip_list=(), global_attributes=None):
        for name in names_skip_list:
            if name not in _SKIPPABLE:
                raise ValueError(f"cannot skip {name!r}; skippable: {_SKIPPABLE}")
        self.skip = set(names_skip_list)
        self.ndens_not_Np = bool(phys_props.ndens_not_Np)
        self.n_species, self.n_cells = phys_props.n_species, phys_props.n_cells
        self.n_moments = int(phys_props.n_moments)
        if len(species_data) != self.n_species:
            raise ValueError("species_data and phys_props disagree on the number of species")
        f = netcdf_file(nc_filename, "w", version=2)
        for k, v in (global_attributes or {}).items():
            setattr(f, str(k), v)
        if self.ndens_not_Np:
            comment = "In this file, the ndens variable stores the number density, NOT # of physical particles in cell"
        else:
            comment = "In this file, the ndens variable stores # of physical particles in cell, NOT number density"
        setattr(f, "COMMENT ndens", comment)
        names = [s.name.encode() for s in species_data]
        strlen = max(1, max(len(n) for n in names))
        f.createDimension("timestep", None)   # the record dimension comes first in classic files
        f.createDimension("vel_components", 3)
        f.createDimension("n_cells", self.n_cells)
        f.createDimension("n_species", self.n_species)
        f.createDimension("name_strlen", strlen)
        self.with_moments = self.n_moments > 0 and "moments" not in self.skip
        if self.with_moments:
            f.createDimension("n_moments", self.n_moments)
        spn = f.createVariable("species_names", "c", ("n_species", "name_strlen"))
        spn[:] = np.array([list(n.ljust(strlen, b"\0").decode("latin-1")) for n in names], dtype="S1")
        self.v_timestep = f.createVariable("timestep", "d", ("timestep",))
        self.vars = {}
        cells = ("timestep", "n_species", "n_cells")
        if "length_particle_array" not in self.skip:
            self.vars["length_particle_array"] = f.createVariable("length_particle_array", "d",
                                                                  ("timestep", "n_species"))
        if self.with_moments:
            mp = f.createVariable("moment_powers", "i", ("n_moments",))
            mp[:] = np.asarray(phys_props.moment_powers, dtype=np.int32)
            self.vars["moments"] = f.createVariable("moments", "d", cells + ("n_moments",))
        for name in ("np", "ndens", "T"):
            if name not in self.skip:
                self.vars[name] = f.createVariable(name, "d", cells)
        if "v" not in self.skip:
            self.vars["v"] = f.createVariable("v", "d", cells + ("vel_components",))
        self.f = f
        self.n_written = 0


_SURF_SKIPS = {"np": "np", "nparticles": "np", "fluxes": "fluxes", "force": "force",
               "normal_pressure": "normal_pressure", "shear_pressure": "shear_pressure",
               "kinetic_energy_flux": "kinetic_energy_flux"}


class NCDataHolderSurf:
    """NCDataHolderSurf(nc_filename[, names_skip_list], species_data,
    surf_props; global_attributes) (io.jl:400-549) for SurfProps time series:
        timestep                                   [timestep]
        species_names                              [n_species, name_strlen] (chars)
        np, flux_incident, flux_reflected,
        normal_pressure, kinetic_energy_flux       [timestep, n_species, n_elements]
        force, shear_pressure                      [timestep, n_species, n_elements, vector_components]
    names_skip_list takes the names IOSkipListSurf takes (io.jl:334-398):
    "np" or "nparticles", "fluxes" (both fluxes), "force", "normal_pressure",
    "shear_pressure", "kinetic_energy_flux"; other names are ignored there
    and rejected here."""

    def __init__(self, nc_filename, species_data, surf_props, names_skip_list=(), global_attributes=None):
        skip = set()
        for name in names_skip_list:
            if name not in _SURF_SKIPS:
                raise ValueError(f"cannot skip {name!r}; skippable: {sorted(_SURF_SKIPS)}")
            skip.add(_SURF_SKIPS[name])
        if len(species_data) != surf_props.n_species:
            raise ValueError("species_data and surf_props disagree on the number of species")
        self.n_elements, self.n_species = surf_props.n_elements, surf_props.n_species
        f = netcdf_file(nc_filename, "w", version=2)
        for k, v in (global_attributes or {}).items():
            setattr(f, str(k), v)
        names = [s.name.encode() for s in species_data]
        strlen = max(1, max(len(n) for n in names))
        f.createDimension("timestep", None)   # the record dimension comes first in classic files
        f.createDimension("vector_components", 3)
        f.createDimension("n_elements", self.n_elements)
        f.createDimension("n_species", self.n_species)
        f.createDimension("name_strlen", strlen)
        spn = f.createVariable("species_names", "c", ("n_species", "name_strlen"))
        spn[:] = np.array([list(n.ljust(strlen, b"\0").decode("latin-1")) for n in names], dtype="S1")
        self.v_timestep = f.createVariable("timestep", "d", ("timestep",))
        el = ("timestep", "n_species", "n_elements")
        vec = el + ("vector_components",)
        wanted = [("np", el, "np"), ("flux_incident", el, "fluxes"), ("flux_reflected", el, "fluxes"),
                  ("force", vec, "force"), ("normal_pressure", el, "normal_pressure"),
                  ("shear_pressure", vec, "shear_pressure"), ("kinetic_energy_flux", el, "kinetic_energy_flux")]
        self.vars = {name: f.createVariable(name, "d", dims) for name, dims, key in wanted if key not in skip}
        self.f = f
        self.n_written = 0


_FLUX_NAMES = ("kinetic_energy_flux", "diagonal_momentum_flux", "off_diagonal_momentum_flux")


class NCDataHolderFlux:
    """NCDataHolderFlux(nc_filename[, names_skip_list], species_data,
    flux_props; global_attributes) (io.jl:654-741) for FluxProps time series:
        timestep                                        [timestep]
        species_names                                   [n_species, name_strlen] (chars)
        kinetic_energy_flux, diagonal_momentum_flux,
        off_diagonal_momentum_flux                      [timestep, n_species, n_cells, vector_components]
    (the reference's [3, n_cells, n_species, timestep] in column-major
    order).  names_skip_list takes the names IOSkipListFlux takes
    (io.jl:612-652); other names are ignored there and rejected here."""

    def __init__(self, nc_filename, species_data, flux_props, names_skip_list=(), global_attributes=None):
        for name in names_skip_list:
            if name not in _FLUX_NAMES:
                raise ValueError(f"cannot skip {name!r}; skippable: {list(_FLUX_NAMES)}")
        if len(species_data) != flux_props.n_species:
            raise ValueError("species_data and flux_props disagree on the number of species")
        self.n_cells, self.n_species = flux_props.n_cells, flux_props.n_species
        f = netcdf_file(nc_filename, "w", version=2)
        for k, v in (global_attributes or {}).items():
            setattr(f, str(k), v)
        names = [s.name.encode() for s in species_data]
        strlen = max(1, max(len(n) for n in names))
        f.createDimension("timestep", None)
        f.createDimension("vector_components", 3)
        f.createDimension("n_cells", self.n_cells)
        f.createDimension("n_species", self.n_species)
        f.createDimension("name_strlen", strlen)
        spn = f.createVariable("species_names", "c", ("n_species", "name_strlen"))
        spn[:] = np.array([list(n.ljust(strlen, b"\0").decode("latin-1")) for n in names], dtype="S1")
        self.v_timestep = f.createVariable("timestep", "d", ("timestep",))
        dims = ("timestep", "n_species", "n_cells", "vector_components")
        self.vars = {name: f.createVariable(name, "d", dims) for name in _FLUX_NAMES if name not in names_skip_list}
        self.f = f
        self.n_written = 0


def _write_flux(ds, flux_props, timestep, sync_freq):
    """write_netcdf(ds, flux_props::FluxProps, timestep; sync_freq) (io.jl:759-780)"""
    if (flux_props.n_cells, flux_props.n_species) != (ds.n_cells, ds.n_species):
        raise ValueError("flux_props does not match the file's dimensions")
    t = ds.n_written
    ds.v_timestep[t] = float(timestep)
    for name, var in ds.vars.items():
        var[t] = np.asarray(getattr(flux_props, name), dtype=np.float64)
    ds.n_written += 1
    if sync_freq > 0 and ds.n_written % sync_freq == 0:
        ds.f.flush()


def _write_surf(ds, surf_props, timestep, sync_freq):
    """write_netcdf(ds, surf_props::SurfProps, timestep; sync_freq) (io.jl:566-610)"""
    if (surf_props.n_elements, surf_props.n_species) != (ds.n_elements, ds.n_species):
        raise ValueError("surf_props does not match the file's dimensions")
    t = ds.n_written
    ds.v_timestep[t] = float(timestep)
    for name, var in ds.vars.items():
        a = np.asarray(getattr(surf_props, name), dtype=np.float64)
        # SurfProps arrays are (element, species) and (component, element, species)
        var[t] = a.T if a.ndim == 2 else np.transpose(a, (2, 1, 0))
    ds.n_written += 1
    if sync_freq > 0 and ds.n_written % sync_freq == 0:
        ds.f.flush()


def _species_particles(src, species):
    """(x, v, w, pia table) of one species from a DeviceStore (downloaded)
    or from a dict with those keys (pia: [n_cells, 7] rows of n_local,
    start1, end1, n_group1, start2, end2, n_group2, positions 1-based)."""
    d = src.download_species(species) if hasattr(src, "download_species") else src[species - 1]
    return np.asarray(d["x"]), np.asarray(d["v"]), np.asarray(d["w"]), np.asarray(d["pia"])


def _cellwise(x, v, w, pia):
    """positions in the reference's output order: per cell, group 1 then
    group 2 (io.jl:864-890); returns (0-based positions, 1-based cells)"""
    pos, cells = [], []
    for c, row in enumerate(pia, start=1):
        _, s1, e1, _, s2, e2, n2 = (int(t) for t in row)
        seg = list(range(s1 - 1, e1)) if e1 >= s1 > 0 else []
        if n2 > 0:
            seg += list(range(s2 - 1, e2))
        pos += seg
        cells += [c] * len(seg)
    return np.array(pos, dtype=np.int64), np.array(cells, dtype=np.int32)


def write_particles_netcdf(nc_filename, particles, species_data, species=None, global_attributes=None):
    """write_netcdf(nc_filename, pv, pia, species, species_data) (io.jl:842-892)
    when species is given, else write_netcdf(nc_filename, particles, pia,
    species_data) (io.jl:906-984): the particles cell by cell (group 1, then
    group 2).  particles is a DeviceStore or a list of per-species dicts.
    Dimension and variable names are the reference's ("nparticles_[\"Ar\"]",
    "w", "v", "x", "cell" for one species; "nparticles_Ar", "w_Ar", ... for
    all); vectors are stored (3d, nparticles), the reference's (nparticles,
    3d) Julia arrays in file order.  The cell index is int32: the classic
    format has no 64-bit integers."""
    f = netcdf_file(nc_filename, "w", version=2)
    for k, val in (global_attributes or {}).items():
        setattr(f, str(k), val)
    single = species is not None
    spec = [species] if single else list(range(1, len(species_data) + 1))
    names = [species_data[s - 1].name.encode() for s in spec]
    strlen = max(1, max(len(n) for n in names))
    f.createDimension("3d", 3)
    f.createDimension("1d" if single else "n_species", len(spec))
    f.createDimension("name_strlen", strlen)
    spn = f.createVariable("species_names", "c", ("1d" if single else "n_species", "name_strlen"))
    spn[:] = np.array([list(n.ljust(strlen, b"\0").decode("latin-1")) for n in names], dtype="S1")
    for s in spec:
        x, v, w, pia = _species_particles(particles, s)
        pos, cells = _cellwise(x, v, w, pia)
        name = species_data[s - 1].name
        dim = f'nparticles_["{name}"]' if single else f"nparticles_{name}"
        sfx = "" if single else f"_{name}"
        f.createDimension(dim, len(pos))
        f.createVariable("w" + sfx, "d", (dim,))[:] = w[pos]
        f.createVariable("v" + sfx, "d", ("3d", dim))[:] = v[pos].T
        f.createVariable("x" + sfx, "d", ("3d", dim))[:] = x[pos].T
        f.createVariable("cell" + sfx, "i", (dim,))[:] = cells
    f.close()


def write_netcdf(ds, phys_props, timestep=None, sync_freq=0, **kw):
    """write_netcdf(ds, phys_props, timestep; sync_freq) (io.jl:157-219):
    append one output time; with an NCDataHolderSurf and a SurfProps, the
    surface-properties method (io.jl:566-610).  With a file name first,
    write_netcdf(nc_filename, particles, species_data[, species]) writes a
    particle dump (write_particles_netcdf)."""
    if isinstance(ds, str):
        return write_particles_netcdf(ds, phys_props, timestep, **kw)
    if isinstance(ds, NCDataHolderSurf):
        return _write_surf(ds, phys_props, timestep, sync_freq)
    if isinstance(ds, NCDataHolderFlux):
        return _write_flux(ds, phys_props, timestep, sync_freq)
    if (phys_props.n_species, phys_props.n_cells) != (ds.n_species, ds.n_cells):
        raise ValueError("phys_props does not match the file's dimensions")
    t = ds.n_written
    ds.v_timestep[t] = float(timestep)
    src = {"length_particle_array": phys_props.lpa, "np": phys_props.np, "ndens": phys_props.n,
           "T": phys_props.T, "v": phys_props.v, "moments": phys_props.moments}
    for name, var in ds.vars.items():
        var[t] = np.asarray(src[name], dtype=np.float64)
    ds.n_written += 1
    if sync_freq > 0 and ds.n_written % sync_freq == 0:
        ds.f.flush()


def close_netcdf(ds):
    """close_netcdf(ds) (io.jl:151-155)"""
    ds.f.close()
