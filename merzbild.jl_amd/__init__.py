"""merzbild.jl_amd -- host side of the MI355X particle path.

Python mirror of the Merzbild.jl operator API for the per-timestep particle
path (sort_particles!, convect_particles!, ntc!, compute_props!,
compute_props_sorted!, merge_octree_N2_based!, merge_nnls_based!), backed by
the gfx950 HIP library libmzhip.so through its C ABI (include/mzhip.h).  The
Julia names drop the trailing "!"; every operation mutates the device store in
place as the Julia functions mutate their arguments.

There is no CPU fallback: loading fails loudly if libmzhip.so is absent.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmzhip.so")

K_B = 1.380649e-23          # src/constants.jl:4
TWOPI = 2 * np.pi           # src/constants.jl:29
DELTA_PARTICLES = 256       # src/Merzbild.jl:23

i32, i64, f64, u64 = C.c_int32, C.c_int64, C.c_double, C.c_uint64
P = C.POINTER

_SIGS = {
    "mz_last_error": (C.c_char_p, []),
    "mz_device_info": (C.c_int, [C.c_char_p, C.c_int, P(C.c_int), P(f64)]),
    "mz_create": (C.c_int, [i32, i32, P(i64), P(C.c_void_p)]),
    "mz_destroy": (C.c_int, [C.c_void_p]),
    "mz_sync": (C.c_int, [C.c_void_p]),
    "mz_upload_species": (C.c_int, [C.c_void_p, i32, i64, P(f64), P(f64), P(f64), P(i64), P(i64), i64, i32,
                                    P(i64), i64, i64]),
    "mz_upload_cell_keys": (C.c_int, [C.c_void_p, i32, i64, P(i64)]),
    "mz_get_scalars": (C.c_int, [C.c_void_p, i32, P(i64)]),
    "mz_download_species": (C.c_int, [C.c_void_p, i32, i64, P(f64), P(f64), P(f64), P(i64), P(i64), P(i64),
                                      P(i64), i64]),
    "mz_set_grid_1d": (C.c_int, [C.c_void_p, f64, i64, f64]),
    "mz_set_walls_1d": (C.c_int, [C.c_void_p, P(f64), P(f64), P(f64), P(f64)]),
    "mz_set_interaction": (C.c_int, [C.c_void_p, i32, i32, f64, f64, f64, f64]),
    "mz_set_sigma_g_w_max": (C.c_int, [C.c_void_p, i32, i32, P(f64)]),
    "mz_get_collision_factors": (C.c_int, [C.c_void_p, i32, i32, P(f64), P(i64), P(i64), P(i64), P(i64),
                                           P(i64)]),
    "mz_set_seed": (C.c_int, [C.c_void_p, u64]),
    "mz_set_replay": (C.c_int, [C.c_void_p, i32, P(f64), i64, i64]),
    "mz_sort": (C.c_int, [C.c_void_p, i32, i32]),
    "mz_sort_noalloc": (C.c_int, [C.c_void_p, i32, i32]),
    "mz_restore_particle_ordering": (C.c_int, [C.c_void_p, i32]),
    "mz_ntc_cells": (C.c_int, [C.c_void_p, i32, i32, f64, P(i32), P(f64), i32, f64, i32, u64, i32]),
    "mz_merge_octree_cells": (C.c_int, [C.c_void_p, i32, P(i32), i32, i32, i32, i32, i32, i32, i32, i32, u64, i32,
                                        P(i64)]),
    "mz_merge_nnls_cells": (C.c_int, [C.c_void_p, i32, P(i32), i32, P(i32), i32, P(i32), i32, f64, i32, i32, i32,
                                      P(f64), i32, f64, i32, f64, i32, i32, u64, i32, P(i32)]),
    "mz_compute_flux_props": (C.c_int, [C.c_void_p, P(f64), i32, i32]),
    "mz_flux_avg": (C.c_int, [C.c_void_p, f64]),
    "mz_flux_clear": (C.c_int, [C.c_void_p, i32]),
    "mz_get_flux": (C.c_int, [C.c_void_p, i32, P(f64), P(f64), P(f64)]),
    "mz_merge_grid": (C.c_int, [C.c_void_p, i32, i32, i32, i32, i32, P(f64), P(f64), f64, i32, i32, u64, i32,
                                 P(i64)]),
    "mz_set_sort_path": (C.c_int, [C.c_void_p, i32]),
    "mz_squash": (C.c_int, [C.c_void_p, i32]),
    "mz_convect": (C.c_int, [C.c_void_p, i32, f64, u64, i32]),
    "mz_convect_surf": (C.c_int, [C.c_void_p, i32, f64, u64, i32, f64]),
    "mz_set_surf_areas": (C.c_int, [C.c_void_p, P(f64)]),
    "mz_surf_avg": (C.c_int, [C.c_void_p, f64]),
    "mz_surf_clear": (C.c_int, [C.c_void_p, i32]),
    "mz_get_surf": (C.c_int, [C.c_void_p, i32, P(f64)]),
    "mz_set_surf": (C.c_int, [C.c_void_p, i32, P(f64)]),
    "mz_ntc": (C.c_int, [C.c_void_p, i32, i32, f64, P(f64), f64, i32, u64, i32, i32, i32]),
    "mz_swpm": (C.c_int, [C.c_void_p, i32, f64, f64, P(f64), u64, i32, i32, i32]),
    "mz_compute_props": (C.c_int, [C.c_void_p, i32, i32, P(f64)]),
    "mz_get_props": (C.c_int, [C.c_void_p, P(f64), P(f64), P(f64), P(f64), P(f64)]),
    "mz_compute_props_moments": (C.c_int, [C.c_void_p, P(i32), i32, f64, P(f64)]),
    "mz_get_props_moments": (C.c_int, [C.c_void_p, P(f64)]),
    "mz_compute_mixed_moment": (C.c_int, [C.c_void_p, i32, i32, i32, i32, i32, i32, f64, f64, P(f64)]),
    "mz_merge_octree": (C.c_int, [C.c_void_p, i32, i32, i32, i32, i32, i32, i32, i32, i32, i32, u64, i32,
                                  P(i64)]),
    "mz_merge_octree_targets": (C.c_int, [C.c_void_p, i32, P(i32), P(i32), i32, i32, i32, i32, i32, i32, i32, u64,
                                          i32, P(i64)]),
    "mz_merge_nnls": (C.c_int, [C.c_void_p, i32, i32, P(i32), i32, P(i32), i32, f64, i32, i32, i32, P(f64), i32,
                                f64, i32, f64, i32, u64, i32, P(i32)]),
    "mz_get_nnls_info": (C.c_int, [C.c_void_p, P(i32)]),
    "mz_merge_roulette": (C.c_int, [C.c_void_p, i32, i32, i32, i32, u64, i32, P(i64)]),
    "mz_merge_nnls_cand": (C.c_int, [C.c_void_p, i32, i32, P(i32), i32, P(i32), i32, f64, i32, i32, i32, P(f64),
                                     i32, f64, i32, f64, i32, i32, i32, u64, i32, P(i32)]),
    "mz_couette_steps": (C.c_int, [C.c_void_p, i32, u64, f64, P(f64), i32, i32, P(f64), i32]),
    "mz_couette_avg_steps": (C.c_int, [C.c_void_p, i32, u64, f64, P(f64), P(f64), i32, f64]),
    "mz_props_avg": (C.c_int, [C.c_void_p, f64]),
    "mz_props_avg_clear": (C.c_int, [C.c_void_p]),
    "mz_get_props_avg": (C.c_int, [C.c_void_p, P(f64), P(f64), P(f64), P(f64), P(f64)]),
    "mz_couette_last_path": (C.c_int, [C.c_void_p]),
    "mz_set_slab_shape": (C.c_int, [C.c_void_p, i32, i32]),
    "mz_get_slab_info": (C.c_int, [C.c_void_p, P(i32)]),
    "mz_slab_occupancy": (C.c_int, [i32, P(i32), P(i32)]),
    "mz_batch0d_steps": (C.c_int, [C.c_void_p, i32, i32, u64, f64, P(f64), i32, P(f64), i32]),
    "mz_set_batch_shape": (C.c_int, [C.c_void_p, i32, i32]),
    "mz_get_batch_info": (C.c_int, [C.c_void_p, P(i32)]),
    "mz_timing_enable": (C.c_int, [C.c_void_p, i32]),
    "mz_timing_get": (C.c_int, [C.c_void_p, C.c_char_p, P(f64), P(i64)]),
    "mz_timing_reset": (C.c_int, [C.c_void_p]),
}

EXPORTED_SYMBOLS = tuple(_SIGS)

_lib = None


class MzError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    """Load libmzhip.so; raises if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise MzError(f"libmzhip.so not found at {path}; build it with merzbild.jl_amd/build.py")
        L = C.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise MzError(f"libmzhip error {rc}: {load_library().mz_last_error().decode()}")


def _d(a):
    return None if a is None else a.ctypes.data_as(P(f64))


def _l(a):
    return None if a is None else a.ctypes.data_as(P(i64))


def device_info():
    name = C.create_string_buffer(64)
    ncu = C.c_int(0)
    bw = f64(0)
    _check(load_library().mz_device_info(name, 64, C.byref(ncu), C.byref(bw)))
    return {"arch": name.value.decode(), "n_cu": ncu.value, "hbm_peak_gbps": bw.value}


# ---------------------------------------------------------------------------
# Setup data (src/particles.jl:580-608, src/collisions/collision_utils.jl)

class Species:
    """Species (src/particles.jl:30-35)."""

    def __init__(self, name, mass, charge=0.0):
        self.name, self.mass, self.charge = name, mass, charge


def _parse_toml(path):
    """Minimal reader for the flat [section] key = value files of data/*.toml."""
    out, cur = {}, None
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            if line.startswith("["):
                cur = line.strip("[]").strip().strip('"')
                out[cur] = {}
            elif "=" in line and cur is not None:
                k, v = (t.strip() for t in line.split("=", 1))
                try:
                    out[cur][k] = float(v)
                except ValueError:
                    out[cur][k] = v.strip('"')
    return out


def load_species_data(path, names):
    """load_species_data (src/particles.jl:580-608)."""
    if isinstance(names, str):
        names = [names]
    data = _parse_toml(path)
    return [Species(n, data[n]["mass"], data[n].get("charge", 0.0)) for n in names]


class Interaction:
    """Interaction (src/collisions/collision_utils.jl:73-101, 218-223)."""

    def __init__(self, m1, m2, d, omega, Tref):
        from math import gamma, pi, sqrt
        self.m_r = m1 * m2 / (m1 + m2)
        self.mu1 = m1 / (m1 + m2)
        self.mu2 = m2 / (m1 + m2)
        self.vhs_d, self.vhs_o, self.vhs_Tref = d, omega, Tref
        mref = 0.5 * (m1 + m2)
        self.vhs_muref = 30.0 * sqrt(mref * K_B * Tref) / (4.0 * sqrt(pi) * (5.0 - 2.0 * omega) *
                                                            (7.0 - 2.0 * omega) * d * d)
        self.vhs_factor = pi * d ** 2 * (2 * K_B * Tref / self.m_r) ** (omega - 0.5) / gamma(2.5 - omega)

    def sigma_vhs(self, g):
        """sigma_vhs (src/collisions/collision_cross_sections.jl:179-181)"""
        return self.vhs_factor * g ** (1.0 - 2 * self.vhs_o)


def load_interaction_data(path, species_data):
    """load_interaction_data (src/collisions/collision_utils.jl:159-202): [i][k]
    holds mu1 = m_i/(m_i+m_k); KeyError when a pair is missing."""
    data = _parse_toml(path)
    n = len(species_data)
    out = [[None] * n for _ in range(n)]
    for i, a in enumerate(species_data):
        for k, b in enumerate(species_data):
            if i >= k:
                key = f"{a.name},{b.name}"
                if key not in data:
                    key = f"{b.name},{a.name}"
                e = data[key]
                out[i][k] = Interaction(a.mass, b.mass, e["vhs_d"], e["vhs_o"], e["vhs_Tref"])
                out[k][i] = Interaction(b.mass, a.mass, e["vhs_d"], e["vhs_o"], e["vhs_Tref"])
    return out


def mean_free_path(interaction, species, n, T):
    """mean_free_path(interaction, species, n, T)
    (src/properties/collisional_props.jl:11-15): the VHS single-species
    mean free path 1 / (sqrt(2) pi d^2 n (Tref / T)^(omega - 1/2))."""
    it = interaction[species - 1][species - 1]
    lam = np.sqrt(2.0) * np.pi * (it.vhs_d ** 2) * n
    lam *= (it.vhs_Tref / T) ** (it.vhs_o - 0.5)
    return 1.0 / lam


def mean_collision_frequency(interaction, species_data, species, n, T):
    """mean_collision_frequency(interaction, species_data, species, n, T)
    (src/properties/collisional_props.jl:29-34): 4 d^2 n sqrt(pi k Tref / m)
    (T / Tref)^(1 - omega)."""
    it = interaction[species - 1][species - 1]
    nu = 4.0 * (it.vhs_d ** 2) * n
    nu *= np.sqrt(np.pi * K_B * it.vhs_Tref / species_data[species - 1].mass)
    nu *= (T / it.vhs_Tref) ** (1.0 - it.vhs_o)
    return nu


def estimate_sigma_g_w_max(interaction, m1, m2, T1, T2, Fnum, mult_factor=1.0):
    """estimate_sigma_g_w_max (src/collisions/collision_utils.jl:418-424)"""
    g1 = np.sqrt(2 * T1 * K_B / m1)
    g2 = np.sqrt(2 * T2 * K_B / m2)
    g = 0.5 * (g1 + g2)
    return mult_factor * interaction.sigma_vhs(g) * g * Fnum


class Grid1DUniform:
    """Grid1DUniform(L, nx; wall_offset) (src/grids/grid_uniform1D.jl:49-86)."""

    def __init__(self, L, nx, wall_offset=1e-12):
        self.L = float(L)
        self.n_cells = int(nx)
        self.dx = self.L / nx
        self.inv_dx = 1.0 / self.dx
        self.wall_offset = wall_offset
        self.min_x = self.dx * wall_offset
        self.max_x = self.L - self.dx * wall_offset
        self.V = np.full(nx, self.dx)
        self.cells = _OneBased([_GridCell(i * self.dx, (i + 1) * self.dx, self.dx) for i in range(nx)])


class _GridCell:
    """Cell1D(xlo, xhi, V) of Grid1DUniform (src/grids/grid_uniform1D.jl:14-18)."""
    __slots__ = ("xlo", "xhi", "V", "inv_V")

    def __init__(self, xlo, xhi, V):
        self.xlo, self.xhi, self.V, self.inv_V = xlo, xhi, V, 1.0 / V


class _OneBased:
    """grid.cells[cell] with the reference's 1-based cell numbers."""

    def __init__(self, items):
        self._items = items

    def __getitem__(self, i):
        if not 1 <= i <= len(self._items):
            raise IndexError(f"cell {i} out of range 1..{len(self._items)}")
        return self._items[i - 1]

    def __len__(self):
        return len(self._items)


class MaxwellWalls1D:
    """MaxwellWalls1D (src/convection/boundary_conditions.jl:49-53)."""

    def __init__(self, species_data, T_l, T_r, vy_l, vy_r, acc_l, acc_r):
        self.T = np.array([T_l, T_r], dtype=np.float64)
        self.v = np.array([[0.0, vy_l, 0.0], [0.0, vy_r, 0.0]], dtype=np.float64)
        self.acc = np.array([acc_l, acc_r], dtype=np.float64)
        self.vsq = np.array([[2 * K_B * T / s.mass for s in species_data] for T in (T_l, T_r)], dtype=np.float64)


# ---------------------------------------------------------------------------
# Device store

PIA_FIELDS = ("n_local", "start1", "end1", "n_group1", "start2", "end2", "n_group2")


class DeviceStore:
    """n_species ParticleVectors and their ParticleIndexerArray on the GPU.

    Slot p of a species holds the particle at logical position p+1 of the
    reference's index array, with its physical id; see include/mzhip.h.
    """

    def __init__(self, n_cells, n_species, capacity):
        lib = load_library()
        cap = np.ascontiguousarray(np.broadcast_to(np.asarray(capacity, dtype=np.int64), (n_species,)))
        h = C.c_void_p()
        _check(lib.mz_create(n_cells, n_species, _l(cap), C.byref(h)))
        self.h = h
        self.n_cells, self.n_species = n_cells, n_species
        self.capacity = cap
        self.grid = None
        self.masses = None

    def close(self):
        if self.h:
            load_library().mz_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- data movement ---------------------------------------------------------
    def upload_species(self, species, x, v, w, ids, pia, n_total, contiguous=True, buffer=None, length=None):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        w = np.ascontiguousarray(w, dtype=np.float64)
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        pia = np.ascontiguousarray(pia, dtype=np.int64).reshape(self.n_cells, 7)
        buf = np.ascontiguousarray(buffer if buffer is not None else np.zeros(0), dtype=np.int64)
        n_pos = len(w)
        if length is None:
            length = max(n_pos, int(ids.max()) if n_pos else 0) + len(buf)
        _check(load_library().mz_upload_species(self.h, species, n_pos, _d(x), _d(v), _d(w), _l(ids), _l(pia),
                                                int(n_total), 1 if contiguous else 0, _l(buf), len(buf),
                                                int(length)))

    def upload_cell_keys(self, species, cells):
        """pv.cell[1:n] for the precomputed-cell sort."""
        k = np.ascontiguousarray(cells, dtype=np.int64)
        _check(load_library().mz_upload_cell_keys(self.h, species, len(k), _l(k)))

    def scalars(self, species=1):
        out = np.zeros(6, dtype=np.int64)
        _check(load_library().mz_get_scalars(self.h, species, _l(out)))
        return dict(zip(("n_total", "contiguous", "nbuffer", "length", "n_pos", "err"), out.tolist()))

    def download_species(self, species=1):
        sc = self.scalars(species)
        n = sc["n_pos"]
        x = np.zeros((max(n, 1), 3)); v = np.zeros((max(n, 1), 3)); w = np.zeros(max(n, 1))
        ids = np.zeros(max(n, 1), dtype=np.int64); keys = np.zeros(max(n, 1), dtype=np.int64)
        pia = np.zeros((self.n_cells, 7), dtype=np.int64)
        buf = np.zeros(max(sc["nbuffer"], 1), dtype=np.int64)
        _check(load_library().mz_download_species(self.h, species, n, _d(x), _d(v), _d(w), _l(ids), _l(keys),
                                                  _l(pia), _l(buf), len(buf)))
        return {"x": x[:n], "v": v[:n], "w": w[:n], "index": ids[:n], "cell": keys[:n], "pia": pia,
                "buffer": buf[:sc["nbuffer"]], **sc}

    def pia(self, species=1):
        """The ParticleIndexerArray rows of a species: int64 [n_cells, 7] in
        PIA_FIELDS order."""
        return self.download_species(species)["pia"]

    # -- parameters ------------------------------------------------------------
    def set_grid(self, grid):
        _check(load_library().mz_set_grid_1d(self.h, grid.L, grid.n_cells, grid.wall_offset))
        self.grid = grid

    def set_walls(self, walls):
        _check(load_library().mz_set_walls_1d(self.h, _d(walls.T), _d(np.ascontiguousarray(walls.v.ravel())),
                                              _d(walls.acc), _d(np.ascontiguousarray(walls.vsq.ravel()))))

    def set_interactions(self, interactions):
        for i in range(self.n_species):
            for k in range(self.n_species):
                it = interactions[i][k]
                _check(load_library().mz_set_interaction(self.h, i + 1, k + 1, it.mu1, it.mu2, it.vhs_factor,
                                                         it.vhs_o))

    def set_species_data(self, species_data):
        self.masses = np.array([s.mass for s in species_data], dtype=np.float64)

    def set_sigma_g_w_max(self, s1, s2, values):
        vals = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.float64), (self.n_cells,)))
        _check(load_library().mz_set_sigma_g_w_max(self.h, s1, s2, _d(vals)))

    def collision_factors(self, s1, s2):
        """CollisionFactors fields per cell for the pair (s1, s2)."""
        nc = self.n_cells
        sg = np.zeros(nc)
        arrs = [np.zeros(nc, dtype=np.int64) for _ in range(5)]
        _check(load_library().mz_get_collision_factors(self.h, s1, s2, _d(sg), *[_l(a) for a in arrs]))
        return dict(sigma_g_w_max=sg, n1=arrs[0], n2=arrs[1], n_coll=arrs[2], n_coll_performed=arrs[3],
                    n_eq_w_coll_performed=arrs[4])

    def set_seed(self, seed):
        _check(load_library().mz_set_seed(self.h, int(seed) & 0xFFFFFFFFFFFFFFFF))

    def set_replay(self, kind, table):
        t = np.ascontiguousarray(table, dtype=np.float64)
        rows, cols = t.shape if t.size else (0, 0)
        _check(load_library().mz_set_replay(self.h, {"ntc": 0, "convect": 1, "merge": 2}[kind], _d(t), rows, cols))

    def sync(self):
        _check(load_library().mz_sync(self.h))

    # -- timing ----------------------------------------------------------------
    def timing(self, on=True):
        _check(load_library().mz_timing_enable(self.h, 1 if on else 0))

    def timing_get(self, family):
        ms, n = f64(0), i64(0)
        _check(load_library().mz_timing_get(self.h, family.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def timing_reset(self):
        _check(load_library().mz_timing_reset(self.h))

    # -- slab-resident step shape ------------------------------------------------
    def props_avg(self, n_avg_timesteps):
        """avg_props! of PhysProps on the device: time average += props / n."""
        _check(load_library().mz_props_avg(self.h, float(n_avg_timesteps)))

    def props_avg_clear(self):
        _check(load_library().mz_props_avg_clear(self.h))

    def surf_avg(self, n_avg_timesteps):
        """avg_props! on the device: time-average set += instantaneous set / n."""
        _check(load_library().mz_surf_avg(self.h, float(n_avg_timesteps)))

    def flux_avg(self, n_avg_timesteps):
        """avg_props! of the store's FluxProps sets (flux_props.jl:152-186)."""
        _check(load_library().mz_flux_avg(self.h, float(n_avg_timesteps)))

    def flux_clear(self, which=0):
        """clear_props! of the FluxProps set `which` (0 instantaneous, 1 average)."""
        _check(load_library().mz_flux_clear(self.h, which))

    def surf_clear(self, which=0):
        """clear_props! of the device set `which` (0 instantaneous, 1 average)."""
        _check(load_library().mz_surf_clear(self.h, which))

    def set_sort_path(self, path=0):
        """0 chooses by size, 1 the count-matrix sort, 2 the radix sort (same order)."""
        _check(load_library().mz_set_sort_path(self.h, {"auto": 0, "matrix": 1, "radix": 2}.get(path, path)))

    def set_slab_shape(self, cells_per_slab=0, particles_per_thread=0):
        _check(load_library().mz_set_slab_shape(self.h, cells_per_slab, particles_per_thread))

    def set_batch_shape(self, max_cells=0, target_particles=0):
        _check(load_library().mz_set_batch_shape(self.h, max_cells, target_particles))

    def batch_info(self):
        """(items, max cells per item, target particles per item, workgroups)
        of the last persistent batched 0-D launch."""
        out = (i32 * 4)()
        _check(load_library().mz_get_batch_info(self.h, out))
        return tuple(out)

    def nnls_info(self):
        """(cells tried, cells handed to the FP64 method) of the last
        merge_nnls_candidates call."""
        out = (i32 * 2)()
        _check(load_library().mz_get_nnls_info(self.h, out))
        return tuple(out)

    def slab_info(self):
        """(particles per thread, cells per slab, slabs, outcome) of the last
        slab call; outcome 1 = ran, -1 = aborted and replayed on the graph,
        -2 = the state was not the sorted state, -3 = no shape fits in LDS."""
        out = (i32 * 4)()
        _check(load_library().mz_get_slab_info(self.h, out))
        return tuple(out)


# ---------------------------------------------------------------------------
# Operations with the reference's names

RNG_PHILOX, RNG_REPLAY = 0, 1


def sort_particles(store, species, compute_cells=True):
    """sort_particles!(gridsort, grid, pv, pia, species) (compute_cells) or
    sort_particles!(gridsort, pv, pia, species) (src/grids/grid_sorting.jl:58-182)."""
    _check(load_library().mz_sort(store.h, species, 1 if compute_cells else 0))


def sort_particles_noalloc(store, species, compute_cells=True):
    """sort_particles! as sort_particles, with the sort's scratch in a fixed
    device array instead of allocated memory (for callers that cannot
    allocate); same results."""
    _check(load_library().mz_sort_noalloc(store.h, species, 1 if compute_cells else 0))


def squash_pia(store, species):
    """squash_pia!(pv, pia, species) (src/particles.jl:622-701)"""
    _check(load_library().mz_squash(store.h, species))


def restore_particle_ordering(store, species):
    """restore_particle_ordering!(pv, inv_map) (src/particles.jl:1086-1137):
    the index becomes the identity, the free-id buffer length, length - 1, ..."""
    _check(load_library().mz_restore_particle_ordering(store.h, species))


def convect_particles(store, species, dt, tag=0, rng_mode=RNG_PHILOX, species_data=None, surf_props=None,
                      fetch=True):
    """convect_particles!(rng, grid, walls, pv, pia, species, species_data, dt)
    (src/convection/convection_1D.jl:130-157), and with surf_props
    convect_particles!(..., species_data, surf_props, dt) (176-207): the
    store's instantaneous surface properties are cleared, accumulated over
    the wall contacts and scaled on the device, then copied into surf_props
    (unless fetch is False)."""
    L = load_library()
    if surf_props is None:
        _check(L.mz_convect(store.h, species, dt, tag, rng_mode))
        return None
    if surf_props.n_elements != 2 or surf_props.n_species != store.n_species:
        raise MzError("convect_particles: surf_props must be SurfProps(pia, grid) of the store's species")
    mass = float(_masses(store, species_data)[species - 1])
    areas = np.ascontiguousarray(surf_props.areas, dtype=np.float64)
    _check(L.mz_set_surf_areas(store.h, _d(areas)))
    _check(L.mz_convect_surf(store.h, species, dt, tag, rng_mode, mass))
    return surf_props.fetch(store) if fetch else None


class SurfProps:
    """SurfProps(n_elements, n_species, areas, normals)
    (src/properties/surface_props.jl:22-58), host copies in the reference's
    shapes: np, flux_incident, flux_reflected, normal_pressure and
    kinetic_energy_flux are (n_elements, n_species); force and
    shear_pressure (3, n_elements, n_species); normals (3, n_elements).
    SurfProps.grid1d(n_species) is SurfProps(pia, grid::Grid1DUniform)
    (surface_props.jl:63-64), the layout the device computes."""

    FIELDS = ("np", "flux_incident", "flux_reflected", "force", "normal_pressure", "shear_pressure",
              "kinetic_energy_flux")

    def __init__(self, n_elements, n_species, areas, normals):
        self.n_elements, self.n_species = int(n_elements), int(n_species)
        self.areas = np.array(areas, dtype=np.float64)
        self.inv_areas = np.array([1.0 / a for a in self.areas])
        self.normals = np.array(normals, dtype=np.float64).reshape(3, self.n_elements)
        ne, ns = self.n_elements, self.n_species
        self.np = np.zeros((ne, ns))
        self.flux_incident = np.zeros((ne, ns))
        self.flux_reflected = np.zeros((ne, ns))
        self.force = np.zeros((3, ne, ns))
        self.normal_pressure = np.zeros((ne, ns))
        self.shear_pressure = np.zeros((3, ne, ns))
        self.kinetic_energy_flux = np.zeros((ne, ns))

    @classmethod
    def grid1d(cls, n_species):
        return cls(2, n_species, [1.0, 1.0], np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]]).T)

    def _from_device(self, q):
        q = q.reshape(self.n_species, 2, 11)
        self.np[:] = q[:, :, 0].T
        self.flux_incident[:] = q[:, :, 1].T
        self.flux_reflected[:] = q[:, :, 2].T
        self.force[:] = np.transpose(q[:, :, 3:6], (2, 1, 0))
        self.normal_pressure[:] = q[:, :, 6].T
        self.shear_pressure[:] = np.transpose(q[:, :, 7:10], (2, 1, 0))
        self.kinetic_energy_flux[:] = q[:, :, 10].T
        return self

    def _to_device(self):
        q = np.zeros((self.n_species, 2, 11))
        q[:, :, 0] = self.np.T
        q[:, :, 1] = self.flux_incident.T
        q[:, :, 2] = self.flux_reflected.T
        q[:, :, 3:6] = np.transpose(self.force, (2, 1, 0))
        q[:, :, 6] = self.normal_pressure.T
        q[:, :, 7:10] = np.transpose(self.shear_pressure, (2, 1, 0))
        q[:, :, 10] = self.kinetic_energy_flux.T
        return np.ascontiguousarray(q)

    def fetch(self, store, which=0):
        """Copy the store's set `which` (0: last convection, 1: time average)."""
        q = np.zeros(self.n_species * 2 * 11)
        _check(load_library().mz_get_surf(store.h, which, _d(q)))
        return self._from_device(q)

    def push(self, store, which=0):
        q = self._to_device()
        _check(load_library().mz_set_surf(store.h, which, _d(q)))


_FLUX_FIELDS = ("kinetic_energy_flux", "diagonal_momentum_flux", "off_diagonal_momentum_flux")


def clear_props(props):
    """clear_props! of host copies: SurfProps (surface_props.jl:171-189) or
    FluxProps (flux_props.jl:133-137); for the store's device sets use
    DeviceStore.surf_clear / flux_clear."""
    for f in (_FLUX_FIELDS if isinstance(props, FluxProps) else SurfProps.FIELDS):
        getattr(props, f)[...] = 0.0


def surface_props_scale(species, surf_props, species_data, dt):
    """surface_props_scale!(species, surf_props, species_data, dt)
    (surface_props.jl:144-160)"""
    factor_base = species_data[species - 1].mass / dt
    s = species - 1
    for e in range(surf_props.n_elements):
        factor = factor_base * surf_props.inv_areas[e]
        for f in ("flux_incident", "flux_reflected", "normal_pressure", "kinetic_energy_flux"):
            getattr(surf_props, f)[e, s] *= factor
        surf_props.force[:, e, s] *= factor
        surf_props.shear_pressure[:, e, s] *= factor


def avg_props(props_avg, props, n_avg_timesteps):
    """avg_props! on host copies, for SurfProps (surface_props.jl:202-222)
    and PhysProps (physical_props.jl:281-299): avg + props / n, within one
    rounding of the reference's fused form.  DeviceStore.surf_avg and
    DeviceStore.props_avg do the same on the store's device sets."""
    inv = 1.0 / n_avg_timesteps
    if isinstance(props, PhysProps):
        if props_avg.ndens_not_Np != props.ndens_not_Np:
            raise MzError("Inconsistent computation of ndens/number of physical particles in cell")
        fields = ("lpa", "np", "n", "v", "T")
    elif isinstance(props, FluxProps):
        fields = _FLUX_FIELDS           # flux_props.jl:152-186
    else:
        fields = SurfProps.FIELDS
    for f in fields:
        a = getattr(props_avg, f)
        a[...] = a + getattr(props, f) * inv


def reduce_surf_props(surf_props_target, surf_props_chunks):
    """reduce_surf_props!(surf_props_target, surf_props_chunks)
    (surface_props.jl:232-254)"""
    clear_props(surf_props_target)
    for sp in surf_props_chunks:
        for f in SurfProps.FIELDS:
            getattr(surf_props_target, f)[...] += getattr(sp, f)


def ntc(store, s1, s2, dt, V=None, dw_tol=1e-16, equal_weight=False, tag=0, rng_mode=RNG_PHILOX, cells=None):
    """ntc! / ntc_equal_weight! over cells (1-based inclusive range)
    (src/collisions/collision_ntc.jl:338-595)."""
    lo, hi = cells if cells is not None else (1, store.n_cells)
    Vd = None if V is None else np.ascontiguousarray(np.broadcast_to(np.asarray(V, dtype=np.float64),
                                                                      (store.n_cells,)))
    _check(load_library().mz_ntc(store.h, s1, s2, dt, _d(Vd), dw_tol, 1 if equal_weight else 0, tag, rng_mode,
                                 lo, hi))


def swpm(store, species, G, dt, V=None, tag=0, rng_mode=RNG_PHILOX, cells=None):
    """swpm!(rng, collision_factors_swpm, collision_data, interaction, pv,
    pia, cell, species, G, dt, V) over cells (1-based inclusive range)
    (src/collisions/collision_swpm.jl:201-288).  (sigma g)_max per cell is
    set with DeviceStore.set_sigma_g_w_max and read back by
    collision_factors (its "sigma_g_w_max" entry)."""
    lo, hi = cells if cells is not None else (1, store.n_cells)
    Vd = None if V is None else np.ascontiguousarray(np.broadcast_to(np.asarray(V, dtype=np.float64),
                                                                      (store.n_cells,)))
    _check(load_library().mz_swpm(store.h, species, float(G), dt, _d(Vd), tag, rng_mode, lo, hi))


def ntc_equal_weight(store, s1, s2, dt, V=None, tag=0, rng_mode=RNG_PHILOX, cells=None):
    return ntc(store, s1, s2, dt, V=V, equal_weight=True, tag=tag, rng_mode=rng_mode, cells=cells)


class PhysProps:
    """PhysProps (src/properties/physical_props.jl:24-90), host copies.
    moments[species, cell, k] holds the total moment of power
    moment_powers[k] (compute_props_with_total_moments!)."""

    def __init__(self, n_cells, n_species, moments_list=(), ndens_not_Np=False, Tref=1.0):
        if isinstance(moments_list, bool):     # PhysProps(n_cells, n_species, ndens_not_Np)
            moments_list, ndens_not_Np = (), moments_list
        self.n_cells, self.n_species, self.ndens_not_Np = n_cells, n_species, ndens_not_Np
        self.moment_powers = np.ascontiguousarray(moments_list, dtype=np.int32)
        self.n_moments = len(self.moment_powers)
        self.Tref = Tref
        self.lpa = np.zeros(n_species)
        self.np = np.zeros((n_species, n_cells))
        self.n = np.zeros((n_species, n_cells))
        self.v = np.zeros((n_species, n_cells, 3))
        self.T = np.zeros((n_species, n_cells))
        self.moments = np.zeros((n_species, n_cells, self.n_moments))

    def fetch_avg(self, store):
        """Copy the store's time-averaged props (DeviceStore.props_avg)."""
        _check(load_library().mz_get_props_avg(store.h, _d(self.lpa), _d(self.np), _d(self.n), _d(self.v),
                                               _d(self.T)))
        return self

    def fetch(self, store, moments=False):
        _check(load_library().mz_get_props(store.h, _d(self.lpa), _d(self.np), _d(self.n), _d(self.v), _d(self.T)))
        if moments and self.n_moments:
            _check(load_library().mz_get_props_moments(store.h, _d(self.moments)))
        return self


def _masses(store, species_data):
    if species_data is not None:
        return np.array([s.mass for s in species_data], dtype=np.float64)
    if store.masses is None:
        raise MzError("species masses unknown: pass species_data or call set_species_data")
    return store.masses


def compute_props(store, species_data=None, phys_props=None):
    """compute_props!(particles, pia, species_data, phys_props)
    (src/properties/physical_props.jl:104-154)"""
    m = _masses(store, species_data)
    _check(load_library().mz_compute_props(store.h, 0, 0, _d(m)))
    pp = phys_props or PhysProps(store.n_cells, store.n_species)
    return pp.fetch(store)


def compute_props_with_total_moments(store, species_data, phys_props, fetch=True):
    """compute_props_with_total_moments!(particles, pia, species_data, phys_props)
    (src/properties/physical_props.jl:168-245); without moments it is
    compute_props!, as in the reference."""
    if phys_props.n_moments == 0:
        return compute_props(store, species_data, phys_props)
    m = _masses(store, species_data)
    _check(load_library().mz_compute_props_moments(store.h, phys_props.moment_powers.ctypes.data_as(P(i32)),
                                                    phys_props.n_moments, phys_props.Tref, _d(m)))
    return phys_props.fetch(store, moments=True) if fetch else None


def compute_mixed_moment(store, species, powers, cells=None, sum_scaler=1.0, res_scaler=1.0):
    """compute_mixed_moment(particles, pia, cell, species, powers; sum_scaler,
    res_scaler) (src/properties/physical_props.jl:473-501).  cells is one
    1-based cell (returns a float) or an inclusive (lo, hi) range (returns an
    array, one value per cell); None means every cell."""
    px, py, pz = (int(p) for p in powers)
    single = cells is not None and np.isscalar(cells)
    lo, hi = (int(cells), int(cells)) if single else ((1, store.n_cells) if cells is None else map(int, cells))
    out = np.zeros(max(hi - lo + 1, 1))
    _check(load_library().mz_compute_mixed_moment(store.h, int(species), lo, hi, px, py, pz, float(sum_scaler),
                                                   float(res_scaler), _d(out)))
    return float(out[0]) if single else out


def compute_props_sorted(store, species_data=None, phys_props=None, ndens_not_Np=False, fetch=True):
    """compute_props_sorted!(particles, pia, species_data, phys_props[, grid])
    (src/properties/physical_props.jl:317-454)"""
    m = _masses(store, species_data)
    _check(load_library().mz_compute_props(store.h, 1, 1 if ndens_not_Np else 0, _d(m)))
    if not fetch:
        return None
    pp = phys_props or PhysProps(store.n_cells, store.n_species, ndens_not_Np)
    return pp.fetch(store)


OCTREE_MID, OCTREE_MEAN = 1, 2
OCTREE_INIT_MINMAX, OCTREE_INIT_MINMAX_SYM, OCTREE_INIT_C = 1, 2, 3
OCTREE_BOUNDS_INHERIT, OCTREE_BOUNDS_RECOMPUTE = 1, 2


def merge_octree_N2_based(store, species, threshold, target_np, split=OCTREE_MID,
                          init_bin_bounds=OCTREE_INIT_MINMAX, bin_bounds_compute=OCTREE_BOUNDS_INHERIT,
                          max_Nbins=4096, max_depth=10, grid_clamp=False, squash=True, tag=0,
                          rng_mode=RNG_PHILOX):
    """merge_octree_N2_based! on every cell with n_local > threshold
    (src/merging/merging_octree_N2.jl:1060-1094); returns the merged-cell count."""
    n = i64(0)
    _check(load_library().mz_merge_octree(store.h, species, threshold, target_np, split, init_bin_bounds,
                                          bin_bounds_compute, max_Nbins, max_depth, 1 if grid_clamp else 0,
                                          1 if squash else 0, tag, rng_mode, C.byref(n)))
    return n.value


def merge_octree_targets(store, species, thresholds, targets, split=OCTREE_MID, init_bin_bounds=OCTREE_INIT_MINMAX,
                         bin_bounds_compute=OCTREE_BOUNDS_INHERIT, max_Nbins=4096, max_depth=10, grid_clamp=False,
                         squash=True, tag=0, rng_mode=RNG_PHILOX):
    """merge_octree_N2_based!(rng, octree, pv, pia, cell, species,
    targets[cell]) on every cell whose n_local exceeds thresholds[cell]
    (per-cell arrays of n_cells entries; src/merging/merging_octree_N2.jl:1060-1094);
    returns the merged-cell count."""
    th = np.ascontiguousarray(thresholds, dtype=np.int32)
    tg = np.ascontiguousarray(targets, dtype=np.int32)
    if th.shape != (store.n_cells,) or tg.shape != (store.n_cells,):
        raise MzError("merge_octree_targets: thresholds and targets need one entry per cell")
    n = i64(0)
    _check(load_library().mz_merge_octree_targets(store.h, species, th.ctypes.data_as(P(i32)),
                                                  tg.ctypes.data_as(P(i32)), split, init_bin_bounds,
                                                  bin_bounds_compute, max_Nbins, max_depth, 1 if grid_clamp else 0,
                                                  1 if squash else 0, tag, rng_mode, C.byref(n)))
    return n.value


def merge_octree_cells(store, species, cells, target_np, split=OCTREE_MID, init_bin_bounds=OCTREE_INIT_MINMAX,
                       bin_bounds_compute=OCTREE_BOUNDS_INHERIT, max_Nbins=4096, max_depth=10, grid_clamp=False,
                       tag=0, rng_mode=RNG_PHILOX):
    """merge_octree_N2_based! on each listed cell (1-based), in list order,
    without squash (src/merging/merging_octree_N2.jl:1060-1094); returns
    the merged-cell count."""
    cl = np.ascontiguousarray(cells, dtype=np.int32)
    n = i64(0)
    _check(load_library().mz_merge_octree_cells(store.h, species, cl.ctypes.data_as(P(i32)), len(cl), target_np,
                                                split, init_bin_bounds, bin_bounds_compute, max_Nbins, max_depth,
                                                1 if grid_clamp else 0, tag, rng_mode, C.byref(n)))
    return n.value


def _mim3(m):
    return np.ascontiguousarray(np.asarray(m, dtype=np.int32).reshape(-1, 3))


def merge_nnls_cells(store, species, cells, multi_index_moments, vref=1.0, scaling="variance", n_rand_pairs=0,
                     centered_at_mean=False, v_multipliers=(), max_err=1e-11, iteration_mult=2, w_threshold=0.0,
                     max_pcols=-1, method="reference", multi_index_moments_pos=(), tag=0, rng_mode=RNG_PHILOX):
    """merge_nnls_based! on each listed cell (1-based), in list order,
    without squash (src/merging/merging_nnls.jl:1457-1510); returns the
    reference's return value per listed cell (1 merged, -1 failed)."""
    cl = np.ascontiguousarray(cells, dtype=np.int32)
    mim, mp = _mim3(multi_index_moments), _mim3(multi_index_moments_pos)
    vm = np.ascontiguousarray(v_multipliers, dtype=np.float64)
    st = np.zeros(len(cl), dtype=np.int32)
    _check(load_library().mz_merge_nnls_cells(store.h, species, cl.ctypes.data_as(P(i32)), len(cl),
                                              mim.ctypes.data_as(P(i32)), len(mim), mp.ctypes.data_as(P(i32)),
                                              len(mp), vref, 1 if scaling == "vref" else 0, n_rand_pairs,
                                              1 if centered_at_mean else 0, _d(vm), len(vm), max_err,
                                              iteration_mult, w_threshold, max_pcols, NNLS_METHODS[method], tag,
                                              rng_mode, st.ctypes.data_as(P(i32))))
    return st


class FluxProps:
    """FluxProps (src/properties/flux_props.jl:27-58), host copies shaped
    [species, cell, component] (the reference's arrays are (component, cell,
    species) in column-major order, the same memory order)."""

    def __init__(self, n_cells, n_species):
        self.n_cells, self.n_species = n_cells, n_species
        self.kinetic_energy_flux = np.zeros((n_species, n_cells, 3))
        self.diagonal_momentum_flux = np.zeros((n_species, n_cells, 3))
        self.off_diagonal_momentum_flux = np.zeros((n_species, n_cells, 3))

    def fetch(self, store, which=0):
        """Copy the store's instantaneous (which=0) or averaged (1) set."""
        _check(load_library().mz_get_flux(store.h, which, _d(self.kinetic_energy_flux),
                                          _d(self.diagonal_momentum_flux), _d(self.off_diagonal_momentum_flux)))
        return self


def compute_flux_props(store, species_data=None, flux_props=None, cells=None, fetch=True):
    """compute_flux_props!(particles, pia, species_data, phys_props,
    flux_props, grid) (src/properties/flux_props.jl:77-131), or with cells =
    (lo, hi) the cell_chunk form of compute_flux_props_sorted! (:199-249).
    Uses the mean velocities of the store's last compute_props call."""
    m = _masses(store, species_data)
    lo, hi = (1, store.n_cells) if cells is None else cells
    _check(load_library().mz_compute_flux_props(store.h, _d(m), int(lo), int(hi)))
    if not fetch:
        return None
    fp = flux_props if flux_props is not None else FluxProps(store.n_cells, store.n_species)
    return fp.fetch(store, 0)


def compute_flux_props_sorted(store, species_data=None, flux_props=None, cells=None, fetch=True):
    """compute_flux_props_sorted! (src/properties/flux_props.jl:199-249); the
    device pass reads groups 1 and 2, so it serves both forms."""
    return compute_flux_props(store, species_data, flux_props, cells, fetch)


class GridN2Merge:
    """GridN2Merge(Nx, Ny, Nz, extent_multiplier) (src/merging/merging_grid.jl:72-115)."""

    def __init__(self, Nx, Ny=None, Nz=None, extent_multiplier=3.5):
        Ny = Nx if Ny is None else Ny
        Nz = Nx if Nz is None else Nz
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        m = np.broadcast_to(np.asarray(extent_multiplier, dtype=np.float64), (3,))
        self.extent_multiplier = np.ascontiguousarray(m)
        self.Ntotal = self.Nx * self.Ny * self.Nz + 8


def merge_grid_based(store, species, threshold, merging_grid, species_data=None, extents=None, grid_clamp=False,
                     squash=True, tag=0, rng_mode=RNG_PHILOX):
    """merge_grid_based! (src/merging/merging_grid.jl:597-703) on every cell
    with n_local > threshold.  extents = (vx_extent, vy_extent, vz_extent)
    gives the velocity grid's bounds; without it the bounds come from each
    cell's T and v of the last compute_props call on the store (species_data
    gives the mass).  Returns the number of cells merged."""
    mg = merging_grid
    n = i64(0)
    if extents is not None:
        ext = np.ascontiguousarray(np.asarray(extents, dtype=np.float64).reshape(6))
        mass, mult = 0.0, None
    else:
        if species_data is None:
            raise ValueError("merge_grid_based: extents from the props need species_data")
        ext, mult, mass = None, mg.extent_multiplier, float(species_data[species - 1].mass)
    _check(load_library().mz_merge_grid(store.h, species, int(threshold), mg.Nx, mg.Ny, mg.Nz,
                                        _d(mult) if mult is not None else None, _d(ext) if ext is not None else None,
                                        mass, 1 if grid_clamp else 0, 1 if squash else 0, tag, rng_mode,
                                        C.byref(n)))
    return n.value


def merge_roulette(store, species, threshold, target_np, conservative=False, tag=0, rng_mode=RNG_PHILOX):
    """merge_roulette!(rng, pv, pia, cell, species, target_np; conservative)
    (src/merging/merging_roulette.jl:25-184) on every cell above `threshold`;
    returns the number of cells merged."""
    n = i64(0)
    _check(load_library().mz_merge_roulette(store.h, species, threshold, target_np, 1 if conservative else 0, tag,
                                            rng_mode, C.byref(n)))
    return n.value


def merge_nnls_based(store, species, threshold, multi_index_moments, vref=1.0, scaling="variance",
                     centered_at_mean=True, v_multipliers=(0.5, 1.0), max_err=1e-11, iteration_mult=2,
                     w_threshold=0.0, squash=True, multi_index_moments_pos=(), n_rand_pairs=0, tag=0,
                     rng_mode=RNG_PHILOX):
    """merge_nnls_based! on every cell with n_local > threshold
    (src/merging/merging_nnls.jl:1457-1510), with the spatial moments of
    NNLSMerge's multi_index_moments_pos and n_rand_pairs random-pair
    columns (drawn with tag); returns the per-cell status (1 merged, -1
    failed, 0 not attempted)."""
    mim, mp = _mim3(multi_index_moments), _mim3(multi_index_moments_pos)
    vm = np.ascontiguousarray(v_multipliers, dtype=np.float64)
    status = np.zeros(store.n_cells, dtype=np.int32)
    _check(load_library().mz_merge_nnls(store.h, species, threshold, mim.ctypes.data_as(P(i32)), len(mim),
                                        mp.ctypes.data_as(P(i32)), len(mp), vref, 1 if scaling == "vref" else 0,
                                        n_rand_pairs, 1 if centered_at_mean else 0, _d(vm), len(vm), max_err,
                                        iteration_mult, w_threshold, 1 if squash else 0, tag, rng_mode,
                                        status.ctypes.data_as(P(i32))))
    return status


NNLS_METHODS = {"reference": 0, "lds": 1, "mixed": 2, "wave": 3}


def merge_nnls_candidates(store, species, threshold, multi_index_moments, max_pcols, method="lds", vref=1.0,
                          scaling="variance", centered_at_mean=True, v_multipliers=(0.5, 1.0), max_err=1e-11,
                          iteration_mult=2, w_threshold=0.0, squash=True, multi_index_moments_pos=(),
                          n_rand_pairs=0, tag=0, rng_mode=RNG_PHILOX):
    """merge_nnls_based! with the columns cut to the first max_pcols particles
    plus the fictitious ones; the moments matched are those of all the cell's
    particles.  method "reference" keeps the reference's layout (bit-exact
    with the CPU restatement), "lds" runs one workgroup per cell with the
    matrix in LDS, "mixed" runs the active-set search on the matrix stored in
    FP32 and solves the final passive set again in FP64 (cells it cannot
    confirm run "lds"), "wave" does the same with one wavefront per cell and
    the FP32 matrix in registers.  Returns the per-cell status as
    merge_nnls_based."""
    mim, mp = _mim3(multi_index_moments), _mim3(multi_index_moments_pos)
    vm = np.ascontiguousarray(v_multipliers, dtype=np.float64)
    status = np.zeros(store.n_cells, dtype=np.int32)
    _check(load_library().mz_merge_nnls_cand(store.h, species, threshold, mim.ctypes.data_as(P(i32)), len(mim),
                                             mp.ctypes.data_as(P(i32)), len(mp), vref,
                                             1 if scaling == "vref" else 0, n_rand_pairs,
                                             1 if centered_at_mean else 0, _d(vm), len(vm), max_err,
                                             iteration_mult, w_threshold, 1 if squash else 0, max_pcols,
                                             NNLS_METHODS[method], tag, rng_mode, status.ctypes.data_as(P(i32))))
    return status


COUETTE_MODES = {"eager": 0, "graph": 1, "slab": 2}


def couette_steps(store, n_steps, step0, dt, species_data=None, V=None, equal_weight=True, props_every=1,
                  graph=True, mode=None):
    """n_steps of the 1-D Couette loop (test/test_1D_couette.jl:66-97): ntc! in
    every cell, convect_particles!, sort_particles!, compute_props_sorted!.

    mode "eager" launches kernel by kernel, "graph" replays a captured
    hipGraph of the step, "slab" runs all steps in one launch with each slab
    of cells resident in LDS (the chunked loop of
    simulations/1D/couette_multithreaded.jl:97-161); "slab" falls back to
    "graph" where it cannot run.  Without mode, graph=True/False picks
    "graph"/"eager".  Returns the path taken."""
    m = _masses(store, species_data)
    Vd = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
    code = COUETTE_MODES[mode] if mode is not None else (1 if graph else 0)
    L = load_library()
    _check(L.mz_couette_steps(store.h, n_steps, step0, dt, _d(Vd), 1 if equal_weight else 0, props_every, _d(m),
                              code))
    return {v: k for k, v in COUETTE_MODES.items()}.get(L.mz_couette_last_path(store.h))


def couette_avg_steps(store, n_steps, step0, dt, n_avg_timesteps, species_data=None, V=None, surf_props=None,
                      mode="graph"):
    """n_steps of the averaging phase of
    simulations/1D/couette_with_surface_quantities.jl:73-115: ntc!,
    convect_particles! with surface properties, avg_props! of SurfProps,
    sort_particles!, compute_props_sorted!, avg_props! of PhysProps.  The
    averages stay on the device (PhysProps.fetch_avg, SurfProps.fetch(store,
    1)); surf_props, if given, sets the element areas.  Returns the path taken."""
    m = _masses(store, species_data)
    if surf_props is not None:
        areas = np.ascontiguousarray(surf_props.areas, dtype=np.float64)
        _check(load_library().mz_set_surf_areas(store.h, _d(areas)))
    Vd = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
    L = load_library()
    _check(L.mz_couette_avg_steps(store.h, n_steps, step0, dt, _d(Vd), _d(m), COUETTE_MODES[mode],
                                  float(n_avg_timesteps)))
    return {v: k for k, v in COUETTE_MODES.items()}.get(L.mz_couette_last_path(store.h))


def slab_occupancy(particles_per_thread):
    """(workgroups per CU, LDS bytes per workgroup) of the slab kernel."""
    b, l = i32(0), i32(0)
    _check(load_library().mz_slab_occupancy(particles_per_thread, C.byref(b), C.byref(l)))
    return b.value, l.value


def couette_autotune(store, step0, dt, species_data=None, steps=200,
                     shapes=((0, 0), (1, 1), (1, 2), (2, 2), (2, 4), (3, 4), (4, 4))):
    """Pick the slab shape (cells per slab, particles per thread) of the
    slab-resident Couette step by running `steps` steps with each candidate
    on the store itself (the steps advance the simulation, as warmup steps
    would) and keeping the fastest one that ran without falling back; then
    two steps with that shape.  Returns (best shape, {shape: ms per step},
    next step number)."""
    times = {}
    s = step0
    for shp in shapes:
        store.set_slab_shape(*shp)
        store.sync()
        store.timing_reset()
        store.timing(True)
        path = couette_steps(store, steps, s, dt, species_data, mode="slab")
        store.sync()
        ms, _ = store.timing_get("slab")
        store.timing(False)
        s += steps
        if path == "slab" and ms > 0:
            times[tuple(shp)] = ms
    best = min(times, key=times.get) if times else (0, 0)
    store.set_slab_shape(*best)
    # two more steps with the chosen shape: the shape's launch set-up
    # (occupancy query, scratch sizes) is then done before the caller's run
    couette_steps(store, 2, s, dt, species_data, mode="slab")
    return best, times, s + 2


def batch0d_autotune(store, step0, dt, species_data=None, steps=3,
                     shapes=((64, 4096), (64, 1024), (16, 4096), (8, 8192), (32, 16384), (4, 2048))):
    """Pick the item packing (max cells, target particles) of the persistent
    batched 0-D path by timing `steps` steps with each candidate on the store
    itself.  Returns (best shape, {shape: ms per step}, next step number)."""
    times = {}
    s = step0
    for shp in shapes:
        store.set_batch_shape(*shp)
        store.sync()
        store.timing_reset()
        store.timing(True)
        batch0d_steps(store, steps, s, dt, species_data, mode="batched")
        store.sync()
        ms, _ = store.timing_get("batch0d")
        store.timing(False)
        s += steps
        times[tuple(shp)] = ms
    best = min(times, key=times.get)
    store.set_batch_shape(*best)
    return best, times, s


def batch0d_steps(store, n_steps, step0, dt, species_data=None, V=None, props_every=1, species=1, mode="batched"):
    """n_steps of ntc_equal_weight! + compute_props! in every 0-D cell (the
    per-cell loops of test/test_bkw.jl:88-93 over a batch of cells).  mode
    "plain" launches the NTC kernel over all cells and the props kernel per
    step, "per-cell" launches NTC once per cell (the reference's loop of
    ntc! calls, one drop-in launch each) and then props, "batched" runs all
    steps in one persistent launch; all give identical results."""
    m = _masses(store, species_data)
    Vd = None if V is None else np.ascontiguousarray(np.broadcast_to(np.asarray(V, dtype=np.float64),
                                                                      (store.n_cells,)))
    _check(load_library().mz_batch0d_steps(store.h, species, n_steps, step0, dt, _d(Vd), props_every, _d(m),
                                           {"plain": 0, "batched": 1, "per-cell": 2}[mode]))


def multi_index_moments(n):
    """compute_multi_index_moments(n) (src/merging/merging_nnls.jl:341-352)"""
    return [(i, j, n - i - j) for i in range(n + 1) for j in range(n - i + 1) if n - i - j >= 0]


# NetCDF output in the reference's NCDataHolder layout (io.py), loaded on use
_IO_NAMES = ("NCDataHolder", "NCDataHolderSurf", "NCDataHolderFlux", "write_netcdf", "write_particles_netcdf",
             "close_netcdf")


def __getattr__(name):
    if name in _IO_NAMES:
        from . import io as _io
        return getattr(_io, name)
    if name == "dropin":
        import importlib
        return importlib.import_module(__name__ + ".dropin")
    raise AttributeError(name)
