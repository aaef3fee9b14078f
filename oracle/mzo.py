"""ctypes binding of the CPU oracle (oracle/libmzo.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by ``bench.py``'s cpu_baseline leg, always as the checker.  The product
path (merzbild.jl_amd) never imports this module.

Cells, species and logical positions are 1-based, as in the reference
(Merzbild.jl ``src/particles.jl``).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmzo.so")

i64 = C.c_int64
i32 = C.c_int32
f64 = C.c_double
P = C.POINTER


class Rng(C.Structure):
    _fields_ = [("mode", i32), ("err", i32), ("seed", C.c_uint64), ("tag", C.c_uint64),
                ("stream", C.c_uint64), ("k", C.c_uint64), ("table", P(f64)),
                ("rows", i64), ("cols", i64)]


class Ind(C.Structure):
    _fields_ = [(n, i64) for n in ("n_local", "start1", "end1", "n_group1", "start2", "end2", "n_group2")]

    def astuple(self):
        return (self.n_local, self.start1, self.end1, self.n_group1, self.start2, self.end2, self.n_group2)


class PiaS(C.Structure):
    _fields_ = [("n_cells", i64), ("n_species", i64), ("ind", P(Ind)), ("n_total", P(i64)),
                ("contiguous", P(C.c_int8))]


class PvS(C.Structure):
    _fields_ = [("length", i64), ("cap", i64), ("w", P(f64)), ("v", P(f64)), ("x", P(f64)),
                ("index", P(i64)), ("cell", P(i64)), ("buffer", P(i64)), ("nbuffer", i64)]


class Grid1D(C.Structure):
    _fields_ = [("L", f64), ("n_cells", i64), ("dx", f64), ("inv_dx", f64), ("min_x", f64), ("max_x", f64)]


class Walls(C.Structure):
    _fields_ = [("T", f64 * 2), ("v", (f64 * 3) * 2), ("acc", f64 * 2), ("n_species", i64), ("vsq", f64 * 32)]


class SurfS(C.Structure):
    _fields_ = [("n_elements", i64), ("n_species", i64), ("inv_areas", P(f64)), ("normals", P(f64)),
                ("q", P(f64))]


class CFSwpm(C.Structure):
    _fields_ = [("n1", i64), ("n2", i64), ("sigma_g_max", f64), ("n_coll", i64), ("n_coll_performed", i64)]


class Interaction(C.Structure):
    _fields_ = [(n, f64) for n in ("m_r", "mu1", "mu2", "vhs_d", "vhs_o", "vhs_Tref", "vhs_muref", "vhs_factor")]


class CF(C.Structure):
    _fields_ = [("n1", i64), ("n2", i64), ("sigma_g_w_max", f64), ("n_coll", i64),
                ("n_coll_performed", i64), ("n_eq_w_coll_performed", i64)]


class PropsS(C.Structure):
    _fields_ = [("n_cells", i64), ("n_species", i64), ("n_moments", i64), ("ndens_not_Np", i32),
                ("Tref", f64), ("lpa", P(f64)), ("np", P(f64)), ("n", P(f64)), ("v", P(f64)),
                ("T", P(f64)), ("moments", P(f64)), ("moment_powers", P(i32))]


class OBin(C.Structure):
    _fields_ = [("np", i64), ("w", f64), ("v_min", f64 * 3), ("v_max", f64 * 3), ("depth", i64),
                ("can_be_refined", i32)]


class OFull(C.Structure):
    _fields_ = [("v_mean", f64 * 3), ("v_std_sq", f64 * 3), ("x_mean", f64 * 3), ("x_std_sq", f64 * 3),
                ("particle_index1", i64), ("particle_index2", i64), ("w1", f64), ("w2", f64),
                ("v1", f64 * 3), ("v2", f64 * 3), ("x1", f64 * 3), ("x2", f64 * 3)]


class OctreeS(C.Structure):
    _fields_ = [("max_Nbins", i64), ("Nbins", i64), ("n_particles", i64), ("bins", P(OBin)),
                ("full", P(OFull)), ("bin_start", P(i64)), ("bin_end", P(i64)), ("pis", P(i64)),
                ("poct", P(i64)), ("psort", P(i64)), ("buf_len", i64), ("bin_bounds_compute", i32),
                ("split", i32), ("init_bin_bounds", i32), ("max_depth", i64),
                ("total_post_merge_np", i64), ("vel_middle", f64 * 3), ("v_min_parent", f64 * 3),
                ("v_max_parent", f64 * 3), ("direction_vec", f64 * 3)]


class NNLSMergeS(C.Structure):
    _fields_ = [("v0", f64 * 3), ("x0", f64 * 3), ("vref", f64), ("inv_vref", f64),
                ("Ex", f64), ("Ey", f64), ("Ez", f64), ("Epx", f64), ("Epy", f64), ("Epz", f64),
                ("w_total", f64), ("minvx", f64), ("maxvx", f64), ("minvy", f64), ("maxvy", f64),
                ("minvz", f64), ("maxvz", f64), ("scalevx", f64), ("scalevy", f64), ("scalevz", f64),
                ("scalex", f64), ("scaley", f64), ("scalez", f64), ("n_total_conserved", i64),
                ("n_moments_vel", i64), ("n_moments_pos", i64), ("mim", (i32 * 3) * 64),
                ("tot_order", i32 * 64), ("mim_pos", (i32 * 3) * 64), ("pos_i_x", i64),
                ("pos_i_y", i64), ("pos_i_z", i64), ("rhs", f64 * 128), ("rnorm", f64),
                ("nsetp", i64), ("lhs_ncols", i64), ("mode", i32), ("reset_x0", i32)]


def build():
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "mzo_philox4x32_10": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
            "mzo_rand": (f64, [P(Rng)]),
            "mzo_pv_new": (P(PvS), [i64]),
            "mzo_pv_free": (None, [P(PvS)]),
            "mzo_pv_resize": (None, [P(PvS), i64]),
            "mzo_pv_buffer_pop_to": (None, [P(PvS), i64]),
            "mzo_pia_new": (P(PiaS), [i64, i64]),
            "mzo_pia_free": (None, [P(PiaS)]),
            "mzo_update_buffer_index_new_particle": (None, [P(PvS), P(PiaS), i64, i64]),
            "mzo_delete_particle_end": (None, [P(PvS), P(PiaS), i64, i64]),
            "mzo_squash_pia": (None, [P(PvS), P(PiaS), i64]),
            "mzo_check_pia_is_correct": (C.c_int, [P(PiaS), i64, P(i64)]),
            "mzo_check_unique_index": (C.c_int, [P(PvS), P(PiaS), i64, P(i64)]),
            "mzo_check_unique_buffer": (C.c_int, [P(PvS), P(i64)]),
            "mzo_count_disordered_particles": (i64, [P(PvS), P(PiaS), i64, C.c_int]),
            "mzo_restore_particle_ordering": (None, [P(PvS)]),
            "mzo_grid1d_init": (None, [P(Grid1D), f64, i64, f64]),
            "mzo_sort_particles": (None, [P(Grid1D), P(PvS), P(PiaS), i64]),
            "mzo_sort_particles_precomputed": (None, [P(PvS), P(PiaS), i64]),
            "mzo_walls_init": (None, [P(Walls), P(f64), i64, f64, f64, f64, f64, f64, f64]),
            "mzo_convect_particles": (None, [P(Rng), P(Grid1D), P(Walls), P(PvS), P(PiaS), i64, f64]),
            "mzo_props_avg": (None, [P(PropsS), P(PropsS), f64]),
            "mzo_delete_particle": (None, [P(PvS), P(PiaS), i64, i64, i64]),
            "mzo_swpm": (None, [P(Rng), P(CFSwpm), P(Interaction), P(PvS), P(PiaS), i64, i64, f64, f64, f64]),
            "mzo_merge_roulette": (None, [P(Rng), P(PvS), P(PiaS), i64, i64, i64, C.c_int]),
            "mzo_gridmerge_new": (P(GridMergeS), [i64, i64, i64, P(f64)]),
            "mzo_gridmerge_free": (None, [P(GridMergeS)]),
            "mzo_gridmerge_extent_props": (None, [P(GridMergeS), f64, P(f64), f64]),
            "mzo_gridmerge_extent": (None, [P(GridMergeS), P(f64), P(f64), P(f64)]),
            "mzo_gridmerge_index": (i64, [P(GridMergeS), P(f64)]),
            "mzo_merge_grid_based": (None, [P(Rng), P(GridMergeS), P(PvS), P(PiaS), i64, i64, P(Grid1D)]),
            "mzo_surf_update_incident": (None, [P(SurfS), f64, P(f64), i64, i64]),
            "mzo_surf_update_reflected": (None, [P(SurfS), f64, P(f64), i64, i64]),
            "mzo_surf_scale": (None, [P(SurfS), i64, f64, f64]),
            "mzo_surf_clear": (None, [P(SurfS)]),
            "mzo_surf_avg": (None, [P(SurfS), P(SurfS), f64]),
            "mzo_convect_particles_surf": (None, [P(Rng), P(Grid1D), P(Walls), P(PvS), P(PiaS), i64, f64, P(SurfS), f64]),
            "mzo_interaction_init": (None, [P(Interaction), P(Interaction), f64, f64, f64, f64, f64]),
            "mzo_sigma_vhs": (f64, [P(Interaction), f64]),
            "mzo_estimate_sigma_g_w_max": (f64, [P(Interaction), f64, f64, f64, f64, f64, f64]),
            "mzo_ntc_single": (None, [P(Rng), P(CF), P(Interaction), P(PvS), P(PiaS), i64, i64, f64, f64, f64, C.c_int]),
            "mzo_ntc_two": (None, [P(Rng), P(CF), P(Interaction), P(PvS), P(PvS), P(PiaS), i64, i64, i64, f64, f64, f64, C.c_int]),
            "mzo_collide_pair": (None, [P(Rng), P(CF), P(Interaction), P(PvS), P(PvS), P(PiaS), i64, i64, i64, i64, i64, f64, C.c_int]),
            "mzo_compute_props": (None, [P(P(PvS)), P(PiaS), P(f64), P(PropsS)]),
            "mzo_compute_flux_props": (None, [P(P(PvS)), P(PiaS), P(f64), P(PropsS), P(f64), P(f64), i64, i64]),
            "mzo_compute_props_sorted": (None, [P(P(PvS)), P(PiaS), P(f64), P(PropsS), P(f64), i64, i64]),
            "mzo_compute_props_with_total_moments": (None, [P(P(PvS)), P(PiaS), P(f64), P(PropsS)]),
            "mzo_compute_mixed_moment": (f64, [P(PvS), P(PiaS), i64, i64, C.c_int, C.c_int, C.c_int, f64, f64]),
            "mzo_octree_new": (P(OctreeS), [i32, i32, i32, i64, i64]),
            "mzo_octree_free": (None, [P(OctreeS)]),
            "mzo_octree_init": (None, [P(OctreeS), P(PvS), P(PiaS), i64, i64]),
            "mzo_octree_split_bin": (None, [P(OctreeS), i64, P(PvS)]),
            "mzo_octree_compute_bin_props": (None, [P(OctreeS), i64, P(PvS)]),
            "mzo_octree_compute": (None, [P(OctreeS), P(PvS), i64]),
            "mzo_merge_octree_N2_based": (None, [P(Rng), P(OctreeS), P(PvS), P(PiaS), i64, i64, i64, P(Grid1D)]),
            "mzo_nnls_solve": (C.c_int, [P(f64), i64, i64, P(f64), P(f64), P(f64), P(f64), P(i64), i64, f64, P(f64), P(i64)]),
            "mzo_scale_columns": (None, [P(f64), i64, i64, P(f64)]),
            "mzo_nnls_merge_init": (None, [P(NNLSMergeS), P(i32), i64, P(i32), i64]),
            "mzo_merge_nnls_based": (C.c_int, [P(Rng), P(NNLSMergeS), P(PvS), P(PiaS), i64, i64, f64, i32, i64, f64, i32, P(f64), i64, i64, f64]),
            "mzo_merge_nnls_cand": (C.c_int, [P(Rng), P(NNLSMergeS), P(PvS), P(PiaS), i64, i64, f64, i32, i64, f64, i32, P(f64), i64, i64, f64, i64]),
            "mzo_sample_maxwellian": (None, [P(Rng), P(PvS), i64, i64, f64, f64, P(f64)]),
            "mzo_sample_particles_equal_weight": (None, [P(Rng), P(PvS), P(PiaS), i64, i64, i64, f64, f64, f64, f64, f64, f64, f64, f64, f64]),
            "mzo_sample_bkw": (None, [P(Rng), P(PvS), i64, i64, f64, f64, P(f64)]),
            "mzo_omp_max_threads": (C.c_int, []),
            "mzo_couette_steps": (None, [P(Rng), P(Grid1D), P(Walls), P(Interaction), P(CF), P(PvS), P(PiaS), f64, P(f64), f64, P(PropsS), i64, C.c_uint64, i32, i32, i32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(P(f64))


# --------------------------------------------------------------------------
class RNG:
    """Draw source: Philox (mode 0) or a replay table [stream, draw] (mode 1)."""

    def __init__(self, seed=0, tag=0, table=None):
        self.s = Rng()
        self.s.seed = seed
        self.s.tag = tag
        self._table = None
        if table is not None:
            self._table = np.ascontiguousarray(table, dtype=np.float64)
            self.s.mode = 1
            self.s.table = _dptr(self._table)
            self.s.rows, self.s.cols = self._table.shape

    @property
    def ref(self):
        return C.byref(self.s)

    def draws(self, stream, n):
        self.s.stream = stream
        self.s.k = 0
        return np.array([lib().mzo_rand(self.ref) for _ in range(n)])


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().mzo_philox4x32_10(c, k, o)
    return list(o)


class ParticleVector:
    """Reference-layout particle store (data by physical id, index by logical position)."""

    def __init__(self, np_):
        self.p = lib().mzo_pv_new(np_)

    def __del__(self):
        try:
            lib().mzo_pv_free(self.p)
        except Exception:
            pass

    @property
    def s(self):
        return self.p.contents

    def _view(self, ptr, n, dt, shape=None):
        if n == 0:
            return np.zeros(0 if shape is None else (0,) + shape[1:], dtype=dt)
        a = np.ctypeslib.as_array(ptr, shape=(n,) if shape is None else shape)
        return a

    @property
    def length(self):
        return self.s.length

    @property
    def w(self):
        return self._view(self.s.w, self.length, np.float64)

    @property
    def v(self):
        return self._view(self.s.v, 3 * self.length, np.float64, (self.length, 3))

    @property
    def x(self):
        return self._view(self.s.x, 3 * self.length, np.float64, (self.length, 3))

    @property
    def index(self):
        return self._view(self.s.index, self.length, np.int64)

    @property
    def cell(self):
        return self._view(self.s.cell, self.length, np.int64)

    @property
    def buffer(self):
        return self._view(self.s.buffer, self.length, np.int64)

    @property
    def nbuffer(self):
        return self.s.nbuffer

    def pop_to(self, position):
        lib().mzo_pv_buffer_pop_to(self.p, position)

    def set(self, position, w, v, x):
        """ParticleVector[position] = Particle(w, v, x)"""
        pid = self.index[position - 1]
        self.w[pid - 1] = w
        self.v[pid - 1] = v
        self.x[pid - 1] = x

    def get(self, position):
        pid = self.index[position - 1]
        return self.w[pid - 1], self.v[pid - 1].copy(), self.x[pid - 1].copy()

    def resize(self, n):
        lib().mzo_pv_resize(self.p, n)

    def logical_arrays(self, n_total):
        """(x, v, w, id) in logical order for positions 1..n_total."""
        ids = self.index[:n_total].copy()
        return self.x[ids - 1].copy(), self.v[ids - 1].copy(), self.w[ids - 1].copy(), ids


class PIA:
    def __init__(self, n_cells, n_species=1):
        self.p = lib().mzo_pia_new(n_cells, n_species)
        self.n_cells = n_cells
        self.n_species = n_species

    def __del__(self):
        try:
            lib().mzo_pia_free(self.p)
        except Exception:
            pass

    def ind(self, cell, species=1):
        return self.p.contents.ind[(species - 1) * self.n_cells + (cell - 1)]

    def set(self, cell, species, n_local, start1, end1, n_group1, start2=0, end2=-1, n_group2=0):
        d = self.ind(cell, species)
        d.n_local, d.start1, d.end1, d.n_group1 = n_local, start1, end1, n_group1
        d.start2, d.end2, d.n_group2 = start2, end2, n_group2

    def n_total(self, species=1):
        return self.p.contents.n_total[species - 1]

    def set_n_total(self, species, n):
        self.p.contents.n_total[species - 1] = n

    def contiguous(self, species=1):
        return bool(self.p.contents.contiguous[species - 1])

    def set_contiguous(self, species, flag):
        self.p.contents.contiguous[species - 1] = 1 if flag else 0

    def table(self, species=1):
        """int64 array [n_cells, 7] of (n_local, start1, end1, n_group1, start2, end2, n_group2)."""
        return np.array([self.ind(c, species).astuple() for c in range(1, self.n_cells + 1)], dtype=np.int64)


def grid1d(L, nx, wall_offset=1e-12):
    g = Grid1D()
    lib().mzo_grid1d_init(C.byref(g), L, nx, wall_offset)
    return g


def walls(masses, T_l, T_r, vy_l, vy_r, acc_l, acc_r):
    w = Walls()
    m = np.ascontiguousarray(masses, dtype=np.float64)
    lib().mzo_walls_init(C.byref(w), _dptr(m), len(m), T_l, T_r, vy_l, vy_r, acc_l, acc_r)
    return w


class Surf:
    """SurfProps (surface_props.jl:22-64) over the oracle's flat layout
    q[species-1, element-1, k], k = np, flux_incident, flux_reflected,
    force x3, normal_pressure, shear_pressure x3, kinetic_energy_flux."""

    Q = 11

    def __init__(self, n_elements, n_species, areas, normals):
        self.n_elements, self.n_species = n_elements, n_species
        self.inv_areas = np.ascontiguousarray([1.0 / a for a in areas], dtype=np.float64)
        self.normals = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).reshape(n_elements, 3))
        self.q = np.zeros((n_species, n_elements, self.Q))
        self.s = SurfS(n_elements, n_species, _dptr(self.inv_areas), _dptr(self.normals), _dptr(self.q))

    @classmethod
    def grid1d(cls, n_species):
        """SurfProps(pia, grid::Grid1DUniform): two unit-area walls, normals +x and -x."""
        return cls(2, n_species, [1.0, 1.0], [[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])

    @property
    def ref(self):
        return C.byref(self.s)

    # the reference's arrays: (n_elements, n_species) and (3, n_elements, n_species)
    def _f(self, k):
        return self.q[:, :, k].T

    np = property(lambda self: self._f(0))
    flux_incident = property(lambda self: self._f(1))
    flux_reflected = property(lambda self: self._f(2))
    force = property(lambda self: np.transpose(self.q[:, :, 3:6], (2, 1, 0)))
    normal_pressure = property(lambda self: self._f(6))
    shear_pressure = property(lambda self: np.transpose(self.q[:, :, 7:10], (2, 1, 0)))
    kinetic_energy_flux = property(lambda self: self._f(10))

    def update_incident(self, w, v, species, element):
        vv = np.ascontiguousarray(v, dtype=np.float64)
        lib().mzo_surf_update_incident(self.ref, w, _dptr(vv), species, element)

    def update_reflected(self, w, v, species, element):
        vv = np.ascontiguousarray(v, dtype=np.float64)
        lib().mzo_surf_update_reflected(self.ref, w, _dptr(vv), species, element)

    def scale(self, species, mass, dt):
        lib().mzo_surf_scale(self.ref, species, mass, dt)

    def clear(self):
        lib().mzo_surf_clear(self.ref)

    def avg_from(self, inst, n_avg_timesteps):
        lib().mzo_surf_avg(self.ref, inst.ref, float(n_avg_timesteps))


def convect_particles_surf(rng, grid, walls, pv, pia, species, dt, surf, mass):
    """convect_particles!(..., surf_props, dt) (convection_1D.jl:176-207)"""
    lib().mzo_convect_particles_surf(rng.ref, C.byref(grid), C.byref(walls), pv.p, pia.p, species, dt, surf.ref,
                                     mass)


def interaction_pair(m_i, m_k, d, omega, Tref):
    a, b = Interaction(), Interaction()
    lib().mzo_interaction_init(C.byref(a), C.byref(b), m_i, m_k, d, omega, Tref)
    return a, b


class Props:
    def __init__(self, n_cells, n_species, moment_powers=(), ndens_not_Np=False, Tref=300.0):
        self.lpa = np.zeros(n_species)
        self.np = np.zeros((n_species, n_cells))
        self.n = np.zeros((n_species, n_cells))
        self.v = np.zeros((n_species, n_cells, 3))
        self.T = np.zeros((n_species, n_cells))
        nm = len(moment_powers)
        self.moments = np.zeros((n_species, n_cells, max(nm, 1)))
        self.mp = np.ascontiguousarray(moment_powers if nm else [0], dtype=np.int32)
        self.s = PropsS(n_cells, n_species, nm, 1 if ndens_not_Np else 0, Tref, _dptr(self.lpa),
                        _dptr(self.np), _dptr(self.n), _dptr(self.v), _dptr(self.T), _dptr(self.moments),
                        self.mp.ctypes.data_as(P(i32)))

    @property
    def ref(self):
        return C.byref(self.s)


def pv_array(pvs):
    arr = (P(PvS) * len(pvs))(*[p.p for p in pvs])
    return arr


def compute_props(pvs, pia, masses, props):
    m = np.ascontiguousarray(masses, dtype=np.float64)
    lib().mzo_compute_props(pv_array(pvs), pia.p, _dptr(m), props.ref)


def compute_flux_props(pvs, pia, masses, props, inv_V, cells=None):
    """[species, cell, 9] array: energy flux xyz, diagonal xyz, off-diagonal xy xz yz"""
    m = np.ascontiguousarray(masses, dtype=np.float64)
    iv = np.ascontiguousarray(inv_V, dtype=np.float64)
    out = np.zeros((pia.n_species, pia.n_cells, 9))
    lo, hi = (1, pia.n_cells) if cells is None else cells
    lib().mzo_compute_flux_props(pv_array(pvs), pia.p, _dptr(m), props.ref, _dptr(iv), _dptr(out), lo, hi)
    return out


def compute_props_sorted(pvs, pia, masses, props, inv_V=None, cells=None):
    m = np.ascontiguousarray(masses, dtype=np.float64)
    iv = None if inv_V is None else np.ascontiguousarray(inv_V, dtype=np.float64)
    lo, hi = (1, pia.n_cells) if cells is None else cells
    lib().mzo_compute_props_sorted(pv_array(pvs), pia.p, _dptr(m), props.ref,
                                   None if iv is None else _dptr(iv), lo, hi)


def compute_props_with_total_moments(pvs, pia, masses, props):
    m = np.ascontiguousarray(masses, dtype=np.float64)
    lib().mzo_compute_props_with_total_moments(pv_array(pvs), pia.p, _dptr(m), props.ref)


class GCell(C.Structure):
    _fields_ = [("np", i64), ("w", f64), ("v_mean", f64 * 3), ("v_std_sq", f64 * 3), ("x_mean", f64 * 3),
                ("x_std_sq", f64 * 3), ("particle_index1", i64), ("particle_index2", i64)]


class GridMergeS(C.Structure):
    _fields_ = [("Nx", i64), ("Ny", i64), ("Nz", i64), ("NyNz", i64), ("Ntotal", i64), ("mult", f64 * 3),
                ("lower", f64 * 3), ("upper", f64 * 3), ("mid", f64 * 3), ("dv", f64 * 3), ("dv_inv", f64 * 3),
                ("cells", P(GCell)), ("out_of_range", i64)]


class GridMerge:
    """GridN2Merge (src/merging/merging_grid.jl:72-115)."""

    def __init__(self, Nx, Ny, Nz, mult):
        m = (C.c_double * 3)(*([float(mult)] * 3 if np.isscalar(mult) else [float(x) for x in mult]))
        self.p = lib().mzo_gridmerge_new(Nx, Ny, Nz, m)

    def __del__(self):
        try:
            lib().mzo_gridmerge_free(self.p)
        except Exception:
            pass

    @property
    def s(self):
        return self.p.contents

    def extent_props(self, T, v, mass):
        lib().mzo_gridmerge_extent_props(self.p, T, (C.c_double * 3)(*v), mass)

    def extent(self, vx, vy, vz):
        lib().mzo_gridmerge_extent(self.p, (C.c_double * 2)(*vx), (C.c_double * 2)(*vy), (C.c_double * 2)(*vz))

    def index(self, v):
        return lib().mzo_gridmerge_index(self.p, (C.c_double * 3)(*v))

    def cell(self, k):
        """velocity cell k (1-based)"""
        return self.s.cells[k - 1]


class Octree:
    MID, MEAN = 1, 2
    INIT_MINMAX, INIT_MINMAX_SYM, INIT_C = 1, 2, 3
    INHERIT, RECOMPUTE = 1, 2

    def __init__(self, split=1, init_bin_bounds=1, bin_bounds_compute=1, max_Nbins=4096, max_depth=10):
        self.p = lib().mzo_octree_new(split, init_bin_bounds, bin_bounds_compute, max_Nbins, max_depth)

    def __del__(self):
        try:
            lib().mzo_octree_free(self.p)
        except Exception:
            pass

    @property
    def s(self):
        return self.p.contents

    def bin(self, b):
        return self.s.bins[b]

    def full(self, b):
        return self.s.full[b]

    def pis(self, n):
        return np.array([self.s.pis[i] for i in range(n)], dtype=np.int64)

    def bin_start(self, n):
        return np.array([self.s.bin_start[i] for i in range(1, n + 1)], dtype=np.int64)

    def bin_end(self, n):
        return np.array([self.s.bin_end[i] for i in range(1, n + 1)], dtype=np.int64)


def multi_index_moments(n):
    """compute_multi_index_moments(n) (src/merging/merging_nnls.jl:341-352)"""
    out = []
    for i in range(n + 1):
        for j in range(n - i + 1):
            k = n - i - j
            if k >= 0:
                out.append((i, j, k))
    return out


class NNLSMerge:
    def __init__(self, mim, mim_pos=()):
        self.s = NNLSMergeS()
        a = np.ascontiguousarray(np.array(mim, dtype=np.int32).reshape(-1, 3)) if len(mim) else np.zeros((1, 3), np.int32)
        b = np.ascontiguousarray(np.array(mim_pos, dtype=np.int32).reshape(-1, 3)) if len(mim_pos) else np.zeros((1, 3), np.int32)
        lib().mzo_nnls_merge_init(C.byref(self.s), a.ctypes.data_as(P(i32)), len(mim), b.ctypes.data_as(P(i32)), len(mim_pos))

    def merge(self, rng, pv, pia, cell, species, vref=1.0, scaling="variance", n_rand_pairs=0,
              max_err=1e-11, centered_at_mean=True, v_multipliers=(0.5, 1.0), iteration_mult=2,
              w_threshold=0.0, max_pcols=-1):
        """merge_nnls_based!; max_pcols >= 0 keeps only the first max_pcols
        particles (logical order) as columns, all particles still define the
        moments."""
        vm = np.ascontiguousarray(v_multipliers, dtype=np.float64)
        return lib().mzo_merge_nnls_cand(rng.ref, C.byref(self.s), pv.p, pia.p, cell, species, vref,
                                         1 if scaling == "vref" else 0, n_rand_pairs, max_err,
                                         1 if centered_at_mean else 0, _dptr(vm), len(vm),
                                         iteration_mult, w_threshold, max_pcols)


def nnls_solve(A, b, max_iter=None, kkt_tol=1e-16):
    A = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    m, n = A.shape
    bb = np.ascontiguousarray(b, dtype=np.float64).copy()
    x = np.zeros(n); w = np.zeros(n); zz = np.zeros(m); idx = np.zeros(n, dtype=np.int64)
    rn = f64(0); ns = i64(0)
    mode = lib().mzo_nnls_solve(_dptr(A), m, n, _dptr(bb), _dptr(x), _dptr(w), _dptr(zz),
                                idx.ctypes.data_as(P(i64)), max_iter if max_iter else 3 * n, kkt_tol,
                                C.byref(rn), C.byref(ns))
    return x, rn.value, ns.value, mode


def scale_columns(A):
    A = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    m, n = A.shape
    cn = np.zeros(n)
    lib().mzo_scale_columns(_dptr(A), m, n, _dptr(cn))
    return A, cn


# ---- species / interaction data restated from the reference's data/*.toml ----
MASS = {"Ar": 66.3e-27, "He": 6.65e-27}          # data/particles.toml
VHS = {("Ar", "Ar"): (4.11e-10, 0.81, 273.0),    # data/vhs.toml
       ("Ar", "He"): (3.25e-10, 0.735, 273.0),
       ("He", "He"): (2.33e-10, 0.66, 273.0)}
PSEUDO_MAXWELL = {("Ar", "Ar"): (4.11e-10, 1.0, 273.0),   # data/pseudo_maxwell.toml
                  ("Ar", "He"): (3.25e-10, 1.0, 273.0),
                  ("He", "He"): (2.33e-10, 1.0, 273.0)}
KB = 1.380649e-23


def interactions(names, table=VHS):
    """n x n list of Interaction as load_interaction_data (collision_utils.jl:159-202)."""
    n = len(names)
    out = [[None] * n for _ in range(n)]
    for i in range(n):
        for k in range(n):
            if i >= k:
                key = (names[i], names[k]) if (names[i], names[k]) in table else (names[k], names[i])
                d, o, tr = table[key]
                a, b = interaction_pair(MASS[names[i]], MASS[names[k]], d, o, tr)
                out[i][k] = a
                out[k][i] = b
    return out
