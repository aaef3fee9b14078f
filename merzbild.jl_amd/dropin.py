"""The reference's call signatures over the device store.

A Merzbild time loop calls, per step (test/test_1D_couette.jl:66-97),

    for cell in 1:grid.n_cells
        ntc!(rng, collision_factors[1, 1, cell], collision_data, interaction_data,
             particles[1], pia, cell, 1, dt, grid.cells[cell].V)
    end
    convect_particles!(rng, grid, boundaries, particles[1], pia, 1, species_data, dt)
    sort_particles!(gridsorter, grid, particles[1], pia, 1)
    compute_props_sorted!(particles, pia, species_data, phys_props)

This module keeps those positional signatures (names without the "!",
which Python does not allow) over device-backed stand-ins for the
reference's structs: DeviceParticleVector, DevicePIA, CollisionFactorsArray,
CollisionData, GridSortInPlace, OctreeN2Merge, NNLSMerge and DeviceRNG.

Batching.  A per-cell ntc! call does not launch: it queues the cell on the
store.  The next call of any other kind (convection, sort, props, a merge,
or reading the particles, the PIA or the collision factors) first flushes
the queue as one mz_ntc_cells call, so the reference's loop over cells
becomes one launch.  Queued calls with the same pair and parameters that
follow one another batch together; the flush keeps the calls' order, so the
result equals the per-cell calls one by one.  merge_octree_N2_based! and
merge_nnls_based! calls queue the same way and flush as one launch per run
of calls with the same parameters.

Random numbers.  The device draws counter-based Philox streams keyed by
(seed, tag, cell) for collisions and merges and (seed, tag, particle id)
for convection, not a sequential generator.  A DeviceRNG hands out tags
(step << 4) | phase in the order of the calls: within a step, NTC of the
first pair gets phase 1, convection phase 2, the first merge kind phase 3,
further pairs and merge kinds phases 4-15; a step ends when a call would
draw a stream it already drew in this step (the same pair in the same cell,
convection of the same species, the same merge in the same cell).  So the
loop above draws exactly what DeviceStore-level calls with those tags
draw, whether its calls flush per cell or per step.  Julia's StableRNG
stream itself is not reproduced; parity with the reference is through
replay tables and statistics.

Errors are raised as MzError (the reference's exceptions); merge_nnls_based
returns 1 or -1 per call as the reference does.
"""
import numpy as np

from . import (MzError, PhysProps as _PhysProps, _check, _d, estimate_sigma_g_w_max, i32, load_library)
from . import merge_octree_cells as _merge_octree_cells
from . import merge_octree_targets as _merge_octree_targets
from . import merge_nnls_cells as _merge_nnls_cells

P = __import__("ctypes").POINTER

__all__ = ["DeviceRNG", "DeviceParticleVector", "DevicePIA", "CollisionFactorsArray", "CollisionData",
           "GridSortInPlace", "OctreeN2Merge", "NNLSMerge", "create_collision_factors_array", "ntc",
           "ntc_equal_weight", "convect_particles", "sort_particles", "compute_props", "compute_props_sorted",
           "compute_mixed_moment",
           "merge_octree_N2_based", "merge_nnls_based", "squash_pia", "restore_particle_ordering", "PhysProps", "flush", "MergeStatus"]

PHASE_NTC, PHASE_CONVECT, PHASE_MERGE = 1, 2, 3
_NTC_EXTRA = tuple(range(4, 16))


class DeviceRNG:
    """The rng argument: a Philox seed and the tag counter described above."""

    _PREFERRED = {"ntc": PHASE_NTC, "convect": PHASE_CONVECT, "merge": PHASE_MERGE}

    def __init__(self, seed, step=1):
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.step = int(step)
        self._used = set()
        self._phases = {}

    def _advance(self):
        self.step += 1
        self._used.clear()
        self._phases = {}

    def tag(self, kind, key=None, cells=None):
        """The tag of a launch of `kind` ("ntc", "convect", "merge") with
        `key` (the NTC pair, the species, the merge kind) over `cells`
        (None: every particle of the species)."""
        streams = [(kind, key, c) for c in (cells if cells is not None else (None,))]
        if any(st in self._used for st in streams):
            self._advance()
        phase = self._phases.get((kind, key))
        if phase is None:
            taken = set(self._phases.values())
            phase = self._PREFERRED[kind]
            if phase in taken:
                free = [p for p in _NTC_EXTRA if p not in taken]
                if free:
                    phase = free[0]
                else:
                    self._advance()
                    phase = self._PREFERRED[kind]
            self._phases[(kind, key)] = phase
        self._used.update(streams)
        return (self.step << 4) | phase


# ---------------------------------------------------------------------------
# the per-store queue of per-cell calls

class _Queue:
    def __init__(self, store):
        self.store = store
        self.items = []   # (kind, key, cell, value, rng)
        self.results = {}

    def push(self, kind, key, cell, value, rng):
        self.items.append((kind, key, cell, value, rng))

    def flush(self):
        items, self.items = self.items, []
        k0 = 0
        while k0 < len(items):
            kind, key, _, _, rng = items[k0]
            k1 = k0 + 1
            while k1 < len(items) and items[k1][0] == kind and items[k1][1] == key and items[k1][4] is rng:
                k1 += 1
            run = items[k0:k1]
            if kind == "ntc":
                self._ntc(key, run, rng)
            elif kind == "octree":
                self._octree(key, run, rng)
            else:
                self._nnls(key, run, rng)
            k0 = k1

    def _seed(self, rng):
        if getattr(self.store, "_dropin_seed", None) != rng.seed:
            self.store.set_seed(rng.seed)
            self.store._dropin_seed = rng.seed

    def _ntc(self, key, run, rng):
        s1, s2, dt, dw_tol, equal_weight = key
        self._seed(rng)
        cells = np.ascontiguousarray([it[2] for it in run], dtype=np.int32)
        V = np.ascontiguousarray([it[3] for it in run], dtype=np.float64)
        tag = rng.tag("ntc", (s1, s2), [int(c) for c in cells])
        _check(load_library().mz_ntc_cells(self.store.h, s1, s2, float(dt), cells.ctypes.data_as(P(i32)), _d(V),
                                           len(cells), float(dw_tol), 1 if equal_weight else 0, tag, 0))
        _invalidate(self.store)

    def _octree(self, key, run, rng):
        oc, species, use_grid = key
        self._seed(rng)
        cells = [it[2] for it in run]
        targets = [it[3] for it in run]
        tag = rng.tag("merge", ("octree", species), cells)
        if len(set(targets)) == 1:
            _merge_octree_cells(self.store, species, cells, targets[0], oc.split, oc.init_bin_bounds,
                                oc.bin_bounds_compute, oc.max_Nbins, oc.max_depth, use_grid, tag=tag)
        else:
            # per-cell targets: each run of ascending cells is one launch that
            # merges exactly its cells (threshold -1), each down to its own target
            nc = self.store.n_cells
            k0 = 0
            while k0 < len(cells):
                k1 = k0 + 1
                while k1 < len(cells) and cells[k1] > cells[k1 - 1]:
                    k1 += 1
                thr = np.full(nc, np.iinfo(np.int32).max, dtype=np.int32)
                tgt = np.zeros(nc, dtype=np.int32)
                for c, t in zip(cells[k0:k1], targets[k0:k1]):
                    thr[c - 1] = -1
                    tgt[c - 1] = t
                _merge_octree_targets(self.store, species, thr, tgt, oc.split, oc.init_bin_bounds,
                                      oc.bin_bounds_compute, oc.max_Nbins, oc.max_depth, use_grid, squash=False,
                                      tag=tag)
                k0 = k1
        _invalidate(self.store)

    def _nnls(self, key, run, rng):
        nm, species, opts = key
        o = dict(opts)
        self._seed(rng)
        cells = [it[2] for it in run]
        tag = rng.tag("merge", ("nnls", species), cells) if o["n_rand_pairs"] else 0
        status = _merge_nnls_cells(self.store, species, cells, nm.mim, vref=o["vref"], scaling=o["scaling"],
                                   n_rand_pairs=o["n_rand_pairs"], centered_at_mean=o["centered_at_mean"],
                                   v_multipliers=o["v_multipliers"], max_err=o["max_err"],
                                   iteration_mult=o["iteration_mult"], w_threshold=o["w_threshold"],
                                   multi_index_moments_pos=nm.mim_pos, tag=tag)
        for it, st in zip(run, status):
            it[3].value = int(st)
        _invalidate(self.store)


def _queue(store):
    q = getattr(store, "_dropin_queue", None)
    if q is None:
        q = _Queue(store)
        store._dropin_queue = q
    return q


def _invalidate(store):
    store._dropin_cache = {}


def flush(store):
    """Launch every queued per-cell call of the store."""
    _queue(store).flush()


def _cached(store, key, fetch):
    flush(store)
    c = getattr(store, "_dropin_cache", None)
    if c is None:
        c = store._dropin_cache = {}
    if key not in c:
        c[key] = fetch()
    return c[key]


# ---------------------------------------------------------------------------
# device-backed stand-ins for the reference's structs

class Particle:
    """Particle(w, v, x) (src/particles.jl:14-18), a host copy."""
    __slots__ = ("w", "v", "x")

    def __init__(self, w, v, x):
        self.w, self.v, self.x = w, v, x


class DeviceParticleVector:
    """ParticleVector of one species (src/particles.jl:194-243) held in a
    DeviceStore.  pv[i] (1-based logical position) is particles[index[i]];
    reading flushes queued calls and downloads the species once until the
    next call that changes it."""

    def __init__(self, store, species):
        self.store, self.species = store, species

    def _d(self):
        return _cached(self.store, ("species", self.species), lambda: self.store.download_species(self.species))

    def __getitem__(self, i):
        d = self._d()
        return Particle(d["w"][i - 1], d["v"][i - 1].copy(), d["x"][i - 1].copy())

    def __len__(self):
        return int(self._d()["length"])

    @property
    def index(self):
        return self._d()["index"]

    @property
    def cell(self):
        return self._d()["cell"]

    @property
    def buffer(self):
        return self._d()["buffer"]

    @property
    def nbuffer(self):
        return int(self._d()["nbuffer"])


class _Indexer:
    """ParticleIndexer (src/particles.jl:56-84), a host copy of one row."""
    __slots__ = ("n_local", "start1", "end1", "n_group1", "start2", "end2", "n_group2")

    def __init__(self, row):
        (self.n_local, self.start1, self.end1, self.n_group1, self.start2, self.end2,
         self.n_group2) = (int(v) for v in row)


class _IndexerView:
    def __init__(self, pia):
        self.pia = pia

    def __getitem__(self, idx):
        cell, species = idx
        return _Indexer(self.pia._rows(species)[cell - 1])


class DevicePIA:
    """ParticleIndexerArray (src/particles.jl:104-174) of a DeviceStore:
    pia.indexer[cell, species], pia.n_total[species],
    pia.contiguous[species] read the device state (after a flush)."""

    def __init__(self, store):
        self.store = store
        self.indexer = _IndexerView(self)

    def _sp(self, species):
        return _cached(self.store, ("species", species), lambda: self.store.download_species(species))

    def _rows(self, species):
        return self._sp(species)["pia"]

    @property
    def n_total(self):
        return _Lazy(lambda s: int(self._sp(s)["n_total"]))

    @property
    def contiguous(self):
        return _Lazy(lambda s: bool(self._sp(s)["contiguous"]))


class _Lazy:
    def __init__(self, f):
        self.f = f

    def __getitem__(self, species):
        return self.f(species)


class _CFView:
    """CollisionFactors (src/collisions/collision_ntc.jl:18-32) of one
    (s1, s2, cell)."""

    def __init__(self, arr, s1, s2, cell):
        self.arr, self.s1, self.s2, self.cell = arr, s1, s2, cell

    def _get(self, name):
        return _cached(self.arr.store, ("cf", self.s1, self.s2),
                       lambda: self.arr.store.collision_factors(self.s1, self.s2))[name][self.cell - 1]

    sigma_g_w_max = property(lambda self: float(self._get("sigma_g_w_max")))
    n1 = property(lambda self: int(self._get("n1")))
    n2 = property(lambda self: int(self._get("n2")))
    n_coll = property(lambda self: int(self._get("n_coll")))
    n_coll_performed = property(lambda self: int(self._get("n_coll_performed")))
    n_eq_w_coll_performed = property(lambda self: int(self._get("n_eq_w_coll_performed")))


class CollisionFactorsArray:
    """The array collision_factors[s1, s2, cell] of CollisionFactors, kept on
    the device (the running (sigma g w)_max and the counters)."""

    def __init__(self, store):
        self.store = store

    def __getitem__(self, idx):
        s1, s2, cell = idx
        return _CFView(self, s1, s2, cell)


def create_collision_factors_array(pia, interactions, species_data, T, Fnum, mult_factor=1.0):
    """create_collision_factors_array(pia, interactions, species_data, T,
    Fnum; mult_factor) (src/collisions/collision_ntc.jl:100-155): every
    (s1, s2, cell) starts at estimate_sigma_g_w_max with T for both species."""
    store = pia.store
    ns = store.n_species
    for s1 in range(ns):
        for s2 in range(ns):
            v = estimate_sigma_g_w_max(interactions[s1][s2], species_data[s1].mass, species_data[s2].mass, T, T,
                                       Fnum, mult_factor)
            store.set_sigma_g_w_max(s1 + 1, s2 + 1, np.full(store.n_cells, v))
    _invalidate(store)
    return CollisionFactorsArray(store)


class CollisionData:
    """CollisionData (src/collisions/collision_utils.jl:22-31): per-call
    scratch in the reference; the device keeps its own."""


class GridSortInPlace:
    """GridSortInPlace(grid, n_particles) (src/grids/grid_sorting.jl:10-41):
    the reference's sort buffers; the device store holds its own."""

    def __init__(self, grid, n_particles=0):
        self.grid, self.n_particles = grid, n_particles


OCTREE_SPLIT = {"OctreeBinMidSplit": 1, "OctreeBinMeanSplit": 2}
OCTREE_INIT = {"OctreeInitBinMinMaxVel": 1, "OctreeInitBinMinMaxVelSymmetric": 2, "OctreeInitBinC": 3}
OCTREE_BOUNDS = {"OctreeBinBoundsInherit": 1, "OctreeBinBoundsRecompute": 2}


class OctreeN2Merge:
    """OctreeN2Merge(split; init_bin_bounds, bin_bounds_compute, max_Nbins,
    max_depth) (src/merging/merging_octree_N2.jl:131-246)."""

    def __init__(self, split="OctreeBinMidSplit", init_bin_bounds="OctreeInitBinMinMaxVel",
                 bin_bounds_compute="OctreeBinBoundsInherit", max_Nbins=4096, max_depth=10):
        self.split = OCTREE_SPLIT[split] if isinstance(split, str) else int(split)
        self.init_bin_bounds = OCTREE_INIT[init_bin_bounds] if isinstance(init_bin_bounds, str) \
            else int(init_bin_bounds)
        self.bin_bounds_compute = OCTREE_BOUNDS[bin_bounds_compute] if isinstance(bin_bounds_compute, str) \
            else int(bin_bounds_compute)
        self.max_Nbins, self.max_depth = int(max_Nbins), int(max_depth)


class NNLSMerge:
    """NNLSMerge(multi_index_moments, init_np; rhs_vector_size,
    multi_index_moments_pos) (src/merging/merging_nnls.jl:50-207)."""

    def __init__(self, multi_index_moments, init_np=0, multi_index_moments_pos=()):
        self.mim = np.ascontiguousarray(np.asarray(multi_index_moments, dtype=np.int32).reshape(-1, 3))
        self.mim_pos = np.ascontiguousarray(np.asarray(multi_index_moments_pos, dtype=np.int32).reshape(-1, 3))
        self.init_np = init_np


def PhysProps(pia, moments_list=(), Tref=1.0, ndens_not_Np=False):
    """PhysProps(pia[, moments_list[, Tref]]) (src/properties/physical_props.jl:47-90)."""
    s = pia.store
    return _PhysProps(s.n_cells, s.n_species, moments_list, ndens_not_Np, Tref)


# ---------------------------------------------------------------------------
# operations with the reference's positional signatures

def _store_of(pv):
    if not isinstance(pv, DeviceParticleVector):
        raise MzError("expected a DeviceParticleVector")
    return pv.store


def ntc(rng, collision_factors, collision_data, interaction, *args, dw_tol=1e-16):
    """ntc!(rng, cf, cd, interaction, pv, pia, cell, species, dt, V; dw_tol)
    and ntc!(rng, cf, cd, interaction, pv1, pv2, pia, cell, species1,
    species2, dt, V; dw_tol) (src/collisions/collision_ntc.jl:338, 412):
    queues the cell."""
    _ntc_queue(rng, args, dw_tol, False)


def ntc_equal_weight(rng, collision_factors, collision_data, interaction, *args):
    """ntc_equal_weight!(...) (src/collisions/collision_ntc.jl:479, 554),
    same signatures as ntc!: queues the cell."""
    _ntc_queue(rng, args, 1e-16, True)


def _ntc_queue(rng, args, dw_tol, equal_weight):
    if len(args) == 6:
        pv, pia, cell, species, dt, V = args
        s1 = s2 = species
    elif len(args) == 8:
        pv, pv2, pia, cell, s1, s2, dt, V = args
        if _store_of(pv2) is not _store_of(pv):
            raise MzError("ntc: both species must live in one device store")
    else:
        raise MzError("ntc: expected (pv, pia, cell, species, dt, V) or (pv1, pv2, pia, cell, s1, s2, dt, V)")
    store = _store_of(pv)
    if not 1 <= cell <= store.n_cells:
        raise MzError(f"ntc: cell {cell} out of range 1..{store.n_cells}")
    _queue(store).push("ntc", (int(s1), int(s2), float(dt), float(dw_tol), bool(equal_weight)), int(cell),
                       float(V), rng)


def _ensure_grid(store, grid):
    if store.grid is not grid:
        store.set_grid(grid)


def convect_particles(rng, grid, boundaries, pv, pia, species, species_data, dt):
    """convect_particles!(rng, grid, boundaries, particles, pia, species,
    species_data, dt) (src/convection/convection_1D.jl:130-157)."""
    from . import convect_particles as _convect
    store = _store_of(pv)
    flush(store)
    _ensure_grid(store, grid)
    if getattr(store, "_dropin_walls", None) is not boundaries:
        store.set_walls(boundaries)
        store._dropin_walls = boundaries
    q = _queue(store)
    q._seed(rng)
    _convect(store, species, dt, tag=rng.tag("convect", species), species_data=species_data)
    _invalidate(store)


def sort_particles(gridsort, *args):
    """sort_particles!(gridsort, grid, particles, pia, species) and
    sort_particles!(gridsort, particles, pia, species) with the stored cells
    (src/grids/grid_sorting.jl:58, 128)."""
    from . import sort_particles as _sort
    if len(args) == 4:
        grid, pv, pia, species = args
        store = _store_of(pv)
        flush(store)
        _ensure_grid(store, grid)
        _sort(store, species, compute_cells=True)
    elif len(args) == 3:
        pv, pia, species = args
        store = _store_of(pv)
        flush(store)
        _sort(store, species, compute_cells=False)
    else:
        raise MzError("sort_particles: expected (gridsort, grid, pv, pia, species) or (gridsort, pv, pia, species)")
    _invalidate(store)


def compute_props(particles, pia, species_data, phys_props):
    """compute_props!(particles, pia, species_data, phys_props)
    (src/properties/physical_props.jl:104-154)."""
    from . import compute_props as _cp, compute_props_with_total_moments as _cpm
    store = pia.store
    flush(store)
    if phys_props.n_moments:
        _cpm(store, species_data, phys_props)
    else:
        _cp(store, species_data, phys_props)


def compute_props_sorted(particles, pia, species_data, phys_props):
    """compute_props_sorted!(particles, pia, species_data, phys_props)
    (src/properties/physical_props.jl:317-454)."""
    from . import compute_props_sorted as _cps
    store = pia.store
    flush(store)
    _cps(store, species_data, phys_props, ndens_not_Np=phys_props.ndens_not_Np)


def compute_mixed_moment(particles, pia, cell, species, powers, sum_scaler=1.0, res_scaler=1.0):
    """compute_mixed_moment(particles, pia, cell, species, powers; sum_scaler,
    res_scaler) (src/properties/physical_props.jl:473-501); returns the moment."""
    from . import compute_mixed_moment as _cmm
    store = pia.store
    flush(store)
    return _cmm(store, species, powers, cells=int(cell), sum_scaler=sum_scaler, res_scaler=res_scaler)


def squash_pia(particles, pia, species=None):
    """squash_pia!(particles, pia, species) and squash_pia!(particles, pia)
    for every species (src/particles.jl:622-701)."""
    from . import squash_pia as _sq
    store = pia.store
    flush(store)
    for s in ([species] if species is not None else range(1, store.n_species + 1)):
        _sq(store, s)
    _invalidate(store)


def restore_particle_ordering(pv, inv_map=None):
    """restore_particle_ordering!(pv, inv_map) (src/particles.jl:1086-1137);
    inv_map is the reference's scratch and is not used."""
    from . import restore_particle_ordering as _rpo
    store = _store_of(pv)
    flush(store)
    _rpo(store, pv.species)
    _invalidate(store)


def merge_octree_N2_based(rng, octree, pv, pia, cell, species, target_np, grid=None):
    """merge_octree_N2_based!(rng, octree, particles, pia, cell, species,
    target_np[, grid]) (src/merging/merging_octree_N2.jl:1060, 1088):
    queues the cell (grid: the grid-clamped variant)."""
    store = _store_of(pv)
    _queue(store).push("octree", (octree, int(species), grid is not None), int(cell), int(target_np), rng)


def merge_nnls_based(rng, nnls_merging, pv, pia, cell, species, vref=1.0, scaling="variance", n_rand_pairs=0,
                     max_err=1e-11, centered_at_mean=False, v_multipliers=(), iteration_mult=2,
                     w_threshold=0.0):
    """merge_nnls_based!(rng, nnls_merging, particles, pia, cell, species;
    vref, scaling, n_rand_pairs, max_err, centered_at_mean, v_multipliers,
    iteration_mult, w_threshold) (src/merging/merging_nnls.jl:1457-1510).
    Returns 1 on success and -1 on failure, as the reference; it flushes
    the queue at once, since the caller reads the result."""
    store = _store_of(pv)
    opts = (("vref", float(vref)), ("scaling", scaling), ("n_rand_pairs", int(n_rand_pairs)),
            ("max_err", float(max_err)), ("centered_at_mean", bool(centered_at_mean)),
            ("v_multipliers", tuple(float(v) for v in v_multipliers)), ("iteration_mult", int(iteration_mult)),
            ("w_threshold", float(w_threshold)))
    out = MergeStatus(store)
    _queue(store).push("nnls", (nnls_merging, int(species), opts), int(cell), out, rng)
    return out


class MergeStatus:
    """The return value of merge_nnls_based (1 merged, -1 failed).  The call
    stays queued until the value is read (int(), ==, <, ...) or the queue
    flushes for another reason, so unread results batch."""

    def __init__(self, store):
        self.store, self.value = store, None

    def _v(self):
        if self.value is None:
            flush(self.store)
        return self.value

    def __int__(self):
        return self._v()

    __index__ = __int__

    def __eq__(self, other):
        return self._v() == int(other)

    def __ne__(self, other):
        return self._v() != int(other)

    def __lt__(self, other):
        return self._v() < int(other)

    def __gt__(self, other):
        return self._v() > int(other)

    def __hash__(self):
        return hash(self._v())

    def __repr__(self):
        return repr(self._v())
