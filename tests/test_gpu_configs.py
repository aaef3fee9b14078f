"""End-to-end parity of the batched 0-D and variable-weight Couette
workloads at reduced sizes, step by step against the oracle.

Batched uneven 0-D (65,536 cells at full size; here 192): cells of log-uniform occupancy,
pseudo-Maxwell argon, BKW velocities at 273 K, weights uniform in
[0.5, 2] Fnum, ntc! with weight splitting in every cell, then the sort
(which folds the created particles into their cells) and compute_props!.

Variable-weight Couette (test/test_1D_couette_varweight.jl:78-101): weights
lognormal(0, 1) around Fnum, ntc! with splitting, octree N:2 merging of every
cell above the threshold (grid-clamped), squash_pia!, convection, sort,
compute_props_sorted!.  The device runs each phase for all cells at once, so
both sides here use that order (all NTC, then all merges, then one squash);
the reference's driver interleaves NTC and merge per cell, which changes
which free ids the splits reuse, so its trajectories agree statistically, not
bit for bit."""
import ctypes as C
import math

import numpy as np
import pytest

import analytic_cases as ac
import helpers

pytestmark = pytest.mark.gpu


def _interactions(hip, st, its):
    n = len(its)
    st.set_interactions([[type("I", (), dict(mu1=its[i][k].mu1, mu2=its[i][k].mu2, vhs_factor=its[i][k].vhs_factor,
                                             vhs_o=its[i][k].vhs_o))() for k in range(n)] for i in range(n)])


def _batched_0d(hip, mzo, nc, nmax, seed=7, cap_mult=2, centred=False, counts=None):
    """Cells of log-uniform occupancy in [20, nmax) (or the given counts),
    BKW velocities at 273 K, w ~ U(0.5, 2) Fnum; the device store and the
    oracle's state.  centred: every particle at its cell's centre (a merge's
    new positions, mean +- the standard deviation of its bin's positions,
    then stay in the cell)."""
    rng = np.random.default_rng(seed)
    if counts is None:
        counts = np.floor(np.exp(rng.uniform(math.log(20), math.log(nmax), size=nc))).astype(int)
    counts = np.asarray(counts, dtype=int)
    n = int(counts.sum())
    T0, m = 273.0, ac.M_AR
    grid = mzo.grid1d(float(nc), nc)              # one unit-length cell per 0-D cell, V = 1
    pv, pia = mzo.ParticleVector(n), mzo.PIA(nc, 1)
    r = mzo.RNG(seed=7, tag=55)
    pos = 1
    Fnum = 1e23 / counts
    for c in range(1, nc + 1):
        k = int(counts[c - 1])
        for j in range(k):
            pv.pop_to(pos + j)
            w, xr = Fnum[c - 1] * rng.uniform(0.5, 2.0), rng.uniform(0.01, 0.99)
            pv.set(pos + j, w, [0.0, 0.0, 0.0], [c - 0.5 if centred else c - 1 + xr, 0.5, 0.5])
        r.s.stream, r.s.k = c, 0
        mzo.lib().mzo_sample_bkw(r.ref, pv.p, k, pos - 1, m, T0, (C.c_double * 3)(0.0, 0.0, 0.0))
        pia.set(c, 1, k, pos, pos + k - 1, k)
        pos += k
    pia.set_n_total(1, n)
    its = mzo.interactions(["Ar"], table=mzo.PSEUDO_MAXWELL)
    it = its[0][0]
    d = mzo.PSEUDO_MAXWELL[("Ar", "Ar")][0]
    tref = 1.0 / (1e23 * math.pi * d * d) / math.sqrt(2 * ac.KB * T0 / m)
    dt = 0.025 * tref
    cfs = (mzo.CF * nc)()
    sg = np.array([mzo.lib().mzo_estimate_sigma_g_w_max(C.byref(it), m, m, T0, T0, float(f), 1.0) for f in Fnum])
    for c in range(nc):
        cfs[c].sigma_g_w_max = sg[c]
    st = hip.DeviceStore(nc, 1, cap_mult * n + 1024)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    st.set_grid(hip.Grid1DUniform(float(nc), nc))
    _interactions(hip, st, its)
    st.set_sigma_g_w_max(1, 1, sg)
    st.set_seed(7)
    return dict(st=st, pv=pv, pia=pia, cfs=cfs, it=it, grid=grid, dt=dt, m=m, n=n, counts=counts)


def test_batched_0d_cells_with_splitting(hip, mzo):
    b = _batched_0d(hip, mzo, 192, 2000)
    st, pv, pia, cfs, it, grid, dt, m, n = (b[k] for k in ("st", "pv", "pia", "cfs", "it", "grid", "dt", "m", "n"))
    nc = 192
    sp = [hip.Species("Ar", m)]
    V = np.ones(nc)
    ref = mzo.Props(nc, 1)
    mzo.compute_props([pv], pia, [m], ref)
    n_cell0 = ref.n[0].copy()
    n_prev = n
    for step in range(1, 5):
        tag = (step << 4) | 1
        hip.ntc(st, 1, 1, dt, V=V, tag=tag)
        rr = mzo.RNG(seed=7, tag=tag)
        for c in range(1, nc + 1):
            mzo.lib().mzo_ntc_single(rr.ref, C.byref(cfs[c - 1]), C.byref(it), pv.p, pia.p, c, 1, dt, 1.0, 1e-16, 0)
        got = st.download_species(1)
        assert got["n_total"] > n_prev
        n_prev = got["n_total"]
        helpers.assert_state_matches(mzo, got, pv, pia, rtol=1e-11)
        hip.sort_particles(st, 1)
        mzo.lib().mzo_sort_particles(C.byref(grid), pv.p, pia.p, 1)
    got = st.download_species(1)
    assert got["err"] == 0 and got["n_total"] > n
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=1e-11)
    pp = hip.compute_props(st, sp)
    mzo.compute_props([pv], pia, [m], ref)
    assert np.array_equal(pp.np, ref.np)
    assert np.allclose(pp.n, ref.n, rtol=1e-12) and np.allclose(pp.T, ref.T, rtol=1e-10)
    # weight splitting conserves each cell's total weight (up to rounding)
    assert np.allclose(pp.n[0], n_cell0, rtol=1e-12)
    st.close()


def test_batched_0d_many_cells_splitting(hip, mzo):
    """ntc! with splitting over more than 4,096 cells (one lane per cell;
    4,100 cells of 20 to 40 particles and four of 8,000 to 16,000, whose
    candidate loops run hundreds of candidates): three steps of NTC and the
    sort against the oracle, index and PIA bit-exact, values to 1e-11."""
    rng = np.random.default_rng(3)
    counts = list(rng.integers(20, 41, size=4100)) + [8000, 12000, 16000, 9000]
    nc = len(counts)
    b = _batched_0d(hip, mzo, nc, 0, seed=5, counts=counts)
    st, pv, pia, cfs, it, grid, dt = (b[k] for k in ("st", "pv", "pia", "cfs", "it", "grid", "dt"))
    V = np.ones(nc)
    for step in range(1, 4):
        tag = (step << 4) | 1
        hip.ntc(st, 1, 1, dt, V=V, tag=tag)
        rr = mzo.RNG(seed=7, tag=tag)
        for c in range(1, nc + 1):
            mzo.lib().mzo_ntc_single(rr.ref, C.byref(cfs[c - 1]), C.byref(it), pv.p, pia.p, c, 1, dt, 1.0, 1e-16, 0)
        got = st.download_species(1)
        assert got["err"] == 0
        helpers.assert_state_matches(mzo, got, pv, pia, rtol=1e-11)
        hip.sort_particles(st, 1)
        mzo.lib().mzo_sort_particles(C.byref(grid), pv.p, pia.p, 1)
    cf = st.collision_factors(1, 1)
    assert max(cf["n_coll"][-4:]) >= 256
    st.close()


def test_batched_0d_cells_split_sort_merge(hip, mzo):
    """The batched 0-D step with the growth remedy: ntc! with splitting, the
    sort, octree N:2 merging of each cell above its own threshold down to its
    own initial count (merge_octree_N2_based! with a per-cell target_np, the
    cells above 2,048 particles on the device-memory path), squash_pia!, five
    steps against the oracle: index and PIA bit-exact, values to 1e-11."""
    nc = 96
    b = _batched_0d(hip, mzo, nc, 6000, seed=11, centred=True)
    st, pv, pia, cfs, it, grid, dt, m, counts = (b[k] for k in ("st", "pv", "pia", "cfs", "it", "grid", "dt", "m",
                                                                 "counts"))
    # thresholds a little above the start, so merges begin within the five steps
    thresholds = (counts + counts // 40 + 1).astype(np.int32)
    targets = counts.astype(np.int32)
    V = np.ones(nc)
    merged = 0
    for step in range(1, 6):
        tag = (step << 4) | 1
        hip.ntc(st, 1, 1, dt, V=V, tag=tag)
        rr = mzo.RNG(seed=7, tag=tag)
        for c in range(1, nc + 1):
            mzo.lib().mzo_ntc_single(rr.ref, C.byref(cfs[c - 1]), C.byref(it), pv.p, pia.p, c, 1, dt, 1.0, 1e-16, 0)
        hip.sort_particles(st, 1)
        mzo.lib().mzo_sort_particles(C.byref(grid), pv.p, pia.p, 1)
        mtag = (step << 4) | 3
        got_m = hip.merge_octree_targets(st, 1, thresholds, targets, max_Nbins=16384, max_depth=10, squash=True,
                                         tag=mtag)
        rm = mzo.RNG(seed=7, tag=mtag)
        exp_m = 0
        for c in range(1, nc + 1):
            if pia.ind(c).n_local > thresholds[c - 1]:
                oc = mzo.Octree(max_Nbins=16384, max_depth=10)
                mzo.lib().mzo_merge_octree_N2_based(rm.ref, oc.p, pv.p, pia.p, c, 1, int(targets[c - 1]), None)
                exp_m += 1
        mzo.lib().mzo_squash_pia(pv.p, pia.p, 1)
        assert got_m == exp_m
        merged += got_m
        got = st.download_species(1)
        assert got["err"] == 0
        helpers.assert_state_matches(mzo, got, pv, pia, rtol=1e-11)
    assert merged > nc // 2
    assert any(k > 2048 for k in counts)
    st.close()


def test_varweight_couette_with_octree_merges(hip, mzo):
    nc, ppc, thr, target = 20, 100, 102, 64
    case = helpers.couette_case(mzo, n_cells=nc, ppc=ppc, seed=13, dt=2e-10, sgwm_mult=40.0,
                                weights=lambda rng, k: rng.lognormal(0.0, 1.0, size=k) / math.exp(0.5))
    st = helpers.store_from_case(hip, case, capacity=4 * nc * ppc + 4096)
    pv, pia, g = case["pv"], case["pia"], case["grid"]
    sp = case["species"]
    merged = 0
    for step in range(1, 11):
        # all NTC, then all merges, then one squash (the device's phase order)
        hip.ntc(st, 1, 1, case["dt"], tag=(step << 4) | 1)
        r = mzo.RNG(seed=case["seed"], tag=(step << 4) | 1)
        for c in range(1, nc + 1):
            mzo.lib().mzo_ntc_single(r.ref, C.byref(case["cfs"][c - 1]), C.byref(case["it"]), pv.p, pia.p, c, 1,
                                     case["dt"], g.dx, 1e-16, 0)
        mtag = (step << 4) | 3
        merged += hip.merge_octree_N2_based(st, 1, thr, target, grid_clamp=True, squash=True, tag=mtag)
        r = mzo.RNG(seed=case["seed"], tag=mtag)
        oc = mzo.Octree(split=1, init_bin_bounds=1, bin_bounds_compute=1, max_Nbins=4096, max_depth=10)
        for c in range(1, nc + 1):
            if pia.ind(c).n_local > thr:
                mzo.lib().mzo_merge_octree_N2_based(r.ref, oc.p, pv.p, pia.p, c, 1, target, C.byref(g))
        mzo.lib().mzo_squash_pia(pv.p, pia.p, 1)
        hip.convect_particles(st, 1, case["dt"], tag=(step << 4) | 2)
        r = mzo.RNG(seed=case["seed"], tag=(step << 4) | 2)
        mzo.lib().mzo_convect_particles(r.ref, C.byref(g), C.byref(case["walls"]), pv.p, pia.p, 1, case["dt"])
        hip.sort_particles(st, 1)
        mzo.lib().mzo_sort_particles(C.byref(g), pv.p, pia.p, 1)
        got = st.download_species(1)
        assert got["err"] == 0
        helpers.assert_state_matches(mzo, got, pv, pia, rtol=1e-11)
    assert merged > 0
    pp = hip.compute_props_sorted(st, sp)
    assert np.allclose(pp.T[0], helpers.oracle_props_sorted(mzo, case), rtol=1e-10)
    st.close()


def _batch_case(hip, mzo, nc, nmax, seed):
    rng = np.random.default_rng(seed)
    counts = np.floor(np.exp(rng.uniform(math.log(20), math.log(nmax), size=nc))).astype(int)
    n = int(counts.sum())
    T0, m = 273.0, ac.M_AR
    pv, pia = mzo.ParticleVector(n), mzo.PIA(nc, 1)
    r = mzo.RNG(seed=seed, tag=55)
    pos = 1
    Fnum = 1e23 / counts
    for c in range(1, nc + 1):
        k = int(counts[c - 1])
        for j in range(k):
            pv.pop_to(pos + j)
            pv.set(pos + j, Fnum[c - 1] * rng.uniform(0.5, 2.0), [0.0, 0.0, 0.0], [c - 0.5, 0.5, 0.5])
        r.s.stream, r.s.k = c, 0
        mzo.lib().mzo_sample_bkw(r.ref, pv.p, k, pos - 1, m, T0, (C.c_double * 3)(0.0, 0.0, 0.0))
        pia.set(c, 1, k, pos, pos + k - 1, k)
        pos += k
    pia.set_n_total(1, n)
    its = mzo.interactions(["Ar"], table=mzo.PSEUDO_MAXWELL)
    d = mzo.PSEUDO_MAXWELL[("Ar", "Ar")][0]
    tref = 1.0 / (1e23 * math.pi * d * d) / math.sqrt(2 * ac.KB * T0 / m)
    sg = np.array([mzo.lib().mzo_estimate_sigma_g_w_max(C.byref(its[0][0]), m, m, T0, T0, float(f), 1.0)
                   for f in Fnum])

    def store():
        st = hip.DeviceStore(nc, 1, n + 256)
        helpers.upload_from_oracle(st, mzo, pv, pia)
        st.set_grid(hip.Grid1DUniform(float(nc), nc))
        _interactions(hip, st, its)
        st.set_sigma_g_w_max(1, 1, sg)
        st.set_seed(seed)
        st.set_species_data([hip.Species("Ar", m)])
        return st
    return dict(pv=pv, pia=pia, it=its[0][0], sg=sg, dt=0.025 * tref, n=n, nc=nc, store=store, m=m, seed=seed)


def test_batched_0d_per_cell_launches_equal_plain(hip, mzo):
    """One NTC launch per cell per step (the reference's ntc! loop as drop-in
    calls) leaves the same particles, properties and collision factors as the
    NTC kernel over all cells."""
    case = _batch_case(hip, mzo, nc=40, nmax=2000, seed=23)
    a, b = case["store"](), case["store"]()
    hip.batch0d_steps(a, 3, 1, case["dt"], mode="plain")
    hip.batch0d_steps(b, 3, 1, case["dt"], mode="per-cell")
    da, db = a.download_species(1), b.download_species(1)
    n = case["n"]
    for k in ("v", "w", "index"):
        assert np.array_equal(da[k][:n], db[k][:n]), k
    pa, pb = hip.PhysProps(case["nc"], 1).fetch(a), hip.PhysProps(case["nc"], 1).fetch(b)
    for k in ("np", "n", "v", "T"):
        assert np.array_equal(getattr(pa, k), getattr(pb, k)), k
    ca, cb = a.collision_factors(1, 1), b.collision_factors(1, 1)
    for k in ca:
        assert np.array_equal(ca[k], cb[k]), k
    a.close()
    b.close()


@pytest.mark.parametrize("shape", [(0, 0), (1, 1), (8, 512), (64, 100000)])
def test_batched_0d_persistent_equals_per_step(hip, mzo, shape):
    """The persistent batched launch and the per-step kernels leave identical
    particles, properties and collision factors; both equal the oracle's
    ntc_equal_weight! + compute_props! to 1e-11."""
    case = _batch_case(hip, mzo, nc=300, nmax=3000, seed=21)
    a, b = case["store"](), case["store"]()
    steps = 6
    hip.batch0d_steps(a, steps, 1, case["dt"], mode="plain")
    b.set_batch_shape(*shape)
    hip.batch0d_steps(b, steps, 1, case["dt"], mode="batched")
    info = b.batch_info()
    assert info[0] >= 1
    da, db = a.download_species(1), b.download_species(1)
    assert da["err"] == 0 and db["err"] == 0
    n = case["n"]
    for k in ("x", "v", "w", "index"):
        assert np.array_equal(da[k][:n], db[k][:n]), k
    pa, pb = hip.PhysProps(case["nc"], 1).fetch(a), hip.PhysProps(case["nc"], 1).fetch(b)
    for k in ("np", "n", "v", "T"):
        assert np.array_equal(getattr(pa, k), getattr(pb, k)), k
    ca, cb = a.collision_factors(1, 1), b.collision_factors(1, 1)
    for k in ca:
        assert np.array_equal(ca[k], cb[k]), k
    # the oracle
    pv, pia = case["pv"], case["pia"]
    cfs = (mzo.CF * case["nc"])()
    for c in range(case["nc"]):
        cfs[c].sigma_g_w_max = case["sg"][c]
    for step in range(1, steps + 1):
        rr = mzo.RNG(seed=case["seed"], tag=(step << 4) | 1)
        for c in range(1, case["nc"] + 1):
            mzo.lib().mzo_ntc_single(rr.ref, C.byref(cfs[c - 1]), C.byref(case["it"]), pv.p, pia.p, c, 1,
                                     case["dt"], 1.0, 1e-16, 1)
    helpers.assert_state_matches(mzo, db, pv, pia, rtol=1e-11)
    ref = mzo.Props(case["nc"], 1)
    mzo.compute_props([pv], pia, [case["m"]], ref)
    assert np.allclose(pb.T, ref.T, rtol=1e-10) and np.array_equal(pb.np, ref.np)
    assert sum(cb["n_coll_performed"]) > 0
    a.close()
    b.close()
