"""Benchmark: the 1-D Couette flow step on one MI355X.

Workload (the reference's Couette case, test/test_1D_couette.jl:66-97, at the
scale of the headline configuration): argon VHS (T_ref 273 K, omega 0.81),
500 cells over L = 5e-4 m, diffuse walls at 300 K moving at -/+500 m/s, 200
equal-weight particles per cell (100,000 particles) sampled Maxwellian at
300 K, number density 1e23 m^-3, FP64, dt = 2e-10 s (CFL 0.17).  One step is
ntc! in every cell, convection, the bit-exact counting sort, and
compute_props_sorted!.  Inputs are synthetic (sampled here with numpy) and
resident on the GPU before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode slab|graph|eager] [--no-cpu-baseline]

Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))

K_B = 1.380649e-23
N_CELLS, PPC, L = 500, 200, 5e-4
T_WALL, V_WALL, NDENS, DT = 300.0, 500.0, 1e23, 2.0e-10

# algorithmic bytes per particle (reference method) of each kernel family:
# convection reads x, v_x and writes x (+ the cell key) 28 B; the sort reads
# the key twice and reads/writes the index 16 B; props read w, v and the index
# 36 B; NTC touches < 1 B per particle at this collision rate.
ALG_BYTES = {"convect": 28.0, "sort": 16.0, "props": 36.0, "ntc": 1.0, "step": 80.0}
# HBM bytes the slab kernel moves per particle-step, from rocprofv3 FETCH_SIZE
# (x 2.0) and WRITE_SIZE (x 1.0) counters of a 2000-step launch on this
# workload, the factors calibrated on a 256 MiB device copy
# (profiles/couette_100k_slab_step.md)
SLAB_HBM_BYTES_PER_PARTICLE_STEP = 16.9


def sample_case(seed=12345, n_cells=N_CELLS, ppc=PPC, length=L, ndens=NDENS):
    """Particles in the reference layout: cell c holds logical positions
    [(c-1) ppc + 1, c ppc], ids 1..N, positions uniform in the cell,
    Maxwellian velocities (src/distributions_and_sampling.jl:432-509)."""
    rng = np.random.default_rng(seed)
    n = n_cells * ppc
    dx = length / n_cells
    m = 66.3e-27
    vs = np.sqrt(2 * K_B * T_WALL / m)
    cells = np.repeat(np.arange(n_cells), ppc)
    x = np.empty((n, 3))
    x[:, 0] = (cells + rng.uniform(size=n)) * dx
    x[:, 1:] = rng.uniform(size=(n, 2))
    vn = np.sqrt(-np.log(1.0 - rng.uniform(size=n)))
    vr = np.sqrt(-np.log(1.0 - rng.uniform(size=n)))
    t1 = 2 * np.pi * rng.uniform(size=n)
    t2 = 2 * np.pi * rng.uniform(size=n)
    v = vs * np.stack([vn * np.cos(t1), vr * np.cos(t2), vr * np.sin(t2)], axis=1)
    Fnum = ndens * dx / ppc
    w = np.full(n, Fnum)
    pia = np.zeros((n_cells, 7), dtype=np.int64)
    c1 = np.arange(n_cells)
    pia[:, 0] = ppc
    pia[:, 1] = c1 * ppc + 1
    pia[:, 2] = (c1 + 1) * ppc
    pia[:, 3] = ppc
    pia[:, 4] = 0
    pia[:, 5] = -1
    return dict(x=x, v=v, w=w, ids=np.arange(1, n + 1), pia=pia, n=n, Fnum=Fnum, mass=m, n_cells=n_cells, L=length)


def load_pkg():
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    return ge.load_package()


def setup_store(pkg, case, capacity=None):
    data = os.path.join(ROOT, "merzbild.jl_amd", "data")
    sp = pkg.load_species_data(os.path.join(data, "particles.toml"), ["Ar"])
    it = pkg.load_interaction_data(os.path.join(data, "vhs.toml"), sp)
    st = pkg.DeviceStore(case["n_cells"], 1, capacity or case["n"] + 4096)
    st.upload_species(1, case["x"], case["v"], case["w"], case["ids"], case["pia"], case["n"], True,
                      np.zeros(0, dtype=np.int64), case["n"])
    grid = pkg.Grid1DUniform(case["L"], case["n_cells"])
    st.set_grid(grid)
    st.set_walls(pkg.MaxwellWalls1D(sp, T_WALL, T_WALL, -V_WALL, V_WALL, 1.0, 1.0))
    st.set_interactions(it)
    st.set_species_data(sp)
    # create_collision_factors_array(pia, interactions, species_data, T, Fnum)
    st.set_sigma_g_w_max(1, 1, pkg.estimate_sigma_g_w_max(it[0][0], sp[0].mass, sp[0].mass, T_WALL, T_WALL,
                                                           case["Fnum"]))
    st.set_seed(20000)
    return st, sp, it


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, it, target_s=12.0, serial_s=6.0):
    """The C oracle (a restatement of the reference method, not the Julia
    reference), -O3 with OpenMP over cells and a threaded counting sort, on
    this host's cores, same workload, bounded to a sample of steps; also a
    single-thread sample.  Returns (threaded particle-timesteps/s, threads,
    steps, serial particle-timesteps/s, serial steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mzo
    mzo.build()
    n = case["n"]
    pv = mzo.ParticleVector(n)
    pia = mzo.PIA(N_CELLS, 1)
    for i in range(1, n + 1):
        pv.pop_to(i)
    pv.x[:n] = case["x"]
    pv.v[:n] = case["v"]
    pv.w[:n] = case["w"]
    for c in range(N_CELLS):
        pia.set(c + 1, 1, *case["pia"][c])
    pia.set_n_total(1, n)
    grid = mzo.grid1d(L, N_CELLS)
    walls = mzo.walls([case["mass"]], T_WALL, T_WALL, -V_WALL, V_WALL, 1.0, 1.0)
    oit = mzo.interactions(["Ar"])[0][0]
    cfs = (mzo.CF * N_CELLS)()
    sg = mzo.lib().mzo_estimate_sigma_g_w_max(C.byref(oit), case["mass"], case["mass"], T_WALL, T_WALL,
                                              case["Fnum"], 1.0)
    for c in range(N_CELLS):
        cfs[c].sigma_g_w_max = sg
    V = np.full(N_CELLS, L / N_CELLS)
    pp = mzo.Props(N_CELLS, 1)
    rng = mzo.RNG(seed=20000)
    threads = mzo.lib().mzo_omp_max_threads()
    args = lambda k, s0, nt: (rng.ref, C.byref(grid), C.byref(walls), C.byref(oit), cfs, pv.p, pia.p, DT,
                              V.ctypes.data_as(C.POINTER(C.c_double)), case["mass"], pp.ref, k, s0, 1, nt, 1)
    step = 1

    def timed(nt, budget):
        nonlocal step
        t0 = time.perf_counter()
        mzo.lib().mzo_couette_steps(*args(5, step, nt))
        step += 5
        k = max(10, int(budget / max((time.perf_counter() - t0) / 5, 1e-6)))
        t0 = time.perf_counter()
        mzo.lib().mzo_couette_steps(*args(k, step, nt))
        el = time.perf_counter() - t0
        step += k
        return n * k / el, k

    serial, serial_steps = timed(1, serial_s)
    threaded, steps = timed(threads, target_s)
    return threaded, threads, steps, serial, serial_steps


# Batched uneven 0-D cells: 65,536 independent cells of V = 1 with log-uniform
# occupancy in [20, 20000] (Philox seed 7), argon pseudo-Maxwell molecules,
# BKW velocities at 273 K (|v| ~ Chi(5) sqrt(0.3 2kT/m), polar angle uniform
# on [0, pi] as the reference's sample_bkw! draws it), weights uniform in
# [0.5, 2] Fnum with Fnum = 1e23 / N_cell, dt = 0.025 t_ref.  A step is
# ntc_equal_weight! in every cell (no splitting, so the particle count stays
# fixed) and compute_props!.
B0_CELLS, B0_NMIN, B0_NMAX, B0_T0, B0_NDENS = 65536, 20, 20000, 273.0, 1e23


def sample_batched_0d(seed=7):
    g = np.random.Generator(np.random.Philox(seed))
    counts = np.floor(np.exp(g.uniform(np.log(B0_NMIN), np.log(B0_NMAX), size=B0_CELLS))).astype(np.int64)
    n = int(counts.sum())
    m = 66.3e-27
    vscale = np.sqrt(2 * K_B * B0_T0 / m) * np.sqrt(0.3)
    v = np.empty((n, 3))
    chunk = 1 << 24
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        k = b - a
        vabs = np.sqrt(g.chisquare(5, size=k))
        th = g.uniform(size=k) * np.pi
        ph = g.uniform(size=k) * 2 * np.pi
        st = np.sin(th)
        v[a:b, 0] = vscale * vabs * st * np.cos(ph)
        v[a:b, 1] = vscale * vabs * st * np.sin(ph)
        v[a:b, 2] = vscale * vabs * np.cos(th)
    fnum = B0_NDENS / counts
    w = np.repeat(fnum, counts) * g.uniform(0.5, 2.0, size=n)
    x = np.zeros((n, 3))
    x[:, 0] = np.repeat(np.arange(B0_CELLS, dtype=np.float64), counts) + 0.5
    starts = np.concatenate(([0], np.cumsum(counts)[:-1]))
    pia = np.zeros((B0_CELLS, 7), dtype=np.int64)
    pia[:, 0] = counts
    pia[:, 1] = starts + 1
    pia[:, 2] = starts + counts
    pia[:, 3] = counts
    pia[:, 5] = -1
    return dict(x=x, v=v, w=w, ids=np.arange(1, n + 1, dtype=np.int64), pia=pia, n=n, counts=counts, fnum=fnum,
                mass=m)


def run_batched_0d(a, pkg):
    case = sample_batched_0d()
    n = case["n"]
    data = os.path.join(ROOT, "merzbild.jl_amd", "data")
    sp = pkg.load_species_data(os.path.join(data, "particles.toml"), ["Ar"])
    it = pkg.load_interaction_data(os.path.join(data, "pseudo_maxwell.toml"), sp)
    st = pkg.DeviceStore(B0_CELLS, 1, n + 1024)
    st.upload_species(1, case["x"], case["v"], case["w"], case["ids"], case["pia"], n, True,
                      np.zeros(0, dtype=np.int64), n)
    st.set_grid(pkg.Grid1DUniform(float(B0_CELLS), B0_CELLS))     # V = 1 per cell
    st.set_interactions(it)
    st.set_species_data(sp)
    ia = it[0][0]
    sg = np.array([pkg.estimate_sigma_g_w_max(ia, sp[0].mass, sp[0].mass, B0_T0, B0_T0, f) for f in case["fnum"]])
    st.set_sigma_g_w_max(1, 1, sg)
    st.set_seed(7)
    sigma_ref = np.pi * ia.vhs_d ** 2
    tref = 1.0 / (B0_NDENS * sigma_ref) / np.sqrt(2 * K_B * B0_T0 / sp[0].mass)
    dt = 0.025 * tref
    del case["x"], case["v"], case["w"]
    mode = a.mode if a.mode in ("batched", "plain") else "batched"
    steps = a.steps if a.steps != 20000 else 200

    def run(k, s0):
        return pkg.batch0d_steps(st, k, s0, dt, sp, mode=mode)

    tune = None
    s0 = 1000
    if mode == "batched" and not a.no_autotune:
        best, times, s0 = pkg.batch0d_autotune(st, 1, dt, sp)
        tune = {"best": list(best), "ms_per_step": {f"{k[0]}x{k[1]}": round(v, 4) for k, v in times.items()}}
    else:
        run(a.warmup if a.warmup != 200 else 5, 1)
    st.sync()
    st.timing_reset()
    st.timing(True)
    t0 = time.perf_counter()
    run(steps, s0)
    st.sync()
    el = time.perf_counter() - t0
    st.timing(False)
    sc = st.scalars(1)
    assert sc["err"] == 0 and sc["n_total"] == n, sc
    # the same cells under the other launch patterns, continuing the run
    # (a few steps each; the per-step cost barely changes as the state relaxes)
    others = {}
    if mode == "batched" and not a.no_compare:
        s1 = s0 + steps
        for other, k in (("plain", 20), ("per-cell", 2)):
            st.sync()
            t1 = time.perf_counter()
            pkg.batch0d_steps(st, k, s1, dt, sp, mode=other)
            st.sync()
            others[other] = {"steps": k, "ms_per_step": (time.perf_counter() - t1) * 1e3 / k}
            s1 += k
        for other in others:
            others[other]["batched_speedup"] = others[other]["ms_per_step"] / (el * 1e3 / steps)
    out = {"metric": "particle-timesteps/s (NTC+props), batched uneven 0-D cells", "value": n * steps / el,
           "unit": "particle-timesteps/s", "n_gpus": 1, "steps": steps, "ms_per_step": el * 1e3 / steps,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "65536 0-D cells, log-uniform 20..20000 particles (Philox seed 7), "
                                  "pseudo-Maxwell Ar, BKW at 273 K, w ~ U(0.5, 2) Fnum; ntc_equal_weight! + "
                                  "compute_props! per step", "n_particles": n, "mode": mode,
                      "batch_shape": st.batch_info() if mode == "batched" else None, "autotune": tune},
           "compare": others}
    print(json.dumps(out), flush=True)
    st.close()


def run_batched_0d_split(a, pkg):
    """The batched uneven 0-D cells with weight splitting and the survey's
    growth remedy: per step ntc! with splitting in every cell (its created
    particles appended, as collision_ntc.jl:223-309 does), the sort that folds
    them into their cells (without it a cell's second group would run into
    the next cell's appended particles), octree N:2 merging of every cell
    above 1.5x its initial count back down to that count
    (merge_octree_N2_based! with the cell's own target_np, then squash_pia!),
    and compute_props!.  The rate counts the particles each step holds; the
    merge time is reported on its own and left out of `value`."""
    case = sample_batched_0d()
    n = case["n"]
    counts = case["pia"][:, 0].astype(np.int64)
    data = os.path.join(ROOT, "merzbild.jl_amd", "data")
    sp = pkg.load_species_data(os.path.join(data, "particles.toml"), ["Ar"])
    it = pkg.load_interaction_data(os.path.join(data, "pseudo_maxwell.toml"), sp)
    steps = a.steps if a.steps != 20000 else 60
    # merging caps every cell at 1.5x its start plus one step's growth
    st = pkg.DeviceStore(B0_CELLS, 1, int(n * 1.6) + 4096)
    st.upload_species(1, case["x"], case["v"], case["w"], case["ids"], case["pia"], n, True,
                      np.zeros(0, dtype=np.int64), n)
    del case["x"], case["v"], case["w"]
    st.set_grid(pkg.Grid1DUniform(float(B0_CELLS), B0_CELLS))
    st.set_interactions(it)
    st.set_species_data(sp)
    ia = it[0][0]
    sg = np.array([pkg.estimate_sigma_g_w_max(ia, sp[0].mass, sp[0].mass, B0_T0, B0_T0, f) for f in case["fnum"]])
    st.set_sigma_g_w_max(1, 1, sg)
    st.set_seed(7)
    tref = 1.0 / (B0_NDENS * np.pi * ia.vhs_d ** 2) / np.sqrt(2 * K_B * B0_T0 / sp[0].mass)
    dt = 0.025 * tref
    masses = np.array([sp[0].mass])
    lib = pkg.load_library()
    thresholds = (3 * counts // 2).astype(np.int32)
    targets = counts.astype(np.int32)
    merged = [0]

    def step(k):
        pkg.ntc(st, 1, 1, dt, V=1.0, equal_weight=False, tag=(k << 4) | 1)
        pkg.sort_particles(st, 1)
        merged[0] += pkg.merge_octree_targets(st, 1, thresholds, targets, max_Nbins=1 << 20, max_depth=10,
                                              squash=True, tag=(k << 4) | 3)
        pkg._check(lib.mz_compute_props(st.h, 0, 0, pkg._d(masses)))

    step(1)   # warm-up (sizes the creation and sort scratch)
    st.sync()
    merged[0] = 0
    counts_held = []
    st.timing_reset()
    st.timing(True)
    t0 = time.perf_counter()
    for k in range(2, 2 + steps):
        counts_held.append(st.scalars(1)["n_total"])
        step(k)
    st.sync()
    el = time.perf_counter() - t0
    ms = {name: st.timing_get(name) for name in ("ntc", "sort", "merge_octree", "props")}
    st.timing(False)
    sc = st.scalars(1)
    assert sc["err"] == 0, sc
    merge_s = ms["merge_octree"][0] * ms["merge_octree"][1] * 1e-3
    held = float(np.sum(counts_held))
    out = {"metric": "particle-timesteps/s (NTC with splitting + sort + props; octree merge timed apart), "
                     "batched uneven 0-D cells",
           "value": held / (el - merge_s), "unit": "particle-timesteps/s", "n_gpus": 1, "steps": steps,
           "ms_per_step": (el - merge_s) * 1e3 / steps, "higher_is_better": True, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "65536 0-D cells, log-uniform 20..20000 particles (Philox seed 7), "
                                  "pseudo-Maxwell Ar, BKW at 273 K, w ~ U(0.5, 2) Fnum; per step ntc! with weight "
                                  "splitting, sort, octree N:2 merge of each cell above 1.5x its initial count "
                                  "down to that count (mid split, min/max bounds, max depth 10) + squash, "
                                  "compute_props!",
                      "n_particles_start": n, "n_particles_end": sc["n_total"],
                      "n_particles_max": int(max(counts_held)), "cells_merged": merged[0],
                      "value_including_merge": held / el, "ms_per_step_including_merge": el * 1e3 / steps,
                      "kernel_ms_per_call": {k: v[0] for k, v in ms.items()},
                      "calls": {k: v[1] for k, v in ms.items()}}}
    print(json.dumps(out), flush=True)
    st.close()


def run_couette_varweight(a, pkg):
    """The variable-weight Couette flow with octree merging
    (test/test_1D_couette_varweight.jl:78-105 at the headline size): 500 cells
    x 200 particles, weights Fnum lognormal(0, 1) / e^0.5, per step ntc! with
    splitting, octree N:2 merging (midpoint split, min/max bounds, max depth
    8, grid-clamped) of every cell above 400 particles down to 64, squash,
    convection, sort and compute_props_sorted!.  The rate counts the
    particles each step holds; merge time is reported on its own."""
    case = sample_case()
    n = case["n"]
    g = np.random.default_rng(99)
    case["w"] = case["Fnum"] * g.lognormal(0.0, 1.0, size=n) / np.exp(0.5)
    # splitting grows a cell up to the merge threshold (400 of 200 at the start)
    st, sp, it = setup_store(pkg, case, capacity=3 * n + 4096)
    steps = a.steps if a.steps != 20000 else 4000
    warm = a.warmup if a.warmup != 200 else 200
    masses = np.array([sp[0].mass])
    lib = pkg.load_library()
    merged = [0]

    def step(k):
        pkg.ntc(st, 1, 1, DT, V=L / N_CELLS, equal_weight=False, tag=(k << 4) | 1)
        merged[0] += pkg.merge_octree_N2_based(st, 1, threshold=400, target_np=64, max_depth=8, grid_clamp=True,
                                               squash=True, tag=(k << 4) | 3)
        pkg.convect_particles(st, 1, DT, tag=(k << 4) | 2, fetch=False)
        pkg.sort_particles(st, 1)
        pkg._check(lib.mz_compute_props(st.h, 1, 0, pkg._d(masses)))

    for k in range(1, warm + 1):
        step(k)
    st.sync()
    merged[0] = 0
    n0 = st.scalars(1)["n_total"]
    st.timing_reset()
    st.timing(True)
    t0 = time.perf_counter()
    for k in range(warm + 1, warm + 1 + steps):
        step(k)
    st.sync()
    el = time.perf_counter() - t0
    fam = {f: st.timing_get(f)[0] for f in ("ntc", "merge_octree", "convect", "sort", "props")}
    st.timing(False)
    sc = st.scalars(1)
    assert sc["err"] == 0, sc
    n_mean = 0.5 * (n0 + sc["n_total"])
    merge_ms = fam["merge_octree"] * steps
    out = {"metric": "particle-timesteps/s (ntc! with splitting + octree merge + convection + sort + props), "
                     "variable-weight Couette",
           "value": n_mean * steps / el, "unit": "particle-timesteps/s", "n_gpus": 1, "steps": steps,
           "warmup": warm, "ms_per_step": el * 1e3 / steps, "higher_is_better": True, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "1D Couette, argon VHS, 500 cells x 200 particles, lognormal(0,1)/e^0.5 weights, "
                                  "dt 2e-10 s; octree N:2 merge above 400 to 64 (mid split, max depth 8, grid clamp)",
                      "n_particles_start": n0, "n_particles_end": sc["n_total"], "cells_merged": merged[0],
                      "kernel_ms_per_step": fam,
                      "value_excluding_merge": n_mean * steps / max(el - merge_ms * 1e-3, 1e-9)}}
    print(json.dumps(out), flush=True)
    st.close()


# The averaging phase of the reference's large Couette run with wall
# properties (simulations/1D/couette_with_surface_quantities.jl:130-131):
# argon, L = 5e-4 m in 2,000 cells, 250 particles per cell (500,000),
# number density 5e22 m^-3, walls at 300 K moving at -/+500 m/s, dt =
# 2.59e-9 s.  A step: ntc!, convection with wall properties, their time
# average, sort, compute_props_sorted!, the props' time average.  Particles
# cross up to ~15 cells per step, so the slab kernel (moves within one slab)
# does not apply and the step graph runs it.
CS_CELLS, CS_PPC, CS_NDENS, CS_DT = 2000, 250, 5e22, 2.59e-9


def run_couette_surf(a, pkg):
    """The averaging phase at the reference run's size, or (--surf-headline)
    at the headline Couette configuration, where the slab kernel runs it."""
    if a.surf_headline:
        n_cells, ppc, ndens, dt = N_CELLS, PPC, NDENS, DT
    else:
        n_cells, ppc, ndens, dt = CS_CELLS, CS_PPC, CS_NDENS, CS_DT
    case = sample_case(seed=1234, n_cells=n_cells, ppc=ppc, length=L, ndens=ndens)
    st, sp, it = setup_store(pkg, case)
    n = case["n"]
    steps = a.steps if a.steps != 20000 else 2000
    warm = a.warmup if a.warmup != 200 else 50
    mode = a.mode if a.mode in ("slab", "graph", "eager") else "graph"
    n_avg = warm + steps
    s0, tune = 1, None
    if mode == "slab" and not a.no_autotune:
        # the slab shape, picked on plain steps (which serve as warmup)
        best, times, s0 = pkg.couette_autotune(st, 1, dt, steps=20)
        tune = {"best": list(best), "ms_per_step": {f"{k[0]}x{k[1]}": round(v, 5) for k, v in times.items()}}
    st.surf_clear(1)
    st.props_avg_clear()
    pkg.couette_avg_steps(st, warm, s0, dt, n_avg, species_data=sp, mode=mode)
    st.sync()
    st.timing_reset()
    st.timing(True)
    t0 = time.perf_counter()
    path = pkg.couette_avg_steps(st, steps, s0 + warm, dt, n_avg, species_data=sp, mode=mode)
    st.sync()
    el = time.perf_counter() - t0
    fam = "slab" if path == "slab" else "step"
    launch_ms, _ = st.timing_get(fam)
    st.timing(False)
    sc = st.scalars(1)
    assert sc["err"] == 0 and sc["n_total"] == n, sc
    # the convection-with-wall-properties kernel on its own (HIP events)
    st.timing_reset()
    st.timing(True)
    for t in range(100):
        pkg.convect_particles(st, 1, dt, tag=((s0 + warm + steps + t) << 4) | 2, species_data=sp,
                              surf_props=pkg.SurfProps.grid1d(1), fetch=False)
    conv_ms = st.timing_get("convect")[0]
    st.timing(False)
    surf = pkg.SurfProps.grid1d(1).fetch(st, 1)
    pp = pkg.PhysProps(n_cells, 1).fetch_avg(st)
    conv_gbs = ALG_BYTES["convect"] * n / (conv_ms * 1e-3) / 1e9
    out = {"metric": "particle-timesteps/s (averaging phase with wall properties)", "value": n * steps / el,
           "unit": "particle-timesteps/s", "n_gpus": 1, "steps": steps, "warmup": warm,
           "ms_per_step": el * 1e3 / steps, "higher_is_better": True, "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": f"1D Couette with wall properties, argon VHS, {n_cells} cells x {ppc} "
                                  f"particles ({n}), ndens {ndens:g}, dt {dt:g} s, FP64; per step ntc!, "
                                  "convection + wall properties, avg_props! (wall), counting sort, "
                                  "compute_props_sorted!, avg_props! (cells)", "n_particles": n, "mode": path,
                      "launch_ms_per_step": launch_ms, "slab_shape": st.slab_info() if path == "slab" else None,
                      "autotune": tune},
           "convect_surf": {"ms": conv_ms, "alg_bytes_per_particle": ALG_BYTES["convect"],
                            "achieved_GBs": conv_gbs, "frac_of_8000": conv_gbs / 8000.0},
           "wall_averages": {"normal_pressure_Pa": surf.normal_pressure[:, 0].tolist(),
                             "shear_y_Pa": surf.force[1, :, 0].tolist(),
                             "contacts_per_step": surf.np[:, 0].tolist(),
                             "T_mid_K": float(pp.T[0, n_cells // 2])}}
    print(json.dumps(out), flush=True)
    st.close()


# Batched NNLS merging: 32,768 cells of U{200..3000} particles (Philox seed
# 5) from bi-Maxwellians (T_perp = 300 K, T_par / T_perp ~ U[0.5, 2] along a
# random axis, drift uniform in the ball |u| <= 500 m/s), weights
# lognormal(0, 0.5); every cell merged by merge_nnls_based! matching the base
# moments and all multi-index moments to total order 4, :variance scaling.
NN_CELLS, NN_NMIN, NN_NMAX = 32768, 200, 3000


def _mim(order):
    return [(a, b, i - a - b) for i in range(1, order + 1) for a in range(i + 1) for b in range(i - a + 1)]


def _moments(w, v, mim):
    return np.array([np.sum(w * v[:, 0] ** a * v[:, 1] ** b * v[:, 2] ** c) for (a, b, c) in mim])


def _abs_moments(w, v, mim):
    return np.array([np.sum(w * np.abs(v[:, 0] ** a * v[:, 1] ** b * v[:, 2] ** c)) for (a, b, c) in mim])


def sample_nnls_cells():
    """The batched NNLS workload on the host: counts, positions, velocities,
    weights of all cells (Philox seed 5)."""
    g = np.random.Generator(np.random.Philox(5))
    counts = g.integers(NN_NMIN, NN_NMAX + 1, size=NN_CELLS)
    n = int(counts.sum())
    m = 66.3e-27
    s_perp = np.sqrt(K_B * 300.0 / m)
    v = g.normal(size=(n, 3))
    cell = np.repeat(np.arange(NN_CELLS), counts)
    ratio = g.uniform(0.5, 2.0, size=NN_CELLS)
    axis = g.normal(size=(NN_CELLS, 3))
    axis /= np.linalg.norm(axis, axis=1)[:, None]
    drift = g.normal(size=(NN_CELLS, 3))
    drift *= (500.0 * g.uniform(size=NN_CELLS) ** (1.0 / 3.0) / np.linalg.norm(drift, axis=1))[:, None]
    # stretch the component along the cell's axis by sqrt(T_par / T_perp)
    ax = axis[cell]
    par = np.sum(v * ax, axis=1)
    v = s_perp * (v + (np.sqrt(ratio[cell]) - 1.0)[:, None] * par[:, None] * ax) + drift[cell]
    w = g.lognormal(0.0, 0.5, size=n)
    x = g.uniform(size=(n, 3))
    starts = np.concatenate(([0], np.cumsum(counts)[:-1]))
    return dict(counts=counts, starts=starts, n=n, x=x, v=v, w=w)


def _nnls_warm(pkg, cs, max_pcols, method, n_cells=64):
    """One untimed merge of the first cells with the same method, so that the
    timed call does not include loading the kernels' code."""
    counts = cs["counts"][:n_cells]
    n = int(counts.sum())
    starts = np.concatenate(([0], np.cumsum(counts)[:-1]))
    pia = np.zeros((n_cells, 7), dtype=np.int64)
    pia[:, 0] = counts
    pia[:, 1] = starts + 1
    pia[:, 2] = starts + counts
    pia[:, 3] = counts
    pia[:, 5] = -1
    st = pkg.DeviceStore(n_cells, 1, n + 1024)
    st.upload_species(1, cs["x"][:n], cs["v"][:n], cs["w"][:n], np.arange(1, n + 1, dtype=np.int64), pia, n, True,
                      np.zeros(0, dtype=np.int64), n)
    if max_pcols is None:
        pkg.merge_nnls_based(st, 1, 1, _mim(4), scaling="variance", squash=False)
    else:
        pkg.merge_nnls_candidates(st, 1, 1, _mim(4), max_pcols, method=method, scaling="variance", squash=False)
    st.sync()
    st.close()


def nnls_merge_once(pkg, cs, max_pcols, method, n_check=256):
    """Upload the cells, merge every cell once (timed by the library's event
    timer around the merge kernels; an untimed merge of a few cells with the
    same method runs first), then the largest relative moment error
    |M_after - M_before| / sum w |v^k| over a sample of merged cells."""
    _nnls_warm(pkg, cs, max_pcols, method)
    n, counts, starts = cs["n"], cs["counts"], cs["starts"]
    pia = np.zeros((NN_CELLS, 7), dtype=np.int64)
    pia[:, 0] = counts
    pia[:, 1] = starts + 1
    pia[:, 2] = starts + counts
    pia[:, 3] = counts
    pia[:, 5] = -1
    st = pkg.DeviceStore(NN_CELLS, 1, n + 1024)
    st.upload_species(1, cs["x"], cs["v"], cs["w"], np.arange(1, n + 1, dtype=np.int64), pia, n, True,
                      np.zeros(0, dtype=np.int64), n)
    mim = _mim(4)
    st.sync()
    st.timing_reset()
    st.timing(True)
    t0 = time.perf_counter()
    if max_pcols is None:
        status = pkg.merge_nnls_based(st, 1, 1, mim, scaling="variance", squash=False)
    else:
        status = pkg.merge_nnls_candidates(st, 1, 1, mim, max_pcols, method=method, scaling="variance",
                                           squash=False)
    st.sync()
    el = time.perf_counter() - t0
    ms, _ = st.timing_get("merge_nnls")
    st.timing(False)
    merged = int(np.sum(status == 1))
    redo = st.nnls_info()[1] if max_pcols is not None else 0
    got = st.download_species(1)
    st.close()
    sample = np.linspace(0, NN_CELLS - 1, n_check).astype(int)
    full = _mim(4) + [(0, 0, 0)]
    worst = 0.0
    w, v = cs["w"], cs["v"]
    for c in sample:
        if status[c] != 1:
            continue
        a0, a1 = starts[c], starts[c] + counts[c]
        pre = _moments(w[a0:a1], v[a0:a1], full)
        scale = _abs_moments(w[a0:a1], v[a0:a1], full)
        r = got["pia"][c]
        post = _moments(got["w"][r[1] - 1:r[2]], got["v"][r[1] - 1:r[2]], full)
        worst = max(worst, float(np.max(np.abs(post - pre) / scale)))
    return dict(merges_per_s=merged / (ms * 1e-3), kernel_ms=ms, wall_s=el, merged=merged,
                failed=int(np.sum(status == -1)), fp64_redo=redo, max_rel_moment_error_sampled=worst)


# The candidate set: the first 239 particles of a cell plus the reference's
# 17 fictitious particles (centre, 8 octants at 0.5 and 1.0 sigma), 256
# columns.  64 columns (47 particles) leave the 35-moment system infeasible
# (the residual stays near 10% of |b|), so those merges fail.
NN_PCOLS = 239


def run_nnls(a, pkg):
    cs = sample_nnls_cells()
    wav = nnls_merge_once(pkg, cs, NN_PCOLS, "wave")
    mix = nnls_merge_once(pkg, cs, NN_PCOLS, "mixed")
    lds = nnls_merge_once(pkg, cs, NN_PCOLS, "lds")
    ref = nnls_merge_once(pkg, cs, NN_PCOLS, "reference")
    k64 = {m: nnls_merge_once(pkg, cs, 47, m, n_check=0) for m in ("lds", "wave")}
    best = wav
    out = {"metric": "NNLS merges/s at <= 1e-10 moment error", "value": best["merges_per_s"],
           "unit": "merges/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32 search, f64 solve",
           "data": "synthetic", "wall_s": best["wall_s"], "kernel_ms": best["kernel_ms"], "vs_baseline": None,
           "config": {"workload": "32768 cells x U{200..3000} bi-Maxwellian particles, lognormal(0.5) weights; "
                                  "merge_nnls_based! with base + order <= 4 multi-index moments (35 rows), "
                                  ":variance, columns = first 239 particles + 17 fictitious",
                      "n_particles": cs["n"], "max_pcols": NN_PCOLS, "method": "wave",
                      "merged": best["merged"], "failed": best["failed"], "refined_in_fp64": best["fp64_redo"],
                      "max_rel_moment_error_sampled": best["max_rel_moment_error_sampled"],
                      "rows": len(set(_mim(4)) | {(0, 0, 0)})},
           "mixed_lds": mix, "lds_fp64": lds, "reference_layout_fp64": ref,
           "speedup_vs_lds_fp64": best["merges_per_s"] / lds["merges_per_s"],
           "speedup_vs_mixed_lds": best["merges_per_s"] / mix["merges_per_s"],
           "speedup_vs_reference_layout": best["merges_per_s"] / ref["merges_per_s"],
           "k64_columns": {m: {"merged": r["merged"], "failed": r["failed"]} for m, r in k64.items()}}
    print(json.dumps(out), flush=True)


def run_roulette(a, pkg):
    """Roulette merging (merge_roulette!) of the NNLS workload's 32,768 cells
    (U{200..3000} particles) to 64 particles each, plain and conservative."""
    cs = sample_nnls_cells()
    n, counts, starts = cs["n"], cs["counts"], cs["starts"]
    pia = np.zeros((NN_CELLS, 7), dtype=np.int64)
    pia[:, 0] = counts
    pia[:, 1] = starts + 1
    pia[:, 2] = starts + counts
    pia[:, 3] = counts
    pia[:, 5] = -1
    res = {}
    for conservative in (False, True):
        st = pkg.DeviceStore(NN_CELLS, 1, n + 1024)
        st.upload_species(1, cs["x"], cs["v"], cs["w"], np.arange(1, n + 1, dtype=np.int64), pia, n, True,
                          np.zeros(0, dtype=np.int64), n)
        st.set_seed(5)
        st.sync()
        st.timing_reset()
        st.timing(True)
        t0 = time.perf_counter()
        merged = pkg.merge_roulette(st, 1, 64, 64, conservative=conservative, tag=3)
        st.sync()
        el = time.perf_counter() - t0
        ms, _ = st.timing_get("merge_roulette")
        st.timing(False)
        sc = st.scalars(1)
        assert sc["err"] == 0 and sc["n_total"] == 64 * NN_CELLS, sc
        res["conservative" if conservative else "plain"] = {
            "merged": merged, "kernel_ms": ms, "wall_s": el, "merges_per_s": merged / (ms * 1e-3),
            "deleted_per_s": (n - 64 * NN_CELLS) / (ms * 1e-3)}
        st.close()
    out = {"metric": "roulette merges/s", "value": res["plain"]["merges_per_s"], "unit": "merges/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic", "vs_baseline": None,
           "config": {"workload": "32768 cells x U{200..3000} bi-Maxwellian particles, lognormal(0.5) weights; "
                                  "merge_roulette! of every cell to 64 particles", "n_particles": n},
           "plain": res["plain"], "conservative": res["conservative"]}
    print(json.dumps(out), flush=True)


# SWPM collisions with octree N:2 merging (the loop of
# test/test_bkw_varweight_octree_swpm.jl) over many independent BKW boxes:
# 12,288 cells of V = 1, each 1,000 equal-weight argon particles (Fnum =
# 1e23 / 1,000, pseudo-Maxwell molecules) with BKW velocities at 273 K, at
# the cell centre.  A step: swpm! with G = 1 in every cell, octree N:2
# merging of every cell above 1,250 particles to 1,000 (512 bins), and the
# sort that puts the new particles back in their cells' runs.
SW_CELLS, SW_PPC, SW_THR, SW_TARGET, SW_T0, SW_NDENS = 12288, 1000, 1250, 1000, 273.0, 1e23


def sample_bkw_boxes(n_cells, ppc, seed=11):
    g = np.random.Generator(np.random.Philox(seed))
    n = n_cells * ppc
    m = 66.3e-27
    vscale = np.sqrt(2 * K_B * SW_T0 / m) * np.sqrt(0.3)
    vabs = np.sqrt(g.chisquare(5, size=n))
    th = g.uniform(size=n) * np.pi
    ph = g.uniform(size=n) * 2 * np.pi
    v = np.empty((n, 3))
    v[:, 0] = vscale * vabs * np.sin(th) * np.cos(ph)
    v[:, 1] = vscale * vabs * np.sin(th) * np.sin(ph)
    v[:, 2] = vscale * vabs * np.cos(th)
    x = np.full((n, 3), 0.5)
    x[:, 0] = np.repeat(np.arange(n_cells, dtype=np.float64), ppc) + 0.5
    w = np.full(n, SW_NDENS / ppc)
    pia = np.zeros((n_cells, 7), dtype=np.int64)
    pia[:, 0] = ppc
    pia[:, 1] = np.arange(n_cells) * ppc + 1
    pia[:, 2] = np.arange(1, n_cells + 1) * ppc
    pia[:, 3] = ppc
    pia[:, 5] = -1
    return dict(x=x, v=v, w=w, pia=pia, n=n, mass=m)


def swpm_cpu_baseline(sp, ia, dt, n_cells=128, target_s=12.0):
    """The C oracle's swpm!, octree merge and sort on one core, over n_cells
    boxes of the same workload; returns (particle-timesteps/s, steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mzo
    mzo.build()
    case = sample_bkw_boxes(n_cells, SW_PPC, seed=12)
    n = case["n"]
    pv, pia = mzo.ParticleVector(n), mzo.PIA(n_cells, 1)
    for i in range(1, n + 1):
        pv.pop_to(i)
    pv.x[:n], pv.v[:n], pv.w[:n] = case["x"], case["v"], case["w"]
    for c in range(n_cells):
        pia.set(c + 1, 1, *case["pia"][c])
    pia.set_n_total(1, n)
    grid = mzo.grid1d(float(n_cells), n_cells)
    oit = mzo.interactions(["Ar"], table=mzo.PSEUDO_MAXWELL)[0][0]
    cfs = [mzo.CFSwpm() for _ in range(n_cells)]
    for cf in cfs:
        cf.sigma_g_max = mzo.lib().mzo_estimate_sigma_g_w_max(C.byref(oit), sp[0].mass, sp[0].mass, SW_T0, SW_T0,
                                                              1.0, 1.0)
    oc = mzo.Octree(split=1, init_bin_bounds=1, bin_bounds_compute=1, max_Nbins=512, max_depth=10)

    def step(k):
        r = mzo.RNG(seed=11, tag=(k << 4) | 1)
        done = 0
        for c in range(1, n_cells + 1):
            done += pia.ind(c).n_local
            mzo.lib().mzo_swpm(r.ref, C.byref(cfs[c - 1]), C.byref(oit), pv.p, pia.p, c, 1, 1.0, dt, 1.0)
        r = mzo.RNG(seed=11, tag=(k << 4) | 3)
        for c in range(1, n_cells + 1):
            if pia.ind(c).n_local > SW_THR:
                mzo.lib().mzo_merge_octree_N2_based(r.ref, oc.p, pv.p, pia.p, c, 1, SW_TARGET, None)
        mzo.lib().mzo_squash_pia(pv.p, pia.p, 1)
        mzo.lib().mzo_sort_particles(C.byref(grid), pv.p, pia.p, 1)
        return done

    for k in range(1, 6):
        step(k)
    t0 = time.perf_counter()
    done, k = 0, 6
    while time.perf_counter() - t0 < target_s:
        done += step(k)
        k += 1
    return done / (time.perf_counter() - t0), k - 6


def run_bkw_swpm(a, pkg):
    case = sample_bkw_boxes(SW_CELLS, SW_PPC)
    n = case["n"]
    data = os.path.join(ROOT, "merzbild.jl_amd", "data")
    sp = pkg.load_species_data(os.path.join(data, "particles.toml"), ["Ar"])
    it = pkg.load_interaction_data(os.path.join(data, "pseudo_maxwell.toml"), sp)
    st = pkg.DeviceStore(SW_CELLS, 1, 2 * n)
    st.upload_species(1, case["x"], case["v"], case["w"], np.arange(1, n + 1, dtype=np.int64), case["pia"], n, True,
                      np.zeros(0, dtype=np.int64), n)
    del case["x"], case["v"], case["w"]
    st.set_grid(pkg.Grid1DUniform(float(SW_CELLS), SW_CELLS))     # V = 1 per cell
    st.set_interactions(it)
    st.set_species_data(sp)
    ia = it[0][0]
    st.set_sigma_g_w_max(1, 1, np.full(SW_CELLS, pkg.estimate_sigma_g_w_max(ia, sp[0].mass, sp[0].mass, SW_T0,
                                                                            SW_T0, 1.0)))
    st.set_seed(11)
    tref = 1.0 / (SW_NDENS * np.pi * ia.vhs_d ** 2) / np.sqrt(2 * K_B * SW_T0 / sp[0].mass)
    dt = 0.025 * tref
    steps = a.steps if a.steps != 20000 else 40
    warm = a.warmup if a.warmup != 200 else 10

    def step(k):
        pkg.swpm(st, 1, 1.0, dt, tag=(k << 4) | 1)
        merged = pkg.merge_octree_N2_based(st, 1, SW_THR, SW_TARGET, max_Nbins=512, squash=True, tag=(k << 4) | 3)
        pkg.sort_particles(st, 1)
        return merged

    for k in range(1, warm + 1):
        step(k)
    st.sync()
    st.timing_reset()
    st.timing(True)
    done, merged = 0, 0
    t0 = time.perf_counter()
    for k in range(warm + 1, warm + steps + 1):
        done += st.scalars(1)["n_total"]
        merged += step(k)
    st.sync()
    el = time.perf_counter() - t0
    fam = {f: st.timing_get(f)[0] for f in ("swpm", "merge_octree", "sort")}
    st.timing(False)
    cf = st.collision_factors(1, 1)
    sc = st.scalars(1)
    assert sc["err"] == 0, sc
    out = {"metric": "particle-timesteps/s (SWPM + octree merge + sort), BKW boxes", "value": done / el,
           "unit": "particle-timesteps/s", "n_gpus": 1, "steps": steps, "warmup": warm, "ms_per_step": el * 1e3 / steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "12288 0-D BKW boxes x 1000 Ar particles (pseudo-Maxwell, 273 K), swpm! G = 1, "
                                  "octree N:2 merge above 1250 to 1000 (512 bins), sort; dt = 0.025 t_ref",
                      "n_particles_end": sc["n_total"], "cells_merged_per_step": merged / steps,
                      "collisions_per_step": int(np.sum(cf["n_coll_performed"]))},
           "kernel_ms_per_call": fam}
    if not a.no_cpu_baseline:
        v, k = swpm_cpu_baseline(sp, ia, dt)
        out["cpu_baseline"] = {"value": v, "unit": "particle-timesteps/s", "cores": 1, "kind": "port",
                               "sample": f"{k} steps of the same loop over 128 of the boxes on the C oracle"}
    print(json.dumps(out), flush=True)
    st.close()


def grid_merge_cpu_baseline(sp, n_cells=512, target_s=10.0):
    """The C oracle's merge_grid_based! on one core over n_cells of the BKW
    boxes, re-sampled between passes; returns the cpu_baseline object."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mzo
    mzo.build()
    case = sample_bkw_boxes(n_cells, SW_PPC, seed=12)
    n = case["n"]
    mg = mzo.GridMerge(4, 4, 4, 3.5)
    done, el = 0, 0.0
    while el < target_s:
        pv, pia = mzo.ParticleVector(n), mzo.PIA(n_cells, 1)
        for i in range(1, n + 1):
            pv.pop_to(i)
        pv.x[:n], pv.v[:n], pv.w[:n] = case["x"], case["v"], case["w"]
        for c in range(n_cells):
            pia.set(c + 1, 1, *case["pia"][c])
        pia.set_n_total(1, n)
        pp = mzo.Props(n_cells, 1)
        mzo.compute_props([pv], pia, [sp[0].mass], pp)
        r = mzo.RNG(seed=11, tag=3)
        t0 = time.perf_counter()
        for c in range(1, n_cells + 1):
            mg.extent_props(pp.T[0, c - 1], pp.v[0, c - 1], sp[0].mass)
            mzo.lib().mzo_merge_grid_based(r.ref, mg.p, pv.p, pia.p, c, 1, None)
        el += time.perf_counter() - t0
        done += n_cells
    return {"value": done / el, "unit": "merges/s", "cores": 1, "kind": "port",
            "sample": f"{done} box merges ({n_cells} boxes per pass) on the C oracle"}


def run_grid_merge(a, pkg):
    """merge_grid_based! (GridN2Merge(4, 4, 4, 3.5), extents from each cell's
    props) of the 12,288 BKW boxes of 1,000 particles: one call merges every
    box; the boxes are re-uploaded between repetitions."""
    case = sample_bkw_boxes(SW_CELLS, SW_PPC)
    n = case["n"]
    data = os.path.join(ROOT, "merzbild.jl_amd", "data")
    sp = pkg.load_species_data(os.path.join(data, "particles.toml"), ["Ar"])
    st = pkg.DeviceStore(SW_CELLS, 1, n + 1024)
    mg = pkg.GridN2Merge(4, extent_multiplier=3.5)
    reps = 5
    times, out_np = [], 0
    for r in range(reps + 1):
        st.upload_species(1, case["x"], case["v"], case["w"], np.arange(1, n + 1, dtype=np.int64), case["pia"], n,
                          True, np.zeros(0, dtype=np.int64), n)
        st.set_seed(11)
        pkg.compute_props(st, sp)
        st.sync()
        st.timing_reset()
        st.timing(True)
        merged = pkg.merge_grid_based(st, 1, 2, mg, species_data=sp, squash=False, tag=(r << 4) | 3)
        st.sync()
        ms, _ = st.timing_get("merge_grid")
        st.timing(False)
        sc = st.scalars(1)
        assert sc["err"] == 0 and merged == SW_CELLS, sc
        out_np = sc["n_total"]
        if r > 0:      # the first call allocates the scratch
            times.append(ms)
    ms = float(np.median(times))
    cpu = None
    if not a.no_cpu_baseline:
        cpu = grid_merge_cpu_baseline(sp)
    out = {"metric": "grid N:2 merges/s", "value": SW_CELLS / (ms * 1e-3), "unit": "merges/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic", "vs_baseline": None,
           "ms_per_call": ms, "particles_per_s": n / (ms * 1e-3),
           "config": {"workload": "12288 BKW boxes x 1000 Ar particles, merge_grid_based! with GridN2Merge(4, 3.5), "
                                  "extents from the props", "n_particles": n, "n_after": out_np}}
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    st.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--mode", choices=("slab", "graph", "eager", "batched", "plain"), default="slab",
                    help="slab: all steps in one launch, cells resident in LDS; graph: replayed step graph; "
                         "eager: kernel by kernel")
    ap.add_argument("--graph-steps", type=int, default=400, help="steps of the step-graph path timed alongside")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-nnls", action="store_true", help="couette: skip the NNLS merges/s half of the metric")
    ap.add_argument("--replicas", type=int, default=16, help="couette: independent replicas run side by side (0: skip)")
    ap.add_argument("--replica-steps", type=int, default=400, help="couette: steps of the replica run")
    ap.add_argument("--no-compare", action="store_true",
                    help="batched-0d: skip timing the plain and per-cell launch patterns")
    ap.add_argument("--no-autotune", action="store_true", help="slab mode: keep the library's default shape")
    ap.add_argument("--workload", choices=("couette", "batched-0d", "batched-0d-split", "couette-varweight",
                                           "nnls-merge", "couette-surf", "roulette-merge", "bkw-swpm", "grid-merge"),
                    default="couette",
                    help="couette: the headline 1-D Couette line; batched-0d: 65,536 uneven 0-D cells; "
                         "nnls-merge: 32,768 cells merged by NNLS; couette-surf: the averaging phase of the "
                         "2,000-cell Couette run with wall properties")
    ap.add_argument("--surf-headline", action="store_true",
                    help="couette-surf: run the averaging phase on the headline 500-cell configuration")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run a plain device copy of MB megabytes three times (to scale the HBM byte "
                         "counters of a profiler run) and exit")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        n = a.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n * 4, "launches": 3}))
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1", "n_gpus": a.gpus}))
        sys.exit(2)
    pkg = load_pkg()
    if a.workload == "batched-0d":
        run_batched_0d(a, pkg)
        return
    if a.workload == "batched-0d-split":
        run_batched_0d_split(a, pkg)
        return
    if a.workload == "couette-varweight":
        run_couette_varweight(a, pkg)
        return
    if a.workload == "nnls-merge":
        run_nnls(a, pkg)
        return
    if a.workload == "couette-surf":
        run_couette_surf(a, pkg)
        return
    if a.workload == "roulette-merge":
        run_roulette(a, pkg)
        return
    if a.workload == "bkw-swpm":
        run_bkw_swpm(a, pkg)
        return
    if a.workload == "grid-merge":
        run_grid_merge(a, pkg)
        return
    case = sample_case()
    st, sp, it = setup_store(pkg, case)
    n = case["n"]

    # warmup (also captures the step graph / sizes the slab scratch); in slab
    # mode the warmup is the shape autotuner's trial runs
    tune = None
    if a.mode == "slab" and not a.no_autotune:
        best, times, step0 = pkg.couette_autotune(st, 1, DT, steps=a.warmup)
        tune = {"best": list(best), "ms_per_step": {f"{k[0]}x{k[1]}": round(v, 5) for k, v in times.items()}}
    else:
        pkg.couette_steps(st, a.warmup, 1, DT, props_every=1, mode=a.mode)
        step0 = 1 + a.warmup
    st.sync()
    st.timing_reset()
    st.timing(True)
    t0 = time.perf_counter()
    path = pkg.couette_steps(st, a.steps, step0, DT, props_every=1, mode=a.mode)
    st.sync()
    t1 = time.perf_counter()
    fam = "slab" if path == "slab" else "step"
    launch_ms, launches = st.timing_get(fam)
    st.timing(False)
    slab_info = st.slab_info() if path == "slab" else None
    elapsed = t1 - t0
    value = n * a.steps / elapsed
    sc = st.scalars(1)
    assert sc["err"] == 0 and sc["n_total"] == n, sc

    # the step-by-step graph path on the same store, for comparison
    st.timing_reset()
    st.timing(True)
    s1 = step0 + a.steps
    g0 = time.perf_counter()
    pkg.couette_steps(st, a.graph_steps, s1, DT, props_every=1, mode="graph")
    st.sync()
    graph_ms = (time.perf_counter() - g0) * 1e3 / max(a.graph_steps, 1)
    # per-kernel durations of the step-by-step kernels, HIP events around each
    kern = {}
    st.timing_reset()
    s2 = s1 + a.graph_steps
    for t in range(100):
        stp = s2 + t
        pkg.ntc(st, 1, 1, DT, equal_weight=True, tag=(stp << 4) | 1)
        pkg.convect_particles(st, 1, DT, tag=(stp << 4) | 2)
        pkg.sort_particles(st, 1)
        pkg.compute_props_sorted(st, sp, fetch=False)
    for f in ("ntc", "convect", "sort", "props"):
        kern[f] = round(st.timing_get(f)[0], 5)
    st.timing(False)
    sc = st.scalars(1)
    assert sc["err"] == 0 and sc["n_total"] == n, sc

    # roofline of the dominant kernel: algorithmic bytes of the steps one launch
    # runs over that launch's duration (HIP events on the library's stream; the
    # library reports the launch duration divided by the steps it ran)
    steps_per_launch = a.steps if path == "slab" else 1
    kernel = "k_slab_couette" if path == "slab" else "step graph"
    step_s = launch_ms * 1e-3
    achieved = ALG_BYTES["step"] * n / step_s / 1e9
    peak = 8000.0
    traffic = SLAB_HBM_BYTES_PER_PARTICLE_STEP * n * steps_per_launch if path == "slab" else None
    roof = {"bound": "hbm", "achieved": achieved, "peak": peak, "unit": "GB/s", "frac": achieved / peak,
            "traffic": traffic, "traffic_bytes_per_particle_step": SLAB_HBM_BYTES_PER_PARTICLE_STEP
            if path == "slab" else None, "kernel": kernel, "launch_ms": launch_ms * steps_per_launch,
            "kernel_ms_per_step": launch_ms, "steps_per_launch": steps_per_launch,
            "alg_bytes_per_particle_step": ALG_BYTES["step"]}

    # 16 independent replicas of the same Couette case (seeds differ), each in
    # its own store on its own stream: the ensemble the survey's C2 row asks
    # for.  Each replica runs its steps in one slab launch with the shape
    # chosen above; a slab launch needs its slabs resident at once, so a
    # store's launch waits for the other stores' work and the replicas take
    # the device in turn.
    replicas = None
    if a.replicas > 0:
        stores = []
        shape = tuple(tune["best"]) if tune else (0, 0)
        for r in range(a.replicas):
            rc = sample_case(seed=12345 + 1000 * (r + 1))
            rs, _, _ = setup_store(pkg, rc)
            rs.set_seed(20000 + r)
            rs.set_slab_shape(*shape)
            pkg.couette_steps(rs, 4, 1, DT, props_every=1, mode=a.mode)
            stores.append(rs)
        for rs in stores:
            rs.sync()
        k = a.replica_steps
        paths = set()
        t0 = time.perf_counter()
        for rs in stores:
            paths.add(pkg.couette_steps(rs, k, 5, DT, props_every=1, mode=a.mode))
        for rs in stores:
            rs.sync()
        el_r = time.perf_counter() - t0
        n_r = sum(rs.scalars(1)["n_total"] for rs in stores)
        for rs in stores:
            assert rs.scalars(1)["err"] == 0
            rs.close()
        ach = ALG_BYTES["step"] * n_r * k / el_r / 1e9
        replicas = {"replicas": a.replicas, "steps": k, "n_particles": n_r, "value": n_r * k / el_r,
                    "unit": "particle-timesteps/s", "ms_per_step": el_r * 1e3 / k,
                    "path": ", ".join(sorted(paths)) + ", one store and stream per replica, in turn",
                    "roofline": {"bound": "hbm", "achieved": ach, "peak": 8000.0, "unit": "GB/s", "frac": ach / 8000.0,
                                 "alg_bytes_per_particle_step": ALG_BYTES["step"],
                                 "note": "wall clock of all replicas' steps (host enqueue included)"}}

    # the second half of the metric: NNLS merges/s at <= 1e-10 moment error on
    # the batched NNLS workload, the wave method beside the plain FP64 one
    nnls = None
    if not a.no_nnls:
        cs = sample_nnls_cells()
        wav = nnls_merge_once(pkg, cs, NN_PCOLS, "wave")
        lds = nnls_merge_once(pkg, cs, NN_PCOLS, "lds")
        nnls = {"metric": "NNLS merges/s at <= 1e-10 moment error", "value": wav["merges_per_s"],
                "unit": "merges/s",
                "method": "wave (one wavefront per cell: FP32 active-set search in registers, FP64 passive-set solve)",
                "kernel_ms": wav["kernel_ms"], "merged": wav["merged"], "failed": wav["failed"],
                "refined_in_fp64": wav["fp64_redo"],
                "max_rel_moment_error_sampled": wav["max_rel_moment_error_sampled"],
                "plain_fp64": {"merges_per_s": lds["merges_per_s"], "kernel_ms": lds["kernel_ms"],
                               "max_rel_moment_error_sampled": lds["max_rel_moment_error_sampled"]},
                "speedup_vs_plain_fp64": wav["merges_per_s"] / lds["merges_per_s"],
                "workload": "32768 cells x U{200..3000} bi-Maxwellian particles, lognormal(0.5) weights, 35 "
                            "moments (order <= 4), :variance, columns = first 239 particles + 17 fictitious",
                "n_particles": cs["n"]}
        del cs
    cpu = None
    if not a.no_cpu_baseline:
        v_cpu, threads, steps, v_serial, serial_steps = cpu_baseline(case, it)
        cpu = {"value": v_cpu, "unit": "particle-timesteps/s", "cores": threads, "kind": "port",
               "sample": f"{steps} Couette steps x {n} particles (C oracle -O3, OpenMP over cells, threaded "
                         f"counting sort); serial: {serial_steps} steps on 1 thread",
               "serial_value": v_serial, "cpu_model": _cpu_model(),
               "reference_published": {"serial_i9_13900K_50k": 8.8e7, "threads32_epyc_9374F": 1.2e8,
                                       "source": "BENCHMARKS.md:23,240 (derived particle-timesteps/s)"}}
    out = {
        # the metric's first half is "value"; its second half is in "nnls"
        "metric": "particle-timesteps/s (sort+NTC+props); NNLS merges/s at \u22641e-10 moment error",
        "value": value,
        "unit": "particle-timesteps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "1D Couette, argon VHS, 500 cells x 200 particles (100k), dt 2e-10 s, FP64; "
                               "per step ntc!, convection, counting sort, compute_props_sorted!",
                   "n_particles": n, "n_cells": N_CELLS, "mode": path, "slab_shape": slab_info, "autotune": tune,
                   "graph_ms_per_step": graph_ms, "step_kernels_ms": kern},
        "roofline": roof,
        "cpu_baseline": cpu,
        "nnls": nnls,
        "replicas": replicas,
    }
    print(json.dumps(out), flush=True)
    st.close()


if __name__ == "__main__":
    main()
