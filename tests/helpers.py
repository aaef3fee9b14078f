"""Shared test helpers: build reference-layout cases with the CPU oracle, move
them onto the device store, and compare device state with the oracle's.

Test infrastructure only (imports the oracle)."""
import ctypes as C

import numpy as np

MASS_AR = 66.3e-27   # data/particles.toml
KB = 1.380649e-23


def oracle_upload_arrays(mzo, pv, pia, species=1):
    """The device upload arrays of an oracle ParticleVector + PIA."""
    tab = pia.table(species)
    n_total = pia.n_total(species)
    n_pos = int(max(tab[:, 2].max(initial=0), tab[:, 5].max(initial=0), n_total if pia.contiguous(species) else 0))
    ids = pv.index[:n_pos].copy()
    x, v, w = pv.x[ids - 1].copy(), pv.v[ids - 1].copy(), pv.w[ids - 1].copy()
    return dict(x=x, v=v, w=w, ids=ids, pia=tab, n_total=n_total, contiguous=pia.contiguous(species),
                buffer=pv.buffer[:pv.nbuffer].copy(), length=pv.length)


def upload_from_oracle(store, mzo, pv, pia, species=1):
    a = oracle_upload_arrays(mzo, pv, pia, species)
    store.upload_species(species, a["x"], a["v"], a["w"], a["ids"], a["pia"], a["n_total"], a["contiguous"],
                         a["buffer"], a["length"])


def assert_state_matches(mzo, got, pv, pia, species=1, rtol=1e-12, check_buffer=True):
    """Integer state bit for bit; x, v, w per logical position within rtol
    (relative to each component's scale)."""
    a = oracle_upload_arrays(mzo, pv, pia, species)
    assert got["n_total"] == a["n_total"], (got["n_total"], a["n_total"])
    assert bool(got["contiguous"]) == bool(a["contiguous"])
    assert np.array_equal(got["pia"], a["pia"]), np.argwhere(got["pia"] != a["pia"])[:10]
    n = len(a["ids"])
    assert got["n_pos"] >= n
    # compare only positions referenced by some cell
    used = np.zeros(n, dtype=bool)
    for row in a["pia"]:
        if row[3] > 0:
            used[row[1] - 1:row[2]] = True
        if row[6] > 0:
            used[row[4] - 1:row[5]] = True
    gi = got["index"][:n]
    assert np.array_equal(gi[used], a["ids"][used]), np.argwhere(gi[used] != a["ids"][used])[:10]
    if check_buffer:
        assert got["nbuffer"] == pv.nbuffer and got["length"] == pv.length
        assert np.array_equal(got["buffer"], a["buffer"])
    for name in ("x", "v"):
        g, r = got[name][:n][used], a[name][used]
        scale = np.maximum(np.abs(r).max(axis=0, initial=0.0), 1e-300)
        err = np.abs(g - r) / scale
        assert err.max(initial=0.0) <= rtol, (name, err.max(), np.unravel_index(err.argmax(), err.shape))
    gw, rw = got["w"][:n][used], a["w"][used]
    assert np.all(np.abs(gw - rw) <= rtol * np.abs(rw)), "weights differ"


def couette_case(mzo, n_cells=8, ppc=64, seed=5, T=300.0, ndens=1e23, v_wall=500.0, dt=2.0e-10,
                 weights=None, sgwm_mult=1.0, dx=1e-6):
    """A 1-D Couette setup in the reference layout (test/test_1D_couette.jl):
    argon VHS, cells of dx, diffuse walls at T moving at -/+ v_wall,
    ppc particles per cell sampled Maxwellian, Fnum = ndens dx / ppc."""
    L = dx * n_cells
    grid = mzo.grid1d(L, n_cells)
    walls = mzo.walls([MASS_AR], T, T, -v_wall, v_wall, 1.0, 1.0)
    it = mzo.interactions(["Ar"])[0][0]
    Fnum = ndens * grid.dx / ppc
    n = n_cells * ppc
    pv = mzo.ParticleVector(n)
    pia = mzo.PIA(n_cells, 1)
    r = mzo.RNG(seed=seed, tag=0xABCD)
    for c in range(1, n_cells + 1):
        r.s.stream, r.s.k = c - 1, 0
        mzo.lib().mzo_sample_particles_equal_weight(r.ref, pv.p, pia.p, c, 1, ppc, MASS_AR, T, Fnum,
                                                    (c - 1) * grid.dx, c * grid.dx, 0.0, 1.0, 0.0, 1.0)
    if weights is not None:
        rng = np.random.default_rng(seed)
        ids = pv.index[:n]
        pv.w[ids - 1] = Fnum * weights(rng, n)
    sg = sgwm_mult * mzo.lib().mzo_estimate_sigma_g_w_max(C.byref(it), MASS_AR, MASS_AR, T, T, Fnum, 1.0)
    cfs = (mzo.CF * n_cells)()
    for c in range(n_cells):
        cfs[c].sigma_g_w_max = sg
    return dict(pv=pv, pia=pia, grid=grid, walls=walls, it=it, cfs=cfs, dt=dt, Fnum=Fnum, sgwm=sg, seed=seed,
                T=T, v_wall=v_wall, L=L, n_cells=n_cells, mass=MASS_AR, ppc=ppc)


def store_from_case(pkg, case, capacity=None):
    nc = case["n_cells"]
    n = case["pia"].n_total()
    cap = capacity or (2 * n + 1024)
    store = pkg.DeviceStore(nc, 1, cap)
    upload_from_oracle(store, None, case["pv"], case["pia"])
    grid = pkg.Grid1DUniform(case["L"], nc)
    store.set_grid(grid)
    sp = [pkg.Species("Ar", case["mass"])]
    case["species"] = sp
    store.set_species_data(sp)
    store.set_walls(pkg.MaxwellWalls1D(sp, case["T"], case["T"], -case["v_wall"], case["v_wall"], 1.0, 1.0))
    it = case["it"]
    store.set_interactions([[type("I", (), dict(mu1=it.mu1, mu2=it.mu2, vhs_factor=it.vhs_factor,
                                               vhs_o=it.vhs_o))()]])
    store.set_sigma_g_w_max(1, 1, np.array([case["cfs"][c].sigma_g_w_max for c in range(nc)]))
    store.set_seed(case["seed"])
    return store


def oracle_couette_step(mzo, case, step, equal_weight=False, dw_tol=1e-16):
    """One step of test/test_1D_couette.jl:66-92 on the oracle, with the
    library's tag layout (step << 4) | phase."""
    pv, pia, g = case["pv"], case["pia"], case["grid"]
    r = mzo.RNG(seed=case["seed"], tag=(step << 4) | 1)
    for c in range(1, case["n_cells"] + 1):
        mzo.lib().mzo_ntc_single(r.ref, C.byref(case["cfs"][c - 1]), C.byref(case["it"]), pv.p, pia.p, c, 1,
                                 case["dt"], g.dx, dw_tol, 1 if equal_weight else 0)
    r = mzo.RNG(seed=case["seed"], tag=(step << 4) | 2)
    mzo.lib().mzo_convect_particles(r.ref, C.byref(g), C.byref(case["walls"]), pv.p, pia.p, 1, case["dt"])
    mzo.lib().mzo_sort_particles(C.byref(g), pv.p, pia.p, 1)


def oracle_props_sorted(mzo, case):
    pp = mzo.Props(case["n_cells"], 1)
    mzo.compute_props_sorted([case["pv"]], case["pia"], [case["mass"]], pp)
    return pp.T[0].copy()
