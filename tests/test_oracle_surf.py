"""Surface properties: the oracle pinned to the reference's own known answers
(test/test_surface_props_1D_uniform.jl) and to what the physics fixes
(specular walls reflect without energy or tangential momentum exchange; the
force on the walls is the momentum the particles lose there), and the
package's host-side SurfProps helpers held to the same answers.  CPU only.
"""
import numpy as np
import pytest

from test_oracle_kat import make_pv

EPS = np.finfo(float).eps
M_AR = 66.3e-27   # data/particles.toml


def test_surf_layout(mzo):
    """test_surface_props_1D_uniform.jl:17-33"""
    sp = mzo.Surf.grid1d(1)
    assert sp.n_elements == 2 and sp.n_species == 1
    assert list(sp.inv_areas) == [1.0, 1.0]
    assert list(sp.normals[0]) == [1.0, 0.0, 0.0] and list(sp.normals[1]) == [-1.0, 0.0, 0.0]
    assert sp.np.shape == (2, 1) and sp.force.shape == (3, 2, 1) and sp.shear_pressure.shape == (3, 2, 1)


def test_surf_incident_reflected_scale_kat(mzo):
    """test_surface_props_1D_uniform.jl:69-167: particle 3 hits the left wall
    and particle 8 the right one; specular reflection on the left, a new
    velocity and weight on the right; then the scaling by mass / dt / area."""
    sp = mzo.Surf.grid1d(1)
    sp.clear()
    sp.update_incident(3.5, [-9.0, 9.0, 0.0], 1, 1)
    sp.update_incident(10.5, [10.0, 0.0, 3.0], 1, 2)
    assert list(sp.np[:, 0]) == [1.0, 1.0]
    assert list(sp.flux_incident[:, 0]) == [3.5, 10.5] and list(sp.flux_reflected[:, 0]) == [0.0, 0.0]
    assert abs(sp.force[0, 0, 0] - 3.5 * -9) < 2 * EPS and abs(sp.force[0, 1, 0] - 10.5 * 10) < 2 * EPS
    assert abs(sp.force[1, 0, 0] - 3.5 * 9) < 2 * EPS and sp.force[1, 1, 0] == 0.0
    assert sp.force[2, 0, 0] == 0.0 and abs(sp.force[2, 1, 0] - 10.5 * 3) < 2 * EPS
    assert abs(sp.normal_pressure[0, 0] - 9 * 3.5) < 2 * EPS
    assert abs(sp.normal_pressure[1, 0] - 10 * 10.5) < 2 * EPS
    assert abs(sp.shear_pressure[0, 0, 0]) < 2 * EPS and abs(sp.shear_pressure[1, 0, 0] - 9 * 3.5) < 2 * EPS
    assert abs(sp.shear_pressure[2, 0, 0]) < 2 * EPS
    assert abs(sp.shear_pressure[0, 1, 0]) < 2 * EPS and abs(sp.shear_pressure[1, 1, 0]) < 2 * EPS
    assert abs(sp.shear_pressure[2, 1, 0] - 3 * 10.5) < 2 * EPS
    assert abs(sp.kinetic_energy_flux[0, 0] - 0.5 * 3.5 * (81 + 81)) < 2 * EPS
    assert abs(sp.kinetic_energy_flux[1, 0] - 0.5 * 10.5 * (100 + 9)) < 2 * EPS

    sp.update_reflected(3.5, [9.0, 9.0, 0.0], 1, 1)
    sp.update_reflected(10.0, [-1.0, -2.0, -4.0], 1, 2)
    assert list(sp.np[:, 0]) == [1.0, 1.0]
    assert list(sp.flux_incident[:, 0]) == [3.5, 10.5] and list(sp.flux_reflected[:, 0]) == [-3.5, -10.0]
    assert abs(sp.force[0, 0, 0] - (3.5 * -9 - 3.5 * 9)) < 2 * EPS
    assert abs(sp.force[0, 1, 0] - (10.5 * 10 - 10.0 * -1)) < 2 * EPS
    assert abs(sp.force[1, 0, 0]) < 2 * EPS and abs(sp.force[1, 1, 0] - 20.0) < 2 * EPS
    assert sp.force[2, 0, 0] == 0.0 and abs(sp.force[2, 1, 0] - (10.5 * 3 + 40.0)) < 2 * EPS
    assert abs(sp.normal_pressure[0, 0] - (9 * 3.5 + 9 * 3.5)) < 2 * EPS
    assert abs(sp.normal_pressure[1, 0] - (10 * 10.5 + 1 * 10.0)) < 2 * EPS
    assert abs(sp.shear_pressure[1, 0, 0]) < 2 * EPS
    assert abs(sp.shear_pressure[1, 1, 0] - 20.0) < 2 * EPS
    assert abs(sp.shear_pressure[2, 1, 0] - (3 * 10.5 + 40.0)) < 2 * EPS
    assert abs(sp.kinetic_energy_flux[0, 0]) < 2 * EPS
    ke_r = 0.5 * 10.5 * 109 - 0.5 * 10.0 * 21
    assert abs(sp.kinetic_energy_flux[1, 0] - ke_r) < 2 * EPS

    before = sp.q.copy()
    sp.inv_areas[:] = [0.5, 0.25]
    dt = 1e-20
    sp.scale(1, M_AR, dt)
    f = M_AR * np.array([0.5, 0.25]) / dt
    assert list(sp.np[:, 0]) == [1.0, 1.0]   # np is not scaled
    for e in range(2):
        assert np.all(np.abs(sp.q[0, e, 1:] - f[e] * before[0, e, 1:]) <= 2 * EPS * np.abs(f[e] * before[0, e, 1:]))
    assert abs(sp.flux_incident[0, 0] - f[0] * 3.5) < 2 * EPS
    assert abs(sp.flux_reflected[1, 0] + f[1] * 10.0) < 2 * EPS
    sp.clear()
    assert not sp.q.any()


def test_surf_avg_kat(mzo):
    """test_surface_props_1D_uniform.jl:200-300: two avg_props! calls with n = 2."""
    inst, avg = mzo.Surf.grid1d(1), mzo.Surf.grid1d(1)
    first = {"np": (20, 10), "fi": (6, 0), "fr": (0, 11), "f": ((1, 3, -8), (4, 9, 5)), "pn": (11, 9),
             "s": ((4, 5, 6), (11, 9, 2)), "ke": (3, 7)}
    second = {"np": (10, 20), "fi": (5, 0), "fr": (0, 3), "f": ((2, 4, 8), (3, 1, 9)), "pn": (12, 6),
              "s": ((1, 9, 8), (3, 4, 8)), "ke": (2, 6)}

    def fill(d):
        inst.clear()
        for e in range(2):
            q = inst.q[0, e]
            q[0], q[1], q[2] = d["np"][e], d["fi"][e], d["fr"][e]
            q[3:6] = d["f"][e]
            q[6] = d["pn"][e]
            q[7:10] = d["s"][e]
            q[10] = d["ke"][e]

    fill(first)
    avg.avg_from(inst, 2)
    fill(second)
    avg.avg_from(inst, 2)
    a = avg.q[0]
    assert list(a[:, 0]) == [15.0, 15.0]
    assert a[0, 1] == 5.5 and a[1, 2] == 7.0
    assert list(a[0, 3:6]) == [1.5, 3.5, 0.0] and list(a[1, 3:6]) == [3.5, 5.0, 7.0]
    assert list(a[:, 6]) == [11.5, 7.5]
    assert list(a[0, 7:10]) == [2.5, 7.0, 7.0] and list(a[1, 7:10]) == [7.0, 6.5, 5.0]
    assert list(a[:, 10]) == [2.5, 6.5]


def _two_cell_case(mzo):
    """test_surface_props_1D_uniform.jl:312-350: grid L = 4 with 8 cells,
    specular walls; 4 particles at x = 0.1 moving at -1 (all reach the left
    wall within dt = 0.2), 4 at x = 3.9 of which one moves at +1."""
    parts = [(1.0, [-1.0, 0.0, 0.0], [0.1, 0.0, 0.0]) for _ in range(4)]
    parts += [(3.0, [1.0 if i == 5 else 0.0, 0.0, 3.0], [3.9, 0.0, 0.0]) for i in range(5, 9)]
    pv = make_pv(mzo, parts)
    pia = mzo.PIA(8, 1)
    pia.set(1, 1, 4, 1, 4, 4)
    pia.set(8, 1, 4, 5, 8, 4)
    pia.set_n_total(1, 8)
    return pv, pia


def test_convection_surf_kat(mzo):
    """test_surface_props_1D_uniform.jl:312-366 (contiguous) and 420-471
    (two particles of cell 1 deleted, PIA not contiguous)."""
    grid = mzo.grid1d(4.0, 8)
    wl = mzo.walls([M_AR], 1.0, 1.0, 0.0, 0.0, 0.0, 0.0)
    f = M_AR / 0.2
    for deleted in (0, 2):
        pv, pia = _two_cell_case(mzo)
        for _ in range(deleted):
            mzo.lib().mzo_delete_particle_end(pv.p, pia.p, 1, 1)
        if deleted:
            pia.set_contiguous(1, False)
        sp = mzo.Surf.grid1d(1)
        mzo.convect_particles_surf(mzo.RNG(seed=1234), grid, wl, pv, pia, 1, 0.2, sp, M_AR)
        n_left = 4 - deleted
        assert list(sp.np[:, 0]) == [float(n_left), 1.0]
        assert abs(sp.flux_incident[0, 0] / f - n_left) < 3 * EPS
        assert abs(sp.flux_incident[1, 0] / f - 3.0) < 3 * EPS
        assert abs(sp.flux_reflected[0, 0] / f + n_left) < 3 * EPS
        assert abs(sp.flux_reflected[1, 0] / f + 3.0) < 3 * EPS


def _random_case(mzo, n, seed, L=1e-3, nx=20):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, L, n)
    v = rng.normal(0.0, 400.0, (n, 3))
    w = rng.uniform(0.5, 2.0, n)
    parts = [(w[i], v[i], [x[i], 0.0, 0.0]) for i in range(n)]
    pv = make_pv(mzo, parts)
    grid = mzo.grid1d(L, nx)
    pia = mzo.PIA(nx, 1)
    pia.set(1, 1, n, 1, n, n)
    pia.set_n_total(1, n)
    mzo.lib().mzo_sort_particles(mzo.C.byref(grid), pv.p, pia.p, 1)
    return pv, pia, grid


def test_specular_walls_invariants(mzo):
    """Specular walls flip v_x only: no energy and no tangential momentum reach
    the walls, every contact returns its weight, and the normal pressure is
    twice the normal momentum of the incident particles."""
    pv, pia, grid = _random_case(mzo, 3000, 5)
    wl = mzo.walls([M_AR], 300.0, 300.0, 0.0, 0.0, 0.0, 0.0)
    dt = 4e-7     # a few percent of the particles reach a wall
    xs, vs, ws, _ = pv.logical_arrays(3000)
    sp = mzo.Surf.grid1d(1)
    mzo.convect_particles_surf(mzo.RNG(seed=3), grid, wl, pv, pia, 1, dt, sp, M_AR)
    f = M_AR / dt
    xe = xs[:, 0] + vs[:, 0] * dt
    hit_l, hit_r = xe <= 0.0, xe >= grid.L
    assert sp.np[0, 0] == hit_l.sum() > 0 and sp.np[1, 0] == hit_r.sum() > 0
    assert np.array_equal(sp.flux_reflected[:, 0], -sp.flux_incident[:, 0])
    assert np.all(sp.kinetic_energy_flux == 0.0)
    for e, hit in ((0, hit_l), (1, hit_r)):
        pn = 2.0 * np.sum(ws[hit] * np.abs(vs[hit, 0])) * f
        assert abs(sp.normal_pressure[e, 0] - pn) <= 1e-13 * pn
        assert abs(sp.flux_incident[e, 0] - f * ws[hit].sum()) <= 1e-13 * f * ws[hit].sum()
        scale = f * np.sum(ws[hit][:, None] * np.abs(vs[hit])) + 1e-300
        assert abs(sp.force[1, e, 0]) <= 1e-13 * scale and abs(sp.force[2, e, 0]) <= 1e-13 * scale
        assert abs(sp.shear_pressure[0, e, 0]) <= 1e-13 * scale


@pytest.mark.parametrize("acc", [1.0, 0.5])
def test_wall_force_is_momentum_lost(mzo, acc):
    """Diffuse (and half-diffuse) moving walls: the summed force and energy
    flux on both walls equal mass / dt times the momentum and kinetic energy
    the particles lose in the step (only wall contacts change velocities)."""
    n = 3000
    pv, pia, grid = _random_case(mzo, n, 11)
    wl = mzo.walls([M_AR], 250.0, 350.0, -300.0, 300.0, acc, acc)
    dt = 4e-7
    _, v0, w0, _ = pv.logical_arrays(n)
    sp = mzo.Surf.grid1d(1)
    mzo.convect_particles_surf(mzo.RNG(seed=7, tag=2), grid, wl, pv, pia, 1, dt, sp, M_AR)
    _, v1, w1, _ = pv.logical_arrays(n)
    assert np.array_equal(w0, w1)
    f = M_AR / dt
    dp = np.sum(w0[:, None] * (v0 - v1), axis=0) * f
    scale = np.sum(w0[:, None] * (np.abs(v0) + np.abs(v1)), axis=0) * f
    force = sp.force[:, 0, 0] + sp.force[:, 1, 0]
    assert np.all(np.abs(force - dp) <= 1e-12 * scale)
    de = 0.5 * np.sum(w0 * (np.sum(v0 * v0, axis=1) - np.sum(v1 * v1, axis=1))) * f
    escale = 0.5 * np.sum(w0 * (np.sum(v0 * v0, axis=1) + np.sum(v1 * v1, axis=1))) * f
    assert abs(sp.kinetic_energy_flux.sum() - de) <= 1e-12 * escale
    # each element's normal pressure is the normal momentum it takes, sign-fixed
    assert sp.normal_pressure[0, 0] > 0 and sp.normal_pressure[1, 0] > 0


# --------------------------------------------------------- package host helpers
def _pkg():
    import os
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import __graft_entry__ as ge
    return ge.load_package()


class _Sp:
    def __init__(self, mass):
        self.mass = mass


def test_package_surfprops_helpers_kat():
    """The package's SurfProps host helpers (shapes, scale, avg, reduce,
    clear) against test_surface_props_1D_uniform.jl's answers."""
    pkg = _pkg()
    sp = pkg.SurfProps.grid1d(1)
    assert sp.n_elements == 2 and sp.n_species == 1
    assert list(sp.normals[:, 0]) == [1.0, 0.0, 0.0] and list(sp.normals[:, 1]) == [-1.0, 0.0, 0.0]
    assert sp.np.shape == (2, 1) and sp.force.shape == (3, 2, 1)
    sp.flux_incident[:, 0] = [3.5, 10.5]
    sp.flux_reflected[:, 0] = [-3.5, -10.0]
    sp.np[:, 0] = [1.0, 1.0]
    sp.areas[:] = [2.0, 4.0]
    sp.inv_areas[:] = [0.5, 0.25]
    pkg.surface_props_scale(1, sp, [_Sp(M_AR)], 1e-20)
    f = M_AR * np.array([0.5, 0.25]) / 1e-20
    assert abs(sp.flux_incident[0, 0] - f[0] * 3.5) < 2 * EPS
    assert abs(sp.flux_reflected[1, 0] + f[1] * 10.0) < 2 * EPS
    assert list(sp.np[:, 0]) == [1.0, 1.0]
    pkg.clear_props(sp)
    assert not any(getattr(sp, k).any() for k in pkg.SurfProps.FIELDS)
    avg = pkg.SurfProps.grid1d(1)
    sp.np[:, 0] = [20, 10]
    pkg.avg_props(avg, sp, 2)
    sp.np[:, 0] = [10, 20]
    pkg.avg_props(avg, sp, 2)
    assert list(avg.np[:, 0]) == [15.0, 15.0]
    tot = pkg.SurfProps.grid1d(1)
    pkg.reduce_surf_props(tot, [avg, sp])
    assert list(tot.np[:, 0]) == [25.0, 35.0]
    # device layout round trip of the host copy
    sp.force[:, 1, 0] = [1.0, 2.0, 3.0]
    sp.shear_pressure[:, 0, 0] = [4.0, 5.0, 6.0]
    q = sp._to_device()
    back = pkg.SurfProps.grid1d(1)._from_device(q.ravel())
    for k in pkg.SurfProps.FIELDS:
        assert np.array_equal(getattr(back, k), getattr(sp, k))


def _equilibrium(n, L, T, seed):
    """n argon particles uniform in (0, L), Maxwellian at T, at rest."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, L, n)
    v = rng.normal(0.0, np.sqrt(1.380649e-23 * T / M_AR), (n, 3))
    return x, v


def equilibrium_expectations(n, L, T, w):
    """Kinetic theory of a gas at rest and in equilibrium with diffuse walls
    at its own temperature: pressure n k T on each wall, incident number flux
    n sqrt(k T / (2 pi m)), no net energy or tangential momentum exchange."""
    kb = 1.380649e-23
    ndens = n * w / L
    flux = ndens * np.sqrt(kb * T / (2 * np.pi * M_AR))
    return dict(p=ndens * kb * T, mass_flux=M_AR * flux, energy_flux_in=2 * kb * T * flux)


def check_equilibrium(avg, e, tol):
    """avg: time-averaged (np, flux_incident, flux_reflected, force[3],
    normal_pressure, shear[3], ke) per element, rows = elements."""
    for el, sgn in ((0, -1.0), (1, 1.0)):
        q = avg[el]
        assert abs(q[6] / e["p"] - 1) < tol, (el, q[6] / e["p"])
        assert abs(sgn * q[3] / e["p"] - 1) < tol, (el, q[3] / e["p"])
        assert abs(q[1] / e["mass_flux"] - 1) < tol, (el, q[1] / e["mass_flux"])
        assert q[2] == -q[1]
        assert abs(q[10]) < 3 * tol * e["energy_flux_in"], (el, q[10] / e["energy_flux_in"])
        for k in (4, 5, 8, 9):
            assert abs(q[k]) < 2 * tol * e["p"], (el, k, q[k] / e["p"])


def test_equilibrium_wall_pressure(mzo):
    """20,000 particles, 200 steps of convection between diffuse walls at the
    gas temperature: the time-averaged wall properties match kinetic theory
    (about 20,000 contacts per wall; 4 % is ten standard deviations)."""
    n, L, T, dt = 20000, 1e-3, 300.0, 5e-8
    w = 1e23 * L / n
    x, v = _equilibrium(n, L, T, 21)
    pv = make_pv(mzo, [(w, v[i], [x[i], 0.0, 0.0]) for i in range(n)])
    pia = mzo.PIA(10, 1)
    pia.set(1, 1, n, 1, n, n)
    pia.set_n_total(1, n)
    grid = mzo.grid1d(L, 10)
    mzo.lib().mzo_sort_particles(mzo.C.byref(grid), pv.p, pia.p, 1)
    wl = mzo.walls([M_AR], T, T, 0.0, 0.0, 1.0, 1.0)
    inst, avg = mzo.Surf.grid1d(1), mzo.Surf.grid1d(1)
    steps = 200
    for t in range(steps):
        mzo.convect_particles_surf(mzo.RNG(seed=9, tag=(t << 4) | 2), grid, wl, pv, pia, 1, dt, inst, M_AR)
        avg.avg_from(inst, steps)
    check_equilibrium(avg.q[0], equilibrium_expectations(n, L, T, w), 0.04)


def test_props_avg_kat(mzo):
    """avg_props! of PhysProps, test/test_computes.jl:82-116: the average of
    two PhysProps with n = 2, then clear; the package's host helper gives the
    same and refuses mixed ndens conventions."""
    nc, ns = 40, 3
    p3, p4, avg = mzo.Props(nc, ns), mzo.Props(nc, ns), mzo.Props(nc, ns)
    for p, T, n, npart, lpa, v in ((p3, 100.0, 2e11, 100, 1e7, 1e3), (p4, 200.0, 4e11, 200, 2e7, 2e3)):
        p.T[:], p.n[:], p.np[:], p.lpa[:], p.v[:] = T, n, npart, lpa, v
    mzo.lib().mzo_props_avg(avg.ref, p3.ref, 2.0)
    mzo.lib().mzo_props_avg(avg.ref, p4.ref, 2.0)
    assert np.max(np.abs(avg.T - 150.0)) < 2 * EPS and np.max(np.abs(avg.n - 3e11)) < 2 * EPS
    assert np.max(np.abs(avg.np - 150)) < 2 * EPS and np.max(np.abs(avg.lpa - 1.5e7)) < 2 * EPS
    assert np.max(np.abs(avg.v - 1.5e3)) < 2 * EPS

    pkg = _pkg()
    h3, h4, havg = pkg.PhysProps(nc, ns), pkg.PhysProps(nc, ns), pkg.PhysProps(nc, ns)
    for p, T, n, npart, lpa, v in ((h3, 100.0, 2e11, 100, 1e7, 1e3), (h4, 200.0, 4e11, 200, 2e7, 2e3)):
        p.T[:], p.n[:], p.np[:], p.lpa[:], p.v[:] = T, n, npart, lpa, v
    pkg.avg_props(havg, h3, 2)
    pkg.avg_props(havg, h4, 2)
    for k in ("T", "n", "np", "lpa", "v"):
        assert np.array_equal(getattr(havg, k), getattr(avg, k)), k
    with pytest.raises(pkg.MzError):
        pkg.avg_props(havg, pkg.PhysProps(nc, ns, ndens_not_Np=True), 2)
