"""Flux properties (compute_flux_props!, src/properties/flux_props.jl): the
oracle pinned to the configuration of test/test_flux_computes.jl with its
values worked out by hand, and to symmetric particle sets whose odd fluxes
vanish exactly.  The reference's own asserts compare values of order 1e-24
at an absolute eps, so they hold for any result, and its expected vectors
carry a 1/N factor the code does not apply; the hand values here are the
code's.  CPU only.
"""
import numpy as np

from test_oracle_kat import make_pv

M_AR = 66.3e-27


def _kat(mzo):
    """test_flux_computes.jl:3-56: cell 1 holds 4 particles (two in group 1,
    two in group 2), cell 6 two; Grid1DUniform(4.0, 8), so V = 0.5."""
    parts = [(1.0, [vx, vy, 0.0], [0.3, 0.0, 0.0]) for vx, vy in zip([1.0, -1.0, -2.0, 6.0], [-3.0, -4.0, 3.0, 4.0])]
    parts += [(w, [vx, 0.0, vz], [2.8, 0.0, 0.0]) for w, vx, vz in zip([2.0, 8.0], [8.0, -2.0], [2.0, -2.0])]
    pv = make_pv(mzo, parts)
    pia = mzo.PIA(8, 1)
    pia.set(1, 1, 4, 1, 2, 2, 3, 4, 2)
    pia.set(6, 1, 2, 5, 6, 2)
    pia.set_n_total(1, 6)
    return pv, pia


def test_flux_kat(mzo):
    pv, pia = _kat(mzo)
    pp = mzo.Props(8, 1)
    mzo.compute_props([pv], pia, [M_AR], pp)
    f = mzo.compute_flux_props([pv], pia, [M_AR], pp, np.full(8, 2.0))
    k = M_AR * 2.0
    # cell 1: C = (0, -3), (-2, -4), (-3, 3), (5, 4); |C|^2 = 9, 20, 18, 41
    assert np.allclose(f[0, 0, 0:3], 0.5 * np.array([111.0, 111.0, 0.0]) * k, rtol=1e-15, atol=0)
    assert np.allclose(f[0, 0, 3:6], np.array([38.0, 50.0, 0.0]) * k, rtol=1e-15, atol=0)
    assert np.allclose(f[0, 0, 6:9], np.array([19.0, 0.0, 0.0]) * k, rtol=1e-15, atol=0)
    # cell 6: mean (0, 0, -1.2); C = (8, 0, 3.2) w 2 and (-2, 0, -0.8) w 8
    assert np.allclose(f[0, 5, 0:3], 0.5 * np.array([1113.6, 0.0, 445.44]) * k, rtol=1e-13, atol=0)
    assert np.allclose(f[0, 5, 3:6], np.array([160.0, 0.0, 25.6]) * k, rtol=1e-13, atol=0)
    assert np.allclose(f[0, 5, 6:9], np.array([0.0, 64.0, 0.0]) * k, rtol=1e-13, atol=1e-30)
    for c in (1, 2, 3, 4, 6, 7):
        assert np.all(f[0, c] == 0.0)
    # the cell_chunk form (3:7) computes cell 6 only
    g = mzo.compute_flux_props([pv], pia, [M_AR], pp, np.full(8, 2.0), cells=(3, 7))
    assert np.array_equal(g[0, 5], f[0, 5]) and np.all(g[0, 0] == 0.0)


def test_flux_symmetric_sets(mzo):
    """Particles in +/- pairs about the mean: the energy flux and the
    off-diagonal momentum flux vanish exactly; the diagonal is the pressure
    tensor m / V sum w C_k^2."""
    rng = np.random.default_rng(4)
    base = rng.normal(0, 300.0, (50, 3))
    u = np.array([120.0, -40.0, 7.0])
    parts = [(1.5, u + b, [0.5, 0, 0]) for b in base] + [(1.5, u - b, [0.5, 0, 0]) for b in base]
    pv = make_pv(mzo, parts)
    pia = mzo.PIA(1, 1)
    pia.set(1, 1, 100, 1, 100, 100)
    pia.set_n_total(1, 100)
    pp = mzo.Props(1, 1)
    mzo.compute_props([pv], pia, [M_AR], pp)
    f = mzo.compute_flux_props([pv], pia, [M_AR], pp, [4.0])
    scale = M_AR * 4.0 * 1.5 * (base ** 2).sum() * 300.0
    assert np.all(np.abs(f[0, 0, 0:3]) < 1e-12 * scale)
    assert np.allclose(f[0, 0, 3:6], M_AR * 4.0 * 1.5 * 2 * (base ** 2).sum(0), rtol=1e-12)
    off = np.array([(base[:, 0] * base[:, 1]).sum(), (base[:, 0] * base[:, 2]).sum(), (base[:, 1] * base[:, 2]).sum()])
    assert np.allclose(f[0, 0, 6:9], M_AR * 4.0 * 1.5 * 2 * off, rtol=1e-10, atol=1e-12 * scale / 300.0)


def test_flux_host_clear_and_avg(mzo):
    """clear_props! and avg_props! of FluxProps (flux_props.jl:133-186) on
    host copies, with the reference test's values (test_flux_computes.jl:86-110)."""
    from test_oracle_surf import _pkg
    mz = _pkg()
    fp, avg = mz.FluxProps(8, 1), mz.FluxProps(8, 1)
    fp.kinetic_energy_flux[...] = 1.0
    fp.diagonal_momentum_flux[...] = 2.0
    fp.off_diagonal_momentum_flux[...] = 3.0
    mz.avg_props(avg, fp, 2)
    fp.kinetic_energy_flux[...] = 3.0
    fp.diagonal_momentum_flux[...] = 5.0
    fp.off_diagonal_momentum_flux[...] = -3.0
    mz.avg_props(avg, fp, 2)
    assert np.all(avg.kinetic_energy_flux == 2.0) and np.all(avg.diagonal_momentum_flux == 3.5)
    assert np.all(avg.off_diagonal_momentum_flux == 0.0)
    mz.clear_props(fp)
    assert not fp.kinetic_energy_flux.any() and not fp.off_diagonal_momentum_flux.any()


def test_mean_free_path_and_collision_frequency():
    """mean_free_path and mean_collision_frequency (collisional_props.jl):
    at T = Tref the hard-sphere values 1 / (sqrt(2) pi d^2 n) and n pi d^2
    <g> with <g> = sqrt(16 k T / (pi m)); their ratio is the mean speed
    sqrt(8 k T / (pi m)); and the VHS temperature exponents."""
    from test_oracle_surf import _pkg
    mz = _pkg()
    import os
    data = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "merzbild.jl_amd", "data")
    sp = mz.load_species_data(os.path.join(data, "particles.toml"), ["Ar"])
    it = mz.load_interaction_data(os.path.join(data, "vhs.toml"), sp)
    i0 = it[0][0]
    d, Tref, om, m, n, kb = i0.vhs_d, i0.vhs_Tref, i0.vhs_o, sp[0].mass, 1e23, 1.380649e-23
    lam = mz.mean_free_path(it, 1, n, Tref)
    nu = mz.mean_collision_frequency(it, sp, 1, n, Tref)
    assert abs(lam * np.sqrt(2) * np.pi * d * d * n - 1) < 1e-14
    assert abs(nu / (n * np.pi * d * d * np.sqrt(16 * kb * Tref / (np.pi * m))) - 1) < 1e-14
    assert abs(nu * lam / np.sqrt(8 * kb * Tref / (np.pi * m)) - 1) < 1e-14
    assert abs(mz.mean_free_path(it, 1, n, 4 * Tref) / lam - 4.0 ** (om - 0.5)) < 1e-13
    assert abs(mz.mean_collision_frequency(it, sp, 1, n, 4 * Tref) / nu - 4.0 ** (1 - om)) < 1e-13
