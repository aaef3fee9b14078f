"""Analytic checks of the oracle's NTC + props path (CPU): the two-gas
relaxation to the mixture temperature (test/test_2species.jl) and the BKW
moment evolution of pseudo-Maxwell molecules (test/test_bkw.jl).  These pin
the collision and property code against closed forms rather than against a
restatement of the same formulas."""
import ctypes as C

import numpy as np

import analytic_cases as ac


def test_two_gas_relaxes_to_mixture_temperature(mzo):
    """Both species end within 12% of T_eq = 600 K (the reference's bound,
    test_2species.jl:104-106) after relaxing monotonically; the mixture
    temperature keeps its sampled value (600 K up to sampling noise); particle
    counts and densities never change;
    the total energy is conserved at every step."""
    case = ac.two_gas_case(mzo)
    pvs, pia, masses = case["pvs"], case["pia"], case["masses"]
    pp = mzo.Props(1, 2)
    mzo.compute_props(pvs, pia, list(masses), pp)
    E0 = ac.total_energy(pp.n[:, 0], pp.v[:, 0], pp.T[:, 0], masses)
    counts = np.array(ac.TWO_GAS["n"], dtype=float)
    T_mix0 = pp.T[:, 0] @ counts / counts.sum()      # the sampled state's mixture temperature
    T_hist = []
    for step in range(1, ac.TWO_GAS["steps"] + 1):
        ac.two_gas_oracle_step(mzo, case, step)
        mzo.compute_props(pvs, pia, list(masses), pp)
        T_hist.append(pp.T[:, 0].copy())
        assert list(pp.np[:, 0]) == list(ac.TWO_GAS["n"])
        for s in range(2):
            n_s = ac.TWO_GAS["n"][s] * ac.TWO_GAS["Fnum"]
            assert abs(pp.n[s, 0] - n_s) / n_s < np.finfo(float).eps
        E = ac.total_energy(pp.n[:, 0], pp.v[:, 0], pp.T[:, 0], masses)
        assert abs(E / E0 - 1) < 1e-11
    T_hist = np.array(T_hist)
    T_eq = case["T_eq"]
    assert abs(T_eq - 600.0) < 1e-9
    assert np.all(np.abs(T_hist[-1] - T_eq) / T_eq < 0.12)
    T_mix = T_hist @ counts / counts.sum()
    assert abs(T_mix0 - T_eq) / T_eq < 0.02
    assert np.max(np.abs(T_mix - T_mix0)) / T_mix0 < 0.002
    # Ar cools and He heats, monotonically over 100-step windows
    win = T_hist.reshape(8, 100, 2).mean(axis=1)
    assert np.all(np.diff(win[:, 0]) < 0) and np.all(np.diff(win[:, 1]) > 0)


def test_bkw_moments_follow_analytic_solution(mzo):
    """The normalised 4th, 6th and 8th total moments stay within 5%, 5.5% and
    15% of the BKW solution over 500 steps (test_bkw.jl:125-135)."""
    case = ac.bkw_case(mzo)
    pv, pia, it, cf = case["pv"], case["pia"], case["it"], case["cf"]
    pp = mzo.Props(1, 1, moment_powers=ac.BKW["moments"], Tref=ac.BKW["T0"])
    mzo.compute_props_with_total_moments([pv], pia, [ac.M_AR], pp)
    mom = [pp.moments[0, 0, :3].copy()]
    for step in range(1, ac.BKW["steps"] + 1):
        rr = mzo.RNG(seed=case["seed"], tag=(step << 4) | 1)
        mzo.lib().mzo_ntc_single(rr.ref, C.byref(cf), C.byref(it), pv.p, pia.p, 1, 1, case["dt"], 1.0, 1e-16, 0)
        mzo.compute_props_with_total_moments([pv], pia, [ac.M_AR], pp)
        mom.append(pp.moments[0, 0, :3].copy())
        assert pp.np[0, 0] == ac.BKW["n"]
    mom = np.array(mom)
    t = np.arange(ac.BKW["steps"] + 1) * ac.BKW["dt_scaled"]
    for j, p in enumerate((4, 6, 8)):
        ref = ac.bkw_analytic(t, case["magic"], p)
        assert np.max(np.abs(ref - mom[:, j]) / ref) < ac.BKW_TOL[p], (p, np.max(np.abs(ref - mom[:, j]) / ref))
    # the moments start at the BKW value (0.84 for the 4th) and relax towards 1
    assert abs(mom[0, 0] - 0.84) < 0.03 and abs(mom[-1, 0] - 1.0) < 0.05
