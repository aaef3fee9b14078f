"""The analytic 0-D cases on the device: two-gas relaxation (ntc! for Ar-Ar,
He-Ar and He-He, then compute_props!, 800 steps) and the BKW moment
evolution (ntc! with pseudo-Maxwell molecules, then
compute_props_with_total_moments!, 500 steps).  The device trajectory equals
the oracle's (same Philox streams) to 1e-9 and satisfies the same analytic
bounds as the reference's tests."""
import numpy as np
import pytest

import analytic_cases as ac
import helpers

pytestmark = pytest.mark.gpu


def _species(hip):
    return [hip.Species("Ar", ac.M_AR), hip.Species("He", ac.M_HE)]


def _set_interactions(hip, st, its):
    n = len(its)
    st.set_interactions([[type("I", (), dict(mu1=its[i][k].mu1, mu2=its[i][k].mu2, vhs_factor=its[i][k].vhs_factor,
                                             vhs_o=its[i][k].vhs_o))() for k in range(n)] for i in range(n)])


def test_props_moments_vs_oracle(hip, mzo):
    """compute_props_with_total_moments! over groups 1 and 2, with an empty
    cell (its moments are 0/0 in the reference as in the oracle)."""
    rng = np.random.default_rng(17)
    nc, n = 5, 1500
    pv = mzo.ParticleVector(n)
    for i in range(1, n + 1):
        pv.pop_to(i)
        pv.set(i, rng.uniform(0.5, 2.0) * 1e12, rng.normal([50.0, -20.0, 5.0], 400.0), [0, 0, 0])
    pia = mzo.PIA(nc, 1)
    pia.set(1, 1, 300, 1, 300, 300)
    pia.set(2, 1, 500, 301, 700, 400, 1401, 1500, 100)
    pia.set(3, 1, 0, 0, -1, 0)
    pia.set(4, 1, 700, 701, 1400, 700)
    pia.set(5, 1, 0, 0, -1, 0)
    pia.set_n_total(1, n)
    st = hip.DeviceStore(nc, 1, n + 256)
    helpers.upload_from_oracle(st, mzo, pv, pia)
    powers = (4, 6, 8, 10, 3)
    pp = hip.PhysProps(nc, 1, powers, Tref=273.0)
    hip.compute_props_with_total_moments(st, [hip.Species("Ar", ac.M_AR)], pp)
    ref = mzo.Props(nc, 1, moment_powers=powers, Tref=273.0)
    mzo.compute_props_with_total_moments([pv], pia, [ac.M_AR], ref)
    full = [0, 1, 3]
    assert np.array_equal(pp.np, ref.np)
    assert np.allclose(pp.n[0, full], ref.n[0, full], rtol=1e-13)
    assert np.allclose(pp.T[0, full], ref.T[0, full], rtol=1e-12)
    assert np.allclose(pp.moments[0, full], ref.moments[0, full], rtol=1e-12)
    assert np.all(np.isnan(pp.moments[0, [2, 4]])) and np.all(np.isnan(ref.moments[0, [2, 4]]))
    st.close()


def test_two_gas_relaxation_on_device(hip, mzo):
    case = ac.two_gas_case(mzo)
    st = hip.DeviceStore(1, 2, [2 * ac.TWO_GAS["n"][0], 2 * ac.TWO_GAS["n"][1]])
    for s in (1, 2):
        helpers.upload_from_oracle(st, mzo, case["pvs"][s - 1], case["pia"], species=s)
    _set_interactions(hip, st, case["its"])
    for (i, k), cf in case["cfs"].items():
        st.set_sigma_g_w_max(i, k, [cf.sigma_g_w_max])
    st.set_seed(case["seed"])
    sp = _species(hip)
    pp = hip.PhysProps(1, 2)
    hip.compute_props(st, sp, pp)
    counts = np.array(ac.TWO_GAS["n"], dtype=float)
    E0 = ac.total_energy(pp.n[:, 0], pp.v[:, 0], pp.T[:, 0], case["masses"])
    T_mix0 = pp.T[:, 0] @ counts / counts.sum()
    T_dev, T_ref = [], []
    ref = mzo.Props(1, 2)
    for step in range(1, ac.TWO_GAS["steps"] + 1):
        for (s1, s2) in ac.TWO_GAS_PAIRS:
            hip.ntc(st, s1, s2, ac.TWO_GAS["dt"], V=[ac.TWO_GAS["V"]], tag=ac.two_gas_tag(step, s1, s2))
        hip.compute_props(st, sp, pp)
        T_dev.append(pp.T[:, 0].copy())
        assert list(pp.np[:, 0]) == list(ac.TWO_GAS["n"])
        E = ac.total_energy(pp.n[:, 0], pp.v[:, 0], pp.T[:, 0], case["masses"])
        assert abs(E / E0 - 1) < 1e-11
        ac.two_gas_oracle_step(mzo, case, step)
        mzo.compute_props(case["pvs"], case["pia"], list(case["masses"]), ref)
        T_ref.append(ref.T[:, 0].copy())
    T_dev, T_ref = np.array(T_dev), np.array(T_ref)
    assert np.max(np.abs(T_dev - T_ref) / T_ref) < 1e-9
    T_eq = case["T_eq"]
    assert np.all(np.abs(T_dev[-1] - T_eq) / T_eq < 0.12)
    assert np.max(np.abs(T_dev @ counts / counts.sum() - T_mix0)) / T_mix0 < 0.002
    st.close()


def test_bkw_moments_on_device(hip, mzo):
    case = ac.bkw_case(mzo)
    n = ac.BKW["n"]
    st = hip.DeviceStore(1, 1, n + 256)
    helpers.upload_from_oracle(st, mzo, case["pv"], case["pia"])
    its = mzo.interactions(["Ar"], table=mzo.PSEUDO_MAXWELL)
    _set_interactions(hip, st, its)
    st.set_sigma_g_w_max(1, 1, [case["cf"].sigma_g_w_max])
    st.set_seed(case["seed"])
    sp = [hip.Species("Ar", ac.M_AR)]
    pp = hip.PhysProps(1, 1, ac.BKW["moments"], Tref=ac.BKW["T0"])
    ref = mzo.Props(1, 1, moment_powers=ac.BKW["moments"], Tref=ac.BKW["T0"])
    hip.compute_props_with_total_moments(st, sp, pp)
    mom = [pp.moments[0, 0, :3].copy()]
    import ctypes as C
    for step in range(1, ac.BKW["steps"] + 1):
        tag = (step << 4) | 1
        hip.ntc(st, 1, 1, case["dt"], V=[1.0], tag=tag)
        hip.compute_props_with_total_moments(st, sp, pp)
        mom.append(pp.moments[0, 0, :3].copy())
        rr = mzo.RNG(seed=case["seed"], tag=tag)
        mzo.lib().mzo_ntc_single(rr.ref, C.byref(case["cf"]), C.byref(case["it"]), case["pv"].p, case["pia"].p, 1,
                                 1, case["dt"], 1.0, 1e-16, 0)
        if step % 50 == 0:
            mzo.compute_props_with_total_moments([case["pv"]], case["pia"], [ac.M_AR], ref)
            assert np.allclose(pp.moments[0, 0], ref.moments[0, 0], rtol=1e-9), step
            assert abs(pp.T[0, 0] / ref.T[0, 0] - 1) < 1e-9
    mom = np.array(mom)
    t = np.arange(ac.BKW["steps"] + 1) * ac.BKW["dt_scaled"]
    for j, p in enumerate((4, 6, 8)):
        a = ac.bkw_analytic(t, case["magic"], p)
        assert np.max(np.abs(a - mom[:, j]) / a) < ac.BKW_TOL[p]
    cf = st.collision_factors(1, 1)
    assert cf["n_coll_performed"][0] > 0
    st.close()
