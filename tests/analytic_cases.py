"""The two analytic 0-D cases of the reference's test suite, set up on the
oracle: two-gas relaxation (test/test_2species.jl) and the BKW solution for
pseudo-Maxwell molecules (test/test_bkw.jl).  Test infrastructure only (uses
the oracle)."""
import ctypes as C
import math

import numpy as np

KB = 1.380649e-23
M_AR, M_HE = 66.3e-27, 6.65e-27     # data/particles.toml

# test/test_2species.jl:13-22
TWO_GAS = dict(n=(400, 4000), T0=(3000.0, 360.0), Fnum=5e12, dt=2.5e-3, V=1.0, steps=800)
TWO_GAS_PAIRS = ((1, 1), (2, 1), (2, 2))          # the loop order of test_2species.jl:52-62

# test/test_bkw.jl:47-60
BKW = dict(n=20000, T0=273.0, ndens=1e23, dt_scaled=0.025, steps=500, moments=(4, 6, 8, 10))


def two_gas_case(mzo, seed=1234):
    """Ar (400, 3000 K) and He (4000, 360 K) Maxwellian in one cell of V = 1."""
    pvs, pia = [mzo.ParticleVector(TWO_GAS["n"][0]), mzo.ParticleVector(TWO_GAS["n"][1])], mzo.PIA(1, 2)
    r = mzo.RNG(seed=seed, tag=99)
    for s, (n, m, T) in enumerate(zip(TWO_GAS["n"], (M_AR, M_HE), TWO_GAS["T0"]), start=1):
        r.s.stream, r.s.k = s, 0
        mzo.lib().mzo_sample_particles_equal_weight(r.ref, pvs[s - 1].p, pia.p, 1, s, n, m, T, TWO_GAS["Fnum"],
                                                    0, 1, 0, 1, 0, 1)
    its = mzo.interactions(["Ar", "He"])
    cfs = {}
    masses = (M_AR, M_HE)
    for (i, k) in ((1, 1), (2, 1), (2, 2), (1, 2)):
        cf = mzo.CF()
        # estimate_sigma_g_w_max! with the initial temperatures (collision_utils.jl:440-504)
        cf.sigma_g_w_max = mzo.lib().mzo_estimate_sigma_g_w_max(C.byref(its[i - 1][k - 1]), masses[i - 1],
                                                                masses[k - 1], TWO_GAS["T0"][i - 1],
                                                                TWO_GAS["T0"][k - 1], TWO_GAS["Fnum"], 1.0)
        cfs[(i, k)] = cf
    T_eq = sum(n * T for n, T in zip(TWO_GAS["n"], TWO_GAS["T0"])) / sum(TWO_GAS["n"])
    return dict(pvs=pvs, pia=pia, its=its, cfs=cfs, masses=masses, T_eq=T_eq, seed=seed)


def two_gas_tag(step, s1, s2):
    return (step << 4) | (s1 * 4 + s2)


def two_gas_oracle_step(mzo, case, step):
    pvs, pia, its, cfs = case["pvs"], case["pia"], case["its"], case["cfs"]
    for (s1, s2) in TWO_GAS_PAIRS:
        rr = mzo.RNG(seed=case["seed"], tag=two_gas_tag(step, s1, s2))
        if s1 == s2:
            mzo.lib().mzo_ntc_single(rr.ref, C.byref(cfs[(s1, s1)]), C.byref(its[s1 - 1][s1 - 1]), pvs[s1 - 1].p,
                                     pia.p, 1, s1, TWO_GAS["dt"], TWO_GAS["V"], 1e-16, 0)
        else:
            mzo.lib().mzo_ntc_two(rr.ref, C.byref(cfs[(s1, s2)]), C.byref(its[s1 - 1][s2 - 1]), pvs[s1 - 1].p,
                                  pvs[s2 - 1].p, pia.p, 1, s1, s2, TWO_GAS["dt"], TWO_GAS["V"], 1e-16, 0)


def total_energy(n, v, T, masses):
    """sum over species of n (m |vbar|^2 / 2 + 3 k T / 2), per cell 1."""
    return sum(n[s] * (0.5 * masses[s] * float(np.dot(v[s], v[s])) + 1.5 * KB * T[s]) for s in range(len(masses)))


def bkw_case(mzo, seed=1234):
    """20000 Ar particles, BKW velocities at t = 0 (distributions_and_sampling.jl:195-213, with its
    polar angle uniform on [0, pi]), pseudo-Maxwell interaction, one cell of V = 1."""
    n, T0 = BKW["n"], BKW["T0"]
    its = mzo.interactions(["Ar"], table=mzo.PSEUDO_MAXWELL)
    it = its[0][0]
    Fnum = BKW["ndens"] / n
    pv, pia = mzo.ParticleVector(n), mzo.PIA(1, 1)
    r = mzo.RNG(seed=seed, tag=77)
    mzo.lib().mzo_sample_particles_equal_weight(r.ref, pv.p, pia.p, 1, 1, n, M_AR, T0, Fnum, 0, 1, 0, 1, 0, 1)
    r.s.stream, r.s.k = 1, 0
    v0 = (C.c_double * 3)(0.0, 0.0, 0.0)
    mzo.lib().mzo_sample_bkw(r.ref, pv.p, n, 0, M_AR, T0, v0)
    d = mzo.PSEUDO_MAXWELL[("Ar", "Ar")][0]
    sigma_ref = math.pi * d * d
    vref = math.sqrt(2 * KB * T0 / M_AR)
    Lref = 1.0 / (BKW["ndens"] * sigma_ref)
    tref = Lref / vref
    # test/test_bkw.jl:60-62
    kappa_mult = sigma_ref * (it.m_r / (2 * KB * T0)) ** (-0.5) / math.gamma(2.5 - 1.0)
    ttt_bkw = 1.0 / (4 * math.pi * BKW["ndens"] * kappa_mult)
    magic = tref / ttt_bkw / (4 * math.pi)
    cf = mzo.CF()
    cf.sigma_g_w_max = mzo.lib().mzo_estimate_sigma_g_w_max(C.byref(it), M_AR, M_AR, T0, T0, Fnum, 1.0)
    return dict(pv=pv, pia=pia, it=it, cf=cf, dt=BKW["dt_scaled"] * tref, magic=magic, Fnum=Fnum, seed=seed)


def bkw_analytic(scaled_time, magic, N):
    """The normalised N-th total moment of the BKW solution (test/test_bkw.jl:23-27)."""
    Cc = 1.0 - 0.4 * np.exp(-np.asarray(scaled_time) * magic / 6)
    kk = N / 2
    return Cc ** (kk - 1) * (kk - (kk - 1) * Cc)


BKW_TOL = {4: 0.05, 6: 0.055, 8: 0.15}            # test/test_bkw.jl:125-135
