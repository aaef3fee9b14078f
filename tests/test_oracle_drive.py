"""The oracle's multi-step Couette driver (the CPU baseline): its threaded
sort and per-cell collisions leave the state bit for bit equal to the
single-threaded run, which uses the serial sort_particles! restatement."""
import ctypes as C

import numpy as np

import helpers


def _run(mzo, nthreads, steps=12):
    case = helpers.couette_case(mzo, n_cells=40, ppc=50, seed=13, dt=2e-10, sgwm_mult=30.0)
    V = np.full(case["n_cells"], case["grid"].dx)
    pp = mzo.Props(case["n_cells"], 1)
    rng = mzo.RNG(seed=case["seed"])
    mzo.lib().mzo_couette_steps(rng.ref, C.byref(case["grid"]), C.byref(case["walls"]), C.byref(case["it"]),
                                case["cfs"], case["pv"].p, case["pia"].p, case["dt"],
                                V.ctypes.data_as(C.POINTER(C.c_double)), case["mass"], pp.ref, steps, 1, 1,
                                nthreads, 1)
    return case, pp


def test_threaded_driver_equals_serial(mzo):
    a, pa = _run(mzo, 1)
    b, pb = _run(mzo, 4)
    ua = helpers.oracle_upload_arrays(mzo, a["pv"], a["pia"])
    ub = helpers.oracle_upload_arrays(mzo, b["pv"], b["pia"])
    for k in ("pia", "ids", "x", "v", "w"):
        assert np.array_equal(ua[k], ub[k]), k
    assert np.array_equal(a["pv"].cell[:ua["n_total"]], b["pv"].cell[:ub["n_total"]])
    assert np.array_equal(pa.T, pb.T)
    # and the serial run equals the step-by-step oracle calls
    c = helpers.couette_case(mzo, n_cells=40, ppc=50, seed=13, dt=2e-10, sgwm_mult=30.0)
    for step in range(1, 13):
        helpers.oracle_couette_step(mzo, c, step, equal_weight=True)
    uc = helpers.oracle_upload_arrays(mzo, c["pv"], c["pia"])
    for k in ("pia", "ids", "x", "v", "w"):
        assert np.array_equal(ua[k], uc[k]), k
