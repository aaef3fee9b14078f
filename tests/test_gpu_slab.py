"""The slab-resident Couette step (one launch for all steps, each
slab of cells in LDS) against the oracle and against the step-by-step graph:
the index, the PIA, pv.cell, the collision factors and the properties agree
bit for bit with the graph path, and with the oracle as the other Couette
tests require."""
import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def _state(hip, st, nc):
    d = st.download_species(1)
    pp = hip.PhysProps(nc, 1).fetch(st)
    cf = st.collision_factors(1, 1)
    return d, pp, cf


def _assert_same(a, b):
    da, pa, ca = a
    db, pb, cb = b
    assert da["err"] == 0 and db["err"] == 0
    for k in ("n_total", "n_pos", "contiguous", "nbuffer", "length"):
        assert da[k] == db[k], k
    n = da["n_total"]
    assert np.array_equal(da["pia"], db["pia"])
    assert np.array_equal(da["index"][:n], db["index"][:n])
    assert np.array_equal(da["cell"][:n], db["cell"][:n])
    for k in ("x", "v", "w"):
        assert np.array_equal(da[k][:n], db[k][:n]), k
    for k in ("np", "n", "v", "T"):
        assert np.array_equal(getattr(pa, k), getattr(pb, k)), k
    for k in ca:
        assert np.array_equal(np.asarray(ca[k]), np.asarray(cb[k])), k


@pytest.mark.parametrize("shape", [(0, 0), (3, 2), (1, 1), (6, 4)])
def test_slab_couette_vs_oracle(hip, mzo, shape):
    """Thirty Couette steps (CFL 0.17) on the slab kernel against the oracle:
    index and PIA bit-exact, velocities and T to 1e-11 (as the graph path).
    Slabs of up to four cells collide each cell on a whole wavefront, larger
    ones (6 cells here) one lane per cell."""
    case = helpers.couette_case(mzo, n_cells=20, ppc=100, seed=3, dt=2e-10, sgwm_mult=50.0)
    st = helpers.store_from_case(hip, case)
    st.set_slab_shape(*shape)
    path = hip.couette_steps(st, 30, 1, case["dt"], equal_weight=True, props_every=1, mode="slab")
    assert path == "slab", st.slab_info()
    info = st.slab_info()
    assert info[3] == 1
    if shape[0]:
        assert info[1] == shape[0] and info[0] == shape[1]
    for step in range(1, 31):
        helpers.oracle_couette_step(mzo, case, step, equal_weight=True)
    got = st.download_species(1)
    assert got["err"] == 0
    helpers.assert_state_matches(mzo, got, case["pv"], case["pia"], rtol=1e-11)
    pp = hip.PhysProps(case["n_cells"], 1).fetch(st)
    assert np.allclose(pp.T[0], helpers.oracle_props_sorted(mzo, case), rtol=1e-11)
    cf = st.collision_factors(1, 1)
    assert sum(cf["n_coll_performed"]) > 0
    st.close()


def test_slab_couette_full_size_vs_oracle(hip, mzo):
    """The benchmarked configuration itself, 500 cells x 200 particles
    (test/test_1D_couette.jl's 100k-particle flow, equal weights, the
    collision call the bench runs), 20 steps in one slab launch with the
    library's shape, against the
    oracle: index and PIA bit-exact, positions, velocities and weights to
    1e-11, the sorted properties' T to 1e-11."""
    case = helpers.couette_case(mzo, n_cells=500, ppc=200, seed=5, dt=2e-10)
    st = helpers.store_from_case(hip, case)
    path = hip.couette_steps(st, 20, 1, case["dt"], props_every=1, mode="slab")
    assert path == "slab", st.slab_info()
    assert st.slab_info()[3] == 1
    for step in range(1, 21):
        helpers.oracle_couette_step(mzo, case, step, equal_weight=True)
    got = st.download_species(1)
    assert got["err"] == 0
    assert got["n_total"] == case["pia"].n_total()
    helpers.assert_state_matches(mzo, got, case["pv"], case["pia"], rtol=1e-11)
    pp = hip.PhysProps(case["n_cells"], 1).fetch(st)
    assert np.allclose(pp.T[0], helpers.oracle_props_sorted(mzo, case), rtol=1e-11)
    cf = st.collision_factors(1, 1)
    assert sum(cf["n_coll_performed"]) > 0
    st.close()


# a requested shape runs only with room for its largest slab to grow by 25%
# plus four Poisson deviations: (2, 4) holds 600 particles in 1,024 places
@pytest.mark.parametrize("n_cells,ppc,steps,shape", [(500, 200, 3000, (0, 0)), (97, 300, 400, (2, 4)),
                                                     (64, 50, 800, (2, 1))])
def test_slab_equals_graph(hip, mzo, n_cells, ppc, steps, shape):
    """The slab kernel and the captured step graph leave identical state,
    over enough steps for neighbouring slabs to drift a step apart many
    times."""
    case = helpers.couette_case(mzo, n_cells=n_cells, ppc=ppc, seed=21, dt=2e-10)
    a = helpers.store_from_case(hip, case)
    b = helpers.store_from_case(hip, case)
    assert hip.couette_steps(a, steps, 7, case["dt"], props_every=1, mode="graph") == "graph"
    b.set_slab_shape(*shape)
    assert hip.couette_steps(b, steps, 7, case["dt"], props_every=1, mode="slab") == "slab", b.slab_info()
    _assert_same(_state(hip, a, n_cells), _state(hip, b, n_cells))
    # a second call continues from the state the first left
    assert hip.couette_steps(a, 5, 7 + steps, case["dt"], props_every=1, mode="graph") == "graph"
    assert hip.couette_steps(b, 5, 7 + steps, case["dt"], props_every=1, mode="slab") == "slab"
    _assert_same(_state(hip, a, n_cells), _state(hip, b, n_cells))
    a.close()
    b.close()


def test_slab_falls_back_on_long_moves(hip, mzo):
    """With a time step that carries particles across more than one slab the
    launch aborts, the saved state is restored and the graph runs the same
    steps: the result equals the graph path's."""
    case = helpers.couette_case(mzo, n_cells=24, ppc=40, seed=8, dt=1.2e-8)
    a = helpers.store_from_case(hip, case)
    b = helpers.store_from_case(hip, case)
    assert hip.couette_steps(a, 4, 1, case["dt"], props_every=1, mode="graph") == "graph"
    b.set_slab_shape(1, 1)
    assert hip.couette_steps(b, 4, 1, case["dt"], props_every=1, mode="slab") == "graph"
    assert b.slab_info()[3] == -1
    _assert_same(_state(hip, a, 24), _state(hip, b, 24))
    a.close()
    b.close()
