"""The radix form of sort_particles! (grid_sorting.jl:58-182) that the
library runs for many cells: index, PIA and keys bit for bit against the
oracle's counting sort, at cell counts beyond the count-matrix form's bound
(14,336) and with one, two and three 8-bit digit passes; the count-matrix
form on the same input gives the same state; and the Couette step graph
with the radix sort forced equals the one with the count-matrix sort."""
import ctypes as C

import numpy as np
import pytest

import helpers
from test_gpu_parity import _store

pytestmark = pytest.mark.gpu


def _scattered(mzo, n, nc, seed, nearly_sorted=False):
    """n particles over [0, 1) in random logical order (or sorted by x with
    a jitter of two cells, the order a convection step leaves)."""
    rng = np.random.default_rng(seed)
    pv = mzo.ParticleVector(n)
    pia = mzo.PIA(nc, 1)
    for i in range(1, n + 1):
        pv.pop_to(i)
    if nearly_sorted:
        x = np.clip(np.sort(rng.uniform(0, 1.0, n)) + rng.uniform(-2.0, 2.0, n) / nc, 0.0, 1.0 - 1e-9)
        pv.x[:n, 0] = x
    else:
        pv.index[:n] = rng.permutation(n) + 1
        pv.x[:n, 0] = rng.uniform(0, 1.0, n)
    pv.x[:n, 1:] = rng.normal(size=(n, 2))
    pv.v[:n] = rng.normal(size=(n, 3))
    pv.w[:n] = rng.uniform(0.5, 2, n)
    pia.set(1, 1, n, 1, n, n)
    pia.set_n_total(1, n)
    return pv, pia


@pytest.mark.parametrize("n,nc,path,nearly", [(20000, 200, "radix", False), (20000, 700, "radix", True),
                                              (20000, 700, "matrix", True), (200000, 40000, "auto", False),
                                              (120000, 70000, "auto", True)])
def test_radix_sort_vs_oracle(hip, mzo, n, nc, path, nearly):
    pv, pia = _scattered(mzo, n, nc, seed=nc, nearly_sorted=nearly)
    grid = mzo.grid1d(1.0, nc)
    st = _store(hip, mzo, pv, pia, nc, cap=n + 5000)
    st.set_grid(hip.Grid1DUniform(1.0, nc))
    st.set_sort_path(path)
    hip.sort_particles(st, 1)
    mzo.lib().mzo_sort_particles(C.byref(grid), pv.p, pia.p, 1)
    got = st.download_species(1)
    assert got["err"] == 0 and got["contiguous"]
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=0.0)
    assert np.array_equal(got["cell"], pv.cell[:n])
    # a second sort on the stored keys (sort_particles!(gridsort, pv, pia, species)),
    # which the reference leaves in the pre-sort order
    hip.sort_particles(st, 1, compute_cells=False)
    mzo.lib().mzo_sort_particles_precomputed(pv.p, pia.p, 1)
    got = st.download_species(1)
    helpers.assert_state_matches(mzo, got, pv, pia, rtol=0.0)
    st.close()


def test_radix_and_matrix_sorts_agree_with_empty_cells(hip, mzo):
    """3,000 cells of which every third is empty, 30,000 particles: both
    forms leave the same index, data and PIA."""
    rng = np.random.default_rng(5)
    n, nc = 30000, 3000
    pv = mzo.ParticleVector(n)
    pia = mzo.PIA(nc, 1)
    for i in range(1, n + 1):
        pv.pop_to(i)
    cells = rng.choice(np.arange(nc)[np.arange(nc) % 3 != 0], size=n)
    pv.x[:n, 0] = (cells + rng.uniform(0.01, 0.99, n)) / nc
    pv.v[:n] = rng.normal(size=(n, 3))
    pv.w[:n] = 1.0
    pia.set(1, 1, n, 1, n, n)
    pia.set_n_total(1, n)
    outs = []
    for path in ("matrix", "radix"):
        st = _store(hip, mzo, pv, pia, nc, cap=n + 1024)
        st.set_grid(hip.Grid1DUniform(1.0, nc))
        st.set_sort_path(path)
        hip.sort_particles(st, 1)
        outs.append(st.download_species(1))
        st.close()
    a, b = outs
    for k in ("index", "pia", "x", "v", "w", "cell"):
        assert np.array_equal(a[k], b[k]), k
    assert int((a["pia"][:, 0] == 0).sum()) == nc // 3


def test_couette_graph_radix_equals_matrix(hip, mzo):
    """Twelve Couette steps on the step graph with each sort form: the same
    particles and props bit for bit."""
    res = []
    for path in ("matrix", "radix"):
        case = helpers.couette_case(mzo, n_cells=64, ppc=60, seed=9, dt=2e-9, sgwm_mult=50.0)
        st = helpers.store_from_case(hip, case)
        st.set_sort_path(path)
        hip.couette_steps(st, 12, 1, case["dt"], equal_weight=True, props_every=1, graph=True)
        got = st.download_species(1)
        assert got["err"] == 0
        res.append((got, hip.PhysProps(case["n_cells"], 1).fetch(st)))
        st.close()
    (a, pa), (b, pb) = res
    for k in ("index", "pia", "x", "v", "w"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(pa.T, pb.T) and np.array_equal(pa.n, pb.n)
