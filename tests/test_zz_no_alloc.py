"""mz_sort_noalloc, the sort whose scratch is a fixed device array, against
mz_sort through the C-ABI: two calls in a row on one GPU, the second with a
smaller input, each leaving state bit for bit equal to mz_sort's (and to the
oracle's).  The second case starts non-contiguous, so the squash also runs
in the fixed array, over contents the first call left there."""
import ctypes as C

import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def _assert_same(a, b):
    da, db = a.download_species(1), b.download_species(1)
    assert da["err"] == 0 and db["err"] == 0
    for k in ("n_total", "n_pos", "contiguous", "nbuffer", "length"):
        assert da[k] == db[k], k
    n = da["n_total"]
    assert np.array_equal(da["pia"], db["pia"])
    for k in ("index", "cell", "x", "v", "w"):
        assert np.array_equal(da[k][:n], db[k][:n]), k
    return db


def _holey_case(mzo, nc, per_cell, seed):
    """A store layout with holes after deletions at cell ends, non-contiguous
    (the state a merge leaves before squash_pia!)."""
    rng = np.random.default_rng(seed)
    n = nc * per_cell
    pv = mzo.ParticleVector(n)
    pia = mzo.PIA(nc, 1)
    for i in range(1, n + 1):
        pv.pop_to(i)
        pv.set(i, 1.0 + i, rng.normal(size=3), [rng.uniform(0, float(nc)), 0.0, 0.0])
    for c in range(1, nc + 1):
        pia.set(c, 1, per_cell, per_cell * (c - 1) + 1, per_cell * c, per_cell)
    pia.set_n_total(1, n)
    for c, k in ((2, 9), (5, 23), (nc, 4)):
        for _ in range(k):
            mzo.lib().mzo_delete_particle_end(pv.p, pia.p, c, 1)
    pia.set_contiguous(1, False)
    return pv, pia


def test_sort_noalloc_equals_sort(hip, mzo):
    lib = hip.load_library()
    # call 1: 120 cells x 150 particles after one convection step
    big = helpers.couette_case(mzo, n_cells=120, ppc=150, seed=31, dt=2e-9)
    a = helpers.store_from_case(hip, big)
    b = helpers.store_from_case(hip, big)
    for st in (a, b):
        hip.convect_particles(st, 1, big["dt"], tag=(1 << 4) | 2)
    hip.sort_particles(a, 1)
    hip._check(lib.mz_sort_noalloc(b.h, 1, 1))
    _assert_same(a, b)
    r = mzo.RNG(seed=big["seed"], tag=(1 << 4) | 2)
    mzo.lib().mzo_convect_particles(r.ref, C.byref(big["grid"]), C.byref(big["walls"]), big["pv"].p,
                                    big["pia"].p, 1, big["dt"])
    mzo.lib().mzo_sort_particles(C.byref(big["grid"]), big["pv"].p, big["pia"].p, 1)
    helpers.assert_state_matches(mzo, b.download_species(1), big["pv"], big["pia"], rtol=1e-12)
    a.close()
    b.close()

    # call 2: a smaller, non-contiguous store (squash, then sort)
    nc = 9
    pv, pia = _holey_case(mzo, nc, 70, seed=5)
    stores = []
    for _ in range(2):
        st = hip.DeviceStore(nc, 1, 2 * pv.length + 64)
        helpers.upload_from_oracle(st, mzo, pv, pia)
        st.set_grid(hip.Grid1DUniform(float(nc), nc))
        stores.append(st)
    a, b = stores
    hip.sort_particles(a, 1)
    hip._check(lib.mz_sort_noalloc(b.h, 1, 1))
    _assert_same(a, b)
    mzo.lib().mzo_squash_pia(pv.p, pia.p, 1)
    mzo.lib().mzo_sort_particles(C.byref(mzo.grid1d(float(nc), nc)), pv.p, pia.p, 1)
    helpers.assert_state_matches(mzo, b.download_species(1), pv, pia, rtol=0.0)
    a.close()
    b.close()


def test_sort_noalloc_rejects_oversize(hip):
    """A count matrix beyond the fixed array is refused before any launch."""
    st = hip.DeviceStore(2000, 1, 1200 * 1024)
    st.set_grid(hip.Grid1DUniform(1.0, 2000))
    rc = hip.load_library().mz_sort_noalloc(st.h, 1, 1)
    assert rc != 0
    assert b"sort_noalloc" in hip.load_library().mz_last_error()
    st.close()
