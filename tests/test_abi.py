"""CPU checks of the drop-in boundary: the library builds, loads, and exports
every entry point include/mzhip.h declares; the Python host mirrors the
reference's setup helpers.  No GPU calls."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "mzhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mz_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    import ctypes
    import __graft_entry__ as ge
    pkg = ge.load_package()
    if not os.path.exists(pkg.LIB_PATH):
        ge.build()
    lib = ctypes.CDLL(pkg.LIB_PATH)
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes signatures cover the same set
    assert sorted(pkg.EXPORTED_SYMBOLS) == names


def test_species_and_interaction_loading():
    """load_species_data / load_interaction_data restated from the reference's
    data files; the checks are test_collision_utils.jl:13-18, 55-62."""
    import __graft_entry__ as ge
    pkg = ge.load_package()
    data = os.path.join(ROOT, "merzbild.jl_amd", "data")
    sp = pkg.load_species_data(os.path.join(data, "particles.toml"), ["Ar", "He"])
    assert sp[0].mass == 66.3e-27 and sp[1].mass == 6.65e-27
    it = pkg.load_interaction_data(os.path.join(data, "vhs.toml"), sp)
    eps = np.finfo(float).eps
    assert abs(it[0][0].mu1 - 0.5) < eps and abs(it[1][1].mu2 - 0.5) < eps
    assert abs(it[0][1].mu1 - it[1][0].mu2) < eps
    assert abs(it[0][1].mu1 + it[0][1].mu2 - 1.0) < eps
    assert it[0][1].vhs_o == 0.735


def test_host_interaction_matches_oracle(mzo):
    """The host-side VHS factor equals the oracle's (same formula, same data)."""
    import __graft_entry__ as ge
    pkg = ge.load_package()
    for (a, b), (d, o, tr) in mzo.VHS.items():
        ho = pkg.Interaction(mzo.MASS[a], mzo.MASS[b], d, o, tr)
        oo, _ = mzo.interaction_pair(mzo.MASS[a], mzo.MASS[b], d, o, tr)
        assert abs(ho.vhs_factor / oo.vhs_factor - 1) < 4e-16
        assert ho.mu1 == oo.mu1 and ho.mu2 == oo.mu2
