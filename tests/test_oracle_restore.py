"""The oracle's restore_particle_ordering! against the reference's known
answers (test/test_particle_index_sorting.jl:1-163): after the call the
index is the identity, the first nbuffer buffer entries count down from
the vector's length, and every logical position keeps its particle."""
import numpy as np
import pytest

# (index before, particles in use, buffer before, nbuffer, expected buffer[:nbuffer],
#  whether the testset zeroes the positions not in use)
CASES = [
    ([1, 2, 3, 4, 5, 6, 7, 8, 9, 10], 0, [10, 9, 8, 7, 6, 5, 4, 3, 2, 1], 10, [10, 9, 8, 7, 6, 5, 4, 3, 2, 1],
     False),
    ([5, 1, 7, 3, 2, 4, 10, 8, 6, 9], 2, [7, 3, 2, 4, 10, 8, 6, 9, 5, 1], 8, [10, 9, 8, 7, 6, 5, 4, 3], True),
    ([10, 9, 5, 6, 1, 8, 3, 7, 2, 4], 7, [4, 7, 2, 10, 9, 5, 6, 1, 8, 3], 3, [10, 9, 8], True),
    ([4, 10, 7, 3, 1, 2, 9, 8, 5, 6], 10, [4, 7, 2, 10, 9, 5, 6, 1, 8, 3], 0, [], False),
    # the last five index entries are duplicate rubbish
    ([10, 5, 7, 6, 2, 1, 3, 5, 4, 6], 5, [3, 7, 4, 1, 9, 10, 8, 2, 6, 5], 5, [10, 9, 8, 7, 6], False),
]


def _particle(i):
    return float(i), [i, 2 * i, 3 * i], [-i, -4 * i, -5 * i]


def reference_case_state(mzo, case):
    """A ParticleVector(10) in the state one of the reference's testsets
    builds: the given index, particles i = 1..n_used set through the index,
    the positions not in use zeroed, the given buffer and nbuffer."""
    index, n_used, buffer, nbuffer, _, zero_rest = case
    pv = mzo.ParticleVector(10)
    pv.index[:] = index
    for i in range(1, (n_used or 10) + 1):
        pv.set(i, *_particle(i))
    if zero_rest:
        for i in range(n_used + 1, 11):
            pv.set(i, 0.0, [0, 0, 0], [0, 0, 0])
    pv.buffer[:] = buffer
    pv.s.nbuffer = nbuffer
    return pv


def test_fresh_vector_layout(mzo):
    """test_particle_index_sorting.jl:5-9: a new ParticleVector(10)."""
    pv = mzo.ParticleVector(10)
    assert list(pv.index) == list(range(1, 11))
    assert list(pv.buffer) == list(range(10, 0, -1))
    assert pv.nbuffer == 10


@pytest.mark.parametrize("case", CASES)
def test_restore_particle_ordering_reference_cases(mzo, case):
    pv = reference_case_state(mzo, case)
    n_used = case[1]
    mzo.lib().mzo_restore_particle_ordering(pv.p)
    assert list(pv.index) == list(range(1, 11))
    assert pv.nbuffer == case[3]
    assert list(pv.buffer[:case[3]]) == case[4]
    for i in range(1, (n_used or 10) + 1):
        w, v, x = pv.get(i)
        ew, ev, ex = _particle(i)
        assert w == ew and np.array_equal(v, ev) and np.array_equal(x, ex)
