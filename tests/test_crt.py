"""CPU checks of the residue-number (CRT) W-tilde modes' constants, without a GPU.

The library's recombination constants (aj_crt_constants) for 12 moduli (int8crt) and 11 moduli
(int8crt11) are compared with exact Python integer arithmetic: the moduli are pairwise coprime,
their product P covers the largest integer Gram entry of a chunk on the mode's fixed-point grid,
and the split CRT basis recovers integers exactly when evaluated the way the combine kernel does
(FP64, exact high part).
"""

import math

import numpy as np
import pytest

# (moduli, largest chunk in visibilities) per mode; the largest feature integer Q comes from the library
MODES = {12: 32768, 11: 16384}


def constants(nm):
    import ctypes

    from autojax_amd import _lib

    m, yh, yl, pq = (np.zeros(nm), np.zeros(nm), np.zeros(nm), np.zeros(4))
    _lib.call("aj_crt_constants", nm, *(a.ctypes.data_as(ctypes.c_void_p) for a in (m, yh, yl, pq)))
    return m, yh, yl, pq


@pytest.mark.parametrize("nm", [12, 11])
def test_moduli_and_range(nm):
    m, _, _, pq = constants(nm)
    mi = [int(v) for v in m]
    assert all(v <= 256 for v in mi) and len(set(mi)) == nm
    for i in range(nm):
        for j in range(i):
            assert math.gcd(mi[i], mi[j]) == 1
    P = math.prod(mi)
    Q = int(pq[3])
    assert Q == pq[3]
    kc = MODES[nm]
    # |G_ij| <= G_ii <= kc (Q^2 + sqrt(2) Q + 1/2) must lie inside (-P/2, P/2) with room for the FP64
    # estimate of the quotient (relative error about 1e-12)
    assert kc * (Q * Q + 2 * Q) < 0.45 * P
    # int32 accumulation of 2 kc products of residues |r| <= 128 cannot overflow
    assert 2 * kc * 128 * 128 < 2**31


@pytest.mark.parametrize("nm", [12, 11])
def test_basis_is_exact_crt_basis(nm):
    m, yh, yl, pq = constants(nm)
    mi = [int(v) for v in m]
    P = math.prod(mi)
    for i in range(nm):
        hi, lo = int(yh[i]), int(yl[i])
        assert hi == yh[i] and lo == yl[i] and hi % 2**53 == 0 and 0 <= lo < 2**53
        y = hi + lo
        assert 0 <= y < P
        for j in range(nm):
            assert y % mi[j] == (1 if i == j else 0)
    assert int(pq[0]) + int(pq[1]) == P and int(pq[0]) % 2**53 == 0
    assert abs(pq[2] * P - 1) < 1e-15


def _balanced(x, mod):
    if mod == 256:
        return ((x + 128) % 256) - 128
    r = x % mod
    return r - mod if r > (mod - 1) // 2 else r


def _recombine(x, mi, yh, yl, pq):
    r = np.array([float(_balanced(x, v)) for v in mi])
    assert np.all(np.abs(r) <= 128)
    shi = 0.0
    slo = 0.0
    for i in range(len(mi)):
        shi = shi + r[i] * yh[i]
        slo = slo + r[i] * yl[i]
    n = np.rint((shi + slo) * pq[2])
    return (shi - n * pq[0]) + (slo - n * pq[1])


@pytest.mark.parametrize("nm", [12, 11])
def test_recombination_recovers_integers(nm):
    """The combine kernel's FP64 evaluation returns the integer to within 2^20 of |X| <= kc (Q^2 + 2Q)."""
    m, yh, yl, pq = constants(nm)
    mi = [int(v) for v in m]
    Q = int(pq[3])
    rng = np.random.default_rng(7)
    bound = MODES[nm] * (Q * Q + 2 * Q)
    xs = [0, 1, -1, bound, -bound, bound - 1, -(bound - 1)]
    xs += [int(v) for v in (rng.random(200) * 2 - 1) * float(bound)]
    xs += [int(rng.integers(-2**40, 2**40)) for _ in range(50)]
    for x in xs:
        G = _recombine(x, mi, yh, yl, pq)
        assert abs(int(G) - x) <= 2**20, (x, int(G))


def test_recombination_fails_with_a_wrong_basis():
    """A plant: swapping two basis vectors must break the recovery (the check above has teeth)."""
    m, yh, yl, pq = constants(12)
    mi = [int(v) for v in m]
    yh2, yl2 = yh.copy(), yl.copy()
    yh2[[3, 4]] = yh2[[4, 3]]
    yl2[[3, 4]] = yl2[[4, 3]]
    x = 123456789123456789123
    assert abs(int(_recombine(x, mi, yh, yl, pq)) - x) <= 2**20
    assert abs(int(_recombine(x, mi, yh2, yl2, pq)) - x) > 2**40
