"""CPU checks of the residue-number (CRT) W-tilde mode's constants, without a GPU.

The library's recombination constants (aj_crt_constants) are compared with exact Python integer
arithmetic: the moduli are pairwise coprime, their product P covers the largest integer Gram entry
of a 16384-visibility chunk on the five-digit fixed-point grid, and the split CRT basis recovers
integers exactly when evaluated the way the combine kernel does (FP64, exact high part).
"""

import math

import numpy as np

KCRT = 16384                  # visibilities per exact integer Gram matrix (wtilde_crt.hip)
QMAX = 127 * 256**4           # largest feature integer (five signed 8-bit digits)


def constants():
    import ctypes

    from autojax_amd import _lib

    m, yh, yl, pq = (np.zeros(12), np.zeros(12), np.zeros(12), np.zeros(3))
    _lib.call("aj_crt_constants", *(a.ctypes.data_as(ctypes.c_void_p) for a in (m, yh, yl, pq)))
    return m, yh, yl, pq


def test_moduli_and_range():
    m, _, _, _ = constants()
    mi = [int(v) for v in m]
    assert all(v <= 256 for v in mi)
    for i in range(12):
        for j in range(i):
            assert math.gcd(mi[i], mi[j]) == 1
    P = math.prod(mi)
    # |G_ij| <= G_ii <= KCRT (Q^2 + sqrt(2) Q + 1/2): must lie inside (-P/2, P/2) with room for the
    # FP64 estimate of the quotient
    assert KCRT * (QMAX * QMAX + 2 * QMAX) < 0.2 * P
    # int32 accumulation of 2 KCRT products of residues |r| <= 128 cannot overflow
    assert 2 * KCRT * 128 * 128 < 2**31


def test_basis_is_exact_crt_basis():
    m, yh, yl, pq = constants()
    mi = [int(v) for v in m]
    P = math.prod(mi)
    for i in range(12):
        hi, lo = int(yh[i]), int(yl[i])
        assert hi == yh[i] and lo == yl[i] and hi % 2**53 == 0 and 0 <= lo < 2**53
        y = hi + lo
        assert 0 <= y < P
        for j in range(12):
            assert y % mi[j] == (1 if i == j else 0)
    assert int(pq[0]) + int(pq[1]) == P and int(pq[0]) % 2**53 == 0
    assert abs(pq[2] * P - 1) < 1e-15


def _balanced(x, mod):
    r = x % mod
    return r - mod if r > (mod - 1) // 2 else r


def test_recombination_recovers_integers():
    """The combine kernel's FP64 evaluation returns the integer to within 2^20 of |X| <= 2^93."""
    m, yh, yl, pq = constants()
    mi = [int(v) for v in m]
    rng = np.random.default_rng(7)
    bound = KCRT * (QMAX * QMAX + 2 * QMAX)
    xs = [0, 1, -1, bound, -bound, bound - 1, -(bound - 1)]
    xs += [int(v) for v in (rng.random(200) * 2 - 1) * float(bound)]
    xs += [int(rng.integers(-2**40, 2**40)) for _ in range(50)]
    for x in xs:
        r = np.array([float(_balanced(x, v)) if v != 256 else float(((x + 128) % 256) - 128) for v in mi])
        assert np.all(np.abs(r) <= 128)
        shi = 0.0
        slo = 0.0
        for i in range(12):
            shi = shi + r[i] * yh[i]
            slo = slo + r[i] * yl[i]
        n = np.rint((shi + slo) * pq[2])
        G = (shi - n * pq[0]) + (slo - n * pq[1])
        assert abs(int(G) - x) <= 2**20, (x, int(G))


def test_recombination_fails_with_a_wrong_basis():
    """A plant: swapping two basis vectors must break the recovery (the check above has teeth)."""
    m, yh, yl, pq = constants()
    mi = [int(v) for v in m]
    yh2, yl2 = yh.copy(), yl.copy()
    yh2[[3, 4]] = yh2[[4, 3]]
    yl2[[3, 4]] = yl2[[4, 3]]
    x = 123456789123456789123
    r = np.array([float(_balanced(x, v)) if v != 256 else float(((x + 128) % 256) - 128) for v in mi])
    shi = float(np.sum(r * yh2))
    slo = float(np.sum(r * yl2))
    n = np.rint((shi + slo) * pq[2])
    G = (shi - n * pq[0]) + (slo - n * pq[1])
    assert abs(int(G) - x) > 2**40
