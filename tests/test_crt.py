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


def _sk_schedule(ntiles, nm, pairs, G):
    from autojax_amd import _lib

    units = nm * ntiles
    seg_cap = units + 256
    segs = np.zeros((seg_cap, 8), np.int32)
    beg = np.zeros(G + 1, np.int32)
    fix = np.zeros((units, 4), np.int32)
    counts = np.zeros(3, np.int64)
    _lib.call("aj_crt_sk_schedule", ntiles, nm, pairs, G, segs.ctypes.data, seg_cap, beg.ctypes.data,
              fix.ctypes.data, units, counts.ctypes.data)
    return segs[:counts[0]], beg, fix[:counts[1]], int(counts[2])


@pytest.mark.parametrize("G", [8, 64, 256])
@pytest.mark.parametrize("pairs", [1, 7, 256])
@pytest.mark.parametrize("nm", [11, 12])
@pytest.mark.parametrize("ntiles", [1, 3, 10, 55, 210, 861])
def test_persistent_syrk_schedule(ntiles, nm, pairs, G):
    """The persistent residue SYRK's schedule as the library builds it: every (tile, modulus) unit's
    stage pairs covered exactly once by contiguous segments, a unit written whole or in pieces whose
    partial-residue slots are distinct and listed for the fix-up, and the workgroups' loads within one
    stage pair of each other (when every XCD runs the same number of full rounds)."""
    segs, beg, fix, nslots = _sk_schedule(ntiles, nm, pairs, G)
    t, mi, k0, ns, slot = (segs[:, c].astype(np.int64) for c in range(5))
    assert beg[0] == 0 and beg[G] == len(segs) and np.all(np.diff(beg) >= 0)
    assert np.all((t >= 0) & (t < ntiles) & (mi >= 0) & (mi < nm))
    assert np.all((ns >= 2) & (ns % 2 == 0) & (k0 % 2 == 0) & (k0 + ns <= 2 * pairs))
    whole = slot < 0
    assert np.all(k0[whole] == 0) and np.all(ns[whole] == 2 * pairs)
    # coverage: per unit, segments sorted by first stage start at 0, abut, and end at 2 pairs
    unit = mi * ntiles + t
    o = np.lexsort((k0, unit))
    u, a, e = unit[o], k0[o], (k0 + ns)[o]
    assert np.array_equal(np.unique(u), np.arange(nm * ntiles))
    first = np.r_[True, u[1:] != u[:-1]]
    last = np.r_[u[1:] != u[:-1], True]
    assert np.all(a[first] == 0) and np.all(e[last] == 2 * pairs)
    assert np.all(a[1:][~last[:-1]] == e[:-1][~last[:-1]])
    # slots: the pieces' slots are 0 .. nslots-1 once each; the fix-up list names each split unit's run
    assert np.array_equal(np.sort(slot[~whole]), np.arange(nslots))
    unit_of_slot = np.empty(nslots, np.int64)
    unit_of_slot[slot[~whole]] = unit[~whole]
    assert int(fix[:, 3].sum()) == nslots and np.all(fix[:, 3] >= 2)
    for ft, fm, s0, n in fix:
        assert np.all(unit_of_slot[s0:s0 + n] == fm * ntiles + ft)
    # balance: stage pairs per workgroup.  Workgroup j * 8 + x runs on XCD x, whose tiles
    # [x ntiles / 8, (x + 1) ntiles / 8) give F_x = (their units) // (G / 8) full rounds; the leftover pairs
    # level the rest, so the loads agree to one pair when all F_x are equal and otherwise differ by at
    # most the difference of full rounds
    load = np.array([ns[beg[w]:beg[w + 1]].sum() // 2 for w in range(G)])
    assert load.sum() == nm * ntiles * pairs
    F = [nm * ((x + 1) * ntiles // 8 - x * ntiles // 8) // (G // 8) for x in range(8)]
    assert load.max() - load.min() <= (max(F) - min(F)) * pairs + 1


def test_persistent_syrk_schedule_rejects_bad_arguments():
    from autojax_amd import _lib

    buf = np.zeros(4096, np.int64)
    for ntiles, nm, pairs, G in ((10, 11, 4, 12), (10, 10, 4, 8), (0, 11, 4, 8), (10, 12, 0, 8), (10, 12, 4, 264)):
        with pytest.raises(ValueError):
            _lib.call("aj_crt_sk_schedule", ntiles, nm, pairs, G, buf.ctypes.data, 1, buf.ctypes.data, buf.ctypes.data,
                      1, buf.ctypes.data)
    with pytest.raises(ValueError, match="exceed"):  # tables too small for the schedule
        _lib.call("aj_crt_sk_schedule", 10, 11, 4, 8, buf.ctypes.data, 1, buf.ctypes.data, buf.ctypes.data, 1,
                  buf.ctypes.data)
