"""Pins the oracle's FP32 ILU(0) apply (orc_ilu0_apply_f32, the bitwise
checker of the device FP32 triangular solves) -- CPU only.

* against the reference-pinned FP64 apply (orc_ilu0_apply, bitwise equal to
  ILUp_precond.c:26-46 on the reference's own fixtures): on factors, pivots
  and right-hand sides whose every intermediate is a small integer, both
  precisions compute exactly, so any difference in loop bounds, indices,
  signs or the order of the two sweeps shows up as unequal bits;
* against the mathematics: on the ILU(0) factors of a real SUPG matrix the
  FP32 result satisfies the componentwise backward-error bound of a
  triangular solve (Higham, Accuracy and Stability of Numerical
  Algorithms, Thm 8.5): the computed y of L y = p has
  |p - L y| <= gamma_m |L| |y|, gamma_m = m u / (1 - m u), u = 2^-24, with m
  the terms of the row, and likewise for the backward sweep with U and the
  inverted pivots.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, load_golden


@pytest.fixture(scope="module")
def o():
    return O.Oracle()


def _random_pattern(rng, n, per_row):
    """sorted CSR pattern with the diagonal and per_row random off-diagonal
    columns in each row"""
    IA = np.zeros(n + 1, np.int32)
    JA = []
    for i in range(n):
        cols = set(rng.choice(n, size=min(per_row, n - 1), replace=False).tolist()) | {i}
        JA.extend(sorted(cols))
        IA[i + 1] = len(JA)
    return IA, np.array(JA, np.int32)


def _exact_case(rng, n, per_row):
    """factors in {-1, 0, 1}, inverted pivots in {+-1, +-2}, p in [-3, 3]:
    every partial sum of both sweeps is an integer; accept the draw only when
    every intermediate stays below 2^20 (exact in FP32 and FP64)"""
    for _ in range(200):
        IA, JA = _random_pattern(rng, n, per_row)
        LU = rng.integers(-1, 2, len(JA)).astype(np.float64)
        D = rng.choice([-2.0, -1.0, 1.0, 2.0], n)
        p = np.append(rng.integers(-3, 4, n).astype(np.float64), 0.0)
        # integer replay of the two sweeps, recording the largest magnitude
        z = np.zeros(n + 1, dtype=object)
        big = 0
        for i in range(n):
            x = int(p[i])
            for q in range(IA[i], IA[i + 1]):
                if JA[q] < i:
                    x -= z[JA[q]] * int(LU[q])
                    big = max(big, abs(x))
            z[i] = x
        for i in range(n - 1, -1, -1):
            x = z[i]
            for q in range(IA[i], IA[i + 1]):
                if JA[q] > i:
                    x -= z[JA[q]] * int(LU[q])
                    big = max(big, abs(x))
            z[i] = x * int(D[i])
            big = max(big, abs(z[i]))
        if big < 2 ** 20:
            return IA, JA, LU, D, p
    raise AssertionError("no exact draw found")


@pytest.mark.parametrize("seed,n,per_row", [(1, 8, 2), (2, 20, 3), (3, 40, 2), (4, 64, 4)])
def test_ilu0_apply_f32_equals_pinned_f64_on_exact_inputs(o, seed, n, per_row):
    rng = np.random.default_rng(seed)
    IA, JA, LU, D, p = _exact_case(rng, n, per_row)
    z64 = o.ilu0_apply(n, LU, JA, IA, D, p)
    z32 = o.ilu0_apply_f32(n, LU, JA, IA, D, p)
    assert np.all(np.abs(z64[:n]) < 2 ** 20)
    assert np.array_equal(z32[:n].astype(np.float64), z64[:n])
    # and the pinned FP64 apply really inverts L U D^-1 here
    L = np.eye(n)
    U = np.zeros((n, n))
    for i in range(n):
        for q in range(IA[i], IA[i + 1]):
            if JA[q] < i:
                L[i, JA[q]] = LU[q]
            elif JA[q] > i:
                U[i, JA[q]] = LU[q]
        U[i, i] = 1.0 / D[i]
    assert np.array_equal(L @ (U @ z64[:n]), p[:n])


def test_ilu0_apply_f32_plants_fail(o):
    """a wrong operand in the FP32 call (the factors of another pattern, the
    pivots reversed) is caught by the exact comparison"""
    rng = np.random.default_rng(11)
    n = 24
    IA, JA, LU, D, p = _exact_case(rng, n, 3)
    z64 = o.ilu0_apply(n, LU, JA, IA, D, p)
    assert not np.array_equal(o.ilu0_apply_f32(n, LU, JA, IA, D[::-1].copy(), p)[:n], z64[:n])
    assert not np.array_equal(o.ilu0_apply_f32(n, -LU, JA, IA, D, p)[:n], z64[:n])


def _gamma(m):
    u = 2.0 ** -24
    return m * u / (1 - m * u)


@pytest.mark.parametrize("name", ["ss3d_kuhn4.npz", "ss3d_difconv_125_384.npz"])
def test_ilu0_apply_f32_backward_error_bound(o, name):
    d = load_golden(os.path.join(GOLDEN, name))
    n = int(d["neq"])
    IA, JA, AA = d["IA"], d["JA"], d["AA"]
    rc, LU, D = o.ilu0_factor(n, AA, JA, IA)
    assert rc == 0
    L32, D32 = LU.astype(np.float32), D.astype(np.float32)
    rng = np.random.default_rng(5)
    for trial in range(4):
        p = np.append(rng.standard_normal(n), 0.0).astype(np.float32)
        z = o.ilu0_apply_f32(n, LU, JA, IA, D, p)[:n].astype(np.float64)
        # the intermediate y of the forward sweep, recovered from z exactly
        # enough in FP64: y_i = z_i / D_i + sum_{j>i} U_ij z_j
        y = np.zeros(n)
        for i in range(n):
            s = z[i] / float(D32[i])
            for q in range(IA[i], IA[i + 1]):
                if JA[q] > i:
                    s += float(L32[q]) * z[JA[q]]
            y[i] = s
        for i in range(n):
            m = IA[i + 1] - IA[i] + 1
            lo = [(float(L32[q]), y[JA[q]]) for q in range(IA[i], IA[i + 1]) if JA[q] < i]
            # forward: |p_i - y_i - sum L_ij y_j| <= gamma_m (|y_i| + sum |L_ij y_j|)
            r = float(p[i]) - y[i] - sum(a * b for a, b in lo)
            bound = _gamma(m) * (abs(y[i]) + sum(abs(a * b) for a, b in lo)) * 1.01 + 1e-300
            # the recovered y carries the backward sweep's rounding as well
            up = [(float(L32[q]), z[JA[q]]) for q in range(IA[i], IA[i + 1]) if JA[q] > i]
            bound += _gamma(m + 1) * (abs(z[i] / float(D32[i])) +
                                      sum(abs(a * b) for a, b in up)) * 2.02
            assert abs(r) <= bound, (trial, i, r, bound)
        # and the FP32 result against the FP64 one: relative error at FP32 level
        z64 = o.ilu0_apply(n, LU, JA, IA, D, p.astype(np.float64))[:n]
        assert np.max(np.abs(z - z64)) <= 1e-5 * np.max(np.abs(z64))
