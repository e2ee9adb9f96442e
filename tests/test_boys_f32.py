"""The FP32 kernels' Boys evaluation (mixed precision).

The library tabulates F_m(T_k), T_k = k/8, in FP64 and rounds to FP32; the
kernels stage the five orders they expand in LDS and evaluate F_LT(T) by a
5-term Taylor expansion about the nearest grid point, then recur downwards (F_m = (2T F_{m+1} + e^{-T}) / (2m+1)), and use the
asymptotic form sqrt(pi/4T) with upward recursion from T = 36 on.  Both the
table and that evaluation (emulated here in numpy float32, operation by
operation) are checked against the incomplete-gamma closed form
F_m(T) = Gamma(m+1/2) P(m+1/2, T) / (2 T^(m+1/2)) from scipy, which is
independent of the library and of the oracle.
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest
from scipy.special import gamma, gammainc

LIB = Path(__file__).resolve().parents[1] / "quick_amd" / "libquick_mi355x.so"
pytestmark = pytest.mark.skipif(not LIB.exists(), reason="library not built")


def boys_closed_form(m, T):
    T = np.asarray(T, dtype=np.float64)
    out = np.empty_like(T)
    small = T < 1e-12
    out[small] = 1.0 / (2 * m + 1)
    a = m + 0.5
    t = T[~small]
    out[~small] = gamma(a) * gammainc(a, t) / (2.0 * t ** a)
    return out


@pytest.fixture(scope="module")
def table():
    L = ctypes.CDLL(str(LIB))
    rows, cols, per = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L.qk_boys_table_shape(ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(per))
    tab = np.zeros((rows.value, cols.value), dtype=np.float32)
    rc = L.qk_boys_table_f32(tab.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), rows, cols)
    assert rc == 0
    return tab, per.value


def test_table_matches_closed_form(table):
    tab, per = table
    T = np.arange(tab.shape[0]) / per
    for m in range(tab.shape[1]):
        ref = boys_closed_form(m, T)
        rel = np.abs(tab[:, m].astype(np.float64) - ref) / ref
        assert rel.max() < 6.0e-8, (m, rel.max())   # FP32 rounding of an accurate value


def test_table_rejects_wrong_shape():
    L = ctypes.CDLL(str(LIB))
    buf = np.zeros(4, dtype=np.float32)
    assert L.qk_boys_table_f32(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 2, 2) != 0


def boys_f32_emulated(LT, T, tab, per):
    """Operation-by-operation float32 restatement of the device evaluation."""
    f = np.float32
    T = f(T)
    F = [f(0)] * (LT + 1)
    tmax = f(tab.shape[0] - 1) / f(per)
    if T < tmax:
        k = int(np.rint(T * f(per)))
        d = f(k) * (f(1) / f(per)) - T
        r = tab[k, LT:LT + 5]
        v = (((r[4] * (d * f(0.25)) + r[3]) * (d * (f(1) / f(3))) + r[2]) * (d * f(0.5)) + r[1]) * d + r[0]
        F[LT] = f(v)
        e = f(np.exp(-np.float64(T)))   # __expf is within 2 ulp
        t2 = T + T
        for m in range(LT - 1, -1, -1):
            F[m] = (t2 * F[m + 1] + e) * (f(1) / f(2 * m + 1))
    else:
        F[0] = np.sqrt(f(0.78539816339744831) / T)
        h = f(0.5) / T
        for m in range(1, LT + 1):
            F[m] = F[m - 1] * (f(2 * m - 1) * h)
    return np.array(F, dtype=np.float64)


@pytest.mark.parametrize("LT", [0, 1, 2, 4, 6, 8])
def test_fp32_evaluation_error_bound(table, LT):
    tab, per = table
    rng = np.random.default_rng(100 + LT)
    Ts = np.concatenate([rng.uniform(0, 1, 300), rng.uniform(0, 40, 600), rng.uniform(30, 200, 200),
                         [0.0, 1e-7, 1.0 / 32, 35.99, 36.0, 36.01]])
    worst = 0.0
    for T in Ts:
        F = boys_f32_emulated(LT, T, tab, per)
        for m in range(LT + 1):
            ref = boys_closed_form(m, np.array([T]))[0]
            worst = max(worst, abs(F[m] - ref) / ref)
    # a few FP32 ulps: the Taylor truncation (|d|^5/120 <= 8e-9, |d| <= 1/16) is below
    assert worst < 1.0e-6, worst
