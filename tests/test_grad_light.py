"""CPU check of the light classes' generated gradient code (grad_light.inc).

The straight-line code is plain templated C++; it is compiled here for the host
with g++ (the device qualifiers defined away), driven over the primitive
quartets of one shell quartet on four distinct oxygen centres, and compared
with the table-driven gradient program of the same class (checked against
finite differences in test_grad_tables.py); its FP32 instantiation, used by the
mixed-precision kernels, is held to 2e-5 of the largest derivative.
"""
import ctypes
import math
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "quick_amd" / "csrc"
sys.path.insert(0, str(CSRC))
import gen_eri  # noqa: E402
import gen_grad  # noqa: E402
from tests.test_grad_tables import qshell_data, run_blob, X0  # noqa: E402

from oracle.oracle import fmt  # noqa: E402


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    inc = CSRC / "grad_light.inc"
    gen_grad.write_light(str(inc))
    d = tmp_path_factory.mktemp("gl")
    src = ["#define __device__", "#define __forceinline__ inline", '#include "grad_light.inc"', """
// One pass over the primitive quartets for the contracted sets named by mask M
// (bit 0: S0, 1: SA, 2: SB, 3: SC), then its contraction into g; the device
// kernel runs the passes of GradSets_<class>::kMask in the same way.
template <class S, int M, class R, class VRG, class GCON>
static void one_pass(int np, const R* F, int nf, const R* geo, const R* wts, const R* ABCD,
                     const R* K, R* g, VRG vrg, GCON gcon) {
    R a0[S::n0] = {}, aA[S::nA] = {}, aB[S::nB] = {}, aC[S::nC] = {};
    for (int p = 0; p < np; ++p) {
        const R* G = geo + 17 * p;
        vrg(F + nf * p, G[0], G[1], G[2], G[3], G[4], G[5], G[6], G[7], G[8], G[9], G[10], G[11],
            G[12], G[13], G[14], G[15], G[16], wts[3 * p], wts[3 * p + 1], wts[3 * p + 2], a0, aA, aB, aC);
    }
    gcon(a0, aA, aB, aC, ABCD[0], ABCD[1], ABCD[2], ABCD[3], ABCD[4], ABCD[5], K, g);
}"""]
    for cls in gen_grad.LIGHT_CLASSES:
        t = gen_eri.tag(cls)
        for R, suf in (("double", ""), ("float", "_f32")):
            passes = "\n".join(
                f"    if constexpr (S::kPasses > {k}) one_pass<S, S::kMask[{k}], {R}>(np, F, nf, geo, wts, ABCD, K, g, "
                f"vrg_{t}<S::kMask[{k}], {R}>, gcon_{t}<S::kMask[{k}], {R}>);" for k in range(4))
            src.append(f"""extern "C" void run{suf}_{t}(int np, const {R}* F, const {R}* geo, const {R}* wts,
                        const {R}* ABCD, const {R}* K, {R}* g) {{
    using S = GradSets_{t};
    const int nf = {sum(cls) + 3};
    for (int k = 0; k < 9; ++k) g[k] = 0;
{passes}
}}""")
    (d / "shim.cpp").write_text("\n".join(src) + "\n")
    so = d / "shim.so"
    subprocess.run(["g++", "-O1", "-fPIC", "-shared", "-std=c++17", "-I", str(CSRC), str(d / "shim.cpp"),
                    "-o", str(so)], check=True)
    return ctypes.CDLL(str(so))


@pytest.fixture(scope="module")
def o4():
    from quick_amd.basis import build_basis
    xyz = np.array([[0.0, 0.0, 0.0], [1.3, 0.1, 0.0], [0.5, 1.2, 0.2], [0.6, 0.5, 1.1]])
    return build_basis(["O"] * 4, xyz)


@pytest.mark.parametrize("cls", gen_grad.LIGHT_CLASSES)
def test_light_gradient_code_matches_table_program(shim, o4, cls):
    fx = o4
    rng = np.random.default_rng(3 + sum(cls))
    shells = []
    for atom, l in enumerate(cls):
        cand = [s for s in range(fx.nshell) if fx.katom[s] - 1 == atom and fx.Qstart[s] <= l <= fx.Qfinal[s]]
        shells.append(cand[len(cand) // 2])   # a contracted shell: several primitive quartets
    qs = [qshell_data(fx, s, L) for s, L in zip(shells, cls)]
    Mb = sum(fx.Qfinal[s] for s in shells) + 2
    na, nb, nc, nd = [(l + 1) * (l + 2) // 2 for l in cls]
    K = rng.normal(size=(na, nb, nc, nd))
    ref = run_blob(cls, qs, Mb, K)
    (A, ea, ca, _), (B, eb, cb, _), (C, ec, cc, _), (D, ed, cd, _) = qs
    LT2 = sum(cls) + 2
    Fs, geo, wts = [], [], []
    for x, ax in zip(ea, ca):
        for y, by in zip(eb, cb):
            zab = x + y
            P = (A * x + B * y) / zab
            Kab = math.exp(-x * y / zab * np.sum((A - B) ** 2)) / zab * ax * by
            for u, cu in zip(ec, cc):
                for v, dv in zip(ed, cd):
                    zcd = u + v
                    Q = (C * u + D * v) / zcd
                    Kcd = math.exp(-u * v / zcd * np.sum((C - D) ** 2)) / zcd * cu * dv
                    inv = 1.0 / (zab + zcd)
                    W = (P * zab + Q * zcd) * inv
                    T = np.sum((P - Q) ** 2) * zab * zcd * inv
                    Fs.append(fmt(Mb, T, exact=True)[:LT2 + 1] * X0 * Kab * Kcd * math.sqrt(inv))
                    geo.append(np.concatenate([P - A, W - P, Q - C, W - Q,
                                               [0.5 / zab, 0.5 / zcd, 0.5 * inv, zcd * inv, zab * inv]]))
                    wts.append([2 * x, 2 * y, 2 * u])
    F = np.ascontiguousarray(np.array(Fs))
    G = np.ascontiguousarray(np.array(geo))
    Wt = np.ascontiguousarray(np.array(wts))
    ABCD = np.concatenate([A - B, C - D])
    g = np.zeros(9)
    dp = ctypes.POINTER(ctypes.c_double)
    p = lambda arr: arr.ctypes.data_as(dp)
    Kf = np.ascontiguousarray(K.ravel())
    getattr(shim, f"run_{gen_eri.tag(cls)}")(len(Fs), p(F), p(G), p(Wt), p(ABCD), p(Kf), p(g))
    scale = np.abs(ref).max()
    assert scale > 1e-10
    assert np.abs(g - ref).max() <= 1e-11 * max(1.0, scale), (cls, g, ref)
    # the same code instantiated in FP32 (the mixed-precision kernels)
    fp = ctypes.POINTER(ctypes.c_float)
    q = lambda arr: arr.ctypes.data_as(fp)
    f32 = lambda arr: np.ascontiguousarray(arr, dtype=np.float32)
    F32, G32, W32, ABCD32, K32 = f32(F), f32(G), f32(Wt), f32(ABCD), f32(Kf)
    g32 = np.zeros(9, dtype=np.float32)
    getattr(shim, f"run_f32_{gen_eri.tag(cls)}")(len(Fs), q(F32), q(G32), q(W32), q(ABCD32), q(K32), q(g32))
    err = np.abs(g32.astype(np.float64) - ref).max()
    assert err <= 2e-5 * scale, (cls, err, scale)
