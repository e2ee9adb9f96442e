"""CPU check of the gradient kernel's generated tables (csrc/gen_grad.py).

The gradient kernel (quick_amd/csrc/qk_grad.hip) is driven entirely by one
int32 blob per quartet class: a raised VRR node program, four target lists, and
per derivative term a bra HRR, a ket HRR and a contraction table.  This test
executes a class's blob in numpy the way the kernel does (same header, same
encodings, same weights) for one Qshell quartet on four different atoms and a
sparse random density factor K, and compares the nine sums
sum_abcd K_abcd d(ab|cd)/dA,B,C with central differences of the oracle's
contracted ERIs at displaced centres -- so an encoding or indexing mistake in
the generator is caught before the tables run on a device.
"""
import copy
import math
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "quick_amd" / "csrc"))
import gen_eri  # noqa: E402
import gen_grad  # noqa: E402

from oracle.oracle import Oracle, fmt  # noqa: E402

X0 = 2.0 * math.pi ** 2.5
H = gen_grad.HDR
SF = gen_grad.N_STAGE_FIELDS
NT = gen_grad.MAX_TERMS


def qshell_data(fx, shell, L):
    npr = fx.kprim[shell]
    col = fx.ksumtype[shell] - 1 + L
    f0 = fx.ksumtype[shell] - 1 + fx.Qsbasis[shell, L]
    return fx.xyz[:, fx.katom[shell] - 1], fx.gcexpo[:npr, col], fx.gccoeff[:npr, col], f0


def run_blob(cls, qs, Mb, K):
    blob, _ = gen_grad.class_blob(cls)
    b = np.array(blob)
    nslots, LT2 = b[2], b[3] + 2
    n0, nA, nB, nC = b[4:8]
    nstages, nout = b[12], b[13]
    nodes = b[b[18]:b[18] + 4 * b[0]].reshape(-1, 4)
    lev = b[b[19]:b[19] + b[1] + 1]
    tgt = b[b[20]:b[20] + n0 + nA + nB + nC]
    (A, ea, ca, _), (B, eb, cb, _), (C, ec, cc, _), (D, ed, cd, _) = qs
    S = np.zeros(n0 + nA + nB + nC)
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
                    F = fmt(Mb, T, exact=True)[:LT2 + 1] * X0 * Kab * Kcd * math.sqrt(inv)
                    PA, WP, QC, WQ = P - A, W - P, Q - C, W - Q
                    Vs = np.zeros(nslots)
                    for k in range(lev[0], lev[1]):
                        Vs[nodes[k][0] & 0xffff] = F[(nodes[k][3] >> 8) & 0xff]
                    for L in range(1, len(lev) - 1):
                        for k in range(lev[L], lev[L + 1]):
                            x0, x1, x2, fl = (int(t) for t in nodes[k])
                            isbra = (fl & 3) == 1
                            d = (fl >> 2) & 3
                            n1, nee = (fl >> 4) & 15, (fl >> 8) & 15
                            c1 = PA[d] if isbra else QC[d]
                            c2 = WP[d] if isbra else WQ[d]
                            c3 = n1 * (0.5 / zab if isbra else 0.5 / zcd)
                            c4 = zcd * inv if isbra else zab * inv
                            Vs[x0 & 0xffff] = (c1 * Vs[x0 >> 16] + c2 * Vs[x1 & 0xffff] +
                                               c3 * (Vs[x1 >> 16] - c4 * Vs[x2 & 0xffff]) + nee * 0.5 * inv * Vs[x2 >> 16])
                    w = np.concatenate([np.ones(n0), np.full(nA, 2 * x), np.full(nB, 2 * y), np.full(nC, 2 * u)])
                    S += w * Vs[tgt]
    pw = np.zeros(24)
    for k in range(24):
        ax, e = (k % 12) // 4, k % 4
        pw[k] = ((A[ax] - B[ax]) if k < 12 else (C[ax] - D[ax])) ** e
    setoff = [0, n0, n0 + nA, n0 + nA + nB]
    g = np.zeros(9)
    for st in range(nstages):
        dfield, src, nR1, nCol, nR2, ob, ok, oc = b[H + SF * st: H + SF * (st + 1)]
        # direction (bits 0-7), most bra / ket terms of a row (bits 8-15 / 16-23):
        # only those leading slots of each row are read, as on the device
        dirn, nb, nk = dfield & 0xff, (dfield >> 8) & 0xff, (dfield >> 16) & 0xff
        Ss = S[setoff[src]:]
        # packed entries: term = row | pow << 12 | coef << 18; contraction = index | factor << 16
        unpack = lambda e: (int(e) & 0xfff, (int(e) >> 12) & 63, (int(e) >> 18) & 15)
        bt = b[ob:ob + nR1 * NT].reshape(nR1, NT)[:, :nb]
        kt = b[ok:ok + nR2 * NT].reshape(nR2, NT)[:, :nk]
        con = b[oc:oc + nout * 4].reshape(nout, 4)
        o1 = np.zeros(nR1 * nCol)
        for idx in range(nR1 * nCol):
            r1, col = divmod(idx, nCol)
            for e in bt[r1]:
                row, pc, cf = unpack(e)
                o1[idx] += cf * pw[pc & 3] * pw[4 + ((pc >> 2) & 3)] * pw[8 + ((pc >> 4) & 3)] * Ss[row * nCol + col]
        o2 = np.zeros(nR1 * nR2)
        for idx in range(nR1 * nR2):
            r1, r2 = divmod(idx, nR2)
            for e in kt[r2]:
                colk, pc, cf = unpack(e)
                o2[idx] += cf * pw[12 + (pc & 3)] * pw[16 + ((pc >> 2) & 3)] * pw[20 + ((pc >> 4) & 3)] * o1[r1 * nCol + colk]
        for o in range(nout):
            for x in range(3):
                e = int(con[o, x])
                ix, fac = e & 0xffff, e >> 16
                if ix != 0xffff:
                    g[3 * dirn + x] += fac * K.ravel()[o] * o2[ix]
    return g


def eri_raw(fx, xyz, i, j, k, l):
    f = copy.copy(fx)
    f.xyz = xyz
    v = Oracle(f).eri(i, j, k, l)
    return v / (fx.cons[i] * fx.cons[j] * fx.cons[k] * fx.cons[l])


@pytest.fixture(scope="module")
def o4():
    """Four O atoms on a distorted tetrahedron, ~1.3 A apart: every s/p/d Qshell
    on four distinct centres with sizeable overlaps."""
    from quick_amd.basis import build_basis
    xyz = np.array([[0.0, 0.0, 0.0], [1.3, 0.1, 0.0], [0.5, 1.2, 0.2], [0.6, 0.5, 1.1]])
    return build_basis(["O"] * 4, xyz)


@pytest.mark.parametrize("cls", gen_eri.quartet_classes())
def test_grad_blob_matches_finite_differences(o4, cls):
    fx = o4
    rng = np.random.default_rng(7 + sum(cls) * 5 + cls[0])
    # the last (most diffuse) shell of each angular momentum on atoms 0..3
    shells = []
    for atom, l in enumerate(cls):
        cand = [s for s in range(fx.nshell) if fx.katom[s] - 1 == atom and fx.Qstart[s] <= l <= fx.Qfinal[s]]
        shells.append(cand[-1])
    atoms = [0, 1, 2, 3]
    qs = [qshell_data(fx, s, L) for s, L in zip(shells, cls)]
    Mb = sum(fx.Qfinal[s] for s in shells) + 2
    na, nb, nc, nd = [(l + 1) * (l + 2) // 2 for l in cls]
    K = np.zeros((na, nb, nc, nd))
    if K.size <= 36:
        picks = [tuple(int(v) for v in p) for p in np.ndindex(K.shape)]
    else:
        picks = sorted({tuple(int(rng.integers(n)) for n in K.shape) for _ in range(12)})
    for p in picks:
        K[p] = rng.normal()
    g = run_blob(cls, qs, Mb, K)
    h = 1e-5
    f0 = [q[3] for q in qs]
    ref = np.zeros(9)
    for c in range(3):                      # centres A, B, C
        for x in range(3):
            for sgn in (1, -1):
                xyz = fx.xyz.copy()
                xyz[x, atoms[c]] += sgn * h
                for p in picks:
                    i, j, k, l = (f0[t] + p[t] for t in range(4))
                    ref[3 * c + x] += sgn * K[p] * eri_raw(fx, xyz, i, j, k, l) / (2 * h)
    scale = np.abs(ref).max()
    assert scale > 1e-6                      # the check is not vacuous
    assert np.abs(g - ref).max() <= 1e-6 * scale, (cls, g, ref)
