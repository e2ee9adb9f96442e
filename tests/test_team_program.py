"""CPU check of the wavefront-cooperative heavy-class programs.

The team kernel (quick_amd/csrc/qk_team.h) executes generated node programs
(VRR) and index tables (HRR) instead of straight-line code.  This test runs the
same programs in numpy -- same node encoding, same coefficient selection, same
table layout -- on real shell quartets and compares the integrals with the
oracle, so an encoding or indexing mistake in the generator is caught before the
programs run on a device.
"""
import math
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "quick_amd" / "csrc"))
import gen_eri  # noqa: E402

from oracle.oracle import Oracle, fmt  # noqa: E402

X0 = 2.0 * math.pi ** 2.5


def qshell_data(fx, shell, L):
    """(centre, exponents, coefficients, first function) as the library builds them."""
    npr = fx.kprim[shell]
    col = fx.ksumtype[shell] - 1 + L
    f0 = fx.ksumtype[shell] - 1 + fx.Qsbasis[shell, L]
    return fx.xyz[:, fx.katom[shell] - 1], fx.gcexpo[:npr, col], fx.gccoeff[:npr, col], f0


def run_program(cls, qa, qb, qc, qd, Mb):
    nodes, lev, tslots, nslots, bra, ket, ne, nf = gen_eri.team_tables(cls)
    la, lb, lc, ld = cls
    (A, ea, ca, _), (B, eb, cb, _), (C, ec, cc, _), (D, ed, cd, _) = qa, qb, qc, qd
    LT = la + lb + lc + ld
    acc = np.zeros(ne * nf)
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
                    F = fmt(Mb, T)[:LT + 1] * X0 * Kab * Kcd * math.sqrt(inv)
                    PA, WP, QC, WQ = P - A, W - P, Q - C, W - Q
                    Vs = np.zeros(nslots)
                    for k in range(lev[0], lev[1]):
                        dst = nodes[k][0] & 0xffff
                        Vs[dst] = F[(nodes[k][3] >> 8) & 0xff]
                    for L in range(1, len(lev) - 1):
                        for k in range(lev[L], lev[L + 1]):
                            x0, x1, x2, fl = nodes[k]
                            isbra = (fl & 3) == 1
                            d = (fl >> 2) & 3
                            n1, nee = (fl >> 4) & 15, (fl >> 8) & 15
                            c1 = PA[d] if isbra else QC[d]
                            c2 = WP[d] if isbra else WQ[d]
                            c3 = n1 * (0.5 / zab if isbra else 0.5 / zcd)
                            c4 = zcd * inv if isbra else zab * inv
                            Vs[x0 & 0xffff] = (c1 * Vs[x0 >> 16] + c2 * Vs[x1 & 0xffff] +
                                               c3 * (Vs[x1 >> 16] - c4 * Vs[x2 & 0xffff]) + nee * 0.5 * inv * Vs[x2 >> 16])
                    acc += Vs[tslots]
    na, nb, nc, nd = [(l + 1) * (l + 2) // 2 for l in cls]
    pw = np.zeros(18)
    for ax in range(3):
        for k in range(3):
            pw[ax * 3 + k] = (A[ax] - B[ax]) ** k
            pw[9 + ax * 3 + k] = (C[ax] - D[ax]) ** k
    hb = np.zeros(na * nb * nf)
    for h in range(na * nb * nf):
        ab, f = divmod(h, nf)
        for u in range(4):
            e, p3, cf = bra[ab][u]   # the generator packs each term as e | p3 << 12 | cf << 18
            packed = e | (p3 << 12) | (cf << 18)
            assert (packed & 0xfff, (packed >> 12) & 63, (packed >> 18) & 15) == (e, p3, cf)
            hb[h] += cf * pw[p3 & 3] * pw[3 + ((p3 >> 2) & 3)] * pw[6 + ((p3 >> 4) & 3)] * acc[e * nf + f]
    out = np.zeros(na * nb * nc * nd)
    for o in range(out.size):
        ab, cdk = divmod(o, nc * nd)
        for u in range(4):
            f, p3, cf = ket[cdk][u]
            packed = f | (p3 << 12) | (cf << 18)
            assert (packed & 0xfff, (packed >> 12) & 63, (packed >> 18) & 15) == (f, p3, cf)
            out[o] += cf * pw[9 + (p3 & 3)] * pw[12 + ((p3 >> 2) & 3)] * pw[15 + ((p3 >> 4) & 3)] * hb[ab * nf + f]
    return out.reshape(na, nb, nc, nd)


@pytest.mark.parametrize("cls", gen_eri.TEAM_CLASSES)
def test_team_program_matches_oracle(golden, cls):
    fx = golden("water2")
    o = Oracle(fx)
    # Qshells by angular momentum on the two waters' atoms
    by_L = {0: [], 1: [], 2: []}
    for s in range(fx.nshell):
        for L in range(fx.Qstart[s], fx.Qfinal[s] + 1):
            by_L[L].append((s, L))
    rng = np.random.default_rng(sum(cls))
    for trial in range(2):
        sh = [by_L[l][rng.integers(len(by_L[l]))] for l in cls]
        data = [qshell_data(fx, s, L) for s, L in sh]
        Mb = sum(fx.Qfinal[s] for s, _ in sh)
        got = run_program(cls, *data, Mb)
        f0 = [d[3] for d in data]
        for idx in [(0, 0, 0, 0), tuple(x - 1 for x in got.shape), (1, 0, 2 % got.shape[2], 0)]:
            i, j, k, l = (f0[t] + idx[t] for t in range(4))
            ref = o.eri(i, j, k, l) / (fx.cons[i] * fx.cons[j] * fx.cons[k] * fx.cons[l])
            assert abs(got[idx] - ref) <= 1e-11 * max(1.0, abs(ref)), (cls, sh, idx, got[idx], ref)


def test_bench_names_the_team_kernels():
    """bench.py labels each class's kernel by name (for the rocprofv3 figures);
    the classes it calls team_kernel are exactly the generator's team classes."""
    sys.path.insert(0, str(ROOT))
    import bench
    for cls in gen_eri.quartet_classes():
        want = "team_kernel" if cls in gen_eri.TEAM_CLASSES else "fock_kernel"
        assert bench.kernel_of(gen_eri.tag(cls)) == want, cls
