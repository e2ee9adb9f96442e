"""ctypes binding of oracle/liboracle.so (the C restatement in qk_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the thing measured
or shipped.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build(force: bool = False) -> Path:
    srcs = [HERE / "qk_oracle.c", HERE / "qk_xc_oracle.c", HERE / "qk_oneint_oracle.c"]
    if force or not LIB.exists() or any(LIB.stat().st_mtime < s.stat().st_mtime for s in srcs):
        subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-o", str(LIB),
                        *map(str, srcs), "-lm"], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(str(LIB))
        P = C.c_void_p
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.qko_create.restype = P
        L.qko_create.argtypes = [C.c_int] * 4 + [dp] + [ip] * 10 + [dp] * 3
        L.qko_destroy.argtypes = [P]
        L.qko_set_boys_exact.argtypes = [P, C.c_int]
        L.qko_fmt.argtypes = [C.c_int, C.c_double, dp]
        L.qko_fmt_exact.argtypes = [C.c_int, C.c_double, dp]
        L.qko_schwarz.argtypes = [P, dp, dp]
        L.qko_set_cutoffs.argtypes = [P, dp, dp]
        L.qko_density_cutoff.argtypes = [P, dp, dp]
        L.qko_fock.restype = C.c_long
        L.qko_fock.argtypes = [P, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                               C.POINTER(C.c_long)]
        L.qko_eri.restype = C.c_double
        L.qko_eri.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int]
        L.qko_grad.restype = C.c_long
        L.qko_grad.argtypes = [P, dp, dp, C.c_double, C.c_double, C.POINTER(C.c_long)]
        dptr = C.POINTER(C.c_double)
        L.qkx_b3lyp.argtypes = [C.c_double, C.c_double, dptr, dptr, dptr]
        L.qkx_sigrad2.argtypes = [C.c_int, ip, dp, ip, C.c_double, dp]
        L.qkx_sg1_grid.restype = C.c_long
        L.qkx_sg1_grid.argtypes = [C.c_int, dp, ip, ip, C.POINTER(dptr), C.c_double, dp, C.c_long]
        L.qkx_xc.argtypes = [C.c_int, dp, C.c_int, ip, ip, ip, dp, dp, dp, dp, C.c_long, dp, C.c_double, dp,
                             dptr, dptr, dptr]
        L.qko1_hcore.argtypes = [C.c_int, dp, dp, C.c_int, C.c_int, ip, ip, ip, ip, dp, dp, dp, dp]
        L.qko1_grad.argtypes = [C.c_int, dp, dp, C.c_int, C.c_int, ip, ip, ip, ip, dp, dp, dp, dp, dp]
        _lib = L
    return _lib


def _f(a, dt):
    """Fortran-ordered array flattened to a C-contiguous buffer."""
    return np.ascontiguousarray(np.asarray(a, dt).ravel(order="F"))


def fmt(M: int, T: float, exact: bool = False) -> np.ndarray:
    out = np.zeros(M + 1)
    (lib().qko_fmt_exact if exact else lib().qko_fmt)(M, float(T), out)
    return out


class Oracle:
    """Holds one system's shell tables (any object with the Basis attributes)."""

    def __init__(self, b):
        L = lib()
        self.nbasis, self.nshell, self.jbasis = int(b.nbasis), int(b.nshell), int(b.jbasis)
        self.natom = int(np.asarray(b.xyz).shape[1])
        self._keep = [_f(b.xyz, np.float64)] + [_f(getattr(b, n), np.int32) for n in
                      ("kstart", "kprim", "katom", "ktype", "Qstart", "Qfinal", "Qsbasis",
                       "Qfbasis", "ksumtype", "KLMN")] + [_f(getattr(b, n), np.float64) for n in
                                                          ("cons", "gccoeff", "gcexpo")]
        self.h = L.qko_create(self.natom, self.nshell, self.nbasis, self.jbasis, *self._keep)

    def __del__(self):
        if getattr(self, "h", None):
            lib().qko_destroy(self.h)
            self.h = None

    def boys_exact(self, on: bool):
        lib().qko_set_boys_exact(self.h, int(on))

    def schwarz(self):
        Y = np.zeros(self.nshell * self.nshell)
        cp = np.zeros(self.jbasis * self.jbasis)
        lib().qko_schwarz(self.h, Y, cp)
        return Y.reshape(self.nshell, self.nshell, order="F"), cp.reshape(self.jbasis, self.jbasis, order="F")

    def set_cutoffs(self, Y, cutprim):
        lib().qko_set_cutoffs(self.h, _f(Y, np.float64), _f(cutprim, np.float64))

    def fock(self, D, tau=1e-11, prim_limit=1e-20, pair_cut=1e-10, hybrid=1.0, mode="gpu"):
        """Two-electron operator G(D).  mode 'cpu' follows hfoperator's get2e
        enumeration; 'gpu' the reference GPU pair list (pair_cut on Ycutoff) with
        the GPU's Boys order; 'gpu_cpuboys' the same list with the CPU's Boys order."""
        G = np.zeros(self.nbasis * self.nbasis)
        npq = C.c_long(0)
        nq = lib().qko_fock(self.h, _f(D, np.float64), G, tau, prim_limit, pair_cut, hybrid,
                            {"cpu": 0, "gpu": 1, "gpu_cpuboys": 2}[mode], C.byref(npq))
        return G.reshape(self.nbasis, self.nbasis, order="F"), int(nq), int(npq.value)

    def grad(self, D, cutoff=1e-11, hybrid=1.0):
        """Four-centre gradient dE2/dR (3 x natom) at fixed D, E2 = 1/2 sum D G(D),
        with hfgrad's enumeration and screening (cutoff = gradCutoff)."""
        g = np.zeros(3 * self.natom)
        npq = C.c_long(0)
        nq = lib().qko_grad(self.h, _f(D, np.float64), g, cutoff, hybrid, C.byref(npq))
        return g.reshape(3, self.natom, order="F"), int(nq), int(npq.value)

    def eri(self, i, j, k, l):
        return lib().qko_eri(self.h, i, j, k, l)

    def density_cutoff(self, D):
        cm = np.zeros(self.nshell * self.nshell)
        lib().qko_density_cutoff(self.h, _f(D, np.float64), cm)
        return cm.reshape(self.nshell, self.nshell, order="F")


class Fixture:
    """A golden .npz as an object with the Basis attribute names."""

    def __init__(self, path):
        z = np.load(path)
        self.__dict__.update({k: z[k] for k in z.files})
        self.nbasis, self.nshell, self.jbasis = int(self.nbasis), int(self.nshell), int(self.jbasis)


# ---------------------------------------------------------------- XC (B3LYP)

def b3lyp(rho: float, sigma: float):
    """Closed-shell B3LYP f(rho, sigma), df/drho, df/dsigma (qkx_b3lyp)."""
    f, vr, vs = C.c_double(), C.c_double(), C.c_double()
    lib().qkx_b3lyp(float(rho), float(sigma), C.byref(f), C.byref(vr), C.byref(vs))
    return f.value, vr.value, vs.value


def _xc_arrays(b):
    itype = np.ascontiguousarray(np.asarray(b.KLMN, np.int32).T.ravel())   # (l, m, n) per function
    return (itype, np.ascontiguousarray(b.ncontract, np.int32), np.ascontiguousarray(np.asarray(b.ncenter, np.int32)),
            _f(b.dcoeff, np.float64), _f(b.aexp, np.float64))


def sigrad2(b, cutoff: float = 1e-10) -> np.ndarray:
    itype, ncon, _, _, aexp = _xc_arrays(b)
    out = np.zeros(int(b.nbasis))
    lib().qkx_sigrad2(int(b.nbasis), ncon, aexp, itype, cutoff, out)
    return out


def lebedev_tables(orders=(3, 9, 15, 23)):
    """Lebedev rules (x, y, z, w; sum w = 4 pi) of 6, 38, 86, 194 points (scipy)."""
    from scipy.integrate import lebedev_rule
    out = []
    for o in orders:
        x, w = lebedev_rule(o)
        out.append(np.ascontiguousarray(np.vstack([x, w[None, :]]).T))
    return out


def sg1_grid(xyz_bohr, Z, cut: float = 1e-10) -> np.ndarray:
    """The SG-1 grid (npts x 5: x, y, z, w, parent atom), qkx_sg1_grid."""
    xyz = np.ascontiguousarray(np.asarray(xyz_bohr, np.float64).T.ravel())   # (natom, 3)
    Z = np.ascontiguousarray(Z, np.int32)
    tabs = lebedev_tables()
    ns = np.array([len(t) for t in tabs], np.int32)
    arr = (C.POINTER(C.c_double) * 4)(*[t.ctypes.data_as(C.POINTER(C.c_double)) for t in tabs])
    cap = len(Z) * 50 * 194
    out = np.zeros(cap * 5)
    n = lib().qkx_sg1_grid(len(Z), xyz, Z, ns, arr, cut, out, cap)
    return out[:5 * n].reshape(n, 5)


def xc(b, D, pts, sig2=None, dmcut: float = 1e-10):
    """B3LYP V_xc (n x n), E_xc, alpha and beta electron counts on the grid
    pts (npts x >=4: x, y, z, w), restating dftoperator's XC loop."""
    itype, ncon, ncen, dco, aexp = _xc_arrays(b)
    n = int(b.nbasis)
    sig2 = sigrad2(b) if sig2 is None else np.ascontiguousarray(sig2, np.float64)
    p5 = np.zeros((len(pts), 5))
    p5[:, :4] = np.asarray(pts)[:, :4]
    V = np.zeros(n * n)
    E, a, bb = C.c_double(), C.c_double(), C.c_double()
    xyz = np.ascontiguousarray(np.asarray(b.xyz, np.float64).T.ravel())
    lib().qkx_xc(len(b.symbols) if hasattr(b, "symbols") else xyz.size // 3, xyz, n, ncen, itype, ncon, dco, aexp,
                 sig2, _f(D, np.float64), len(p5), np.ascontiguousarray(p5.ravel()), dmcut, V, C.byref(E),
                 C.byref(a), C.byref(bb))
    return V.reshape(n, n, order="F"), E.value, a.value, bb.value


# ---------------------------------------------------------------- one-electron


def _oneint_arrays(b):
    """per-function arrays of a Basis: atoms, (l, m, n), contraction lengths, the
    Qfinal of each function's shell, exponents and coefficients (10 x n)"""
    n = int(b.nbasis)
    itype = np.ascontiguousarray(np.asarray(b.KLMN, np.int32).T.ravel())
    qf = np.zeros(n, np.int32)
    ks = np.asarray(b.ksumtype)
    for s in range(len(b.Qfinal)):
        qf[ks[s] - 1:ks[s + 1] - 1] = b.Qfinal[s]
    Z = np.asarray(b.Z, np.float64)
    xyz = _f(np.asarray(b.xyz, np.float64), np.float64)
    maxc = int(np.asarray(b.dcoeff).shape[0])
    return (len(Z), xyz, np.ascontiguousarray(Z), n, maxc, np.ascontiguousarray(b.ncenter, np.int32), itype,
            np.ascontiguousarray(b.ncontract, np.int32), qf, _f(b.aexp, np.float64), _f(b.dcoeff, np.float64))


def hcore(b):
    """S and H = T + V (qko1_hcore)"""
    args = _oneint_arrays(b)
    n = args[3]
    S, H = np.zeros(n * n), np.zeros(n * n)
    lib().qko1_hcore(*args, S, H)
    return S.reshape(n, n, order="F"), H.reshape(n, n, order="F")


def oneint_grad(b, D, W):
    """nuclear repulsion + one-electron part of dE/dR (3 x natom), qko1_grad"""
    args = _oneint_arrays(b)
    g = np.zeros(3 * args[0])
    lib().qko1_grad(*args, _f(D, np.float64), _f(W, np.float64), g)
    return g.reshape(3, args[0], order="F")
