"""ctypes binding of the library's exchange-correlation step (qk_xc_*).

``XcEngine`` mirrors QUICK's XC inputs: the atoms and the Cartesian basis
functions as gpu_upload_basis_ carries them to gpu_getxc_ (gpu.cu:1070-1390,
1866-1975), a grid (SG-1 as QUICK's default, or a radial x Lebedev product
grid), and per SCF cycle the density.  The library is required; there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from quick_amd.fock import _check, lib

QK_GRID_SG1, QK_GRID_PRODUCT = 1, 2


class XcOpts(C.Structure):
    _fields_ = [("grid", C.c_int), ("nrad", C.c_int), ("nang", C.c_int), ("weight_cut", C.c_double),
                ("basis_cut", C.c_double), ("density_cut", C.c_double)]


_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    P = C.c_void_p
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.qk_xc_default_opts.argtypes = [C.POINTER(XcOpts)]
    L.qk_xc_create.restype = P
    L.qk_xc_create.argtypes = [C.c_int]
    L.qk_xc_destroy.argtypes = [P]
    L.qk_xc_set_system.argtypes = [P, C.c_int, dp, ip, C.c_int, C.c_int, ip, ip, dp, dp, ip, dp, C.c_double]
    L.qk_xc_sigrad2.argtypes = [P, dp]
    L.qk_xc_build_grid.argtypes = [P, C.POINTER(XcOpts)]
    L.qk_xc_set_grid.argtypes = [P, C.c_longlong, dp]
    L.qk_xc_grid_size.restype = C.c_longlong
    L.qk_xc_grid_size.argtypes = [P]
    L.qk_xc_get_grid.argtypes = [P, dp, ip]
    L.qk_xc_batch_info.argtypes = [P, ip, dp, dp]
    L.qk_xc_device.argtypes = [P, C.c_void_p, C.c_void_p, dp, dp, C.c_void_p]
    L.qk_xc_host.argtypes = [P, dp, dp, dp, dp]
    _bound = True
    return L


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def make_xc_opts(**kw) -> XcOpts:
    o = XcOpts()
    _bind().qk_xc_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class XcEngine:
    """B3LYP XC of one system on one device."""

    def __init__(self, b, device: int = 0, sigrad2=None, basis_cut: float = 1e-10):
        L = _bind()
        self.h = L.qk_xc_create(device)
        if not self.h:
            raise RuntimeError(f"qk_xc_create failed: {L.qk_last_error().decode()}")
        self.nbasis = int(b.nbasis)
        xyz = np.ascontiguousarray(np.asarray(b.xyz, np.float64).ravel(order="F"))
        Z = np.ascontiguousarray(b.Z, np.int32)
        itype = np.ascontiguousarray(np.asarray(b.KLMN, np.int32).ravel(order="F"))
        ncon = np.ascontiguousarray(b.ncontract, np.int32)
        ncen = np.ascontiguousarray(b.ncenter, np.int32)
        maxc = int(b.dcoeff.shape[0])
        dco = np.ascontiguousarray(np.asarray(b.dcoeff, np.float64).ravel(order="F"))
        aex = np.ascontiguousarray(np.asarray(b.aexp, np.float64).ravel(order="F"))
        s2 = None if sigrad2 is None else np.ascontiguousarray(sigrad2, np.float64)
        self._keep = (xyz, Z, itype, ncon, ncen, dco, aex, s2)
        _check(L.qk_xc_set_system(self.h, len(Z), _p(xyz, C.c_double), _p(Z, C.c_int), self.nbasis, maxc,
                                  _p(ncon, C.c_int), _p(itype, C.c_int), _p(aex, C.c_double), _p(dco, C.c_double),
                                  _p(ncen, C.c_int), None if s2 is None else _p(s2, C.c_double), basis_cut),
               "qk_xc_set_system")

    def __del__(self):
        if getattr(self, "h", None):
            lib().qk_xc_destroy(self.h)
            self.h = None

    def sigrad2(self) -> np.ndarray:
        out = np.zeros(self.nbasis)
        _check(_bind().qk_xc_sigrad2(self.h, _p(out, C.c_double)), "qk_xc_sigrad2")
        return out

    def build_grid(self, opts: XcOpts | None = None) -> int:
        opts = opts or make_xc_opts()
        _check(_bind().qk_xc_build_grid(self.h, C.byref(opts)), "qk_xc_build_grid")
        return self.grid_size()

    def set_grid(self, pts) -> int:
        a = np.ascontiguousarray(np.asarray(pts, np.float64)[:, :4])
        _check(_bind().qk_xc_set_grid(self.h, len(a), _p(a, C.c_double)), "qk_xc_set_grid")
        return self.grid_size()

    def grid_size(self) -> int:
        return int(_bind().qk_xc_grid_size(self.h))

    def grid(self):
        n = self.grid_size()
        pts = np.zeros((n, 4))
        atom = np.zeros(n, np.int32)
        _check(_bind().qk_xc_get_grid(self.h, _p(pts, C.c_double), _p(atom, C.c_int)), "qk_xc_get_grid")
        return pts, atom

    def batch_info(self):
        nb, mf, fl = C.c_int(), C.c_double(), C.c_double()
        _check(_bind().qk_xc_batch_info(self.h, C.byref(nb), C.byref(mf), C.byref(fl)), "qk_xc_batch_info")
        return nb.value, mf.value, fl.value

    def xc(self, D):
        """V_xc (n x n), E_xc and the electron count of the density D (host)."""
        Df = np.ascontiguousarray(np.asarray(D, np.float64).ravel(order="F"))
        V = np.zeros(self.nbasis * self.nbasis)
        E, N = C.c_double(), C.c_double()
        _check(_bind().qk_xc_host(self.h, _p(Df, C.c_double), _p(V, C.c_double), C.byref(E), C.byref(N)),
               "qk_xc_host")
        return V.reshape(self.nbasis, self.nbasis, order="F"), E.value, N.value

    def xc_device(self, D_ptr: int, V_ptr: int, stream: int = 0, want_scalars: bool = True):
        E, N = C.c_double(), C.c_double()
        _check(_bind().qk_xc_device(self.h, C.c_void_p(D_ptr), C.c_void_p(V_ptr),
                                    C.byref(E) if want_scalars else None, C.byref(N) if want_scalars else None,
                                    C.c_void_p(stream)), "qk_xc_device")
        return E.value, N.value
