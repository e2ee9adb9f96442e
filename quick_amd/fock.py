"""ctypes binding of libquick_mi355x.so: the host-side mirror of QUICK's
Fock-build boundary.

``FockEngine`` mirrors the sequence QUICK's Fortran drives through the GPU
boundary (main.f90:139-177, hfoperator.f90:78-103): upload the system
(gpu_upload_xyz_/gpu_upload_basis_), upload or compute the Schwarz bounds
(gpu_upload_cutoff_matrix_), then per SCF cycle build G(D) (gpu_get2e_).  The
explicit device-pointer entry points let torch tensors stay resident in HBM.

The library is required: there is no CPU fallback, and loading fails loudly if
it has not been built (``python -m quick_amd.build``).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# the in-tree build; QUICK_MI355X_LIB selects another build of the same library
LIB_PATH = Path(os.environ.get("QUICK_MI355X_LIB", Path(__file__).resolve().parent / "libquick_mi355x.so"))

QK_FP64, QK_MIXED = 0, 1
QK_BOYS_ORDER_SHELL, QK_BOYS_ORDER_QSHELL = 0, 1


class QkOpts(C.Structure):
    _fields_ = [("tau", C.c_double), ("prim_limit", C.c_double), ("pair_cut", C.c_double),
                ("hybrid", C.c_double), ("boys_order", C.c_int), ("precision", C.c_int),
                ("fp32_threshold", C.c_double)]


class QkScfOpts(C.Structure):
    _fields_ = [("max_cycles", C.c_int), ("max_diis", C.c_int), ("dens_rms", C.c_double), ("fock", QkOpts)]


class QkStats(C.Structure):
    _fields_ = [("quartets", C.c_longlong), ("quartets_fp32", C.c_longlong),
                ("prim_quartets", C.c_longlong), ("flops_fp64", C.c_double),
                ("flops_fp32", C.c_double)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        # One HIP runtime per process: PyTorch's ROCm wheel carries its own
        # libamdhip64.so.7; if torch is present it is loaded first so that this
        # library binds to the same runtime instead of a second copy.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m quick_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.qk_last_error.restype = C.c_char_p
        L.qk_create.restype = P
        L.qk_create.argtypes = [C.c_int]
        L.qk_destroy.argtypes = [P]
        L.qk_default_opts.argtypes = [C.POINTER(QkOpts)]
        L.qk_set_system.argtypes = [P, C.c_int, dp, C.c_int, C.c_int, C.c_int] + [ip] * 10 + [dp] * 3
        L.qk_set_schwarz.argtypes = [P, dp, dp]
        L.qk_compute_schwarz.argtypes = [P, dp, dp]
        L.qk_prepare.argtypes = [P, C.POINTER(QkOpts)]
        L.qk_fock_device.argtypes = [P, C.c_void_p, C.c_void_p, C.POINTER(QkOpts), C.c_void_p]
        L.qk_fock_device_noalloc.argtypes = [P, C.c_void_p, C.c_void_p, C.POINTER(QkOpts), C.c_void_p]
        L.qk_set_charges.argtypes = [P, dp]
        L.qk_scf_default_opts.argtypes = [C.POINTER(QkScfOpts)]
        L.qk_scf.argtypes = [P, C.c_int, C.POINTER(QkScfOpts), dp, C.POINTER(C.c_double), C.POINTER(C.c_int),
                             C.POINTER(C.c_double)]
        L.qk_oneint_device.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p]
        L.qk_nuclear_repulsion.argtypes = [P, dp]
        L.qk_oneint_grad_device.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.qk_oneint_host.argtypes = [P, dp, dp]
        L.qk_oneint_grad_host.argtypes = [P, dp, dp, dp]
        L.qk_fock_host.argtypes = [P, dp, dp, C.POINTER(QkOpts)]
        L.qk_grad_device.argtypes = [P, C.c_void_p, C.c_void_p, C.POINTER(QkOpts), C.c_void_p]
        L.qk_grad_host.argtypes = [P, dp, dp, C.POINTER(QkOpts)]
        L.qk_last_stats.argtypes = [P, C.POINTER(QkStats)]
        L.qk_check_range.argtypes = [P]
        L.qk_check_range.restype = C.c_int
        L.qk_num_classes.restype = C.c_int
        L.qk_class_info.argtypes = [P, C.c_int, ip, C.POINTER(C.c_longlong)]
        L.qk_set_timing.argtypes = [P, C.c_int]
        L.qk_class_lane_fill.restype = C.c_double
        L.qk_class_lane_fill.argtypes = [P, C.c_int]
        L.qk_class_stats.argtypes = [P, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                     C.POINTER(C.c_double), C.POINTER(C.c_float)]
        L.qk_grad_class_stats.argtypes = [P, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                          C.POINTER(C.c_float)]
        L.qk_fock_build.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(QkOpts)]
        L.qk_fock_grad.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(QkOpts)]
        L.qk_grad_class_stats_fp32.argtypes = [P, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                               C.POINTER(C.c_float)]
        L.qk_class_stats_fp32.argtypes = [P, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                          C.POINTER(C.c_double), C.POINTER(C.c_float)]
        _lib = L
    return _lib


def last_error() -> str:
    return lib().qk_last_error().decode()


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


def _f(a, dt):
    return np.ascontiguousarray(np.asarray(a, dt).ravel(order="F"))


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def make_opts(tau=1e-11, prim_limit=1e-20, pair_cut=0.0, hybrid=1.0,
              boys_order=QK_BOYS_ORDER_SHELL, precision=QK_FP64, fp32_threshold=1e-5) -> QkOpts:
    return QkOpts(tau, prim_limit, pair_cut, hybrid, boys_order, precision, fp32_threshold)


class FockEngine:
    """One system on one device.  ``basis`` has QUICK's shell-table attributes
    (quick_amd.basis.Basis, or a golden fixture)."""

    def __init__(self, basis, device: int = 0):
        L = lib()
        self.h = L.qk_create(device)
        if not self.h:
            raise RuntimeError(f"qk_create failed: {last_error()}")
        self.nbasis = int(basis.nbasis)
        self.nshell = int(basis.nshell)
        self.jbasis = int(basis.jbasis)
        self.natom = int(np.asarray(basis.xyz).shape[1])
        keep = [_f(basis.xyz, np.float64)]
        keep += [_f(getattr(basis, k), np.int32) for k in
                 ("kstart", "kprim", "katom", "ktype", "Qstart", "Qfinal", "Qsbasis", "Qfbasis",
                  "ksumtype", "KLMN")]
        keep += [_f(getattr(basis, k), np.float64) for k in ("cons", "gccoeff", "gcexpo")]
        args = [_ptr(keep[0], C.c_double)] + [_ptr(a, C.c_int) for a in keep[1:11]] + \
               [_ptr(a, C.c_double) for a in keep[11:]]
        _check(L.qk_set_system(self.h, self.natom, args[0], self.nshell, self.nbasis, self.jbasis,
                               *args[1:]), "qk_set_system")

    def close(self):
        if getattr(self, "h", None):
            lib().qk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_schwarz(self, Y, cutprim):
        Yf, cf = _f(Y, np.float64), _f(cutprim, np.float64)
        _check(lib().qk_set_schwarz(self.h, _ptr(Yf, C.c_double), _ptr(cf, C.c_double)), "qk_set_schwarz")

    def compute_schwarz(self):
        Y = np.zeros(self.nshell * self.nshell)
        cp = np.zeros(self.jbasis * self.jbasis)
        _check(lib().qk_compute_schwarz(self.h, _ptr(Y, C.c_double), _ptr(cp, C.c_double)),
               "qk_compute_schwarz")
        return Y.reshape(self.nshell, self.nshell, order="F"), cp.reshape(self.jbasis, self.jbasis, order="F")

    def prepare(self, opts: QkOpts | None = None):
        opts = opts or make_opts()
        _check(lib().qk_prepare(self.h, C.byref(opts)), "qk_prepare")

    def fock_device(self, D_ptr: int, G_ptr: int, opts: QkOpts | None = None, stream: int = 0):
        """Enqueue G <- G(D) on ``stream``; D and G are device pointers to n*n doubles."""
        opts = opts or make_opts()
        _check(lib().qk_fock_device(self.h, C.c_void_p(D_ptr), C.c_void_p(G_ptr), C.byref(opts),
                                    C.c_void_p(stream)), "qk_fock_device")

    def fock_device_noalloc(self, D_ptr: int, G_ptr: int, opts: QkOpts | None = None, stream: int = 0) -> int:
        """qk_fock_device with the per-build scratch in the library's fixed device
        array (no allocation); the work lists must be prepared for ``opts``.
        Returns the entry point's code (0, or -6 / -7 with last_error())."""
        opts = opts or make_opts()
        return int(lib().qk_fock_device_noalloc(self.h, C.c_void_p(D_ptr), C.c_void_p(G_ptr), C.byref(opts),
                                                C.c_void_p(stream)))

    def set_charges(self, charges):
        """nuclear charges (natom) for the one-electron integrals"""
        self._chg = np.ascontiguousarray(np.asarray(charges, np.float64))
        _check(lib().qk_set_charges(self.h, _ptr(self._chg, C.c_double)), "qk_set_charges")

    def scf(self, nocc: int, fock_opts: QkOpts | None = None, dens_rms: float = 1e-7, max_cycles: int = 100,
            require_convergence: bool = True):
        """RHF SCF on the device (qk_scf): returns (D, E, cycles, last RMS change of D);
        an unconverged run at the cycle limit raises unless require_convergence is False"""
        o = QkScfOpts()
        lib().qk_scf_default_opts(C.byref(o))
        o.dens_rms, o.max_cycles = dens_rms, max_cycles
        if fock_opts is not None:
            o.fock = fock_opts
        D = np.zeros(self.nbasis * self.nbasis)
        E, cyc, rms = C.c_double(), C.c_int(), C.c_double()
        rc = lib().qk_scf(self.h, int(nocc), C.byref(o), _ptr(D, C.c_double), C.byref(E), C.byref(cyc),
                          C.byref(rms))
        if not (rc == -4 and not require_convergence):
            _check(rc, "qk_scf")
        return D.reshape(self.nbasis, self.nbasis, order="F"), E.value, cyc.value, rms.value

    def oneint(self):
        """S and H_core = T + V (n x n, host)"""
        n = self.nbasis
        S, H = np.zeros(n * n), np.zeros(n * n)
        _check(lib().qk_oneint_host(self.h, _ptr(S, C.c_double), _ptr(H, C.c_double)), "qk_oneint_host")
        return S.reshape(n, n, order="F"), H.reshape(n, n, order="F")

    def nuclear_repulsion(self) -> float:
        e = C.c_double()
        _check(lib().qk_nuclear_repulsion(self.h, C.byref(e)), "qk_nuclear_repulsion")
        return e.value

    def oneint_grad(self, D, W) -> np.ndarray:
        """nuclear repulsion + one-electron part of dE/dR (3 x natom) for the
        density D and energy-weighted density W"""
        Df, Wf = _f(D, np.float64), _f(W, np.float64)
        g = np.zeros(3 * self.natom)
        _check(lib().qk_oneint_grad_host(self.h, _ptr(Df, C.c_double), _ptr(Wf, C.c_double), _ptr(g, C.c_double)),
               "qk_oneint_grad_host")
        return g.reshape(3, self.natom, order="F")

    def fock(self, D, opts: QkOpts | None = None) -> np.ndarray:
        """Host arrays: returns G(D) (n x n)."""
        opts = opts or make_opts()
        Df = _f(D, np.float64)
        F = np.zeros(self.nbasis * self.nbasis)
        _check(lib().qk_fock_host(self.h, _ptr(Df, C.c_double), _ptr(F, C.c_double), C.byref(opts)),
               "qk_fock_host")
        return F.reshape(self.nbasis, self.nbasis, order="F")

    def grad_device(self, D_ptr: int, grad_ptr: int, opts: QkOpts | None = None, stream: int = 0):
        """Enqueue grad <- grad + dE2/dR on ``stream`` (device pointers: D n*n, grad 3*natom)."""
        opts = opts or make_opts()
        _check(lib().qk_grad_device(self.h, C.c_void_p(D_ptr), C.c_void_p(grad_ptr), C.byref(opts),
                                    C.c_void_p(stream)), "qk_grad_device")

    def grad(self, D, opts: QkOpts | None = None) -> np.ndarray:
        """Four-centre gradient dE2/dR (3 x natom) at fixed D, E2 = 1/2 sum D G(D);
        opts.tau is the gradient cutoff."""
        opts = opts or make_opts()
        Df = _f(D, np.float64)
        g = np.zeros(3 * self.natom)
        _check(lib().qk_grad_host(self.h, _ptr(Df, C.c_double), _ptr(g, C.c_double), C.byref(opts)),
               "qk_grad_host")
        return g.reshape(3, self.natom, order="F")

    def grad_class_stats(self, fp32: bool = False):
        """[(tag, quartets, prim_quartets, ms)] of the last gradient (ms when timing
        is on); fp32: the FP32 part of a mixed-precision gradient."""
        out = []
        L = lib()
        fn = L.qk_grad_class_stats_fp32 if fp32 else L.qk_grad_class_stats
        l4 = (C.c_int * 4)()
        w = C.c_longlong()
        nq, npq, ms = C.c_longlong(), C.c_longlong(), C.c_float()
        for k in range(L.qk_num_classes()):
            _check(L.qk_class_info(self.h, k, l4, C.byref(w)), "qk_class_info")
            _check(fn(self.h, k, C.byref(nq), C.byref(npq), C.byref(ms)), "qk_grad_class_stats")
            out.append(("".join("spd"[x] for x in l4), nq.value, npq.value, ms.value))
        return out

    def stats(self) -> QkStats:
        s = QkStats()
        _check(lib().qk_last_stats(self.h, C.byref(s)), "qk_last_stats")
        return s

    def set_timing(self, on: bool = True):
        _check(lib().qk_set_timing(self.h, int(on)), "qk_set_timing")

    def class_stats(self):
        """[(tag, quartets, prim_quartets, flops, ms, lane_fill)] of the FP64 kernels in the last build."""
        out = []
        L = lib()
        l4 = (C.c_int * 4)()
        w = C.c_longlong()
        nq, npq, fl, ms = C.c_longlong(), C.c_longlong(), C.c_double(), C.c_float()
        for k in range(L.qk_num_classes()):
            _check(L.qk_class_info(self.h, k, l4, C.byref(w)), "qk_class_info")
            _check(L.qk_class_stats(self.h, k, C.byref(nq), C.byref(npq), C.byref(fl), C.byref(ms)),
                   "qk_class_stats")
            out.append(("".join("spd"[x] for x in l4), nq.value, npq.value, fl.value, ms.value,
                        L.qk_class_lane_fill(self.h, k)))
        return out

    def class_stats_fp32(self):
        """[(tag, quartets, prim_quartets, flops, ms)] of the FP32 kernels in the last build."""
        out = []
        L = lib()
        l4 = (C.c_int * 4)()
        w = C.c_longlong()
        nq, npq, fl, ms = C.c_longlong(), C.c_longlong(), C.c_double(), C.c_float()
        for k in range(L.qk_num_classes()):
            _check(L.qk_class_info(self.h, k, l4, C.byref(w)), "qk_class_info")
            _check(L.qk_class_stats_fp32(self.h, k, C.byref(nq), C.byref(npq), C.byref(fl), C.byref(ms)),
                   "qk_class_stats_fp32")
            out.append(("".join("spd"[x] for x in l4), nq.value, npq.value, fl.value, ms.value))
        return out

    def class_work(self):
        out = []
        L = lib()
        l4 = (C.c_int * 4)()
        w = C.c_longlong()
        for k in range(L.qk_num_classes()):
            _check(L.qk_class_info(self.h, k, l4, C.byref(w)), "qk_class_info")
            out.append((tuple(l4), int(w.value)))
        return out
