"""QUICK's Fortran call sequence through the gpu_*_ symbols, run as a script.

The tests start this in a child process, so that a fatal error inside the
library (which aborts, as the reference's exit(1) does) fails one test with its
stderr instead of ending the test session.

It drives, in order, the set-up of main.f90:86-177 (gpu_startup_,
gpu_set_device_, gpu_init_, gpu_setup_, gpu_upload_xyz_,
gpu_upload_atom_and_chg_, gpu_upload_basis_, gpu_upload_cutoff_matrix_), the
Fock build of hfoperator.f90:77-103 (gpu_upload_method_, gpu_upload_calculated_,
gpu_upload_cutoff_ with three arguments, gpu_get2e_) and the gradient of
gradient.f90:279-287 (gpu_upload_calculated_, gpu_upload_cutoff_,
gpu_upload_grad_, gpu_grad_), then gpu_cleanup_ and gpu_shutdown_.

usage: python tests/abi_sequence.py <fixture.npz> <out.npz> [method]
"""
from __future__ import annotations

import ctypes
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def shell_block_max(D, ksumtype, Qsbasis, Qfbasis, Qstart, Qfinal):
    """cutMatrix(II,JJ) = max |D| over the functions of shells II and JJ (the
    caller's DNscreen, schwarz.f90:550-577)."""
    ns = len(Qstart)
    f0 = [ksumtype[s] - 1 + Qsbasis[s, Qstart[s]] for s in range(ns)]
    f1 = [ksumtype[s] - 1 + Qfbasis[s, Qfinal[s]] + 1 for s in range(ns)]
    A = np.abs(D)
    return np.array([[A[f0[i]:f1[i], f0[j]:f1[j]].max() for j in range(ns)] for i in range(ns)])


def main(fixture: str, out: str, method: int = 0) -> None:
    from quick_amd.fock import lib
    z = np.load(fixture)
    natom, nbasis, nshell, jbasis = (int(z[k]) for k in ("natom", "nbasis", "nshell", "jbasis"))
    L = lib()
    i = lambda v: ctypes.byref(ctypes.c_int(v))
    d = lambda v: ctypes.byref(ctypes.c_double(v))
    arr = lambda a, t: np.ascontiguousarray(np.asarray(a, t).ravel(order="F"))
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    print(f"abi_sequence: {Path(fixture).name} nbasis={nbasis} method={method}", flush=True)
    L.gpu_startup_()
    L.gpu_set_device_(i(-1))
    L.gpu_init_()
    L.gpu_setup_(i(natom), i(nbasis), i(10), i(1), i(0), i(0))
    xyz = arr(z["xyz"], np.float64)
    L.gpu_upload_xyz_(P(xyz))
    Z = {"H": 1, "C": 6, "N": 7, "O": 8}
    atype = arr([Z[str(s)] for s in z["symbols"]], np.int32)
    chg = arr([float(Z[str(s)]) for s in z["symbols"]], np.float64)
    L.gpu_upload_atom_and_chg_(P(atype), P(chg))
    ints = {k: arr(z[k], np.int32) for k in ("kstart", "katom", "ktype", "kprim", "ksumtype",
                                             "Qstart", "Qfinal", "Qsbasis", "Qfbasis", "KLMN")}
    dbl = {k: arr(z[k], np.float64) for k in ("gccoeff", "cons", "gcexpo")}
    zi = np.zeros(max(16, 10 * nbasis), np.int32)
    zd = np.zeros(max(16, 10 * nbasis))
    L.gpu_upload_basis_(i(nshell), i(jbasis), i(nshell), i(jbasis), i(10),
                        P(zi), P(zi), P(zd), P(zd), P(zi), P(zi), P(zi), P(zi), P(zi),
                        P(ints["kstart"]), P(ints["katom"]), P(ints["ktype"]), P(ints["kprim"]), P(zi),
                        P(ints["ksumtype"]), P(zi), P(ints["Qstart"]), P(ints["Qfinal"]), P(ints["Qsbasis"]),
                        P(ints["Qfbasis"]), P(dbl["gccoeff"]), P(dbl["cons"]), P(dbl["gcexpo"]), P(ints["KLMN"]))
    Yc, cp = arr(z["Ycutoff"], np.float64), arr(z["cutprim"], np.float64)
    L.gpu_upload_cutoff_matrix_(P(Yc), P(cp))
    D = np.asarray(z["D"], np.float64)
    cm = arr(shell_block_max(D, z["ksumtype"], z["Qsbasis"], z["Qfbasis"], z["Qstart"], z["Qfinal"]), np.float64)
    rng = np.random.default_rng(5)
    H = rng.normal(size=(nbasis, nbasis))
    H = H + H.T
    O = arr(H, np.float64)
    dense = arr(D, np.float64)
    # hfoperator.f90:77-103
    L.gpu_upload_method_(i(method))
    L.gpu_upload_calculated_(P(O), P(zd), P(zd), P(dense))
    L.gpu_upload_cutoff_(P(cm), d(1e-11), d(1e-20))
    L.gpu_get2e_(P(O))
    # gradient.f90:279-287
    seed = rng.normal(size=3 * natom)
    grad = seed.copy()          # gpu_grad_ writes into grad; seed keeps the input
    L.gpu_upload_calculated_(P(O.copy()), P(zd), P(zd), P(dense))
    L.gpu_upload_cutoff_(P(cm), d(1e-11), d(1e-20))
    L.gpu_upload_grad_(P(grad), d(1e-11))
    L.gpu_grad_(P(grad))
    L.gpu_cleanup_()
    L.gpu_shutdown_()
    np.savez(out, H=H, O=O.reshape(nbasis, nbasis, order="F"), seed=seed.reshape(3, natom, order="F"),
             grad=grad.reshape(3, natom, order="F"))
    print("abi_sequence: done", flush=True)


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 0)
