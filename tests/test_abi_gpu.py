"""The reference's Fortran -> C boundary (gpu_*_ symbols) on the device.

Each call sequence runs in a child process: a fatal error inside the library
aborts the process (as the reference's exit(1) does), and the child's return
code and stderr then fail one test instead of ending the session.

Two callers drive the same sequence (set-up of main.f90:84-177, Fock build of
hfoperator.f90:77-103, gradient of gradient.f90:279-287):
  * tests/abi_sequence.py through ctypes;
  * tests/fortran/abi_sequence, a Fortran program with implicit interfaces built
    with flang, i.e. exactly how QUICK's own Fortran links and calls the symbols.
The results are compared with the reference CPU path's G(D) and four-centre
gradient (tests/golden/water4.npz) and, for the B3LYP exchange weight, with the
oracle.
"""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from tests.abi_sequence import shell_block_max

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
FORTRAN = ROOT / "tests" / "fortran" / "abi_sequence"
TOL_F = 1e-10
TOL_G = 1e-9


def _run(cmd, timeout=300):
    print("[gpu-test] " + " ".join(map(str, cmd)), flush=True)
    r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True, timeout=timeout, cwd=ROOT)
    assert r.returncode == 0, f"exit {r.returncode}\nstdout:\n{r.stdout[-2000:]}\nstderr:\n{r.stderr[-4000:]}"
    return r


def test_abi_sequence_ctypes_matches_reference(tmp_path):
    out = tmp_path / "out.npz"
    _run([sys.executable, ROOT / "tests" / "abi_sequence.py", GOLDEN / "water4.npz", out])
    z, r = np.load(GOLDEN / "water4.npz"), np.load(out)
    dF = np.abs(r["O"] - (r["H"] + z["G"])).max()
    dg = np.abs(r["grad"] - (r["seed"] + z["grad4c"])).max()
    print(f"ctypes ABI: max|dF| {dF:.2e} max|dgrad| {dg:.2e}")
    assert dF <= TOL_F
    assert dg <= TOL_G


def test_abi_sequence_b3lyp_exchange_weight(tmp_path, golden):
    """gpu_upload_method_(1): K scaled by 0.2 (gpu.cu:269-278) in both the Fock
    build and the gradient, against the oracle at hybrid 0.2."""
    from oracle.oracle import Oracle
    out = tmp_path / "out.npz"
    _run([sys.executable, ROOT / "tests" / "abi_sequence.py", GOLDEN / "water2.npz", out, "1"])
    fx = golden("water2")
    o = Oracle(fx)
    o.set_cutoffs(fx.Ycutoff, fx.cutprim)
    # the oracle's CPU enumeration hard-codes the HF exchange weight, as the
    # reference's CPU iclass does; the GPU enumeration (no pair cut) carries it
    Gref, _, _ = o.fock(fx.D, hybrid=0.2, mode="gpu_cpuboys", pair_cut=0.0)
    gref, _, _ = o.grad(fx.D, hybrid=0.2)
    r = np.load(out)
    assert np.abs(r["O"] - (r["H"] + Gref)).max() <= TOL_F
    assert np.abs(r["grad"] - (r["seed"] + gref)).max() <= TOL_G


def _fortran_input(z, path, method=0, seed=5):
    n, ns = int(z["nbasis"]), int(z["nshell"])
    Zn = {"H": 1, "C": 6, "N": 7, "O": 8}
    at = np.array([Zn[str(s)] for s in z["symbols"]], np.int32)
    D = np.asarray(z["D"], np.float64)
    cm = shell_block_max(D, z["ksumtype"], z["Qsbasis"], z["Qfbasis"], z["Qstart"], z["Qfinal"])
    rng = np.random.default_rng(seed)
    H = rng.normal(size=(n, n))
    H = H + H.T
    g0 = rng.normal(size=(3, int(z["natom"])))
    f = lambda a, t: np.asarray(a, t).ravel(order="F").tobytes()
    with open(path, "wb") as fh:
        fh.write(np.array([int(z["natom"]), n, ns, int(z["jbasis"]), int(z["maxprim"]), 10, method],
                          np.int32).tobytes())
        fh.write(f(z["xyz"], np.float64) + f(at, np.int32) + f(at.astype(np.float64), np.float64))
        for k in ("kstart", "kprim", "katom", "ktype", "Qstart", "Qfinal", "Qsbasis", "Qfbasis", "ksumtype",
                  "KLMN"):
            fh.write(f(z[k], np.int32))
        for k in ("cons", "gccoeff", "gcexpo", "Ycutoff", "cutprim"):
            fh.write(f(z[k], np.float64))
        fh.write(f(D, np.float64) + f(cm, np.float64) + f(H, np.float64) + f(g0, np.float64))
    return H, g0


@pytest.mark.skipif(not FORTRAN.exists(), reason="Fortran test program not built (flang absent)")
def test_abi_sequence_fortran_matches_reference(tmp_path):
    z = np.load(GOLDEN / "water4.npz")
    n, na = int(z["nbasis"]), int(z["natom"])
    H, g0 = _fortran_input(z, tmp_path / "in.bin")
    r = _run([FORTRAN, tmp_path / "in.bin", tmp_path / "out.bin"])
    assert "abi_sequence: done" in r.stdout
    raw = np.fromfile(tmp_path / "out.bin", np.float64)
    O = raw[:n * n].reshape(n, n, order="F")
    g = raw[n * n:].reshape(3, na, order="F")
    dF = np.abs(O - (H + z["G"])).max()
    dg = np.abs(g - (g0 + z["grad4c"])).max()
    print(f"Fortran ABI: max|dF| {dF:.2e} max|dgrad| {dg:.2e}")
    assert dF <= TOL_F
    assert dg <= TOL_G
