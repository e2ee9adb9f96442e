"""The SCF step beside the path (qk_scf): RHF from the core-Hamiltonian guess
with the library's Fock build, one-electron integrals, rocBLAS DGEMMs and the
FP64 rocSOLVER eigensolver, to the reference's convergence (DENSERMS 1e-7).

* (H2O)4 and (H2O)32: the SCF energy against the reference's own SCF
  (tests/golden/*_scf.npz) to 1e-8 Eh, and the converged density;
* (H2O)32 with mixed-precision Fock builds (threshold 1e-3): the energy against
  the FP64 SCF to 1 uEh.
"""
import numpy as np
import pytest

from oracle.oracle import Fixture
from quick_amd.basis import build_basis
from quick_amd.fock import FockEngine, QK_MIXED, make_opts
from quick_amd.geometry import CONFIGS

pytestmark = pytest.mark.gpu

ZOF = {"H": 1.0, "C": 6.0, "N": 7.0, "O": 8.0}


def _banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def test_scf_water4_matches_reference():
    _banner("scf water4")
    z = np.load("tests/golden/water4_scf.npz")
    fx = Fixture("tests/golden/water4_scf.npz")
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    eng.set_charges([ZOF[str(s)] for s in z["symbols"]])
    nocc = z["C_occ"].shape[1]
    D, E, cyc, rms = eng.scf(nocc)
    Dref = 2.0 * z["C_occ"] @ z["C_occ"].T
    print(f"water4: E {E:.10f} ref {float(z['E_total']):.10f} |dE| {abs(E - float(z['E_total'])):.2e} "
          f"cycles {cyc} (reference {int(z['scf_cycles'])}) rms {rms:.1e} max|dD| {np.abs(D - Dref).max():.1e}")
    assert abs(E - float(z["E_total"])) <= 1e-8
    assert np.abs(D - Dref).max() <= 1e-5


def test_scf_water32_fp64_and_mixed():
    _banner("scf water32")
    z = np.load("tests/golden/water32_scf.npz")
    syms, xyz = CONFIGS["water32"]()
    b = build_basis(syms, xyz)
    b.xyz = np.asarray(z["xyz"], np.float64).copy()
    eng = FockEngine(b)
    eng.compute_schwarz()
    eng.set_charges([ZOF[s] for s in syms])
    nocc = z["C_occ"].shape[1]
    D, E, cyc, rms = eng.scf(nocc)
    print(f"water32 FP64: E {E:.10f} |dE| {abs(E - float(z['E_total'])):.2e} cycles {cyc} "
          f"(reference {int(z['scf_cycles'])})", flush=True)
    assert abs(E - float(z["E_total"])) <= 1e-8
    Dm, Em, cm, _ = eng.scf(nocc, make_opts(precision=QK_MIXED, fp32_threshold=1e-3))
    print(f"water32 mixed 1e-3: |dE| vs FP64 SCF {abs(Em - E):.2e}, cycles {cm}", flush=True)
    assert abs(Em - E) <= 1e-6
