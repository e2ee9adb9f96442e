"""The path at converged RHF densities: the reference's own SCF (driver mode
scfgrad, tests/golden/*_scf.npz) ran to DENSERMS 1e-7, and its density
D = 2 C_occ C_occ^T, two-electron operator G(D), one-electron operator H,
energies and four-centre gradient were kept.

* H2O, (H2O)2, (H2O)4: G(D) element-wise (1e-10), the total energy
  E = sum D H + 1/2 sum D G + E_nn against the reference SCF energy (1e-8 Eh,
  H taken from the fixture), and the four-centre gradient (1e-9 Eh/bohr);
* (H2O)32, the benchmark system: Schwarz bounds, 40k seeded samples of G, its
  diagonal and row norms, E2 and the total energy, the FP64 gradient against
  grad4c (1e-9), the mixed-precision build (threshold 1e-3) and gradient
  (threshold 1e-2) against FP64 (1 uEh, 1e-6 Eh/bohr) and the reference.
"""
import numpy as np
import pytest

from oracle.oracle import Fixture
from quick_amd.basis import build_basis
from quick_amd.fock import FockEngine, QK_MIXED, make_opts
from quick_amd.geometry import CONFIGS

pytestmark = pytest.mark.gpu

TOL_F = 1e-10
TOL_E = 1e-8
TOL_G = 1e-9


def _banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


@pytest.mark.parametrize("name", ["h2o", "water2", "water4"])
def test_converged_density_small(name):
    _banner(f"converged density {name}")
    fx = Fixture(f"tests/golden/{name}_scf.npz")
    D = 2.0 * fx.C_occ @ fx.C_occ.T
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    G = eng.fock(D)
    dG = np.abs(G - fx.G).max()
    E = float(np.sum(D * fx.H)) + 0.5 * float(np.sum(D * G)) + float(fx.E_nn)
    g = eng.grad(D)
    dg = np.abs(g - fx.grad4c).max()
    print(f"{name}: max|dG| {dG:.2e} |dE_total| {abs(E - float(fx.E_total)):.2e} max|dgrad4c| {dg:.2e}")
    assert dG <= TOL_F
    assert abs(E - float(fx.E_total)) <= TOL_E
    assert dg <= TOL_G


def test_converged_density_water32():
    _banner("converged density water32")
    z = np.load("tests/golden/water32_scf.npz")
    syms, xyz = CONFIGS["water32"]()
    b = build_basis(syms, xyz)
    b.xyz = np.asarray(z["xyz"], np.float64).copy()
    C = z["C_occ"]
    D = 2.0 * C @ C.T
    eng = FockEngine(b)
    Y, _ = eng.compute_schwarz()
    assert np.abs(Y - z["Ycutoff"]).max() <= 1e-12
    G = eng.fock(D)
    dg = np.abs(G[z["G_sample_i"], z["G_sample_j"]] - z["G_sample"]).max()
    dd = np.abs(np.diag(G) - z["G_diag"]).max()
    dn = np.abs(np.linalg.norm(G, axis=1) - z["G_rownorm"]).max()
    E2 = 0.5 * float(np.sum(D * G))
    E = float(z["E1"]) + E2 + float(z["E_nn"])
    print(f"water32: max|dG| samples {dg:.2e} diag {dd:.2e} rownorm {dn:.2e} |dE2| {abs(E2 - float(z['E2'])):.2e} "
          f"|dE_total| {abs(E - float(z['E_total'])):.2e}", flush=True)
    assert dg <= TOL_F and dd <= TOL_F and dn <= 1e-9
    assert abs(E2 - float(z["E2"])) <= TOL_E
    assert abs(E - float(z["E_total"])) <= TOL_E
    g64 = eng.grad(D)
    dgr = np.abs(g64 - z["grad4c"]).max()
    print(f"water32: FP64 gradient max|d grad4c| {dgr:.2e}", flush=True)
    assert dgr <= TOL_G
    # mixed precision at the converged density
    Gm = eng.fock(D, make_opts(precision=QK_MIXED, fp32_threshold=1e-3))
    dEm = abs(0.5 * float(np.sum(D * (Gm - G))))
    dFm = np.abs(Gm - G).max()
    gm = eng.grad(D, make_opts(precision=QK_MIXED, fp32_threshold=1e-2))
    dgm = np.abs(gm - g64).max()
    print(f"water32 mixed: Fock 1e-3 |dE| {dEm:.2e} max|dF| {dFm:.2e}; gradient 1e-2 max|dgrad| {dgm:.2e} "
          f"(vs reference {np.abs(gm - z['grad4c']).max():.2e})", flush=True)
    assert dEm <= 1e-6 and dFm <= 1e-6
    assert dgm <= 1e-6 and np.abs(gm - z["grad4c"]).max() <= 1e-6
