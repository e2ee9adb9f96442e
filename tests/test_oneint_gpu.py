"""HIP one-electron integrals and their gradient terms (qk_oneint.hip) against
the reference and the oracle, and the complete HF energy and gradient.

* S and H_core of H2O, (H2O)2, (H2O)4 against the reference's get1e output
  (tests/golden/*_scf.npz) and against the oracle;
* the total energy sum D H + 1/2 sum D G + E_nn from the device's H, G and
  E_nn at the reference's converged density against its SCF energy (1e-8 Eh);
* the complete gradient from one backward call, qk_fock_grad(D, W): nuclear
  repulsion, the one-electron terms and the four-centre term, against the
  reference's full hfgrad output (1e-9 Eh/bohr);
* (H2O)32 (800 functions): H_core samples, the total energy and the complete
  gradient at the converged density; bitwise repeatability;
* n-icosane (510 functions) at the density of the reference's own SCF: G and
  H_core samples, the total energy, the four-centre gradient (FP64 and mixed)
  and the complete gradient.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from oracle.oracle import Fixture
from quick_amd.basis import build_basis
from quick_amd.fock import FockEngine, lib, make_opts
from quick_amd.geometry import CONFIGS, small_system

pytestmark = pytest.mark.gpu

ZOF = {"H": 1.0, "C": 6.0, "N": 7.0, "O": 8.0}


def _banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def _dens(z):
    C_ = z["C_occ"]
    eps = z["eps"][:C_.shape[1]]
    return 2.0 * C_ @ C_.T, 2.0 * (C_ * eps[None, :]) @ C_.T


def _complete_grad(eng, D, W):
    Df, Wf = np.asfortranarray(D), np.asfortranarray(W)
    g = np.zeros(3 * eng.natom)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    rc = lib().qk_fock_grad(eng.h, dp(Df), dp(Wf), dp(g), C.byref(make_opts()))
    assert rc == 0
    return g.reshape(3, eng.natom, order="F")


@pytest.mark.parametrize("name", ["h2o", "water2", "water4"])
def test_oneint_energy_and_complete_gradient(name):
    _banner(f"one-electron {name}")
    z = np.load(f"tests/golden/{name}_scf.npz")
    fx = Fixture(f"tests/golden/{name}_scf.npz")
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    eng.set_charges([ZOF[str(s)] for s in z["symbols"]])
    S, H = eng.oneint()
    b = build_basis(*small_system(name))
    b.xyz = np.asarray(z["xyz"], np.float64).copy()
    So, Ho = O.hcore(b)
    print(f"{name}: max|dS| {np.abs(S - z['S']).max():.2e} max|dH| {np.abs(H - z['H']).max():.2e} "
          f"(oracle {np.abs(H - Ho).max():.2e})", flush=True)
    assert np.abs(S - z["S"]).max() <= 1e-12 and np.abs(H - z["H"]).max() <= 1e-10
    assert np.abs(S - So).max() <= 1e-12 and np.abs(H - Ho).max() <= 1e-10
    assert np.array_equal(H, H.T) and np.array_equal(S, S.T)
    D, W = _dens(z)
    G = eng.fock(D)
    Enn = eng.nuclear_repulsion()
    E = float(np.sum(D * H)) + 0.5 * float(np.sum(D * G)) + Enn
    assert abs(Enn - float(z["E_nn"])) <= 1e-10
    assert abs(E - float(z["E_total"])) <= 1e-8
    g1 = eng.oneint_grad(D, W)
    assert np.abs(g1 - O.oneint_grad(b, D, W)).max() <= 1e-9
    g = _complete_grad(eng, D, W)
    dg = np.abs(g - z["grad_total"]).max()
    print(f"{name}: |dE_total| {abs(E - float(z['E_total'])):.2e} max|dgrad_total| {dg:.2e}", flush=True)
    assert dg <= 1e-9
    assert np.array_equal(g1, eng.oneint_grad(D, W))      # bitwise repeatable


def test_oneint_water32_converged():
    _banner("one-electron water32")
    z = np.load("tests/golden/water32_scf.npz")
    syms, xyz = CONFIGS["water32"]()
    b = build_basis(syms, xyz)
    b.xyz = np.asarray(z["xyz"], np.float64).copy()
    eng = FockEngine(b)
    eng.compute_schwarz()
    eng.set_charges([ZOF[s] for s in syms])
    S, H = eng.oneint()
    dh = np.abs(H[z["H_sample_i"], z["H_sample_j"]] - z["H_sample"]).max()
    dd = np.abs(np.diag(H) - z["H_diag"]).max()
    dn = np.abs(np.linalg.norm(H, axis=1) - z["H_rownorm"]).max()
    D, W = _dens(z)
    G = eng.fock(D)
    E = float(np.sum(D * H)) + 0.5 * float(np.sum(D * G)) + eng.nuclear_repulsion()
    g = _complete_grad(eng, D, W)
    dg = np.abs(g - z["grad_total"]).max()
    print(f"water32: max|dH| samples {dh:.2e} diag {dd:.2e} rownorm {dn:.2e} |dE_total| "
          f"{abs(E - float(z['E_total'])):.2e} max|dgrad_total| {dg:.2e}", flush=True)
    assert dh <= 1e-10 and dd <= 1e-10 and dn <= 1e-9
    assert abs(E - float(z["E_total"])) <= 1e-8
    assert dg <= 1e-9
    assert np.abs(g.sum(axis=1)).max() <= 1e-8           # translational invariance


def test_icosane_at_reference_scf_density():
    """n-icosane C20H42 (510 functions) at the density of the reference's own
    SCF (tests/golden/icosane_scf.npz: 201 cycles, last RMS density change
    5e-9): H_core samples, the total energy, the four-centre gradient in FP64
    and in mixed precision against the reference's grad4c, and the complete
    gradient from one backward call against its full hfgrad output."""
    from quick_amd.fock import QK_MIXED
    _banner("icosane at the reference SCF density")
    z = np.load("tests/golden/icosane_scf.npz")
    syms, xyz = CONFIGS["icosane"]()
    b = build_basis(syms, xyz)
    b.xyz = np.asarray(z["xyz"], np.float64).copy()
    eng = FockEngine(b)
    Y, _ = eng.compute_schwarz()
    assert np.abs(Y - z["Ycutoff"]).max() <= 1e-12
    eng.set_charges([ZOF[s] for s in syms])
    S, H = eng.oneint()
    dh = np.abs(H[z["H_sample_i"], z["H_sample_j"]] - z["H_sample"]).max()
    D, W = _dens(z)
    G = eng.fock(D)
    dG = np.abs(G[z["G_sample_i"], z["G_sample_j"]] - z["G_sample"]).max()
    E = float(np.sum(D * H)) + 0.5 * float(np.sum(D * G)) + eng.nuclear_repulsion()
    g4 = eng.grad(D)
    g4m = eng.grad(D, make_opts(precision=QK_MIXED, fp32_threshold=1e-2))
    g = _complete_grad(eng, D, W)
    d4, d4m, dg = (np.abs(g4 - z["grad4c"]).max(), np.abs(g4m - z["grad4c"]).max(),
                   np.abs(g - z["grad_total"]).max())
    print(f"icosane: max|dH| samples {dh:.2e} max|dG| samples {dG:.2e} |dE_total| {abs(E - float(z['E_total'])):.2e} "
          f"max|dgrad4c| {d4:.2e} (mixed {d4m:.2e}) max|dgrad_total| {dg:.2e}", flush=True)
    assert dh <= 1e-10 and dG <= 1e-10
    assert abs(E - float(z["E_total"])) <= 1e-8
    assert d4 <= 1e-9 and d4m <= 1e-6
    assert dg <= 1e-9
    assert np.abs(g.sum(axis=1)).max() <= 1e-8
