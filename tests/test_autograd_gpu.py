"""The forward/backward C-ABI pair (qk_fock_build, qk_fock_grad) through the
test harness's torch.autograd.Function: the backward against central
differences of the forward, and the W argument's error path."""
import ctypes as C

import numpy as np
import pytest
import torch

from quick_amd.basis import build_basis
from quick_amd.fock import FockEngine, lib, last_error, make_opts, QK_MIXED
from quick_amd.geometry import ANGSTROM_TO_BOHR, small_system, synthetic_density
from tests.autograd_shim import two_electron_energy

pytestmark = pytest.mark.gpu


def _system(name):
    syms, xyz = small_system(name)
    b = build_basis(syms, xyz)
    D = synthetic_density(b.xyz, b.katom, b.ksumtype, b.nbasis, 7)
    return syms, xyz, D


def test_autograd_backward_matches_central_differences():
    """d E2 / d R from backward (qk_fock_grad) against central differences of
    forward (qk_fock_build), step 1e-4 bohr, on six seeded coordinates of (H2O)2.
    Both sides use QUICK's piecewise-fitted Boys function (with upward
    recursion), and the analytic derivative takes F_{m+1} for -dF_m/dT, which
    the fit honours only approximately: differences of up to ~1e-6 relative
    remain here (a tighter cutoff does not change them), on derivatives of
    5-40 Eh/A from the synthetic density.  The bar is 2e-6 relative; the
    backward is pinned to 1e-9 against the reference's own hfgrad elsewhere."""
    print("[gpu-test] autograd fp64", flush=True)
    syms, xyz, D = _system("water2")
    opts = make_opts()
    x = torch.tensor(xyz, dtype=torch.float64, requires_grad=True)
    E = two_electron_energy(x, syms, D, opts)
    E.backward()
    g = x.grad.numpy()
    h = 1e-4 / ANGSTROM_TO_BOHR       # 1e-4 bohr in Angstrom
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(6):
        a, k = int(rng.integers(len(syms))), int(rng.integers(3))
        e = {}
        for st in (1, -1):
            x2 = xyz.copy()
            x2[a, k] += st * h
            e[st] = float(two_electron_energy(torch.tensor(x2), syms, D, opts))
        fd = (e[1] - e[-1]) / (2 * h)
        worst = max(worst, abs(fd - g[a, k]) / max(1.0, abs(g[a, k])))
        print(f"atom {a} axis {k}: backward {g[a, k]:+.9f} fd {fd:+.9f} (Eh/A)", flush=True)
    assert worst <= 2e-6, worst
    assert np.abs(g.sum(axis=0)).max() <= 1e-8     # translational invariance


def test_autograd_mixed_backward_against_fp64():
    """the mixed-precision pair through the same Function: E2 and dE2/dR
    against the FP64 ones (FP32 rounding makes finite differences of the mixed
    forward too noisy to check against)"""
    print("[gpu-test] autograd mixed", flush=True)
    syms, xyz, D = _system("water2")
    out = {}
    for name, opts in (("fp64", make_opts()), ("mixed", make_opts(precision=QK_MIXED, fp32_threshold=1e-2))):
        x = torch.tensor(xyz, dtype=torch.float64, requires_grad=True)
        E = two_electron_energy(x, syms, D, opts)
        E.backward()
        out[name] = (float(E), x.grad.numpy() / ANGSTROM_TO_BOHR)
    assert abs(out["mixed"][0] - out["fp64"][0]) <= 1e-6
    assert np.abs(out["mixed"][1] - out["fp64"][1]).max() <= 1e-6


def test_fock_grad_with_energy_weighted_density_needs_charges():
    print("[gpu-test] qk_fock_grad W", flush=True)
    syms, xyz, D = _system("h2o")
    eng = FockEngine(build_basis(syms, xyz))
    eng.compute_schwarz()
    Df = np.asfortranarray(D)
    W = np.zeros_like(Df)
    g = np.zeros(3 * eng.natom)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    rc = lib().qk_fock_grad(eng.h, dp(Df), dp(W), dp(g), C.byref(make_opts()))
    assert rc != 0 and "charges" in last_error()
    assert not g.any()
    # the forward of the pair is the host build
    F = np.zeros_like(Df)
    assert lib().qk_fock_build(eng.h, dp(Df), dp(F), C.byref(make_opts())) == 0
    assert np.allclose(F, eng.fock(D), rtol=0, atol=1e-13)
