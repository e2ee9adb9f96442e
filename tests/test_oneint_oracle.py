"""The oracle's one-electron restatement (oracle/qk_oneint_oracle.c) against
the reference's own output and the mathematics.

tests/golden/{h2o,water2,water4}_scf.npz hold the reference's S and H_core
(get1e) and, at its converged density, its full hfgrad gradient and the
four-centre term; their difference is the one-electron part (nuclear repulsion,
-2 W dS/dR + 2 D dT/dR, D dV/dR) the oracle restates, with W = 2 sum_occ e_i
c_i c_i^T.  Independently of the reference: that gradient against central
differences of E1(R) = sum D H(R) - sum W S(R) + E_nn(R) at fixed D and W, and
translational invariance.
"""
import numpy as np
import pytest

from oracle import oracle as O
from quick_amd.basis import build_basis
from quick_amd.geometry import small_system

SYSTEMS = ["h2o", "water2", "water4"]


def _load(name):
    z = np.load(f"tests/golden/{name}_scf.npz")
    b = build_basis(*small_system(name))
    b.xyz = np.asarray(z["xyz"], np.float64).copy()
    C = z["C_occ"]
    eps = z["eps"][:C.shape[1]]
    return z, b, 2.0 * C @ C.T, 2.0 * (C * eps[None, :]) @ C.T


@pytest.mark.parametrize("name", SYSTEMS)
def test_hcore_and_overlap_match_reference(name):
    z, b, _, _ = _load(name)
    S, H = O.hcore(b)
    assert np.abs(S - z["S"]).max() <= 1e-13
    assert np.abs(H - z["H"]).max() <= 1e-12


@pytest.mark.parametrize("name", SYSTEMS)
def test_oneint_gradient_matches_reference(name):
    z, b, D, W = _load(name)
    g = O.oneint_grad(b, D, W)
    assert np.abs(g - (z["grad_total"] - z["grad4c"])).max() <= 1e-11
    assert np.abs(g.sum(axis=1)).max() <= 1e-11          # translational invariance


def _e1(b, D, W):
    S, H = O.hcore(b)
    R = b.xyz.T
    Z = np.asarray(b.Z, np.float64)
    enn = sum(Z[i] * Z[j] / np.linalg.norm(R[i] - R[j]) for i in range(len(Z)) for j in range(i + 1, len(Z)))
    return float(np.sum(D * H) - np.sum(W * S)) + enn


def test_oneint_gradient_central_differences():
    z, b, D, W = _load("water2")
    g = O.oneint_grad(b, D, W)
    h = 1e-4
    rng = np.random.default_rng(2)
    for _ in range(6):
        a, x = int(rng.integers(0, b.xyz.shape[1])), int(rng.integers(0, 3))
        xp = b.xyz.copy()
        b.xyz[x, a] = xp[x, a] + h
        ep = _e1(b, D, W)
        b.xyz[x, a] = xp[x, a] - h
        em = _e1(b, D, W)
        b.xyz[:] = xp
        assert abs((ep - em) / (2 * h) - g[x, a]) <= 1e-7


def test_planted_mistake_is_seen():
    """Dropping the energy-weighted density term moves the gradient by far more than the bar."""
    z, b, D, W = _load("h2o")
    g = O.oneint_grad(b, D, np.zeros_like(W))
    assert np.abs(g - (z["grad_total"] - z["grad4c"])).max() > 1e-3
