"""Pins the CPU oracle (oracle/qk_oracle.c) before it is trusted as the checker.

Pins, per restated function:
  * basis tables (basis.f90:377-599): bit-equal to the reference's own dump;
  * Boys FmT (subs/fmt.f90): closed forms F_m(0)=1/(2m+1), F_0 = sqrt(pi/T)erf(sqrt T)/2,
    scipy's incomplete gamma, and the reference's known instability of upward
    recursion at small T with high m (which the reference-path parity depends on);
  * VRR/HRR/normalisation: (ss|ss) closed form, an independent McMurchie-Davidson
    brute force (tests/md_reference.py), 8-fold permutational symmetry;
  * Schwarz bounds and G(D): element-wise against the reference CPU path's own
    outputs for 5 seeded systems (tests/golden/*.npz, written by
    tests/golden/make_golden.py from oracle/_ref);
  * a planted mistake (K with the wrong weight) is caught by the golden check;
  * the four-centre gradient: against the reference's hfgrad output (grad4c) for
    the same 5 systems, translational invariance, and central finite differences
    of the oracle's own E2 = 1/2 sum D G(D) (HF and 0.2 exchange weight).
"""
import math

import numpy as np
import pytest
from scipy.special import erf, gammainc, gamma

from oracle.oracle import Oracle, fmt
from quick_amd.basis import build_basis
from quick_amd.geometry import small_system

SYSTEMS = ["h2o", "ch4", "water2", "ethane", "water4"]


def test_basis_tables_match_reference(golden):
    for name in SYSTEMS:
        fx = golden(name)
        syms, xyz = small_system(name)
        b = build_basis(syms, xyz)
        for k in ("kstart", "kprim", "katom", "ktype", "Qstart", "Qfinal", "Qsbasis",
                  "Qfbasis", "ksumtype", "KLMN"):
            assert np.array_equal(getattr(b, k), getattr(fx, k)), (name, k)
        for k in ("cons", "gccoeff", "gcexpo"):
            assert np.array_equal(getattr(b, k), getattr(fx, k)), (name, k)
        assert np.allclose(b.xyz, fx.xyz, rtol=0, atol=1e-15)


def test_boys_closed_forms():
    for m in range(0, 13):
        v = fmt(m, 0.0)
        assert np.allclose(v, [1.0 / (2 * k + 1) for k in range(m + 1)], rtol=0, atol=1e-15)
    for T in np.concatenate([np.linspace(1e-4, 50.0, 400), [0.9, 1.0, 1.1, 2.9, 3.0, 3.1, 4.99, 5.01, 9.99, 10.01, 15.01, 33.01]]):
        f0 = 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
        assert abs(fmt(0, T)[0] - f0) < 5e-14 * max(1.0, f0), T   # the fit is good to ~2e-14


def test_boys_exact_mode_matches_incomplete_gamma():
    for T in [1e-6, 1e-3, 0.5, 3.0, 12.0, 30.0, 40.0]:
        ref = [gamma(m + 0.5) * gammainc(m + 0.5, T) / (2 * T ** (m + 0.5)) for m in range(11)]
        got = fmt(10, T, exact=True)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-15), T


def test_boys_reference_fit_regimes():
    # T >= 5: the fit plus upward recursion is good to ~1e-13 for all orders used
    for T in [6.0, 12.0, 20.0, 40.0]:
        assert np.allclose(fmt(8, T), fmt(8, T, exact=True), rtol=2e-13, atol=0)
    # T = 2: upward recursion already costs digits at m = 8 (~7e-11 relative)
    assert np.allclose(fmt(8, 2.0), fmt(8, 2.0, exact=True), rtol=1e-9, atol=0)
    # small T with high order: the reference's upward recursion loses digits;
    # the oracle must reproduce that (it is what the reference computes)
    bad = abs(fmt(8, 1e-3)[8] - fmt(8, 1e-3, exact=True)[8])
    assert bad > 1.0


def _ss_closed_form(fx, i, j, k, l):
    from tests.md_reference import contracted
    fs = [contracted(fx, x) for x in (i, j, k, l)]
    tot = 0.0
    (A, _, ea, ca), (B, _, eb, cb), (C, _, ec, cc), (D, _, ed, cd) = fs
    for a, x1 in zip(ea, ca):
        for b, x2 in zip(eb, cb):
            for c, x3 in zip(ec, cc):
                for d, x4 in zip(ed, cd):
                    z, e = a + b, c + d
                    P, Q = (a * A + b * B) / z, (c * C + d * D) / e
                    Kab = math.exp(-a * b / z * np.sum((A - B) ** 2))
                    Kcd = math.exp(-c * d / e * np.sum((C - D) ** 2))
                    T = z * e / (z + e) * np.sum((P - Q) ** 2)
                    F0 = 1.0 if T < 1e-15 else 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
                    tot += x1 * x2 * x3 * x4 * 2 * math.pi ** 2.5 / (z * e * math.sqrt(z + e)) * Kab * Kcd * F0
    return tot


def test_ssss_closed_form(golden):
    fx = golden("water2")
    o = Oracle(fx)
    o.boys_exact(True)
    s_funcs = [i for i in range(fx.nbasis) if fx.KLMN[:, i].sum() == 0]
    rng = np.random.default_rng(0)
    for _ in range(40):
        i, j, k, l = rng.choice(s_funcs, 4)
        ref = _ss_closed_form(fx, i, j, k, l)
        assert abs(o.eri(i, j, k, l) - ref) <= 1e-13 * max(1.0, abs(ref))


def test_eri_vs_mcmurchie_davidson(golden):
    from tests.md_reference import eri as md_eri
    fx = golden("h2o")
    o = Oracle(fx)
    o.boys_exact(True)
    rng = np.random.default_rng(1)
    d_funcs = [i for i in range(fx.nbasis) if fx.KLMN[:, i].sum() == 2]
    p_funcs = [i for i in range(fx.nbasis) if fx.KLMN[:, i].sum() == 1]
    picks = [tuple(rng.integers(0, fx.nbasis, 4)) for _ in range(25)]
    picks += [tuple(rng.choice(d_funcs, 4)) for _ in range(6)]                  # (dd|dd)
    picks += [(d_funcs[1], p_funcs[2], d_funcs[4], p_funcs[5]) for _ in range(1)]  # (dp|dp)
    for i, j, k, l in picks:
        ref = md_eri(fx, i, j, k, l)
        got = o.eri(int(i), int(j), int(k), int(l))
        assert abs(got - ref) <= 1e-11 * max(1.0, abs(ref)), (i, j, k, l, got, ref)


def test_eri_eightfold_symmetry(golden):
    fx = golden("ch4")
    o = Oracle(fx)
    rng = np.random.default_rng(2)
    for _ in range(30):
        i, j, k, l = (int(v) for v in rng.integers(0, fx.nbasis, 4))
        v = o.eri(i, j, k, l)
        for p in [(j, i, k, l), (i, j, l, k), (k, l, i, j), (l, k, j, i)]:
            assert abs(o.eri(*p) - v) <= 1e-12 * max(1.0, abs(v))


@pytest.mark.parametrize("name", SYSTEMS)
def test_schwarz_matches_reference(golden, name):
    fx = golden(name)
    Y, cp = Oracle(fx).schwarz()
    assert np.abs(Y - fx.Ycutoff).max() <= 1e-13
    assert np.abs(cp - fx.cutprim).max() <= 1e-12


@pytest.mark.parametrize("name", SYSTEMS)
def test_fock_matches_reference(golden, name):
    fx = golden(name)
    o = Oracle(fx)
    o.set_cutoffs(fx.Ycutoff, fx.cutprim)
    G, nq, _ = o.fock(fx.D, mode="cpu")
    assert nq > 0
    assert np.abs(G - fx.G).max() <= 1e-12
    assert abs(0.5 * np.sum(fx.D * (G - fx.G))) <= 1e-11
    # the HIP path's enumeration (all Qshell pairs, CPU Boys order) is the same operator
    G2, _, _ = o.fock(fx.D, mode="gpu_cpuboys", pair_cut=0.0)
    assert np.abs(G2 - fx.G).max() <= 1e-12


def test_planted_mistake_is_caught(golden):
    fx = golden("water2")
    o = Oracle(fx)
    o.set_cutoffs(fx.Ycutoff, fx.cutprim)
    G, _, _ = o.fock(fx.D, mode="gpu_cpuboys", pair_cut=0.0, hybrid=0.9)   # wrong K weight
    assert np.abs(G - fx.G).max() > 1e-4
    G, _, _ = o.fock(fx.D.T * 1.0, mode="gpu_cpuboys", pair_cut=0.0, tau=1e-3)   # wrong cutoff
    assert np.abs(G - fx.G).max() > 1e-6


# ---------------------------------------------------------------------------
# four-centre gradient (gradient.f90:261-320, shell.f90:1863-2535, hrr.f90:862-1014)

@pytest.mark.parametrize("name", SYSTEMS)
def test_grad_matches_reference(golden, name):
    """Against the reference's own hfgrad four-centre loop (grad4c, gradCutoff 1e-11)."""
    fx = golden(name)
    o = Oracle(fx)
    o.set_cutoffs(fx.Ycutoff, fx.cutprim)
    g, nq, _ = o.grad(fx.D, cutoff=1e-11)
    assert nq > 0
    assert np.abs(g - fx.grad4c).max() <= 1e-11
    # translational invariance: the forces on the atoms sum to zero
    assert np.abs(g.sum(axis=1)).max() <= 1e-11


def _e2_unscreened(fx, xyz, hybrid):
    import copy
    f = copy.copy(fx)
    f.xyz = xyz
    o = Oracle(f)
    o.boys_exact(True)
    Y, cp = o.schwarz()
    o.set_cutoffs(Y, cp)
    G, _, _ = o.fock(fx.D, tau=0.0, prim_limit=0.0, pair_cut=-1.0, hybrid=hybrid, mode="gpu_cpuboys")
    return 0.5 * float(np.sum(fx.D * G))


@pytest.mark.parametrize("hybrid", [1.0, 0.2])
def test_grad_matches_central_differences(golden, hybrid):
    """dE2/dR at fixed D against central differences (step 1e-4 bohr) of
    E2 = 1/2 sum D G(D) from the oracle's own Fock build, with exact Boys
    values and no screening so that E2 is smooth in R."""
    fx = golden("h2o")
    o = Oracle(fx)
    o.boys_exact(True)
    Y, cp = o.schwarz()
    o.set_cutoffs(Y, cp)
    g, _, _ = o.grad(fx.D, cutoff=0.0, hybrid=hybrid)
    h = 1e-4
    worst = 0.0
    for a in range(fx.xyz.shape[1]):
        for x in range(3):
            xp, xm = fx.xyz.copy(), fx.xyz.copy()
            xp[x, a] += h
            xm[x, a] -= h
            fd = (_e2_unscreened(fx, xp, hybrid) - _e2_unscreened(fx, xm, hybrid)) / (2 * h)
            worst = max(worst, abs(g[x, a] - fd))
    assert worst <= 2e-7, worst


def test_grad_planted_mistake_is_caught(golden):
    """A wrong exchange weight moves the gradient well outside the golden tolerance."""
    fx = golden("water2")
    o = Oracle(fx)
    o.set_cutoffs(fx.Ycutoff, fx.cutprim)
    g, _, _ = o.grad(fx.D, cutoff=1e-11, hybrid=0.9)
    assert np.abs(g - fx.grad4c).max() > 1e-4
