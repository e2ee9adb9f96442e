"""HIP four-centre gradient (dE2/dR at fixed D, E2 = 1/2 sum D G(D)) against
the reference's hfgrad output, the oracle, and size-independent properties."""
import numpy as np
import pytest

from quick_amd.basis import build_basis
from quick_amd.fock import FockEngine, make_opts
from quick_amd.geometry import CONFIGS, synthetic_density

pytestmark = pytest.mark.gpu

SYSTEMS = ["h2o", "ch4", "water2", "ethane", "water4"]
TOL_G = 1e-9     # |d grad|, Eh/bohr, FP64 path against the reference


def _banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


@pytest.mark.parametrize("name", SYSTEMS)
def test_grad_matches_reference(golden, name):
    _banner(f"grad parity {name}")
    fx = golden(name)
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    g = eng.grad(fx.D)
    d = np.abs(g - fx.grad4c).max()
    print(f"{name}: max|dgrad| {d:.2e} (max|grad| {np.abs(fx.grad4c).max():.2f})")
    assert d <= TOL_G
    assert np.abs(g.sum(axis=1)).max() <= 1e-9


def test_grad_bitwise_deterministic_and_accumulates(golden):
    _banner("grad determinism water4")
    fx = golden("water4")
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    g1 = eng.grad(fx.D)
    g2 = eng.grad(fx.D)
    assert np.array_equal(g1, g2)
    # the device entry point adds into the caller's gradient
    import torch
    D = torch.from_numpy(np.asfortranarray(fx.D).ravel(order="F").copy()).cuda()
    base = torch.arange(3 * fx.natom, dtype=torch.float64, device="cuda")
    out = base.clone()
    eng.grad_device(D.data_ptr(), out.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = (out - base).cpu().numpy().reshape(3, -1, order="F")
    assert np.abs(got - g1).max() <= 1e-12


def test_grad_hybrid_weight_against_oracle(golden):
    from oracle.oracle import Oracle
    _banner("grad hybrid 0.2 water2")
    fx = golden("water2")
    o = Oracle(fx)
    o.set_cutoffs(fx.Ycutoff, fx.cutprim)
    gref, _, _ = o.grad(fx.D, cutoff=1e-11, hybrid=0.2)
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    g = eng.grad(fx.D, make_opts(hybrid=0.2))
    assert np.abs(g - gref).max() <= TOL_G


def _icosane():
    syms, xyz = CONFIGS["icosane"]()
    b = build_basis(syms, xyz)
    D = synthetic_density(b.xyz, b.katom, b.ksumtype, b.nbasis, 7)
    return syms, xyz, b, D


def test_icosane_full_size_matches_reference():
    """n-icosane C20H42, 510 functions (the gradient benchmark system): the
    device's Schwarz bounds and the FP64 four-centre gradient at the seeded
    density against the reference's own hfgrad four-centre term
    (tests/golden/icosane.npz, on the coordinates exactly as the reference read
    them), and the mixed-precision gradient (FP32 below 1e-2) within 1e-6."""
    from quick_amd.fock import QK_MIXED
    _banner("icosane grad vs reference")
    z = np.load("tests/golden/icosane.npz")
    syms, xyz = CONFIGS["icosane"]()
    b = build_basis(syms, xyz)
    b.xyz = np.asarray(z["xyz"], np.float64).copy()
    assert b.nbasis == int(z["nbasis"]) == 510
    D = synthetic_density(b.xyz, b.katom, b.ksumtype, b.nbasis, int(z["density_seed"]))
    eng = FockEngine(b)
    Y, _ = eng.compute_schwarz()
    assert np.abs(Y - z["Ycutoff"]).max() <= 1e-12
    g = eng.grad(D)
    d = np.abs(g - z["grad4c"]).max()
    gm = eng.grad(D, make_opts(precision=QK_MIXED, fp32_threshold=1e-2))
    dm = np.abs(gm - z["grad4c"]).max()
    print(f"icosane: max|dgrad| FP64 {d:.2e}, mixed 1e-2 {dm:.2e} (max|grad| {np.abs(z['grad4c']).max():.1f})",
          flush=True)
    assert d <= TOL_G
    assert dm <= 1e-6


def test_icosane_grad_invariances():
    """n-icosane, 510 bf (the gradient benchmark system): forces sum to zero,
    and translating the molecule leaves the gradient unchanged."""
    _banner("icosane grad invariances")
    syms, xyz, b, D = _icosane()
    e1 = FockEngine(b)
    e1.compute_schwarz()
    g1 = e1.grad(D)
    e1.close()
    assert np.abs(g1.sum(axis=1)).max() <= 1e-8
    b2 = build_basis(syms, xyz + np.array([0.7, -1.1, 0.4]))
    e2 = FockEngine(b2)
    e2.compute_schwarz()
    g2 = e2.grad(D)
    assert np.abs(g1 - g2).max() <= 1e-8


def test_icosane_grad_central_differences():
    """dE2/dR on 8 seeded Cartesian coordinates of n-icosane against central
    differences (step 1e-4 bohr) of E2 = 1/2 sum D G(D) from the FP64 Fock
    build at the displaced geometries (same D)."""
    _banner("icosane grad vs central differences")
    from quick_amd.geometry import ANGSTROM_TO_BOHR
    syms, xyz, b, D = _icosane()
    eng = FockEngine(b)
    eng.compute_schwarz()
    g = eng.grad(D)
    eng.close()
    rng = np.random.default_rng(8)
    picks = [(int(rng.integers(len(syms))), int(rng.integers(3))) for _ in range(8)]
    h = 1e-4   # bohr
    worst = 0.0
    for atom, ax in picks:
        e = []
        for sgn in (1, -1):
            x2 = xyz.copy()
            x2[atom, ax] += sgn * h / ANGSTROM_TO_BOHR
            bb = build_basis(syms, x2)
            en = FockEngine(bb)
            en.compute_schwarz()
            e.append(0.5 * float(np.sum(D * en.fock(D))))
            en.close()
        fd = (e[0] - e[1]) / (2 * h)
        print(f"atom {atom} axis {ax}: analytic {g[ax, atom]:+.9f} fd {fd:+.9f} diff {g[ax, atom] - fd:.2e}",
              flush=True)
        worst = max(worst, abs(g[ax, atom] - fd))
    assert worst <= 1e-5


# ---------------------------------------------------------------------------
# mixed precision: quartets with Schwarz bound below the threshold differentiated
# in FP32, accumulated in FP64 fixed point

TOL_G_MIXED = 1e-6   # Eh/bohr, mixed against FP64


def _fp32_share(eng):
    q64 = sum(c[1] for c in eng.grad_class_stats())
    q32 = sum(c[1] for c in eng.grad_class_stats(fp32=True))
    return q32 / max(q64 + q32, 1), q64 + q32


def test_grad_mixed_threshold_zero_is_fp64(golden):
    """threshold 0: every quartet stays FP64, bitwise the FP64 gradient"""
    from quick_amd.fock import QK_MIXED
    _banner("grad mixed threshold 0 water4")
    fx = golden("water4")
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    g64 = eng.grad(fx.D)
    g0 = eng.grad(fx.D, make_opts(precision=QK_MIXED, fp32_threshold=0.0))
    assert _fp32_share(eng)[0] == 0.0
    assert np.array_equal(g0, g64)


@pytest.mark.parametrize("name", ["water2", "water4"])
def test_grad_mixed_against_reference(golden, name):
    """an FP32 share of most quartets (threshold 1e-2) stays within 1e-6 of the
    reference gradient, and the same quartets are counted as in FP64"""
    from quick_amd.fock import QK_MIXED
    _banner(f"grad mixed {name}")
    fx = golden(name)
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    g64 = eng.grad(fx.D)
    n64 = sum(c[1] for c in eng.grad_class_stats())
    gm = eng.grad(fx.D, make_opts(precision=QK_MIXED, fp32_threshold=1e-2))
    share, nq = _fp32_share(eng)
    d = np.abs(gm - fx.grad4c).max()
    print(f"{name}: fp32 share {share:.3f} max|dgrad| vs reference {d:.2e}, vs FP64 {np.abs(gm - g64).max():.2e}")
    assert share > 0.3
    assert nq == n64
    assert d <= TOL_G_MIXED
    assert np.abs(gm.sum(axis=1)).max() <= 1e-9   # D = -(A+B+C) holds per quartet in any precision
    # reproducible
    assert np.array_equal(gm, eng.grad(fx.D, make_opts(precision=QK_MIXED, fp32_threshold=1e-2)))


@pytest.mark.parametrize("system", ["water32", "icosane"])
def test_grad_mixed_full_size(system):
    """(H2O)32 and n-icosane: the mixed gradient (threshold 1e-3) against the
    FP64 gradient, max error <= 1e-6 Eh/bohr"""
    from quick_amd.fock import QK_MIXED
    _banner(f"grad mixed {system}")
    syms, xyz = CONFIGS[system]()
    b = build_basis(syms, xyz)
    D = synthetic_density(b.xyz, b.katom, b.ksumtype, b.nbasis, 7)
    eng = FockEngine(b)
    eng.compute_schwarz()
    g64 = eng.grad(D)
    gm = eng.grad(D, make_opts(precision=QK_MIXED, fp32_threshold=1e-3))
    share, _ = _fp32_share(eng)
    d = np.abs(gm - g64).max()
    print(f"{system}: fp32 share {share:.3f} max|dgrad| {d:.2e} (max|grad| {np.abs(g64).max():.3f})", flush=True)
    assert share > 0.5
    assert d <= TOL_G_MIXED
    # threshold 0: all FP64, but most light classes then run from their tiles
    # where the FP64 gradient runs them density-compacted: bit for bit equal
    g0 = eng.grad(D, make_opts(precision=QK_MIXED, fp32_threshold=0.0))
    assert _fp32_share(eng)[0] == 0.0
    assert np.array_equal(g0, g64)
    # both streams in use (FP64: alternating classes; mixed at 1e-2 on (H2O)32:
    # light FP64 lists with FP32 lists after them): repeated calls bit for bit
    assert np.array_equal(eng.grad(D), g64)
    o2 = make_opts(precision=QK_MIXED, fp32_threshold=1e-2)
    g2 = eng.grad(D, o2)
    assert np.array_equal(eng.grad(D, o2), g2)
    assert np.abs(g2 - g64).max() <= TOL_G_MIXED
