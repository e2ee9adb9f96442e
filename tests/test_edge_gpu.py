"""Edge cases of the HIP path: a single atom (every quartet one-centre), a
molecule with s and p shells only, a zero density, a cutoff that screens every
quartet, and a contraction longer than the kernels stage (rejected)."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from quick_amd.basis import build_basis
from quick_amd.fock import FockEngine, QK_MIXED, make_opts
from quick_amd.geometry import small_system, synthetic_density

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def _setup(symbols, xyz):
    b = build_basis(symbols, np.asarray(xyz, float))
    D = synthetic_density(b.xyz, b.katom, b.ksumtype, b.nbasis, 3)
    eng = FockEngine(b)
    Y, cp = eng.compute_schwarz()
    o = Oracle(b)
    o.set_cutoffs(Y, cp)
    return b, D, eng, o


@pytest.mark.parametrize("symbols,xyz", [(["O"], [[0.0, 0.0, 0.0]]),
                                         (["H", "H"], [[0.0, 0.0, 0.0], [0.0, 0.0, 0.74]])],
                         ids=["O-atom", "H2"])
def test_small_molecules_against_oracle(symbols, xyz):
    """G(D) against the oracle's restatement of the reference CPU path; the
    gradient against the oracle's hfgrad loop (a single atom: exactly zero,
    every quartet is one-centre)"""
    _banner(f"edge {symbols}")
    b, D, eng, o = _setup(symbols, xyz)
    G = eng.fock(D)
    Gref, nq, _ = o.fock(D, mode="cpu")
    assert nq > 0
    assert np.abs(G - Gref).max() <= 1e-10
    assert np.array_equal(G, G.T)
    g = eng.grad(D)
    gref, _, _ = o.grad(D)
    assert np.abs(g - gref).max() <= 1e-10
    if len(symbols) == 1:
        assert not g.any()


def test_zero_density_and_full_screening():
    """D = 0 and a cutoff above every Schwarz product leave no quartet: G and
    the gradient are exactly zero, in FP64 and mixed precision"""
    _banner("edge zero density / full screening")
    syms, xyz = small_system("water2")
    b, D, eng, o = _setup(syms, xyz)
    Z = np.zeros_like(D)
    for opts in (make_opts(), make_opts(precision=QK_MIXED, fp32_threshold=1e-2)):
        G = eng.fock(Z, opts)
        assert not G.any()
        assert eng.stats().quartets == 0
        assert not eng.grad(Z, opts).any()
    G = eng.fock(D, make_opts(tau=1e6))
    assert not G.any() and eng.stats().quartets == 0
    assert not eng.grad(D, make_opts(tau=1e6)).any()
    # and the build is whole again with the default cutoff
    Gref, _, _ = o.fock(D, mode="cpu")
    assert np.abs(eng.fock(D) - Gref).max() <= 1e-10


def test_contraction_longer_than_staged_is_rejected():
    """a shell of 9 primitives (the kernels stage at most 8 per shell): the
    library refuses the system with a message instead of computing"""
    _banner("edge oversize contraction")
    b = build_basis(["O"], np.zeros((1, 3)))
    b.kprim = np.array(b.kprim).copy()
    b.kprim[0] = 9
    with pytest.raises(RuntimeError, match="at most 8"):
        FockEngine(b)


def test_fixed_point_range_guard(golden):
    """A density scaled so that |G| reaches 2e4 (the accumulator's range is about
    3.3e4 for off-diagonal elements): the build reports an error instead of
    returning a silently wrapped operator; at the normal scale the check passes."""
    from quick_amd.fock import FockEngine, lib
    print("[gpu-test] fixed-point range guard water2", flush=True)
    fx = golden("water2")
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    G = eng.fock(fx.D)
    assert lib().qk_check_range(eng.h) == 0
    scale = 2.0e4 / np.abs(G).max()
    with pytest.raises(RuntimeError, match="fixed-point"):
        eng.fock(fx.D * scale)
    assert lib().qk_check_range(eng.h) == -5
    G2 = eng.fock(fx.D)   # the next build clears the flag
    assert np.array_equal(G2, G)
    assert lib().qk_check_range(eng.h) == 0
