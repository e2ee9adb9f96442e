"""The no-allocation Fock entry point (qk_fock_device_noalloc) through the
C-ABI, against qk_fock_device on the same inputs: the operators are bitwise
equal (integer fixed-point accumulation of the same quartets) on two calls in
a row, (H2O)32 (800 functions) and then the smaller (H2O)4 (100 functions),
which shares the fixed scratch array with the first, in FP64 and in mixed
precision; and it refuses to run without work lists prepared for its options
(it would have to allocate them)."""
import numpy as np
import pytest

from quick_amd.basis import build_basis
from quick_amd.fock import FockEngine, QK_MIXED, last_error, make_opts
from quick_amd.geometry import CONFIGS, synthetic_density

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def _dev(torch, a):
    return torch.from_numpy(np.asarray(a, np.float64).ravel(order="F").copy()).cuda()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_fock_device_two_calls(golden, precision):
    import torch
    _banner(f"no-alloc Fock entry point, {precision}")
    opts = make_opts(precision=QK_MIXED, fp32_threshold=1e-3) if precision == "mixed" else make_opts()
    syms, xyz = CONFIGS["water32"]()
    b32 = build_basis(syms, xyz)
    e32 = FockEngine(b32)
    e32.compute_schwarz()
    D32 = _dev(torch, synthetic_density(b32.xyz, b32.katom, b32.ksumtype, b32.nbasis, 3))
    fx4 = golden("water4")
    e4 = FockEngine(fx4)
    e4.set_schwarz(fx4.Ycutoff, fx4.cutprim)
    D4 = _dev(torch, fx4.D)
    stream = torch.cuda.current_stream().cuda_stream
    # the existing entry point (it prepares the work lists for opts)
    R32, R4 = torch.empty_like(D32), torch.empty_like(D4)
    e32.fock_device(D32.data_ptr(), R32.data_ptr(), opts, stream)
    e4.fock_device(D4.data_ptr(), R4.data_ptr(), opts, stream)
    # the no-allocation one: two calls in a row, the second smaller
    G32, G4 = torch.full_like(D32, float("nan")), torch.full_like(D4, float("nan"))
    assert e32.fock_device_noalloc(D32.data_ptr(), G32.data_ptr(), opts, stream) == 0, last_error()
    assert e4.fock_device_noalloc(D4.data_ptr(), G4.data_ptr(), opts, stream) == 0, last_error()
    torch.cuda.synchronize()
    assert torch.equal(G32, R32)
    assert torch.equal(G4, R4)
    # and once more on the large system after the small one
    G32b = torch.full_like(D32, float("nan"))
    assert e32.fock_device_noalloc(D32.data_ptr(), G32b.data_ptr(), opts, stream) == 0, last_error()
    torch.cuda.synchronize()
    assert torch.equal(G32b, R32)
    g4 = G4.cpu().numpy().reshape(fx4.nbasis, fx4.nbasis, order="F")
    if precision == "fp64":
        assert np.abs(g4 - fx4.G).max() <= 1e-10      # and the reference's G(D)
    print(f"{precision}: (H2O)32 and (H2O)4 bitwise equal to qk_fock_device", flush=True)


def test_noalloc_requires_prepared_work_lists(golden):
    import torch
    _banner("no-alloc entry point without prepared work lists")
    fx = golden("water2")
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    D = _dev(torch, fx.D)
    G = torch.zeros_like(D)
    stream = torch.cuda.current_stream().cuda_stream
    assert eng.fock_device_noalloc(D.data_ptr(), G.data_ptr(), make_opts(), stream) == -6
    assert "qk_prepare" in last_error()
    eng.prepare(make_opts())
    # prepared for FP64 only: a mixed build would need other work lists
    assert eng.fock_device_noalloc(D.data_ptr(), G.data_ptr(), make_opts(precision=QK_MIXED), stream) == -6
    assert eng.fock_device_noalloc(D.data_ptr(), G.data_ptr(), make_opts(), stream) == 0, last_error()
    torch.cuda.synchronize()
    g = G.cpu().numpy().reshape(fx.nbasis, fx.nbasis, order="F")
    assert np.abs(g - fx.G).max() <= 1e-10
