"""HIP Fock build vs the oracle and the reference's golden vectors (MI355X)."""
import ctypes
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from quick_amd.basis import build_basis
from quick_amd.fock import FockEngine, make_opts, QK_BOYS_ORDER_QSHELL, QK_MIXED
from quick_amd.geometry import CONFIGS, small_system

pytestmark = pytest.mark.gpu

SYSTEMS = ["h2o", "ch4", "water2", "ethane", "water4"]
TOL_F = 1e-10      # |dF_ij|, Eh (FP64 parity bar)
TOL_E = 1e-8       # |dE|, Eh


def _banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


@pytest.mark.parametrize("name", SYSTEMS)
def test_fock_matches_reference_cpu_path(golden, name):
    _banner(f"fock parity {name}")
    fx = golden(name)
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    G = eng.fock(fx.D)
    dG = np.abs(G - fx.G).max()
    dE = abs(0.5 * np.sum(fx.D * (G - fx.G)))
    assert dG <= TOL_F, dG
    assert dE <= TOL_E, dE
    assert np.array_equal(G, G.T)


@pytest.mark.parametrize("name", SYSTEMS)
def test_schwarz_on_device_matches_reference(golden, name):
    _banner(f"schwarz {name}")
    fx = golden(name)
    eng = FockEngine(fx)
    Y, cp = eng.compute_schwarz()
    assert np.abs(Y - fx.Ycutoff).max() <= 1e-12
    assert np.abs(cp - fx.cutprim).max() <= 1e-12


def test_reference_gpu_screening_mode_matches_oracle(golden):
    _banner("reference-GPU screening mode water4")
    fx = golden("water4")
    o = Oracle(fx)
    o.set_cutoffs(fx.Ycutoff, fx.cutprim)
    Gref, _, _ = o.fock(fx.D, mode="gpu", pair_cut=1e-10)
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    G = eng.fock(fx.D, make_opts(pair_cut=1e-10, boys_order=QK_BOYS_ORDER_QSHELL))
    assert np.abs(G - Gref).max() <= TOL_F


def test_bitwise_deterministic(golden):
    _banner("determinism water4")
    fx = golden("water4")
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    G1 = eng.fock(fx.D)
    G2 = eng.fock(fx.D)
    assert np.array_equal(G1, G2)


def test_device_pointer_entry_point(golden):
    import torch
    _banner("device entry point water2")
    fx = golden("water2")
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    D = torch.from_numpy(np.asfortranarray(fx.D).ravel(order="F").copy()).cuda()
    G = torch.empty_like(D)
    eng.fock_device(D.data_ptr(), G.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    Gh = G.cpu().numpy().reshape(fx.nbasis, fx.nbasis, order="F")
    assert np.abs(Gh - fx.G).max() <= TOL_F


def test_hybrid_exchange_weight(golden):
    _banner("hybrid weight water2")
    fx = golden("water2")
    o = Oracle(fx)
    o.set_cutoffs(fx.Ycutoff, fx.cutprim)
    Gref, _, _ = o.fock(fx.D, mode="gpu_cpuboys", pair_cut=0.0, hybrid=0.2)
    eng = FockEngine(fx)
    eng.set_schwarz(fx.Ycutoff, fx.cutprim)
    G = eng.fock(fx.D, make_opts(hybrid=0.2))
    assert np.abs(G - Gref).max() <= TOL_F


def _water32_engine():
    syms, xyz = CONFIGS["water32"]()
    b = build_basis(syms, xyz)
    eng = FockEngine(b)
    eng.compute_schwarz()
    return b, eng


def _density(b, seed=3):
    from quick_amd.geometry import synthetic_density
    return synthetic_density(b.xyz, b.katom, b.ksumtype, b.nbasis, seed)


def test_water32_translation_and_permutation_invariance():
    """Full-size checks that need no oracle: translating the cluster leaves G
    unchanged; reversing the order of the waters permutes G accordingly."""
    _banner("water32 invariances")
    syms, xyz = CONFIGS["water32"]()
    b = build_basis(syms, xyz)
    D = _density(b)
    e1 = FockEngine(b)
    e1.compute_schwarz()
    G1 = e1.fock(D)
    assert np.array_equal(G1, G1.T)
    e1.close()
    b2 = build_basis(syms, xyz + np.array([1.25, -0.5, 2.0]))
    e2 = FockEngine(b2)
    e2.compute_schwarz()
    G2 = e2.fock(D)
    assert np.abs(G1 - G2).max() <= 1e-9
    e2.close()
    nw = len(syms) // 3
    order = np.concatenate([[3 * w, 3 * w + 1, 3 * w + 2] for w in reversed(range(nw))])
    b3 = build_basis([syms[i] for i in order], xyz[order])
    per_w = b.nbasis // nw
    fperm = np.concatenate([np.arange(w * per_w, (w + 1) * per_w) for w in reversed(range(nw))])
    D3 = D[np.ix_(fperm, fperm)]
    e3 = FockEngine(b3)
    e3.compute_schwarz()
    G3 = e3.fock(D3)
    assert np.abs(G3 - G1[np.ix_(fperm, fperm)]).max() <= 1e-9


def test_water32_full_size_matches_reference_samples(golden):
    """(H2O)32, 800 bf: the benchmark system against the reference CPU path's
    own full build (40k seeded element samples, diagonal, row norms, E2)."""
    from quick_amd.geometry import synthetic_density
    _banner("water32 vs reference samples")
    fx = golden("water32")
    syms, xyz = CONFIGS["water32"]()
    b = build_basis(syms, xyz)
    b.xyz = fx.xyz.copy()                      # the coordinates exactly as the reference read them
    D = synthetic_density(b.xyz, b.katom, b.ksumtype, b.nbasis, int(fx.density_seed))
    eng = FockEngine(b)
    Y, _ = eng.compute_schwarz()
    assert np.abs(Y - fx.Ycutoff).max() <= 1e-12
    G = eng.fock(D)
    dg = np.abs(G[fx.G_sample_i, fx.G_sample_j] - fx.G_sample).max()
    dd = np.abs(np.diag(G) - fx.G_diag).max()
    dn = np.abs(np.linalg.norm(G, axis=1) - fx.G_rownorm).max()
    dE = abs(0.5 * np.sum(D * G) - float(fx.E2))
    print(f"water32: max|dG| samples {dg:.2e} diag {dd:.2e} rownorm {dn:.2e} |dE2| {dE:.2e}")
    assert dg <= TOL_F and dd <= TOL_F
    assert dn <= 1e-9
    assert dE <= TOL_E


def test_water64_full_size_matches_reference_samples():
    """(H2O)64, 1,600 functions (the B3LYP configuration's ERI half): the
    device's Schwarz bounds and G(D) against the reference CPU path's own full
    build (tests/golden/water64.npz: 40k seeded samples of Ycutoff and of G at a
    seeded density, the diagonals, row norms and E2), FP64, with every class's
    exchange rows in LDS; then at the configuration's exchange weight 0.2 the
    mixed-precision build (threshold 1e-3) against the FP64 one."""
    from quick_amd.geometry import synthetic_density
    _banner("water64 vs reference samples")
    z = np.load("tests/golden/water64.npz")
    syms, xyz = CONFIGS["water64"]()
    b = build_basis(syms, xyz)
    b.xyz = np.asarray(z["xyz"], np.float64).copy()
    assert b.nbasis == 1600 and int(z["nbasis"]) == 1600
    D = synthetic_density(b.xyz, b.katom, b.ksumtype, b.nbasis, int(z["density_seed"]))
    eng = FockEngine(b)
    Y, _ = eng.compute_schwarz()
    assert np.abs(Y[z["Y_sample_i"], z["Y_sample_j"]] - z["Y_sample"]).max() <= 1e-12
    assert np.abs(np.diag(Y) - z["Y_diag"]).max() <= 1e-12
    G = eng.fock(D)
    dg = np.abs(G[z["G_sample_i"], z["G_sample_j"]] - z["G_sample"]).max()
    dd = np.abs(np.diag(G) - z["G_diag"]).max()
    dn = np.abs(np.linalg.norm(G, axis=1) - z["G_rownorm"]).max()
    dE = abs(0.5 * np.sum(D * G) - float(z["E2"]))
    print(f"water64: max|dG| samples {dg:.2e} diag {dd:.2e} rownorm {dn:.2e} |dE2| {dE:.2e}", flush=True)
    assert dg <= TOL_F and dd <= TOL_F
    assert dn <= 1e-9
    assert dE <= TOL_E
    assert np.array_equal(G, G.T)
    G64 = eng.fock(D, make_opts(hybrid=0.2))
    Gm = eng.fock(D, make_opts(hybrid=0.2, precision=QK_MIXED, fp32_threshold=1e-3))
    dEm = abs(0.5 * np.sum(D * (Gm - G64)))
    dFm = np.abs(Gm - G64).max()
    print(f"water64 hybrid 0.2: mixed 1e-3 vs FP64 |dE2| {dEm:.2e} max|dF| {dFm:.2e}", flush=True)
    assert dEm <= 1e-6 and dFm <= 1e-6
