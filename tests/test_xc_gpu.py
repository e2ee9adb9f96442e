"""HIP B3LYP XC (FP64 MFMA contractions) against the reference and the oracle.

* on the reference's own SG-1 grid (tests/golden/*_xc.npz) the device's V_xc and
  E_xc match the reference's dftoperator output;
* the device-built SG-1 grid is the reference's grid, and XC on it matches;
* at 1600 functions ((H2O)64, the B3LYP configuration) on the 75 x 302 product
  grid: the electron count of a density with a known integral, a symmetric
  V_xc, bitwise repeatability, and the oracle on a sample of the grid.
"""
import numpy as np
import pytest

from oracle import oracle as O
from quick_amd.basis import build_basis
from quick_amd.geometry import CONFIGS, small_system
from quick_amd.xc import QK_GRID_PRODUCT, QK_GRID_SG1, XcEngine, make_xc_opts

pytestmark = pytest.mark.gpu

TOL_V = 1e-8
TOL_E = 1e-8


def _banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def _load(name):
    z = np.load(f"tests/golden/{name}_xc.npz")
    b = build_basis(*small_system(name))
    C = z["C_occ"]
    return z, b, 2.0 * C @ C.T


@pytest.mark.parametrize("name", ["h2o", "water2"])
def test_xc_on_reference_grid_matches_reference(name):
    _banner(f"xc on the reference grid {name}")
    z, b, D = _load(name)
    eng = XcEngine(b, sigrad2=z["sigrad2"])
    assert np.array_equal(eng.sigrad2(), z["sigrad2"])
    eng.set_grid(np.column_stack([z["grid_xyz"], z["grid_w"]]))
    V, E, N = eng.xc(D)
    dV = np.abs(V - z["Vxc"]).max()
    print(f"{name}: |dExc| {abs(E - float(z['Exc'])):.2e} max|dV| {dV:.2e} N {N:.10f}")
    assert abs(E - float(z["Exc"])) <= TOL_E
    assert dV <= TOL_V
    assert abs(0.5 * N - float(z["aelec"])) <= 1e-9
    assert np.array_equal(V, V.T)


@pytest.mark.parametrize("name", ["h2o", "water2"])
def test_device_sg1_grid_is_the_reference_grid(name):
    _banner(f"device SG-1 grid {name}")
    z, b, D = _load(name)
    eng = XcEngine(b)
    assert np.array_equal(eng.sigrad2(), z["sigrad2"])
    eng.build_grid(make_xc_opts(grid=QK_GRID_SG1))
    pts, atom = eng.grid()
    # the device drops points no function reaches (they contribute nothing)
    ref = np.column_stack([z["grid_xyz"], z["grid_w"]])
    key = lambda a: np.lexsort((np.round(a[:, 2], 9), np.round(a[:, 1], 9), np.round(a[:, 0], 9)))
    s = O.sigrad2(b)
    reach = np.zeros(len(b.Z))
    for f in range(b.nbasis):
        reach[b.ncenter[f] - 1] = max(reach[b.ncenter[f] - 1], s[f])
    d2 = ((ref[:, None, :3] - b.xyz.T[None, :, :]) ** 2).sum(-1)
    ref = ref[(d2 <= reach[None, :]).any(1)]
    assert len(pts) == len(ref)
    ps, rs = pts[key(pts)], ref[key(ref)]
    assert np.abs(ps[:, :3] - rs[:, :3]).max() <= 1e-12
    # weights as the oracle is held to them: absolute, on the scale of the largest
    # (near a cell boundary 1 - g(mu) cancels, so small weights carry few digits)
    assert np.abs(ps[:, 3] - rs[:, 3]).max() <= 1e-12 * np.abs(rs[:, 3]).max()
    V, E, N = eng.xc(D)
    assert abs(E - float(z["Exc"])) <= TOL_E
    assert np.abs(V - z["Vxc"]).max() <= TOL_V


def _water64():
    syms, xyz = CONFIGS["water64"]()
    b = build_basis(syms, xyz)
    # a density with a known integral: D diagonal and non-negative, so rho >= 0
    # everywhere and int rho = sum_i D_ii (functions normalised to one)
    rng = np.random.default_rng(3)
    d = rng.uniform(0.2, 1.0, b.nbasis)
    d *= b.nelec / d.sum()
    return b, np.diag(d)


def test_xc_water64_product_grid_full_size():
    _banner("xc (H2O)64 75x302 product grid")
    b, D = _water64()
    assert b.nbasis == 1600
    eng = XcEngine(b)
    n = eng.build_grid(make_xc_opts(grid=QK_GRID_PRODUCT, nrad=75, nang=302))
    nb, mf, fl = eng.batch_info()
    print(f"points {n} batches {nb} mean functions/batch {mf:.0f} MFMA GFLOP {fl / 1e9:.1f}", flush=True)
    assert n > 1_000_000
    V, E, N = eng.xc(D)
    print(f"E_xc {E:.10f} N {N:.8f} (exact {b.nelec})", flush=True)
    assert abs(N - b.nelec) <= 1e-6 * b.nelec
    assert np.array_equal(V, V.T)
    V2, E2, N2 = eng.xc(D)
    assert np.array_equal(V, V2) and E == E2 and N == N2      # bitwise repeatable
    # the oracle on a sample of the grid points (every 2999th), same density
    pts, _ = eng.grid()
    sub = pts[::2999]
    eng.set_grid(sub)
    Vs, Es, Ns = eng.xc(D)
    Vo, Eo, ao, _ = O.xc(b, D, sub, eng.sigrad2())
    print(f"sample {len(sub)} points: |dE| {abs(Es - Eo):.2e} max|dV| {np.abs(Vs - Vo).max():.2e}", flush=True)
    assert abs(Es - Eo) <= TOL_E
    assert np.abs(Vs - Vo).max() <= TOL_V
