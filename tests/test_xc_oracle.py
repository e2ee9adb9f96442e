"""The oracle's XC restatement against the reference's own dftoperator output.

tests/golden/{h2o,water2}_xc.npz hold, for the converged HF density of each
system, the reference's SG-1 grid (every point dftoperator visits with weight
>= 1e-10), its radii of significance, its V_xc (the operator after the B3LYP XC
terms minus the operator before them), E_xc and the alpha/beta electron counts
(tests/golden/make_golden.py --xc).
"""
import numpy as np
import pytest

from oracle import oracle as O
from quick_amd.basis import build_basis
from quick_amd.geometry import small_system

SYSTEMS = ["h2o", "water2"]


def _load(name):
    z = np.load(f"tests/golden/{name}_xc.npz")
    b = build_basis(*small_system(name))
    C = z["C_occ"]
    return z, b, 2.0 * C @ C.T


@pytest.mark.parametrize("name", SYSTEMS)
def test_sigrad_matches_reference(name):
    z, b, _ = _load(name)
    assert np.array_equal(O.sigrad2(b), z["sigrad2"])


@pytest.mark.parametrize("name", SYSTEMS)
def test_sg1_grid_matches_reference(name):
    z, b, _ = _load(name)
    g = O.sg1_grid(b.xyz, b.Z)
    ref = np.column_stack([z["grid_xyz"], z["grid_w"], z["grid_atom"] - 1])
    assert len(g) == len(ref)
    # same enumeration order (atom, radial shell, Lebedev point): compare as sets
    key = lambda a: np.lexsort((np.round(a[:, 2], 9), np.round(a[:, 1], 9), np.round(a[:, 0], 9), a[:, 4]))
    gs, rs = g[key(g)], ref[key(ref)]
    assert np.abs(gs[:, :3] - rs[:, :3]).max() <= 1e-12
    assert np.abs(gs[:, 3] - rs[:, 3]).max() <= 1e-12 * np.abs(rs[:, 3]).max()
    assert np.array_equal(gs[:, 4], rs[:, 4])


@pytest.mark.parametrize("name", SYSTEMS)
def test_xc_on_reference_grid_matches_reference(name):
    z, b, D = _load(name)
    pts = np.column_stack([z["grid_xyz"], z["grid_w"]])
    V, E, a, bb = O.xc(b, D, pts, z["sigrad2"])
    assert abs(E - float(z["Exc"])) <= 1e-12
    assert np.abs(V - z["Vxc"]).max() <= 1e-12
    assert abs(a - float(z["aelec"])) <= 1e-12 and abs(bb - float(z["belec"])) <= 1e-12
    # the electron count of the closed-shell density: N/2 per spin
    assert abs(a - 0.5 * int(b.Z.sum())) <= 1e-4


def test_b3lyp_derivatives_against_finite_differences():
    """df/drho and df/dsigma against central differences of f."""
    for rho, sigma in [(1e-4, 1e-8), (0.05, 0.01), (0.3, 0.2), (2.0, 1.5), (50.0, 800.0)]:
        f, vr, vs = O.b3lyp(rho, sigma)
        hr, hs = 1e-6 * rho, 1e-6 * sigma
        fr = (O.b3lyp(rho + hr, sigma)[0] - O.b3lyp(rho - hr, sigma)[0]) / (2 * hr)
        fs = (O.b3lyp(rho, sigma + hs)[0] - O.b3lyp(rho, sigma - hs)[0]) / (2 * hs)
        assert abs(vr - fr) <= 1e-7 * max(1.0, abs(vr))
        assert abs(vs - fs) <= 1e-7 * max(1.0, abs(vs))


def test_b3lyp_uniform_limits():
    """At sigma = 0 the B88 correction vanishes, the LYP term reduces to
    -a rho / (1 + d rho^-1/3) and the rest is 0.8 Slater + 0.19 VWN-RPA; at
    rho -> large the Slater term dominates: f / rho^(4/3) -> -0.8 (3/4)(3/pi)^(1/3)."""
    cx = 0.75 * (3.0 / np.pi) ** (1.0 / 3.0)
    rho = 1e12
    f, _, _ = O.b3lyp(rho, 0.0)
    assert abs(f / rho ** (4.0 / 3.0) + 0.8 * cx) <= 1e-3
    # the gradient terms lower the energy density at fixed rho (B88 is negative definite)
    assert O.b3lyp(0.1, 0.05)[0] < O.b3lyp(0.1, 0.0)[0] + 0.81 * 0.1 + 1.0   # finite
    f0 = O.b3lyp(0.1, 1e-30)[0]
    assert np.isfinite(f0)


def test_oracle_planted_mistake_is_seen():
    """A wrong density (scaled by 1.01) moves E_xc by far more than the bar."""
    z, b, D = _load("h2o")
    pts = np.column_stack([z["grid_xyz"], z["grid_w"]])
    _, E, _, _ = O.xc(b, 1.01 * D, pts, z["sigrad2"])
    assert abs(E - float(z["Exc"])) > 1e-3


def test_device_functional_matches_oracle(tmp_path):
    """The generated functional the XC kernels evaluate (xc_b3lyp.inc, from
    gen_xc.py's symbolic derivatives) compiled for the host with g++, against
    the oracle's complex-step functional over the density range of real grids."""
    import ctypes
    import subprocess
    import sys
    from pathlib import Path
    csrc = Path(__file__).resolve().parents[1] / "quick_amd" / "csrc"
    sys.path.insert(0, str(csrc))
    import gen_xc
    inc = tmp_path / "xc_b3lyp.inc"
    gen_xc.emit(str(inc))
    src = tmp_path / "f.cpp"
    src.write_text('#include <cmath>\nusing std::pow; using std::sqrt; using std::exp; using std::log;\n'
                   'using std::asinh; using std::atan;\n#define QK_HD\n#include "xc_b3lyp.inc"\n'
                   'extern "C" void run(double r, double s, double* o) { b3lyp_closed<double>(r, s, o, o + 1, o + 2); }\n')
    so = tmp_path / "f.so"
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-I", str(tmp_path), str(src), "-o", str(so)], check=True)
    L = ctypes.CDLL(str(so))
    out = (ctypes.c_double * 3)()
    rng = np.random.default_rng(1)
    for _ in range(400):
        rho = 10.0 ** rng.uniform(-9, 3)
        sigma = rho ** (8.0 / 3.0) * 10.0 ** rng.uniform(-4, 3)
        L.run(ctypes.c_double(rho), ctypes.c_double(sigma), out)
        f, vr, vs = O.b3lyp(rho, sigma)
        assert abs(out[0] - f) <= 1e-12 * max(abs(f), 1e-300) + 1e-300
        assert abs(out[1] - vr) <= 1e-10 * abs(vr)
        assert abs(out[2] - vs) <= 1e-10 * abs(vs)
