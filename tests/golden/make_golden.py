"""Write the golden fixtures from the reference CPU path (container-only).

Runs ``oracle/_ref/qk_ref_driver`` (the reference Fortran built from its own
sources by ``oracle/ref_build.mk``) on small seeded systems and stores, per
system, one compressed ``.npz`` beside this script:

  * the reference's shell tables (kstart, kprim, katom, ktype, Qstart, Qfinal,
    Qsbasis, Qfbasis, ksumtype, KLMN, cons, gccoeff, gcexpo) and xyz in bohr;
  * Schwarz bounds ``Ycutoff`` (nshell^2) and primitive bounds ``cutprim``
    (jbasis^2) from schwarzoff (schwarz.f90:12-42);
  * a seeded synthetic density ``D`` and the two-electron operator ``G = G(D)``
    returned by hfoperator with a zero one-electron part (hfoperator.f90:24-138);
  * optionally the four-centre gradient term ``grad4c`` (gradient.f90:261-320).

Systems above 400 functions keep seeded samples of G instead of the full
matrices, and the density's seed instead of the density.

With ``--scf`` it instead runs the driver's ``scfgrad`` mode and stores, as
``<system>_scf.npz``, the reference at a converged RHF density: the occupied
orbitals (D = 2 C_occ C_occ^T), their energies, the total energy and its parts,
H_core, S and G(D) (sampled for systems above 400 functions), the full analytic
gradient of hfgrad and its four-centre term.

With ``--xc`` it runs the driver's ``xc`` mode (job card B3LYP) at the system's
converged HF density (from ``<system>_scf.npz``) and stores, as
``<system>_xc.npz``, the reference's SG-1 grid (points, weights >= 1e-10, parent
atom), the significance radii, V_xc (dftoperator's operator minus its operator
before the XC terms), E_xc, the alpha/beta electron counts, and the B3LYP
two-electron part G(D) with exchange weight 0.2 (the operator before the XC
terms minus the one-electron operator H of the SCF fixture).

usage (repository root):  python tests/golden/make_golden.py [--scf|--xc] [system ...]
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from quick_amd.geometry import small_system, synthetic_density, to_quick_input  # noqa: E402

DRIVER = ROOT / "oracle" / "_ref" / "qk_ref_driver"
BASIS_DIR = "/root/reference/basis"
KEYWORDS = "HF BASIS=6-31GSS CUTOFF=1.0e-11 DENSERMS=1.0e-7"

# system -> (do_fock, do_grad)
SYSTEMS = {
    "h2o": (True, True),
    "ch4": (True, True),
    "water2": (True, True),
    "ethane": (True, True),
    "water4": (True, True),
    "water32": (True, False),   # full benchmark size: ~10 min of reference CPU time
    "water64": (True, False),   # 1600 functions: ~1 h of reference CPU time
    "icosane": (False, True),   # 510 functions: the four-centre gradient only
}
SCF_SYSTEMS = ["h2o", "water2", "water4", "icosane", "water32"]
XC_SYSTEMS = ["h2o", "water2"]


def run_xc(name: str) -> Path:
    """The reference B3LYP operator (driver mode xc) at the converged HF density."""
    import re
    syms, xyz = small_system(name)
    scf = np.load(Path(__file__).with_name(f"{name}_scf.npz"))
    C = scf["C_occ"]
    D = 2.0 * C @ C.T
    env = dict(os.environ, QUICK_BASIS=BASIS_DIR)
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        (td / "job.in").write_text(to_quick_input(syms, xyz, KEYWORDS.replace("HF", "B3LYP")))
        od = td / "xc"
        od.mkdir()
        D.ravel(order="F").tofile(td / "D.bin")
        r = subprocess.run(f"ulimit -s unlimited && {DRIVER} job.in xc D.bin {od}", shell=True, check=True,
                           cwd=td, env=env, capture_output=True, text=True)
        m = re.search(r"Eelxc\s*=\s*([-+0-9.EeDd]+)", r.stdout)
        out = _read_setup(od, syms)
        n = int(out["nbasis"])
        npts = int(np.fromfile(od / "npts.bin", np.int32)[0])
        g = np.fromfile(od / "grid.bin").reshape(npts, 5)
        O = np.fromfile(od / "O.bin").reshape(n, n, order="F")
        Os = np.fromfile(od / "Osave.bin").reshape(n, n, order="F")
        sc = np.fromfile(od / "xc_scalars.bin")
        sig2 = np.fromfile(od / "sigrad2.bin")
    L = np.tril(O - Os)
    V = L + np.tril(L, -1).T
    Lg = np.tril(Os - scf["H"])
    G02 = Lg + np.tril(Lg, -1).T
    out.update(C_occ=C, grid_xyz=g[:, :3].copy(), grid_w=g[:, 3].copy(), grid_atom=g[:, 4].astype(np.int16),
               sigrad2=sig2, Vxc=V, Exc=np.array(float(m.group(1).replace("D", "E"))),
               aelec=np.array(sc[0]), belec=np.array(sc[1]), E_el=np.array(sc[2]), time_dft=np.array(sc[3]),
               G_b3lyp=G02)
    dst = Path(__file__).with_name(f"{name}_xc.npz")
    np.savez_compressed(dst, **out)
    return dst


def _read_setup(od: Path, syms) -> dict:
    natom, nbasis, nshell, _, jbasis, maxprim = np.fromfile(od / "sizes.bin", np.int32)
    rd = lambda n, dt=np.float64: np.fromfile(od / f"{n}.bin", dt)
    return dict(
        natom=natom, nbasis=nbasis, nshell=nshell, jbasis=jbasis, maxprim=maxprim,
        symbols=np.array(syms), xyz=rd("xyz").reshape(3, natom, order="F"),
        kstart=rd("kstart", np.int32), kprim=rd("kprim", np.int32),
        katom=rd("katom", np.int32), ktype=rd("ktype", np.int32),
        Qstart=rd("Qstart", np.int32), Qfinal=rd("Qfinal", np.int32),
        Qsbasis=rd("Qsbasis", np.int32).reshape(nshell, 4, order="F"),
        Qfbasis=rd("Qfbasis", np.int32).reshape(nshell, 4, order="F"),
        ksumtype=rd("ksumtype", np.int32),
        KLMN=rd("KLMN", np.int32).reshape(3, nbasis, order="F"),
        cons=rd("cons"),
        gccoeff=rd("gccoeff").reshape(maxprim, nbasis, order="F"),
        gcexpo=rd("gcexpo").reshape(maxprim, nbasis, order="F"),
        Ycutoff=rd("Ycutoff").reshape(nshell, nshell, order="F"),
        cutprim=rd("cutprim").reshape(jbasis, jbasis, order="F"),
    )


def _sample(out: dict, name: str, M: np.ndarray, rng_seed: int = 12345) -> None:
    """Seeded element samples, the diagonal and row norms of an n x n matrix."""
    n = M.shape[0]
    rng = np.random.default_rng(rng_seed)
    ii, jj = rng.integers(0, n, 40000), rng.integers(0, n, 40000)
    out.update({f"{name}_sample_i": ii.astype(np.int32), f"{name}_sample_j": jj.astype(np.int32),
                f"{name}_sample": M[ii, jj], f"{name}_diag": np.diag(M).copy(),
                f"{name}_rownorm": np.linalg.norm(M, axis=1)})


def run_scf(name: str) -> Path:
    """The reference at a converged RHF density (driver mode scfgrad)."""
    syms, xyz = small_system(name)
    env = dict(os.environ, QUICK_BASIS=BASIS_DIR)
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        (td / "job.in").write_text(to_quick_input(syms, xyz, KEYWORDS + " OPT"))
        od = td / "scf"
        od.mkdir()
        subprocess.run(f"ulimit -s unlimited && {DRIVER} job.in scfgrad none {od}",
                       shell=True, check=True, cwd=td, env=env, stdout=subprocess.DEVNULL)
        out = _read_setup(od, syms)
        n, natom = int(out["nbasis"]), int(out["natom"])
        rd = lambda f: np.fromfile(od / f"{f}.bin")
        E = rd("energy")
        nocc = int(sum({"H": 1, "C": 6, "N": 7, "O": 8}[x] for x in syms)) // 2
        C = rd("C").reshape(n, n, order="F")[:, :nocc].copy()
        D, H, S, G = (rd(f).reshape(n, n, order="F") for f in ("D", "H", "S", "G"))
        out.update(C_occ=C, eps=rd("eps")[:nocc + 1], E_total=np.array(E[0]), E_el=np.array(E[1]),
                   E_nn=np.array(E[2]), scf_cycles=np.array(E[3]), scf_rms=np.array(E[4]),
                   time_scf=np.array(E[5]), E2=np.array(0.5 * np.sum(D * G)), E1=np.array(np.sum(D * H)),
                   D_ref_minus_CCt=np.array(np.abs(D - 2.0 * C @ C.T).max()),
                   grad_total=rd("gradtotal").reshape(3, natom, order="F"),
                   grad4c=rd("grad4c").reshape(3, natom, order="F"),
                   time_grad=np.array(rd("time_grad")[0]))
    if n > 400:
        for k in ("cutprim", "gccoeff", "gcexpo"):
            out.pop(k)
        _sample(out, "G", G)
        _sample(out, "H", H, 54321)
    else:
        out.update(G=G, H=H, S=S)
    dst = Path(__file__).with_name(f"{name}_scf.npz")
    np.savez_compressed(dst, **out)
    return dst


def run(name: str, fock: bool, grad: bool, seed: int = 7) -> Path:
    syms, xyz = small_system(name)
    env = dict(os.environ, QUICK_BASIS=BASIS_DIR)
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        out = {}
        for mode in ["setup"] + (["fock"] if fock else []) + (["grad"] if grad else []):
            kw = KEYWORDS + (" OPT" if mode == "grad" else "")
            (td / "job.in").write_text(to_quick_input(syms, xyz, kw))
            od = td / mode
            od.mkdir()
            if mode != "setup":
                out["D"].ravel(order="F").tofile(td / "D.bin")
            subprocess.run(f"ulimit -s unlimited && {DRIVER} job.in {mode} D.bin {od}",
                           shell=True, check=True, cwd=td, env=env,
                           stdout=subprocess.DEVNULL)
            if mode == "setup":
                out.update(_read_setup(od, syms))
                out["D"] = synthetic_density(out["xyz"], out["katom"], out["ksumtype"],
                                             int(out["nbasis"]), seed)
            elif mode == "fock":
                out["G"] = np.fromfile(od / "G.bin").reshape(out["nbasis"], out["nbasis"], order="F")
                out["time_fock"] = np.fromfile(od / "time_fock.bin")[0]
            elif mode == "grad":
                out["grad4c"] = np.fromfile(od / "grad4c.bin").reshape(3, out["natom"], order="F")
                out["time_grad"] = np.fromfile(od / "time_grad.bin")[0]
    if int(out["nbasis"]) > 400:
        # full-size system: keep the fixture small -- seeded element samples, the
        # diagonal, row norms and the energy term instead of the full matrices;
        # the density is regenerated from its seed by synthetic_density
        D = out.pop("D")
        for k in ("cutprim", "gccoeff", "gcexpo"):
            out.pop(k)
        if "G" in out:
            G = out.pop("G")
            _sample(out, "G", G)
            out["E2"] = np.array(0.5 * np.sum(D * G))
        out["density_seed"] = np.array(seed)
        if int(out["nbasis"]) > 1200:
            Y = out.pop("Ycutoff")
            _sample(out, "Y", Y, 777)
    dst = Path(__file__).with_name(f"{name}.npz")
    np.savez_compressed(dst, **out)
    return dst


if __name__ == "__main__":
    args = sys.argv[1:]
    if args and args[0] == "--xc":
        for n in args[1:] or XC_SYSTEMS:
            p = run_xc(n)
            z = np.load(p)
            print(p.name, "nbasis", int(z["nbasis"]), "points", len(z["grid_w"]), "Exc", float(z["Exc"]),
                  "aelec", float(z["aelec"]), f"{p.stat().st_size/1024:.0f} KiB")
        sys.exit(0)
    if args and args[0] == "--scf":
        for n in args[1:] or SCF_SYSTEMS:
            p = run_scf(n)
            z = np.load(p)
            print(p.name, "nbasis", int(z["nbasis"]), "E", float(z["E_total"]), "cycles", int(z["scf_cycles"]),
                  "scf s", float(z["time_scf"]), "grad s", float(z["time_grad"]),
                  f"{p.stat().st_size/1024:.0f} KiB")
        sys.exit(0)
    names = args or list(SYSTEMS)
    for n in names:
        f, g = SYSTEMS[n]
        p = run(n, f, g)
        z = np.load(p)
        print(p.name, "nbasis", int(z["nbasis"]), "fock s", z.get("time_fock"),
              "grad s", z.get("time_grad"), f"{p.stat().st_size/1024:.0f} KiB")
