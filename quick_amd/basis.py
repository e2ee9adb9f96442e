"""Contracted Cartesian basis in QUICK's shell-table form.

This restates what QUICK's ``readbasis`` builds (basis.f90:377-599 together with
``xnorm``/``xnewnorm``, subs/xnorm.f90:21-97): the per-shell tables ``kstart``,
``kprim``, ``katom``, ``ktype``, ``Qstart``/``Qfinal``, ``Qsbasis``/``Qfbasis``,
``ksumtype`` and the per-function ``KLMN``, ``cons``, ``gcexpo`` and ``gccoeff``.
These are exactly the arrays QUICK's Fortran hands to ``gpu_upload_basis_``, so
the host code that feeds the C-ABI in tests and benchmarks produces the same
bytes the Fortran callers would.

Conventions kept from the reference (1-based, Fortran column-major):
  * shell types: S -> ktype 1 (Q 0), P -> 3 (Q 1), SP -> 4 (Q 0..1), D -> 6 (Q 2);
  * d components in the order xx, xy, yy, xz, yz, zz with ``cons`` = 1, sqrt3, 1,
    sqrt3, sqrt3, 1 (basis.f90:551-574);
  * the d contraction is renormalised in two groups: (xx, xy, yy) with the
    (2,0,0) factor computed from the yy coefficients and (xz, yz, zz) with the
    (1,1,0) factor computed from the zz coefficients (basis.f90:588-599).
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .geometry import ANGSTROM_TO_BOHR, SYMBOL_TO_Z

MAXPRIM = 10  # quick_size_module.f90:37
PI = 3.1415926535897932384626433832795
PITO3HALF = PI ** 1.5

_DATA = Path(__file__).with_name("data") / "basis_6-31gss.json"


def xnorm(a: float, i: int, j: int, k: int) -> float:
    """Primitive normalisation of x^i y^j z^k exp(-a r^2) (subs/xnorm.f90:21-58)."""
    v = (2.0 * a) ** (float(i + j + k) / 2.0 + 0.75)
    g = 1.0
    for n in (i, j, k):
        for L in range(1, n + 1):
            g *= float(n - L) + 0.5
    return v * PI ** (-0.75) * g ** (-0.5)


def xnewnorm(i: int, j: int, k: int, coef, expo) -> float:
    """Contraction renormalisation factor (subs/xnorm.f90:61-97)."""
    s = 0.0
    xp = -1.5 + float(-i - j - k)
    n = len(coef)
    for c1 in range(n):
        for c2 in range(n):
            s += coef[c1] * coef[c2] * (expo[c1] + expo[c2]) ** xp
    g = 1.0
    for m in (i, j, k):
        for L in range(1, m + 1):
            g *= float(m - L) + 0.5
    return (s * g * PITO3HALF) ** (-0.5)


@dataclass
class Basis:
    symbols: list
    Z: np.ndarray           # (natom,)
    xyz: np.ndarray         # (3, natom) bohr, Fortran layout
    nshell: int
    nbasis: int
    jbasis: int             # total primitives
    kstart: np.ndarray      # (nshell,) 1-based first primitive
    kprim: np.ndarray
    katom: np.ndarray       # 1-based atom
    ktype: np.ndarray
    Qnumber: np.ndarray
    Qstart: np.ndarray
    Qfinal: np.ndarray
    Qsbasis: np.ndarray     # (nshell, 4)
    Qfbasis: np.ndarray
    ksumtype: np.ndarray    # (nshell+1,) 1-based first function
    KLMN: np.ndarray        # (3, nbasis)
    cons: np.ndarray        # (nbasis,)
    gccoeff: np.ndarray     # (MAXPRIM, nbasis)
    gcexpo: np.ndarray      # (MAXPRIM, nbasis)
    nelec: int = 0
    meta: dict = field(default_factory=dict)
    # per-function contraction of the Cartesian functions the XC step evaluates
    # (normalize_basis, quick_basis_module_.f90:331-373): each component
    # normalised on its own, coefficients dcoeff over exponents aexp
    dcoeff: np.ndarray | None = None   # (MAXPRIM, nbasis)
    aexp: np.ndarray | None = None     # (MAXPRIM, nbasis)
    ncontract: np.ndarray | None = None   # (nbasis,)
    ncenter: np.ndarray | None = None     # (nbasis,) 1-based atom

    @property
    def natom(self) -> int:
        return len(self.symbols)


def load_basis_data(path: Path = _DATA) -> dict:
    return json.loads(Path(path).read_text())["elements"]


def build_basis(symbols, xyz_ang, charge: int = 0, data: dict | None = None) -> Basis:
    """Shell tables for atoms ``symbols`` at ``xyz_ang`` (Angstrom).

    Coordinates are rounded to the 10 decimals QUICK input decks carry and then
    converted with QUICK's own factor, so the bohr values are bit-identical to
    what the reference reads from the same deck.
    """
    data = data or load_basis_data()
    xyz_ang = np.asarray(xyz_ang, float)
    xyz = np.array([[float(f"{v:.10f}") * ANGSTROM_TO_BOHR for v in row] for row in xyz_ang]).T.copy()
    Z = np.array([SYMBOL_TO_Z[s] for s in symbols], np.int32)

    shells = []  # (atom, type, prims)
    for ia, s in enumerate(symbols):
        for sh in data[s]:
            shells.append((ia + 1, sh["type"], sh["prims"]))
    nshell = len(shells)
    nbf_of = {"S": 1, "P": 3, "SP": 4, "D": 6}
    nbasis = sum(nbf_of[t] for _, t, _ in shells)
    jbasis = sum(len(p) for _, _, p in shells)

    I = lambda *shape: np.zeros(shape, np.int32)
    kstart, kprim, katom, ktype = I(nshell), I(nshell), I(nshell), I(nshell)
    Qnumber, Qstart, Qfinal = I(nshell), I(nshell), I(nshell)
    Qsbasis, Qfbasis, ksumtype = I(nshell, 4), I(nshell, 4), I(nshell + 1)
    KLMN = I(3, nbasis)
    cons = np.zeros(nbasis)
    gccoeff = np.zeros((MAXPRIM, nbasis))
    gcexpo = np.zeros((MAXPRIM, nbasis))
    dcoeff = np.zeros((MAXPRIM, nbasis))
    aexp = np.zeros((MAXPRIM, nbasis))
    ncontract = I(nbasis)
    ncenter = I(nbasis)

    jb = 1  # running primitive index (1-based)
    ni = 0  # running function count (Ninitial)
    for js, (atom, t, prims) in enumerate(shells):
        npr = len(prims)
        aa = [p[0] for p in prims]
        kprim[js], katom[js], kstart[js] = npr, atom, jb
        ksumtype[js] = ni + 1
        if t == "S":
            ktype[js], Qnumber[js], Qstart[js], Qfinal[js] = 1, 0, 0, 0
            bb = [p[1] for p in prims]
            ni += 1
            cons[ni - 1] = 1.0
            c = [bb[k] * xnorm(aa[k], 0, 0, 0) for k in range(npr)]
            f = xnewnorm(0, 0, 0, c, aa)
            gccoeff[:npr, ni - 1] = [f * v for v in c]
            gcexpo[:npr, ni - 1] = aa
        elif t == "P":
            ktype[js], Qnumber[js], Qstart[js], Qfinal[js] = 3, 6, 1, 1
            Qsbasis[js, 1], Qfbasis[js, 1] = 0, 2
            bb = [p[1] for p in prims]
            c = [bb[k] * xnorm(aa[k], 1, 0, 0) for k in range(npr)]
            f = xnewnorm(1, 0, 0, c, aa)
            for d in range(3):
                ni += 1
                cons[ni - 1] = 1.0
                KLMN[d, ni - 1] = 1
                gccoeff[:npr, ni - 1] = [f * v for v in c]
                gcexpo[:npr, ni - 1] = aa
        elif t == "SP":
            ktype[js], Qnumber[js], Qstart[js], Qfinal[js] = 4, 1, 0, 1
            Qsbasis[js, 0], Qfbasis[js, 0], Qsbasis[js, 1], Qfbasis[js, 1] = 0, 0, 1, 3
            bb = [p[1] for p in prims]
            cc = [p[2] for p in prims]
            ni += 1
            cons[ni - 1] = 1.0
            c = [bb[k] * xnorm(aa[k], 0, 0, 0) for k in range(npr)]
            f = xnewnorm(0, 0, 0, c, aa)
            gccoeff[:npr, ni - 1] = [f * v for v in c]
            gcexpo[:npr, ni - 1] = aa
            c = [cc[k] * xnorm(aa[k], 1, 0, 0) for k in range(npr)]
            f = xnewnorm(1, 0, 0, c, aa)
            for d in range(3):
                ni += 1
                cons[ni - 1] = 1.0
                KLMN[d, ni - 1] = 1
                gccoeff[:npr, ni - 1] = [f * v for v in c]
                gcexpo[:npr, ni - 1] = aa
        elif t == "D":
            ktype[js], Qnumber[js], Qstart[js], Qfinal[js] = 6, 2, 2, 2
            Qsbasis[js, 2], Qfbasis[js, 2] = 0, 5
            bb = [p[1] for p in prims]
            comps = [((2, 0, 0), 1.0), ((1, 1, 0), math.sqrt(3.0)), ((0, 2, 0), 1.0),
                     ((1, 0, 1), math.sqrt(3.0)), ((0, 1, 1), math.sqrt(3.0)), ((0, 0, 2), 1.0)]
            first = ni
            raw = []
            for (l, m, n), cn in comps:
                ni += 1
                KLMN[:, ni - 1] = (l, m, n)
                cons[ni - 1] = cn
                raw.append([bb[k] * xnorm(aa[k], l, m, n) for k in range(npr)])
                gcexpo[:npr, ni - 1] = aa
            f1 = xnewnorm(2, 0, 0, raw[2], aa)   # from the yy coefficients
            f2 = xnewnorm(1, 1, 0, raw[5], aa)   # from the zz coefficients
            for q in range(6):
                f = f1 if q < 3 else f2
                gccoeff[:npr, first + q] = [f * v for v in raw[q]]
        else:
            raise ValueError(f"shell type {t} not supported")
        # the XC step's per-function contraction: raw coefficient x xnorm of the
        # function's own (l, m, n), renormalised per function
        sp_first = ksumtype[js] - 1
        for f in range(sp_first, ni):
            l, m, n = (int(v) for v in KLMN[:, f])
            col = 2 if (t == "SP" and f > sp_first) else 1
            c = [p[col] * xnorm(aa[k], l, m, n) for k, p in enumerate(prims)]
            fn = xnewnorm(l, m, n, c, aa)
            dcoeff[:npr, f] = [fn * v for v in c]
            aexp[:npr, f] = aa
            ncontract[f] = npr
            ncenter[f] = atom
        jb += npr
    ksumtype[nshell] = ni + 1
    nelec = int(Z.sum()) - charge
    return Basis(list(symbols), Z, xyz, nshell, nbasis, jbasis, kstart, kprim, katom, ktype,
                 Qnumber, Qstart, Qfinal, Qsbasis, Qfbasis, ksumtype, KLMN, cons, gccoeff,
                 gcexpo, nelec, dcoeff=dcoeff, aexp=aexp, ncontract=ncontract, ncenter=ncenter)


def qshells(b: Basis):
    """Split shells into single-angular-momentum Qshells.

    Returns a list of (shell0, L, first_function0, nfunc) with 0-based indices,
    ordered by shell and then by L (the S part of an SP shell first).
    """
    out = []
    nf = {0: 1, 1: 3, 2: 6}
    for s in range(b.nshell):
        for L in range(b.Qstart[s], b.Qfinal[s] + 1):
            f0 = b.ksumtype[s] - 1 + b.Qsbasis[s, L]
            out.append((s, L, f0, nf[L]))
    return out


def write_quick_basis_file(path, data: dict | None = None) -> None:
    """Write basis data in the text layout QUICK's readbasis parses
    (basis.f90:386-400): an "<element>  0" line, then per shell "<type> <nprim>
    1.00" followed by one line per primitive, and "****" after each element."""
    data = data or load_basis_data()
    lines = []
    for el, shells in data.items():
        lines.append(f"{el:<6s}0")
        for sh in shells:
            lines.append(f"{sh['type']:<6s}{len(sh['prims'])}     1.00")
            for pr in sh["prims"]:
                lines.append("    " + "".join(f"{float(v)!r:>24}" for v in pr))
        lines.append("****")
    Path(path).write_text("\n".join(lines) + "\n")
