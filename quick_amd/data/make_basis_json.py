"""Extract the H, C, N, O blocks of a QUICK basis file into JSON.

The committed ``basis_6-31gss.json`` holds the standard 6-31G** exponents and
contraction coefficients (Hehre, Ditchfield & Pople 1972; Hariharan & Pople
1973) exactly as QUICK's ``basis/6-31GSS`` file lists them, so that the package
does not need the QUICK tree at run time.  Run once:

    python quick_amd/data/make_basis_json.py /root/reference/basis/6-31GSS
"""
import json
import sys
from pathlib import Path

KEEP = ("H", "C", "N", "O")


def parse(path):
    out, cur, lines = {}, None, Path(path).read_text().splitlines()
    i = 0
    while i < len(lines):
        tok = lines[i].split()
        i += 1
        if len(tok) == 2 and tok[1] == "0" and tok[0].isalpha():
            cur = tok[0]
            out.setdefault(cur, [])
            continue
        if not tok or tok[0] == "****" or cur is None:
            continue
        kind, nprim = tok[0], int(tok[1])
        prims = []
        for _ in range(nprim):
            prims.append([float(v.replace("D", "E")) for v in lines[i].split()])
            i += 1
        out[cur].append({"type": kind, "prims": prims})
    return {k: v for k, v in out.items() if k in KEEP}


if __name__ == "__main__":
    data = parse(sys.argv[1])
    dst = Path(__file__).with_name("basis_6-31gss.json")
    dst.write_text(json.dumps({"name": "6-31G**", "source": "QUICK basis/6-31GSS",
                               "elements": data}, indent=1))
    print(dst, {k: [s["type"] for s in v] for k, v in data.items()})
