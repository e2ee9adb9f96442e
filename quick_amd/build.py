"""Builds the HIP library libquick_mi355x.so in-tree for gfx950.

    python -m quick_amd.build        (or __graft_entry__.build())

Steps: regenerate the per-class VRR/HRR device code (csrc/gen_eri.py) and the
gradient tables (csrc/gen_grad.py), compile every translation unit with hipcc
(in parallel, into _build/), and link the shared library next to this file.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libquick_mi355x.so"
ARCH = os.environ.get("QK_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed-FP32 instructions (v_pk_*_f32) in any generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    # FP32 division and square root by the hardware approximations (the FP32
    # kernels serve the mixed-precision path, whose error budget is 1e-6;
    # FP64 code is unaffected)
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-Wno-unused-result",
]


OBJ = HERE / "_build"


def _compile(src: Path, verbose: bool) -> Path:
    obj = OBJ / (src.stem + ".o")
    deps = [src, CSRC / "qk_kernels.h", CSRC / "qk_internal.h", CSRC / "eri_generated.inc",
            CSRC / "qk_team.h", CSRC / "team_tables.inc", CSRC / "grad_tables.inc", CSRC / "grad_light.inc",
            CSRC / "qk_grad_light.h", CSRC / "xc_b3lyp.inc", CSRC / "lebedev.inc",
            HERE.parent / "include" / "quick_mi355x.h", Path(__file__)]
    if obj.exists() and all(obj.stat().st_mtime >= d.stat().st_mtime for d in deps if d.exists()):
        return obj
    flags = [f for f in HIP_FLAGS if f != "-shared"]
    cmd = ["hipcc", *flags, "-I", str(HERE.parent / "include"), "-c", str(src), "-o", str(obj)]
    if verbose:
        print("  hipcc", src.name, flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose: bool = True, jobs: int = 8) -> Path:
    from concurrent.futures import ThreadPoolExecutor
    OBJ.mkdir(exist_ok=True)
    subprocess.run([sys.executable, str(CSRC / "gen_eri.py"), str(CSRC / "eri_generated.inc"),
                    str(CSRC / "gen")], check=True, stdout=subprocess.DEVNULL)
    subprocess.run([sys.executable, str(CSRC / "gen_grad.py"), str(CSRC / "grad_tables.inc")], check=True,
                   stdout=subprocess.DEVNULL)
    subprocess.run([sys.executable, str(CSRC / "gen_xc.py"), str(CSRC / "xc_b3lyp.inc"), str(CSRC / "lebedev.inc")],
                   check=True, stdout=subprocess.DEVNULL)
    srcs = sorted(list((CSRC / "gen").glob("fock_*.hip")) + list((CSRC / "gen").glob("grad_*.hip")),
                  key=lambda p: -p.stat().st_size)
    srcs += [CSRC / "qk_xc.hip", CSRC / "qk_oneint.hip", CSRC / "qk_scf.hip", CSRC / "qk_lists.hip", CSRC / "qk_fock.hip",
             CSRC / "qk_grad.hip", CSRC / "qk_host.cpp"]
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    # rocBLAS and rocSOLVER serve the SCF step (qk_scf.hip)
    cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp),
           "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB, flush=True)
    return LIB


if __name__ == "__main__":
    build()
