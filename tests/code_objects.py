"""Kernel resource metadata of the built library, read from its own gfx950 code
objects on the CPU (no device needed).

The HIP fat binary sits in the library's ``.hip_fatbin`` section as a sequence of
clang offload bundles (one per translation unit).  Each bundle's gfx950 entry
is an ELF code object whose AMDGPU metadata note lists, per kernel, its private
segment (scratch) size per lane, its VGPR count and spills; ``llvm-readelf
--notes`` prints that note.
"""
from __future__ import annotations

import shutil
import struct
import subprocess
import tempfile
from pathlib import Path

MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
LLVM = Path("/opt/rocm/llvm/bin")


def _tool(name: str) -> str | None:
    p = LLVM / name
    return str(p) if p.exists() else shutil.which(name)


def code_objects(so: Path) -> list[bytes]:
    objcopy = _tool("llvm-objcopy")
    with tempfile.TemporaryDirectory() as td:
        fb = Path(td) / "fatbin"
        subprocess.run([objcopy, f"--dump-section=.hip_fatbin={fb}", str(so), str(Path(td) / "lib")],
                       check=True, capture_output=True)
        data = fb.read_bytes()
    out = []
    start = data.find(MAGIC)
    while start >= 0:
        n = struct.unpack_from("<Q", data, start + len(MAGIC))[0]
        off = start + len(MAGIC) + 8
        for _ in range(n):
            o, size, ts = struct.unpack_from("<QQQ", data, off)
            off += 24
            triple = data[off:off + ts].decode()
            off += ts
            if "amdgcn" in triple and size:
                out.append(data[start + o:start + o + size])
        start = data.find(MAGIC, start + 1)
    return out


def kernel_metadata(so: Path) -> dict[str, dict[str, int]]:
    """{mangled kernel name: {scratch, vgpr, vgpr_spill, sgpr_spill, wg_size}}."""
    readelf = _tool("llvm-readelf")
    kernels: dict[str, dict[str, int]] = {}
    keys = {".private_segment_fixed_size:": "scratch", ".vgpr_count:": "vgpr", ".vgpr_spill_count:": "vgpr_spill",
            ".sgpr_spill_count:": "sgpr_spill", ".max_flat_workgroup_size:": "wg_size",
            ".group_segment_fixed_size:": "lds"}
    with tempfile.TemporaryDirectory() as td:
        for i, co in enumerate(code_objects(so)):
            if not co.startswith(b"\x7fELF"):
                continue   # a compressed bundle entry: not produced by this build's flags
            p = Path(td) / f"co{i}.elf"
            p.write_bytes(co)
            txt = subprocess.run([readelf, "--notes", str(p)], check=True, capture_output=True, text=True).stdout
            cur: dict[str, int] = {}
            name = None
            for line in txt.splitlines():
                s = line.strip()
                if s.startswith("- .agpr_count:") or s.startswith("- .args:"):
                    if name:
                        kernels[name] = cur
                    cur, name = {}, None
                if ".name:" in s and not s.startswith("- .name") and "_ZN" in s:
                    name = s.split(":", 1)[1].strip()
                for k, v in keys.items():
                    if s.startswith(k) or s.startswith("- " + k):
                        cur[v] = int(s.split(":", 1)[1].strip())
            if name:
                kernels[name] = cur
    return kernels
