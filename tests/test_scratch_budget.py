"""Scratch (private segment) backing of every kernel the library launches.

The runtime backs a dispatch's scratch with per-lane scratch x 64 lanes x the
waves the grid can have in flight.  The library bounds the grid of every
spilling kernel (qk_scratch_grid_cap) so that this stays under 256 MiB per
dispatch; every build and gradient runs its class kernels on two streams, one
after another on each, so at most two backings are live at a time.  Here the per-lane scratch is read
from the built gfx950 code objects (no device needed) and the bound is checked
for every kernel.
"""
import ctypes
import re
from pathlib import Path

import pytest

from tests.code_objects import kernel_metadata

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "quick_amd" / "libquick_mi355x.so"
BUDGET = 256 * 1024 * 1024
MAX_WAVES = 256 * 32
# thread-per-quartet instantiations of the classes the wavefront-cooperative
# team kernel evaluates: compiled, never launched
UNLAUNCHED = {(2, 1, 2, 1), (2, 2, 1, 1), (2, 2, 2, 0), (2, 2, 2, 1), (2, 2, 2, 2)}


@pytest.fixture(scope="module")
def meta():
    if not LIB.exists():
        pytest.skip("library not built")
    return kernel_metadata(LIB)


def _class_of(name):
    m = re.search(r"fock_kernelILi(\d)ELi(\d)ELi(\d)ELi(\d)E", name)
    return tuple(int(x) for x in m.groups()) if m else None


def test_code_objects_list_the_kernels(meta):
    names = " ".join(meta)
    for k in ("fock_kernel", "fock_kernel_f32", "team_kernel", "grad_kernel", "grad_light_kernel",
              "schwarz_kernel", "compact_kernel", "fock_finalize"):
        assert k in names, k
    assert all("scratch" in v and "wg_size" in v for v in meta.values())


def test_every_launched_kernel_keeps_scratch_under_budget(meta):
    L = ctypes.CDLL(str(LIB))
    L.qk_scratch_grid_cap.restype = ctypes.c_longlong
    L.qk_scratch_grid_cap.argtypes = [ctypes.c_longlong, ctypes.c_int]
    spilling = 0
    for name, m in meta.items():
        if _class_of(name) in UNLAUNCHED and "f32" not in name:
            continue
        lane = m["scratch"]
        if lane == 0:
            continue
        spilling += 1
        wpb = max(1, m["wg_size"] // 64)
        cap = L.qk_scratch_grid_cap(lane, wpb)
        backing = lane * 64 * min(cap * wpb, MAX_WAVES)
        assert backing <= BUDGET, (name, lane, cap, backing)
        # a kernel never launches fewer than one block, whatever its scratch
        assert cap >= 1
    assert spilling > 0   # the check saw the spilling kernels


def test_scratch_cap_arithmetic():
    L = ctypes.CDLL(str(LIB))
    L.qk_scratch_grid_cap.restype = ctypes.c_longlong
    L.qk_scratch_grid_cap.argtypes = [ctypes.c_longlong, ctypes.c_int]
    assert L.qk_scratch_grid_cap(0, 4) >= 1 << 40          # no scratch: no bound
    assert L.qk_scratch_grid_cap(4096, 4) == BUDGET // (4096 * 64 * 4)
    assert L.qk_scratch_grid_cap(1 << 30, 4) == 1           # at least one block
