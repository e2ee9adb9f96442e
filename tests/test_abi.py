"""The C-ABI library loads on the CPU and exports every symbol the header declares."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "quick_mi355x.h"
LIB = ROOT / "quick_amd" / "libquick_mi355x.so"


def declared_symbols():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    names = re.findall(r"\b((?:gpu|qk)_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_reference_boundary():
    names = declared_symbols()
    for ref in ["gpu_startup_", "gpu_set_device_", "gpu_init_", "gpu_get_device_info_", "gpu_setup_",
                "gpu_upload_xyz_", "gpu_upload_atom_and_chg_", "gpu_upload_basis_",
                "gpu_upload_cutoff_matrix_", "gpu_upload_method_", "gpu_upload_calculated_",
                "gpu_upload_cutoff_", "gpu_get2e_", "gpu_upload_grad_", "gpu_grad_", "gpu_getxc_",
                "gpu_aoint_", "gpu_addint_", "gpu_cleanup_", "gpu_shutdown_"]:
        assert ref in names


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_no_device_is_an_error_not_a_crash():
    import os
    import subprocess
    import sys
    # a child process with no visible device: qk_create must fail cleanly
    code = ("import ctypes; L=ctypes.CDLL(%r); L.qk_create.restype=ctypes.c_void_p; "
            "L.qk_last_error.restype=ctypes.c_char_p; h=L.qk_create(0); "
            "print('NULL' if not h else 'HANDLE', L.qk_last_error().decode())" % str(LIB))
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("NULL")


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_abi_build_before_upload_is_fatal_with_message():
    """gpu_get2e_ before any upload: the library aborts with a message, as the
    reference exits (gpu_common.h:142-146); a child process with no visible
    device, so that nothing reaches a GPU."""
    import os
    import subprocess
    import sys
    code = ("import ctypes, numpy as np; L = ctypes.CDLL(%r); o = np.zeros(4); "
            "L.gpu_get2e_(o.ctypes.data_as(ctypes.c_void_p))" % str(LIB))
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "gpu_get2e: density not uploaded" in r.stderr
