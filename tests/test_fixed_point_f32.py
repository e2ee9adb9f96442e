"""The FP32 kernels' fixed-point conversion (fx_round_f in qk_kernels.h)
against the FP64 one it replaces, __double2ll_rn(double(v) * 2^48), i.e.
llrint(double(v) * 2^48): the function's own text compiled for the host with
g++, over random magnitudes of both signs, tiny values (whose fraction is all
of the value), ties at the last place and integers."""
import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np

CSRC = Path(__file__).resolve().parents[1] / "quick_amd" / "csrc"


def _compile(tmp_path):
    src = (CSRC / "qk_kernels.h").read_text()
    m = re.search(r"__device__ __forceinline__ long long fx_round_f\(float v\) \{.*?\n\}", src, re.S)
    assert m, "fx_round_f not found"
    cpp = tmp_path / "fx.cpp"
    cpp.write_text("#include <math.h>\n"
                   "#define __device__\n#define __forceinline__ inline\n" + m.group(0) +
                   "\nextern \"C\" void run(const float* v, long long* o, long long n) {"
                   " for (long long i = 0; i < n; ++i) o[i] = fx_round_f(v[i]); }\n")
    so = tmp_path / "fx.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", str(cpp), "-o", str(so)], check=True)
    return ctypes.CDLL(str(so))


def test_fx_round_f_equals_fp64_rounding(tmp_path):
    L = _compile(tmp_path)
    rng = np.random.default_rng(5)
    mag = 10.0 ** rng.uniform(-30, 4.4, 200000)
    v = (mag * rng.choice([-1.0, 1.0], mag.size)).astype(np.float32)
    # ties: v 2^48 = k + 1/2, and tiny values of both signs
    k = rng.integers(-2 ** 20, 2 ** 20, 2000)
    ties = ((k + 0.5) / 2.0 ** 48).astype(np.float32)
    tiny = np.array([1e-20, -1e-20, 2 ** -49, -2 ** -49, 3 * 2 ** -50, -3 * 2 ** -50, 0.0, -0.0,
                     2 ** -48, -2 ** -48, 1.0, -1.0, 32767.0, -32767.0], np.float32)
    v = np.concatenate([v, ties, tiny]).astype(np.float32)
    out = np.zeros(v.size, np.int64)
    L.run(v.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), ctypes.c_longlong(v.size))
    ref = np.rint(v.astype(np.float64) * 2.0 ** 48).astype(np.int64)
    bad = np.nonzero(out != ref)[0]
    assert bad.size == 0, (v[bad[:5]], out[bad[:5]], ref[bad[:5]])
