"""The two int8 product kernels of 6 digits (128x64 direct-to-LDS
double-buffered tiles, and the 64x64 kernel with register-staged copies,
forced by QT_I8_KERNEL=1, read once per process, hence one short child
process per setting) add the same integer digit products into int32
accumulators, where every order gives the same sums, and fold them
identically: the mixed products are bit-identical whichever kernel runs, on
shapes that leave partial tiles in both kernels.  7 digits run the 64x64
kernel in both settings, and are checked the same way."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[2])
from quantum_transport_amd import rgf
rgf.init(0)
rng = np.random.default_rng(8)
out = {}
for i, (b, m, n, k) in enumerate([(5, 192, 160, 200), (3, 512, 512, 512), (4, 130, 70, 64), (9, 64, 64, 1000)]):
    A = rng.standard_normal((b, m, k)) + 1j * rng.standard_normal((b, m, k))
    B = rng.standard_normal((b, k, n)) + 1j * rng.standard_normal((b, k, n))
    Bc = rng.standard_normal((b, n, k)) + 1j * rng.standard_normal((b, n, k))
    for p in (rgf.QT_PREC_MIXED, rgf.QT_PREC_MIXED47):
        out[f"nn{i}p{p}"] = rgf.zgemm_batched(A, B, 0, 0, precision=p)
        out[f"nc{i}p{p}"] = rgf.zgemm_batched(A, Bc, 0, 2, precision=p)
np.savez(sys.argv[1], **out)
print("i8 child done", flush=True)
"""


def test_i8_kernels_bit_identical(tmp_path):
    res = {}
    for name, val in (("lds", None), ("staged", "1")):
        print("i8 child", name, flush=True)
        e = dict(os.environ)
        e.pop("QT_I8_KERNEL", None)
        if val is not None:
            e["QT_I8_KERNEL"] = val
        path = tmp_path / f"{name}.npz"
        r = subprocess.run([sys.executable, "-c", CHILD, str(path), ROOT], env=e, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        res[name] = np.load(path)
    for key in res["lds"].files:
        a, b = res["lds"][key], res["staged"][key]
        assert np.array_equal(a, b), key
        assert np.isfinite(a).all(), key
