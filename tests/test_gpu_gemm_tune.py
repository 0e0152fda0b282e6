"""The GEMM tile autotuner may pick any of three kernels per shape; the 64x64,
128x64 and 128x64-with-8-waves kernels accumulate every output element in
the same order, so the products are bit-identical whichever is used: the
same batched products
(NN and NC, shapes that the row-waste rule and the tuner may split either
way) with each kernel forced (QT_GEMM_TILE, read once per process, hence one
short child process per setting) and with the tuner; likewise for the
three-multiplication form (QT_PREC_FP64_3M)."""
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
rng = np.random.default_rng(5)
out = {}
for i, (b, m, n, k) in enumerate([(9, 192, 160, 200), (4, 512, 512, 512), (3, 130, 70, 64), (16, 64, 64, 64)]):
    A = rng.standard_normal((b, m, k)) + 1j * rng.standard_normal((b, m, k))
    B = rng.standard_normal((b, k, n)) + 1j * rng.standard_normal((b, k, n))
    Bc = rng.standard_normal((b, n, k)) + 1j * rng.standard_normal((b, n, k))
    out[f"nn{i}"] = rgf.zgemm_batched(A, B, 0, 0)
    out[f"nc{i}"] = rgf.zgemm_batched(A, Bc, 0, 2)
    out[f"nn3m{i}"] = rgf.zgemm_batched(A, B, 0, 0, precision=rgf.QT_PREC_FP64_3M)
    out[f"nc3m{i}"] = rgf.zgemm_batched(A, Bc, 0, 2, precision=rgf.QT_PREC_FP64_3M)
np.savez(sys.argv[1], **out)
print("gemm child done", flush=True)
"""


def test_tile_choice_does_not_change_products(tmp_path):
    res = {}
    for name, env in (("narrow", {"QT_GEMM_TILE": "64x64"}), ("wide", {"QT_GEMM_TILE": "128x64"}),
                      ("w8", {"QT_GEMM_TILE": "128x64w8"}), ("tuned", {"QT_GEMM_TUNE_LOG": "1"})):
        print("gemm child", name, flush=True)
        e = dict(os.environ)
        e.pop("QT_GEMM_TILE", None)
        e.update(env)
        path = tmp_path / f"{name}.npz"
        r = subprocess.run([sys.executable, "-c", CHILD, str(path), ROOT], env=e, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        if name == "tuned":
            print(r.stderr)
            assert "qt gemm tune" in r.stderr
        res[name] = np.load(path)
    for key in res["narrow"].files:
        a = res["narrow"][key]
        for other in ("wide", "w8", "tuned"):
            assert np.array_equal(a, res[other][key]), (key, other)
