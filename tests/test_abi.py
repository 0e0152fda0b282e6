"""The C ABI without a GPU: the library loads, exports every function
include/qt_rgf.h declares, the ctypes and Fortran mirrors of its structs
have the C layout, and calls fail cleanly before qt_init / without a device."""
import ctypes as C
import os
import re
import subprocess

import pytest

from quantum_transport_amd import rgf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qt_rgf.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qt_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("qt_init", "qt_finish", "qt_ham_create", "qt_rgf_solve", "qt_sancho",
                     "qt_zgemm_batched", "qt_invert_batched"):
        assert required in names


def test_library_exports_every_declared_symbol():
    L = rgf.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", rgf.LIBPATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (qt_\w+)", out))
    assert set(declared_functions()) <= exported


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(f"""
#include <stdio.h>
#include <stddef.h>
#include "{HEADER}"
int main(void) {{
  printf("%zu %zu %zu %zu\\n", sizeof(qt_rgf_params), offsetof(qt_rgf_params, mul),
         offsetof(qt_rgf_params, lead_tol), sizeof(qt_rgf_outputs));
  printf("%zu %zu %zu\\n", offsetof(qt_rgf_outputs, gr_diag), offsetof(qt_rgf_outputs, Gln),
         sizeof(qt_timing));
  printf("%zu %zu\\n", offsetof(qt_timing, gemm_launches), offsetof(qt_timing, batch));
  return 0;
}}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    assert vals[0] == C.sizeof(rgf.Params)
    assert vals[1] == rgf.Params.mul.offset
    assert vals[2] == rgf.Params.lead_tol.offset
    assert vals[3] == C.sizeof(rgf.Outputs)
    assert vals[4] == rgf.Outputs.gr_diag.offset
    assert vals[5] == rgf.Outputs.Gln.offset
    assert vals[6] == C.sizeof(rgf.Timing)
    assert vals[7] == rgf.Timing.gemm_launches.offset
    assert vals[8] == rgf.Timing.batch.offset


def test_calls_before_init_fail_cleanly():
    L = rgf.lib()
    p = rgf.make_params()
    o = rgf.Outputs()
    assert L.qt_rgf_solve(None, C.byref(p), C.c_int32(1), None, None, None, None, C.byref(o)) == 2
    assert L.qt_last_timing(None) == 2
    assert L.qt_status_string(2) == b"qt_init() has not been called"


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a HIP device is visible")
def test_init_without_device_reports_no_device():
    assert rgf.lib().qt_init(C.c_int(0)) == 7
    with pytest.raises(rgf.QTError):
        rgf.init(0)


def test_fortran_modules_export_reference_names():
    lib = os.path.join(ROOT, "quantum_transport_amd", "fortran", "_build", "libqtfortran.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.dirname(os.path.dirname(lib)), "-s"], check=True)
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True,
                         check=True).stdout.lower()
    for mod, sub in (("rgf_mod", "rgf_variableblock_forward"), ("rgf_mod", "rgf_variableblock_backward"),
                     ("cuda_rgf_mod", "cuda_rgf_constblocksize"), ("cuda_rgf_mod", "cuda_rgf_init"),
                     ("cuda_rgf_mod", "cuda_rgf_finish"), ("negf_mod", "negf_solve"),
                     ("selfenergy_mod", "sancho"), ("output", "write_spectrum_summed_over_k"),
                     ("devicehammod", "devh_build_fromwannierfile")):
        assert sub in out, sub


def _header_struct_fields(name):
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + r"\s*;", text).group(1)
    fields = []
    for decl in body.split(";"):
        parts = [p.strip() for p in decl.strip().split(",")] if decl.strip() else []
        fields += [re.split(r"[\s*]+", p)[-1] for p in parts]
    return fields


def _fortran_type_fields(text, name):
    body = re.search(r"type, bind\(C\) :: " + name + r"(.*?)end type", text, flags=re.S | re.I).group(1)
    fields = []
    for line in body.splitlines():
        line = line.split("!")[0]
        if "::" not in line:
            continue
        for item in line.split("::", 1)[1].split(","):
            fields.append(item.split("=")[0].strip())
    return fields


@pytest.mark.parametrize("where", ["INTEGRATION.md", "quantum_transport_amd/fortran/qt_capi.f90"])
def test_documented_fortran_binding_matches_header(where):
    """The Fortran mirror of qt_rgf_outputs a maintainer copies has every
    pointer of the C struct, in the C order (a shorter type would let
    qt_rgf_solve read past the caller's struct)."""
    text = open(os.path.join(ROOT, where)).read()
    c_fields = [f.lower() for f in _header_struct_fields("qt_rgf_outputs")]
    f_fields = [f.lower() for f in _fortran_type_fields(text, "qt_rgf_outputs")]
    assert f_fields == c_fields, (where, f_fields, c_fields)
    c_par = [f.lower() for f in _header_struct_fields("qt_rgf_params")]
    f_par = [f.lower() for f in _fortran_type_fields(text, "qt_rgf_params")]
    assert f_par == c_par, (where, f_par, c_par)
