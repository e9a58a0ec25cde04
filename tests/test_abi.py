"""The C-ABI library loads without a GPU and exports exactly what
include/curvami_gpu.h declares; the binding covers every declaration."""
import ctypes
import os
import re
import subprocess

import pytest

import curvami_gpu as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "curvami_gpu.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"CVM_API\s+[\w\s\*]+?\b(cvm_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "cvm_assemble" in names and "cvm_shape_functions" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(G.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(G.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), f"{name} not exported"
    out = subprocess.run(["nm", "-D", "--defined-only", G.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l and l.split()[-1].startswith("cvm_")}
    assert exported == set(declared()), exported ^ set(declared())


def test_binding_covers_header():
    assert set(G.SIGNATURES) == set(declared())


def test_no_device_fails_loudly():
    """On a host without a HIP device the product path refuses to run (no CPU fallback)."""
    if not os.path.exists(G.LIB_PATH):
        pytest.skip("library not built")
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("GPU present")
    with pytest.raises(G.CvmError):
        G.Plate(0)


KKT_HEADER = os.path.join(ROOT, "include", "curvami_kkt.h")


def test_kkt_library_exports_every_declared_symbol():
    """libcurvami_kkt.so exports exactly what include/curvami_kkt.h declares."""
    names = set(re.findall(r"CVM_KKT_API\s+[\w\s\*]+?\b(cvm_kkt_\w+)\s*\(", open(KKT_HEADER).read()))
    assert {"cvm_kkt_create", "cvm_kkt_factorize", "cvm_kkt_solve"} <= names
    path = os.path.join(ROOT, "curvami-official_amd", "libcurvami_kkt.so")
    if not os.path.exists(path):
        pytest.skip("library not built")
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l and l.split()[-1].startswith("cvm_")}
    assert exported == names, exported ^ names


def test_h5_library_exports_every_declared_symbol():
    """libcurvami_h5.so exports exactly what include/curvami_h5.h declares."""
    header = os.path.join(ROOT, "include", "curvami_h5.h")
    names = set(re.findall(r"CVM_H5_API\s+[\w\s\*]+?\b(cvm_h5_\w+)\s*\(", open(header).read()))
    assert {"cvm_h5_write_step", "cvm_h5_read_dataset", "cvm_h5_list"} <= names
    path = os.path.join(ROOT, "curvami-official_amd", "libcurvami_h5.so")
    if not os.path.exists(path):
        pytest.skip("library not built")
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l and l.split()[-1].startswith("cvm_")}
    assert exported == names, exported ^ names
