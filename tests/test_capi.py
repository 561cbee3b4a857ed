"""The C ABI library loads and exports every symbol its headers declare
(no device calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "modfem_amd", "libmodfem_hip.so")


def declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(\w+)\s*\(", src, flags=re.M))


def test_library_exports_header_symbols():
    lib = ctypes.CDLL(LIB)
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            names |= declared(h)
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_error_path_without_device():
    import modfem_amd as mf
    L = mf.lib()
    assert L.mf_problem_create(12345) < 0
    assert L.mf_last_error()
    assert L.mf_mesh_box_prism(0, 1, 1, 1.0, 1.0, 1.0, 0, 0.0, 1) < 0
