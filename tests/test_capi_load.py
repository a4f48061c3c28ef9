"""The C-ABI library loads and exports every entry point the header declares.

No compute call is made here (no GPU needed).
"""
import ctypes
import os
import re
import subprocess

from smartcore_amd import _capi as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    names = S.declared_symbols()
    assert "sc_knn_search" in names and "sc_kmeans_run" in names
    lib = ctypes.CDLL(S.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_exported_symbols_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", S.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (sc_[a-z0-9_]+)\b", out))
    assert set(S.declared_symbols()) <= exported


def test_header_compiles_as_c():
    src = os.path.join(ROOT, "include", "smartcore_amd.h")
    r = subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-x", "c", src], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr


def test_last_error_without_gpu_call():
    # sc_last_error is host-only; a fresh thread has no error
    assert S.last_error() == ""
