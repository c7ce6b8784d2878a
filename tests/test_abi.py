"""The C-ABI library loads and exports every entry point include/*.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re
import subprocess

from fvmhd3d_amd import gpu as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(fvm_\w+)\s*\(", text, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    L = G.lib()
    names = declared_functions()
    assert len(names) >= 25, names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and nothing is declared that the binding does not know the signature of
    assert names - set(G.SIGNATURES) <= {"fvm_last_error", "fvm_sg_last_error"}


def test_exports_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", G.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (fvm_\w+)$", out, flags=re.M))
    assert declared_functions() <= exported
