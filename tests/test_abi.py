"""CPU checks of the C ABI: the HIP library loads and exports every entry point that
include/*.h declares, and the Python binding knows each one.  No GPU calls."""
import os
import re
import subprocess

import pytest

import dpd_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dpd_\w+)\s*\(", txt)))


def headers():
    return sorted(h for h in os.listdir(os.path.join(ROOT, "include")) if h.endswith(".h"))


def test_library_built():
    assert os.path.exists(dpd_amd.LIB_PATH), "run __graft_entry__.build()"


@pytest.mark.parametrize("header", headers())
def test_exports_every_declared_symbol(header):
    lib = dpd_amd.lib()
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", dpd_amd.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r" T (dpd_\w+)", out))
    assert set(names) <= exported


def test_bindings_cover_headers():
    from dpd_amd import host, mlp
    assert set(declared("dpd_engine.h")) == set(dpd_amd.SIGNATURES)
    assert set(declared("dpd_host.h")) == set(host.HOST_SIGNATURES)
    assert set(declared("dpd_mlp.h")) == set(mlp.MLP_SIGNATURES)


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(dpd_amd, "_lib", None)
    monkeypatch.setattr(dpd_amd, "LIB_PATH", "/nonexistent/libospreydpd_amd.so")
    with pytest.raises(dpd_amd.EngineMissing):
        dpd_amd.lib()
