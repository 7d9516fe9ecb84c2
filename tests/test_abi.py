"""The C-ABI library: it loads, exports every function include/mgsolve.h
declares, and refuses to run without a GPU (there is no CPU fallback).
These run without a GPU; no compute call is made."""
import ctypes as C
import os
import subprocess

import pytest

import mgsolve as M
from conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "mgsolve.h")


def test_library_is_built():
    assert os.path.exists(M.LIB_PATH), "run __graft_entry__.build() first"


def test_every_declared_symbol_is_exported():
    names = M.exported_symbols_declared_in_header(HEADER)
    assert len(names) >= 29
    lib = C.CDLL(M.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and nm agrees they are dynamic text symbols
    out = subprocess.run(["nm", "-D", "--defined-only", M.LIB_PATH], capture_output=True, text=True).stdout
    dyn = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= dyn


def test_python_mirror_binds_every_symbol():
    L = M.load()
    for n in M.exported_symbols_declared_in_header(HEADER):
        fn = getattr(L, n)
        assert fn.restype is not None or n == "mgs_last_error", n


def test_header_compiles_as_c():
    src = f'#include "{HEADER}"\nint main(void){{ mgs_params p; (void)p; return 0; }}\n'
    r = subprocess.run(["gcc", "-x", "c", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-"], input=src,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_no_gpu_is_an_error_not_a_fallback():
    with pytest.raises(M.MgsError, match="no HIP device"):
        M.Solver(M.Config(N=16))


def test_invalid_parameters_are_rejected_before_any_device_call():
    L = M.load()
    h = C.c_void_p()
    for bad in (M.Config(N=15), M.Config(N=16, model=4), M.Config(N=16, fr_n=3),
                M.Config(N=16, precision=M.PREC_MIXED, fr0=1e-5),
                M.Config(N=18, precision=M.PREC_MIXED), M.Config(N=16, box=-1.0),
                M.Config(N=1280)):  # beyond the 32-bit level indexing (n <= 1024)
        p = bad.c()
        rc = L.mgs_create(C.byref(p), C.byref(h))
        assert rc == -1, bad
        assert L.mgs_last_error(None)
