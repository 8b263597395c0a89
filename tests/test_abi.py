"""CPU-side checks of the C ABI boundary: the library loads and exports every
entry point include/spicier_hip.h declares, the ctypes structures match the
header, and host-side validation (no GPU calls) behaves."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import circuit as CK

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spicier_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(A.LIB_PATH):
        pytest.fail("libspicier_hip.so not built (run __graft_entry__.build())")
    lib = C.CDLL(A.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_header():
    # compile a tiny C program against the header and compare sizeof/offsetof
    import subprocess
    import tempfile
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "spicier_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(spx_element), sizeof(spx_circuit),
         sizeof(spx_param_spec), sizeof(spx_sweep_spec), sizeof(spx_nr_config),
         sizeof(spx_sweep_result), offsetof(spx_nr_config, kcl_vntol));
  return 0;
}'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                              check=True).stdout.split()]
    want = [C.sizeof(A.SpxElement), C.sizeof(A.SpxCircuit), C.sizeof(A.SpxParamSpec),
            C.sizeof(A.SpxSweepSpec), C.sizeof(A.SpxNrConfig), C.sizeof(A.SpxSweepResult),
            A.SpxNrConfig.kcl_vntol.offset]
    assert got == want


def test_default_config_matches_reference_defaults():
    lib = A.load_library()
    c = A.SpxNrConfig()
    lib.spx_nr_config_default(C.byref(c))
    d = A.default_nr_config()
    for name, _ in A.SpxNrConfig._fields_:
        assert getattr(c, name) == getattr(d, name), name
    # newton.rs:24-34
    assert (c.v_abstol, c.v_reltol, c.i_abstol, c.max_iterations) == (1e-6, 1e-3, 1e-12, 50)


def test_circuit_size_and_host_validation():
    lib = A.load_library()
    c = CK.cs_amp()
    ckt, keep = c.to_c()
    assert lib.spx_circuit_size(C.byref(ckt)) == 5 == c.size
    assert CK.inverter_chain().size == 64
    # null handle -> invalid argument, without touching a device
    x = np.zeros(4)
    rc = lib.spx_solve_batch_f64(None, None, None, 2, 1, A.ptr(x), None, None)
    assert rc == A.SPX_ERR_INVALID_ARGUMENT
    assert b"null handle" in lib.spx_last_error()


def test_engine_raises_without_library(monkeypatch):
    # the product path fails loudly instead of falling back to the CPU
    monkeypatch.setattr(A, "LIB_PATH", "/nonexistent/libspicier_hip.so")
    monkeypatch.setattr(A, "_lib", None)
    with pytest.raises(RuntimeError, match="missing"):
        A.load_library()
