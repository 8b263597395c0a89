"""The Rust adapter crate (rust/spicier-backend-hip) against the C ABI: no
Rust toolchain exists in this image, so its ffi.rs is checked as text
against include/spicier_hip.h — every function with the same argument count
and types, every #[repr(C)] struct with the header's fields in the same
order and types, every constant with the same value — and the wrappers in
lib.rs and the batched-sweep adapter call only declared functions and map
every status code."""
import os
import re

import pytest

import cheader as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = open(os.path.join(ROOT, "include", "spicier_hip.h")).read()
CRATE = os.path.join(ROOT, "rust", "spicier-backend-hip")
FFI = open(os.path.join(CRATE, "src", "ffi.rs")).read()
LIB = open(os.path.join(CRATE, "src", "lib.rs")).read()
ADAPTER = open(os.path.join(ROOT, "rust", "batched-sweep-integration", "src", "hip.rs")).read()

PRIM = {"int": "c_int", "int32_t": "i32", "int64_t": "i64", "uint32_t": "u32",
        "uint64_t": "u64", "uint8_t": "u8", "double": "f64", "float": "f32", "char": "c_char",
        "void": "c_void", "size_t": "usize"}


def camel(name):
    return "".join(w.capitalize() for w in name.split("_"))


def rust_type(c):
    """The Rust spelling of a C type: const T* -> *const T, T** -> *mut *mut T."""
    c = c.strip()
    const = c.startswith("const ")
    base = (c[6:] if const else c).replace("*", "").strip()
    stars = c.count("*")
    out = PRIM.get(base, camel(base))
    for i in range(stars):
        out = f"*{'const' if const and i == 0 else 'mut'} {out}"
    return out


def _norm(t):
    return " ".join(t.replace("*const", "*const ").replace("*mut", "*mut ").split())


def rust_functions():
    body = FFI[FFI.index('extern "C" {'):]
    out = {}
    for m in re.finditer(r"pub fn (spx_\w+)\((.*?)\)(?:\s*->\s*([^;]+))?;", body, flags=re.S):
        args = [a.strip() for a in m.group(2).split(",") if a.strip()]
        out[m.group(1)] = ([_norm(a.split(":", 1)[1]) for a in args],
                           _norm(m.group(3)) if m.group(3) else None)
    return out


def rust_structs():
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\]\s*(?:#\[derive\([^)]*\)\]\s*)?pub struct (\w+) \{(.*?)\}",
                         FFI, flags=re.S):
        fields = []
        for f in re.finditer(r"pub (\w+): ([^,\n]+),", m.group(2)):
            fields.append((f.group(1), _norm(f.group(2))))
        out[m.group(1)] = fields
    return out


def test_every_function_is_bound_with_its_signature():
    rf = rust_functions()
    cf = H.functions(HEADER)
    assert len(cf) >= 60
    assert sorted(rf) == sorted(n for _, n, _ in cf)
    for ret, name, params in cf:
        args, rret = rf[name]
        assert len(args) == len(params), name
        for (ct, cn), ra in zip(params, args):
            assert ra == _norm(rust_type(ct)), (name, cn, ct, ra)
        assert rret == (None if ret == "void" else _norm(rust_type(ret))), (name, ret, rret)


def test_every_struct_has_the_header_fields_in_order():
    rs = rust_structs()
    cs = H.structs(HEADER)
    for cname, fields in cs.items():
        rname = camel(cname)
        assert rname in rs, rname
        want = []
        for ty, fn, n in fields:
            t = _norm(rust_type(ty))
            want.append((fn, f"[{t}; {n}]"))
        got = [(f, t.replace("SPX_MAX_ELEMENT_PARAMS", str(H.defines(HEADER)["SPX_MAX_ELEMENT_PARAMS"])))
               for f, t in rs[rname]]
        want = [(f, t) for f, t in want]
        assert [f for f, _ in got] == [f for f, _ in want], rname
        for (f, t), (_, wt) in zip(got, want):
            if wt.startswith("[") and wt.endswith("; None]"):
                wt = wt[1:-7]
            assert t == wt, (rname, f, t, wt)
    for o in H.opaque(HEADER):
        assert re.search(rf"pub struct {camel(o)} \{{\s*_opaque: \[u8; 0\],", FFI), o


def test_every_constant_has_the_header_value():
    got = {m.group(1): int(m.group(2)) for m in
           re.finditer(r"pub const (\w+): (?:c_int|usize) = (-?\d+);", FFI)}
    want = dict(H.constants(HEADER))
    want.update({k: v for k, v in H.defines(HEADER).items() if k != "SPICIER_HIP_H"})
    assert got == want


def test_wrappers_call_declared_functions_and_map_every_status():
    declared = set(rust_functions())
    for src in (LIB, ADAPTER):
        used = set(re.findall(r"ffi::(spx_\w+)\(", src))
        assert used <= declared, used - declared
    statuses = [k for k in H.constants(HEADER) if k.startswith("SPX_ERR_")]
    for s in statuses:
        assert f"ffi::{s} =>" in LIB, s
    variants = re.findall(r"^\s{4}(\w+)(?:\(String\)|\s*\{[^}]*\})?,\s*$",
                          LIB[LIB.index("pub enum HipError"):LIB.index("impl fmt::Display")],
                          flags=re.M)
    assert len(variants) == len(statuses)
    for v in variants:
        assert f"HipError::{v}" in ADAPTER or v not in ("NoDevice", "DeviceInit", "MemoryAlloc",
                                                        "Transfer", "InvalidDimension",
                                                        "BatchTooLarge"), v
    assert "impl BatchedLuSolver for HipBatchedSolver" in ADAPTER
    for m in ("fn solve_batch(", "fn should_use_gpu(", "fn backend_type(", "fn config("):
        assert m in ADAPTER, m


def test_crate_files_exist():
    for f in ("Cargo.toml", "build.rs", "src/ffi.rs", "src/lib.rs"):
        assert os.path.exists(os.path.join(CRATE, f)), f
    assert "rustc-link-lib=dylib=spicier_hip" in open(os.path.join(CRATE, "build.rs")).read()
