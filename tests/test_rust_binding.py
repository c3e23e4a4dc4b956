"""The vxl-phys-hip crate (rust-wl_amd/vxl-phys-hip) against the C ABI it
binds. There is no Rust toolchain here, so the crate's raw declarations are
checked as text: every function its `extern "C"` block declares exists in
include/vxl_sph.h with the same parameter types and in libvxl_sph.so's
exports, and its #[repr(C)] structures list the header's fields in the
header's order with the same types. The call sequence the crate makes is the
one tests/harness/abi_harness.c makes (tests/test_gpu_harness.py)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CRATE = os.path.join(ROOT, "rust-wl_amd", "vxl-phys-hip")
HEADER = os.path.join(ROOT, "include", "vxl_sph.h")

SCALARS = {"uint32_t": "u32", "int32_t": "i32", "uint64_t": "u64", "float": "f32", "char": "c_char",
           "VxlSphStatus": "Status", "void": "c_void"}
STRUCTS = {"VxlSphHeader": "Header", "VxlSphFluidConfig": "FluidConfig", "VxlSphSpan": "Span",
           "VxlSphReaction": "Reaction", "VxlSphStepper": "Stepper"}


def strip_c_comments(s):
    return re.sub(r"/\*.*?\*/", "", s, flags=re.S)


def c_type_to_rust(decl):
    """A C parameter declaration ('const float *pos', 'float lo[3]') as the Rust type it binds to."""
    decl = decl.strip()
    array = decl.endswith("]")
    decl = re.sub(r"\[\d*\]$", "", decl)
    m = re.match(r"(const\s+)?(\w+)\s*(\**)\s*\w*$", decl)
    assert m, decl
    const, base, stars = m.group(1), m.group(2), m.group(3)
    rust = SCALARS.get(base) or STRUCTS.get(base)
    assert rust, f"unmapped C type {base}"
    for level in range(len(stars) + (1 if array else 0)):
        rust = ("*const " if const and level == 0 else "*mut ") + rust
    return rust


def header_functions():
    """name -> (return type, [C parameter declarations])."""
    text = strip_c_comments(open(HEADER).read())
    out = {}
    for m in re.finditer(r"(\w+)\s+(\**)\s*(vxl_sph_\w+)\s*\(([^)]*)\)\s*;", text, flags=re.S):
        ret, name, params = m.group(1) + m.group(2), m.group(3), " ".join(m.group(4).split())
        out[name] = (ret, [] if params in ("", "void") else [p.strip() for p in params.split(",")])
    return out


def rust_functions():
    text = open(os.path.join(CRATE, "src", "ffi.rs")).read()
    block = text[text.index('extern "C" {'):]
    out = {}
    for m in re.finditer(r"pub fn (vxl_sph_\w+)\s*\(([^)]*)\)\s*(?:->\s*([^;]+))?;", block, flags=re.S):
        params = " ".join(m.group(2).split())
        types = [] if not params else [p.split(":", 1)[1].strip() for p in params.split(",") if p.strip()]
        out[m.group(1)] = ((m.group(3) or "()").strip(), types)
    return out


def test_crate_files_present():
    for f in ("Cargo.toml", "build.rs", "src/lib.rs", "src/ffi.rs", "examples/hip_facade_probe.rs"):
        assert os.path.exists(os.path.join(CRATE, f)), f


def test_declarations_match_the_header():
    hdr = header_functions()
    rs = rust_functions()
    assert len(rs) >= 13
    for name, (ret, types) in rs.items():
        assert name in hdr, f"{name} is not in vxl_sph.h"
        hret, hparams = hdr[name]
        htypes = [c_type_to_rust(p) for p in hparams]
        assert types == htypes, f"{name}: crate {types} vs header {htypes}"
        want_ret = {"uint32_t": "u32", "int32_t": "i32", "VxlSphStatus": "Status", "char*": "*const c_char"}[hret]
        assert ret == want_ret, f"{name}: returns {ret}, header {hret}"


def test_declared_functions_are_exported():
    lib = os.path.join(ROOT, "rust-wl_amd", "libvxl_sph.so")
    if not os.path.exists(lib):
        pytest.fail("libvxl_sph.so is not built")
    syms = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True).stdout.split()
    missing = [n for n in rust_functions() if n not in syms]
    assert not missing, missing


def c_struct_fields(name):
    text = strip_c_comments(open(HEADER).read())
    m = re.search(r"typedef struct %s\s*\{(.*?)\}\s*%s;" % (name, name), text, flags=re.S)
    assert m, name
    fields = []
    for decl in m.group(1).split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        t, names = decl.split(" ", 1)
        for n in names.split(","):
            n = n.strip()
            arr = re.match(r"(\w+)\[(\d+)\]", n)
            rt = SCALARS.get(t) or STRUCTS[t]
            fields.append((arr.group(1), f"[{rt}; {arr.group(2)}]") if arr else (n, rt))
    return fields


def rust_struct_fields(name):
    text = open(os.path.join(CRATE, "src", "ffi.rs")).read()
    m = re.search(r"pub struct %s \{(.*?)\n\}" % name, text, flags=re.S)
    assert m, name
    return [(f.group(1), f.group(2).strip()) for f in re.finditer(r"pub (\w+): ([^,\n]+),", m.group(1))]


@pytest.mark.parametrize("c_name", ["VxlSphHeader", "VxlSphFluidConfig", "VxlSphSpan", "VxlSphReaction"])
def test_structures_match_the_header(c_name):
    assert rust_struct_fields(STRUCTS[c_name]) == c_struct_fields(c_name)
