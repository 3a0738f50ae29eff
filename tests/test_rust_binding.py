"""The Rust FFI crate (crates/chemsim-hip-sys/src/lib.rs) against the C
header it binds (include/chemsim.h), without a Rust toolchain: every function
the header declares is in the extern block with the same name, arity and
argument types (C type -> Rust type), nothing else is; the #[repr(C)] structs
have the C structs' fields in the same order with the same types; and every
enum constant has the header's value."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "chemsim.h")
LIB_RS = os.path.join(ROOT, "crates", "chemsim-hip-sys", "src", "lib.rs")

C_TO_RUST = {
    "int": "c_int", "int32_t": "i32", "int64_t": "i64", "uint64_t": "u64", "uint32_t": "u32", "uint16_t": "u16",
    "uint8_t": "u8",
    "size_t": "usize", "float": "f32", "double": "f64", "void": "c_void", "char": "c_char",
}


def strip_c_comments(text):
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def c_type_to_rust(ctype):
    """'const float*' -> '*const f32'; 'cs_lattice**' -> '*mut *mut cs_lattice'."""
    t = ctype.replace("*", " * ").split()
    const = "const" in t
    t = [x for x in t if x != "const"]
    stars = t.count("*")
    base = " ".join(x for x in t if x != "*")
    rust = C_TO_RUST.get(base, base)
    if stars == 0:
        return rust
    out = rust
    for k in range(stars):
        # the const qualifies the pointee (the innermost pointer)
        out = ("*const " if (const and k == 0) else "*mut ") + out
    return out


def header_functions():
    text = strip_c_comments(open(HEADER).read())
    text = re.sub(r"typedef\s+int\s*\(\*cs_frame_fn\)\s*\([^)]*\)\s*;", "", text)
    funcs = {}
    for m in re.finditer(r"^\s*((?:const\s+)?\w+\s*\**)\s*(cs_\w+)\s*\(([^)]*)\)\s*;", text, flags=re.M):
        ret, name, args = m.group(1).strip(), m.group(2), m.group(3).strip()
        params = []
        if args and args != "void":
            for a in args.split(","):
                a = a.strip()
                a = re.sub(r"(\w+)\s*\[\s*\d+\s*\]$", r"* \1", a)        # double out[3] -> double* out
                mm = re.match(r"^(.*?)(\w+)$", a)
                ctype = mm.group(1).strip()
                params.append(c_type_to_rust(ctype) if ctype != "cs_frame_fn" else "cs_frame_fn")
        funcs[name] = (c_type_to_rust(ret), params)
    return funcs


def rust_functions():
    text = open(LIB_RS).read()
    block = text[text.index('extern "C" {'):]
    block = re.sub(r"//[^\n]*", "", block)
    funcs = {}
    for m in re.finditer(r"pub fn (cs_\w+)\((.*?)\)\s*(?:->\s*([^;]+))?;", block, flags=re.S):
        name, args, ret = m.group(1), m.group(2), (m.group(3) or "()").strip()
        params = []
        for a in [x.strip() for x in args.split(",") if x.strip()]:
            params.append(re.sub(r"\s+", " ", a.split(":", 1)[1].strip()))
        funcs[name] = (ret, params)
    return funcs


def test_header_function_count():
    assert len(header_functions()) >= 60


def test_every_function_bound_with_same_signature():
    h, r = header_functions(), rust_functions()
    assert sorted(h) == sorted(r), (sorted(set(h) - set(r)), sorted(set(r) - set(h)))
    for name, (ret, params) in h.items():
        rret, rparams = r[name]
        assert rret == ret, (name, rret, ret)
        assert rparams == params, (name, rparams, params)


def c_struct_fields(name):
    text = strip_c_comments(open(HEADER).read())
    m = re.search(r"typedef struct " + name + r"\s*\{(.*?)\}\s*" + name + ";", text, flags=re.S)
    fields = []
    for line in m.group(1).split(";"):
        line = line.strip()
        if line:
            ctype, fname = line.rsplit(None, 1)
            fields.append((fname, C_TO_RUST[ctype]))
    return fields


def rust_struct_fields(name):
    text = open(LIB_RS).read()
    m = re.search(r"#\[repr\(C\)\][^\n]*\n(?:#\[[^\n]*\n)*pub struct " + name + r" \{(.*?)\n\}", text, flags=re.S)
    assert m, name
    return [(a.strip(), b.strip()) for a, b in re.findall(r"pub (\w+):\s*([\w:]+),", m.group(1))]


@pytest.mark.parametrize("name", ["cs_params", "cs_batch_tuning", "cs_record_stats"])
def test_struct_layout(name):
    assert rust_struct_fields(name) == c_struct_fields(name)


def header_constants():
    text = strip_c_comments(open(HEADER).read())
    consts = {}
    for body in re.findall(r"enum\s*\w*\s*\{(.*?)\}", text, flags=re.S):
        for item in body.split(","):
            if "=" in item:
                k, v = item.split("=")
                consts[k.strip()] = eval(v.strip())  # small integer expressions such as 3 << 20
    consts["CS_ABI_VERSION"] = int(re.search(r"#define CS_ABI_VERSION (\d+)", text).group(1))
    return consts


def test_constants_match():
    rs = dict((k, eval(v)) for k, v in re.findall(r"pub const (CS_\w+): \w+ = ([^;]+);", open(LIB_RS).read()))
    hc = header_constants()
    assert sorted(rs) == sorted(hc), (sorted(set(hc) - set(rs)), sorted(set(rs) - set(hc)))
    for k, v in hc.items():
        assert rs[k] == v, k
