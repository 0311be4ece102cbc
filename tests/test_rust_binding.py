"""The Rust binding (rust/lbflow-hip-sys/src/lib.rs) against the C headers and
the Python ctypes mirror (lbflow_amd/abi.py), on the CPU: no rustc here, so
the Rust source is parsed. Every #[repr(C)] struct must list the header's
fields in the header's order with the corresponding types, every extern
function must have the header's name, argument count and argument types,
and every constant must have the header's value."""
import ctypes as C
import os
import re

import pytest

from lbflow_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUST = os.path.join(ROOT, "rust", "lbflow-hip-sys", "src", "lib.rs")
HEADERS = [os.path.join(ROOT, "include", h) for h in ("lbflow_hip.h", "lbflow_serve.h")]

C_SCALARS = {"int32_t": "i32", "uint32_t": "u32", "uint64_t": "u64", "double": "f64",
             "int": "c_int", "void": "()", "uint8_t": "u8", "char": "c_char", "size_t": "usize"}


def strip_c(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def c_type_to_rust(ctype, name):
    """(C declaration type, declarator) -> Rust type string."""
    ctype = " ".join(ctype.split())
    dims = re.findall(r"\[(\w+)\]", name)
    const = ctype.startswith("const ")
    base = ctype.replace("const ", "").replace("*", "").strip()
    ptrs = ctype.count("*") + name.count("*")
    rust = C_SCALARS.get(base, base)
    for _ in range(ptrs):
        rust = f"*{'const' if const else 'mut'} {rust}"
    for d in reversed(dims):
        rust = f"[{rust}; {d}]"
    return rust


def c_structs():
    out = {}
    for h in HEADERS:
        text = strip_c(open(h).read())
        for body, name in re.findall(r"typedef struct \w+ \{(.*?)\} (\w+);", text, flags=re.S):
            fields = []
            for decl in body.split(";"):
                decl = " ".join(decl.split())
                if not decl:
                    continue
                # `double a, b;` declares two fields of one type
                first, *more = [p.strip() for p in decl.split(",")]
                m = re.match(r"(.*?)(\*?\s*)(\w+)((?:\[\w+\])*)$", first)
                ctype, star, fname, dims = m.groups()
                fields.append((fname, c_type_to_rust(ctype + star, fname + dims)))
                for part in more:
                    m2 = re.match(r"(\*?)\s*(\w+)((?:\[\w+\])*)$", part)
                    fields.append((m2.group(2), c_type_to_rust(ctype + m2.group(1), m2.group(2) + m2.group(3))))
            out[name] = fields
    return out


def rust_structs():
    text = re.sub(r"//[^\n]*", "", open(RUST).read())
    out = {}
    for name, body in re.findall(r"#\[repr\(C\)\][^{;]*?pub struct (\w+) \{(.*?)\n\}", text,
                                 flags=re.S):
        fields = re.findall(r"pub (\w+): ([^,\n]+),", body)
        if fields:
            out[name] = [(f, " ".join(t.split())) for f, t in fields]
    return out


def test_every_header_struct_has_a_matching_rust_struct():
    cs, rs = c_structs(), rust_structs()
    assert set(cs) == set(rs), (set(cs) ^ set(rs))
    for name, fields in cs.items():
        assert rs[name] == fields, name


CTYPES_TO_RUST = {C.c_int32: "i32", C.c_uint32: "u32", C.c_uint64: "u64", C.c_double: "f64",
                  C.c_uint8: "u8"}


def ctypes_to_rust(t):
    if t in CTYPES_TO_RUST:
        return CTYPES_TO_RUST[t]
    if hasattr(t, "_length_"):
        inner = ctypes_to_rust(t._type_)
        length = {abi.MAX_Q: "LBF_MAX_Q"}.get(t._length_, str(t._length_))
        if inner == "f64" and t._length_ == abi.MAX_Q:
            return f"[f64; {length}]"
        return f"[{inner}; {t._length_}]"
    if hasattr(t, "_type_") and issubclass(t, C._Pointer):
        return f"*const {ctypes_to_rust(t._type_)}"
    return {abi.FaceBC: "lbf_face_bc", abi.Collision: "lbf_collision",
            abi.ScalarFaceBC: "lbf_scalar_face_bc", abi.ScalarDesc: "lbf_scalar_desc",
            abi.CaseDesc: "lbf_case_desc", abi.ServerConfig: "lbf_server_config",
            abi.RequestInfo: "lbf_request_info", abi.ServerStats: "lbf_server_stats"}[t]


@pytest.mark.parametrize("pyname,rname", [
    ("FaceBC", "lbf_face_bc"), ("Collision", "lbf_collision"),
    ("ScalarFaceBC", "lbf_scalar_face_bc"), ("ScalarDesc", "lbf_scalar_desc"),
    ("CaseDesc", "lbf_case_desc"), ("ServerConfig", "lbf_server_config"),
    ("RequestInfo", "lbf_request_info"), ("ServerStats", "lbf_server_stats")])
def test_rust_structs_equal_the_ctypes_mirror(pyname, rname):
    py = [(n, ctypes_to_rust(t)) for n, t in getattr(abi, pyname)._fields_]
    rs = rust_structs()[rname]
    assert [n for n, _ in py] == [n for n, _ in rs]
    for (n, tp), (_, tr) in zip(py, rs):
        tr = tr.replace("*mut", "*const")  # ctypes has one pointer kind
        assert tp == tr, (rname, n, tp, tr)


def c_functions():
    out = {}
    for h in HEADERS:
        text = strip_c(open(h).read())
        for ret, name, args in re.findall(r"\n\s*([\w\s\*]+?)\s*\b(lbf_\w+)\s*\(([^)]*)\)\s*;", text):
            params = [a.strip() for a in " ".join(args.split()).split(",") if a.strip() not in ("", "void")]
            types = []
            for p in params:
                m = re.match(r"(.*?)(\w+)$", p)
                types.append(c_type_to_rust(m.group(1), m.group(2)).replace("()", "c_void"))
            out[name] = (ret.strip(), types)
    return out


def rust_functions():
    text = re.sub(r"//[^\n]*", "", open(RUST).read())
    out = {}
    for name, args in re.findall(r"pub fn (lbf_\w+)\((.*?)\)", text, flags=re.S):
        types = [" ".join(a.split(":", 1)[1].split()) for a in args.split(",") if a.strip()]
        out[name] = types
    return out


def test_every_header_function_is_bound_with_its_argument_types():
    cf, rf = c_functions(), rust_functions()
    assert set(cf) == set(rf), set(cf) ^ set(rf)
    for name, (_, ctypes_) in cf.items():
        rtypes = rf[name]
        assert len(rtypes) == len(ctypes_), name
        for tc, tr in zip(ctypes_, rtypes):
            # opaque handles: C `lbf_case*` is `*mut lbf_case` in Rust
            assert tc == tr, (name, tc, tr)


def test_constants_equal_header_values():
    defines = {}
    for h in HEADERS:
        for name, value in re.findall(r"#define (LBF_\w+)\s+\(?(-?[0-9]+)u?\)?", strip_c(open(h).read())):
            defines[name] = int(value)
    consts = dict(re.findall(r"pub const (LBF_\w+): \w+ = (-?[0-9]+);", open(RUST).read()))
    assert set(consts) <= set(defines), set(consts) - set(defines)
    for name, value in consts.items():
        assert int(value) == defines[name], name
    # every enumerated value of the headers is bound
    assert set(defines) <= set(consts), set(defines) - set(consts)
