"""The Rust hip_api crate (impeller_amd/rust/hip_api) against the C header:
no Rust toolchain runs here, so this checks on the CPU that the crate's
extern declarations name exactly the header's impeller_rt_* functions with
the same parameter counts, pointer-ness and integer widths, and that the
library exports them."""
import ctypes
import os
import re

from impeller_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "impeller_hip.h")
CRATE = os.path.join(ROOT, "impeller_amd", "rust", "hip_api")

# C parameter type -> the Rust type the binding must use
C_TO_RUST = {"int": "c_int", "size_t": "usize", "uint64_t": "u64", "float": "f32"}


def c_decls():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*([\w\s\*]+?)\b(impeller_rt_\w+)\s*\(([^;]*?)\)\s*;", text, re.M):
        params = [p.strip() for p in m.group(3).replace("\n", " ").split(",") if p.strip()]
        out[m.group(2)] = (m.group(1).strip(), params)
    return out


def rust_decls():
    text = open(os.path.join(CRATE, "src", "bindings.rs")).read()
    body = text[text.index('extern "C" {'):]
    out = {}
    for m in re.finditer(r"pub fn (impeller_rt_\w+)\((.*?)\)\s*(?:->\s*([^;]+))?;", body, re.S):
        params = [p.strip() for p in m.group(2).split(",") if p.strip()]
        out[m.group(1)] = ((m.group(3) or "").strip(), params)
    return out


def c_kind(param):
    t = param.rsplit(None, 1)[0] if " " in param else param
    if "*" in param or t in ("hipStream_t", "hipEvent_t"):
        return "ptr"
    base = " ".join(w for w in t.split() if w != "const")
    return C_TO_RUST.get(base, base)


def rust_kind(param):
    t = param.split(":", 1)[1].strip()
    if t.startswith("*") or t in ("hipStream_t", "hipEvent_t"):
        return "ptr"
    return t


def test_crate_declares_the_header_runtime_layer():
    c, r = c_decls(), rust_decls()
    assert len(c) >= 30
    assert set(c) == set(r), sorted(set(c) ^ set(r))
    for name, (_, cparams) in c.items():
        rparams = r[name][1]
        assert len(cparams) == len(rparams), name
        for cp, rp in zip(cparams, rparams):
            assert c_kind(cp) == rust_kind(rp), (name, cp, rp)
        cret, rret = c[name][0], r[name][0]
        want = "ptr" if "*" in cret else {"size_t": "usize"}.get(cret, cret)
        got = "ptr" if rret.startswith("*") else rret
        assert want == got, (name, cret, rret)


def test_crate_modules_mirror_the_reference_crate():
    # cuda_api/src/lib.rs modules: bindings, device_info, error, file, mem, stream
    lib_rs = open(os.path.join(CRATE, "src", "lib.rs")).read()
    for m in ("bindings", "device_info", "error", "file", "mem", "stream"):
        assert f"pub mod {m};" in lib_rs
        assert os.path.exists(os.path.join(CRATE, "src", f"{m}.rs"))
    cargo = open(os.path.join(CRATE, "Cargo.toml")).read()
    assert 'name = "zkpoly_hip_api"' in cargo
    build = open(os.path.join(CRATE, "build.rs")).read()
    assert "rustc-link-lib=dylib=impeller_hip" in build


def test_library_exports_the_crate_symbols():
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in rust_decls():
        assert hasattr(L, name), name
