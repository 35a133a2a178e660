"""The Rust binding (lxmf-vanity_amd/rust/hip.rs, the `Backend` trait of
src/gpu/mod.rs:40-52 over the C ABI) checked against the C header, since no
Rust compiler is available to build it: every `extern "C"` declaration names
a function of include/lxmf_vanity_hip.h with the same arity and argument and
return types, every #[repr(C)] struct has the field offsets and size of the
C struct it mirrors (measured by gcc on the header), and the binding uses
the reference's LaunchArgs fields (src/gpu/mod.rs:28-38) and trait methods
with their signatures."""

import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RS = os.path.join(ROOT, "lxmf-vanity_amd", "rust", "hip.rs")
HDR = os.path.join(ROOT, "include", "lxmf_vanity_hip.h")

# Rust struct name -> C struct it mirrors
STRUCTS = {"OpenOpts": "lxv_open_opts", "CLaunchArgs": "lxv_launch_args", "Hit": "lxv_hit"}
SCALARS = {"u8": ("uint8_t", 1), "u32": ("uint32_t", 4), "i32": ("int32_t", 4), "u64": ("uint64_t", 8),
           "usize": ("size_t", 8), "c_int": ("int", 4), "c_char": ("char", 1)}


def rust_source():
    with open(RS) as f:
        return f.read()


def rust_structs(src):
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\]\s*(?:#\[derive\([^)]*\)\]\s*)?struct\s+(\w+)\s*\{(.*?)\}", src, re.S):
        fields = []
        for name, ty in re.findall(r"(\w+)\s*:\s*(\[\s*\w+\s*;\s*\d+\s*\]|\w+)", m.group(2)):
            fields.append((name, ty.replace(" ", "")))
        out[m.group(1)] = fields
    return out


def rust_layout(fields):
    """Offsets and size of a #[repr(C)] struct of scalars and arrays."""
    off, align, offs = 0, 1, {}
    for name, ty in fields:
        arr = re.fullmatch(r"\[(\w+);(\d+)\]", ty)
        base, count = (arr.group(1), int(arr.group(2))) if arr else (ty, 1)
        size = SCALARS[base][1]
        off = (off + size - 1) // size * size
        offs[name] = off
        off += size * count
        align = max(align, size)
    return offs, (off + align - 1) // align * align


def c_layout(tmp_path, cstruct, names):
    prog = tmp_path / ("layout_%s.c" % cstruct)
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "lxmf_vanity_hip.h"', "int main(void) {",
             '  printf("sizeof %%zu\\n", sizeof(%s));' % cstruct]
    lines += ['  printf("%s %%zu\\n", offsetof(%s, %s));' % (n, cstruct, n) for n in names]
    lines += ["  return 0;", "}"]
    prog.write_text("\n".join(lines) + "\n")
    exe = tmp_path / ("layout_%s" % cstruct)
    subprocess.run(["gcc", "-I", os.path.dirname(HDR), str(prog), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    vals = dict(ln.split() for ln in out.splitlines())
    return {k: int(v) for k, v in vals.items()}


def rust_type_to_c(ty):
    ty = ty.strip()
    m = re.fullmatch(r"\*(const|mut)\s+(.*)", ty)
    if m:
        inner = rust_type_to_c(m.group(2))
        return ("const " if m.group(1) == "const" else "") + inner + "*"
    if ty == "Ctx":
        return "lxv_ctx"
    if ty in STRUCTS:
        return STRUCTS[ty]
    return SCALARS[ty][0]


def norm_c(ty):
    """A C parameter type (name stripped; arrays as pointers) in one spelling."""
    ty = re.sub(r"\s+", " ", ty).strip()
    ty = re.sub(r"\s*\*\s*", "*", ty)
    return ty


def c_functions():
    with open(HDR) as f:
        src = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
    out = {}
    for m in re.finditer(r"([\w\s\*]+?)\b(lxv_\w+)\s*\(([^)]*)\)\s*;", src):
        ret = norm_c(m.group(1))
        params = []
        for p in m.group(3).split(","):
            p = p.strip()
            if not p or p == "void":
                continue
            arr = re.fullmatch(r"(.*?)(\w+)\s*\[\s*\d*\s*\]", p)
            if arr:
                params.append(norm_c(arr.group(1)) + "*")
            else:
                params.append(norm_c(re.sub(r"\w+$", "", p)))
        out[m.group(2)] = (ret, params)
    return out


def rust_externs(src):
    block = re.search(r'extern "C"\s*\{(.*?)\n\}', src, re.S).group(1)
    out = {}
    for m in re.finditer(r"fn\s+(\w+)\s*\((.*?)\)\s*(?:->\s*([^;]+))?;", block, re.S):
        args = [a.split(":", 1)[1] for a in re.split(r",\s*", m.group(2).strip()) if a.strip()]
        ret = m.group(3).strip() if m.group(3) else None
        out[m.group(1)] = (norm_c(rust_type_to_c(ret)) if ret else "void", [norm_c(rust_type_to_c(a)) for a in args])
    return out


def test_every_extern_matches_the_header():
    externs = rust_externs(rust_source())
    cfun = c_functions()
    # the trait's calls (mod.rs:40-52) and the error/lifetime calls
    for name in ("lxv_open", "lxv_describe", "lxv_threads", "lxv_batch", "lxv_init", "lxv_launch", "lxv_poll",
                 "lxv_close", "lxv_last_error"):
        assert name in externs, name
    for name, (ret, args) in externs.items():
        assert name in cfun, "%s is not declared in the header" % name
        cret, cargs = cfun[name]
        assert ret == cret, (name, ret, cret)
        assert args == cargs, (name, args, cargs)


def test_parser_sees_a_changed_signature():
    # the check above is only as good as its parsing: a binding with one
    # argument type changed, or one dropped, must not match
    cfun = c_functions()
    src = rust_source()
    wrong = src.replace("n_words: usize", "n_words: u32")
    assert wrong != src
    assert rust_externs(wrong)["lxv_init"][1] != cfun["lxv_init"][1]
    wrong = src.replace("cap: u32, n_hits: *mut u32", "n_hits: *mut u32")
    assert wrong != src
    assert rust_externs(wrong)["lxv_poll"][1] != cfun["lxv_poll"][1]


@pytest.mark.parametrize("rname", sorted(STRUCTS))
def test_struct_layouts_match_the_header(rname, tmp_path):
    fields = rust_structs(rust_source())[rname]
    offs, size = rust_layout(fields)
    c = c_layout(tmp_path, STRUCTS[rname], [n for n, _ in fields])
    for name, _ in fields:
        assert offs[name] == c[name], (rname, name, offs[name], c[name])
    if rname == "OpenOpts":
        # a prefix of lxv_open_opts: the library accepts any size that covers
        # the fields up to `subwalks` and defaults the rest
        sub = c_layout(tmp_path, "lxv_open_opts", ["subwalks"])["subwalks"]
        assert sub <= size <= c["sizeof"], (size, sub, c["sizeof"])
        assert [n for n, _ in fields][0] == "size"
    else:
        assert size == c["sizeof"], (rname, size, c["sizeof"])


def test_binding_keeps_the_reference_trait():
    src = rust_source()
    # LaunchArgs fields, in the reference's order (mod.rs:28-38), all copied
    fields = ["iters", "ed", "name_hash", "want", "mask", "q", "max_hits", "dump_threads"]
    assert [n for n, _ in rust_structs(src)["CLaunchArgs"]] == fields
    for f in fields:
        assert re.search(r"\b%s:\s*a\.%s\b" % (f, f), src), f
    # the trait's methods with the reference's signatures (mod.rs:41-51)
    sigs = [r"fn describe\(&self\) -> String", r"fn threads\(&self\) -> u32", r"fn batch\(&self\) -> u64",
            r"fn init\(&mut self, u_prev_cur: &\[u32\]\) -> Result<\(\), String>",
            r"fn launch\(&mut self, a: &LaunchArgs\) -> Result<\(Vec<\(u32, u64\)>, Vec<u32>\), String>"]
    for s in sigs:
        assert re.search(s, src), s
    assert "impl Backend for Hip" in src and "impl Drop for Hip" in src
