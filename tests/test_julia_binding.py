"""The Julia shim's binding against the C declarations, checked on the host
(Julia is not installed here, so the shim cannot run):

* every ccall's argument-type tuple and return type in
  daggerfft.jl_amd/julia/DaggerFFTHIP.jl matches the prototype of the same
  function in include/daggerfft_hip.h, argument by argument (width and
  pointer depth; opaque handles as Ptr{Cvoid});
* every struct the shim mirrors (BatchDesc, ServerConfig, Request,
  RequestInfo) has the C struct's fields in the same order with the same
  sizes and offsets, from a probe compiled against the header that prints
  sizeof/offsetof.  Julia lays out isbits structs with C alignment rules, so
  the same offsets are what the ccall passes.

A planted mistake (a swapped field, a Cint where the C side takes int64_t)
is checked to fail the comparison."""
import json
import os
import re
import subprocess

import pytest

from conftest import ROOT

JL = os.path.join(ROOT, "daggerfft.jl_amd", "julia", "DaggerFFTHIP.jl")
HDR = os.path.join(ROOT, "include", "daggerfft_hip.h")

# Julia mirror -> C typedef
MIRRORS = {"BatchDesc": "dfft_batch_desc", "ServerConfig": "dfft_server_config", "Request": "dfft_request",
           "RequestInfo": "dfft_request_info"}

# size, alignment of the scalar Julia types the shim uses
JL_SCALAR = {"Cint": (4, 4), "UInt32": (4, 4), "Int32": (4, 4), "Int64": (8, 8), "UInt64": (8, 8),
             "Csize_t": (8, 8), "Cdouble": (8, 8), "Float64": (8, 8)}
C_SCALAR = {"int": "Cint", "uint32_t": "UInt32", "int64_t": "Int64", "uint64_t": "UInt64", "size_t": "Csize_t",
            "double": "Cdouble"}
OPAQUE = {"dfft_plan", "dfft_batch", "dfft_server", "void"}


def _strip_comments(s, line="#"):
    if line == "#":
        return re.sub(r"#[^\n]*", "", s)
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    return re.sub(r"//[^\n]*", "", s)


def _split_top(s):
    """Split on commas not inside {} or ()."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "{(":
            depth += 1
        elif ch in "})":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def julia_ccalls(text):
    """[(name, return type, [argument types])] of every ccall."""
    text = _strip_comments(text)
    out = []
    for m in re.finditer(r"ccall\(\(:(dfft_\w+),\s*libdfft\),\s*(\w+),\s*\(", text):
        i, depth = m.end(), 1
        while depth:
            depth += {"(": 1, ")": -1}.get(text[i], 0)
            i += 1
        out.append((m.group(1), m.group(2), _split_top(text[m.end():i - 1])))
    return out


def julia_structs(text):
    """{name: [(field, type)]} of the shim's struct definitions."""
    text = _strip_comments(text)
    out = {}
    for m in re.finditer(r"^(?:mutable\s+)?struct\s+(\w+)\s*\n(.*?)^end", text, flags=re.S | re.M):
        fields = re.findall(r"^\s*(\w+)::([\w{},\s]+?)\s*$", m.group(2), flags=re.M)
        out[m.group(1)] = fields
    return out


def jl_size_align(t):
    t = t.replace(" ", "")
    if t.startswith("Ptr{"):
        return 8, 8
    m = re.fullmatch(r"NTuple\{(\d+),(\w+)\}", t)
    if m:
        s, a = JL_SCALAR[m.group(2)]
        return int(m.group(1)) * s, a
    return JL_SCALAR[t]


def jl_layout(fields):
    """[(field, offset, size)] and total size under C alignment rules."""
    off, al, out = 0, 1, []
    for name, t in fields:
        s, a = jl_size_align(t)
        off = (off + a - 1) // a * a
        out.append((name, off, s))
        off += s
        al = max(al, a)
    return out, (off + al - 1) // al * al


def c_prototypes(text):
    """{name: (return type, [parameter declarations])} of the header."""
    text = _strip_comments(text, line="//")
    out = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(dfft_\w+)\s*\(([^)]*)\)\s*;", text):
        params = [p.strip() for p in m.group(3).split(",") if p.strip() and p.strip() != "void"]
        out[m.group(2)] = (" ".join(m.group(1).split()), params)
    return out


def c_param_as_julia(decl):
    """The Julia type a C parameter declaration corresponds to."""
    d = re.sub(r"\bconst\b", "", decl)
    depth = d.count("*")
    words = d.replace("*", " ").split()
    base = words[0]
    if depth == 0:
        return C_SCALAR[base]
    inner = "Cvoid" if base in OPAQUE else C_SCALAR.get(base) or {v: k for k, v in MIRRORS.items()}.get(base, base)
    if base == "char" and depth == 1:
        inner = "UInt8"
    t = inner
    for _ in range(depth):
        t = "Ptr{%s}" % t
    return t


def c_return_as_julia(ret):
    r = " ".join(ret.replace("const", "").split())
    return {"int": "Cint", "char *": "Cstring", "char*": "Cstring"}[r]


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    d = tmp_path_factory.mktemp("probe")
    src = d / "probe.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HDR, "int main(void) {",
             'printf("{");']
    jl = julia_structs(open(JL).read())
    first = True
    for jname, cname in MIRRORS.items():
        fields = [f for f, _ in jl[jname]]
        lines.append('printf("%s\\"%s\\": {\\"size\\": %%zu, \\"fields\\": [", sizeof(%s));'
                     % ("" if first else ",", cname, cname))
        for i, f in enumerate(fields):
            lines.append('printf("%s[\\"%s\\", %%zu, %%zu]", offsetof(%s, %s), sizeof(((%s *)0)->%s));'
                         % ("," if i else "", f, cname, f, cname, f))
        lines.append('printf("]}");')
        first = False
    lines += ['printf("}\\n");', "return 0;", "}"]
    src.write_text("\n".join(lines))
    exe = d / "probe"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-o", str(exe), str(src)], check=True)
    return json.loads(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout)


def test_ccall_signatures_match_header():
    calls = julia_ccalls(open(JL).read())
    protos = c_prototypes(open(HDR).read())
    assert len(calls) >= 15
    bad = []
    for name, ret, args in calls:
        assert name in protos, name
        cret, params = protos[name]
        want = [c_param_as_julia(p) for p in params]
        got = [a.replace(" ", "") for a in args]
        if got != want or c_return_as_julia(cret) != ret:
            bad.append((name, got, want, ret, cret))
    assert not bad, bad


def test_struct_mirrors_match_c_layout(c_layout):
    jl = julia_structs(open(JL).read())
    for jname, cname in MIRRORS.items():
        lay, size = jl_layout(jl[jname])
        c = c_layout[cname]
        assert size == c["size"], (jname, size, c["size"])
        assert [(f, o, s) for f, o, s in lay] == [tuple(x) for x in c["fields"]], (jname, lay, c["fields"])


def test_planted_mistakes_are_caught(c_layout):
    """A swapped pair of fields and a narrowed argument fail the checks."""
    jl = julia_structs(open(JL).read())
    f = list(jl["Request"])
    f[1], f[2] = f[2], f[1]  # kind <-> sign: same sizes, different names
    lay, _ = jl_layout(f)
    assert [(a, b, c) for a, b, c in lay] != [tuple(x) for x in c_layout["dfft_request"]["fields"]]
    f = list(jl["RequestInfo"])
    f[0] = (f[0][0], "Int64")  # state as Int64 shifts every later offset
    lay, _ = jl_layout(f)
    assert [(a, b, c) for a, b, c in lay] != [tuple(x) for x in c_layout["dfft_request_info"]["fields"]]
    protos = c_prototypes(open(HDR).read())
    want = [c_param_as_julia(p) for p in protos["dfft_server_wait"][1]]
    assert want == ["Ptr{Cvoid}", "Int64", "Cdouble"]
    assert ["Ptr{Cvoid}", "Cint", "Cdouble"] != want


def _flag_values(text, pattern):
    return {m.group(1): int(m.group(2), 16) for m in re.finditer(pattern, text)}


def test_flag_constants_match_header():
    """Every DFFT_* constant the shim defines as UInt32(0x...) has the value
    the header #defines for the same name (plan and batch flags)."""
    jl = _flag_values(open(JL).read(), r"const\s+(DFFT_\w+)\s*=\s*UInt32\(0x([0-9a-fA-F]+)\)")
    hdr = _flag_values(open(HDR).read(), r"#define\s+(DFFT_\w+)\s+0x([0-9a-fA-F]+)u?")
    assert "DFFT_FLAG_BATCH_PERSISTENT" in jl and "DFFT_FLAG_UNNORMALIZED" in hdr
    for name, v in jl.items():
        assert name in hdr, name
        assert hdr[name] == v, (name, hex(v), hex(hdr[name]))
