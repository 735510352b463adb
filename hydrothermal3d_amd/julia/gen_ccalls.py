"""Generate HT3D_ccalls.jl: one Julia method per entry point of include/ht3d.h,
each a single `ccall` with the entry's literal C signature.

Julia's `ccall` needs its argument-type tuple written out, so the bindings are
generated from the header (the single source of truth of the C ABI) rather than
built at run time.  Device-array parameters (every `double *` / `float *` except
the host pointers listed below) go through `devptr`, which accepts a device array
or a raw pointer and refuses host arrays; host pointers, handles and structs
pass through as given (`Ref`s, `Array`s, `Ptr`s).

    python hydrothermal3d_amd/julia/gen_ccalls.py          # rewrite the file
    python hydrothermal3d_amd/julia/gen_ccalls.py --check  # exit 1 if stale
"""
from __future__ import annotations

import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
HEADER = HERE.parent.parent / "include" / "ht3d.h"
OUT = HERE / "HT3D_ccalls.jl"

# double pointers of the header that are host memory
HOST_DOUBLE = {("ht3d_reduce", "out_host"), ("ht3d_lognormal_field", "K"), ("ht3d_lognormal_field", "phi"),
               ("ht3d_lognormal_field", "fault_n"), ("ht3d_lognormal_field", "fault_x0")}
SCALAR = {"int": "Cint", "double": "Cdouble", "float": "Cfloat", "size_t": "Csize_t"}
JULIA_RESERVED = {"end", "function", "type", "global", "local", "let", "begin", "do", "for", "while", "if", "else",
                  "module", "macro", "quote", "return", "try", "catch", "finally", "struct", "abstract", "const"}


def entries(header: Path = HEADER):
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    for ret, name, params in re.findall(r"^(int|const char \*)\s*(ht3d_\w+)\((.*?)\);", text, flags=re.S | re.M):
        ps = [" ".join(p.split()) for p in params.split(",")] if params.strip() not in ("", "void") else []
        yield ret, name, ps


def param(entry: str, decl: str, k: int):
    """(julia name, julia type, argument expression) of one C parameter."""
    if "*" in decl:
        base, pname = decl.rsplit("*", 1)
        stars = decl.count("*")
        base = base.replace("*", "").replace("const", "").strip()
    else:
        base, pname = decl.rsplit(" ", 1)
        stars = 0
    pname = pname.strip() or f"a{k}"
    jname = pname + "_" if pname in JULIA_RESERVED else pname
    if stars == 0:
        return jname, SCALAR[base], jname
    if stars == 1 and base in ("double", "float") and (entry, pname) not in HOST_DOUBLE:
        jt = "Cdouble" if base == "double" else "Cfloat"
        return jname, f"Ptr{{{jt}}}", f"devptr({jt}, {jname})"
    if stars == 1 and base == "double":
        return jname, "Ptr{Cdouble}", jname
    if stars == 1 and base == "char":
        return jname, ("Cstring" if "const" in decl else "Ptr{UInt8}"), jname
    if stars == 1 and base == "int":
        return jname, "Ptr{Cint}", jname
    return jname, "Ptr{Cvoid}", jname


def generate() -> str:
    out = ["# HT3D_ccalls.jl -- generated from include/ht3d.h by gen_ccalls.py; do not edit.",
           "# One method per C entry point, each a ccall with the entry's literal signature.",
           "# Device-array parameters go through devptr (device arrays or raw pointers only).", ""]
    for ret, name, ps in entries():
        cols = [param(name, p, k) for k, p in enumerate(ps)]
        names = [c[0] for c in cols]
        rt = "Cint" if ret == "int" else "Cstring"
        types = ", ".join(c[1] for c in cols)
        tup = f"({types},)" if len(cols) == 1 else f"({types})"
        args = ", ".join(c[2] for c in cols)
        call = f"ccall((:{name}, libht3d), {rt}, {tup}{', ' if cols else ''}{args})"
        pres = " ".join(n for n, t, _ in cols if t.startswith("Ptr"))
        body = f"GC.@preserve {pres} {call}" if pres else call
        out.append(f"{name}({', '.join(names)}) =\n    {body}")
    return "\n".join(out) + "\n"


def main(argv: list[str]) -> int:
    text = generate()
    if "--check" in argv:
        if not OUT.exists() or OUT.read_text() != text:
            print(f"{OUT} is stale: run {Path(__file__).name}", file=sys.stderr)
            return 1
        return 0
    OUT.write_text(text)
    print(f"wrote {OUT}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
