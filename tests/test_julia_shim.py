"""The Julia one-include shim (hydrothermal3d_amd/julia/HT3D_HIP.jl), checked
without a Julia toolchain (runs without a GPU).

* HT3D_ccalls.jl, the literal-signature ccall of every C entry, is what
  gen_ccalls.py writes from include/ht3d.h today.
* Every kernel launch of the three drivers the shim replaces
  (Dimensionless/HydroThermal3D_Alban_v6.jl, Dimensionless/HydroThermal3D_v7.jl,
  Dimensional/HydroThermal3D_v18.jl and the files they include; the committed
  list tests/golden/driver_launches.json) finds an exported shim method of that
  name taking that many arguments.
* Every binding a shim kernel method calls is a C entry of the header, called
  with the entry's number of arguments (cell extents splatted as three).
"""
import json
import re
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
JL = ROOT / "hydrothermal3d_amd" / "julia"


def _call_args(text: str, open_paren: int) -> tuple[list[str], int]:
    """Top-level arguments of the call whose '(' is at text[open_paren], and the
    index of its ')'."""
    depth, args, cur = 0, [], []
    for i in range(open_paren, len(text)):
        ch = text[i]
        if ch in "([{":
            depth += 1
            if depth == 1:
                continue
        elif ch in ")]}":
            depth -= 1
            if depth == 0:
                a = "".join(cur).strip()
                if a:
                    args.append(a)
                return args, i
        elif ch == "," and depth == 1:
            args.append("".join(cur).strip())
            cur = []
            continue
        cur.append(ch)
    raise ValueError("unbalanced call")


def _shim_methods():
    """{name: [(nargs, body)]} of the one-line/short-form methods of HT3D_HIP.jl
    (a definition starts at column 0; its body runs to the next line that does)."""
    text = (JL / "HT3D_HIP.jl").read_text()
    text = re.sub(r"#[^\n]*", lambda m: " " * len(m.group(0)), text)
    out = {}
    for m in re.finditer(r"^(?:function )?([^\s(=#\"]+)\(", text, flags=re.M):
        args, close = _call_args(text, m.end() - 1)
        rest = text[close + 1:]
        if not (m.group(0).startswith("function ") or re.match(r"\s*=(?!=)", rest)):
            continue
        nxt = re.search(r"\n(?=\S)", rest)
        body = rest[: nxt.start()] if nxt else rest
        out.setdefault(m.group(1), []).append((len(args), body))
    return out


def _exports():
    text = (JL / "HT3D_HIP.jl").read_text()
    block = re.search(r"^export (.*?)\n\n", text, flags=re.S | re.M).group(1)
    block = re.sub(r"#[^\n]*", "", block)
    return {e.strip() for e in block.replace("\n", " ").split(",") if e.strip()}


def _header_arity():
    sys.path.insert(0, str(JL))
    import gen_ccalls
    return {name: len(ps) for _, name, ps in gen_ccalls.entries()}


def test_ccalls_file_matches_header():
    out = subprocess.run([sys.executable, str(JL / "gen_ccalls.py"), "--check"], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    text = (JL / "HT3D_ccalls.jl").read_text()
    for name, n in _header_arity().items():
        m = re.search(rf"^{name}\((.*?)\) =\n    (?:GC\.@preserve [^\n]*? )?ccall\(\(:{name}, libht3d\), (Cint|Cstring), "
                      r"\((.*?)\)", text, flags=re.M)
        assert m, name
        types = [t for t in m.group(3).split(",") if t.strip()]
        assert len(types) == n, (name, types)
        assert "..." not in m.group(3)        # a literal type tuple, never a splat


def test_every_driver_launch_has_a_shim_method():
    data = json.loads((ROOT / "tests" / "golden" / "driver_launches.json").read_text())
    methods, exports = _shim_methods(), _exports()
    missing = []
    for row in data["launches"]:
        k, n = row["kernel"], row["nargs"]
        arities = [a for a, _ in methods.get(k, [])]
        if k not in exports or n not in arities:
            missing.append(f"{k}/{n} ({row['at']}): methods {arities}, exported {k in exports}")
    assert not missing, "\n".join(missing)
    assert len(data["launches"]) > 150 and {r["at"].split(":")[0] for r in data["launches"]} >= set(data["drivers"])


def test_shim_methods_call_header_entries_with_their_arity():
    arity = _header_arity()
    checked = 0
    for name, defs in _shim_methods().items():
        if name.startswith("ht3d_"):
            continue
        for _, body in defs:
            for m in re.finditer(r"\b(ht3d_\w+)\(", body):
                entry = m.group(1)
                assert entry in arity, f"{name} calls {entry}, which include/ht3d.h does not declare"
                args, _ = _call_args(body, m.end() - 1)
                n = sum(3 if re.match(r"(cells_\w+|size)\(.*\)\.\.\.$", a) else 1 for a in args)
                assert n == arity[entry], f"{name}: {entry} called with {n} arguments, the header has {arity[entry]}"
                checked += 1
    assert checked > 80


def _c_struct_fields(name: str) -> list[tuple[str, str]]:
    """(member, kind) of `typedef struct { ... } name;` in include/ht3d.h, in
    declaration order; kind is "ptr" for pointers, else the C scalar type
    (comments stripped, one member per declarator)."""
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "ht3d.h").read_text(), flags=re.S)
    m = re.search(r"typedef struct \{([^}]*)\}\s*" + name + r";", text)
    assert m, name
    out = []
    for decl in m.group(1).split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        base = re.match(r"(const )?(int|double|float|size_t)", decl)
        assert base, decl
        for d in decl[base.end():].split(","):
            d = d.strip()
            ptr = d.startswith("*")
            nm = d.lstrip("*").strip()
            arr = re.match(r"(\w+)\[(\d+)\]", nm)
            if arr:
                out += [(arr.group(1), base.group(2))] * int(arr.group(2))
            else:
                out.append((nm, "ptr" if ptr else base.group(2)))
    return out


def _jl_struct_fields(name: str) -> list[tuple[str, str]]:
    text = (JL / "HT3D_HIP.jl").read_text()
    m = re.search(r"^struct " + name + r"\n(.*?)^end", text, flags=re.S | re.M)
    assert m, name
    kinds = {"Cint": "int", "Float64": "double", "Ptr{Cvoid}": "ptr"}
    out = []
    for line in m.group(1).splitlines():
        line = line.split("#")[0]
        for f in line.split(";"):
            f = f.strip()
            if f:
                nm, ty = f.split("::")
                out.append((nm.strip(), kinds[ty.strip()]))
    return out


def test_julia_structs_mirror_the_header():
    """The plain-bits structs the shim hands to the C entries by pointer have
    the header's members in the header's order with matching kinds, so no
    member the library reads lies past the end of the Julia value."""
    for jl, c in (("DyrelParams", "ht3d_dyrel_params"), ("DyrelReport", "ht3d_dyrel_report"),
                  ("V18Field", "ht3d_v18_field")):
        assert _jl_struct_fields(jl) == _c_struct_fields(c), jl
