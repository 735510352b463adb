"""Write driver_launches.json: every kernel launch (`@parallel [blocks threads]
Name(args)`) of the three HydroThermal3D drivers the one-include Julia shim
replaces, and of the files they include, as (kernel name, argument count,
file:line).  Reads the reference's Julia sources as text; run where the
reference is present:

    python tests/golden/make_driver_launches.py /root/reference
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

DRIVERS = ["Dimensionless/HydroThermal3D_Alban_v6.jl", "Dimensionless/HydroThermal3D_v7.jl",
           "Dimensional/HydroThermal3D_v18.jl"]
OUT = Path(__file__).resolve().parent / "driver_launches.json"
LAUNCH = re.compile(r"@parallel\s+(?:[A-Za-z_]\w*\s+[A-Za-z_]\w*\s+)?([^\s(]+)\(")


def top_level_args(text: str, start: int) -> tuple[int, int]:
    """(argument count, end index) of the call whose '(' is at text[start]."""
    depth, n, i, empty = 0, 1, start, True
    while True:
        ch = text[i]
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
            if depth == 0:
                return (0 if empty else n), i
        elif ch == "," and depth == 1:
            n += 1
        elif depth >= 1 and not ch.isspace():
            empty = False
        i += 1


def launches(path: Path, rel: str):
    text = path.read_text()
    code = re.sub(r"#[^\n]*", lambda m: " " * len(m.group(0)), text)    # blank out comments, keep offsets
    out = []
    for m in LAUNCH.finditer(code):
        name = m.group(1)
        if name in ("function",):
            continue
        n, _ = top_level_args(code, m.end() - 1)
        out.append({"kernel": name, "nargs": n, "at": f"{rel}:{code.count(chr(10), 0, m.start()) + 1}"})
    return out


def main(ref: Path) -> int:
    rows, seen = [], set()
    for d in DRIVERS:
        files = [d]
        text = (ref / d).read_text()
        for inc in re.findall(r'^\s*include\("([^"]+)"\)', text, flags=re.M):
            files.append(str((Path(d).parent / inc).as_posix()))
        for f in files:
            p = (ref / f).resolve()
            rel = str(p.relative_to(ref.resolve()))
            if rel in seen or not p.exists():
                continue
            seen.add(rel)
            rows.extend(launches(p, rel))
    OUT.write_text(json.dumps({"drivers": DRIVERS, "launches": rows}, indent=1, ensure_ascii=False) + "\n")
    print(f"wrote {OUT}: {len(rows)} launches, {len({r['kernel'] for r in rows})} kernels")
    return 0


if __name__ == "__main__":
    sys.exit(main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")))
