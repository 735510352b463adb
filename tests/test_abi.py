"""The C ABI: libht3d.so loads and exports every entry point include/ht3d.h
declares (no GPU call is made), and the product path fails loudly without it."""
import ctypes as C
import os
import re
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "ht3d.h").read_text(), flags=re.S)
    return re.findall(r"^(?:int|const char \*)\s*(ht3d_\w+)\(", text, flags=re.M)


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) == len(set(names)) and len(names) > 60
    for must in ("ht3d_init", "ht3d_ComputeFlux", "ht3d_DYRELUpdate", "ht3d_GershgorinPoisson", "ht3d_dyrel_solve",
                 "ht3d_weno5_advect", "ht3d_hrl_step", "ht3d_reduce"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from hydrothermal3d_amd import _lib
    if not _lib.LIBPATH.exists():
        pytest.fail(f"{_lib.LIBPATH} is not built (run __graft_entry__.build())")
    L = C.CDLL(str(_lib.LIBPATH))
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"
    sigs = _lib.parse_header()
    assert set(sigs) == set(_declared())


def test_uninitialised_calls_return_state_error():
    from hydrothermal3d_amd import _lib
    if _lib._initialised:
        pytest.skip("library already initialised in this process")
    L = _lib.lib()
    # no ht3d_init in this process: entry points refuse instead of touching a device
    assert L.ht3d_ComputeFlux(*([None] * 7), 1.0, 1.0, 1.0, 4, 4, 4) == _lib.ERR_STATE
    assert L.ht3d_strerror(_lib.NAN).decode().startswith("NaN")


def test_product_fails_loudly_without_library(tmp_path):
    code = ("import hydrothermal3d_amd as ht\n"
            "try:\n    ht.DeviceArray((2, 2, 2))\nexcept RuntimeError as e:\n    print('RAISED', 'not found' in str(e))\n")
    env = dict(os.environ, HT3D_LIB=str(tmp_path / "missing.so"), PYTHONPATH=str(ROOT))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, cwd=ROOT)
    assert "RAISED True" in out.stdout, out.stderr


def test_product_does_not_import_oracle():
    code = "import sys, hydrothermal3d_amd, hydrothermal3d_amd.solver, hydrothermal3d_amd.inputs; " \
           "print(any(m == 'oracle' or m.startswith('oracle.') for m in sys.modules))"
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT,
                         env=dict(os.environ, PYTHONPATH=str(ROOT)))
    assert out.stdout.strip() == "False", out.stderr


def test_check_interval_rule():
    """The check-interval tuner's cost model (host function, no GPU): L* =
    sqrt(2 I t_chk / t_it) snapped to a divisor of nitout, 50 without history."""
    import math
    from hydrothermal3d_amd import _lib
    L = _lib.lib().ht3d_tune_check_interval
    assert L(-1.0, 100.0, 50.0, 100) == 50
    assert L(2000.0, -1.0, 50.0, 100) == 50
    assert L(2000.0, 100.0, 70.0, 100) == 50          # L* = 52.9
    assert L(400.0, 100.0, 70.0, 100) == 25           # L* = 23.7
    assert L(20000.0, 100.0, 70.0, 1000) == 200       # L* = 167 -> 200 (log-nearest of 100, 200, 250)
    assert L(100.0, 1000.0, 1.0, 100) == 10           # cheap checks: the shortest interval
    for it, ti, tc, nout in ((900.0, 98.0, 60.0, 100), (3500.0, 700.0, 300.0, 500)):
        got = L(it, ti, tc, nout)
        assert nout % got == 0
        best = math.sqrt(2 * it * tc / ti)
        divs = [c for c in (10, 20, 25, 50, 100, 200, 250, 500, 1000) if c <= nout and nout % c == 0]
        assert got == min(divs, key=lambda c: abs(math.log(c / best)))


def test_exceptions_stop_at_the_abi():
    """Every extern "C" entry runs its body through a guard that turns a C++
    exception into a status code.  With the failure hook armed, entries return
    HT3D_ERR_NOMEM (before any device access) and name themselves in the error
    text; once disarmed they behave normally.  No GPU call is made."""
    from hydrothermal3d_amd import _lib
    L = _lib.lib()
    try:
        L.ht3d_debug_inject_host_failure(3)
        assert L.ht3d_tune_check_interval(2000.0, 100.0, 70.0, 100) == _lib.ERR_NOMEM
        assert b"ht3d_tune_check_interval" in L.ht3d_last_error()
        assert L.ht3d_batch_tune_force(0) == _lib.ERR_NOMEM
        assert L.ht3d_ComputeFlux(*([None] * 7), 1.0, 1.0, 1.0, 4, 4, 4) == _lib.ERR_NOMEM
        assert b"host memory allocation failed" in L.ht3d_last_error()
        assert L.ht3d_tune_check_interval(2000.0, 100.0, 70.0, 100) == 50     # the hook is spent
    finally:
        L.ht3d_debug_inject_host_failure(0)
    assert L.ht3d_strerror(_lib.ERR_INTERNAL).decode() == "internal error"


def test_every_entry_body_is_guarded():
    """Source check: each extern "C" entry of csrc/ with a multi-line body opens
    with HT3D_API_BEGIN and closes with HT3D_API_END; one-line entries only
    return a constant or a field."""
    unguarded = []
    for src in sorted((ROOT / "hydrothermal3d_amd" / "csrc").glob("*.hip")):
        lines = src.read_text().split("\n")
        for n, ln in enumerate(lines):
            m = re.match(r'^(?:extern "C" )?int (ht3d_\w+)\(', ln)
            if not m or m.group(1) == "ht3d_debug_inject_host_failure":
                continue
            j = n
            while not lines[j].rstrip().endswith(("{", ";", "}")):
                j += 1
            if lines[j].rstrip().endswith(";"):
                continue
            if lines[j].rstrip().endswith("}"):
                assert re.search(r"\{ return [\w:]+; \}$", lines[j]), f"{src.name}: {m.group(1)}"
                continue
            k = j + 1
            while lines[k] != "}":
                k += 1
            if lines[j + 1].strip() != "HT3D_API_BEGIN" or lines[k - 1].strip() != "HT3D_API_END":
                unguarded.append(f"{src.name}: {m.group(1)}")
    assert not unguarded, unguarded


def _header_struct(name: str) -> list[tuple[str, str]]:
    """(member, kind) of `typedef struct { ... } name;` in include/ht3d.h in
    declaration order: kind "ptr" for pointers, the C type otherwise (arrays
    expanded element by element)."""
    import re
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "ht3d.h").read_text(), flags=re.S)
    m = re.search(r"typedef struct \{([^}]*)\}\s*" + name + r";", text)
    assert m, name
    out = []
    for decl in m.group(1).split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        base = re.match(r"(const )?(long long|int|double|float|size_t|ht3d_\w+)", decl)
        assert base, decl
        for d in decl[base.end():].split(","):
            d = d.strip()
            nm = d.lstrip("*").strip()
            arr = re.match(r"(\w+)\[(\d+)\]", nm)
            if arr:
                out += [(arr.group(1), base.group(2))] * int(arr.group(2))
            else:
                out.append((nm, "ptr" if d.startswith("*") else base.group(2)))
    return out


def _ctypes_struct(cls) -> list[tuple[str, str]]:
    import ctypes as C
    kinds = {C.c_int: "int", C.c_double: "double", C.c_float: "float", C.c_void_p: "ptr", C.c_longlong: "long long",
             C.c_size_t: "size_t"}
    names = {v: k for k, v in STRUCTS.items()}
    out = []
    for nm, ty in cls._fields_:
        if hasattr(ty, "_length_"):
            out += [(nm, kinds[ty._type_])] * ty._length_
        elif ty in names:
            out.append((nm, names[ty]))
        else:
            out.append((nm, kinds[ty]))
    return out


STRUCTS = {}


def test_ctypes_structs_mirror_the_header():
    """The Python layer's ctypes structures have the header's members in the
    header's order with matching types, so every member the library reads or
    writes is where the caller put it (a member appended to the C struct but
    not to its mirror would be read past the caller's struct)."""
    from hydrothermal3d_amd import _lib
    STRUCTS.update({"ht3d_dyrel_params": _lib.DyrelParams, "ht3d_dyrel_report": _lib.DyrelReport,
                    "ht3d_dyrel_field": _lib.DyrelField, "ht3d_v18_field": _lib.V18Field,
                    "ht3d_step_report": _lib.StepReport, "ht3d_batch_stats": _lib.BatchStats})
    for cname, cls in STRUCTS.items():
        assert _ctypes_struct(cls) == _header_struct(cname), cname
