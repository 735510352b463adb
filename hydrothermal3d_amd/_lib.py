"""ctypes binding of libht3d.so.

Signatures are read from include/ht3d.h (the single source of truth for the
C ABI), so the binding cannot drift from the header.  Loading fails loudly when
the library is missing: there is no CPU fallback anywhere in the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
HEADER = ROOT / "include" / "ht3d.h"
LIBPATH = Path(os.environ.get("HT3D_LIB", PKG / "libht3d.so"))

# status codes (include/ht3d.h)
OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_STATE, ERR_INTERNAL = 0, 1, 2, 3, 4, 5
NOT_CONVERGED, NAN, DIVERGED = 10, 11, 12
SUM, SUM_ABS, MAX, MIN, MAX_ABS = 0, 1, 2, 3, 4
FP64, FP32_DC = 0, 1
DT_V18, DT_V18_CONSTANT, DT_V7 = 0, 1, 2


class DyrelParams(C.Structure):
    _fields_ = [("nitmax", C.c_int), ("nitout", C.c_int), ("ncheck", C.c_int), ("norm", C.c_int),
                ("tol", C.c_double), ("CFL", C.c_double), ("cfact", C.c_double), ("lmin_div", C.c_double)]


class DyrelReport(C.Structure):
    _fields_ = [("iters", C.c_int), ("status", C.c_int), ("n_checks", C.c_int), ("n_lambda", C.c_int),
                ("lmax", C.c_double), ("lmin", C.c_double), ("dtau", C.c_double), ("h1", C.c_double),
                ("h2", C.c_double), ("nF_abs", C.c_double), ("nF_rel", C.c_double), ("nF_ini", C.c_double),
                ("ncheck", C.c_int)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class DyrelField(C.Structure):
    _fields_ = [("ncx", C.c_int), ("ncy", C.c_int), ("ncz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("X", C.c_void_p), ("R", C.c_void_p), ("kx", C.c_void_p), ("ky", C.c_void_p),
                ("kz", C.c_void_p), ("transient", C.c_int), ("rhoC", C.c_double), ("dt", C.c_double),
                ("bc_type", C.c_int * 6), ("bc_val", C.c_double * 6), ("kv", C.c_void_p)]


class V18Field(C.Structure):
    """ht3d_v18_field: the dimensional (v18) heat or Darcy DYREL solve."""
    _fields_ = [("ncx", C.c_int), ("ncy", C.c_int), ("ncz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
                ("transient", C.c_double), ("X", C.c_void_p), ("kx", C.c_void_p), ("ky", C.c_void_p),
                ("kz", C.c_void_p), ("phc_ex", C.c_void_p), ("wN", C.c_void_p), ("rhoC", C.c_void_p),
                ("PC", C.c_void_p), ("G", C.c_void_p), ("top", C.c_double), ("norm_scale", C.c_double),
                ("Tc_old", C.c_void_p), ("qyS", C.c_double), ("Qt", C.c_double), ("Tc_ex", C.c_void_p),
                ("phirhof0", C.c_void_p), ("rhofv", C.c_void_p), ("g", C.c_double), ("phi_i", C.c_double),
                ("scale_rho", C.c_double), ("scale_T", C.c_double), ("scale_sigma", C.c_double), ("kv", C.c_void_p)]


class StepReport(C.Structure):
    _fields_ = [("thermal", DyrelReport), ("darcy", DyrelReport), ("dt_used", C.c_double),
                ("dt_next", C.c_double), ("time", C.c_double), ("Nu", C.c_double), ("vmax", C.c_double),
                ("status", C.c_int), ("dc_outer_thermal", C.c_int), ("dc_outer_darcy", C.c_int),
                ("inner_thermal", C.c_int), ("inner_darcy", C.c_int), ("n_adv", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("ms_solve", C.c_double), ("bytes_solve", C.c_double), ("ms_step", C.c_double),
                ("steps", C.c_longlong), ("blocks", C.c_longlong), ("graph_launches", C.c_longlong),
                ("graphs", C.c_longlong), ("repacks", C.c_longlong), ("iterations", C.c_longlong),
                ("tile_rows", C.c_int * 4), ("tile_z", C.c_int * 4), ("tile_cpl", C.c_int * 4),
                ("grid_mode", C.c_int * 4), ("pair_rows", C.c_int * 4), ("pair_z", C.c_int * 4),
                ("pair_cpl", C.c_int * 4), ("vertex", C.c_int * 4)]

    def as_dict(self) -> dict:
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if not isinstance(v, (int, float)) else v
        return d


_CTYPE = {
    "int": C.c_int,
    "float": C.c_float,
    "double": C.c_double,
    "size_t": C.c_size_t,
}

_PROTO = re.compile(r"^(int|const char \*)\s*(ht3d_\w+)\((.*?)\);", re.S | re.M)


def _param_type(decl: str):
    decl = decl.strip()
    if decl == "void":
        return None
    if "*" in decl:
        base = decl.replace("const", "").split("*")[0].strip()
        if base == "char" and decl.count("*") == 1:
            return C.c_char_p
        return C.c_void_p  # device pointers, host pointers and struct pointers alike
    base = decl.rsplit(" ", 1)[0].strip()
    return _CTYPE[base]


def parse_header(path: Path = HEADER) -> dict[str, tuple[type, list]]:
    """Return {name: (restype, [argtypes])} for every entry point of the header."""
    text = path.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in _PROTO.finditer(text):
        ret, name, params = m.group(1), m.group(2), " ".join(m.group(3).split())
        restype = C.c_int if ret == "int" else C.c_char_p
        args = [t for t in (_param_type(p) for p in params.split(",")) if t is not None] if params.strip() else []
        out[name] = (restype, args)
    return out


class HT3DError(RuntimeError):
    def __init__(self, code: int, func: str, detail: str):
        super().__init__(f"{func} failed with status {code}: {detail}")
        self.code = code


_lib = None
_initialised = False


def lib() -> C.CDLL:
    """The loaded library (signatures applied).  Raises if libht3d.so is absent."""
    global _lib
    if _lib is None:
        if not LIBPATH.exists():
            raise RuntimeError(
                f"{LIBPATH} not found: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(str(LIBPATH))
        for name, (restype, args) in parse_header().items():
            fn = getattr(L, name)
            fn.restype = restype
            fn.argtypes = args
        _lib = L
    return _lib


def check(code: int, func: str) -> int:
    if code in (OK, NOT_CONVERGED):
        return code
    detail = lib().ht3d_last_error().decode(errors="replace")
    raise HT3DError(code, func, detail or lib().ht3d_strerror(code).decode())


def call(name: str, *args) -> int:
    ensure_init()
    return check(getattr(lib(), name)(*args), name)


def ensure_init(device: int | None = None) -> None:
    global _initialised
    if not _initialised:
        dev = int(os.environ.get("HT3D_DEVICE", "0")) if device is None else device
        check(lib().ht3d_init(dev), "ht3d_init")
        _initialised = True


def device_name() -> str:
    ensure_init()
    buf = C.create_string_buffer(256)
    check(lib().ht3d_device_name(buf, 256), "ht3d_device_name")
    return buf.value.decode()


def sync() -> None:
    call("ht3d_sync")
