"""ctypes binding of the C ABI in include/mrmd_hip.h.

The argument types are derived from the header itself, so every declared entry point is
bound (and a missing export fails loudly at load time).  There is no CPU fallback: if the
gfx950 library is absent or cannot be loaded, importing this module raises.
"""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HEADER = os.path.join(ROOT, "include", "mrmd_hip.h")
LIB_PATH = os.path.join(HERE, "libmrmd_hip.so")


class Box(C.Structure):
    _fields_ = [("min", C.c_double * 3), ("max", C.c_double * 3), ("ghost", C.c_double * 3)]


class Pred1(C.Structure):
    _fields_ = [("kind", C.c_int32), ("axis", C.c_int32), ("a", C.c_double), ("b", C.c_double),
                ("c", C.c_double), ("tol", C.c_double)]


class Pred2(C.Structure):
    _fields_ = [("op", C.c_int32), ("pad", C.c_int32), ("first", Pred1), ("second", Pred1)]


class Slab(C.Structure):
    _fields_ = [("center", C.c_double * 3), ("atomistic_diameter", C.c_double),
                ("hybrid_diameter", C.c_double), ("nu", C.c_int64), ("interface_type", C.c_int32),
                ("pad", C.c_int32)]


class Replica(C.Structure):
    _fields_ = [("num_atoms", C.c_int64), ("box", C.c_double * 3), ("dt", C.c_double), ("gamma", C.c_double),
                ("temperature", C.c_double), ("mass", C.c_double), ("seed", C.c_uint64), ("adress", C.c_int32),
                ("tf_every", C.c_int32), ("atomistic_diameter", C.c_double), ("hybrid_diameter", C.c_double),
                ("nu", C.c_int64), ("tf_bin_width", C.c_double), ("tf_target_density", C.c_double),
                ("tf_sigma", C.c_double), ("tf_range", C.c_double), ("tf_modulation", C.c_double),
                ("tf_symmetric", C.c_int32), ("tf_periodic", C.c_int32)]


class Spherical(C.Structure):
    _fields_ = [("center", C.c_double * 3), ("atomistic_radius", C.c_double), ("hybrid_diameter", C.c_double),
                ("exponent", C.c_int64)]


class MdLoop(C.Structure):
    _fields_ = [("dt", C.c_double), ("neighbor_radius", C.c_double), ("skin", C.c_double), ("cell_ratio", C.c_double),
                ("grid_min", C.c_double * 3), ("grid_max", C.c_double * 3), ("max_neigh", C.c_int64), ("langevin", C.c_int32), ("reserved", C.c_int32), ("zeta", C.c_double),
                ("temperature", C.c_double), ("seed", C.c_uint64), ("call", C.c_uint64), ("max_disp", C.c_double),
                ("rebuilds", C.c_int64), ("steps", C.c_int64)]


class AdressLoop(C.Structure):
    _fields_ = [("dt", C.c_double), ("neighbor_radius", C.c_double), ("skin", C.c_double), ("zeta", C.c_double),
                ("temperature", C.c_double), ("seed", C.c_uint64), ("tf_every", C.c_int64), ("tf_sigma", C.c_double),
                ("tf_range", C.c_double), ("tf_pred", Pred1), ("call", C.c_uint64), ("max_disp", C.c_double),
                ("rebuilds", C.c_int64), ("steps", C.c_int64), ("energy", C.c_double)]


STRUCTS = {"mrmd_box": Box, "mrmd_pred1": Pred1, "mrmd_pred2": Pred2, "mrmd_slab": Slab, "mrmd_replica": Replica,
           "mrmd_spherical": Spherical, "mrmd_md_loop": MdLoop, "mrmd_adress_loop": AdressLoop}

_SCALARS = {
    "int": C.c_int, "int32_t": C.c_int32, "int64_t": C.c_int64, "uint64_t": C.c_uint64,
    "double": C.c_double,
}


def _ctype(decl):
    decl = decl.strip()
    decl = re.sub(r"\s+", " ", decl)
    # drop the parameter name
    m = re.match(r"^(const )?([A-Za-z_0-9]+)\s*(\*{0,2})\s*[A-Za-z_0-9]*$", decl)
    if not m:
        raise ValueError(f"cannot parse parameter '{decl}'")
    base, stars = m.group(2), m.group(3)
    if stars == "":
        return _SCALARS[base]
    if stars == "*" and base in STRUCTS:
        return C.POINTER(STRUCTS[base])
    return C.c_void_p


def parse_header(path=HEADER):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    decls = {}
    for m in re.finditer(r"\b(int|const char\s*\*)\s+(mrmd_\w+)\s*\(([^;{]*)\)\s*;", text):
        ret, name, params = m.group(1), m.group(2), m.group(3)
        params = [p for p in params.split(",") if p.strip() and p.strip() != "void"]
        decls[name] = (C.c_char_p if "char" in ret else C.c_int, [_ctype(p) for p in params])
    return decls


DECLS = parse_header()


class MrmdError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950)")
    lib = C.CDLL(LIB_PATH)
    missing = []
    for name, (res, args) in DECLS.items():
        try:
            f = getattr(lib, name)
        except AttributeError:
            missing.append(name)
            continue
        f.restype = res
        f.argtypes = args
    if missing:
        raise ImportError("libmrmd_hip.so lacks declared entry points: " + ", ".join(missing))
    return lib


lib = _load()
