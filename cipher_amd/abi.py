"""ctypes mirror of include/cipher_amd.h and the loader of the HIP library.

The structures here match the C declarations field for field; the library is
``cipher_amd/libcipher_amd.so``, built from ``cipher_amd/csrc`` for gfx950.
Loading fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

MAXCP = 5
MAXTSER = 10
MAXAP = 15
TOL = 1.0e-6
R_GAS_CONST = 8.314462

CHEM_QUADRATIC, CHEM_CALPHADDIS, CHEM_CALPHAD2SL, CHEM_NONE = 0, 1, 2, 3
INTERP_LINEAR, INTERP_QUADRATIC, INTERP_CUBIC = 0, 1, 2
BC_NEUMANN, BC_DIRICHLET, BC_NONE = 0, 1, 2
SOURCE_SINK, SOURCE_RANDOM = 0, 1
FP64, FP32_MIXED = 0, 1

STATUS = {0: "CIPHER_OK", 1: "CIPHER_ERR_ARG", 2: "CIPHER_ERR_HIP",
          3: "CIPHER_ERR_CAPACITY", 4: "CIPHER_ERR_NONFINITE",
          5: "CIPHER_ERR_UNSUPPORTED", 6: "CIPHER_ERR_STATE"}
ERR_CAPACITY = 3
ERR_ARG = 1
ERR_UNSUPPORTED = 5


class TSeries(C.Structure):
    _fields_ = [("coeff", C.c_double * MAXTSER), ("logcoeff", C.c_double),
                ("exp", C.c_int32 * MAXTSER), ("n", C.c_int32), ("pad_", C.c_int32)]


class RK(C.Structure):
    _fields_ = [("n", C.c_int32), ("ts", C.c_int32), ("idx", C.c_int32), ("pad_", C.c_int32)]


class Mobility(C.Structure):
    _fields_ = [("m0", C.c_double), ("unary", C.c_int32), ("binary", C.c_int32)]


class Source(C.Structure):
    _fields_ = [("rate", C.c_double * MAXCP), ("ceq", C.c_double * MAXCP),
                ("amplitude", C.c_double * MAXCP), ("model", C.c_int32), ("pad_", C.c_int32)]


class Material(C.Structure):
    _fields_ = [("molarvolume", C.c_double), ("chempot_ex_kineticcoeff", C.c_double),
                ("c0", C.c_double * (2 * MAXCP)), ("stoichiometry", C.c_double * 2),
                ("temperature0", C.c_double),
                ("model", C.c_int32), ("nsites", C.c_int32), ("ref", C.c_int32),
                ("unary", C.c_int32), ("binary", C.c_int32), ("ternary", C.c_int32),
                ("binaryq", C.c_int32), ("ternaryq", C.c_int32),
                ("quad_binary", C.c_int32), ("quad_ceq", C.c_int32), ("mobility", C.c_int32),
                ("specific_heat", C.c_int32), ("latent_heat", C.c_int32),
                ("tconductivity", C.c_int32), ("nsources", C.c_int32), ("source", C.c_int32)]


class Interface(C.Structure):
    _fields_ = [("width", C.c_double), ("e0", C.c_double), ("m0", C.c_double),
                ("potential", C.c_double * MAXCP), ("mobilityc", C.c_double * MAXCP),
                ("e_act", C.c_int32), ("m_act", C.c_int32),
                ("e_naniso", C.c_int32), ("e_aniso", C.c_int32),
                ("m_naniso", C.c_int32), ("m_aniso", C.c_int32),
                ("has_potential", C.c_int32), ("has_mobilityc", C.c_int32),
                ("nsources", C.c_int32), ("source", C.c_int32)]


class BC(C.Structure):
    _fields_ = [("chem_val", C.c_double * MAXCP), ("thermal_val", C.c_double),
                ("boundary_id", C.c_int32), ("chem_type", C.c_int32),
                ("thermal_type", C.c_int32), ("chem_mask", C.c_int32)]


class Config(C.Structure):
    _fields_ = [("tseries", C.POINTER(TSeries)), ("rk", C.POINTER(RK)),
                ("rkindex", C.POINTER(C.c_int32)), ("mobility", C.POINTER(Mobility)),
                ("material", C.POINTER(Material)), ("interface", C.POINTER(Interface)),
                ("source", C.POINTER(Source)), ("bc", C.POINTER(BC)),
                ("aniso", C.POINTER(C.c_double)),
                ("phase_material", C.POINTER(C.c_int32)),
                ("interface_list", C.POINTER(C.c_int32)),
                ("junctionpenalty", C.c_double), ("temperature_rate", C.c_double),
                ("has_temperature_rate", C.c_int32),
                ("dim", C.c_int32), ("npf", C.c_int32), ("ncp", C.c_int32),
                ("nmat", C.c_int32), ("nf", C.c_int32), ("nbcs", C.c_int32),
                ("ntseries", C.c_int32), ("nrk", C.c_int32), ("nrkindex", C.c_int32),
                ("nmobility", C.c_int32), ("nsource", C.c_int32), ("naniso", C.c_int32),
                ("max_phase_neighbours", C.c_int32), ("interpolation", C.c_int32),
                ("gradient_calculation", C.c_int32)]


class Layout(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "maxsites", "sp_size", "sf_size", "as_size", "pf_size", "dp_size", "ex_size",
        "tm_size", "as_offset", "pf_offset", "dp_offset", "ex_offset", "tm_offset", "ndof")]


class Mesh(C.Structure):
    _fields_ = [("face_cells", C.POINTER(C.c_int32)), ("cellgeom", C.POINTER(C.c_double)),
                ("bface_cell", C.POINTER(C.c_int32)), ("bface_id", C.POINTER(C.c_int32)),
                ("ncells", C.c_int32), ("nfaces", C.c_int32), ("nbfaces", C.c_int32),
                ("pad_", C.c_int32), ("centroid", C.POINTER(C.c_double)), ("period", C.c_double * 3)]


class TsParams(C.Structure):
    _fields_ = [("atol", C.c_double), ("rtol", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
                ("safety", C.c_double), ("reject_safety", C.c_double), ("clip_lo", C.c_double),
                ("clip_hi", C.c_double), ("use_graph", C.c_int32), ("method", C.c_int32)]


class TsCtrl(C.Structure):
    _fields_ = [("err_prev", C.c_double), ("safety", C.c_double), ("rejected_last", C.c_int32),
                ("pad_", C.c_int32)]


class TsStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("rejected", C.c_int64), ("rhs_evals", C.c_int64), ("t", C.c_double),
                ("dt", C.c_double), ("err", C.c_double), ("host_syncs", C.c_int64)]


OUT_PHASEID, OUT_MATID, OUT_INTERFACEID, OUT_CHEMPOT, OUT_COMP, OUT_PHI, OUT_TEMPERATURE = range(1, 8)
MAX_OUTPUTS = 32


class ListShape(C.Structure):
    _fields_ = [("lanes", C.c_int32), ("cap", C.c_int32), ("map", C.c_int32), ("pad_", C.c_int32)]

    def tuple(self):
        return (self.lanes, self.cap, self.map)


# TSRK types of cipher_ts_params.method, by petscoptions' -ts_rk_type name
RK_3BS, RK_2A = 0, 1
RK_METHODS = {"3bs": RK_3BS, "2a": RK_2A}


def rk_method(name):
    """cipher_ts_params.method of a '-ts_rk_type' name; ERR_UNSUPPORTED for
    a type the device integrator does not run."""
    if name not in RK_METHODS:
        raise CipherError(ERR_UNSUPPORTED, f"petscoptions -ts_rk_type {name} (3bs and 2a are integrated)")
    return RK_METHODS[name]


def ts_params(atol, rtol, dt_min, dt_max, safety=0.9, reject_safety=0.5, clip=(0.1, 10.0), use_graph=False,
              method=RK_3BS):
    return TsParams(atol, rtol, dt_min, dt_max, safety, reject_safety, clip[0], clip[1], 1 if use_graph else 0,
                    method)


_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcipher_amd.so")
_lib = None


def _declare(lib):
    vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
    sig = {
        "cipher_layout_of": (i32, [C.POINTER(Config), C.POINTER(Layout)]),
        "cipher_ctx_create": (i32, [C.POINTER(Config), i32, C.POINTER(vp)]),
        "cipher_ctx_destroy": (i32, [vp]),
        "cipher_ctx_layout": (i32, [vp, C.POINTER(Layout)]),
        "cipher_set_precision": (i32, [vp, i32]),
        "cipher_set_stream": (i32, [vp, vp]),
        "cipher_set_temperature_rate": (i32, [vp, i32, dbl]),
        "cipher_upload_mesh": (i32, [vp, C.POINTER(Mesh)]),
        "cipher_rhs": (i32, [vp, dbl, vp, vp]),
        "cipher_rhs_active": (i32, [vp, dbl, vp, vp, vp]),
        "cipher_poststep": (i32, [vp, vp]),
        "cipher_setup_problem": (i32, [vp, vp, vp]),
        "cipher_wrms_norm": (i32, [vp, vp, vp, i64, dbl, dbl, C.POINTER(C.c_double)]),
        "cipher_wrms_norm_noalloc": (i32, [vp, vp, i64, dbl, dbl, vp, C.POINTER(C.c_double)]),
        "cipher_axpy_n": (i32, [vp, i64, vp, vp, i32, C.POINTER(C.c_double), C.POINTER(vp)]),
        "cipher_last_error": (C.c_char_p, []),
        "cipher_ts_adapt": (i32, [C.POINTER(TsParams), C.POINTER(TsCtrl), dbl, dbl, C.POINTER(C.c_int),
                                  C.POINTER(C.c_double)]),
        "cipher_ts_setup": (i32, [vp, C.POINTER(TsParams)]),
        "cipher_ts_attempt": (i32, [vp, vp, dbl, C.POINTER(C.c_double)]),
        "cipher_ts_accept": (i32, [vp, vp]),
        "cipher_ts_solve": (i32, [vp, vp, dbl, dbl, C.POINTER(C.c_double), i64, C.POINTER(TsStats)]),
        "cipher_ts_loadcases": (i32, [vp, vp, i32, C.POINTER(C.c_double), C.POINTER(C.c_double), dbl,
                                      C.POINTER(TsStats)]),
        "cipher_batch_create": (i32, [C.POINTER(vp), i32, C.POINTER(vp)]),
        "cipher_batch_destroy": (i32, [vp]),
        "cipher_batch_offsets": (i32, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "cipher_batch_set_precision": (i32, [vp, i32]),
        "cipher_batch_set_stream": (i32, [vp, vp]),
        "cipher_batch_set_temperature_rate": (i32, [vp, i32, i32, dbl]),
        "cipher_batch_rhs": (i32, [vp, dbl, vp, vp]),
        "cipher_batch_rhs_active": (i32, [vp, dbl, vp, vp, vp]),
        "cipher_batch_poststep": (i32, [vp, vp, C.POINTER(C.c_int32)]),
        "cipher_batch_wrms_norm": (i32, [vp, vp, vp, dbl, dbl, C.POINTER(C.c_double)]),
        "cipher_batch_autotune": (i32, [vp, vp, vp, i32, C.POINTER(C.c_double)]),
        "cipher_batch_shape": (i32, [vp, i32, i32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(ListShape)]),
        "cipher_get_shape": (i32, [vp, i32, C.POINTER(ListShape)]),
        "cipher_set_shape": (i32, [vp, i32, C.POINTER(ListShape)]),
        "cipher_autotune": (i32, [vp, vp, vp, i32, C.POINTER(C.c_double)]),
        "cipher_output": (i32, [vp, vp, i32, vp, vp]),
        "cipher_batch_ts_setup": (i32, [vp, C.POINTER(TsParams)]),
        "cipher_batch_ts_solve": (i32, [vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), i64, C.POINTER(TsStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


EXPORTED = ("cipher_layout_of", "cipher_ctx_create", "cipher_ctx_destroy", "cipher_ctx_layout",
            "cipher_set_precision", "cipher_set_stream", "cipher_set_temperature_rate",
            "cipher_upload_mesh", "cipher_rhs", "cipher_rhs_active", "cipher_poststep",
            "cipher_setup_problem", "cipher_wrms_norm", "cipher_wrms_norm_noalloc", "cipher_axpy_n", "cipher_last_error",
            "cipher_ts_adapt", "cipher_ts_setup", "cipher_ts_attempt", "cipher_ts_accept", "cipher_ts_solve",
            "cipher_ts_loadcases",
            "cipher_batch_create", "cipher_batch_destroy", "cipher_batch_offsets", "cipher_batch_set_precision",
            "cipher_batch_set_stream", "cipher_batch_set_temperature_rate", "cipher_batch_rhs",
            "cipher_batch_rhs_active", "cipher_batch_poststep",
            "cipher_batch_wrms_norm", "cipher_batch_autotune", "cipher_batch_shape", "cipher_batch_ts_setup",
            "cipher_batch_ts_solve", "cipher_get_shape", "cipher_set_shape", "cipher_autotune", "cipher_output")


def load_library(path=None):
    """Load libcipher_amd.so; raises ImportError if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"{p} is missing: the HIP library must be built for gfx950 "
            "(python -c 'import __graft_entry__ as g; g.build()'); there is no CPU fallback")
    lib = _declare(C.CDLL(p, mode=C.RTLD_GLOBAL))
    if path is None:
        _lib = lib
    return lib


class CipherError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        msg = ""
        try:
            m = load_library().cipher_last_error()
            msg = m.decode() if m else ""
        except Exception:  # pragma: no cover - error path of the error path
            pass
        super().__init__(f"{what}: {STATUS.get(code, code)} {msg}")


def check(code, what):
    if code != 0:
        raise CipherError(code, what)
