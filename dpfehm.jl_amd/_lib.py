"""ctypes binding of libdpfgpu.so (include/dpfgpu.h).

The library is the only compute path: if it is missing or fails to load,
every call raises -- there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdpfgpu.so")

DPF_OK, DPF_EINVAL, DPF_EHIP, DPF_ENOCONV, DPF_EBREAKDOWN, DPF_ESINGULAR, DPF_ENOMEM = range(7)
DPF_FP64, DPF_MIXED, DPF_MIXED_PC = 0, 1, 2
DPF_HOST, DPF_DEVICE = 0, 1
DPF_JAC_H, DPF_JAC_KS, DPF_JAC_DIRICHLETHS, DPF_JAC_QS = 0, 1, 2, 3
DPF_FLAG_SWEEP_TIMING = 1
DPF_FLAG_GRAPH = 2
DPF_FLAG_KERNEL_TIMING = 4
DPF_FLAG_FP32_STORAGE = 8
DPF_FLAG_AUTOTUNE = 16

_ERRNAMES = {1: "EINVAL", 2: "EHIP", 3: "ENOCONV", 4: "EBREAKDOWN", 5: "ESINGULAR", 6: "ENOMEM"}


class DpfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libdpfgpu {_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class NotConverged(DpfError):
    pass


class SingularSystem(DpfError):
    """DPF_ESINGULAR: a free cell without edges, or (steady solves) a
    connected component with no Dirichlet cell, through the edges of nonzero
    conductance."""


class Opts(ctypes.Structure):
    _fields_ = [
        ("reltol", ctypes.c_double),
        ("abstol", ctypes.c_double),
        ("maxiter", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("inner_reltol", ctypes.c_double),
        ("max_outer", ctypes.c_int32),
        ("cheb_degree", ctypes.c_int32),
        ("memory", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("stream", ctypes.c_void_p),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("outer", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("levels", ctypes.c_int32),
        ("relres", ctypes.c_double),
        ("setup_ms", ctypes.c_double),
        ("solve_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("sweep_ms", ctypes.c_double),
        ("sweep_launches", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class KernelRecord(ctypes.Structure):
    _fields_ = [
        ("kernel", ctypes.c_char * 32),
        ("level", ctypes.c_int32),
        ("value_bytes", ctypes.c_int32),
        ("launches", ctypes.c_int64),
        ("ms", ctypes.c_double),
        ("bytes", ctypes.c_double),
    ]


_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32

# name: (restype, argtypes)
SIGNATURES = {
    "dpf_last_error": (ctypes.c_char_p, []),
    "dpf_opts_init": (None, [ctypes.POINTER(Opts)]),
    "dpf_topology_create": (ctypes.c_int, [_i64, _i64, _p, _p, _i64, _p, ctypes.POINTER(_p)]),
    "dpf_topology_destroy": (ctypes.c_int, [_p]),
    "dpf_batch_topology_create": (ctypes.c_int, [_i32, _p, _p, _p, _p, _p, _p, ctypes.POINTER(_p)]),
    "dpf_member_stats": (ctypes.c_int, [_p, _p, _i32]),
    "dpf_topology_info": (ctypes.c_int, [_p, _p, _i32]),
    "dpf_topology_tuning": (ctypes.c_int, [_p, _i32, _p, _p, _i32]),
    "dpf_topology_set_tuning": (ctypes.c_int, [_p, _i32, _i32, _i32, ctypes.c_double]),
    "dpf_topology_set_smoother": (ctypes.c_int, [_p, _i32, _i32, _i32, _i32, ctypes.c_double]),
    "dpf_topology_set_damping": (ctypes.c_int, [_p, ctypes.c_double, ctypes.c_double]),
    "dpf_topology_lanes": (ctypes.c_int, [_p, _p, _p, _i32]),
    "dpf_topology_set_lanes": (ctypes.c_int, [_p, _i32, _i32]),
    "dpf_twophase_tpfa": (ctypes.c_int, [_p, _p, _p, _p, _p, _p, _i32, _p, _p]),
    "dpf_twophase_upstream": (ctypes.c_int, [_p, _p, _p, _p, ctypes.c_double, _p, _p, _p, _p,
                                             _p]),
    "dpf_solvetwophase": (ctypes.c_int, [_p, _p, _p, _p, _p, _p, _p, _p, ctypes.c_double, _i32,
                                         _i32, _p, _p, _p, _p, _p, _p, _p, _p]),
    "dpf_solvetwophase_pullback": (ctypes.c_int, [_p, _p, _p, _p, _p, _p, _p]),
    "dpf_richards_residuals": (ctypes.c_int, [_p] * 12),
    "dpf_richards_steady": (ctypes.c_int, [_p] * 12),
    "dpf_richards_pullback": (ctypes.c_int, [_p] * 9),
    "dpf_transport_residuals": (ctypes.c_int, [_p] * 12 + [_i32, _p, _p]),
    "dpf_transport_steps": (ctypes.c_int, [_p] * 12 + [_i32, ctypes.c_double, ctypes.c_double,
                                                       _i32, _p, _p, _p, _p, _p]),
    "dpf_transport_pullback": (ctypes.c_int, [_p] * 9),
    "dpf_residuals": (ctypes.c_int, [_p, _p, _p, _p, _p, _p, _p, _p, ctypes.POINTER(Opts)]),
    "dpf_jacobian_pattern": (ctypes.c_int, [_p, _i32, _p, _p, _p]),
    "dpf_jacobian_values": (ctypes.c_int, [_p, _i32, _p, _p, _p, _p, _p, _p,
                                           ctypes.POINTER(Opts)]),
    "dpf_hm_forward": (ctypes.c_int, [_p, _p, _p, ctypes.POINTER(Opts)]),
    "dpf_hm_forward_noalloc": (ctypes.c_int, [_p, _p, _p, ctypes.POINTER(Opts)]),
    "dpf_hm_pullback": (ctypes.c_int, [_p, _p, _p, _p, _p, ctypes.POINTER(Opts)]),
    "dpf_geo_forward": (ctypes.c_int, [_p, _p, _p, ctypes.POINTER(Opts)]),
    "dpf_geo_pullback": (ctypes.c_int, [_p, _p, _p, _p, ctypes.POINTER(Opts)]),
    "dpf_steady_forward": (ctypes.c_int, [_p, _p, _p, _p, _p, ctypes.POINTER(Opts),
                                          ctypes.POINTER(Stats), ctypes.POINTER(_p)]),
    "dpf_steady_pullback": (ctypes.c_int, [_p, _p, _p, _p, _p, ctypes.POINTER(Opts),
                                           ctypes.POINTER(Stats)]),
    "dpf_state_destroy": (ctypes.c_int, [_p]),
    "dpf_transient_forward": (ctypes.c_int, [_p, _p, _p, _p, _p, _p, _p, _i32, ctypes.c_double,
                                             _p, _i32, _p, _p, ctypes.POINTER(Opts),
                                             ctypes.POINTER(Stats), ctypes.POINTER(_p)]),
    "dpf_transient_pullback": (ctypes.c_int, [_p, _p, _p, _p, ctypes.POINTER(Opts),
                                              ctypes.POINTER(Stats)]),
    "dpf_transient_forward_dts": (ctypes.c_int, [_p, _p, _p, _p, _p, _p, _p, _i32, _p, _p, _i32,
                                                 _p, _p, ctypes.POINTER(Opts),
                                                 ctypes.POINTER(Stats), ctypes.POINTER(_p)]),
    "dpf_transient_pullback_seeds": (ctypes.c_int, [_p, _i32, _p, _p, _p, _p, _p, _p,
                                                    ctypes.POINTER(Opts), ctypes.POINTER(Stats)]),
    "dpf_kernel_table": (ctypes.c_int, [_p, _p, _i32, _p, _i32]),
    "dpf_profile_kernel": (ctypes.c_int, [_p, _i32, _i32, _i32, _p, _p, ctypes.POINTER(Opts)]),
    "dpf_profile_grid_barrier": (ctypes.c_int, [_i32, _i32, _i32, _p, _p]),
    "dpf_inspect_last_error": (ctypes.c_char_p, []),
    "dpf_inspect_create": (ctypes.c_int, [_i64, _i64, _p, _p, _i64, _p, _i32, _i32,
                                          ctypes.POINTER(_p)]),
    "dpf_inspect_destroy": (None, [_p]),
    "dpf_inspect_levels": (_i32, [_p]),
    "dpf_inspect_size": (_i64, [_p, _i32, ctypes.c_char_p]),
    "dpf_inspect_get": (ctypes.c_int, [_p, _i32, ctypes.c_char_p, _p, _i64]),
    "dpf_inspect_check_steady": (ctypes.c_int, [_p, _p]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DpfError(-1, f"{LIB_PATH} is not built (run dpfehm.jl_amd/build.py)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != DPF_OK:
        msg = lib().dpf_last_error().decode(errors="replace")
        if code == DPF_ENOCONV:
            raise NotConverged(code, msg)
        if code == DPF_ESINGULAR:
            raise SingularSystem(code, msg)
        raise DpfError(code, msg)


def precision_code(v):
    """DPF_FP64 / DPF_MIXED / DPF_MIXED_PC from "fp64" / "mixed" / "mixed_pc"
    (an int passes through)."""
    if isinstance(v, str):
        return {"fp64": DPF_FP64, "mixed": DPF_MIXED, "mixed_pc": DPF_MIXED_PC}[v]
    return int(v)


def default_opts(**kw):
    o = Opts()
    lib().dpf_opts_init(ctypes.byref(o))
    for k, v in kw.items():
        if k == "precision":
            v = precision_code(v)
        setattr(o, k, v)
    return o
