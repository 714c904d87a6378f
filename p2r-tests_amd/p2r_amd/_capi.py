"""ctypes binding of lib/libp2r_amd.so (C-ABI declared in include/p2r.h, include/p2r_gen.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libp2r_amd.so")

P2R_OK, P2R_ERR_ARG, P2R_ERR_SHAPE, P2R_ERR_HIP, P2R_ERR_UNSUPPORTED, P2R_ERR_NOMEM = range(6)
P2R_REFERENCE, P2R_CHAINED = 0, 1
P2R_EXACT, P2R_FAST = 0, 1
P2R_GRAD_F32, P2R_GRAD_MIXED, P2R_GRAD_F64 = 0, 1, 2
P2R_RAGGED_AUTO, P2R_RAGGED_LAYER_ORDER, P2R_RAGGED_PAIR_STREAM = 0, -1, -2
P2R_OPT_WRAP_PHI, P2R_OPT_CHARGE_FROM_IPT, P2R_OPT_CHI2_PER_LAYER = 1, 2, 4


class p2r_shape(C.Structure):
    _fields_ = [("nevts", C.c_int32), ("ntrks", C.c_int32), ("nlayer", C.c_int32), ("bsize", C.c_int32)]


class p2r_params(C.Structure):
    _fields_ = [("kfact", C.c_float), ("b_scale", C.c_float), ("precision", C.c_int32), ("options", C.c_int32)]


class p2r_gen_config(C.Structure):
    _fields_ = [("mode", C.c_int32), ("seed", C.c_uint64), ("smear", C.c_float), ("r0", C.c_float), ("dr", C.c_float),
                ("hit_var", C.c_float), ("layer_offsets", C.c_void_p), ("b_scale_truth", C.c_float)]


class p2r_tuning(C.Structure):
    _fields_ = [("fwd_threads", C.c_int32), ("fwd_layers_per_launch", C.c_int32), ("ragged_schedule", C.c_int32),
                ("ragged_threads", C.c_int32), ("checkpoint_stride", C.c_int32), ("fwd_waves", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class p2r_tune_entry(C.Structure):
    _fields_ = [("bsize", C.c_int32), ("threads", C.c_int32), ("layers_per_launch", C.c_int32),
                ("schedule", C.c_int32), ("checkpoint_stride", C.c_int32), ("waves", C.c_int32), ("ms", C.c_float)]


class p2r_tune_report(C.Structure):
    _fields_ = [("count", C.c_int32), ("best", C.c_int32), ("entry", p2r_tune_entry * 32)]

    def as_list(self):
        return [{k: getattr(self.entry[i], k) for k, _ in p2r_tune_entry._fields_} for i in range(self.count)]


_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"p2r HIP library not built: {LIB_PATH} (run __graft_entry__.build() or "
                               f"`make -C p2r-tests_amd`)")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_float
        sig = {
            "p2r_create": ([C.c_int, C.POINTER(vp)], C.c_int),
            "p2r_destroy": ([vp], C.c_int),
            "p2r_default_params": ([], p2r_params),
            "p2r_error_string": ([C.c_int], C.c_char_p),
            "p2r_last_error": ([vp], C.c_char_p),
            "p2r_forward": ([vp, vp, vp, vp, C.POINTER(p2r_shape), C.c_int, C.POINTER(p2r_params), vp, vp], C.c_int),
            "p2r_forward_ragged": ([vp, vp, vp, vp, vp, i32, vp, i32, i32, C.c_int, C.POINTER(p2r_params), vp, i32,
                                    vp], C.c_int),
            "p2r_forward_adjoint": ([vp, vp, vp, vp, C.POINTER(p2r_shape), C.POINTER(p2r_params), C.c_int, i32, vp,
                                     vp, vp, vp, vp, vp], C.c_int),
            "p2r_forward_adjoint_noalloc": ([vp, vp, vp, vp, C.POINTER(p2r_shape), C.POINTER(p2r_params), C.c_int,
                                             i32, vp, vp, vp, vp, vp, vp], C.c_int),
            "p2r_propagateToR": ([vp, vp, vp, vp, vp, vp, vp, i32, i64, i64, i64, f32, vp], C.c_int),
            "p2r_KalmanUpdate": ([vp, vp, vp, vp, vp, i32, i64, i64, i64, vp], C.c_int),
            "p2r_gen_reference": ([vp, vp, C.POINTER(p2r_shape)], C.c_int),
            "p2r_reference_hit": ([C.c_int, vp], C.c_int),
            "p2r_gen_default_config": ([], p2r_gen_config),
            "p2r_gen_uniform": ([vp, vp, C.POINTER(p2r_shape), C.POINTER(p2r_gen_config)], C.c_int),
            "p2r_gen_ragged_layout": ([vp, i32, i32, vp, vp], C.c_int),
            "p2r_gen_ragged": ([vp, vp, vp, vp, i32, vp, i32, i32, i32, i32, C.POINTER(p2r_gen_config)], C.c_int),
            "p2r_gen_lognormal_counts": ([vp, i32, C.c_double, C.c_double, i32, i32, C.c_uint64], C.c_int),
            "p2r_forward_streamed": ([vp, vp, vp, vp, vp, vp, vp, C.POINTER(p2r_shape), C.c_int,
                                      C.POINTER(p2r_params), i32], C.c_int),
            "p2r_get_tuning": ([vp, C.POINTER(p2r_tuning)], C.c_int),
            "p2r_set_tuning": ([vp, C.POINTER(p2r_tuning)], C.c_int),
            "p2r_autotune_forward": ([vp, vp, vp, vp, C.POINTER(p2r_shape), C.c_int, C.POINTER(p2r_params), i32,
                                      C.POINTER(p2r_tune_report), vp], C.c_int),
            "p2r_autotune_ragged": ([vp, vp, vp, vp, vp, i32, vp, i32, i32, C.c_int, C.POINTER(p2r_params), i32,
                                     C.POINTER(p2r_tune_report), vp], C.c_int),
            "p2r_autotune_adjoint": ([vp, vp, vp, vp, C.POINTER(p2r_shape), C.POINTER(p2r_params), C.c_int, vp, vp,
                                      vp, vp, vp, i32, C.POINTER(p2r_tune_report), vp], C.c_int),
            "p2r_repack": ([vp, vp, i32, vp, i32, i32, i32, i32, i32, vp], C.c_int),
            "p2r_graph_forward": ([vp, vp, vp, vp, C.POINTER(p2r_shape), C.c_int, C.POINTER(p2r_params), vp,
                                   C.POINTER(vp)], C.c_int),
            "p2r_graph_forward_ragged": ([vp, vp, vp, vp, vp, i32, vp, i32, i32, C.c_int, C.POINTER(p2r_params), vp,
                                          i32, C.POINTER(vp)], C.c_int),
            "p2r_graph_forward_adjoint": ([vp, vp, vp, vp, C.POINTER(p2r_shape), C.POINTER(p2r_params), C.c_int, i32,
                                           vp, vp, vp, vp, vp, C.POINTER(vp)], C.c_int),
            "p2r_graph_launch": ([vp, vp], C.c_int),
            "p2r_graph_destroy": ([vp], C.c_int),
            "p2r_autotune_width": ([vp, vp, vp, C.POINTER(p2r_shape), C.c_int, C.POINTER(p2r_params), vp, i32, i32,
                                    C.POINTER(p2r_tune_report), vp], C.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


class P2RError(RuntimeError):
    def __init__(self, status, detail=""):
        self.status = status
        msg = lib().p2r_error_string(status).decode()
        super().__init__(f"p2r status {status} ({msg}){': ' + detail if detail else ''}")
