"""ctypes binding of the C ABI in include/szsnn_hip.h (libszsnn_hip.so).

The shared library is required: importing this module raises ImportError when
it has not been built (python -m szsnn_amd.build).  There is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libszsnn_hip.so")

SZSNN_OK = 0
SZSNN_ERR_INVALID = -1
SZSNN_ERR_HIP = -2
SZSNN_ERR_OVERFLOW = -3
SZSNN_ERR_NOMEM = -4
SZSNN_ERR_STATE = -5

MODEL_IF_CURR_EXP = 0
MODEL_IZHIKEVICH = 1
MODEL_IF_COND_EXP = 2

SOURCE_NONE = 0
SOURCE_CONST_RATE = 1
SOURCE_POISSON = 2

MODE_GRAPH = 0
MODE_PERSISTENT = 1

KERNEL_GATHER = 0
KERNEL_UPDATE = 1

c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_u16p = ctypes.POINTER(ctypes.c_uint16)
c_f32p = ctypes.POINTER(ctypes.c_float)


class NetDesc(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int32),
        ("n_neurons", ctypes.c_uint32),
        ("n_sources", ctypes.c_uint32),
        ("source_model", ctypes.c_int32),
        ("params", c_f32p),
        ("vars", c_f32p),
        ("src_rate", c_f32p),
        ("src_last", c_f32p),
        ("src_thresh", c_u64p),
        ("n_syn", ctypes.c_uint64),
        ("syn_src", c_u32p),
        ("syn_tgt", c_u32p),
        ("syn_w", c_f32p),
        ("syn_delay", c_u16p),
        ("max_delay", ctypes.c_uint32),
        ("frac_bits", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("noise_amp", c_f32p),
        ("drive_k", ctypes.c_uint32),
        ("drive_cdf", c_u64p),
        ("drive_w", ctypes.c_float),
        ("queue_capacity", ctypes.c_uint32),
        ("pull_layout", ctypes.c_int32),
    ]


class NetInfo(ctypes.Structure):
    _fields_ = [
        ("n_neurons", ctypes.c_uint32),
        ("n_sources", ctypes.c_uint32),
        ("max_delay", ctypes.c_uint32),
        ("n_params", ctypes.c_uint32),
        ("n_vars", ctypes.c_uint32),
        ("frac_bits", ctypes.c_int32),
        ("n_syn_stored", ctypes.c_uint64),
        ("n_rows", ctypes.c_uint64),
        ("pull_groups", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64),
        ("raster_words", ctypes.c_uint64),
        ("pull_entry_bytes", ctypes.c_uint64),
        ("pull_slices", ctypes.c_uint64),
    ]


class Counters(ctypes.Structure):
    _fields_ = [
        ("cycle", ctypes.c_uint64),
        ("events", ctypes.c_uint64),
        ("spikes", ctypes.c_uint64),
        ("drive_events", ctypes.c_uint64),
        ("overflow", ctypes.c_uint64),
        ("rows", ctypes.c_uint64),
        ("deferred", ctypes.c_uint64),
        ("late_deferred", ctypes.c_uint64),
    ]


class Tiling(ctypes.Structure):
    _fields_ = [
        ("update_block", ctypes.c_uint32),
        ("update_ids_per_lane", ctypes.c_uint32),
        ("gather_block", ctypes.c_uint32),
        ("gather_grid", ctypes.c_uint32),
        ("gather_stage", ctypes.c_uint32),
        ("gather_stage_max", ctypes.c_uint32),
        ("gather_batch", ctypes.c_uint32),
        ("mode", ctypes.c_int32),
        ("persist_block", ctypes.c_uint32),
        ("persist_grid", ctypes.c_uint32),
    ]


# (name, restype, argtypes) of every entry point declared in include/szsnn_hip.h
_vp = ctypes.c_void_p
PROTOTYPES = [
    ("szsnn_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    ("szsnn_destroy", ctypes.c_int, [_vp]),
    ("szsnn_last_error", ctypes.c_char_p, [_vp]),
    ("szsnn_device_name", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_size_t]),
    ("szsnn_alloc", ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    ("szsnn_free", ctypes.c_int, [_vp, _vp]),
    ("szsnn_memcpy_h2d", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    ("szsnn_memcpy_d2h", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    ("szsnn_synchronize", ctypes.c_int, [_vp]),
    ("szsnn_stream", _vp, [_vp]),
    ("szsnn_net_upload", ctypes.c_int, [_vp, ctypes.POINTER(NetDesc), ctypes.POINTER(_vp)]),
    ("szsnn_net_destroy", ctypes.c_int, [_vp]),
    ("szsnn_batch_upload", ctypes.c_int, [_vp, ctypes.POINTER(NetDesc), ctypes.c_uint32, ctypes.POINTER(_vp)]),
    ("szsnn_batch_instance", ctypes.c_int, [_vp, ctypes.c_uint32, c_u32p, c_u32p, c_u32p, c_u32p]),
    ("szsnn_net_get_info", ctypes.c_int, [_vp, ctypes.POINTER(NetInfo)]),
    ("szsnn_step", ctypes.c_int, [_vp, c_u32p, ctypes.c_uint32, c_u64p]),
    ("szsnn_run", ctypes.c_int, [_vp, ctypes.c_uint64]),
    ("szsnn_run_async", ctypes.c_int, [_vp, ctypes.c_uint64]),
    ("szsnn_set_volley", ctypes.c_int, [_vp, ctypes.c_uint32]),
    ("szsnn_set_recording", ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_uint64]),
    ("szsnn_read_raster", ctypes.c_int, [_vp, c_u64p]),
    ("szsnn_get_vars", ctypes.c_int, [_vp, c_f32p]),
    ("szsnn_set_vars", ctypes.c_int, [_vp, c_f32p]),
    ("szsnn_get_counters", ctypes.c_int, [_vp, ctypes.POINTER(Counters)]),
    ("szsnn_reset_counters", ctypes.c_int, [_vp]),
    ("szsnn_time_kernels", ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)]),
    ("szsnn_set_tiling", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    ("szsnn_get_tiling", ctypes.c_int, [_vp, c_u32p, c_u32p, c_u32p]),
    ("szsnn_autotune", ctypes.c_int, [_vp, ctypes.c_uint32]),
    ("szsnn_test_kernel", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]),
    ("szsnn_set_tiling_ex", ctypes.c_int, [_vp, ctypes.POINTER(Tiling)]),
    ("szsnn_get_tiling_ex", ctypes.c_int, [_vp, ctypes.POINTER(Tiling)]),
    ("szsnn_set_mode", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32]),
    ("szsnn_set_gather_policy", ctypes.c_int, [_vp, ctypes.c_float]),
    ("szsnn_get_gather_policy", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int32)]),
    ("szsnn_set_heavy_push", ctypes.c_int, [_vp, ctypes.c_uint32]),
    ("szsnn_get_heavy_push", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint32)]),
    ("szsnn_adj_forward_segment", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                 _vp, _vp, _vp, _vp, _vp, _vp]),
    ("szsnn_adj_backward_segment", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                  _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("szsnn_adj_input_currents", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _vp, _vp,
                                                 _vp, _vp]),
    ("szsnn_adj_input_grad_workspace", ctypes.c_size_t, [ctypes.c_uint32, ctypes.c_uint32]),
    ("szsnn_adj_input_grad", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, _vp,
                                             _vp, _vp, _vp, ctypes.c_size_t]),
    ("szsnn_adj_input_grad_noalloc", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _vp,
                                                     _vp, _vp, _vp, _vp]),
    ("szsnn_readout_forward", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, _vp, _vp,
                                             _vp]),
    ("szsnn_readout_backward", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                              ctypes.c_float, _vp, _vp, _vp]),
    ("szsnn_get_mode", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32), c_u32p, c_u32p]),
]

_lib = None


def load():
    """Load libszsnn_hip.so once; raise ImportError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m szsnn_amd.build` "
                          "(the HIP library is required; there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in PROTOTYPES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class SzsnnError(RuntimeError):
    """Raised for a negative return code; mirrors szsnn::SzsnnException."""

    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


def check(rc, ctx=None):
    if rc != SZSNN_OK:
        lib = load()
        msg = lib.szsnn_last_error(ctx)
        raise SzsnnError(rc, msg.decode() if msg else "unknown error")
    return rc


def ptr(arr, ctype):
    """Pointer to a contiguous numpy array (or NULL for None)."""
    if arr is None:
        return ctypes.POINTER(ctype)()
    return arr.ctypes.data_as(ctypes.POINTER(ctype))
