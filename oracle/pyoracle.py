"""Python access to the CPU oracle (oracle/libszsnn_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product path never loads it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libszsnn_oracle.so")

c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_u16p = ctypes.POINTER(ctypes.c_uint16)
c_f32p = ctypes.POINTER(ctypes.c_float)
c_f64p = ctypes.POINTER(ctypes.c_double)
c_u8p = ctypes.POINTER(ctypes.c_uint8)


class OracleNetDesc(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int32), ("n_neurons", ctypes.c_uint32), ("n_sources", ctypes.c_uint32),
        ("source_model", ctypes.c_int32), ("params", c_f32p), ("vars", c_f32p),
        ("src_rate", c_f32p), ("src_last", c_f32p), ("src_thresh", c_u64p),
        ("n_syn", ctypes.c_uint64), ("syn_src", c_u32p), ("syn_tgt", c_u32p), ("syn_w", c_f32p),
        ("syn_delay", c_u16p), ("max_delay", ctypes.c_uint32), ("frac_bits", ctypes.c_int32),
        ("seed", ctypes.c_uint64), ("noise_amp", c_f32p), ("drive_k", ctypes.c_uint32),
        ("drive_cdf", c_u64p), ("drive_w", ctypes.c_float),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.oracle_last_error.restype = ctypes.c_char_p
    lib.oracle_create.restype = ctypes.c_void_p
    lib.oracle_create.argtypes = [ctypes.POINTER(OracleNetDesc), ctypes.c_int]
    lib.oracle_destroy.argtypes = [ctypes.c_void_p]
    lib.oracle_run.restype = ctypes.c_int
    lib.oracle_run.argtypes = [ctypes.c_void_p, ctypes.c_uint64, c_u64p, c_u32p, ctypes.c_uint64, c_u64p,
                               c_f64p, ctypes.c_uint32]
    lib.oracle_run_parallel.restype = ctypes.c_int
    lib.oracle_run_parallel.argtypes = [ctypes.c_void_p, ctypes.c_uint64, c_u64p, c_u32p, ctypes.c_uint64, c_u64p,
                                        c_f64p, ctypes.c_uint32, ctypes.c_int]
    lib.oracle_time_steps.restype = ctypes.c_double
    lib.oracle_time_steps.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    lib.oracle_time_steps_parallel.restype = ctypes.c_double
    lib.oracle_time_steps_parallel.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int]
    for f in ("oracle_words", "oracle_events", "oracle_drive_events"):
        getattr(lib, f).restype = ctypes.c_uint64
        getattr(lib, f).argtypes = [ctypes.c_void_p]
    lib.oracle_get_vars.argtypes = [ctypes.c_void_p, c_f64p]
    lib.oracle_neuron_trace.restype = ctypes.c_int
    lib.oracle_neuron_trace.argtypes = [ctypes.c_int, ctypes.c_int, c_f64p, c_f64p, c_f64p, c_f64p, c_f64p,
                                        c_u8p, ctypes.c_uint32, c_f64p, c_u8p]
    lib.oracle_splitmix_at.restype = ctypes.c_uint64
    lib.oracle_splitmix_at.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    lib.oracle_draw.restype = ctypes.c_uint64
    lib.oracle_draw.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint]
    lib.oracle_gauss_int.restype = ctypes.c_int32
    lib.oracle_gauss_int.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    lib.oracle_gauss_f32.restype = ctypes.c_float
    lib.oracle_gauss_f32.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    lib.oracle_gauss_f64.restype = ctypes.c_double
    lib.oracle_gauss_f64.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    _lib = lib
    return lib


def _p(a, t):
    return ctypes.POINTER(t)() if a is None else a.ctypes.data_as(ctypes.POINTER(t))


N_VARS = {0: 4, 1: 2, 2: 4}


def host_threads():
    """Cores this process may use (the GPU box gives a share of its host)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


class Oracle:
    """One engine instance (precision 32 or 64) over a network description
    with the attributes of szsnn_amd.network.Network."""

    def __init__(self, net, precision=64):
        lib = load()
        self.lib = lib
        keep = []

        def c(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        d = OracleNetDesc()
        d.model = net.model
        d.n_neurons = net.n_neurons
        d.n_sources = net.n_sources
        d.source_model = net.source_model
        d.params = _p(c(net.params, np.float32), ctypes.c_float)
        d.vars = _p(c(net.vars, np.float32), ctypes.c_float)
        d.src_rate = _p(c(net.src_rate, np.float32), ctypes.c_float)
        d.src_last = _p(c(net.src_last, np.float32), ctypes.c_float)
        d.src_thresh = _p(c(net.src_thresh, np.uint64), ctypes.c_uint64)
        d.n_syn = net.syn_src.shape[0]
        d.syn_src = _p(c(net.syn_src, np.uint32), ctypes.c_uint32)
        d.syn_tgt = _p(c(net.syn_tgt, np.uint32), ctypes.c_uint32)
        d.syn_w = _p(c(net.syn_w, np.float32), ctypes.c_float)
        d.syn_delay = _p(c(net.syn_delay, np.uint16), ctypes.c_uint16)
        d.max_delay = net.max_delay
        d.frac_bits = net.frac_bits
        d.seed = net.seed & 0xFFFFFFFFFFFFFFFF
        d.noise_amp = _p(c(net.noise_amp, np.float32), ctypes.c_float)
        cdf = c(net.drive_cdf, np.uint64)
        d.drive_k = 0 if cdf is None else cdf.shape[0]
        d.drive_cdf = _p(cdf, ctypes.c_uint64)
        d.drive_w = net.drive_w
        h = lib.oracle_create(ctypes.byref(d), precision)
        if not h:
            raise RuntimeError("oracle: " + lib.oracle_last_error().decode())
        self.h = ctypes.c_void_p(h)
        self.n = net.n_neurons
        self.nv = N_VARS[net.model]
        self.words = int(lib.oracle_words(self.h))
        self.precision = precision

    def run(self, steps, forced=None, record_v=False, threads=1):
        """forced: list of (step, id), or a pair of arrays (steps, ids).
        threads > 1: the FP32 engine's OpenMP step (identical results).
        Returns (raster [steps, words] u64, v [steps, N] float64 or None)."""
        raster = np.zeros((steps, self.words), np.uint64)
        v = np.zeros((steps, self.n), np.float64) if record_v else None
        if isinstance(forced, tuple):
            ft = np.asarray(forced[0], np.uint64)
            fi = np.asarray(forced[1], np.uint32)
            order = np.lexsort((fi, ft))
            ft = np.ascontiguousarray(ft[order])
            fi = np.ascontiguousarray(fi[order])
        elif forced:
            f = sorted(forced)
            ft = np.array([a for a, _ in f], np.uint64)
            fi = np.array([b for _, b in f], np.uint32)
        else:
            ft = fi = None
        args = (self.h, steps, _p(ft, ctypes.c_uint64), _p(fi, ctypes.c_uint32), 0 if ft is None else ft.shape[0],
                _p(raster, ctypes.c_uint64), _p(v, ctypes.c_double), self.n)
        if threads > 1:
            rc = self.lib.oracle_run_parallel(*args, int(threads))
        else:
            rc = self.lib.oracle_run(*args)
        if rc != 0:
            raise RuntimeError("oracle: " + self.lib.oracle_last_error().decode())
        return raster, v

    def time_steps(self, steps, threads=1):
        """Wall seconds of `steps` steps (no forced input): the serial step,
        or, with threads > 1, the FP32 engine's parallel step."""
        if threads <= 1:
            return self.lib.oracle_time_steps(self.h, steps)
        secs = self.lib.oracle_time_steps_parallel(self.h, steps, int(threads))
        if secs < 0:
            raise RuntimeError("oracle: the parallel step needs the FP32 engine")
        return secs

    def vars(self):
        out = np.zeros((self.nv, self.n), np.float64)
        self.lib.oracle_get_vars(self.h, _p(out, ctypes.c_double))
        return out

    @property
    def events(self):
        return int(self.lib.oracle_events(self.h))

    @property
    def drive_events(self):
        return int(self.lib.oracle_drive_events(self.h))

    def __del__(self):
        try:
            self.lib.oracle_destroy(self.h)
        except Exception:
            pass


def neuron_trace(model, precision, params, vars0, accE, accI, iinj=None, forced=None):
    """Run one neuron with given per-step inputs; returns (trace [steps, nv], spikes [steps])."""
    lib = load()
    steps = len(accE)
    nv = N_VARS[model]
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(vars0, np.float64).copy()
    ae = np.ascontiguousarray(accE, np.float64)
    ai = np.ascontiguousarray(accI, np.float64)
    ij = None if iinj is None else np.ascontiguousarray(iinj, np.float64)
    fo = None if forced is None else np.ascontiguousarray(forced, np.uint8)
    tr = np.zeros((steps, nv), np.float64)
    sp = np.zeros(steps, np.uint8)
    rc = lib.oracle_neuron_trace(model, precision, _p(p, ctypes.c_double), _p(x, ctypes.c_double),
                                 _p(ae, ctypes.c_double), _p(ai, ctypes.c_double), _p(ij, ctypes.c_double),
                                 _p(fo, ctypes.c_uint8), steps, _p(tr, ctypes.c_double), _p(sp, ctypes.c_uint8))
    assert rc == 0
    return tr, sp.astype(bool)
