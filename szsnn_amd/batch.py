"""Batched mode: many independent network instances in one device network.

`BatchSimulation` uploads a list of `Network`s of one model through
szsnn_batch_upload; every step kernel then covers all instances in a single
launch (one graph-captured step loop, one tiling chosen by the autotuner).
Each instance keeps its own seed-derived random streams, so its raster equals
the raster of the same network run alone.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, ptr
from .configs import DEFAULT_SEED, izhikevich2003
from .simulation import Simulation, _desc, _live_nets, raster_to_bool


def _bits(raster, first, count):
    """bool [steps, count] of ids first..first+count-1 from uint64 bitmaps."""
    if count == 0:
        return np.zeros((raster.shape[0], 0), bool)
    w0, w1 = first // 64, (first + count + 63) // 64
    sub = np.ascontiguousarray(raster[:, w0:w1])
    bits = np.unpackbits(sub.view(np.uint8).reshape(sub.shape[0], -1), axis=1, bitorder="little")
    off = first - w0 * 64
    return bits[:, off:off + count].astype(bool)


def batched_izhikevich(seed=DEFAULT_SEED, n_instances=256, n_min=500, n_max=20000, fanout=100,
                       max_delay_steps=20, amp_lo=2.0, amp_hi=8.0):
    """Instances of the Izhikevich 2003 network with log-uniform sizes in
    [n_min, n_max] (80/20 exc/inh), instance seed = seed + i, noise amplitudes
    (exc, inh) ~ U[amp_lo, amp_hi] per instance."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_instances))).astype(int)
    amps = rng.uniform(amp_lo, amp_hi, size=(n_instances, 2))
    nets = []
    for i, n in enumerate(sizes):
        ne = int(round(0.8 * n))
        nets.append(izhikevich2003(seed=seed + 1 + i, n_exc=ne, n_inh=int(n) - ne, fanout=fanout,
                                   max_delay_steps=max_delay_steps, noise_exc=float(amps[i, 0]),
                                   noise_inh=float(amps[i, 1])))
    return nets


class BatchSimulation(Simulation):
    """A batch of instances on one device; the Simulation calls act on all."""

    def __init__(self, context, networks):
        if isinstance(context, int):
            from .simulation import Context
            context = Context(context)
        self.ctx = context
        self.lib = context.lib
        self.networks = list(networks)
        arr = (_lib.NetDesc * len(self.networks))()
        keep = []
        for i, net in enumerate(self.networks):
            d, k = _desc(net)
            arr[i] = d
            keep.append(k)
        h = ctypes.c_void_p()
        check(self.lib.szsnn_batch_upload(context.handle, arr, len(self.networks), ctypes.byref(h)),
              context.handle)
        del keep
        self.handle = h
        info = _lib.NetInfo()
        check(self.lib.szsnn_net_get_info(h, ctypes.byref(info)), context.handle)
        self.info = info
        self.n_neurons = info.n_neurons
        self.words = int(info.raster_words)
        self._wall_ms = 0.0
        self._rec = None
        self._bitmap = np.zeros(self.words, np.uint64)
        self.network = None
        _live_nets.add(self)
        self.offsets = []
        for b in range(len(self.networks)):
            no, n, so, m = (ctypes.c_uint32() for _ in range(4))
            check(self.lib.szsnn_batch_instance(h, b, ctypes.byref(no), ctypes.byref(n), ctypes.byref(so),
                                                ctypes.byref(m)), context.handle)
            self.offsets.append((no.value, n.value, so.value, m.value))

    def instance_raster(self, b, raster=None):
        """bool [steps, n_b + m_b] for instance b (its neurons, then its sources)."""
        r = self.raster() if raster is None else raster
        no, n, so, m = self.offsets[b]
        return np.concatenate([_bits(r, no, n), _bits(r, so, m)], axis=1)

    def instance_rasters(self, raster=None):
        """All instances' rasters, unpacking each only over its own id range."""
        r = self.raster() if raster is None else raster
        return [self.instance_raster(b, r) for b in range(len(self.networks))]

    def instance_vars(self, b, vars_=None):
        v = self.get_vars() if vars_ is None else vars_
        no, n, _, _ = self.offsets[b]
        return v[:, no:no + n]
