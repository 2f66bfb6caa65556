"""CPU tests: the C-ABI library loads and exports every declared symbol, and
the host-side network builders produce what the configurations specify."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(szsnn_[a-z_0-9]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    import ctypes
    from szsnn_amd import _lib
    names = _declared_functions(os.path.join(ROOT, "include", "szsnn_hip.h"))
    assert "szsnn_net_upload" in names and "szsnn_step" in names and len(names) >= 25
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared entry point has a ctypes prototype
    assert sorted(n for n, _, _ in _lib.PROTOTYPES) == names


def test_no_device_gives_clean_error():
    """Without a GPU, szsnn_create fails with a negative code, not a crash."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from szsnn_amd import _lib
    import ctypes
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.szsnn_create(0, ctypes.byref(h))
    assert rc < 0
    assert lib.szsnn_last_error(None)


def test_struct_layout_matches_header():
    """ctypes struct sizes equal the C struct sizes (compiled probe)."""
    import ctypes
    import subprocess
    import tempfile
    from szsnn_amd import _lib
    src = r'''
#include <stdio.h>
#include "szsnn_hip.h"
int main(void){printf("%zu %zu %zu\n", sizeof(szsnn_net_desc), sizeof(szsnn_net_info), sizeof(szsnn_counters));return 0;}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "p.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "p")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], stdout=subprocess.PIPE, text=True, check=True).stdout.split()
    assert [int(x) for x in out] == [ctypes.sizeof(_lib.NetDesc), ctypes.sizeof(_lib.NetInfo),
                                     ctypes.sizeof(_lib.Counters)]


def test_quantize_and_frac_bits():
    from szsnn_amd.network import choose_frac_bits, quantize_weights
    assert choose_frac_bits(1.0) == 30
    assert choose_frac_bits(0.051) == 34
    w = np.array([0.05, -0.25, 0.004, 1e-12, 0.1234567], np.float32)
    F = 34
    q = quantize_weights(w, F)
    fx = q.astype(np.float64) * 2.0 ** F
    assert np.all(fx == np.rint(fx))
    assert q[4] == w[4] and q[0] == w[0] and q[2] == w[2]  # already multiples of 2^-34
    assert q[3] == 0.0


def test_distinct_choice():
    from szsnn_amd.network import distinct_choice
    rng = np.random.default_rng(0)
    x = distinct_choice(rng, 500, 90, 100, exclude=np.arange(500) % 100)
    assert x.shape == (500, 90)
    for r in range(500):
        assert len(set(x[r])) == 90
        assert (r % 100) not in set(x[r])
        assert x[r].min() >= 0 and x[r].max() < 100


def test_izhikevich2003_builder():
    from szsnn_amd.configs import izhikevich2003
    net = izhikevich2003()
    assert net.n_neurons == 1000 and net.n_syn == 100000 and net.max_delay == 20
    assert net.syn_delay.max() == 19 and net.syn_delay.min() == 0
    exc = net.syn_src < 800
    assert np.all(net.syn_w[exc] >= 0) and np.all(net.syn_w[exc] < 0.5)
    assert np.all(net.syn_w[~exc] <= 0) and np.all(net.syn_w[~exc] > -1)
    assert np.all(net.syn_src != net.syn_tgt)
    a, b, c, d = net.params
    assert np.all(a[:800] == np.float32(0.02)) and np.all(c[800:] == -65)
    assert np.all((c[:800] >= -65) & (c[:800] <= -50)) and np.all((d[:800] >= 2) & (d[:800] <= 8))
    np.testing.assert_array_equal(net.vars[1], (net.params[1] * np.float32(-65.0)).astype(np.float32))
    assert set(np.unique(net.noise_amp)) == {2.0, 5.0}


def test_if_curr_exp_10k_builder():
    from szsnn_amd.configs import if_curr_exp_10k
    net = if_curr_exp_10k()
    n = 10000
    assert net.n_neurons == n and net.n_sources == 1000
    rec = net.syn_src < n
    assert rec.sum() == n * 100
    indeg = np.bincount(net.syn_tgt[rec], minlength=n)
    assert np.all(indeg == 100)
    assert np.all(net.syn_src[rec] != net.syn_tgt[rec])
    assert np.bincount(net.syn_src[~rec] - n).tolist() == [50] * 1000
    assert net.max_delay == 16
    # per-step parameters at dt = 0.1 ms
    p = net.params[:, 0]
    assert p[1] == 10.0 and p[2] == 200.0 and p[3] == 20.0 and p[4] == 50.0
    assert net.src_thresh[0] == np.uint64(int(0.001 * 2 ** 64))


def test_if_cond_exp_builder_small():
    from szsnn_amd.configs import if_cond_exp_100k
    net = if_cond_exp_100k(n_exc=800, n_inh=200, k_exc=80, k_inh=20)
    assert net.n_syn == 1000 * 100
    indeg_e = np.bincount(net.syn_tgt[net.syn_w > 0], minlength=1000)
    assert np.all(indeg_e == 80)
    assert np.all(net.syn_w[net.syn_src >= 800] < 0)
    assert net.drive_cdf is not None and len(net.drive_cdf) >= 5
    assert net.params.shape == (11, 1000) and net.params[5, 0] == 100.0  # tau_syn_I 10 ms / 0.1


def test_batched_izhikevich_builder():
    from szsnn_amd.batch import batched_izhikevich
    nets = batched_izhikevich(n_instances=16, fanout=5)
    sizes = np.array([n.n_neurons for n in nets])
    assert len(nets) == 16 and sizes.min() >= 500 and sizes.max() <= 20000
    assert len({n.seed for n in nets}) == 16
    for n in nets:
        assert set(np.unique(n.noise_amp)) <= set(np.unique(n.noise_amp))
        amps = np.unique(n.noise_amp)
        assert len(amps) == 2 and np.all((amps >= 2) & (amps <= 8))
        assert n.meta["n_exc"] == round(0.8 * n.n_neurons)


def test_layout_and_gather_index_check_cpu(tmp_path):
    """Device formats built by the runtime's builder, and the push/pull gather
    index arithmetic replayed on the CPU with bounds checks (ASan/UBSan)."""
    import subprocess
    exe = tmp_path / "layout_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fsanitize=address,undefined", "-I",
                    os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "layout_check.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "layout_check OK" in r.stdout
