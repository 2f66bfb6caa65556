"""Per-kernel timing (szsnn_time_kernels): the steps it runs, one kernel
launch at a time with events recorded by each launch's dispatch, are the
steps a graph run makes (rasters, state and counters bit for bit), and the
durations it returns are those of launches that ran."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from szsnn_amd.simulation import Context
    return Context(0)


def _same(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    for k in ("events", "spikes", "rows", "drive_events"):
        assert a[2][k] == b[2][k], k


@pytest.mark.parametrize("layout", [0, -1])
def test_timed_steps_equal_graph_steps(ctx, layout):
    """The 10k IF_cond_exp network under a volley of 2000 ids (pulled steps
    and pushed ones) with its pull layout, and without one (every step
    pushed, no gather kernel: its events bracket nothing)."""
    from szsnn_amd.configs import if_cond_exp_100k
    from szsnn_amd.simulation import Simulation
    net = dataclasses.replace(if_cond_exp_100k(n_exc=8000, n_inh=2000, k_exc=400, k_inh=100), pull_layout=layout)
    settle, steps = 300, 40
    print(f"case: timed steps against graph steps, pull layout {layout}", flush=True)
    runs = []
    for timed in (False, True):
        s = Simulation(ctx, net)
        s.set_recording(0, settle + steps)
        s.run(settle)
        s.set_volley(2000)
        if timed:
            g_ms, u_ms = s.time_kernels(steps)
        else:
            s.run(steps)
        runs.append((s.raster(), s.get_vars(), s.counters()))
        s.close()
    _same(runs[1], runs[0])
    assert runs[0][2]["events"] > 0
    # every launch ran: positive durations, each below a second
    assert 0.0 < u_ms < 1e3
    if layout == -1:
        assert 0.0 <= g_ms < 1.0
    else:
        assert 0.0 < g_ms < 1e3
