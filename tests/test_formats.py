"""On-disk formats (CPU): .conn, population and spike files."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_conn_fixture_and_round_trip(tmp_path):
    from szsnn_amd.formats import load_conn, save_conn
    s, t, w, d = load_conn(os.path.join(GOLDEN, "izi1.conn"))
    assert len(s) == 77 and s[0] == 0 and t[0] == 13 and w[0] == np.float32(2e-10) and d.max() == 0
    s2, t2, w2, d2 = load_conn(os.path.join(GOLDEN, "izi2.conn"), from_start=80)
    assert np.all(s2 >= 80) and np.all(w2 == np.float32(-0.03))
    p = tmp_path / "x.conn"
    save_conn(p, s, t, w, d)
    lines = p.read_text().splitlines()
    assert lines[0] == "#label p1" and lines[1] == "#label p2" and lines[2].split("\t")[3] == "0"
    r = load_conn(p)
    for a, b in zip((s, t, w, d), r):
        np.testing.assert_array_equal(a, b)


def test_conn_rejects_bad_line(tmp_path):
    from szsnn_amd.formats import FormatError, load_conn
    p = tmp_path / "bad.conn"
    p.write_text("#a\n#b\n0 1 0.5 0\n1 2 0.5\n")
    with pytest.raises(FormatError, match="line #4"):
        load_conn(p)


def test_population_round_trip(tmp_path):
    from szsnn_amd.formats import load_population, load_population_template, save_population
    from szsnn_amd.network import if_curr_exp_params, if_vars
    n = 7
    params = if_curr_exp_params(n, dt=0.1, tau_refrac=2.0)
    params[0] = np.arange(n) - 70.0
    vars_ = if_vars(n, np.linspace(-70, -60, n).astype(np.float32))
    p = tmp_path / "pop.txt"
    save_population(p, 0, params, vars_, first_id=100, label="exc", model_name="IF_curr_exp")
    P, V, first, label = load_population(p, 0)
    np.testing.assert_array_equal(P, params)
    np.testing.assert_array_equal(V, vars_)
    assert first == 100 and label == "exc"
    t = tmp_path / "tmpl.txt"
    t.write_text("# template\n3\n5\ninh\n0.1 0.2 -65 2\n-65 -13\n")
    P, V, first, label = load_population_template(t, 1)
    assert P.shape == (4, 3) and np.all(P[0] == np.float32(0.1)) and np.all(V[1] == -13)


def test_spike_recorder_files(tmp_path):
    from szsnn_amd.formats import FormatError, SpikeRecorder, read_spikes
    r = SpikeRecorder(record_start=1)
    raster = np.zeros((4, 10), bool)
    raster[0, 3] = raster[1, 2] = raster[1, 7] = raster[3, 2] = True
    r.record_raster(raster)
    with pytest.raises(FormatError):
        r.record_cycle_spikes(2, [1])
    p = tmp_path / "s.txt"
    r.save_by_cycles(p)
    lines = p.read_text().splitlines()
    # the header ends with a newline even without a stop cycle (reference D6)
    assert lines[:2] == ["#spikes by cycles", "#record start at: 1"]
    assert lines[2:] == ["1\t2", "1\t7", "3\t2"]
    c, i = read_spikes(p)
    assert c.tolist() == [1, 1, 3] and i.tolist() == [2, 7, 2]
    q = tmp_path / "n.txt"
    r.save_by_neurons(q)
    assert q.read_text().splitlines()[2:] == ["1\t2", "3\t2", "1\t7"]
    with pytest.raises(FormatError):
        SpikeRecorder().record_cycle_spikes(0, [1, 1])
    # the stop cycle itself is recorded (SpikeRecorder.cpp: returns for cycle > stop)
    s = SpikeRecorder(record_start=1, record_stop=2)
    for c in range(4):
        s.record_cycle_spikes(c, [c])
    assert sorted(s.by_cycle) == [1, 2]


def test_population_files_in_the_brian_dumper_orders(tmp_path):
    """The pyszsnn Brian dumper (pyszsnn/brian/dumper.py:12-126) wrote the label
    (templates) or the model name and the size twice (populations) before the
    numbers; the readers take those orders as well as the reference's."""
    from szsnn_amd.formats import FormatError, load_population, load_population_template
    t1 = tmp_path / "b1.tmpl"
    t1.write_text("# population template saved by pyszsnn.brian\n# Izhikevich\nexc\n3\n5\n0.02 0.2 -65 8\n-65 -13\n")
    P, V, first, label = load_population_template(t1, 1)
    assert P.shape == (4, 3) and first == 5 and label == "exc" and np.all(P[3] == 8) and np.all(V[1] == -13)
    t2 = tmp_path / "b2.tmpl"   # written with the default empty label: a blank line
    t2.write_text("# population template saved by pyszsnn.brian\n# Izhikevich\n\n3\n5\n0.02 0.2 -65 8\n-65 -13\n")
    P, V, first, label = load_population_template(t2, 1)
    assert P.shape == (4, 3) and first == 5 and label == ""
    pp = tmp_path / "b.pop"
    pp.write_text("# population saved by pyszsnn.brian\nIzhikevich\n2\n2\n7\n"
                  "0.02 0.2 -65 8\n-65 -13\n0.03 0.2 -65 8\n-64 -13\n")
    P, V, first, label = load_population(pp, 1)
    assert P.shape == (4, 2) and first == 7 and P[0, 1] == np.float32(0.03) and V[0, 1] == -64
    bad = tmp_path / "bad.pop"   # the size twice, but different
    bad.write_text("Izhikevich\n2\n3\n7\n0.02 0.2 -65 8\n-65 -13\n")
    with pytest.raises(FormatError):
        load_population(bad, 1)
