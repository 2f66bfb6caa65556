"""Extract known-answer vectors from the reference's logged traces.

Reads data files shipped with the reference checkout (traces and connection
exports, not source code) and writes small JSON/text fixtures next to this
script.  Run once in a container that has the reference at /root/reference:

    python tests/golden/make_kat.py

Sources (paths relative to the reference checkout):
  szsnn1.1/bin/myfile.txt~       FP32 debug trace of the reference GPU kernel
                                 (IF_curr_exp_update_kernel.cl:238-243,264-270:
                                 one traced neuron of partition 0, one line group per cycle)
  szsnn1.1/bin/pyNN2neurony.vvv~ FP64 v trace of the same neuron from pyNN.brian
  szsnn1.1/bin/izi1.conn~,
  szsnn1.1/bin/izi2.conn~        Brian connection exports (Connection.hpp:76-145)
"""
import json
import os
import re
import sys

REF = "/root/reference/szsnn1.1/bin"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_myfile(path):
    ie_re = re.compile(r">>>>>>IE (\S+) last(\S+) in(\S+) tau(\S+)")
    ii_re = re.compile(r">>>>>>II (\S+) last(\S+) in(\S+) tau(\S+)")
    v_re = re.compile(r">>>>>>V (\S+) last(\S+) rest(\S+) tau(\S+) reset(\S+) I(\S+), cm(\S+), delta(\S+)")
    cycles = []
    cur = {}
    meta = {}
    with open(path, "r", errors="replace") as f:
        for line in f:
            m = ie_re.search(line)
            if m:
                cur = {"ie": m.group(1), "in_e": m.group(3)}
                meta["tau_syn_E"] = m.group(4)
                continue
            m = ii_re.search(line)
            if m:
                cur["ii"] = m.group(1)
                cur["in_i"] = m.group(3)
                meta["tau_syn_I"] = m.group(4)
                continue
            m = v_re.search(line)
            if m:
                cur["v"] = m.group(1)
                cur["v_last"] = m.group(2)
                cur["I"] = m.group(6)
                meta["v_rest"] = m.group(3)
                meta["tau_m"] = m.group(4)
                meta["v_reset"] = m.group(5)
                meta["cm"] = m.group(7)
                cycles.append(cur)
                cur = {}
    return meta, cycles


def parse_vvv(path, limit):
    vals = []
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                continue
            parts = line.split()
            if len(parts) == 2:
                vals.append(parts[0])
            if len(vals) >= limit:
                break
    return vals


def main():
    if not os.path.isdir(REF):
        sys.exit("reference checkout not found at " + REF)
    meta, cycles = parse_myfile(os.path.join(REF, "myfile.txt~"))
    out = {
        "source": "szsnn1.1/bin/myfile.txt~ (reference FP32 GPU debug trace of one neuron)",
        "printed_params": meta,
        "note": "v is printed after the update and before any reset; in_e/in_i are the "
                "gathered currents; values printed with %f",
        "in_e": [c["in_e"] for c in cycles],
        "in_i": [c["in_i"] for c in cycles],
        "ie": [c["ie"] for c in cycles],
        "v": [c["v"] for c in cycles],
        # The log's "fired" printf lines name other neurons; the traced
        # neuron's spikes show as a reset: the next cycle starts from v_reset.
        "fired_cycles": [k for k in range(len(cycles) - 1)
                         if cycles[k + 1]["v_last"] == meta["v_reset"]
                         and float(cycles[k]["v"]) >= -50.0],
    }
    with open(os.path.join(HERE, "kat_if_curr_exp_trace.json"), "w") as f:
        json.dump(out, f, indent=0)
    # Brian FP64 subthreshold run: v0=-70, v_rest=-49, tau_m=20, cm=1, I=0;
    # the first 59 samples precede the threshold crossing.
    vvv = parse_vvv(os.path.join(REF, "pyNN2neurony.vvv~"), 59)
    with open(os.path.join(HERE, "kat_brian_fp64_subthreshold.json"), "w") as f:
        json.dump({"source": "szsnn1.1/bin/pyNN2neurony.vvv~ lines 10-68",
                   "v0": -70.0, "v_rest": -49.0, "tau_m": 20.0, "cm": 1.0, "v": vvv}, f, indent=0)
    for name in ("izi1.conn~", "izi2.conn~"):
        with open(os.path.join(REF, name)) as src, \
                open(os.path.join(HERE, name.rstrip("~")), "w") as dst:
            dst.write(src.read())
    print("cycles:", len(cycles), "fired:", len(out["fired_cycles"]), "vvv:", len(vvv))


if __name__ == "__main__":
    main()
