"""On-disk formats of the reference, read and written from Python.

* Connection files (szsnn1.1/src/Connection.hpp:76-174): two header lines,
  then one synapse per line, `src tgt weight delay`, whitespace separated,
  exactly four values; ids and delay are parsed as float and truncated; ids
  are relative to the projection's (from, to) offsets; the delay is the API
  delay.  The writer emits `#label p1`, `#label p2`, then lines sorted by
  (src, tgt), tab separated, the delay printed as a float.
* Population files (szsnn1.1/src/Population.hpp:118-397): `#` comment lines,
  then the size, the first neuron id and the label, then per neuron one line
  of parameters and one line of variables.  The template variant holds one
  parameter line and one variable line for every neuron.  The reference's
  own writer omitted the label its reader expects, and the pyszsnn Brian
  dumper (pyszsnn/brian/dumper.py:12-126) put the label first (templates:
  label, size, first id; populations: model name, size, size again, first
  id); the reference could read neither.  The readers here take all three
  header orders: a header that starts with a number is size, first id,
  label; one that starts with a word is that word followed by the integers
  up to the first parameter line, of which the first is the size and the
  last the first id.
* Spike files (szsnn1.1/src/SpikeRecorder.cpp:13-104): `#spikes by cycles` or
  `#spikes by neurons`, `#record start at: S[   stop at: E]` (always ended by
  a newline here), then `cycle<TAB>neuron` lines.
"""
import numpy as np

from .network import N_PARAMS, N_VARS


class FormatError(ValueError):
    pass


# ---------------------------------------------------------------- .conn
def load_conn(path, from_start=0, to_start=0):
    """Returns (src, tgt, weight, delay) arrays; ids offset by from/to_start.
    A repeated (src, tgt) pair keeps its last line (the reference stores the
    projection in a map keyed by the pair)."""
    pairs = {}
    with open(path) as f:
        f.readline()
        f.readline()
        for lineno, line in enumerate(f, start=3):
            vals = line.split()
            if not vals:
                continue
            if len(vals) != 4:
                raise FormatError(f"{path}: wrong format at line #{lineno}: {len(vals)} values")
            try:
                s, t, w, d = int(float(vals[0])), int(float(vals[1])), float(vals[2]), int(float(vals[3]))
            except ValueError as e:
                raise FormatError(f"{path}: wrong format at line #{lineno}: {e}") from None
            pairs[(s, t)] = (w, d)
    keys = sorted(pairs)
    src = np.array([k[0] + from_start for k in keys], np.uint32)
    tgt = np.array([k[1] + to_start for k in keys], np.uint32)
    w = np.array([pairs[k][0] for k in keys], np.float32)
    d = np.array([pairs[k][1] for k in keys], np.uint16)
    return src, tgt, w, d


def save_conn(path, src, tgt, weight, delay, labels=("p1", "p2")):
    order = np.lexsort((np.asarray(tgt), np.asarray(src)))
    with open(path, "w") as f:
        f.write(f"#label {labels[0]}\n#label {labels[1]}\n")
        for i in order:
            f.write(f"{int(src[i])}\t{int(tgt[i])}\t{float(np.float32(weight[i])):.9g}\t{float(delay[i]):g}\n")


# ---------------------------------------------------------------- populations
def _data_lines(path):
    with open(path) as f:
        for line in f:
            s = line.strip()
            if s and not s.startswith("#"):
                yield s


def _is_int(s):
    try:
        int(s)
        return True
    except ValueError:
        return False


def _header(lines, path):
    """(size, first id, label, index of the first line after the header) of
    a population or template file's data lines, in any of the three orders."""
    try:
        if _is_int(lines[0]):
            # a label is one word; a longer third line is already the first
            # parameter line (a Brian template written with an empty label)
            if len(lines) > 2 and len(lines[2].split()) == 1:
                return int(lines[0]), int(lines[1]), lines[2].strip(), 3
            return int(lines[0]), int(lines[1]), "", 2
        label, k, ints = lines[0].split()[0], 1, []
        while k < len(lines) and len(lines[k].split()) == 1 and _is_int(lines[k]):
            ints.append(int(lines[k]))
            k += 1
        if len(ints) not in (2, 3) or (len(ints) == 3 and ints[0] != ints[1]):
            raise FormatError(f"{path}: bad header: {len(ints)} integers after `{label}`")
        return ints[0], ints[-1], label, k
    except IndexError:
        raise FormatError(f"{path}: bad header: file too short") from None
    except ValueError as e:
        raise FormatError(f"{path}: bad header: {e}") from None


def load_population(path, model):
    """Returns (params [n_params, n], vars [n_vars, n], first_id, label)."""
    lines = list(_data_lines(path))
    n, first, label, k = _header(lines, path)
    it = iter(lines[k:])
    if n <= 0:
        raise FormatError(f"{path}: population size must be > 0")
    npar, nvar = N_PARAMS[model], N_VARS[model]
    params = np.zeros((npar, n), np.float32)
    vars_ = np.zeros((nvar, n), np.float32)
    for i in range(n):
        try:
            p = [float(x) for x in next(it).split()]
            v = [float(x) for x in next(it).split()]
        except StopIteration:
            raise FormatError(f"{path}: {n} neurons announced, {i} found") from None
        if len(p) != npar or len(v) != nvar:
            raise FormatError(f"{path}: neuron {i}: expected {npar} params and {nvar} vars")
        params[:, i] = p
        vars_[:, i] = v
    return params, vars_, first, label


def save_population(path, model, params, vars_, first_id=0, label="population", model_name=""):
    params = np.asarray(params, np.float32)
    vars_ = np.asarray(vars_, np.float32)
    n = params.shape[1]
    with open(path, "w") as f:
        f.write(f"# Population file for : {label}\n# Model: {model_name}\n# Size: {n}\n"
                f"# First neuron ID: {first_id}\n")
        f.write(f"{n}\n{first_id}\n{label}\n")
        for i in range(n):
            f.write(" ".join(f"{float(x):.9g}" for x in params[:, i]) + "\n")
            f.write(" ".join(f"{float(x):.9g}" for x in vars_[:, i]) + "\n")


def load_population_template(path, model):
    """Template variant: one params line and one vars line for all neurons."""
    lines = list(_data_lines(path))
    n, first, label, k = _header(lines, path)
    try:
        p = [float(x) for x in lines[k].split()]
        v = [float(x) for x in lines[k + 1].split()] if N_VARS[model] else []
    except (IndexError, ValueError) as e:
        raise FormatError(f"{path}: bad template: {e}") from None
    if n <= 0 or len(p) != N_PARAMS[model] or len(v) != N_VARS[model]:
        raise FormatError(f"{path}: bad template sizes")
    params = np.repeat(np.array(p, np.float32)[:, None], n, axis=1)
    vars_ = np.repeat(np.array(v, np.float32)[:, None], n, axis=1)
    return params, vars_, first, label


# ---------------------------------------------------------------- spikes
class SpikeRecorder:
    """recordCycleSpikes / saveToFileByCycles / saveToFileByNeurons of the
    reference (SpikeRecorder.cpp:13-104): a cycle may be recorded once, a
    spike once per cycle."""

    def __init__(self, record_start=0, record_stop=0):
        self.start = int(record_start)
        self.stop = int(record_stop)
        self.by_cycle = {}

    def record_cycle_spikes(self, cycle, ids):
        cycle = int(cycle)
        if cycle < self.start or (self.stop and cycle > self.stop):  # the stop cycle is recorded
            return
        if cycle in self.by_cycle:
            raise FormatError(f"cycle {cycle} already recorded")
        ids = [int(i) for i in ids]
        if len(set(ids)) != len(ids):
            raise FormatError(f"cycle {cycle}: a spike recorded twice")
        self.by_cycle[cycle] = sorted(ids)

    def record_raster(self, raster_bool, first_cycle=0):
        """Rows of a bool raster [steps, ids] as consecutive cycles."""
        for k, row in enumerate(np.asarray(raster_bool)):
            self.record_cycle_spikes(first_cycle + k, np.flatnonzero(row))

    def _header(self, kind):
        h = f"#spikes by {kind}\n#record start at: {self.start}"
        if self.stop:
            h += f"   stop at: {self.stop}"
        return h + "\n"

    def save_by_cycles(self, path):
        with open(path, "w") as f:
            f.write(self._header("cycles"))
            for c in sorted(self.by_cycle):
                for i in self.by_cycle[c]:
                    f.write(f"{c}\t{i}\n")

    def save_by_neurons(self, path):
        by_n = {}
        for c, ids in self.by_cycle.items():
            for i in ids:
                by_n.setdefault(i, []).append(c)
        with open(path, "w") as f:
            f.write(self._header("neurons"))
            for i in sorted(by_n):
                for c in sorted(by_n[i]):
                    f.write(f"{c}\t{i}\n")


def read_spikes(path):
    """(cycles, ids) arrays of a spike file of either ordering."""
    cyc, ids = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            c, i = line.split()
            cyc.append(int(float(c)))
            ids.append(int(i))
    return np.array(cyc, np.int64), np.array(ids, np.int64)
