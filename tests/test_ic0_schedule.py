"""The wait protocol of the device IC(0) sweeps (lfg_ic0.hip, ic0_node_sweep), emulated on
the CPU over the schedules the library builds (lfh_ic0_node_schedule, the code the device
set-up uses).

On the device, four wavefronts take the parts of a sweep in rotation (part l: wavefront
l % 4); the one that owns part l waits until an LDS counter of finished positions reaches the
part's first position, computes its nodes from values in an LDS window of 2,048 node
positions (slot = position % 2048), writes them into the window and sets the counter to the
part's end. Every wait is bounded: a wait that never ends stops the sweep with a fault flag.

The emulation runs that protocol step by step and checks what keeps it progressing and
correct: at every step some wavefront can proceed (otherwise every wavefront polls until its
bound: the device's wait-bound stop), the counter never moves back, every dependency a part
reads is finished and still in its window slot, and the sweep ends with the counter at the
last position. Planted schedule mistakes show which faults each check catches, among them
waiting on part boundaries in a padded numbering while the counter advances in the compact
one: a wait that never ends, which on the device is a wait-bound stop.
"""
import random

import numpy as np
import pytest

from lightfem_amd import meshgen, native, pipeline

WAVES = 4
WINDOW = 2048  # kIcNodeWin


class Stall(AssertionError):
    """No wavefront can proceed: on the device every wavefront polls until its wait bound."""


def emulate(sched, n_nodes, seed=0, window=WINDOW, wait_begin=None):
    """Runs the protocol over `sched`; wait_begin (default: the part boundaries) gives the
    position each part waits for. Returns the number of steps."""
    parts, qptr, key = sched["parts"], sched["qptr"], sched["key"]
    n_parts = len(parts) - 1
    begin = parts[:-1] if wait_begin is None else wait_begin
    owner = np.full(window, -1, dtype=np.int64)
    done_pos = 0
    nxt = list(range(WAVES))  # each wavefront's next part: w, w + 4, ...
    rng = random.Random(seed)
    steps = 0
    while any(p < n_parts for p in nxt):
        runnable = [w for w in range(WAVES) if nxt[w] < n_parts and begin[nxt[w]] <= done_pos]
        if not runnable:
            raise Stall(f"counter at {done_pos}, waiting for {[begin[p] for p in nxt if p < n_parts]}")
        w = rng.choice(runnable)
        part = nxt[w]
        b, e = int(parts[part]), int(parts[part + 1])
        for q in range(b, e):
            for k in key[qptr[q]:qptr[q + 1]]:
                assert k < b, f"position {q} reads {k}, not finished before its part starts at {b}"
                assert owner[k % window] == k, f"position {q} reads {k}, overwritten in the window"
        for q in range(b, e):
            owner[q % window] = q
        assert e >= done_pos, "the counter moved back"
        done_pos = e
        nxt[w] += WAVES
        steps += 1
    assert done_pos == n_nodes
    return steps


def schedules(model, reordering):
    ix = pipeline.index_model(model, reordering)
    row_ptr, col, _ = native.csr_pattern(ix.n_nodes, ix.enodes)
    out = []
    for backward in (False, True):
        s = native.ic0_node_schedule(row_ptr, col, backward)
        assert s is not None  # Q4 factors have the node form
        out.append(s)
    return ix.n_nodes, out


CASES = {
    "6x6": lambda: (meshgen.structured_cantilever(6, 6), "none"),
    "24x10_rcm": lambda: (meshgen.structured_cantilever(24, 10), "rcm"),
    "200x200_rcm": lambda: (meshgen.cantilever_200_rcm(), "rcm"),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_sweep_schedule_progresses(case):
    model, reordering = CASES[case]()
    n_nodes, (fwd, bwd) = schedules(model, reordering)
    for s in (fwd, bwd):
        parts = s["parts"]
        # parts tile the positions in order, at most one wavefront (64 nodes) each
        assert parts[0] == 0 and parts[-1] == n_nodes and np.all(np.diff(parts) > 0)
        assert np.max(np.diff(parts)) <= 64
        # the window condition under which the device takes this form (ic0_nodes_fit)
        assert s["reach"] + 2 * 64 <= WINDOW
        q = np.repeat(np.arange(n_nodes), np.diff(s["qptr"]))
        assert s["reach"] == (int(np.max(q - s["key"])) if len(q) else 0)
        assert np.array_equal(np.sort(s["ord"]), np.arange(n_nodes))  # every node once
        for seed in (0, 1):
            emulate(s, n_nodes, seed)


def test_emulation_catches_planted_schedule_faults():
    n_nodes, (fwd, _) = schedules(meshgen.structured_cantilever(6, 6), "none")
    parts = fwd["parts"]
    # waiting on boundaries padded to 64 positions per part (a slot-major copy's
    # numbering) while the counter advances to the compact ends: some wait never ends
    padded = np.arange(len(parts) - 1) * 64
    with pytest.raises(Stall):
        emulate(fwd, n_nodes, wait_begin=padded)
    # a part boundary moved one position back: its first node reads a dependency in its
    # own part, before that dependency is finished
    bad = dict(fwd)
    p = next(i for i in range(1, len(parts) - 1)
             if any(k == parts[i] - 1 for q in range(parts[i], parts[i + 1])
                    for k in fwd["key"][fwd["qptr"][q]:fwd["qptr"][q + 1]]))
    bad["parts"] = parts.copy()
    bad["parts"][p] -= 1
    with pytest.raises(AssertionError, match="not finished"):
        emulate(bad, n_nodes)
    # a window far shorter than the reach: a value is overwritten before its last reader
    with pytest.raises(AssertionError, match="overwritten"):
        emulate(fwd, n_nodes, window=1)
