"""The grid-wide exchange of the Gram-Schmidt sweep (k_mgs_sweep) and of the
persistent coarsest PCG (k_pcg_persistent), as a model explored on the CPU.

Each of G blocks runs rounds s = 0, 1, ...: wait until its own earlier stores
are acknowledged, publish its partial of round s into slot (s mod 3) of a ring
whose empty entries hold a sentinel, poll the G entries of that slot (one read
at a time, re-reading all of them while any is empty), keep what it read, then
empty its own entry of round s-1.  A store lands at some later moment chosen by
the scheduler, and stores of one block may land out of order; a read sees what
has landed.  Random interleavings must leave every block with exactly the
values published for its round.  Two variants with a planted mistake (no
acknowledgement wait before publishing, a ring of two slots) must be caught,
which shows the search reaches the interleavings that matter.
"""

import random

import pytest

EMPTY = object()


def _run(G, R, seed, slots=3, ack=True, max_steps=200000):
    rng = random.Random(seed)
    ring = [[EMPTY] * G for _ in range(slots)]
    pending = [[] for _ in range(G)]  # stores issued, not yet landed
    # per block: round, phase, poll position, values read in this poll
    st = [dict(s=0, ph="ack", q=0, got=[]) for _ in range(G)]
    seen = [[None] * R for _ in range(G)]

    def value(s, b):
        return (s, b)

    for _ in range(max_steps):
        acts = []
        for b in range(G):
            if pending[b]:
                acts.append(("land", b))
            if st[b]["s"] < R:
                if st[b]["ph"] != "ack" or not ack or not pending[b]:
                    acts.append(("step", b))
        if not acts:
            return seen
        kind, b = rng.choice(acts)
        if kind == "land":
            i = rng.randrange(len(pending[b]))
            (r, q), v = pending[b].pop(i)
            ring[r][q] = v
            continue
        t = st[b]
        r = t["s"] % slots
        if t["ph"] == "ack":  # enabled only once pending[b] is empty (with ack)
            pending[b].append(((r, b), value(t["s"], b)))
            t.update(ph="poll", q=0, got=[])
        elif t["ph"] == "poll":
            v = ring[r][t["q"]]
            t["got"].append(v)
            t["q"] += 1
            if t["q"] == G:
                if any(x is EMPTY for x in t["got"]):
                    t.update(q=0, got=[])
                else:
                    seen[b][t["s"]] = list(t["got"])
                    pending[b].append((((t["s"] + slots - 1) % slots, b), EMPTY))
                    t.update(s=t["s"] + 1, ph="ack")
    return None  # no progress within max_steps: a lost value leaves blocks polling


def _ok(seen, G, R):
    return seen is not None and all(
        seen[b][s] == [(s, q) for q in range(G)] for b in range(G) for s in range(R))


def _stale(seen, G, R):
    return seen is not None and not _ok(seen, G, R)


@pytest.mark.parametrize("G", [1, 2, 3, 4])
def test_ring_exchange_reads_only_its_round(G):
    R = 7
    for seed in range(1500):
        seen = _run(G, R, seed)
        assert _ok(seen, G, R), (G, seed)


def test_ring_without_ack_wait_is_caught():
    # caught as a stale read (a reader of round s+2 sees the value of s-1)
    assert any(_stale(_run(2, 6, seed, ack=False), 2, 6) for seed in range(3000))


def test_two_slot_ring_is_caught():
    # a block that runs one round ahead reads a slot its neighbour still holds
    assert any(_stale(_run(2, 6, seed, slots=2), 2, 6) for seed in range(3000))
