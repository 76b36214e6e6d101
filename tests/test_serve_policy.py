"""Admission order of the serving loop (host logic only, no GPU): which
pending requests form the next unit under arrival order, pairs-first and a
cap on concurrently running sequence units."""
from types import SimpleNamespace

from singleventriclesegmentation_amd.serve import FlowServer, Request


def server(pending, running_kinds=(), sequence_slots=None, pairs_first=False):
    s = FlowServer.__new__(FlowServer)   # no device state needed for the policy
    s.pending = list(pending)
    s.running = [SimpleNamespace(requests=[SimpleNamespace(kind=k)]) for k in running_kinds]
    s.pair_buckets = (1, 2, 4)
    s.sequence_slots = sequence_slots
    s.pairs_first = pairs_first
    return s


def pair(rid, shape=(8, 8, 8)):
    return Request(rid, "pair", shape, 0.0, (None, None))


def seq(rid, shape=(8, 8, 8)):
    return Request(rid, "sequence", shape, 0.0, [None] * 5)


def test_arrival_order():
    s = server([seq(0), pair(1), pair(2)])
    reqs, shape, count = s._next_unit()
    assert [r.rid for r in reqs] == [0] and count == 8
    s = server([pair(1), seq(0), pair(2, (4, 4, 4)), pair(3)])
    reqs, shape, count = s._next_unit()
    assert [r.rid for r in reqs] == [1, 3] and count == 2 and shape == (8, 8, 8)


def test_pairs_first():
    s = server([seq(0), pair(1, (4, 4, 4)), pair(2)], pairs_first=True)
    reqs, shape, count = s._next_unit()
    assert [r.rid for r in reqs] == [1] and shape == (4, 4, 4) and count == 1


def test_sequence_cap():
    s = server([seq(0), pair(1)], running_kinds=("sequence", "sequence"), sequence_slots=2)
    reqs, _, _ = s._next_unit()
    assert [r.rid for r in reqs] == [1]          # the pair goes past the capped sequence
    s = server([seq(0)], running_kinds=("sequence", "sequence"), sequence_slots=2)
    assert s._next_unit() is None
    s = server([seq(0)], running_kinds=("sequence", "pair"), sequence_slots=2)
    assert [r.rid for r in s._next_unit()[0]] == [0]
    assert server([]) ._next_unit() is None


def test_cancel_pending_and_running():
    """A pending request leaves the queue; an admitted one is dropped when its
    unit retires while the request batched with it completes; completed,
    withdrawn and unknown requests cannot be cancelled."""
    import torch
    from collections import deque
    r1, r2, r3, r4 = pair(1), pair(2), seq(3), pair(4)
    s = server([r2, r3])
    s.cancelled, s.completed = [], []
    assert s.cancel(r2, now=1.5) and r2.cancelled and r2.done == 1.5
    assert [r.rid for r in s.pending] == [3] and s.cancelled == [r2]
    assert not s.cancel(r2)                      # already withdrawn
    assert not s.cancel(pair(9))                 # unknown
    out = torch.arange(2 * 2 * 2 * 2 * 3, dtype=torch.float32).reshape(2, 2, 2, 2, 3)
    unit = SimpleNamespace(requests=[r1, r4], event=SimpleNamespace(query=lambda: True), outputs=out, nbytes=7,
                           key=("k",), plan="plan", slot=0)
    s.running = [unit]
    s.used, s.idle_plans, s.free_slots = 7, [], deque()
    assert s.cancel(r1) and r1.done is None      # running: still on the device
    s._retire(now=2.0)
    assert s.completed == [r4] and torch.equal(r4.flows, out[1])
    assert r1.flows is None and r1.done == 2.0 and s.cancelled == [r2, r1]
    assert s.used == 0 and list(s.free_slots) == [0] and s.idle_plans == [(("k",), "plan")]
    assert not s.cancel(r4)                      # completed
