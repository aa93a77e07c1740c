"""CPU model of the one-owner message protocol of the bidiagonalization loop (ralha.hip
solo_publish / solo_receive): partial rows and message slots start as a sentinel; a slice owner
polls its slice of the G partial rows, stores the sum into slot j % 3, puts the sentinel back into
the partials it read and into its slice of slot (j - 1) % 3 (the last slice also into entry len);
every workgroup polls the whole message, then waits for its own stores before its next partial.

Workgroups advance one action at a time in a random order, and every store stays invisible to the
other workgroups for a random time unless its own workgroup waits for it (`vmcnt(0)`), as relaxed
stores to different addresses may land in any order. A workgroup reads its own pending stores.
Every value carries its step, workgroup and entry, so a stale or early read is caught; the run must
also finish (no poll waits forever)."""
import random

import pytest

SENT = "sentinel"


class Memory:
    """Global values plus each workgroup's pending stores. With hold_resets (the adversary of the
    reset ordering), a store of the sentinel lands only when its workgroup waits, or else at the
    worst moment: right after another workgroup's store to the same address becomes visible,
    overwriting it, as relaxed stores to one address from different workgroups may land in
    either order unless some wait orders them."""

    def __init__(self, hold_resets=False):
        self.glob = {}
        self.pending = {}  # wg -> list of (addr, value), in issue order
        self.hold = hold_resets

    def load(self, wg, addr):
        for a, v in reversed(self.pending.get(wg, [])):
            if a == addr:
                return v
        return self.glob.get(addr, SENT)

    def store(self, wg, addr, v):
        self.pending.setdefault(wg, []).append((addr, v))

    def _land(self, wg, addr, v):
        self.glob[addr] = v
        if self.hold and v != SENT:
            for w2, q in self.pending.items():
                if w2 == wg:
                    continue
                for k, (a2, v2) in enumerate(q):
                    if a2 == addr and v2 == SENT:
                        self.glob[addr] = SENT
                        del q[k]
                        break

    def drain(self, wg):  # s_waitcnt vmcnt(0)
        q, self.pending[wg] = self.pending.get(wg, []), []
        for a, v in q:
            self._land(wg, a, v)

    def leak(self, rng):
        """Some pending store of some workgroup becomes visible (any address order)."""
        cands = [(w, k) for w, q in self.pending.items() for k, (_, v) in enumerate(q)
                 if not (self.hold and v == SENT)]
        if not cands:
            return
        w, k = rng.choice(cands)
        q = self.pending[w]
        a, v = q[k]
        # a later store of the same wavefront to the same address stays ordered after this one
        if any(a2 == a for a2, _ in q[:k]):
            return
        del q[k]
        self._land(w, a, v)


def workgroup(w, G, m, mem, log, wait_after_receive=True, wait_before_slice=True):
    """Generator: one workgroup's actions; yields whenever it waits or after each memory action."""
    def slices(j):
        length = m + 3 - j
        valid = (m - 1 - j) + (4 if j > 0 else 0)
        per = -(-length // G)
        r0 = w * per
        return length, valid, r0, min(length, r0 + per)

    last = m - 3
    for j in range(0, last + 1):
        length, valid, r0, r1 = slices(j)
        # reg_share: partial row of step j (entries < valid)
        for e in range(valid):
            mem.store(w, ("P", j & 1, w, e), ("part", j, w, e))
            yield
        # solo_publish, in blocks of 16 entries as the kernel: poll the block's partials, wait for
        # this workgroup's stores, store the block of the slot, reset the partials read. A
        # workgroup whose slice is empty runs no block, and so no wait.
        for base in range(r0, r1, 16):
            blk = range(base, min(r1, base + 16))
            for e in blk:
                for v in range(G):
                    if e >= valid:
                        continue
                    while True:
                        x = mem.load(w, ("P", j & 1, v, e))
                        if x != SENT:
                            break
                        yield
                    assert x == ("part", j, v, e), ("stale partial", w, j, v, e, x)
            if wait_before_slice:
                mem.drain(w)  # stores_done before the block is stored
            for e in blk:
                mem.store(w, ("M", j % 3, e), ("msg", j, e))
                yield
            for e in blk:
                if e < valid:
                    for v in range(G):
                        mem.store(w, ("P", j & 1, v, e), SENT)
                        yield
        if j >= 1 and r0 < r1:
            p1 = length + 1 if r1 == length else r1
            for e in range(r0, p1):
                mem.store(w, ("M", (j - 1) % 3, e), SENT)
                yield
        # solo_receive: the whole message j
        for e in range(length):
            while True:
                x = mem.load(w, ("M", j % 3, e))
                if x != SENT:
                    break
                yield
            assert x == ("msg", j, e), ("stale message", w, j, e, x)
        if wait_after_receive:
            mem.drain(w)  # stores_done at the end of the receive
        log.append((w, j))


def run_model(G, m, seed, hold_resets=False, **flags):
    rng = random.Random(seed)
    mem = Memory(hold_resets)
    log = []
    wgs = {w: workgroup(w, G, m, mem, log, **flags) for w in range(G)}
    steps = 0
    while wgs:
        steps += 1
        if steps > 100_000:
            raise AssertionError("stalled: a poll never ends")
        if rng.random() < 0.35:
            mem.leak(rng)
            continue
        w = rng.choice(list(wgs))
        try:
            next(wgs[w])
        except StopIteration:
            del wgs[w]
    return log


@pytest.mark.parametrize("hold_resets", [False, True])
@pytest.mark.parametrize("G,m", [(1, 6), (2, 7), (3, 9), (4, 12), (5, 8), (7, 15)])
def test_solo_protocol_random_schedules(G, m, hold_resets):
    for seed in range(12):
        log = run_model(G, m, seed * 1000 + G * 31 + m, hold_resets)
        assert len(log) == G * (m - 2)


def test_solo_protocol_model_catches_missing_waits():
    """The model is sensitive: with neither wait, some schedule lets a reset land after the next
    write of the same location, and a poll then stalls or reads stale data."""
    failures = 0
    for seed in range(12):
        try:
            run_model(7, 15, seed, hold_resets=True, wait_after_receive=False, wait_before_slice=False)
        except AssertionError:
            failures += 1
    assert failures > 0


@pytest.mark.parametrize("G,m", [(3, 9), (7, 15), (5, 30)])
def test_solo_protocol_holds_without_the_receive_wait(G, m):
    """The wait at the end of each receive is a margin, not a necessity: a workgroup whose slice
    at step j holds an entry below len_j - 2 (the entries written again two steps later) owns a
    non-empty slice at step j + 1 too, since the slices shrink with len, and that block's wait
    orders its resets before any later write of those locations."""
    for seed in range(12):
        log = run_model(G, m, seed, hold_resets=True, wait_after_receive=False)
        assert len(log) == G * (m - 2)
