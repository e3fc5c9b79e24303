"""The persistent tracer's work queue, checked on the CPU.

A model of the claim and refill state machine of the thin-disk kernel's
wavefront loop (grrt_amd/csrc/trace_thin.hip, trace_thin: lane 0 claims
`chunk` ray indices with one atomic add, one claim ahead; idle lanes are
refilled in lane order once `refill_min` are idle; a cancel flag read with
each claim stops further claims).  Wavefronts are interleaved by a seeded
random scheduler and every ray takes a random number of attempts.

Properties: every ray is traced exactly once; all wavefronts exit; after a
cancel the traced rays are exactly the prefix [0, min(queue_head, n)) of the
queue, since claimed rays always finish.
"""
import random

import pytest

LANES = 64


class Queue:
    def __init__(self):
        self.head = 0
        self.cancel = 0

    def claim(self, ch):
        got = self.head
        self.head += ch
        return got, self.cancel


class Wave:
    """One wavefront of trace_thin's loop, one loop iteration per step()."""

    def __init__(self, q, n, chunk, refill_min, lengths, done):
        self.q, self.n, self.ch, self.rmin = q, n, chunk, refill_min
        self.lengths, self.done = lengths, done
        self.left = [0] * LANES            # attempts left per lane (0: idle)
        self.ray = [-1] * LANES
        self.cur = self.cur_end = 0
        self.pre, self.cancel_seen = q.claim(chunk)
        self.have_pre = True
        self.queue_open = True
        self.exited = False

    def _start(self, lane, mine):
        self.ray[lane] = mine
        if self.lengths[mine] == 0:       # terminates at its start (termination test at init)
            self._finish(lane)
        else:
            self.left[lane] = self.lengths[mine]

    def _finish(self, lane):
        self.done.append(self.ray[lane])
        self.left[lane] = 0
        self.ray[lane] = -1

    def step(self):
        idle = [k for k in range(LANES) if self.left[k] == 0]
        n_idle = len(idle)
        if self.queue_open and (n_idle >= self.rmin or n_idle == LANES):
            avail = self.cur_end - self.cur if self.cur_end > self.cur else 0
            nb = None
            if avail < n_idle and self.have_pre:
                got, stop = self.pre, self.cancel_seen
                self.have_pre = False
                if got < self.n:
                    nb = got
                    if not stop:
                        self.pre, self.cancel_seen = self.q.claim(self.ch)
                        self.have_pre = True
            for k, lane in enumerate(idle):
                mine = None
                if k < avail:
                    mine = self.cur + k
                elif nb is not None:
                    mine = nb + (k - avail)
                if mine is not None and mine < self.n:
                    assert mine < self.cur_end or nb is not None
                    self._start(lane, mine)
            if avail >= n_idle:
                self.cur += n_idle
            elif nb is not None:
                self.cur_end = min(nb + self.ch, self.n)
                self.cur = min(nb + (n_idle - avail), self.cur_end)
            else:
                self.cur = self.cur_end
            if self.cur >= self.cur_end and not self.have_pre:
                self.queue_open = False
        if all(x == 0 for x in self.left):
            if not self.queue_open:
                self.exited = True
            return
        for lane in range(LANES):
            if self.left[lane] > 0:
                self.left[lane] -= 1
                if self.left[lane] == 0:
                    self._finish(lane)


def run(n, waves, chunk, refill_min, seed, cancel_at=None, max_len=12):
    rng = random.Random(seed)
    lengths = [rng.choice([0, 1, 1, 2, 3, rng.randint(1, max_len)]) for _ in range(n)]
    q = Queue()
    done = []
    ws = [Wave(q, n, chunk, refill_min, lengths, done) for _ in range(waves)]
    t = 0
    while not all(w.exited for w in ws):
        if cancel_at is not None and t == cancel_at:
            q.cancel = 1
        w = rng.choice([w for w in ws if not w.exited])
        w.step()
        t += 1
        assert t < 100 * (n + 10) * (max_len + 2), "the queue did not drain"
    return done, q.head


CASES = [(1, 1, 64, 8), (63, 1, 64, 1), (64, 2, 64, 64), (65, 3, 128, 8), (1000, 4, 64, 8),
         (4097, 8, 64, 16), (12345, 8, 192, 4), (777, 5, 100, 3)]


@pytest.mark.parametrize("n,waves,chunk,rmin", CASES)
def test_every_ray_traced_exactly_once(n, waves, chunk, rmin):
    for seed in range(3):
        done, head = run(n, waves, chunk, rmin, seed)
        assert sorted(done) == list(range(n))
        assert head >= n


@pytest.mark.parametrize("n,waves,chunk,rmin", CASES[3:])
def test_cancel_leaves_a_queue_prefix(n, waves, chunk, rmin):
    for seed in range(4):
        cancel_at = random.Random(seed + 100).randint(1, max(2, n // 8))
        done, head = run(n, waves, chunk, rmin, seed, cancel_at=cancel_at)
        assert len(done) == len(set(done))
        assert sorted(done) == list(range(min(head, n)))
        assert head % chunk == 0
