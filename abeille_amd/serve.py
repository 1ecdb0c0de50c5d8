"""Continuous-batching request loop over the batch engine (abx_batch_*).

Requests (transport solves: a Model and a number of batches) arrive over
time.  The loop admits every arrived request whose device footprint fits the
HBM budget next to the problems already on the GPU, advances all admitted
problems step by step in shared launches, retires the problems whose last
batch was recorded (their tallies are read back and their device memory
freed) and admits waiting requests as memory frees up.  A request's results
do not depend on what it was batched with: histories, RNG streams and
ordered banks are the problem's own.

Requests run in lanes, each a batch stepped by a thread of its own on its
own streams, so that the lanes' launches run side by side on the GPU.
Modified-fixed-source generation chains have a lane of their own, so a
chain's generations (and the ordering, cancellation and re-banking between
them) never hold up fixed-source steps; fixed-source requests go to the
fixed-source lane with the least outstanding work, so a request arriving
while a long step runs can start in another lane at once.  One budget
covers all lanes.

The reference has no serving layer; its drivers run one problem per process
(src/fixed_source.cpp:81-173, src/modified_fixed_source.cpp:59-143).
"""
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import Batch, Problem, abi


@dataclass
class Request:
    model: object
    nbatches: int
    arrival: float                 # seconds after the start of the stream
    name: str = ""
    admitted: float = None
    finished: float = None
    stats: object = None
    tallies: list = field(default_factory=list)   # [(avg, std_err)] per tally

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class _Lane:
    """A batch and the thread that steps it."""

    def __init__(self, server, step_histories):
        self.server = server
        self.batch = Batch(0, step_histories)
        self.inbox = deque()       # (request, problem, bytes) to attach
        self.running = {}
        self.outstanding = 0       # histories attached or waiting (for dispatch)
        self.steps = 0
        self.step_seconds = 0.0
        self.error = None
        self.thread = threading.Thread(target=self._loop, daemon=True)

    def _loop(self):
        srv = self.server
        try:
            while True:
                with srv.cv:
                    # an idle lane waits while arrived requests are still
                    # being staged, so that its first step takes them all
                    while not srv.closing and not self.running and (
                            not self.inbox or srv.backlog):
                        srv.cv.wait()
                    if srv.closing and not self.inbox and not self.running:
                        return
                    todo = list(self.inbox)
                    self.inbox.clear()
                for r, p, need in todo:
                    self.batch.add(p, r.nbatches)
                    self.running[p.h.value] = (r, need)
                del todo
                ts = srv.clock()
                done = self.batch.step()
                self.step_seconds += srv.clock() - ts
                self.steps += 1
                out = []
                for p in done:
                    r, need = self.running.pop(p.h.value)
                    r.stats = p.stats()
                    if srv.keep_tallies:
                        r.tallies = [p.tally_stats(t) for t in range(len(r.model.tallies))]
                    r.finished = srv.clock() - srv.t0
                    out.append((r, need))
                # The problems are freed by the admitting thread (freeing
                # device memory waits for the whole device, which would stall
                # this lane): drop this thread's references first, so that
                # the request's own reference is the last.
                del done
                p = None
                if out:
                    with srv.cv:
                        for r, _ in out:
                            self.outstanding -= r._work
                        srv.retired.extend(out)
                        srv.cv.notify_all()
        except BaseException as e:       # surfaced by Server.run
            with srv.cv:
                self.error = e
                srv.cv.notify_all()


class Server:
    """`budget` caps the device bytes of admitted problems over all lanes
    (0 = no cap).  `step_histories` caps the histories of one step of a lane
    (0 = no cap): larger fixed-source batches are then split over several
    steps.  `lanes` is the number of lanes: one runs every request in one
    batch; with more, the last runs the generation chains and the others
    the fixed-source requests."""

    def __init__(self, budget=0, clock=time.perf_counter, keep_tallies=True, step_histories=0,
                 lanes=3):
        self.budget = budget
        self.clock = clock
        self.keep_tallies = keep_tallies
        self.cv = threading.Condition()
        self.used = 0
        self.retired = []           # finished requests whose problems are still allocated
        self.backlog = False        # arrived requests are being staged
        self.free_seconds = 0.0     # time spent freeing finished problems
        self.free_max = 0.0
        self.closing = False
        self.t0 = 0.0
        self.lanes = [_Lane(self, step_histories) for _ in range(max(1, min(8, lanes)))]

    @property
    def steps(self):
        return sum(l.steps for l in self.lanes)

    @property
    def step_seconds(self):
        return sum(l.step_seconds for l in self.lanes)

    def kernel_seconds(self):
        return sum(l.batch.kernel_seconds() for l in self.lanes)

    def _lane_for(self, model):
        if len(self.lanes) == 1:
            return self.lanes[0]
        if model.settings["mode"] == abi.MODIFIED_FIXED_SOURCE:
            return self.lanes[-1]
        return min(self.lanes[:-1], key=lambda l: l.outstanding)

    def _check(self):
        for l in self.lanes:
            if l.error is not None:
                raise l.error

    def _set_backlog(self, on):
        with self.cv:
            if self.backlog != on:
                self.backlog = on
                self.cv.notify_all()

    def _reap(self):
        """Free the device memory of finished requests; returns with cv not
        held (freeing waits for the device)."""
        with self.cv:
            gone = self.retired
            self.retired = []
        t = self.clock()
        for r, _ in gone:
            del r._problem
        if gone:
            dt = self.clock() - t
            self.free_seconds += dt
            self.free_max = max(self.free_max, dt)
        with self.cv:
            for _, need in gone:
                self.used -= need
            if gone:
                self.cv.notify_all()

    def run(self, requests):
        """Serve `requests` (any order; arrival times are honoured) and return
        them completed, with admission / completion times."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        self.closing = False
        self.t0 = self.clock()
        for l in self.lanes:
            if not l.thread.is_alive():
                l.thread = threading.Thread(target=l._loop, daemon=True)
                l.thread.start()
        try:
            while pending:
                r = pending[0]
                self._reap()
                wait = r.arrival - (self.clock() - self.t0)
                if wait > 0:
                    self._set_backlog(False)
                    time.sleep(wait)
                self._set_backlog(True)
                pending.popleft()
                r._problem = p = Problem(r.model)   # staging the tables is part of serving
                lane = self._lane_for(r.model)
                need = lane.batch.need(p, r.nbatches)
                if self.budget and need > self.budget:
                    raise RuntimeError(f"request {r.name} does not fit the budget alone")
                while True:
                    self._reap()
                    with self.cv:
                        self._check()
                        if not self.budget or self.used + need <= self.budget:
                            break
                        self.backlog = False       # lanes with admitted work go on
                        self.cv.notify_all()
                        if not self.retired:
                            self.cv.wait(0.05)
                with self.cv:
                    self.used += need
                    r.admitted = self.clock() - self.t0
                    r._work = int(r.model.settings["nparticles"]) * r.nbatches
                    lane.outstanding += r._work
                    lane.inbox.append((r, p, need))
                    self.cv.notify_all()
        finally:
            with self.cv:
                self.backlog = False
                self.closing = True
                self.cv.notify_all()
            for l in self.lanes:
                l.thread.join()
            self._reap()
        self._check()
        return requests


def latency_percentiles(requests, q=(50, 99)):
    lat = np.array([r.latency for r in requests])
    return {f"p{k}": float(np.percentile(lat, k)) for k in q}


def serial_latencies(arrivals, service):
    """Latencies of the same requests served one at a time, first come first
    served, from measured service times."""
    order = np.argsort(arrivals)
    t = 0.0
    lat = np.zeros(len(arrivals))
    for i in order:
        t = max(t, arrivals[i]) + service[i]
        lat[i] = t - arrivals[i]
    return lat
