"""Continuous-batching request loop over the batch engine (abx_batch_*).

Requests (transport solves: a Model and a number of batches) arrive over
time.  The loop admits every arrived request whose device footprint fits the
HBM budget next to the problems already on the GPU, advances all admitted
problems by one step in shared launches, retires the problems whose last
batch was recorded (their tallies are read back and their device memory
freed) and admits waiting requests as memory frees up.  A request's results
do not depend on what it was batched with: histories, RNG streams and
ordered banks are the problem's own.

The reference has no serving layer; its drivers run one problem per process
(src/fixed_source.cpp:81-173, src/modified_fixed_source.cpp:59-143).
"""
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import Batch, Problem


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


class Server:
    """`budget` caps the device bytes of admitted problems (0 = no cap).
    `step_histories` caps the histories of one step over all problems (0 =
    no cap): larger fixed-source batches are then split over several steps,
    so that a request arriving while a large one runs starts within one
    short step instead of after the large batch."""

    def __init__(self, budget=0, clock=time.perf_counter, keep_tallies=True, step_histories=0):
        self.batch = Batch(budget, step_histories)
        self.clock = clock
        self.keep_tallies = keep_tallies
        self.steps = 0
        self.step_seconds = 0.0     # wall time inside batch steps

    def run(self, requests):
        """Serve `requests` (any order; arrival times are honoured) and return
        them completed, with admission / completion times."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting = deque()
        running = {}
        t0 = self.clock()
        while pending or waiting or running:
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            while waiting:
                r = waiting[0]
                if not hasattr(r, "_problem"):
                    r._problem = Problem(r.model)
                if not self.batch.add(r._problem, r.nbatches):
                    if not running:
                        raise RuntimeError(f"request {r.name} does not fit the budget alone")
                    break
                r.admitted = self.clock() - t0
                running[r._problem.h.value] = r
                waiting.popleft()
            if running:
                ts = self.clock()
                done = self.batch.step()
                self.step_seconds += self.clock() - ts
                for p in done:
                    r = running.pop(p.h.value)
                    r.stats = p.stats()
                    if self.keep_tallies:
                        r.tallies = [p.tally_stats(t) for t in range(len(r.model.tallies))]
                    r.finished = self.clock() - t0
                    del r._problem          # frees the device tables
                self.steps += 1
            elif pending:
                time.sleep(max(0.0, pending[0].arrival - (self.clock() - t0)))
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
