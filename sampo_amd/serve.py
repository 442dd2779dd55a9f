"""Serving loop: scenario runs that arrive over time share one engine.

Continuous batching at day granularity:
  1. requests whose arrival time has passed join a FIFO queue; their inputs
     are generated on a thread pool as they arrive;
  2. from the head of the queue, requests are admitted into free engine
     slots while their device footprint (sampo_engine_footprint) fits in
     the free device memory less a reserve -- admission by HBM footprint;
  3. running instances advance one day (one HIP graph of 24 iterations of
     the batched step), whatever their own iteration counts: each instance
     keeps its own step, horizon, seeds and inputs, so its results are those
     of a run alone.  With policy 'srpt' (the default) only the instances
     with the least work left (agents now x iterations left) advance, as
     many as make up min_active agents; the others are held
     (sampo_engine_hold) for that day.  Small requests then finish first
     instead of all requests sharing every day; with policy 'all' every
     running instance advances every day;
  4. instances that outgrew their buffers are grown (sampo_engine_service);
     finished ones are read out and their slots and memory released.
A request's completion latency is the time from its arrival to the end of
the day in which it finished.

The clock is either wall time (seconds; the serving benchmark) or a count of
loop rounds (deterministic tests: arrival = the round a request arrives in).
"""
import hashlib
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

import sampo_amd as S
from sampo_amd import workloads as W
from sampo_amd.batch import NO_CHUNK_CAP, Tuning

FINISHED = (S.DONE, S.CAPACITY, S.EXTINCT, S.PHYS_FULL)
REJECTED = -2      # result status: could not be admitted
GROWTH_FAILED = -3  # result status: its buffers could not grow for max_deferred days


@dataclass
class Result:
    index: int
    n0: int
    days: int
    arrival: float
    admitted: float = -1.0
    done: float = -1.0
    status: int = -1
    steps: int = 0
    agent_steps: int = 0
    digest: str = ""
    rows: object = None

    @property
    def latency(self):
        return self.done - self.arrival


def rows_digest(rows):
    return hashlib.sha256(np.ascontiguousarray(rows).tobytes()).hexdigest()


class Server:
    def __init__(self, max_slots=1024, reserve_bytes=4 << 30, headroom=1.25, phys_factor=1.25, tuning=None,
                 keep_rows=False, threads=16, log=None, policy="srpt", min_active=64_000_000, max_deferred=16):
        if policy not in ("srpt", "all"):
            raise ValueError(f"unknown policy {policy!r}")
        self.policy = policy
        self.min_active = min_active
        self.max_slots = max_slots
        self.reserve = reserve_bytes
        self.headroom = headroom
        self.phys_factor = phys_factor
        self.tuning = tuning or Tuning()
        self.keep_rows = keep_rows
        self.threads = threads
        self.log = log
        self.max_deferred = max_deferred  # days an instance may wait for memory to grow

    def run(self, specs, clock="wall"):
        """Serve `specs` (arrival times in seconds, or in rounds with
        clock='rounds'); returns (results in spec order, summary)."""
        order = sorted(range(len(specs)), key=lambda i: specs[i].arrival)
        results = [Result(i, specs[i].n0, specs[i].days, specs[i].arrival) for i in range(len(specs))]
        queue = []          # (index, future) in arrival order
        running = {}        # slot -> index
        fresh = set()       # slots admitted since the last poll
        deferred = {}       # slot -> consecutive days paused for a growth that failed
        last = None         # (status, step, agents) of that poll
        nxt = 0
        rounds = 0
        days_stepped = 0
        total_steps = 0
        t_start = time.perf_counter()

        def now():
            return time.perf_counter() - t_start if clock == "wall" else float(rounds)

        with S.Engine(max_slots=self.max_slots) as e, ThreadPoolExecutor(max_workers=self.threads) as pool:
            while True:
                t = now()
                while nxt < len(order) and specs[order[nxt]].arrival <= t:
                    i = order[nxt]
                    queue.append((i, pool.submit(W.materialise, specs[i], self.phys_factor)))
                    nxt += 1
                # admission by footprint, FIFO
                free = None
                while queue and len(running) < self.max_slots:
                    i, fut = queue[0]
                    if clock == "wall" and not fut.done():
                        break
                    kw = fut.result()
                    need = S.footprint(chunk=self.tuning.chunk_for(specs[i].n0), **kw)
                    if free is None:
                        free, _ = S.device_memory()
                    if need + self.reserve > free:
                        if not running:  # cannot fit even with the device to itself: reject
                            queue.pop(0)
                            results[i].status = REJECTED
                            results[i].done = now()
                            if self.log:
                                self.log(f"serve: request {i} needs {need / 2**30:.1f} GiB, rejected")
                            continue
                        break
                    e.set_tile(*self.tuning.tile_for(specs[i].n0))
                    e.set_chunk(self.tuning.chunk_for(specs[i].n0), NO_CHUNK_CAP)
                    try:
                        slot = e.admit(**kw)
                    except S.SampoError as err:  # the device could not take it after all
                        if running:
                            break  # retry when something finishes
                        queue.pop(0)
                        results[i].status, results[i].done = REJECTED, now()
                        if self.log:
                            self.log(f"serve: request {i} rejected ({err})")
                        continue
                    free -= need
                    queue.pop(0)
                    running[slot] = i
                    fresh.add(slot)
                    results[i].admitted = now()
                if not running:
                    if nxt >= len(order) and not queue:
                        break
                    if clock == "wall":  # idle until the next arrival
                        if not queue and nxt < len(order):
                            time.sleep(max(0.0, min(0.05, specs[order[nxt]].arrival - now())))
                        else:
                            time.sleep(0.001)
                    else:
                        rounds += 1
                    continue
                if self.policy == "srpt":
                    self._choose(e, running, specs, last, fresh)
                e.step_days(1)
                days_stepped += 1
                try:
                    e.service(self.headroom)
                except S.SampoError as err:  # out of memory for a growth: stays paused until memory frees
                    if self.log:
                        self.log(f"serve: growth deferred ({err})")
                rounds += 1
                st, step, na = e.poll(max(running) + 1)  # slots past the last running one are free
                last, fresh = (st, step, na), set()
                t = now()
                for slot in list(running):  # a growth that keeps failing ends the request
                    deferred[slot] = deferred.get(slot, 0) + 1 if st[slot] == S.GROW else 0
                    if deferred[slot] > self.max_deferred:
                        i = running.pop(slot)
                        del deferred[slot]
                        results[i].done, results[i].status, results[i].steps = t, GROWTH_FAILED, int(step[slot])
                        e.release(slot)
                        if self.log:
                            self.log(f"serve: request {i} could not grow for {self.max_deferred} days, ended")
                for slot in [s for s in running if st[s] in FINISHED]:
                    i = running.pop(slot)
                    r = results[i]
                    rows = e.rows(slot, 0, int(step[slot]) + 1)
                    r.done, r.status, r.steps = t, int(st[slot]), int(step[slot])
                    r.agent_steps = int(rows["count"][:r.steps].astype(np.uint64).sum())
                    r.digest = rows_digest(rows)
                    if self.keep_rows:
                        r.rows = rows
                    total_steps += r.agent_steps
                    deferred.pop(slot, None)
                    e.release(slot)
                if self.log and days_stepped % 20 == 0:
                    self.log(f"serve: t={t:.1f} running {len(running)} queued {len(queue)} "
                             f"arrived {nxt}/{len(order)} done {sum(1 for x in results if x.done >= 0)}")
        wall = time.perf_counter() - t_start
        lat = np.array([r.latency for r in results if r.done >= 0 and r.status >= 0])
        summary = {"requests": len(specs), "completed": int(len(lat)),
                   "rejected": sum(1 for r in results if r.status == REJECTED),
                   "growth_failed": sum(1 for r in results if r.status == GROWTH_FAILED),
                   "wall_s": wall, "days_stepped": days_stepped,
                   "agent_steps": total_steps, "agent_steps_per_s": total_steps / wall if wall else 0.0}
        if len(lat):
            summary.update(p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)),
                           mean=float(lat.mean()), max=float(lat.max()))
        return results, summary


    def _choose(self, e, running, specs, last, fresh):
        """Hold all but the instances with the least work left (agents now
        x iterations left), keeping at least min_active agents in the step.
        The state is the poll after the last step; instances admitted since
        are at iteration 0 with their initial population."""
        cur = {}
        for s, i in running.items():
            if s in fresh or last is None or s >= last[0].shape[0]:
                cur[s] = (S.RUNNING, 0, specs[i].n0)
            else:
                cur[s] = (int(last[0][s]), int(last[1][s]), int(last[2][s]))
        slots = [s for s in running if cur[s][0] in (S.RUNNING, S.HELD)]
        left = {s: cur[s][2] * max(specs[running[s]].days * 24 - cur[s][1], 1) for s in slots}
        na = {s: cur[s][2] for s in slots}
        mask = np.zeros(max(running) + 1, bool)
        agents = 0
        for k, s in enumerate(sorted(slots, key=lambda s: (left[s], s))):
            if k and agents >= self.min_active:
                mask[s] = True
            else:
                agents += int(na[s])
        e.hold(mask)


def run_alone(spec, phys_factor=0.0, tuning=None):
    """One request alone in its own engine, to its horizon: its rows."""
    tuning = tuning or Tuning()
    kw = W.materialise(spec, phys_factor)
    with S.Engine(max_slots=1) as e:
        e.set_tile(*tuning.tile_for(spec.n0))
        e.set_chunk(tuning.chunk_for(spec.n0), NO_CHUNK_CAP)
        slot = e.admit(**kw)
        while True:
            e.step_days(1)
            e.service(1.25)
            st, step, _ = e.status(slot)
            if st in FINISHED:
                return e.rows(slot, 0, step + 1)
