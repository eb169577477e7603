"""Request loop: range workunits served by continuous batching on one GPU.

The reference processes one workunit per process (BOINC hands it a range;
main.cpp:313-405 runs cl_wilson() on it and exits). Here ranges arrive as a
stream and are served together:

* Admission. Queued requests are admitted in arrival order while the HBM
  budget holds. An admission round (on its own thread and stream) splits
  the requests it takes by the magnitude of their largest factorial target
  (classes a factor 16 wide) and
  turns each class into a cohort: a clw_batch whose accumulators stay
  resident in HBM until its last range is done. Ranges of one cohort need
  about the same sieving primes, so they share every window to the end. A
  request that does not fit waits for memory to free.
* Continuous batching. Cohorts admitted at different times run side by side
  on `streams` HIP streams (one device context each): the event loop keeps a
  window of sieving primes in flight on every stream, so the kernels of
  different cohorts overlap on the GPU, and starts the next window on a
  stream as soon as its last one is done (no round barrier). Stream 0 serves
  the cohorts large enough to fill the GPU by themselves; the others serve
  the smaller ones, whose kernels fill the gaps beside them.
* Scheduling. Shortest work to the next completion first: the cohorts whose
  next range needs the fewest factor applications to finish go first, and a
  window stops at that range's largest factorial target, so a range finishes
  (clw_batch_finish_launch: iterate + check of its primes alone, queued on
  the cohort's stream) as soon as its sieving primes are in, while larger
  ranges of the same cohort go on. Its results are collected when they are
  in (clw_batch_finish_query), so the loop never waits for a finish while
  other streams run out of work.
* Latency is completion time minus arrival time; each request's results are
  those of a single-range run, bit for bit.
"""
from __future__ import annotations

import json
import math
import os
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import Batch, Device, RangeResult, Scratch, Workspace, host_primes
from .workloads import Request, pi_approx

BLOCK_SPAN = 61440


@dataclass
class Completion:
    rid: int
    arrival: float
    admitted: float
    finished: float
    result: RangeResult

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Cohort:
    batch: Batch
    slot: int
    reqs: list[Request]
    maxt: list[int]                      # largest factorial target per range
    T: np.ndarray                        # group targets (one group per range and type)
    n: np.ndarray                        # test primes per group
    admitted: float
    done: set = field(default_factory=set)        # ranges finished and collected
    finishing: set = field(default_factory=set)   # ranges whose finish is in flight
    bytes: int = 0
    rate: float = 0.0                    # factor applications/s it has been running at
    window: tuple = ()                   # (lo, hi, launch time, counter) of the window in flight

    def __post_init__(self):
        self.piT = np.array([pi_approx(float(t)) for t in self.T])
        self.work = self.work_to_next()  # refreshed whenever the cohort moves

    def covered(self) -> int:
        return self.batch.covered()

    def active(self, q: int) -> int:
        """Test primes whose targets reach q."""
        return int(self.n[self.T >= q].sum())

    def next_target(self) -> int | None:
        c = self.covered()
        pending = [m for r, m in enumerate(self.maxt) if r not in self.done and m >= c]
        return min(pending) if pending else None

    def work_to_next(self) -> float:
        """Factor applications until the next range can finish."""
        m = self.next_target()
        if m is None:
            return 0.0
        c = self.covered()
        sel = self.T >= c
        return float((self.n[sel] * (np.minimum(self.piT[sel], pi_approx(m)) - pi_approx(c))).sum())


class Server:
    """Continuous-batching request loop over the batched engine.

    Two threads: the admission thread builds cohorts (host test-prime sieve
    on worker threads, tables, setup) on its own stream every
    `admit_interval` s while the event loop keeps a window in flight on each
    of `streams` streams; built cohorts are bound to a stream and join the
    schedule at once. streams + 1 stays within the 4 hardware queues HIP
    gives a process by default, so no stream waits behind another's kernel
    in a shared queue."""

    def __init__(self, device: int = 0, hbm_budget: int = 64 << 30, streams: int = 3, step_seconds: float = 0.5,
                 apps_per_second: float = 1e13, segment_blocks: int = 2560, max_cohort: int = 64,
                 admit_interval: float = 0.5, host_primes: bool = True, class_bits: int = 4,
                 cohort_primes: int = 0, run_products: int = 2, admit_wait: float = 0.0,
                 min_cohort_primes: int = 1 << 20, max_cohorts: int = 48):
        self.devices = [Device(device) for _ in range(streams)]
        self.admit_dev = Device(device)
        # chunk products of every cohort on a stream live in that stream's
        # scratch (windows on one stream run one after another)
        self.scratch = [Scratch(d, self.SCRATCH_SLOTS) for d in self.devices]
        self.budget = hbm_budget
        self.step_seconds = step_seconds
        self.rate = apps_per_second
        # segments of 2560 blocks (157 M integers, ~5.7 M sieving primes at
        # 1e12): a launch of the matrix-core run products then has some 5,500
        # sub-chunks per slice to spread its per-workgroup costs (table, fold)
        # over; at 640 blocks a replay sustained 5.5e12 factor applications/s,
        # at 2560 8.6e12, and 9.7e12 with windows of 0.5 s instead of 0.2 s
        self.segment_span = segment_blocks * BLOCK_SPAN
        # and the window buffers (record buffer, sieve bitmap, run products)
        # of every cohort on a stream in that stream's workspace: a cohort
        # then holds only its tables and accumulators
        self.workspaces = [Workspace(d, self.segment_span) for d in self.devices]
        self.host_primes = host_primes
        self.max_cohort = max_cohort
        self.class_bits = class_bits          # magnitude classes 2^class_bits wide
        self.cohort_primes = cohort_primes or self.COHORT_PRIMES  # estimated test primes per cohort at most
        self.admit_interval = admit_interval
        self.run_products = run_products      # clw_options.run_products of every cohort
        self.admit_wait = admit_wait          # seconds a class may wait to grow into a cohort
        self.min_cohort_primes = min_cohort_primes
        # cohorts built and not retired at most (0: no cap). Fewer, larger
        # cohorts sieve each segment for more test primes: a 30 s replay ran
        # at 9.1e12 factor applications/s with 48 (75 requests completed,
        # p50 8.8 s), 7.3e12 with 96 (104, 7.3 s), 6.5e12 with no cap (152,
        # 4.3 s)
        self.max_cohorts = max_cohorts
        self.lock = threading.Lock()
        self.queue: list[Request] = []
        self.ready: list[_Cohort] = []       # built, not yet bound to a stream
        self.cohorts: list[_Cohort] = []
        self.running: list[_Cohort | None] = [None] * streams
        # scratch + workspaces + built cohorts (device bytes)
        self.used = len(self.devices) * self.SCRATCH_SLOTS * 16 + sum(w.bytes for w in self.workspaces)
        self.reserved = 0                    # estimated bytes of cohorts being built
        self.peak_used = 0
        self.completed: list[Completion] = []
        self.finishing: list[tuple[_Cohort, int]] = []  # finishes in flight (launched, not collected)
        self.factor_apps = 0
        self.windows = 0
        self.admit_seconds = 0.0
        self._slot = 0
        self._stop = threading.Event()
        self._thread: threading.Thread | None = None
        self._error: BaseException | None = None
        self.t0 = time.perf_counter()

    SCRATCH_SLOTS = 1 << 23    # chunk-product slots per stream (2 x 64 MiB)
    COHORT_PRIMES = 1 << 22    # estimated test primes per cohort at most
    FILL_PRIMES = 1 << 18      # test primes that fill the GPU in one window
    BUILD_THREADS = 4          # cohorts built at once in an admission round

    def close(self) -> None:
        self.stop_admission()
        for c in self.cohorts + self.ready:
            c.batch.wait()
            c.batch.close()
        self.cohorts, self.ready = [], []
        for sc in self.scratch:
            sc.close()
        for w in self.workspaces:
            w.close()
        for d in self.devices + [self.admit_dev]:
            d.close()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, req: Request) -> None:
        with self.lock:
            self.queue.append(req)

    # -- admission (its own thread) ---------------------------------------------
    @staticmethod
    def estimate_primes(req: Request) -> float:
        """~1.1 width / ln p test primes (an upper estimate)."""
        return 1.1 * (req.pmax - req.pmin) / math.log(max(req.pmin, 3)) + 64

    # the 3-limb band (2^60 <= 2 p^2 < 2^90) keeps 8 partial products per
    # test prime for the matrix-core run products (cohorts build their byte
    # tables per workgroup: run_products=2, no table memory)
    MFMA_BAND = (int(2 ** 29.5), int(2 ** 44.5))

    @classmethod
    def estimate_bytes(cls, req: Request) -> int:
        """Upper estimate of a request's device footprint: its test primes x
        (table 88 B + group index, start target and plan ~32 B, + 128 B of
        partial products in the 3-limb band)."""
        n = cls.estimate_primes(req)
        e = n * 120
        if req.pmax > cls.MFMA_BAND[0] and req.pmin < cls.MFMA_BAND[1]:
            e += n * 128
        return int(e)

    @staticmethod
    def magnitude_class(req: Request, bits: int = 4) -> int:
        """Classes of the largest factorial target (~pmax/2), a factor 2^bits
        wide."""
        return int(math.log2(max(req.pmax // 2, 2)) / bits)

    def admit(self) -> int:
        """One admission round; returns the number of requests admitted."""
        # a cohort's own buffers beside its tables (descriptor, counters,
        # plans); its window buffers are its stream's workspace
        fixed = 1 << 20
        classes: dict[int, list[Request]] = {}
        est: dict[int, int] = {}
        primes: dict[int, float] = {}
        with self.lock:
            # at most max_cohorts cohorts at a time: requests queue meanwhile
            # and the next round makes fewer, larger cohorts of them
            room = self.max_cohorts - len(self.cohorts) - len(self.ready) if self.max_cohorts else 1 << 30
            if room <= 0:
                return 0
            # a class waits until its queued requests make a cohort of
            # min_cohort_primes test primes or its oldest has waited
            # admit_wait seconds: small cohorts each sieve every segment for
            # few test primes
            now = self.now()
            total: dict[int, float] = {}
            oldest: dict[int, float] = {}
            for r in self.queue:
                k = self.magnitude_class(r, self.class_bits)
                total[k] = total.get(k, 0.0) + self.estimate_primes(r)
                oldest.setdefault(k, r.arrival)
            ready = {k for k in total if total[k] >= self.min_cohort_primes or now - oldest[k] >= self.admit_wait}
            waiting: list[Request] = []
            i = 0
            while i < len(self.queue):
                r = self.queue[i]
                k = self.magnitude_class(r, self.class_bits)
                if k not in ready:
                    waiting.append(r)
                    i += 1
                    continue
                grp = classes.get(k)
                if grp is not None and (len(grp) >= self.max_cohort or
                                        primes[k] + self.estimate_primes(r) > self.cohort_primes):
                    break
                if grp is None and len(classes) >= room:
                    break
                e = self.estimate_bytes(r) + (0 if grp else fixed)
                if self.used + self.reserved + e > self.budget:
                    break
                classes.setdefault(k, []).append(r)
                est[k] = est.get(k, 0) + e
                primes[k] = primes.get(k, 0.0) + self.estimate_primes(r)
                self.reserved += e
                i += 1
            self.queue = waiting + self.queue[i:]
        t = time.perf_counter()

        def build(k):
            c = self._build(classes[k])
            with self.lock:
                self.ready.append(c)
                self.used += c.bytes
                self.reserved -= est[k]
                self.peak_used = max(self.peak_used, self.used)

        # the classes' cohorts are built side by side (the library releases
        # the interpreter; their uploads share the admission stream)
        ks = sorted(classes)
        if len(ks) > 1:
            with ThreadPoolExecutor(max_workers=min(len(ks), self.BUILD_THREADS)) as ex:
                list(ex.map(build, ks))
        elif ks:
            build(ks[0])
        self.admit_seconds += time.perf_counter() - t
        return sum(len(classes[k]) for k in ks)

    def _build(self, take: list[Request]) -> _Cohort:
        """A cohort of requests: their batch built on the admission stream."""
        b = Batch(self.admit_dev, [(r.pmin, r.pmax) for r in take], segment_span=self.segment_span,
                  segment_apps=int(self.rate * self.step_seconds), host_primes=self.host_primes,
                  external_scratch=True, external_workspace=True,
                  run_products=self.run_products)
        groups = [(st.typeTarget[t], st.tpcnt[t]) for st in b.info for t in range(3) if st.tpcnt[t]]
        c = _Cohort(b, -1, take, [st.maxtarget for st in b.info], np.array([g[0] for g in groups], dtype=np.float64),
                    np.array([g[1] for g in groups], dtype=np.int64), self.now())
        c.bytes = b.device_bytes()
        c.rate = self.rate / 8  # factor applications/s, refined by every window it runs
        return c

    def _admission_loop(self) -> None:
        try:
            while not self._stop.is_set():
                if self.queue:
                    self.admit()
                self._stop.wait(self.admit_interval)
        except BaseException as e:  # surfaced by the event loop
            self._error = e

    def start_admission(self) -> None:
        if self._thread is None:
            self._stop.clear()
            self._thread = threading.Thread(target=self._admission_loop, daemon=True)
            self._thread.start()

    def stop_admission(self) -> None:
        if self._thread is not None:
            self._stop.set()
            self._thread.join()
            self._thread = None

    # -- windows (event loop) -----------------------------------------------------
    def _window(self, c: _Cohort) -> tuple[int, int]:
        lo = c.covered()
        nxt = c.next_target()
        act = max(1, c.active(lo))
        w = int(c.rate * self.step_seconds * math.log(max(lo, 16)) / act) + 1
        if lo < (1 << 20):
            # small sieving primes carry large exponents (a power per prime):
            # their cost is not proportional to the count, so go by bounds
            w = max(w, (1 << 16) - lo) if lo < (1 << 16) else min(max(w, lo), (1 << 20) - lo)
        return lo, min(lo + max(w, 1), nxt + 1)

    def _bind_ready(self) -> None:
        with self.lock:
            ready, self.ready = self.ready, []
        for c in ready:
            # stream 0 takes the cohorts with enough test primes to fill the
            # GPU on their own (throughput), the others the smaller ones
            # (latency), each the least loaded of its kind
            n = int(c.n.sum())
            kinds = [0] if (n >= self.FILL_PRIMES or len(self.devices) == 1) else list(range(1, len(self.devices)))
            load = {s: sum(1 for x in self.cohorts if x.slot == s) for s in kinds}
            c.slot = min(kinds, key=lambda s: load[s])
            c.batch.bind(self.devices[c.slot])
            c.batch.set_scratch(self.scratch[c.slot])
            c.batch.set_workspace(self.workspaces[c.slot])
            self.cohorts.append(c)

    def _launch(self, slot: int) -> bool:
        idle = [c for c in self.cohorts if c.slot == slot and c.next_target() is not None]
        if not idle:
            return False
        c = min(idle, key=lambda c: c.work)
        lo, hi = self._window(c)
        c.window = (lo, hi, time.perf_counter(), c.batch.kernel_stats()["mul_prime_apps"])
        c.batch.launch(lo, hi)
        self.running[slot] = c
        return True

    def _complete(self, slot: int) -> int:
        c = self.running[slot]
        c.batch.wait()
        self.running[slot] = None
        self.windows += 1
        lo, hi, t, before = c.window
        apps = c.batch.kernel_stats()["mul_prime_apps"] - before
        self.factor_apps += apps
        dt = time.perf_counter() - t
        if lo >= (1 << 20) and dt > 0 and apps > 0:
            c.rate = 0.5 * c.rate + 0.5 * apps / dt
        launched = 0
        cov = c.covered()
        for r, m in enumerate(c.maxt):
            if r not in c.done and r not in c.finishing and m < cov:
                c.batch.finish_launch(r)  # queued behind the window on the cohort's stream
                c.finishing.add(r)
                self.finishing.append((c, r))
                launched += 1
        c.work = c.work_to_next()
        return launched

    def _collect(self, block: bool = False) -> int:
        """Results of the finishes that are in (all of them with block);
        a cohort whose ranges are all collected is retired."""
        done = 0
        keep = []
        for c, r in self.finishing:
            if not block and not c.batch.finish_ready(r):
                keep.append((c, r))
                continue
            res = c.batch.finish_range(r)
            c.finishing.discard(r)
            c.done.add(r)
            self.completed.append(Completion(c.reqs[r].rid, c.reqs[r].arrival, c.admitted, self.now(), res))
            done += 1
            if len(c.done) == len(c.reqs):
                c.batch.close()
                with self.lock:
                    self.used -= c.bytes
                self.cohorts.remove(c)
        self.finishing = keep
        return done

    def poll(self) -> bool:
        """One pass of the event loop: finished windows completed, built
        cohorts bound, a new window on every idle stream. Returns whether
        anything happened."""
        if self._error is not None:
            raise self._error
        busy = self._collect() > 0
        if self.ready:
            self._bind_ready()
            busy = True
        for slot in range(len(self.running)):
            c = self.running[slot]
            if c is not None and c.batch.done():
                self._complete(slot)
                busy = True
            if self.running[slot] is None:
                busy |= self._launch(slot)
        return busy

    def drain(self) -> None:
        """Waits for every window in flight (used before shutting down)."""
        for slot, c in enumerate(self.running):
            if c is not None:
                self._complete(slot)

    # -- snapshot / restore -------------------------------------------------------
    def snapshot(self, directory: str) -> None:
        """Checkpoint the server: every cohort's batch (clw_batch_ckpt_write:
        residues, covered prefix, ranges finished on their own) and a
        manifest of cohorts, their requests and the queue. Windows in flight
        are waited for first; admission must be stopped."""
        if self._thread is not None:
            raise RuntimeError("stop admission before a snapshot")
        self.drain()
        self._collect(block=True)
        self._bind_ready()
        os.makedirs(directory, exist_ok=True)
        cohorts = []
        for i, c in enumerate(self.cohorts):
            path = os.path.join(directory, f"cohort{i}.ckpt")
            c.batch.checkpoint(path)
            cohorts.append(dict(file=os.path.basename(path), done=sorted(c.done), admitted=c.admitted,
                                requests=[[r.rid, r.arrival, r.pmin, r.pmax] for r in c.reqs]))
        with self.lock:
            queue = [[r.rid, r.arrival, r.pmin, r.pmax] for r in self.queue]
        with open(os.path.join(directory, "manifest.json"), "w") as f:
            json.dump(dict(cohorts=cohorts, queue=queue, completed=[c.rid for c in self.completed]), f)

    def restore(self, directory: str) -> list[int]:
        """Rebuild the cohorts of a snapshot (same ranges, residues read back)
        and its queue; returns the ids of requests completed before it."""
        with open(os.path.join(directory, "manifest.json")) as f:
            m = json.load(f)
        for entry in m["cohorts"]:
            take = [Request(*r) for r in entry["requests"]]
            c = self._build(take)
            path = os.path.join(directory, entry["file"])
            result, _ = c.batch.resume(path, path)
            if result != 1:
                raise RuntimeError(f"snapshot file {path} is not valid for its cohort")
            c.done = set(entry["done"])
            c.admitted = entry["admitted"]
            c.work = c.work_to_next()
            with self.lock:
                self.ready.append(c)
                self.used += c.bytes
                self.peak_used = max(self.peak_used, self.used)
        with self.lock:
            self.queue.extend(Request(*r) for r in m["queue"])
        self._bind_ready()
        return m["completed"]

    def idle(self) -> bool:
        with self.lock:
            return not (self.queue or self.ready or self.cohorts or self.reserved or self.finishing)

    def serve(self, requests: list[Request], duration: float, progress=None, until_done: bool = False) -> dict:
        """Replays `requests` in real time (arrival seconds after start) for
        `duration` seconds of wall clock (or until every request is done,
        with until_done) and reports latency and throughput."""
        pending = sorted(requests, key=lambda r: r.arrival)
        k = 0
        self.t0 = time.perf_counter()
        self.start_admission()
        last = 0.0
        try:
            while True:
                t = self.now()
                if t >= duration:
                    break
                if until_done and k == len(pending) and self.idle():
                    break
                while k < len(pending) and pending[k].arrival <= t:
                    self.submit(pending[k])
                    k += 1
                if not self.poll():
                    time.sleep(0.0005)
                if progress and self.now() - last >= 10:
                    last = self.now()
                    progress(self.report(k))
        finally:
            self.stop_admission()
            self.drain()
        return self.report(k)

    def report(self, arrived: int) -> dict:
        lat = sorted(c.latency for c in self.completed)

        def pct(p):
            if not lat:
                return None
            return lat[min(len(lat) - 1, int(math.ceil(p / 100 * len(lat))) - 1)]

        wall = self.now()
        tested = sum(len(c.result.p) for c in self.completed)
        with self.lock:
            queued = len(self.queue)
        return dict(wall_s=wall, arrived=arrived, admitted=arrived - queued,
                    completed=len(self.completed),
                    in_flight=sum(len(c.reqs) - len(c.done) for c in self.cohorts + self.ready),
                    queued=queued, p50_latency_s=pct(50), p99_latency_s=pct(99),
                    primes_tested=tested, primes_tested_per_s=tested / wall if wall else 0.0,
                    factor_apps_per_s=self.factor_apps / wall if wall else 0.0, windows=self.windows,
                    admission_thread_s=self.admit_seconds, streams=len(self.devices),
                    hbm_budget=self.budget, hbm_peak=self.peak_used, cohorts=len(self.cohorts))


def alone(dev: Device, req: Request) -> RangeResult:
    """The same request served by itself (a one-range batch)."""
    with Batch(dev, [(req.pmin, req.pmax)]) as b:
        return b.search()[0]


__all__ = ["Server", "Completion", "alone", "host_primes"]
