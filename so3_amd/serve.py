"""Serving loop: continuous batching of SO(3) transform requests on one GPU.

Requests (a forward+inverse pair, or an S^2 -> SO(3) harmonic convolution
followed by the inverse transform) arrive over time.  The loop

* moves arrived requests into a FIFO queue;
* admits queued requests in order while the HBM footprint of everything in
  flight (the request's input/output buffers plus one transform workspace)
  stays within a budget, at most `max_batch` per round;
* launches each admission round as one batch on the next of a few caller
  streams, so that a new round starts while earlier ones still run, and
  takes a completion ticket.  In the persistent mode (the default) the
  round's contractions, whatever the sizes of its requests, run in one
  persistent launch per direction fed by device queues, the other stages
  per request on lanes of HIP streams (so3x_batch_persistent); in the lanes
  mode every request runs whole on one lane (so3x_batch);
* polls the tickets and records each request's latency (completion minus
  arrival) and, on request, its outputs.

The Delta tables, DFT plans and index tables stay resident in the library
between requests (the table cache is bounded; see so3x_cache_limit; one
table per recursion method, at the largest L, serves every smaller L), so a
request of a shape seen before allocates only its own buffers.
"""
import collections
import dataclasses
import time

import numpy as np

from . import api as A
from . import params as P


@dataclasses.dataclass
class Request:
    rid: int
    kind: int                  # A.ITEM_PAIR or A.ITEM_S2TOSO3_INVERSE
    p: object                  # SO3Parameters
    arrival: float             # seconds after the start of the run
    done: float = None         # completion time (seconds after the start)
    outputs: dict = None       # host copies when the run keeps outputs


def request_mix(count, rate, seed=5):
    """The mixed workload: Poisson arrivals at `rate` per second; 70%
    transform pairs with L log-uniform on [16, 512] and N uniform on
    [1, min(L, 32)], 30% s2toso3 + inverse with L log-uniform on [64, 1024]
    and N in {4, 8}; MW, complex, padded, zero-first."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    out = []
    for i in range(count):
        if rng.uniform() < 0.7:
            L = int(round(np.exp(rng.uniform(np.log(16), np.log(512)))))
            N = int(rng.integers(1, min(L, 32) + 1))
            kind = A.ITEM_PAIR
        else:
            L = int(round(np.exp(rng.uniform(np.log(64), np.log(1024)))))
            N = int(rng.choice([4, 8]))
            kind = A.ITEM_S2TOSO3_INVERSE
        p = P.SO3Parameters(L=L, N=N, n_order=P.SO3_N_ORDER_ZERO_FIRST)
        out.append(Request(i, kind, p, float(t[i])))
    return out


class Server:
    """Continuous-batching executor over the library's batch entry point."""

    def __init__(self, hbm_budget=48 << 30, lanes=4, max_batch=32, streams=3, device="cuda:0",
                 mode=A.BATCH_PERSISTENT):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        self.budget = int(hbm_budget)
        self.lanes = lanes
        self.mode = mode
        self.max_batch = max_batch
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(streams)]
        self._next = 0

    def footprint(self, r):
        """HBM bytes a request holds while in flight."""
        p = r.p
        fb = A.f_size(p) * (8 if p.reality else 16)
        nb = A.flmn_size(p) * 16
        ws = A.workspace_bytes(p)  # each request's own in the persistent mode
        if r.kind == A.ITEM_PAIR:
            return fb + 2 * nb + ws
        return fb + nb + 2 * 16 * p.L * p.L + ws

    def _inputs(self, r, input_fn):
        """Device buffers and the batch item of request r."""
        torch = self.torch
        p = r.p
        fdt = torch.float64 if p.reality else torch.complex128
        f = torch.empty(A.f_size(p), dtype=fdt, device=self.dev)
        out = torch.zeros(A.flmn_size(p), dtype=torch.complex128, device=self.dev)
        bufs = {"f": f, "flmn_out": out}
        spec = {"p": p, "kind": r.kind, "f": f.data_ptr(), "flmn_out": out.data_ptr()}
        data = input_fn(r)

        def dev(x):  # pinned host tensors copy asynchronously on the current stream
            if isinstance(x, torch.Tensor):
                return x.to(self.dev, non_blocking=True)
            return torch.as_tensor(x, device=self.dev)

        if r.kind == A.ITEM_PAIR:
            bufs["flmn_in"] = dev(data["flmn"])
            spec["flmn_in"] = bufs["flmn_in"].data_ptr()
        else:
            bufs["flm"] = dev(data["flm"])
            bufs["glm"] = dev(data["glm"])
            spec["flmn_in"] = bufs["flm"].data_ptr()
            spec["glm"] = bufs["glm"].data_ptr()
        return spec, bufs

    def prepare(self, requests):
        """Build the resident tables and plans of every shape in `requests`.
        Largest L first: the Delta table of the largest L serves every smaller
        L, so one table per method is built."""
        seen = set()
        for r in sorted(requests, key=lambda r: -r.p.L):
            key = (r.p.L, r.p.N, r.p.reality, r.p.sampling_scheme, r.p.dl_method)
            if key not in seen:
                seen.add(key)
                A.prepare(r.p)

    def run(self, requests, input_fn, keep_outputs=False, poll_s=5e-5):
        """Serve `requests` in real time (arrival times relative to the call);
        returns them with `done` (and `outputs`) filled in."""
        torch = self.torch
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue = collections.deque()
        inflight = []  # (ticket, [(request, bufs, bytes)])
        used = 0
        t0 = time.perf_counter()
        while pending or queue or inflight:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            # completions
            still = []
            for tk, members in inflight:
                if A.ticket_done(tk):
                    A.ticket_wait(tk)
                    t_done = time.perf_counter() - t0
                    for r, bufs, nbytes in members:
                        r.done = t_done
                        if keep_outputs:
                            r.outputs = {k: bufs[k].cpu().numpy() for k in ("f", "flmn_out")}
                        used -= nbytes
                else:
                    still.append((tk, members))
            inflight = still
            # admission of one round
            members = []
            while queue and len(members) < self.max_batch:
                r = queue[0]
                nbytes = self.footprint(r)
                if used + nbytes > self.budget and (inflight or members):
                    break  # wait for room; a lone oversized request still runs
                queue.popleft()
                used += nbytes
                members.append((r, nbytes))
            if members:
                stream = self.streams[self._next % len(self.streams)]
                self._next += 1
                specs, held = [], []
                with torch.cuda.stream(stream):
                    for r, nbytes in members:
                        spec, bufs = self._inputs(r, input_fn)
                        specs.append(spec)
                        held.append((r, bufs, nbytes))
                items = A.make_items(specs)
                if self.mode == A.BATCH_PERSISTENT:
                    A.batch_persistent(items, lanes=self.lanes, stream=stream.cuda_stream)
                else:
                    A.batch(items, lanes=self.lanes, stream=stream.cuda_stream)
                inflight.append((A.ticket(stream.cuda_stream), held))
            elif not pending or pending[0].arrival > now + poll_s:
                time.sleep(poll_s)
        return requests


def random_inputs(device_rng_seed=0):
    """input_fn drawing U[-1, 1) inputs (numpy, seeded by request id)."""
    def fn(r):
        rng = np.random.default_rng((device_rng_seed, r.rid))
        if r.kind == A.ITEM_PAIR:
            n = A.flmn_size(r.p)
            return {"flmn": rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)}
        n = r.p.L * r.p.L
        return {"flm": rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n),
                "glm": rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)}
    return fn


def device_drawn_pinned_inputs(requests, seed=0, device="cuda:0"):
    """input_fn over `requests` whose U[-1, 1) inputs are drawn on the GPU
    (torch's Philox generator seeded by (seed, request id)) ahead of time and
    held in pinned host memory: the same role as pinned_inputs, without the
    CPU time of drawing gigabytes with numpy."""
    import torch
    dev = torch.device(device)
    g = torch.Generator(device=dev)
    cache = {}

    def draw(n):
        re = torch.rand(n, generator=g, device=dev, dtype=torch.float64) * 2 - 1
        im = torch.rand(n, generator=g, device=dev, dtype=torch.float64) * 2 - 1
        out = torch.empty(n, dtype=torch.complex128, pin_memory=True)
        out.copy_(torch.complex(re, im))
        return out

    for r in requests:
        if r.rid in cache:
            continue
        g.manual_seed(seed * 1000003 + r.rid)
        if r.kind == A.ITEM_PAIR:
            cache[r.rid] = {"flmn": draw(A.flmn_size(r.p))}
        else:
            n = r.p.L * r.p.L
            cache[r.rid] = {"flm": draw(n), "glm": draw(n)}
    return lambda r: cache[r.rid]


def pinned_inputs(requests, input_fn):
    """input_fn over `requests` drawn ahead of time into pinned host memory,
    as clients hand their data over: the serving loop then only enqueues
    asynchronous copies (drawing the larger inputs in the loop took longer
    than transforming them)."""
    import torch
    cache = {}
    for r in requests:
        if r.rid not in cache:
            cache[r.rid] = {k: torch.from_numpy(v).pin_memory() for k, v in input_fn(r).items()}
    return lambda r: cache[r.rid]


def latency_stats(requests):
    lat = np.array([r.done - r.arrival for r in requests]) * 1e3
    span = max(r.done for r in requests) - min(r.arrival for r in requests)
    return {"requests": len(requests), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
            "throughput_rps": len(requests) / span}
