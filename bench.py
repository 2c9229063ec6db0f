#!/usr/bin/env python3
"""FP64 Wigner forward+inverse transforms per second on one MI355X.

Workload (one "step"): an inverse transform flmn -> f followed by a forward
transform f -> flmn, complex FP64, L = N = 256, MW sampling, Risbo Delta,
padded storage, zero-first n order, all n (so3_core_inverse_via_ssht /
so3_core_forward_via_ssht).  Inputs are seeded synthetic coefficients
(Re, Im iid U[-1, 1)) resident in HBM before timing starts; the timed region
is K steps on one stream, bracketed by synchronisation.

Prints one JSON line: value = fwd+inv pairs per second, with p50/p99 step
latency, the round-trip error, a roofline entry for the dominant kernel
(timed live with HIP events on the launch stream by the library's stage
profiler; algorithmic bytes for the DFT stages, flops for the
contractions) and one per stage and the CPU
baseline (the oracle's single-thread restatement of the reference
algorithm, timed on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def contraction_terms(L, N, L0=0, real=False):
    """T = sum over n, l >= max(L0,|n|) of (2l+1)(l+1): Delta-product terms
    (l, m, n, m' >= 0) of one contraction."""
    ns = range(0, N) if real else range(-(N - 1), N)
    T = 0
    for n in ns:
        for l in range(max(L0, abs(n)), L):
            T += (2 * l + 1) * (l + 1)
    return T


def stage_model(p, A, T):
    """Algorithmic work of each pipeline stage of one fwd+inv pair: bytes that
    must cross HBM (read + write once) for the DFT stages, FP64 flops for the
    contractions (5 per Delta-product term)."""
    L, N, real = p.L, p.N, bool(p.reality)
    ss = p.sampling_scheme == 1
    nA, nB = (2 * L, L + 1) if ss else (2 * L - 1, L)
    nG = N if real else 2 * N - 1          # n slots of the complex workspaces
    C = 2 if real else 4                   # sign combinations per (|m|, |n|)
    esz = 8 if real else 16
    x1 = 16 * nG * nA * L                  # contraction output rows
    x2 = 16 * nG * nB * nA                 # beta / alpha workspace
    gp = 16 * N * L * L * C                # folded G' of the forward contraction
    fb = esz * A.f_size(p)
    return {
        "contract_inv": ("mfma", 5.0 * T),
        "beta_inv": ("hbm", x1 + x2),
        "alpha_inv": ("hbm", 2 * x2),
        "gamma_inv": ("hbm", x2 + fb),
        "gamma_fwd": ("hbm", fb + x2),
        "alpha_fwd": ("hbm", 2 * x2),
        "beta_fwd": ("hbm", x2 + gp),
        "contract_fwd": ("mfma", 5.0 * T),
    }


# kernel that carries each stage (names as rocprofv3 reports them)
STAGE_KERNEL = {"contract_inv": "k_contract_inv", "beta_inv": "k_inv_beta_r32", "alpha_inv": "k_rows_r32<1, 1>",
                "gamma_inv": "k_gamma_r32<1>", "gamma_fwd": "k_gamma_r32<-1>", "alpha_fwd": "k_rows_r32<-1, 1>",
                "beta_fwd": "k_fwd_beta_r32", "contract_fwd": "k_contract_fwd"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_l256.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and not line.startswith("#") and parts[0].startswith(kernel.replace(" ", "")):
            return float(parts[1]) + float(parts[2])
    return None


def cpu_baseline(p, n_slices):
    """Single-thread oracle time of one fwd+inv pair, from a bounded sample:
    n_slices evenly spaced n slices timed in full (spin-SHT inverse and
    forward each), scaled to all 2N-1 slices, plus the Delta-table build and
    two gamma DFTs timed on every 16th (a, b) column and scaled."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    N = p.N
    ns = np.unique(np.linspace(-(N - 1), N - 1, n_slices).round().astype(int))
    t0 = time.time()
    t_n, t_dl, t_g = pyoracle.time_sample(p, ns, gamma_every=16)
    wall = time.time() - t0
    total = t_dl + t_n.mean() * (2 * N - 1) + 2 * t_g
    return {
        "value": 1.0 / total,
        "unit": "fwd+inv transforms/s",
        "cores": 1,
        "kind": "port",
        "sample": "%d of %d n-slices (spin-SHT inverse+forward each) at L=N=%d, scaled x%.1f; "
                  "gamma DFT on 1/16 of the columns, scaled; Delta table once; %.1f s of CPU work; "
                  "estimated %.1f s per fwd+inv" % (len(ns), 2 * N - 1, p.L, (2 * N - 1) / len(ns), wall, total),
    }


def hbm_practical(torch, dev, stream, reps=10):
    """The HBM rate a plain streaming kernel reaches on this device: a 1 GiB
    FP64 elementwise add (reads one GiB, writes one), HIP events on the
    launch stream.  Context for the hbm-bound stages' fractions of the
    published peak."""
    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    torch.add(a, 1.0, out=b)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        torch.add(a, 1.0, out=b)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return {"gbs": round(2 * (1 << 30) / (ms * 1e-3) / 1e9, 1), "ms": round(ms, 4),
            "how": "1 GiB FP64 elementwise add (1 GiB read + 1 GiB written), HIP events"}


def batched_workload(P, seed=4, count=96):
    """The uneven batch: L uniform on {32, 48, ..., 384}, N uniform on
    [1, min(L, 64)], MW / MW_SS and complex / real 50/50 each."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        L = int(rng.choice(np.arange(32, 385, 16)))
        N = int(rng.integers(1, min(L, 64) + 1))
        ss = int(rng.integers(0, 2))
        real = int(rng.integers(0, 2))
        out.append(P.SO3Parameters(L=L, N=N, sampling_scheme=ss, reality=real,
                                   n_order=P.SO3_N_ORDER_ZERO_FIRST))
    return out


def _pair_specs(A, torch, ps, seed):
    """device buffers and batch items of fwd+inv pairs for parameter list ps"""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    specs, keep = [], []
    for p in ps:
        nfl = A.flmn_size(p)
        flmn = torch.complex(torch.rand(nfl, generator=g, device=dev, dtype=torch.float64) * 2 - 1,
                             torch.rand(nfl, generator=g, device=dev, dtype=torch.float64) * 2 - 1)
        out = torch.zeros_like(flmn)
        f = torch.empty(A.f_size(p), dtype=torch.float64 if p.reality else torch.complex128, device=dev)
        keep += [flmn, out, f]
        specs.append({"p": p, "kind": A.ITEM_PAIR, "f": f.data_ptr(), "flmn_in": flmn.data_ptr(),
                      "flmn_out": out.data_ptr()})
    return specs, keep


def _batch_legs(A, torch, specs, reps):
    """plain sequential calls, then the batch in both modes, eager and as a
    captured graph (medians of `reps` runs, milliseconds)"""
    items = A.make_items(specs)
    stream = torch.cuda.current_stream()
    sid = stream.cuda_stream

    def timeit(fn):
        fn()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
        ev[0].record(stream)
        for k in range(reps):
            fn()
            ev[k + 1].record(stream)
        torch.cuda.synchronize()
        return np.array([ev[k].elapsed_time(ev[k + 1]) for k in range(reps)])

    def plain():
        for sp in specs:
            A.inverse_device(sp["f"], sp["flmn_in"], sp["p"], A.VIA_SSHT, sid)
            A.forward_device(sp["flmn_out"], sp["f"], sp["p"], A.VIA_SSHT, sid)

    out = {"plain": timeit(plain)}
    lanes = A.batch_tune(items, stream=sid)
    mode, _ = A.batch_choice(items)
    out["lanes"] = timeit(lambda: A.batch(items, lanes=lanes, stream=sid))
    out["persistent"] = timeit(lambda: A.batch_persistent(items, lanes=lanes, stream=sid))
    for m, name in ((A.BATCH_LANES, "lanes_graph"), (A.BATCH_PERSISTENT, "persistent_graph")):
        gh = A.graph_create(items, lanes=lanes, mode=m)
        try:
            out[name] = timeit(lambda: A.graph_launch(gh, sid))
        finally:
            A.graph_destroy(gh)
    med = {k: float(np.median(v)) for k, v in out.items()}
    best = min((v, k) for k, v in med.items() if k != "plain")
    n = len(specs)
    return {
        "items": n, "tuned_lanes": lanes, "tuned_mode": "persistent" if mode == A.BATCH_PERSISTENT else "lanes",
        "ms": {k: round(v, 3) for k, v in med.items()},
        "plain_pairs_per_s": n / (med["plain"] * 1e-3),
        "batched_pairs_per_s": n / (best[0] * 1e-3),
        "best": best[1],
        "speedup": med["plain"] / best[0],
        "p50_ms": float(np.percentile(out[best[1]], 50)),
        "p99_ms": float(np.percentile(out[best[1]], 99)),
    }


def batched_plain_only(A, P, torch, table, reps=2):
    """The 96-item batched workload as plain calls one after another (one
    warm-up pass, then `reps` timed), tuned by the table in `table` (made by
    the autotuner and saved there when the file does not exist)."""
    ps = batched_workload(P)
    specs, keep = _pair_specs(A, torch, ps, 4)
    if os.path.exists(table):
        A.tune_load(table)
        progress("batched plain: tuning table loaded")
    else:
        for p in ps:
            A.tune(p)
        A.tune_save(table)
        progress("batched plain: tuning table made and saved")
    sid = torch.cuda.current_stream().cuda_stream

    def plain():
        for sp in specs:
            A.inverse_device(sp["f"], sp["flmn_in"], sp["p"], A.VIA_SSHT, sid)
            A.forward_device(sp["flmn_out"], sp["f"], sp["p"], A.VIA_SSHT, sid)
    plain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plain()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    del keep
    return {"workload": "96 uneven fwd+inv pairs (seed 4) as plain calls, tuned", "reps": reps,
            "plain_ms": ms}


def batched_bench(A, P, torch, reps=5):
    """96 uneven fwd+inv pairs: plain calls one after another on one stream
    versus one batch call (lanes or persistent mode, tuned) and its captured
    graph, after the (L, N) autotuner has run for every item shape; and the
    same comparison on a stream of 256 small pairs (L 16..63), where the
    batch modes matter most."""
    ps = batched_workload(P)
    specs, keep = _pair_specs(A, torch, ps, 4)
    # plain before tuning, for reference
    t0 = time.perf_counter()
    sid = torch.cuda.current_stream().cuda_stream

    def plain_once():
        for sp in specs:
            A.inverse_device(sp["f"], sp["flmn_in"], sp["p"], A.VIA_SSHT, sid)
            A.forward_device(sp["flmn_out"], sp["f"], sp["p"], A.VIA_SSHT, sid)
    plain_once()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_once()
    torch.cuda.synchronize()
    untuned_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    shapes, defaults = {}, {}
    for p in ps:
        key = (p.L, p.N, p.reality, p.sampling_scheme)
        if key not in shapes:
            defaults[key] = A.tune_get(p)
            shapes[key] = A.tune(p)[0]
    tune_s = time.perf_counter() - t0
    progress("batched: autotuner over %d shapes" % len(shapes))
    out = _batch_legs(A, torch, specs, reps)
    progress("batched: 96 uneven pairs")
    out["workload"] = ("96 uneven fwd+inv pairs: L in {32,48,...,384}, N in [1,min(L,64)], MW/MW_SS, "
                       "complex/real (seed 4), via SSHT, Risbo, padded")
    out["plain_untuned_ms"] = untuned_ms
    out["autotuner"] = {"shapes": len(shapes), "seconds": tune_s,
                        "pw_changed": sum(1 for k, c in shapes.items() if c["pw"] != defaults[k]["pw"]),
                        "inverse_order_1": sum(1 for c in shapes.values() if c["corder_inv"]),
                        "forward_order_1": sum(1 for c in shapes.values() if c["corder_fwd"]),
                        "gamma_register": sum(1 for c in shapes.values() if c["gamma_engine"] == 2),
                        "gamma_workgroup": sum(1 for c in shapes.values() if c["gamma_engine"] == 1),
                        "gamma_changed": sum(1 for k, c in shapes.items()
                                             if c["gamma_engine"] != defaults[k]["gamma_engine"])}
    del keep
    small = [P.SO3Parameters(L=16 + (i * 7) % 48, N=1 + (i * 5) % min(16 + (i * 7) % 48, 16),
                             n_order=P.SO3_N_ORDER_ZERO_FIRST) for i in range(256)]
    for p in small:
        A.prepare(p)
    specs, keep = _pair_specs(A, torch, small, 5)
    out["small_stream"] = _batch_legs(A, torch, specs, reps)
    out["small_stream"]["workload"] = "256 complex MW fwd+inv pairs, L = 16 + 7i mod 48, N = 1 + 5i mod min(L,16)"
    return out


def serving_bench(A, rates=(100, 400, 800, 1600, 3200, 6400), count=120, capacity_count=400):
    """The mixed request stream (Poisson arrivals) through the serving loop at
    rising rates: throughput and p50 / p99 latency per rate, the highest rate
    the server keeps up with (throughput >= 95% of the offered rate), its
    capacity (every request arriving at once) and the plain per-request
    capacity of the same stream (one request at a time through the plain
    entry points: copy in, transform, wait)."""
    import torch
    from so3_amd import serve
    A.cache_limit(96 << 30)
    srv = serve.Server(hbm_budget=48 << 30, lanes=4, max_batch=32, streams=3)
    out = {"workload": "Poisson arrivals; 70%% fwd+inv pairs (L log-uniform [16,512], N in [1,min(L,32)]), "
                       "30%% s2toso3 + inverse (L log-uniform [64,1024], N in {4,8}); MW complex (seed 5); "
                       "%d requests per rate; tables and plans resident (built before timing); inputs "
                       "drawn ahead (on the GPU) into pinned host memory, copied to HBM inside the loop; "
                       "admission rounds run as persistent batches" % count,
           "rates": {}}
    sustained = 0.0
    for rate in rates:
        reqs = serve.request_mix(count, rate)
        srv.prepare(reqs)
        fn = serve.device_drawn_pinned_inputs(reqs, 5)
        warm = [serve.Request(r.rid, r.kind, r.p, 0.0) for r in reqs[:8]]
        srv.run(warm, fn)
        srv.run(reqs, fn)
        st = serve.latency_stats(reqs)
        out["rates"][str(rate)] = st
        progress("serving: %d requests/s offered" % rate)
        if st["throughput_rps"] >= 0.95 * rate:
            sustained = rate
    out["max_sustained_rps"] = sustained
    # capacity: the same mix, every request arriving at t = 0
    reqs = serve.request_mix(capacity_count, 1000.0)
    for r in reqs:
        r.arrival = 0.0
    srv.prepare(reqs)
    fn = serve.device_drawn_pinned_inputs(reqs, 5)
    srv.run([serve.Request(r.rid, r.kind, r.p, 0.0) for r in reqs[:8]], fn)
    srv.run(reqs, fn)
    cap = serve.latency_stats(reqs)
    progress("serving: capacity")
    # plain: one request at a time, copy in, plain calls, wait for the result
    dev = torch.device("cuda:0")
    sid = torch.cuda.current_stream().cuda_stream
    t0 = time.perf_counter()
    for r in reqs:
        d = fn(r)
        p = r.p
        f = torch.empty(A.f_size(p), dtype=torch.complex128, device=dev)
        if r.kind == A.ITEM_PAIR:
            x = d["flmn"].to(dev, non_blocking=True)
            back = torch.empty_like(x)
            A.inverse_device(f.data_ptr(), x.data_ptr(), p, A.VIA_SSHT, sid)
            A.forward_device(back.data_ptr(), f.data_ptr(), p, A.VIA_SSHT, sid)
        else:
            flm = d["flm"].to(dev, non_blocking=True)
            glm = d["glm"].to(dev, non_blocking=True)
            h = torch.empty(A.flmn_size(p), dtype=torch.complex128, device=dev)
            items = A.make_items([{"p": p, "kind": A.ITEM_S2TOSO3_INVERSE, "f": f.data_ptr(),
                                   "flmn_in": flm.data_ptr(), "glm": glm.data_ptr(), "flmn_out": h.data_ptr()}])
            A.batch(items, lanes=1, stream=sid)
        torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    n_tr = sum(2 if r.kind == A.ITEM_PAIR else 1 for r in reqs)
    out["capacity"] = {"requests": len(reqs), "served_rps": cap["throughput_rps"],
                       "served_transforms_per_s": n_tr / len(reqs) * cap["throughput_rps"],
                       "served_p50_ms": cap["p50_ms"], "served_p99_ms": cap["p99_ms"],
                       "plain_rps": len(reqs) / plain_s, "plain_transforms_per_s": n_tr / plain_s,
                       "speedup": cap["throughput_rps"] / (len(reqs) / plain_s),
                       "note": "all requests arriving at once; plain = the same requests one at a time "
                               "(copy in, transform, wait), inputs in pinned host memory in both"}
    return out


def variants_bench(A, P, torch, steps=5):
    """The other configurations the metric is quoted with: C1 variants
    (Trapani Delta, MW_SS grid, real signal) as fwd+inv pairs, and at L=1024
    the s2toso3 convolution followed by the inverse (N=8) and a fwd+inv pair
    (N=4).  Median milliseconds per item over `steps` runs after a warm-up."""
    from so3_amd import synth
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    sid = stream.cuda_stream
    out = {}

    def timeit(fn):
        fn()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        ev[0].record(stream)
        for k in range(steps):
            fn()
            ev[k + 1].record(stream)
        torch.cuda.synchronize()
        return float(np.median([ev[k].elapsed_time(ev[k + 1]) for k in range(steps)]))

    def pair(name, p):
        A.prepare(p)
        flmn = torch.from_numpy(synth.gen_flmn(p, 1)).to(dev)
        back = torch.zeros_like(flmn)
        f = torch.empty(A.f_size(p), dtype=torch.float64 if p.reality else torch.complex128, device=dev)

        def step():
            A.inverse_device(f.data_ptr(), flmn.data_ptr(), p, A.VIA_SSHT, sid)
            A.forward_device(back.data_ptr(), f.data_ptr(), p, A.VIA_SSHT, sid)
        ms = timeit(step)
        err = (back - flmn).abs().max().item() / flmn.abs().max().item()
        out[name] = {"ms_per_pair": ms, "pairs_per_s": 1000.0 / ms, "roundtrip_rel_err": err}

    z = P.SO3_N_ORDER_ZERO_FIRST
    pair("C1_trapani", P.SO3Parameters(L=256, N=256, dl_method=P.SSHT_DL_TRAPANI, n_order=z))
    pair("C1_mw_ss", P.SO3Parameters(L=256, N=256, sampling_scheme=P.SO3_SAMPLING_MW_SS, n_order=z))
    pair("C1_real", P.SO3Parameters(L=256, N=256, reality=1, n_order=z))
    pair("C2b_L1024_N4", P.SO3Parameters(L=1024, N=4, n_order=z))
    # C2: h = s2toso3(flm, glm) at L=1024, N=8, then the inverse transform of h
    hp = P.SO3Parameters(L=1024, N=8, n_order=z)
    A.prepare(hp)
    g = torch.Generator(device=dev)
    g.manual_seed(2)
    L2 = hp.L * hp.L
    flm = torch.complex(torch.rand(L2, generator=g, device=dev, dtype=torch.float64),
                        torch.rand(L2, generator=g, device=dev, dtype=torch.float64))
    glm = torch.complex(torch.rand(L2, generator=g, device=dev, dtype=torch.float64),
                        torch.rand(L2, generator=g, device=dev, dtype=torch.float64))
    h = torch.zeros(A.flmn_size(hp), dtype=torch.complex128, device=dev)
    f = torch.empty(A.f_size(hp), dtype=torch.complex128, device=dev)
    items = A.make_items([{"p": hp, "kind": A.ITEM_S2TOSO3_INVERSE, "f": f.data_ptr(),
                           "flmn_in": flm.data_ptr(), "glm": glm.data_ptr(), "flmn_out": h.data_ptr()}])
    ms = timeit(lambda: A.batch(items, lanes=1, stream=sid))
    out["C2_conv_inverse_L1024_N8"] = {"ms": ms, "per_s": 1000.0 / ms}
    return out


def host_abi_bench(A, p, flmn_h, reps=3):
    """End-to-end fwd+inv through the drop-in host-pointer C ABI
    (so3_core_inverse_via_ssht / so3_core_forward_via_ssht on pageable numpy
    arrays): each call copies its input to HBM, transforms and copies the
    result back, synchronously.  Median wall milliseconds per pair over
    `reps` pairs after one warm-up pair."""
    import ctypes
    lib = A.lib()
    s = ctypes.byref(p.struct())
    f = np.empty(A.f_size(p), dtype=np.float64 if p.reality else np.complex128)
    back = np.zeros(A.flmn_size(p), dtype=np.complex128)
    inv = lib.so3_core_inverse_via_ssht_real if p.reality else lib.so3_core_inverse_via_ssht
    fwd = lib.so3_core_forward_via_ssht_real if p.reality else lib.so3_core_forward_via_ssht
    times = []
    for k in range(reps + 1):
        t0 = time.perf_counter()
        inv(A._ptr(f), A._ptr(flmn_h), s)
        fwd(A._ptr(back), A._ptr(f), s)
        times.append(time.perf_counter() - t0)
    ms = 1e3 * float(np.median(times[1:]))
    moved = 2 * (f.nbytes + flmn_h.nbytes)
    err = float(np.abs(back - flmn_h).max() / np.abs(flmn_h).max())
    return {"ms_per_pair": ms, "pairs_per_s": 1000.0 / ms, "pcie_bytes_per_pair": moved,
            "pcie_gb_s_if_copy_bound": moved / ms / 1e6, "roundtrip_rel_err": err,
            "note": "pageable host arrays; inputs copied to HBM and outputs back inside each call"}


def config0_bench(A, P, torch, reps=10):
    """Configuration 0, the reference harness's case (src/c/so3_test_csv.c:
    132-235): complex inverse then forward at L = N = 16, MW, padded,
    zero-first, all n, Risbo, seed 0.  Average and minimum wall time over
    `reps` repeats and the maximum absolute round-trip error over flmn_size
    entries, on the CPU restatement (single thread) and on the GPU through the
    drop-in host-pointer entry points (copies included, as a C caller sees)."""
    from so3_amd import synth
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    p = P.SO3Parameters(L=16, N=16, n_order=P.SO3_N_ORDER_ZERO_FIRST)
    flmn = synth.gen_flmn(p, 0)
    out = {}
    for name, inv, fwd in (("cpu", pyoracle.inverse, pyoracle.forward), ("gpu", A.inverse, A.forward)):
        ts, err = [], 0.0
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            back = fwd(inv(flmn, p), p)
            ts.append(time.perf_counter() - t0)
            err = max(err, float(np.abs(back - flmn).max()))
        ts = np.array(ts[1:]) * 1e3
        out[name] = {"avg_ms": float(ts.mean()), "min_ms": float(ts.min()), "max_abs_err": err}
    out["cpu"]["cores"] = 1
    out["cpu"]["kind"] = "port"
    out["workload"] = "L=N=16 complex MW, padded, zero-first, all n, Risbo, seed 0; inverse then forward; " \
                      "%d repeats (one warm-up)" % reps
    return out


def error_bounds(A, P, torch):
    """Round-trip error per configuration: the CPU restatement's own round
    trip (inverse then forward, max |error| / max |flmn|) at the sizes it
    finishes in seconds (L = N = 16 and 64, Risbo and Trapani), the GPU's at
    the same sizes, the GPU-vs-CPU inverse difference there, and the GPU's
    round trip at the benchmark sizes.  The bar: GPU error within 10x the
    CPU restatement's round-trip error at the same L."""
    from so3_amd import synth
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    z = P.SO3_N_ORDER_ZERO_FIRST
    out = {"same_L": {}, "gpu_only": {}}
    for L in (16, 64):
        for dl, dn in ((0, "risbo"), (1, "trapani")):
            p = P.SO3Parameters(L=L, N=L, dl_method=dl, n_order=z)
            x = synth.gen_flmn(p, 1)
            scale = np.abs(x).max()
            fc = pyoracle.inverse(x, p)
            e_cpu = float(np.abs(pyoracle.forward(fc, p) - x).max() / scale)
            fg = A.inverse(x, p)
            e_gpu = float(np.abs(A.forward(fg, p) - x).max() / scale)
            d_inv = float(np.abs(fg - fc).max() / np.abs(fc).max())
            out["same_L"]["L=N=%d %s" % (L, dn)] = {
                "cpu_roundtrip_rel": e_cpu, "gpu_roundtrip_rel": e_gpu, "gpu_vs_cpu_inverse_rel": d_inv,
                "gpu_over_cpu": e_gpu / e_cpu, "within_10x": bool(e_gpu <= 10 * e_cpu and d_inv <= 10 * e_cpu)}
    for name, p in (("C1 L=N=256 risbo", P.SO3Parameters(L=256, N=256, n_order=z)),
                    ("C1 L=N=256 trapani", P.SO3Parameters(L=256, N=256, dl_method=1, n_order=z)),
                    ("C1 L=N=256 MW_SS", P.SO3Parameters(L=256, N=256, sampling_scheme=1, n_order=z)),
                    ("C1 L=N=256 real", P.SO3Parameters(L=256, N=256, reality=1, n_order=z)),
                    ("C2b L=1024 N=4", P.SO3Parameters(L=1024, N=4, n_order=z))):
        flmn = torch.from_numpy(synth.gen_flmn(p, 1)).to("cuda:0")
        back = torch.zeros_like(flmn)
        f = torch.empty(A.f_size(p), dtype=torch.float64 if p.reality else torch.complex128, device="cuda:0")
        A.inverse_device(f.data_ptr(), flmn.data_ptr(), p)
        A.forward_device(back.data_ptr(), f.data_ptr(), p)
        torch.cuda.synchronize()
        out["gpu_only"][name] = {"gpu_roundtrip_rel": (back - flmn).abs().max().item() / flmn.abs().max().item()}
        del flmn, back, f
    out["note"] = ("the CPU restatement cannot run L = 256 within the benchmark (about 4 minutes per pair); "
                   "its round trip grows with L (Risbo faster than Trapani: the Risbo table's orthogonality "
                   "defect grows linearly in l, 3.4e-14 at l = 255 against 6e-15 for Trapani)")
    return out


def cpu_baseline_mt(p, per_thread=3):
    """The same oracle work spread over the host's cores: one thread per core
    (ctypes releases the GIL; every oracle call has its own workspace), each
    timing `per_thread` n slices; throughput = slices / wall time, scaled to
    all 2N-1 slices, plus the Delta table (built once) and the gamma DFTs
    divided over the threads."""
    import concurrent.futures
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    N = p.N
    # one GPU's share of the host is 16 cores on the benchmark machines
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    ns_all = np.unique(np.linspace(-(N - 1), N - 1, threads * per_thread).round().astype(int))
    chunks = [ns_all[i::threads] for i in range(threads)]
    stride = (2 * p.L - 1) * p.L
    t_dl, t_g, _ = 0.0, 0.0, None
    t_dl_one = pyoracle.time_sample(p, ns_all[:1], gamma_every=16)
    t_dl, t_g = t_dl_one[1], t_dl_one[2]
    t0 = time.time()
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        res = list(ex.map(lambda c: pyoracle.time_sample(p, c, gamma_every=stride), chunks))
    wall = time.time() - t0
    # each thread also rebuilt the Delta table: take it out of the wall time
    slices_wall = max(0.0, wall - max(r[1] for r in res))
    per_slice = slices_wall / max(1, max(len(c) for c in chunks))
    total = t_dl + per_slice * (2 * N - 1) / threads + 2 * t_g / threads
    return {
        "value": 1.0 / total,
        "unit": "fwd+inv transforms/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d n-slices over %d threads (%d each) at L=N=%d, %.1f s wall; Delta table once, gamma DFTs "
                  "divided over the threads; estimated %.1f s per fwd+inv" % (len(ns_all), threads, per_thread,
                                                                              p.L, wall, total),
    }


_T0 = time.time()


def progress(msg):
    """one line on stderr per finished leg (the JSON line goes to stdout)"""
    print("bench: %s (%.1f s)" % (msg, time.time() - _T0), file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--L", type=int, default=256)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--dl", type=int, default=0, help="0 Risbo, 1 Trapani")
    ap.add_argument("--sampling", type=int, default=0, help="0 MW, 1 MW_SS")
    ap.add_argument("--real", type=int, default=0)
    ap.add_argument("--cpu-slices", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with a 1 GiB device-to-device copy (reference kernel for HBM counters)")
    ap.add_argument("--batched-plain", metavar="TABLE",
                    help="only time the 96-item batched workload as plain calls, with the autotuner's "
                         "table read from TABLE (made and saved there first when absent), so that a "
                         "profiler run of this mode holds the workload's kernels and not the tuner's")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one GPU's "
              "throughput is the reported value).", file=sys.stderr)
        sys.exit(2)

    import torch
    import so3_amd.api as A
    import so3_amd.params as P
    from so3_amd import synth

    if args.batched_plain:
        print(json.dumps(batched_plain_only(A, P, torch, args.batched_plain)), flush=True)
        return

    p = P.SO3Parameters(L=args.L, N=args.N, sampling_scheme=args.sampling, reality=args.real,
                        dl_method=args.dl, n_order=P.SO3_N_ORDER_ZERO_FIRST,
                        storage=P.SO3_STORAGE_PADDED, n_mode=P.N_MODE_ALL)
    dev = torch.device("cuda:0")
    flmn_h = synth.gen_flmn(p, 1)
    flmn = torch.from_numpy(flmn_h).to(dev)
    back = torch.zeros_like(flmn)
    fdt = torch.float64 if args.real else torch.complex128
    f = torch.empty(A.f_size(p), dtype=fdt, device=dev)
    stream = torch.cuda.current_stream()
    sid = stream.cuda_stream
    A.prepare(p)
    torch.cuda.synchronize()

    def step():
        A.inverse_device(f.data_ptr(), flmn.data_ptr(), p, A.VIA_SSHT, sid)
        A.forward_device(back.data_ptr(), f.data_ptr(), p, A.VIA_SSHT, sid)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for k in range(args.steps):
        step()
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([ev[k].elapsed_time(ev[k + 1]) for k in range(args.steps)])
    ms_total = ev[0].elapsed_time(ev[-1])
    ms_step = ms_total / args.steps
    err = (back - flmn).abs().max().item() / flmn.abs().max().item()

    # per-stage kernel times from the library's event profiler (extra steps,
    # outside the timed region)
    A.stage_timing(True)
    A.stage_reset()
    for _ in range(3):
        step()
    torch.cuda.synchronize()
    stages = A.stage_times()
    A.stage_timing(False)
    per = {k: v[0] / v[1] for k, v in stages.items() if v[1] > 0}

    hbm_prac = hbm_practical(torch, dev, stream)
    L, N = p.L, p.N
    T = contraction_terms(L, N, real=bool(args.real))
    fl_contract = 5.0 * T  # 1 mul (Delta x Delta) + 2 FMA (complex accumulate) per term
    model = stage_model(p, A, T)
    stage_roof = {}
    for k, (bound, amount) in model.items():
        if k not in per:
            continue
        if bound == "mfma":
            ach = amount / (per[k] * 1e-3) / 1e12
            stage_roof[k] = {"ms": round(per[k], 4), "bound": "mfma", "achieved": round(ach, 3),
                             "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4)}
        else:
            ach = amount / (per[k] * 1e-3) / 1e9
            stage_roof[k] = {"ms": round(per[k], 4), "bound": "hbm", "achieved": round(ach, 1),
                             "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                             "frac_of_streaming": round(ach / hbm_prac["gbs"], 4)}
    dom = max(stage_roof, key=lambda k: stage_roof[k]["ms"])
    bound, amount = model[dom]
    ach = stage_roof[dom]["achieved"]
    peak = FP64_PEAK_TFLOPS if bound == "mfma" else HBM_PEAK_GBS
    roofline = {"bound": bound, "kernel": STAGE_KERNEL[dom], "stage": dom, "achieved": ach, "peak": peak,
                "unit": stage_roof[dom]["unit"], "frac": ach / peak,
                "traffic": measured_traffic(STAGE_KERNEL[dom]) if bound == "hbm" else None,
                ("flops_per_launch" if bound == "mfma" else "bytes_per_launch"): amount}
    esz = 8 if args.real else 16
    bytes_alg = 2 * (16 * A.flmn_size(p) + esz * A.f_size(p))
    t_min = max(bytes_alg / (HBM_PEAK_GBS * 1e9), 2 * fl_contract / (FP64_PEAK_TFLOPS * 1e12))

    out = {
        "metric": "FP64 Wigner fwd+inv transforms/sec at L=N=256 + p99 latency, batched vs plain",
        "value": 1000.0 / ms_step,
        "unit": "fwd+inv transforms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded U[-1,1) flmn, resident in HBM)",
        "config": {"workload": "plain complex fwd+inv, L=N=%d, MW%s, %s Delta, padded, zero-first, "
                               "all n, via SSHT" % (L, "_SS" if args.sampling else "",
                                                    "Trapani" if args.dl else "Risbo"),
                   "L": L, "N": N, "reality": args.real, "path": "plain"},
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "wall_s": wall,
        "roundtrip_rel_err": err,
        "stage_ms": {k: round(v, 4) for k, v in per.items()},
        # the stage profiler records a pair of events around every stage, so
        # consecutive kernels no longer overlap at their boundaries: its sum
        # runs a few percent above the undisturbed step
        "stage_sum_ms": round(sum(per.values()), 4),
        "stage_sum_over_step": round(sum(per.values()) / ms_step, 4),
        "roofline": roofline,
        "stage_roofline": stage_roof,
        "hbm_practical": hbm_prac,
        "path_roofline": {"t_min_ms": t_min * 1e3, "frac": t_min * 1e3 / ms_step,
                          "bytes_alg": bytes_alg, "contraction_flops": 2 * fl_contract},
    }
    progress("plain fwd+inv at L=%d N=%d: %.3f ms per pair" % (L, N, ms_step))
    if not args.no_cpu:
        out["cpu_baseline_config0"] = config0_bench(A, P, torch)
        progress("configuration 0")
        out["error_bounds"] = error_bounds(A, P, torch)
        progress("error bounds")
    if not args.no_variants:
        out["variants"] = variants_bench(A, P, torch)
        progress("variants")
        out["host_abi"] = host_abi_bench(A, p, flmn_h)
        progress("host-pointer ABI")
    if not args.no_batch:
        out["batched"] = batched_bench(A, P, torch)
        progress("batched")
    if not args.no_serve:
        out["serving"] = serving_bench(A)
        progress("serving")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, args.cpu_slices)
        progress("cpu baseline")
        out["cpu_baseline_mt"] = cpu_baseline_mt(p)
        progress("cpu baseline, threads")
    print(json.dumps(out))
    if args.copy_calibration:
        a = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
