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
(timed live with HIP events by the library's stage profiler) and the CPU
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
    ap.add_argument("--cpu-slices", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(see DESIGN.md for how the transform would shard over n).", file=sys.stderr)
        sys.exit(2)

    import torch
    import so3_amd.api as A
    import so3_amd.params as P
    from so3_amd import synth

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

    L, N = p.L, p.N
    T = contraction_terms(L, N, real=bool(args.real))
    fl_contract = 5.0 * T  # 1 mul (Delta x Delta) + 2 FMA (complex accumulate) per term
    dom = max(("contract_inv", "contract_fwd"), key=lambda k: per.get(k, 0.0))
    achieved = fl_contract / (per[dom] * 1e-3) / 1e12
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
        "roofline": {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "flops_per_launch": fl_contract},
        "path_roofline": {"t_min_ms": t_min * 1e3, "frac": t_min * 1e3 / ms_step,
                          "bytes_alg": bytes_alg, "contraction_flops": 2 * fl_contract},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, args.cpu_slices)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
