#!/usr/bin/env python3
"""Benchmark: photons/s of the photon transport + voxel dose path on one MI355X.

Workload: ct256 (256x256x128 HU phantom at 1 mm, 12 materials, 120 kVp Kramers cone
beam from 600 mm, dose on the phantom grid). One step = one persistent launch of
--histories histories (default 1e9) with inputs resident in HBM; the tally keeps
accumulating across steps (as a run of the reference accumulates events).

Prints ONE JSON line. Fields beyond the common contract:
  roofline: algorithmic HBM bytes of the transport kernel per launch
            (64 B per history + 1 B per Woodcock step + 32 B per (history, voxel)
            deposit) / kernel time from HIP events on the library's stream, against
            8 TB/s; traffic = measured HBM bytes per launch (rocprofv3 FETCH_SIZE +
            WRITE_SIZE, scaled by the copy-kernel calibration) when
            profiles/transport_ct256_traffic.json holds them, else null;
  cpu_baseline: the CPU restatement (oracle/transport_ref.c) on the host cores for a
            bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "photons/s on 1 MI355X: learned vs plain GPU at ≤2% voxel dose diff (2σ)"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(c):
    return 64.0 * c["histories"] + 1.0 * c["steps"] + 32.0 * c["deposits"]


def cpu_baseline(tables, ph, src, tal, target_s=15.0):
    """The CPU restatement on min(16, cores) threads, repeated in chunks until about
    target_s seconds of work have run."""
    from oracle import oracle as O

    threads = max(1, min(16, os.cpu_count() or 1))
    chunk = 200_000 * threads
    O.run(tables, ph, src, tal, 1000, seed=98, threads=1)  # load the library, touch the tables
    done, spent, first = 0, 0.0, 0
    while spent < target_s:
        t = time.perf_counter()
        O.run(tables, ph, src, tal, chunk, seed=100, first=first, threads=threads)
        dt = time.perf_counter() - t
        spent += dt
        done += chunk
        first += chunk
        if dt < target_s / 8:
            chunk *= 2
    return {"value": done / spent, "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": f"{done} histories of the same workload on {threads} threads in {spent:.1f} s "
                      "(CPU restatement oracle/transport_ref.c, not the reference Geant4)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--histories", type=float, default=1e9)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--refill", type=int, default=64)
    ap.add_argument("--slow-min", type=int, default=16)
    ap.add_argument("--refill-min", type=int, default=16)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--workload", default="ct256", choices=("ct256", "batch64"))
    ap.add_argument("--autotune", action="store_true", help="pick launch parameters with gh_autotune first")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy x5 (reference kernel for rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        x = torch.ones(256 * 2**20, device="cuda")  # 1 GiB of float32
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read_per_launch": x.numel() * 4, "copy_bytes_written_per_launch": y.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "histories shard across GPUs by RNG key range, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    from opengate_amd import workloads
    from opengate_amd.transport import GpuEngine

    if args.workload == "batch64":
        return bench_batch(args)
    eng = GpuEngine(0)
    tables, ph, src, tal, info = workloads.ct256(eng)
    eng.set_physics(tables)
    eng.add_simulation(ph, src, tal)
    n = int(args.histories)
    launch = {"refill_interval": args.refill, "grid": args.grid, "slow_min": args.slow_min,
              "refill_min": args.refill_min}
    if args.autotune:
        launch, _ = eng.autotune([min(n, 20_000_000)])
    try:
        import torch

        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731

    first = 0
    for _ in range(args.warmup):
        eng.run([n], 12345, first_history=[first], **launch)
        first += n
    sync()
    t0 = time.perf_counter()
    counters = []
    for _ in range(args.steps):
        counters.append(eng.run([n], 12345, first_history=[first], **launch))
        first += n
    sync()
    elapsed = time.perf_counter() - t0
    value = args.steps * n / elapsed
    kernel_ms = sum(c["kernel_ms"] for c in counters) / len(counters)
    bytes_per_launch = sum(algorithmic_bytes(c) for c in counters) / len(counters)
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    traffic = None
    tf = ROOT / "profiles" / "transport_ct256_traffic.json"
    if tf.exists():
        try:
            t = json.loads(tf.read_text())
            if int(t.get("histories_per_launch", -1)) == n:
                traffic = t.get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    c0 = counters[0]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seed-generated phantom and spectrum)",
        "config": {"workload": info["desc"], "histories_per_step": n, "form": "plain GPU transport",
                   "steps_per_history": c0["steps"] / n, "deposits_per_history": c0["deposits"] / n,
                   "launch": launch},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel_ms": kernel_ms, "algorithmic_bytes_per_launch": bytes_per_launch,
                     "fp64_atomics_per_s": 2.0 * c0["deposits"] / (kernel_ms * 1e-3)},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(tables, ph, src, tal)
    eng.close()
    print(json.dumps(line), flush=True)


def bench_batch(args):
    """64 simulations: one persistent batched launch vs one launch per simulation."""
    from opengate_amd import workloads
    from opengate_amd.transport import GpuEngine

    tables, sims, info = workloads.batch64()
    hist = [h for *_, h in sims]
    total = sum(hist)
    # plain form: each simulation alone in the single-simulation kernel
    eng = GpuEngine(0)
    eng.set_physics(tables)
    plain_ms = 0.0
    for ph, src, tal, h in sims:
        eng.clear()
        eng.add_simulation(ph, src, tal)
        c = eng.run([h], 777)
        plain_ms += c["kernel_ms"]
    # batched form: all simulations in one persistent launch, launch parameters autotuned
    eng.clear()
    for ph, src, tal, _ in sims:
        eng.add_simulation(ph, src, tal)
    launch, _ = eng.autotune([max(1, h // 50) for h in hist])
    eng.run([max(1, h // 50) for h in hist], 1, **launch)  # warm-up
    for k in range(len(sims)):
        eng.reset_tally(k)
    t0 = time.perf_counter()
    c = eng.run(hist, 777, **launch)
    wall = time.perf_counter() - t0
    line = {
        "metric": METRIC, "value": total / wall, "unit": "photons/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic (seed-generated phantoms)",
        "config": {"workload": info["desc"], "histories_total": total, "form": "batched persistent launch",
                   "launch": launch, "plain_separate_launches_photons_per_s": total / (plain_ms * 1e-3),
                   "batched_kernel_photons_per_s": total / (c["kernel_ms"] * 1e-3),
                   "batched_over_plain": (plain_ms / c["kernel_ms"])},
    }
    eng.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
