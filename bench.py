#!/usr/bin/env python3
"""Benchmark: photons/s of the photon transport + voxel dose path on one MI355X.

Default workload ("headline"): the metric's own comparison, learned vs plain GPU.
  value   photons/s of the learned exit-plane form: the 6x256 BF16 generator network
          (noise -> fused MLP forward -> decode) feeding exit-plane photons into the ct256
          phantom (256x256x128 HU phantom at 1 mm, 12 materials), timed per step of
          --ep-chunk photons (default 10 steps of 1e8 after one untimed warm-up step);
  config  the plain form on the same photons (upstream transport through the tungsten
          collimator + the same phantom transport), learned/plain ratio, the dose criterion
          (mean voxel-dose difference over voxels above 10 % of the maximum, and the
          fraction of voxels within 2 sigma), and two legs:
            ct256_plain  plain transport of the 120 kVp cone beam into ct256, 1e9 histories
                         per step, with the transport kernel's HBM and issue rooflines;
            config1      the 64^3 water + bone-sphere phantom, 140.5 keV point source,
                         1e6 histories per run, GPU against the CPU restatement;
  roofline      the dominant kernel of the learned step (the fused generator forward,
                MFMA-bound: FLOP per launch / its average duration from CUDA events on
                the stream it runs on, against 2.5 PFLOP/s dense BF16);
  cpu_baseline  the CPU restatement (oracle/transport_ref.c) on config 1, on the host's
                CPUs (all of them the process may use), for a bounded sample.

Other workloads (--workload): ct256 (plain transport only; roofline = transport kernel:
64 B per history + 1 B per Woodcock step + 32 B per (history, voxel) deposit, against
8 TB/s; traffic from rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a copy-kernel
calibration when profiles/transport_ct256_traffic.json holds them), batch64, exitplane,
spect16, mlp. Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "photons/s on 1 MI355X: learned vs plain GPU at ≤2% voxel dose diff (2σ)"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(c):
    return 64.0 * c["histories"] + 1.0 * c["steps"] + 32.0 * c["deposits"]


def host_threads():
    """CPUs this process may run on (the affinity mask: the host's CPU share)."""
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return max(1, os.cpu_count() or 1)


def cpu_baseline(tables, ph, src, tal, target_s=15.0, threads=None, what="the same workload"):
    """The CPU restatement on `threads` threads (default: every CPU of the affinity mask),
    repeated in chunks until about target_s seconds of work have run."""
    from oracle import oracle as O

    threads = threads or host_threads()
    chunk = 50_000 * threads
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
            "host_cpu_count": os.cpu_count(),
            "sample": f"{done} histories of {what} on {threads} threads in {spent:.1f} s "
                      "(CPU restatement oracle/transport_ref.c, not the reference Geant4)"}


VALU_PEAK_WAVE_INSTR_PER_S = 256 * 2.4e9  # 256 CUs, one wave64 VALU instruction per CU per cycle


def transport_profile_figures(n):
    """Measured HBM bytes and VALU wave-instructions per launch of the ct256 transport kernel
    (rocprofv3 counters, profiles/transport_ct256_traffic.json), if they were taken at n
    histories per launch."""
    tf = ROOT / "profiles" / "transport_ct256_traffic.json"
    if not tf.exists():
        return {}
    try:
        t = json.loads(tf.read_text())
    except Exception:
        return {}
    if int(t.get("histories_per_launch", -1)) != n:
        return {}
    return t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10,
                    help="timed steps (headline / exitplane: chunks of --ep-chunk photons; ct256: launches of "
                         "--histories)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--histories", type=float, default=1e9, help="ct256: histories per launch")
    ap.add_argument("--ct256-steps", type=int, default=5, help="headline: launches of the ct256 plain leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--refill", type=int, default=64)
    ap.add_argument("--slow-min", type=int, default=48)
    ap.add_argument("--refill-min", type=int, default=12)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--block", type=int, default=0,
                    help="ct256: workgroup size (0: automatic staging, all tables in LDS for one simulation)")
    ap.add_argument("--workload", default="headline",
                    choices=("headline", "ct256", "ct256_schneider", "batch64", "exitplane", "spect16", "mlp"))
    ap.add_argument("--config1-runs", type=int, default=20, help="headline: runs of 1e6 histories on config 1")
    ap.add_argument("--arf-train-photons", type=float, default=4e7,
                    help="spect16: photons through the head model that label the ARF training set")
    ap.add_argument("--arf-fit-steps", type=int, default=6000, help="spect16: ARF training steps")
    ap.add_argument("--spect-histories", type=float, default=1e8, help="spect16: histories per phantom")
    ap.add_argument("--ep-train-photons", type=float, default=1e8,
                    help="exitplane: recorded collimator photons used to train the generator")
    ap.add_argument("--ep-fit-steps", type=int, default=4000, help="exitplane: generator training steps")
    ap.add_argument("--ep-chunk", type=float, default=1e8, help="exitplane: photons per step (transport chunk)")
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
              "histories would shard across GPUs by ranges of history ids (the RNG counter)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "headline":
        return bench_headline(args)
    if args.workload == "batch64":
        return bench_batch(args)
    if args.workload == "exitplane":
        line = leg_exitplane(args)
        print(json.dumps(line), flush=True)
        return
    if args.workload == "spect16":
        return bench_spect(args)
    if args.workload == "mlp":
        return bench_mlp(args)
    if args.workload == "ct256_schneider":
        leg = leg_ct256_schneider(args, args.steps)
        print(json.dumps({"metric": METRIC, "value": leg["photons_per_s"], "unit": "photons/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": 1,
                          "ms_per_step": 1e3 * leg["histories_per_step"] / leg["photons_per_s"],
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
                          "data": "synthetic (seed-generated phantom and spectrum)", "config": leg}), flush=True)
        return
    leg = leg_ct256(args, args.steps, cpu=not args.no_cpu_baseline)
    line = {
        "metric": METRIC, "value": leg["photons_per_s"], "unit": "photons/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": leg["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seed-generated phantom and spectrum)",
        "config": {k: v for k, v in leg.items() if k not in ("roofline", "cpu_baseline", "roofline_issue")},
        "roofline": leg["roofline"],
    }
    if "roofline_issue" in leg:
        line["roofline_issue"] = leg["roofline_issue"]
    if "cpu_baseline" in leg:
        line["cpu_baseline"] = leg["cpu_baseline"]
    print(json.dumps(line), flush=True)


def _sync():
    try:
        import torch

        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def leg_ct256(args, steps, cpu=False):
    """Plain transport of the 120 kVp cone beam into ct256: `steps` launches of
    --histories histories after --warmup untimed ones; the transport kernel's rooflines."""
    from opengate_amd import workloads
    from opengate_amd.transport import GpuEngine

    eng = GpuEngine(0)
    tables, ph, src, tal, info = workloads.ct256(eng)
    eng.set_physics(tables)
    eng.add_simulation(ph, src, tal)
    n = int(args.histories)
    launch = {"block": args.block, "refill_interval": args.refill, "grid": args.grid, "slow_min": args.slow_min,
              "refill_min": args.refill_min}
    if args.autotune:
        launch, _ = eng.autotune([min(n, 20_000_000)])
    first = 0
    for _ in range(args.warmup):
        eng.run([n], 12345, first_history=[first], **launch)
        first += n
    _sync()
    t0 = time.perf_counter()
    counters = []
    for _ in range(steps):
        counters.append(eng.run([n], 12345, first_history=[first], **launch))
        first += n
    _sync()
    elapsed = time.perf_counter() - t0
    kernel_ms = sum(c["kernel_ms"] for c in counters) / len(counters)
    bytes_per_launch = sum(algorithmic_bytes(c) for c in counters) / len(counters)
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    prof = transport_profile_figures(n)
    c0 = counters[0]
    leg = {"workload": info["desc"], "histories_per_step": n, "steps": steps, "form": "plain GPU transport",
           "photons_per_s": steps * n / elapsed, "ms_per_step": elapsed / steps * 1e3,
           "steps_per_history": c0["steps"] / n, "deposits_per_history": c0["deposits"] / n, "launch": launch,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": prof.get("hbm_bytes_per_launch"),
                        "kernel_ms": kernel_ms, "algorithmic_bytes_per_launch": bytes_per_launch,
                        "fp64_atomics_per_s": 2.0 * c0["deposits"] / (kernel_ms * 1e-3)}}
    if prof.get("valu_wave_instructions_per_history"):
        wi = prof["valu_wave_instructions_per_history"] * n / (kernel_ms * 1e-3)
        leg["roofline_issue"] = {"bound": "valu_issue", "achieved": wi, "peak": VALU_PEAK_WAVE_INSTR_PER_S,
                                 "unit": "wave64 VALU instructions/s", "frac": wi / VALU_PEAK_WAVE_INSTR_PER_S,
                                 "valu_wave_instructions_per_history": prof["valu_wave_instructions_per_history"],
                                 "valu_lane_utilization": prof.get("valu_lane_utilization")}
    if cpu:
        leg["cpu_baseline"] = cpu_baseline(tables, ph, src, tal, what="ct256")
    eng.close()
    return leg


def leg_ct256_schneider(args, steps=2, tol=0.05):
    """ct256 through the reference's Schneider HU conversion (75 materials at 0.05 g/cm3,
    sharing 27 composition rows): plain transport, `steps` launches of --histories."""
    from opengate_amd import workloads
    from opengate_amd.transport import GpuEngine

    eng = GpuEngine(0)
    tables, ph, src, tal, info = workloads.ct256_schneider(eng, tol)
    eng.set_physics(tables)
    eng.add_simulation(ph, src, tal)
    n = int(args.histories)
    eng.run([n // 10], 7)
    _sync()
    t0 = time.perf_counter()
    cs = [eng.run([n], 12345, first_history=[k * n]) for k in range(steps)]
    _sync()
    elapsed = time.perf_counter() - t0
    kms = sum(c["kernel_ms"] for c in cs) / steps
    eng.close()
    return {"workload": info["desc"], "n_materials": info["n_materials"], "table_rows": info["n_rows"],
            "label_bits": info["label_bits"], "histories_per_step": n, "steps": steps,
            "photons_per_s": steps * n / elapsed, "kernel_ms": kms,
            "steps_per_history": cs[0]["steps"] / n}


def leg_config1(args, cpu=True):
    """Config 1: 64^3 water phantom with three bone spheres, 140.5 keV point source, runs of
    1e6 histories (the tally accumulating), GPU against the CPU restatement."""
    from opengate_amd import workloads
    from opengate_amd.transport import GpuEngine

    tables, ph, src, tal, info = workloads.water64()
    eng = GpuEngine(0)
    eng.set_physics(tables)
    eng.add_simulation(ph, src, tal)
    n = 1_000_000
    eng.run([n], 5)  # warm-up
    eng.reset_tally(0)
    _sync()
    runs = max(1, int(args.config1_runs))
    t0 = time.perf_counter()
    kms = 0.0
    for k in range(runs):
        c = eng.run([n], 2025, first_history=[k * n])
        kms += c["kernel_ms"]
    g = eng.get_tally(0)
    _sync()
    wall = time.perf_counter() - t0
    eng.close()
    leg = {"workload": info["desc"], "histories_per_run": n, "runs": runs,
           "photons_per_s": runs * n / wall, "ms_per_run": wall / runs * 1e3,
           "kernel_photons_per_s": runs * n / (kms * 1e-3),
           "timed": "runs of 1e6 histories plus the final FP64 tally copy to the host",
           "edep_total_mev": float(g["edep"].sum())}
    if cpu:
        cb = cpu_baseline(tables, ph, src, tal, what="config 1 (water64)")
        leg["cpu_baseline"] = cb
        leg["gpu_over_cpu"] = leg["photons_per_s"] / cb["value"]
        leg["target_gpu_over_cpu"] = 50.0
    return leg


def bench_headline(args):
    """The metric's comparison: learned exit-plane form vs plain GPU transport, with the
    ct256 plain leg (transport rooflines) and the config-1 leg (CPU baseline)."""
    t_start = time.perf_counter()
    ep = leg_exitplane(args)
    _log(f"exitplane leg done in {time.perf_counter() - t_start:.0f} s")
    ct = leg_ct256(args, args.ct256_steps, cpu=False)
    _log(f"ct256 leg done: {ct['photons_per_s']:.3e} photons/s")
    cs = leg_ct256_schneider(args)
    _log(f"ct256-schneider leg done: {cs['photons_per_s']:.3e} photons/s")
    c1 = leg_config1(args, cpu=not args.no_cpu_baseline)
    _log(f"config1 leg done: {c1['photons_per_s']:.3e} photons/s")
    cfg = dict(ep["config"])
    cfg["legs"] = {"ct256_plain": ct, "ct256_schneider": cs,
                   "config1": {k: v for k, v in c1.items() if k != "cpu_baseline"}}
    line = dict(ep)
    line["config"] = cfg
    line["roofline_transport"] = ct["roofline"]
    if "roofline_issue" in ct:
        line["roofline_transport_issue"] = ct["roofline_issue"]
    if "cpu_baseline" in c1:
        cb = dict(c1["cpu_baseline"])
        cb["workload"] = "config 1 (water64), compared with config1.photons_per_s"
        line["cpu_baseline"] = cb
        cfg["config1_gpu_over_cpu"] = c1["gpu_over_cpu"]
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


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _collimator_chunks(eng, n_photons, seed, chunk, hist_per_photon, launch=None, first_chunk=0):
    """Plain upstream transport through the collimator in chunks of about `chunk` exit photons
    until at least n_photons are recorded; yields (device records, count, seconds, histories)
    per chunk. Every recorded row of a chunk is kept: the recorder's slot order follows the
    order in which photons finish, so dropping rows from its end would bias the sample."""
    import torch

    got = 0
    k = first_chunk
    while got < n_photons:
        want = int(min(chunk, n_photons - got))
        hist = max(1000, int(round(want * hist_per_photon)))
        eng.reset_tally(0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        eng.run([hist], seed, first_history=[k * (1 << 40)], **(launch or {}))
        rec, n_rec = eng.get_phsp(0, device=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        if n_rec > len(rec[0]):
            raise RuntimeError(f"collimator recorder overflow: {n_rec} photons, capacity {len(rec[0])}")
        got += n_rec
        k += 1
        yield rec, n_rec, dt, hist


def leg_exitplane(args):
    """Learned exit-plane source vs plain transport through the collimator, both driving
    the same exit-plane photons into the ct256 phantom. A step is one chunk: the plain
    form records about --ep-chunk photons behind the collimator and transports them into
    the phantom; the learned form generates the same number with the network and
    transports those. --steps timed steps after --warmup untimed ones."""
    import torch

    from opengate_amd import workloads as WL
    from opengate_amd.learn.exitplane import ExitPlaneGenerator, compare_exit_planes
    from opengate_amd.transport import GpuEngine, SourceSpec

    chunk = int(args.ep_chunk)
    steps = max(1, int(args.steps))
    cap = int(chunk * 1.1) + 100_000
    Tc, phc, srcc, talc, infoc = WL.exitplane(capacity=cap)
    col = GpuEngine(0)
    col.set_physics(Tc)
    col.add_simulation(phc, srcc, talc)
    # histories per recorded photon from a short run
    c = col.run([2_000_000], 99)
    hpp = 2_000_000 / max(1, c["recorded"])
    col.reset_tally(0)
    # launch parameters of each stage from the autotuner, shared by the plain and learned
    # paths (trial histories leave tallies and recorders as they were)
    launch_col, _ = col.autotune([2_000_000])
    # ---- training set and fit
    n_train = int(args.ep_train_photons)
    room = n_train + chunk + 1
    train = [torch.empty(room, device="cuda") for _ in range(8)]
    off = 0
    t_train_rec = 0.0
    for rec, m, dt, _ in _collimator_chunks(col, n_train, 1, chunk, hpp, launch_col):
        for a, b in zip(train, rec):
            a[off:off + m] = b[:m]
        off += m
        t_train_rec += dt
        _log(f"training photons {off}/{n_train}")
    train = [a[:off] for a in train]
    gen = ExitPlaneGenerator(WL.COLLIMATOR_SOURCE, 0.0, seed=0)
    torch.cuda.synchronize()
    t = time.perf_counter()
    hist = gen.fit(train, steps=args.ep_fit_steps, batch=16384, projections=128, lr=1e-3, seed=2, log_every=500)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t
    _log(f"fit {args.ep_fit_steps} steps on {off} photons in {t_fit:.1f} s, final loss "
         f"{float(np.mean(hist[-50:])):.5f}")
    del train
    # ---- phantom stage (ct256), one context
    ct = GpuEngine(0)
    Tp, php, srcp, talp, infop = WL.ct256(ct)
    ct.set_physics(Tp)

    def phsp_source(rec):
        return SourceSpec(pos_type="phsp", phsp=rec, pos_trans=WL.COLLIMATOR_TO_CT256)

    # ---- plain: upstream transport + phantom; warm-up chunks first (untimed, tally reset)
    t_up = t_ph = 0.0
    done = 0
    counts = []
    holdout = None
    launch_ct = None
    n_warm = max(0, int(args.warmup))
    for k, (rec, m, dt, _) in enumerate(_collimator_chunks(col, (steps + n_warm) * chunk, 7, chunk, hpp,
                                                           launch_col)):
        if not ct.sims:
            ct.add_simulation(php, phsp_source([a[:m] for a in rec]), talp)
            launch_ct, _ = ct.autotune([min(m, 20_000_000) // 10])
        else:
            ct.set_source(0, phsp_source([a[:m] for a in rec]))
        torch.cuda.synchronize()
        t = time.perf_counter()
        ct.run([m], 31, first_history=[done], **launch_ct)
        torch.cuda.synchronize()
        dph = time.perf_counter() - t
        if k < n_warm:
            ct.reset_tally(0)
            _log(f"plain warm-up chunk {m}")
            continue
        if holdout is None:
            holdout = [a[:4_000_000].clone() for a in rec]
        t_ph += dph
        t_up += dt
        done += m
        counts.append(m)
        _log(f"plain {done} photons")
    plain = ct.get_tally(0)
    n_total = done
    # ---- learned: generator + phantom, the same photon count per chunk
    ct.reset_tally(0)
    buf = [torch.empty(max(counts), device="cuda") for _ in range(8)]
    events = []
    for k in range(n_warm):
        rec = gen.generate(counts[0], seed=55, first=k * counts[0], out=[b[:counts[0]] for b in buf])
        ct.set_source(0, phsp_source(rec))
        ct.run([counts[0]], 33, first_history=[k * counts[0]], **launch_ct)
    ct.reset_tally(0)
    torch.cuda.synchronize()
    t_gen = t_ph2 = 0.0
    done = 0
    first_rec = None
    ph_counters = []
    for m in counts:
        torch.cuda.synchronize()
        t = time.perf_counter()
        rec = gen.generate(m, seed=5, first=done, out=[b[:m] for b in buf], events=events)
        torch.cuda.synchronize()
        t_gen += time.perf_counter() - t
        if first_rec is None:
            first_rec = [a[:4_000_000].clone() for a in rec]
        ct.set_source(0, phsp_source(rec))
        torch.cuda.synchronize()
        t = time.perf_counter()
        ph_counters.append(ct.run([m], 32, first_history=[done], **launch_ct))
        torch.cuda.synchronize()
        t_ph2 += time.perf_counter() - t
        done += m
        _log(f"learned {done}/{n_total}")
    learned = ct.get_tally(0)
    plane = compare_exit_planes(first_rec, holdout, distance=393.5)
    dose = compare_dose(learned["edep"], plain["edep"], test2=learned["edep2"], ref2=plain["edep2"],
                        n_test=n_total, n_ref=n_total)
    gen_rate = n_total / t_gen
    fwd_ms = sum(a.elapsed_time(b) for a, b in events)
    fwd_launches = len(events)
    fwd_flops = gen.flops_per_photon() * n_total
    fwd_tflops = fwd_flops / (fwd_ms * 1e-3) / 1e12
    ph_kms = sum(c["kernel_ms"] for c in ph_counters)
    ph_bytes = sum(algorithmic_bytes(c) for c in ph_counters)
    learned_s = t_gen + t_ph2
    plain_s = t_up + t_ph
    criterion = {
        "region": "voxels above 10 % of the plain maximum",
        "mean_voxel_dose_diff": dose["voxel_mean_rel_diff"],
        "limit": 0.02,
        "pass_mean_within_2pct": abs(dose["voxel_mean_rel_diff"]) <= 0.02,
        "frac_within_2sigma": dose.get("voxel_frac_within_2sigma"),
        "frac_within_2sigma_if_same_distribution": 0.9545,
        "mean_z2": dose.get("voxel_mean_z2"),
    }
    line = {
        "metric": METRIC, "value": n_total / learned_s, "unit": "photons/s", "n_gpus": 1, "steps": len(counts),
        "warmup": n_warm, "ms_per_step": learned_s / len(counts) * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 generator (fp32 accumulate) + fp32 transport",
        "data": "synthetic (seed-generated collimator and phantom, recorded training photons)",
        "config": {"workload": infoc["desc"] + " -> " + infop["desc"], "exit_photons": n_total,
                   "photons_per_step": counts,
                   "form": "learned exit-plane source (6x256 MLP) driving the ct256 phantom, vs plain transport "
                           "through the collimator driving the same phantom",
                   "learned_photons_per_s": n_total / learned_s,
                   "plain_photons_per_s": n_total / plain_s,
                   "learned_over_plain": plain_s / learned_s,
                   "target_learned_over_plain": 3.0,
                   "dose_criterion": criterion,
                   "training_photons": off, "fit_steps": args.ep_fit_steps, "fit_s": t_fit,
                   "training": dict(gen.fit_stats, frac_of_bf16_peak=gen.fit_stats["gemm_tflops_over_loop"] / 2500.0),
                   "training_record_s": t_train_rec,
                   "learned_over_plain_with_training": plain_s / (learned_s + t_fit + t_train_rec),
                   "generator_photons_per_s": gen_rate,
                   "plain_upstream_photons_per_s": n_total / t_up,
                   "phantom_transport_photons_per_s": n_total / t_ph2,
                   "gain_upstream": t_up / t_gen,
                   "collimator_histories_per_exit_photon": hpp,
                   "launch": {"collimator": launch_col, "phantom": launch_ct},
                   "exit_plane": plane, "dose": dose},
        "roofline": {"bound": "mfma", "kernel": "mlp_forward_kernel (fused 6x256 generator forward)",
                     "achieved": fwd_tflops, "peak": 2500.0, "unit": "TFLOP/s", "frac": fwd_tflops / 2500.0,
                     "traffic": None, "kernel_ms_per_launch": fwd_ms / max(1, fwd_launches),
                     "flop_per_photon": gen.flops_per_photon(), "launches": fwd_launches},
        "roofline_phantom_transport": {"bound": "hbm", "achieved": ph_bytes / (ph_kms * 1e-3) / 1e9,
                                       "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                       "frac": ph_bytes / (ph_kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                       "kernel_ms_per_launch": ph_kms / len(ph_counters)},
        "time_split_s": {"generator": t_gen, "generator_forward_kernels": fwd_ms * 1e-3,
                         "phantom_transport_learned": t_ph2, "plain_upstream": t_up, "plain_phantom": t_ph},
    }
    col.close()
    ct.close()
    return line


def bench_spect(args):
    """16 phantoms in one batched launch, their escaping photons projected on 128 heads by
    the learned detector response and by plain transport through the head model."""
    import math

    import torch

    from opengate_amd import workloads as WL
    from opengate_amd.learn.arf import ArfModel, Heads, HeadSimulator, compare_projections, project
    from opengate_amd.transport import GpuEngine

    theta_max = math.radians(10.0)
    head = HeadSimulator()
    n_train = int(args.arf_train_photons)
    torch.cuda.synchronize()
    t = time.perf_counter()
    rec, E, th, ph = head.training_photons(n_train, 1, theta_max)
    cls = head.windows(rec, n_train, 3)
    torch.cuda.synchronize()
    t_label = time.perf_counter() - t
    del rec
    arf = ArfModel(theta_max)
    X = arf.inputs(E, th, ph)
    del E, th, ph
    t = time.perf_counter()
    arf.fit(X, cls, steps=args.arf_fit_steps, batch=32768, seed=2, log_every=1000)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t
    del X, cls
    _log(f"arf: labelled {n_train} photons in {t_label:.1f} s, fit in {t_fit:.1f} s")
    tables, sims, info = WL.spect16(int(args.spect_histories))
    eng = GpuEngine(0)
    eng.set_physics(tables)
    for ph_, src, tal, _, _ in sims:
        eng.add_simulation(ph_, src, tal)
    hist = [h for _, _, _, h, _ in sims]
    # launch parameters of the batched phantom launch from the autotuner (trial histories
    # leave tallies and recorders as they were), as in the batch64 leg
    t = time.perf_counter()
    launch, _ = eng.autotune([max(1, h // 50) for h in hist])
    t_tune = time.perf_counter() - t
    torch.cuda.synchronize()
    t = time.perf_counter()
    c = eng.run(hist, 4242, **launch)
    torch.cuda.synchronize()
    t_transport = time.perf_counter() - t
    _log(f"phantoms: {sum(hist)} histories in {t_transport:.2f} s")
    t_pairs = t_arf = t_plain = 0.0
    pairs = 0
    per = []
    for k, (_, _, _, h, hk) in enumerate(sims):
        esc, n_esc = eng.get_phsp(k, device=True)
        heads = Heads(theta_max=theta_max, **hk)
        res = project(esc, min(n_esc, len(esc[0])), heads, arf=arf, head_sim=head, seed=100 + k)
        t_pairs += res["t_pairs"]
        t_arf += res["t_arf"]
        t_plain += res["t_plain"]
        pairs += res["pairs"]
        cmp = compare_projections(res["arf"].cpu().numpy(), res["plain"].cpu().numpy(), res["arf2"].cpu().numpy(),
                                  res["plain2"].cpu().numpy(), block=4)
        per.append(cmp)
        del esc, res
        _log(f"phantom {k}: {cmp}")
    total = sum(hist)
    ratios = np.array([p["window_ratio"] for p in per])
    chi2 = np.array([p["chi2_per_pixel"] for p in per if p.get("chi2_per_pixel") is not None])
    learned_s = t_transport + t_pairs + t_arf
    plain_s = t_transport + t_pairs + t_plain
    line = {
        "metric": METRIC, "value": total / learned_s, "unit": "photons/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": learned_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 transport + bf16 response network (fp32 accumulate)",
        "data": "synthetic (seed-generated phantoms and head model)",
        "config": {"workload": info["desc"], "histories_total": total,
                   "form": "batched phantom transport + learned detector response (4x128 MLP) on 128 heads",
                   "arf_training_photons": n_train, "arf_label_s": t_label, "arf_fit_s": t_fit,
                   "phantom_transport_photons_per_s": total / t_transport,
                   "phantom_launch": launch, "autotune_s": t_tune,
                   "pairs": pairs, "pairs_per_s": pairs / t_pairs,
                   "arf_pairs_per_s": pairs / t_arf, "plain_pairs_per_s": pairs / t_plain,
                   "arf_tflops": arf.flops_per_pair() * pairs / t_arf / 1e12,
                   "plain_photons_per_s": total / plain_s, "gain_detector": t_plain / t_arf,
                   "gain_end_to_end": plain_s / learned_s,
                   "window_ratio_mean": ratios.mean(0).tolist(), "window_ratio_min": ratios.min(0).tolist(),
                   "window_ratio_max": ratios.max(0).tolist(),
                   "chi2_per_pixel_median": float(np.median(chi2)) if len(chi2) else None,
                   "chi2_per_pixel_max": float(chi2.max()) if len(chi2) else None},
    }
    head.close()
    eng.close()
    print(json.dumps(line), flush=True)


def bench_mlp(args):
    """Inference throughput of the two networks (random weights, the sizes used by the
    learned components): fused single-kernel forward vs the layer-by-layer GEMMs, in
    TFLOP/s (2 flops per multiply-add of the unpadded layers) against the 2.5 PFLOP/s
    dense BF16 peak."""
    import torch

    from opengate_amd.learn.mlp import MLP

    out = {}
    for name, sizes, M in (("generator_6x256", [8, 256, 256, 256, 256, 256, 6], 1 << 23),
                           ("arf_4x128", [3, 128, 128, 128, 4], 1 << 23)):
        net = MLP(sizes, seed=1)
        X = torch.randn(M, net.psizes[0], device="cuda").to(torch.bfloat16)
        Z = torch.empty(M, net.psizes[-1], device="cuda")
        bufs = [torch.empty(M, max(net.psizes[1:-1]), device="cuda", dtype=torch.bfloat16) for _ in range(2)]
        flops = net.flops_forward(M)
        res = {}
        for form, fn in (("fused", lambda: net.infer_fused(X, Z)), ("layered", lambda: net.infer(X, Z, bufs))):
            fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(args.steps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / args.steps
            res[form] = {"ms": ms, "rows_per_s": M / (ms * 1e-3), "tflops": flops / (ms * 1e-3) / 1e12,
                         "frac_of_bf16_peak": flops / (ms * 1e-3) / 2.5e15}
        res["rows"] = M
        out[name] = res
    g = out["generator_6x256"]["fused"]
    line = {
        "metric": METRIC, "value": g["rows_per_s"], "unit": "photons/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": g["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (fp32 accumulate)", "data": "synthetic (random weights and inputs)",
        "config": {"workload": "MLP inference: exit-plane generator 8->256x5->6 and detector response "
                               "3->128x3->4, 2^23 rows per call", "form": "gh_mlp_forward_bf16 vs gh_gemm_bf16",
                   "networks": out},
        "roofline": {"bound": "mfma", "achieved": g["tflops"], "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": g["tflops"] / 2500.0, "traffic": None},
    }
    print(json.dumps(line), flush=True)


def compare_dose(test, ref, block=4, test2=None, ref2=None, n_test=None, n_ref=None):
    """Dose agreement. Voxel level, over voxels above 10 % of the reference maximum: mean
    signed and mean absolute relative difference, and (with the squared sums and history
    counts) the fraction of voxels within max(2 %, 2 sigma combined). Block level
    (block^3 voxels, blocks above 20 % of the maximum): integral ratio, median / 95th
    percentile relative difference; central-axis depth dose (20 x 20 voxel column) worst
    relative difference."""
    out = {}
    hi = ref > 0.1 * ref.max()
    rel = (test[hi] - ref[hi]) / ref[hi]
    out["voxels_above_10pct"] = int(hi.sum())
    out["voxel_mean_rel_diff"] = float(rel.mean())
    out["voxel_mean_abs_rel_diff"] = float(np.abs(rel).mean())
    if test2 is not None and ref2 is not None and n_test and n_ref:
        def var_of_sum(s, s2, n):
            return np.maximum(s2 - s * s / n, 0.0) * n / max(n - 1, 1)
        sig = np.sqrt(var_of_sum(test[hi], test2[hi], n_test) + var_of_sum(ref[hi], ref2[hi], n_ref))
        d = np.abs(test[hi] - ref[hi])
        out["voxel_frac_within_2pct_or_2sigma"] = float((d <= np.maximum(0.02 * ref[hi], 2.0 * sig)).mean())
        out["voxel_frac_within_2sigma"] = float((d <= 2.0 * sig).mean())
        good = sig > 0
        out["voxel_mean_z2"] = float(np.mean((d[good] / sig[good]) ** 2))
        out["voxel_mean_rel_sigma"] = float(np.mean(sig / ref[hi]))
    nz, ny, nx = ref.shape
    b = block
    crop = lambda d: d[: nz // b * b, : ny // b * b, : nx // b * b]  # noqa: E731
    red = lambda d: crop(d).reshape(nz // b, b, ny // b, b, nx // b, b).sum((1, 3, 5))  # noqa: E731
    R_, T_ = red(ref), red(test)
    core = R_ > 0.2 * R_.max()
    rel = np.abs(T_[core] - R_[core]) / R_[core]
    cy, cx = ny // 2, nx // 2
    col = lambda d: d[:, cy - 10:cy + 10, cx - 10:cx + 10].sum((1, 2))  # noqa: E731
    dr, dt = col(ref), col(test)
    m = dr > 0.2 * dr.max()
    out.update({"block_voxels": b, "core_blocks": int(core.sum()),
                "integral_ratio": float(T_[core].sum() / R_[core].sum()),
                "block_median_rel_diff": float(np.median(rel)), "block_p95_rel_diff": float(np.percentile(rel, 95)),
                "depth_dose_max_rel_diff": float(np.max(np.abs(dt[m] - dr[m]) / dr[m])),
                "total_ratio": float(test.sum() / ref.sum())})
    return out


if __name__ == "__main__":
    main()
