#!/usr/bin/env python
"""FM-DBZP search throughput on one MI355X.

One step = one `acquire_` over the benchmark workload: GPS L1 C/A, 16.368 MHz
real int16 IF samples at 4.092 MHz, 80 ms = 8 noncoherent segments of 10 ms
coherent integration, PRNs 1..32 searched (8 present, CN0 30-40 dB-Hz, 50 bps
data), Doppler +-7 kHz (160 bins of 100 Hz), bit-edge search over 10
alignments. The samples are resident in HBM before the timed region; the
plan (code FFTs, tables, workspace) is built beforehand, as the reference's
benchmarks do (benchmark/benchmarks.jl:120-124).

Prints ONE JSON line. value = search cells/s = sum over PRNs of
get_num_cells (ndop x spc x R) per step / wall time per step.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "acquisition.jl_amd"))

METRIC = "search cells/s + PRN acqs/s on 1 GPU, mixed-prec vs FP32 at equal peak bins"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--edge-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-prns", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the engine runs on a single MI355X; see the sharding notes in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from acquisition import GPSL1CA, acquire_, is_detected, plan_acquire
    from acquisition.device import DeviceArray
    from acquisition.signals import l1ca_if_int16_scene

    t0 = time.time()
    sc = l1ca_if_int16_scene(seed=0x5EED0002)
    log(f"scene: {sc.signal.shape[0]} int16 samples, present {sc.present} ({time.time() - t0:.1f}s)")
    prns = list(range(1, 33))
    plan = plan_acquire(sc.system, sc.sampling_freq, prns, num_coherently_integrated_code_periods=10,
                        bit_edge_search_steps=args.edge_steps, num_noncoherent_accumulations=8,
                        min_doppler_coverage=7000.0)
    dev = DeviceArray(sc.signal)
    ndop, spc, R = plan.num_doppler_bins, plan.samples_per_code, plan.num_secondary_rotations
    cells = len(prns) * ndop * spc * R
    log(f"plan: spc {spc} nb {plan.num_blocks} bs {plan.block_size} F {plan.fft_size} ndop {ndop} "
        f"path {plan.path} E {plan.bit_edge_search_steps} Nnc {plan.num_noncoherent_accumulations}")

    for i in range(args.warmup):
        res = acquire_(plan, dev, prns, interm_freq=sc.interm_freq)
    det = sorted(r.prn for r in res if is_detected(r))
    log(f"warmup done; detected {det} (present {sc.present})")

    plan.native.set_timing(True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        res = acquire_(plan, dev, prns, interm_freq=sc.interm_freq)
    t1 = time.perf_counter()
    kt = plan.native.kernel_times()
    plan.native.set_timing(False)
    dt = (t1 - t0) / args.steps
    log(f"timed {args.steps} steps: {dt * 1e3:.3f} ms/step; kernel ms totals "
        + ", ".join(f"{k} {v[0]:.2f}/{v[1]}" for k, v in kt.items() if v[1]))

    # roofline of the dominant kernel (algorithmic bytes per launch)
    F, nb = plan.fft_size, plan.num_blocks
    nnc = plan.num_noncoherent_accumulations
    spc_eff = plan.samples_per_code_eff
    launches_per_step = {k: v[1] / args.steps for k, v in kt.items()}
    batches = max(1, round(launches_per_step.get("corr_tile", 1) / nnc))
    nbat = len(prns) / batches
    bytes_per_launch = {
        "corr_tile": nbat * (8 * F * ndop + 8 * F * nb + 8 * ndop * spc),
        "column": nbat * (8 * ndop * spc + (8 * ndop * spc_eff if nnc > 1 else 4 * ndop * spc_eff)),
        "signal_fft": nnc * plan.num_coherently_integrated_code_periods * spc * 2 + 8 * F * ndop * nnc,
        "reduce": nbat * 4 * ndop * spc_eff,
    }
    dom = max((k for k in kt if kt[k][1]), key=lambda k: kt[k][0])
    avg_ms = kt[dom][0] / kt[dom][1]
    bpl = bytes_per_launch.get(dom)
    roofline = None
    if bpl:
        ach = bpl / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
                    "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": int(bpl)}
    kernel_share = {k: round(v[0] / args.steps, 4) for k, v in kt.items() if v[1]}

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        sub = list(range(args.cpu_prns))
        t0 = time.perf_counter()
        oracle.acquire(plan, sc.signal, sub, if_hz=sc.interm_freq, nthreads=args.cpu_threads)
        cdt = time.perf_counter() - t0
        cpu = {"value": round(len(sub) * ndop * spc * R / cdt, 1), "unit": "search cells/s",
               "cores": args.cpu_threads, "kind": "port",
               "sample": f"PRNs 1..{len(sub)} of the same workload (all 8 segments), C restatement in FP32, "
                         f"{args.cpu_threads} OpenMP threads over PRNs, {cdt:.1f} s"}
        log(f"cpu baseline: {cpu['value']:.3e} cells/s ({cdt:.1f}s)")

    out = {
        "metric": METRIC,
        "value": round(cells / dt, 1),
        "unit": "search cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "GPS L1 C/A 16.368 MHz real int16 at IF 4.092 MHz, 80 ms; PRNs 1-32 (8 present); "
                               f"N_coh 10, bit-edge steps {plan.bit_edge_search_steps}, N_nc 8, +-7 kHz",
                   "prns": len(prns), "samples_per_code": spc, "num_doppler_bins": ndop,
                   "num_blocks": nb, "fft_size": F, "num_coh": 10,
                   "bit_edge_search_steps": plan.bit_edge_search_steps, "num_noncoherent": nnc,
                   "precision": "fp32", "signal": "device-resident int16"},
        "prn_acqs_per_s": round(len(prns) / dt, 2),
        "cell_dwells_per_s": round(cells * nnc / dt, 1),
        "kernel_ms_per_step": kernel_share,
        "detected": det,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    dev.free()
    plan.close()


if __name__ == "__main__":
    main()
