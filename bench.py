"""Benchmark: focused Msamples/s, Level-0 packets -> SLC, on one MI355X.

Workload: one IW2-sized burst, 1,480 lines x 24,000 complex samples, 400
Rayleigh point targets over clutter sigma 0.1, f_dc = 30 Hz, encoded as
FDBAQ space packets.  One step = the whole chain on device-resident packet
bytes: FDBAQ decode -> range compression -> azimuth FFT -> RCMC + matched
filter -> azimuth IFFT, to a device-resident SLC.  `value` is the FP32
chain; the FP64 path (the reference arithmetic) is timed beside it, and the
oracle (CPU FP64 restatement) is timed on the host cores as the CPU baseline.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
    python bench.py --copy-calibration   # plain copy for HBM counter scaling
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
# OpenMP reads this once, at the first parallel region (scene generation).
os.environ.setdefault("OMP_NUM_THREADS", str(min(16, os.cpu_count() or 1)))

METRIC = "focused Msamples/s L0->SLC, FP32 vs FP64 GPU path, %HBM peak, rel-L2<=1e-5"
HBM_PEAK_GBS = 8000.0


def stage_bytes(prec_bytes, L, N, b_pkt):
    """Algorithmic bytes per launch: one read of the stage's logical input
    plus one write of its logical output (twiddles, filters and tables count
    zero)."""
    c = 2 * prec_bytes  # bytes per complex sample
    return {"decode": b_pkt + c * L * N, "range": 2 * c * L * N, "az_fft": 2 * c * L * N,
            "rcmc_mf": 2 * c * L * N, "az_ifft": 2 * c * L * N}


def run_gpu(burst, steps, warmup):
    """Times the chain per precision three ways: eager launches and the
    burst replayed as one captured hipGraph (both without stage timers; the
    faster is the step time), then an eager pass with per-stage HIP-event
    timers on the context stream for the stage breakdown."""
    import torch
    from sentinel1_decode_amd import capi, pipeline

    ctx = capi.Context(0)
    db = pipeline.from_synth(burst)
    L, N = db.n_lines, db.n_samples
    res = {}
    outs = {}
    for name, prec, k in (("fp32", capi.S1G_F32, steps), ("fp64", capi.S1G_F64, max(2, steps // 2))):
        dt = torch.complex64 if prec == capi.S1G_F32 else torch.complex128
        out = torch.empty((L, N), dtype=dt, device="cuda")
        desc = db.burst_desc(out)
        for _ in range(warmup):
            db.focus(ctx, prec, out=out)
            ctx.focus_batch([desc], prec, graph=True)
        torch.cuda.synchronize()
        times = {}
        for mode in ("eager", "graph"):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(k):
                if mode == "eager":
                    db.focus(ctx, prec, out=out)
                else:
                    ctx.focus_batch([desc], prec, graph=True)
            torch.cuda.synchronize()
            times[mode] = (time.perf_counter() - t0) * 1e3 / k
        mode = min(times, key=times.get)
        ms = times[mode]
        ctx.profile(True)
        ctx.profile_read(reset=True)
        for _ in range(k):
            db.focus(ctx, prec, out=out)
        stages = ctx.profile_read(reset=True)
        ctx.profile(False)
        variants = {st: list(ctx.last_variant(st)) for st in ("range", "az_fft", "rcmc_mf", "az_ifft")}
        res[name] = {"ms_per_step": ms, "msamples_s": L * N / (ms * 1e-3) / 1e6, "steps": k, "mode": mode,
                     "ms_eager": times["eager"], "ms_graph": times["graph"],
                     "stages_ms": {s: v[0] / max(v[1], 1) for s, v in stages.items() if v[1]},
                     "variants": variants}
        status = db.d_status.cpu().numpy()
        res[name]["packets_failed"] = int((status != 0).sum())
        outs[name] = out
        print(f"# {name}: {ms:.3f} ms/step ({mode}; eager {times['eager']:.3f}, graph {times['graph']:.3f}), "
              f"{res[name]['msamples_s']:.1f} Msamples/s, stages "
              + ", ".join(f"{s}={v:.3f}ms" for s, v in res[name]["stages_ms"].items()), file=sys.stderr, flush=True)
    a = outs["fp64"].cpu().numpy()
    b = outs["fp32"].cpu().numpy()
    res["rel_l2_fp32_vs_fp64"] = float(np.linalg.norm(b - a) / np.linalg.norm(a))
    # peak point-target phase error: brightest 64 pixels of the FP64 image
    flat = np.argsort(np.abs(a).ravel())[-64:]
    dphi = np.angle(b.ravel()[flat] * np.conj(a.ravel()[flat]))
    res["peak_phase_err_rad"] = float(np.abs(dphi).max())
    # decode parity on this burst: FP64 decode vs oracle, FP32 = rounded doubles
    dec64, _ = db.decode(ctx, capi.S1G_F64)
    dec32, _ = db.decode(ctx, capi.S1G_F32)
    torch.cuda.synchronize()
    res["_dec64"] = dec64.cpu().numpy()
    res["_dec32"] = dec32.cpu().numpy()
    ctx.close()
    return res, L, N


def measured_traffic(kernel_prefix, L, N):
    """HBM bytes per launch of a kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated against a
    plain copy), or None when the summary is for another shape."""
    path = os.path.join(HERE, "profiles", "chain_iw2_burst_1480x24000_hbm_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except OSError:
        return None
    if t["workload"] != {"lines": L, "samples": N}:
        return None
    for k, v in t["kernels"].items():
        if k.startswith(kernel_prefix):
            return round(v["read_bytes"] + v["write_bytes"])
    return None


def run_cpu_baseline(burst):
    """The oracle (CPU FP64 restatement) on the same packets, full chain."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    from chain_ref import oracle_chain
    t0 = time.perf_counter()
    out = oracle_chain(burst)
    t1 = time.perf_counter()
    return out, t1 - t0


def copy_calibration(mib=256, reps=5):
    """A plain device copy of known size (torch elementwise kernel: read
    mib MiB, write mib MiB per launch), the reference against which
    rocprofv3's FETCH_SIZE / WRITE_SIZE counters are scaled."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n,
                                           "launches": reps}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lines", type=int, default=1480)
    ap.add_argument("--samples", type=int, default=24000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain copy of known size (for counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); bursts shard across GPUs without "
              "communication, so the per-GPU figure is the single-GPU one.", file=sys.stderr)
        sys.exit(2)

    from sentinel1_decode_amd import synth
    t0 = time.perf_counter()
    spec = synth.config1(n_lines=args.lines, n_samples=args.samples)
    burst = synth.make_burst(spec)
    print(f"# scene + encode {time.perf_counter() - t0:.1f} s, {burst.packet_bytes / 1e6:.1f} MB of packets",
          file=sys.stderr, flush=True)

    res, L, N = run_gpu(burst, args.steps, args.warmup)
    b_pkt = burst.packet_bytes

    cpu = None
    dec_mismatch = None
    threads = int(os.environ["OMP_NUM_THREADS"])
    if not args.no_cpu:
        ref, secs = run_cpu_baseline(burst)
        dec = ref["decoded"]
        dec_mismatch = int((res["_dec64"].view(np.uint64) != dec.view(np.uint64)).sum()
                           + (res["_dec32"].view(np.uint32) != dec.astype(np.complex64).view(np.uint32)).sum())
        cpu = {"value": L * N / secs / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
               "sample": f"the same {L}x{N} burst, full FP64 chain (decode, range, azimuth FFT, RCMC+MF, "
                         f"IFFT) by the oracle on {threads} host threads, {secs:.1f} s"}
    fp32, fp64 = res["fp32"], res["fp64"]
    # dominant kernel of the FP32 chain
    sb = stage_bytes(4, L, N, b_pkt)
    dom = max(fp32["stages_ms"], key=lambda s: fp32["stages_ms"][s] if s in sb else -1)
    achieved = sb[dom] / (fp32["stages_ms"][dom] * 1e-3) / 1e9
    fft_frac = {s: stage_bytes(4, L, N, b_pkt)[s] / (fp32["stages_ms"][s] * 1e-3) / 1e9 / HBM_PEAK_GBS
                for s in ("range", "az_fft", "az_ifft") if s in fp32["stages_ms"]}
    kernel_of = {"decode": "fdbaq_decode_kernel<float>", "range": "range_os3_kernel<float",
                 "az_fft": "az_fft_fixed_kernel<float, false", "rcmc_mf": "rcmc_mf_kernel<float",
                 "az_ifft": "az_fft_fixed_kernel<float, true"}
    traffic = measured_traffic(kernel_of.get(dom, "?"), L, N)
    line = {
        "metric": METRIC,
        "value": round(fp32["msamples_s"], 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": fp32["steps"],
        "warmup": args.warmup,
        "ms_per_step": round(fp32["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"IW2 burst {L}x{N} FDBAQ packets -> SLC (400 Rayleigh targets, clutter 0.1, "
                               f"f_dc 30 Hz)", "lines": L, "samples": N, "packet_bytes": b_pkt},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": sb[dom]},
        "cpu_baseline": cpu,
        "fp64_msamples_s": round(fp64["msamples_s"], 3),
        "fp64_ms_per_step": round(fp64["ms_per_step"], 4),
        "fp32_over_fp64": round(fp32["msamples_s"] / fp64["msamples_s"], 3),
        "mode": fp32["mode"],
        "ms_eager_fp32": round(fp32["ms_eager"], 4), "ms_graph_fp32": round(fp32["ms_graph"], 4),
        "rel_l2_fp32_vs_fp64": res["rel_l2_fp32_vs_fp64"],
        "peak_phase_err_rad": res["peak_phase_err_rad"],
        "decode_mismatches": dec_mismatch,
        "packets_failed": fp32["packets_failed"],
        "stages_ms_fp32": {k: round(v, 4) for k, v in fp32["stages_ms"].items()},
        "stages_ms_fp64": {k: round(v, 4) for k, v in fp64["stages_ms"].items()},
        "hbm_frac_fft_passes_fp32": {k: round(v, 4) for k, v in fft_frac.items()},
        "autotuned_variants": {"fp32": fp32["variants"], "fp64": fp64["variants"]},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
