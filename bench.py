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
    python bench.py --workload {iw-batch,iw2-swath-eccm,autofocus,iw-tops}   # the other workloads
"""
from __future__ import annotations

import argparse
import dataclasses
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


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def peak_phase_err(a, b, n=64):
    """Largest phase difference over the n brightest pixels of b (the
    point-target peaks)."""
    flat = np.argsort(np.abs(b).ravel())[-n:]
    return float(np.abs(np.angle(a.ravel()[flat] * np.conj(b.ravel()[flat]))).max())


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
    res["_slc64"], res["_slc32"] = a, b
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
    # the FP32 chain stage by stage (each stage's output kept), for the
    # per-stage error against the oracle's stage outputs
    rc = ctx.range_compress(dec32, db.geo.range_desc())
    res["_stages32"] = {"rc": rc.cpu().numpy()}
    ctx.azimuth_fft(rc, inverse=False)
    res["_stages32"]["rd"] = rc.cpu().numpy()
    mf = ctx.rcmc_matched_filter(rc, db.azimuth_desc())
    res["_stages32"]["mf"] = mf.cpu().numpy()
    ctx.azimuth_fft(mf, inverse=True)
    res["_stages32"]["slc"] = mf.cpu().numpy()
    del dec64, dec32, rc, mf
    ctx.close()
    return res, L, N


# FP32 kernel of each stage variant (the autotuner's numbering in the
# library: range = overlap-save block 16384 or 8192; azimuth FFT = 64-byte
# panels, 32-byte panels, generic, persistent 64-byte panels, the same with
# the first and last passes in registers; RCMC = three
# per-sample tilings, two run kernels, two linear-geometry kernels).
_STAGE_KERNELS = {
    "decode": ["fdbaq_decode_kernel<float"],
    "range": ["range_os3_kernel<float, 16384", "range_os3_kernel<float, 8192"],
    "az_fft": ["az_fft_fixed_kernel<float, false, 512, 8,", "az_fft_fixed_kernel<float, false, 512, 4,",
               "az_fft_generic_kernel<float, false", "az_fft_pipe_kernel<float, false",
               "az_fft_reg_kernel<float, false"],
    "az_ifft": ["az_fft_fixed_kernel<float, true, 512, 8,", "az_fft_fixed_kernel<float, true, 512, 4,",
                "az_fft_generic_kernel<float, true", "az_fft_pipe_kernel<float, true",
                "az_fft_reg_kernel<float, true"],
    "rcmc_mf": ["rcmc_mf_kernel<float, 256, 8>", "rcmc_mf_kernel<float, 256, 4>", "rcmc_mf_kernel<float, 512, 4>",
                "rcmc32_run_kernel<256, 9>", "rcmc32_run_kernel<128, 9>", "rcmc32_lin_kernel<256, 16>",
                "rcmc32_lin_kernel<256, 8>"],
}


def kernel_name(stage, variant):
    names = _STAGE_KERNELS.get(stage, [])
    if stage == "decode":
        return names[0]
    return names[variant] if 0 <= variant < len(names) else "?"


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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_cpu_baseline(burst, threads):
    """The oracle (CPU FP64 restatement, its own mixed-radix FFT, not FFTW)
    on the same packets, full chain: one warm-up then the median of three
    runs on `threads` OpenMP threads, and one run on a single thread.
    Returns (stage outputs of the last run, median seconds, 1-thread
    seconds)."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import oracle
    from chain_ref import oracle_chain
    oracle.set_threads(threads)
    out = oracle_chain(burst)  # warm-up (FFT plans, page faults)
    secs = []
    for _ in range(3):
        t0 = time.perf_counter()
        out = oracle_chain(burst)
        secs.append(time.perf_counter() - t0)
    oracle.set_threads(1)
    t0 = time.perf_counter()
    oracle_chain(burst)
    one = time.perf_counter() - t0
    oracle.set_threads(threads)
    return out, float(np.median(secs)), one


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def run_iw_batch(steps, warmup):
    """27 uneven bursts (9 IW1-like 1420 x 21600, 9 IW2-like 1480 x 24000,
    9 IW3-like 1520 x 20400; 200 Rayleigh targets each, clutter 0.1) through
    the FP32 chain: one captured hipGraph of the batch (the bursts' chains in
    turn, each decode beside the previous burst's focusing) against the
    graph of the stage-ordered batch (one launch per stage and shape class
    over per-burst descriptors) and against launching the bursts one after
    another."""
    import torch
    from sentinel1_decode_amd import capi, pipeline, synth
    t0 = time.perf_counter()
    dbs = []
    for sp in synth.config3_specs():
        dbs.append(pipeline.from_synth(synth.make_burst(sp)))
    print(f"# 27 bursts synthesised and uploaded in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    ctx = capi.Context(0)
    outs = [torch.empty((db.n_lines, db.n_samples), dtype=torch.complex64, device="cuda") for db in dbs]
    descs = [db.burst_desc(o) for db, o in zip(dbs, outs)]
    samples = sum(db.n_lines * db.n_samples for db in dbs)
    ms_graph = _timed(lambda: ctx.focus_batch(descs, capi.S1G_F32, graph=True), steps, warmup)
    ms_staged = _timed(lambda: ctx.focus_batch(descs, capi.S1G_F32, graph=True, staged=True), steps, warmup)
    ms_seq = _timed(lambda: [db.focus(ctx, capi.S1G_F32, out=o) for db, o in zip(dbs, outs)], steps, warmup)
    failed = sum(int((db.d_status.cpu().numpy() != 0).sum()) for db in dbs)
    ctx.close()
    return {"metric": "focused Msamples/s, 27 uneven IW bursts in one hipGraph (FP32)",
            "value": round(samples / (ms_graph * 1e-3) / 1e6, 3), "unit": "Msamples/s", "n_gpus": 1,
            "steps": steps, "warmup": warmup, "ms_per_step": round(ms_graph, 4), "higher_is_better": True,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": "27 bursts: 9x 1420x21600 (fs 50.05 MHz, PRF 1685), 9x 1480x24000 "
                                   "(64.35 MHz, 1717), 9x 1520x20400 (54.60 MHz, 1451)", "samples": samples},
            "staged_graph_ms_per_step": round(ms_staged, 4),
            "staged_graph_msamples_s": round(samples / (ms_staged * 1e-3) / 1e6, 3),
            "sequential_ms_per_step": round(ms_seq, 4),
            "sequential_msamples_s": round(samples / (ms_seq * 1e-3) / 1e6, 3),
            "graph_speedup": round(ms_seq / ms_graph, 3), "packets_failed": failed}


def run_iw_tops(steps, warmup, lines=1480, samples=24000, cpu=True):
    """One IW2 burst (lines x samples, 400 Rayleigh targets, clutter 0.1) in a
    Level-0 stream whose beam sweeps addresses 213..810 (+-0.0105 rad, a
    TOPS steering rate near IW's) through the TOPS branch: range compression
    from packets, Doppler centroid, azimuth-frequency unfolding, line
    filter, azimuth IFFT, azimuth-time unfolding (Level0Product.get_tops_burst)."""
    import torch
    from sentinel1_decode_amd import capi, pipeline, synth
    t0 = time.perf_counter()
    spec = synth.BurstSpec(n_lines=lines, n_samples=samples, rgdec=8, pri_code=21861, n_targets=400,
                           amp_mode="rayleigh", swath_number=11, seed=41, clutter_sigma=0.1)
    s = synth.level0_stream([spec], beam_addresses=(213, 810))
    print(f"# TOPS burst synthesised in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    ctx = capi.Context(0)
    prod = pipeline.Level0Product(s.stream, ctx)
    res = {}
    for prec, name in ((capi.S1G_F64, "f64"), (capi.S1G_F32, "f32")):
        out = prod.get_tops_burst("IW2", 0, prec=prec)
        res[name] = {"ms_per_burst": round(_timed(lambda: prod.get_tops_burst("IW2", 0, prec=prec), steps, warmup), 3),
                     "out_shape": list(out.shape)}
        print(f"# TOPS {name}: {res[name]}", file=sys.stderr, flush=True)
    line = {"metric": "TOPS-focused IW2 burst, ms per burst (packets -> SLC)", "value": res["f64"]["ms_per_burst"],
            "unit": "ms", "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": res["f64"]["ms_per_burst"],
            "higher_is_better": False, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"IW2 burst {lines}x{samples} FDBAQ packets, beam 213..810, TOPS branch"},
            "f32": res["f32"], "f64": res["f64"]}
    ctx.close()
    if cpu:
        # the oracle's TOPS chain on a bounded sample: the same burst cut to
        # 1024 range samples (cost is about linear in samples)
        import oracle
        sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
        import types
        from chain_ref import oracle_tops
        small = dataclasses.replace(spec, n_samples=1024)
        ss = synth.level0_stream([small], beam_addresses=(213, 810))
        offs = oracle.index_level0(ss.stream)
        F = oracle.header_fields(ss.stream, offs)
        uo = np.asarray(offs, np.uint64) + 68
        ob = types.SimpleNamespace(stream=ss.stream, ud_offsets=uo, ud_lengths=(F[:, 7] + 1 - 62).astype(np.uint32),
                                   num_quads=F[:, oracle.SEC["num_quadratures"]].astype(np.int32), spec=small)
        threads = oracle.set_threads(0)
        t1 = time.perf_counter()
        oracle_tops(ob, F, oracle.state_vectors(F))
        dt = time.perf_counter() - t1
        line["cpu_baseline"] = {"value": round(dt * 1e3 * samples / 1024, 1), "unit": "ms per burst (scaled)",
                                "cores": threads, "kind": "port",
                                "sample": f"oracle TOPS chain on {lines}x1024 in {dt:.2f} s, x {samples}/1024"}
    return line


def run_swath_eccm(steps, warmup, n_bursts=9):
    """An IW2 swath of 9 bursts (1480 x 24000 each) with narrowband RFI on
    20 % of the lines (runs of 10-50 lines, 3 tones within +-28 MHz, INR
    15-25 dB); ECCM thresholds from clean-line statistics; FP32 chain with
    ECCM, the 9 bursts in one captured graph."""
    import torch
    from sentinel1_decode_amd import capi, pipeline, synth
    ctx = capi.Context(0)
    dbs, masks = [], []
    t0 = time.perf_counter()
    for k in range(n_bursts):
        sp = synth.config1(seed=2 * 1000 + k)
        raw = sp.raw()
        masks.append(synth.add_rfi(raw, sp, seed=2 * 1000 + k))
        dbs.append(pipeline.from_synth(synth.encode_burst(sp, raw)))
        del raw
    print(f"# swath synthesised in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    # thresholds from the lines the RFI generator left clean (first burst)
    dec, _ = dbs[0].decode(ctx, capi.S1G_F32)
    clean = torch.from_numpy(~masks[0]).cuda()
    spec_max = torch.fft.fft(dec, dim=1).abs().amax(dim=1)
    det = 1.25 * float(spec_max[clean].max())
    win = dec[clean][:64, : (dec.shape[1] // 64 - 1) * 64].unfold(1, 128, 64)
    w = torch.fft.fft(win, dim=-1)
    grad = (w[..., 2:] - w[..., :-2]).abs() / 2
    mit = 2.0 * float(grad.max())
    del dec, win, w, grad
    eccm = capi.EccmDesc(dbs[0].n_lines, dbs[0].n_samples, 128, 64, det, mit)
    outs = [torch.empty((db.n_lines, db.n_samples), dtype=torch.complex64, device="cuda") for db in dbs]
    flags = [torch.zeros(db.n_lines, dtype=torch.uint8, device="cuda") for db in dbs]
    descs = [db.burst_desc(o, eccm=eccm, eccm_flags=f) for db, o, f in zip(dbs, outs, flags)]
    ms = _timed(lambda: ctx.focus_batch(descs, capi.S1G_F32, graph=True), steps, warmup)
    fl = np.concatenate([f.cpu().numpy().astype(bool) for f in flags])
    mk = np.concatenate(masks)
    samples = sum(db.n_lines * db.n_samples for db in dbs)
    ctx.close()
    return {"metric": "focused Msamples/s, IW2 swath of 9 bursts with RFI, ECCM + FP32 chain",
            "value": round(samples / (ms * 1e-3) / 1e6, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": f"{n_bursts} x 1480x24000, RFI on 20% of lines", "samples": samples},
            "eccm_thresholds": {"detection": det, "mitigation": mit},
            "eccm_flag_recall": round(float((fl & mk).sum() / max(mk.sum(), 1)), 4),
            "eccm_flag_precision": round(float((fl & mk).sum() / max(fl.sum(), 1)), 4)}


def run_autofocus(iters=50):
    """Differentiable autofocus on a 4096 x 8192 burst (200 targets, clutter
    0.1): a quadratic phase error (2 rad at the aperture edge) plus a
    Gaussian random walk (0.05 rad / Doppler bin) is injected into the
    matched-filtered range-Doppler spectrum, and `iters` L-BFGS iterations
    driven by the hand-written FP32 adjoint remove it.  The spectrum is the
    checkpoint: each evaluation recomputes the image from it."""
    import torch
    from sentinel1_decode_amd import capi, pipeline, synth
    ctx = capi.Context(0)
    spec = synth.config4()
    db = pipeline.from_synth(synth.make_burst(spec))
    dec, _ = db.decode(ctx, capi.S1G_F32)
    rc = ctx.range_compress(dec, db.geo.range_desc())
    ctx.azimuth_fft(rc, inverse=False)
    rd = ctx.rcmc_matched_filter(rc, db.azimuth_desc())
    del dec, rc
    L = rd.shape[0]
    rng = np.random.default_rng(4)
    f = np.linspace(-1.0, 1.0, L)
    err = 2.0 * f ** 2 + np.cumsum(rng.normal(0.0, 0.05, L))
    rd_err = (rd * torch.from_numpy(np.exp(1j * err)).to(rd.dtype).cuda()[:, None]).contiguous()
    evals = [0]

    def solve(data):
        phase = torch.zeros(L, dtype=torch.float64, device="cuda", requires_grad=True)
        opt = torch.optim.LBFGS([phase], lr=1.0, max_iter=iters, history_size=10, line_search_fn="strong_wolfe")

        def closure():
            opt.zero_grad()
            g = torch.zeros(L, dtype=torch.float64, device="cuda")
            e = ctx.af_adjoint(data, phase.detach(), g)
            phase.grad = g
            evals[0] += 1
            return torch.tensor(e, dtype=torch.float64)

        opt.step(closure)
        return phase.detach()

    zero = torch.zeros(L, dtype=torch.float64, device="cuda")
    e0 = ctx.af_forward(rd_err, zero)
    # the error-free solve runs first: it also warms up the optimiser's
    # library kernels, which would otherwise load inside the timed solve
    est_clean = solve(rd).cpu().numpy()
    evals[0] = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    est = solve(rd_err)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    n_eval = evals[0]
    e1 = ctx.af_forward(rd_err, est)
    est = est.cpu().numpy()
    # one entropy + gradient evaluation on its own: device time of its
    # kernels (HIP events on the context stream) and the call's wall time
    g = torch.zeros(L, dtype=torch.float64, device="cuda")
    d_est0 = torch.from_numpy(est).cuda()
    for _ in range(3):
        ctx.af_adjoint(rd_err, d_est0, g)
    ctx.profile(True)
    ctx.profile_read(reset=True)
    reps = 20
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        ctx.af_adjoint(rd_err, d_est0, g)
    torch.cuda.synchronize()
    call_ms = (time.perf_counter() - t1) * 1e3 / reps
    dev_ms, dev_n = ctx.profile_read(reset=True)["autofocus"]
    ctx.profile(False)
    eval_ms = dev_ms / max(dev_n, 1)
    # gradient check at the solution: FP32 adjoint vs FP64 adjoint and vs
    # central differences (eps 1e-3, 32 seeded bins) of the FP64 entropy
    rd64 = rd_err.to(torch.complex128)
    g32 = torch.zeros(L, dtype=torch.float64, device="cuda")
    g64 = torch.zeros(L, dtype=torch.float64, device="cuda")
    d_est = torch.from_numpy(est).cuda()
    ctx.af_adjoint(rd_err, d_est, g32)
    ctx.af_adjoint(rd64, d_est, g64)
    g32, g64 = g32.cpu().numpy(), g64.cpu().numpy()
    fd_err = 0.0
    for i in np.random.default_rng(9).choice(L, 32, replace=False):
        p1, p2 = est.copy(), est.copy()
        p1[i] += 1e-3
        p2[i] -= 1e-3
        fd = (ctx.af_forward(rd64, torch.from_numpy(p1).cuda()) - ctx.af_forward(rd64, torch.from_numpy(p2).cuda())) / 2e-3
        fd_err = max(fd_err, abs(fd - g64[i]) / max(np.abs(g64).max(), 1e-300))
    ctx.close()

    def rms(x, band=0.75):
        i = np.arange(x.size, dtype=np.float64)
        keep = np.abs(f) <= band
        a = np.vstack([np.ones_like(i), i]).T[keep]
        coef, *_ = np.linalg.lstsq(a, x[keep], rcond=None)
        return float(np.sqrt(np.mean((x[keep] - a @ coef) ** 2)))

    return {"metric": "autofocus: residual azimuth phase RMS after 50 L-BFGS iterations (4096 x 8192)",
            "value": round(rms(err + est, 0.25), 5), "unit": "rad", "higher_is_better": False, "n_gpus": 1,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": "4096x8192, 200 targets, clutter 0.1; error 2 rad quadratic + 0.05 rad/bin "
                                   "random walk", "iterations": iters},
            "residual_rms_rad": {"injected_band0.75": rms(err), "after_band0.75": rms(err + est),
                                 "injected_band0.25": rms(err, 0.25), "after_band0.25": rms(err + est, 0.25),
                                 "differential_vs_errorfree_autofocus_band0.75": rms(err + est - est_clean)},
            "entropy": {"with_error": e0, "autofocused": e1},
            "seconds": round(secs, 3), "entropy_gradient_evaluations": n_eval,
            "ms_per_lbfgs_evaluation": round(secs * 1e3 / max(n_eval, 1), 3),
            "evaluation_device_ms": round(eval_ms, 4), "evaluation_call_ms": round(call_ms, 4),
            "evaluation_bytes": 2 * L * rd.shape[1] * rd.element_size(),
            "evaluation_hbm_frac": round(2 * L * rd.shape[1] * rd.element_size() / (eval_ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
            "gradient_check": {"rel_l2_fp32_vs_fp64_adjoint": float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64)),
                               "max_fd_error_rel_to_max_grad": fd_err}}


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
    ap.add_argument("--workload", default="iw2-burst",
                    choices=["iw2-burst", "iw-batch", "iw2-swath-eccm", "autofocus", "iw-tops"],
                    help="iw2-burst (the default line); the others print their own line")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload != "iw2-burst":
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        fn = {"iw-batch": lambda: run_iw_batch(args.steps, args.warmup),
              "iw2-swath-eccm": lambda: run_swath_eccm(args.steps, args.warmup),
              "autofocus": lambda: run_autofocus(),
              "iw-tops": lambda: run_iw_tops(args.steps, args.warmup, args.lines, args.samples,
                                             cpu=not args.no_cpu)}[args.workload]
        print(json.dumps(fn()), flush=True)
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
    vs_oracle = None
    threads = int(os.environ["OMP_NUM_THREADS"])
    if not args.no_cpu:
        ref, secs, secs1 = run_cpu_baseline(burst, threads)
        dec = ref["decoded"]
        dec_mismatch = int((res["_dec64"].view(np.uint64) != dec.view(np.uint64)).sum()
                           + (res["_dec32"].view(np.uint32) != dec.astype(np.complex64).view(np.uint32)).sum())
        cpu = {"value": L * N / secs / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
               "sample": f"the same {L}x{N} burst, full FP64 chain (decode, range, azimuth FFT, RCMC+MF, "
                         f"IFFT) by the oracle (its own mixed-radix FFT, not FFTW) on {threads} OpenMP threads "
                         f"(the host share of one GPU; nproc {os.cpu_count()}): median of 3 after a warm-up, "
                         f"{secs:.2f} s",
               "cpu_model": cpu_model(), "nproc": os.cpu_count(),
               "one_thread": {"value": L * N / secs1 / 1e6, "seconds": round(secs1, 2)}}
        vs_oracle = {"rel_l2_fp64_vs_oracle": rel_l2(res["_slc64"], ref["slc"]),
                     "rel_l2_fp32_vs_oracle": rel_l2(res["_slc32"], ref["slc"]),
                     "peak_phase_err_fp32_vs_oracle_rad": peak_phase_err(res["_slc32"], ref["slc"]),
                     "stages_fp32_vs_oracle": {
                         name: {"rel_l2": rel_l2(res["_stages32"][key], ref[key]),
                                "peak_phase_err_rad": peak_phase_err(res["_stages32"][key], ref[key])}
                         for name, key in (("range", "rc"), ("az_fft", "rd"), ("rcmc_mf", "mf"),
                                           ("az_ifft", "slc"))}}
        vs_oracle["stages_fp32_vs_oracle"]["decode"] = {
            "rel_l2": rel_l2(res["_dec32"], dec), "bit_mismatches": dec_mismatch}
        del ref
    fp32, fp64 = res["fp32"], res["fp64"]
    # dominant kernel of the FP32 chain
    sb = stage_bytes(4, L, N, b_pkt)
    dom = max(fp32["stages_ms"], key=lambda s: fp32["stages_ms"][s] if s in sb else -1)
    achieved = sb[dom] / (fp32["stages_ms"][dom] * 1e-3) / 1e9
    fft_frac = {s: stage_bytes(4, L, N, b_pkt)[s] / (fp32["stages_ms"][s] * 1e-3) / 1e9 / HBM_PEAK_GBS
                for s in ("range", "az_fft", "az_ifft") if s in fp32["stages_ms"]}
    traffic = measured_traffic(kernel_name(dom, fp32["variants"].get(dom, [0, 1])[0]), L, N)
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
        "vs_oracle": vs_oracle,
        "packets_failed": fp32["packets_failed"],
        "stages_ms_fp32": {k: round(v, 4) for k, v in fp32["stages_ms"].items()},
        "stages_ms_fp64": {k: round(v, 4) for k, v in fp64["stages_ms"].items()},
        "hbm_frac_fft_passes_fp32": {k: round(v, 4) for k, v in fft_frac.items()},
        "autotuned_variants": {"fp32": fp32["variants"], "fp64": fp64["variants"]},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
