"""GPU parity at the full sizes of the benchmark workloads:

* the 1,024 x 8,192 nine-target scene (zero Doppler centroid): decode
  bit-exact, FP64 chain within 1e-10 and FP32 within 1e-5 (relative L2) of
  the CPU FP64 oracle;
* the 27-burst uneven batch (9 each of 1,420 x 21,600, 1,480 x 24,000 and
  1,520 x 20,400) in one captured graph: every burst equal to its own eager
  chain, and one burst of each shape within 1e-5 of the oracle;
* autofocus on 4,096 x 8,192: FP64 entropy gradient (hand-written adjoint)
  within 1e-9 of the oracle's, FP32 adjoint against central differences
  (eps 1e-3, 32 seeded bins) of the FP64 entropy.
"""
import time

import numpy as np
import pytest

import oracle
from chain_ref import oracle_chain
from conftest import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from sentinel1_decode_amd import capi, pipeline, synth  # noqa: E402

F32, F64 = capi.S1G_F32, capi.S1G_F64


def log(msg):
    print(msg, flush=True)


@pytest.mark.timeout(900)
def test_small_scene_full_size_vs_oracle(ctx):
    spec = synth.config0()
    assert (spec.n_lines, spec.n_samples, spec.f_dc) == (1024, 8192, 0.0)
    burst = synth.make_burst(spec)
    t0 = time.perf_counter()
    ref = oracle_chain(burst)
    log(f"1024 x 8192 oracle chain {time.perf_counter() - t0:.1f} s")
    db = pipeline.from_synth(burst)
    dec, st = db.decode(ctx, F64)
    torch.cuda.synchronize()
    assert (st.cpu().numpy() == 0).all()
    assert np.array_equal(dec.cpu().numpy().view(np.uint64), ref["decoded"].view(np.uint64))
    e64 = rel_l2(db.focus(ctx, F64).cpu().numpy(), ref["slc"])
    e32 = rel_l2(db.focus(ctx, F32).cpu().numpy(), ref["slc"])
    log(f"1024 x 8192 rel-L2 vs oracle: FP64 {e64:.2e}, FP32 {e32:.2e}")
    assert e64 < 1e-10
    assert e32 < 1e-5


@pytest.mark.timeout(1200)
def test_uneven_27_burst_batch_full_size(ctx):
    specs = synth.config3_specs()
    assert len(specs) == 27
    t0 = time.perf_counter()
    bursts = [synth.make_burst(s) for s in specs]
    log(f"27 bursts synthesised in {time.perf_counter() - t0:.1f} s")
    dbs = [pipeline.from_synth(b) for b in bursts]
    outs = [torch.empty((db.n_lines, db.n_samples), dtype=torch.complex64, device="cuda") for db in dbs]
    descs = [db.burst_desc(o) for db, o in zip(dbs, outs)]
    for rep in range(2):
        for o in outs:
            o.zero_()
        ctx.focus_batch(descs, F32, graph=True)
        torch.cuda.synchronize()
        for k, (db, o) in enumerate(zip(dbs, outs)):
            assert (db.d_status.cpu().numpy() == 0).all(), k
            assert torch.equal(o, db.focus(ctx, F32)), (rep, k)
        log(f"graph replay {rep}: 27 bursts equal their eager chains")
    # one burst of each shape against the oracle
    for k in range(3):
        t0 = time.perf_counter()
        ref = oracle_chain(bursts[k])
        err = rel_l2(outs[k].cpu().numpy(), ref["slc"])
        log(f"burst {k} {specs[k].n_lines} x {specs[k].n_samples}: FP32 rel-L2 vs oracle {err:.2e} "
            f"({time.perf_counter() - t0:.1f} s)")
        assert err < 1e-5, k


@pytest.mark.timeout(900)
def test_autofocus_gradient_full_size(ctx):
    spec = synth.config4()
    assert (spec.n_lines, spec.n_samples) == (4096, 8192)
    db = pipeline.from_synth(synth.make_burst(spec))
    dec, _ = db.decode(ctx, F64)
    rc = ctx.range_compress(dec, db.geo.range_desc())
    ctx.azimuth_fft(rc, inverse=False)
    rd = ctx.rcmc_matched_filter(rc, db.azimuth_desc())
    del dec, rc
    L = rd.shape[0]
    rng = np.random.default_rng(4)
    f = np.linspace(-1.0, 1.0, L)
    err = 2.0 * f ** 2 + np.cumsum(rng.normal(0.0, 0.05, L))
    phase = -0.5 * err  # half-way to the correction: gradients well away from zero
    rd_err = (rd * torch.from_numpy(np.exp(1j * err)).cuda()[:, None]).contiguous()
    d_ph = torch.from_numpy(phase).cuda()
    g64 = torch.zeros(L, dtype=torch.float64, device="cuda")
    e64 = ctx.af_adjoint(rd_err, d_ph, g64)
    rd_host = rd_err.cpu().numpy()
    t0 = time.perf_counter()
    e_ref, g_ref = oracle.af_entropy_grad(rd_host, phase)
    log(f"oracle entropy + gradient on 4096 x 8192: {time.perf_counter() - t0:.1f} s")
    g64 = g64.cpu().numpy()
    # 1e-10: the oracle sums the 3.4e7 pixel terms serially (rounding up to
    # n eps, about 6e-13 typically); the GPU sums them as a tree
    assert e64 == pytest.approx(e_ref, rel=1e-10)
    log(f"FP64 entropy rel diff {abs(e64 / e_ref - 1):.2e}, gradient rel-L2 vs oracle {rel_l2(g64, g_ref):.2e}")
    assert rel_l2(g64, g_ref) < 1e-9
    rd32 = rd_err.to(torch.complex64)
    g32 = torch.zeros(L, dtype=torch.float64, device="cuda")
    ctx.af_adjoint(rd32, d_ph, g32)
    g32 = g32.cpu().numpy()
    scale = np.abs(g64).max()
    worst = 0.0
    for i in np.random.default_rng(9).choice(L, 32, replace=False):
        p1, p2 = phase.copy(), phase.copy()
        p1[i] += 1e-3
        p2[i] -= 1e-3
        fd = (ctx.af_forward(rd_err, torch.from_numpy(p1).cuda())
              - ctx.af_forward(rd_err, torch.from_numpy(p2).cuda())) / 2e-3
        worst = max(worst, abs(fd - g32[i]) / scale)
        assert abs(fd - g64[i]) <= 1e-5 * scale, (i, fd, g64[i])
    log(f"FP32 adjoint vs finite differences: worst {worst:.2e} of max |grad|; "
        f"FP32 vs FP64 gradient rel-L2 {rel_l2(g32, g64):.2e}")
    assert worst <= 1e-3
    assert rel_l2(g32, g64) < 1e-3
