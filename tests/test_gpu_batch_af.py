"""GPU: graph-captured batches of uneven bursts, and the autofocus forward /
adjoint against the oracle and finite differences."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from sentinel1_decode_amd import capi, pipeline, synth  # noqa: E402

F32, F64 = capi.S1G_F32, capi.S1G_F64


@pytest.fixture(scope="module")
def uneven_bursts():
    specs = [synth.BurstSpec(n_lines=96, n_samples=2160, rgdec=6, pri_code=22276, n_targets=20,
                             amp_mode="rayleigh", seed=31),
             synth.BurstSpec(n_lines=148, n_samples=2400, rgdec=8, pri_code=21861, n_targets=20,
                             amp_mode="rayleigh", f_dc=30.0, seed=32),
             synth.BurstSpec(n_lines=152, n_samples=2040, rgdec=11, pri_code=25868, n_targets=20,
                             amp_mode="rayleigh", seed=33)]
    return [pipeline.from_synth(synth.make_burst(s)) for s in specs]


@pytest.mark.parametrize("prec", [F32, F64])
def test_batch_graph_matches_eager(ctx, uneven_bursts, prec):
    dt = torch.complex64 if prec == F32 else torch.complex128
    eager = [db.focus(ctx, prec) for db in uneven_bursts]
    torch.cuda.synchronize()
    outs = [torch.zeros((db.n_lines, db.n_samples), dtype=dt, device="cuda") for db in uneven_bursts]
    descs = [db.burst_desc(o) for db, o in zip(uneven_bursts, outs)]
    for rep in range(3):
        for o in outs:
            o.zero_()
        ctx.focus_batch(descs, prec, graph=True)
        torch.cuda.synchronize()
        for o, e in zip(outs, eager):
            assert torch.equal(o, e), rep
        # unrelated eager work between replays
        uneven_bursts[rep % 3].focus(ctx, prec)
        torch.cuda.synchronize()
    # the ungraphed batch gives the same bits
    for o in outs:
        o.zero_()
    ctx.focus_batch(descs, prec, graph=False)
    torch.cuda.synchronize()
    for o, e in zip(outs, eager):
        assert torch.equal(o, e)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_shared_output_runs_in_order(ctx, uneven_bursts, graph):
    """Two descriptors that write the same SLC and status words (the same
    burst twice) are not decoded beside each other's focusing: the batch
    runs the chains one after the other and ends with the burst's own
    image."""
    db = uneven_bursts[1]
    eager = db.focus(ctx, F32)
    out = torch.zeros((db.n_lines, db.n_samples), dtype=torch.complex64, device="cuda")
    d = db.burst_desc(out)
    for rep in range(2):
        out.zero_()
        ctx.focus_batch([d, d], F32, graph=graph)
        torch.cuda.synchronize()
        assert torch.equal(out, eager), rep


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("eccm_first", [True, False])
def test_batch_with_eccm_matches_single_bursts(ctx, uneven_bursts, graph, eccm_first):
    """A batch in which one burst runs ECCM (its decode goes through the
    shared workspace, so the chains run in order) gives each burst the image
    and flags of its own single call, with the ECCM burst first or second."""
    a, b = uneven_bursts[1], uneven_bursts[0]
    ed = capi.EccmDesc(a.n_lines, a.n_samples, 128, 64, 1.0e3, 4000.0)
    fa = torch.zeros(a.n_lines, dtype=torch.uint8, device="cuda")
    ea = a.focus(ctx, F32, eccm=ed, eccm_flags=fa)
    eb = b.focus(ctx, F32)
    torch.cuda.synchronize()
    outs = [torch.zeros_like(ea), torch.zeros_like(eb)]
    flags = torch.full_like(fa, 7)
    descs = [a.burst_desc(outs[0], eccm=ed, eccm_flags=flags), b.burst_desc(outs[1])]
    if not eccm_first:
        descs = descs[::-1]
    for rep in range(2):
        for o in outs:
            o.zero_()
        ctx.focus_batch(descs, F32, graph=graph)
        torch.cuda.synchronize()
        assert torch.equal(outs[0], ea), rep
        assert torch.equal(outs[1], eb), rep
        assert torch.equal(flags, fa), rep


@pytest.fixture(scope="module")
def iw_class_bursts():
    """Bursts of the IW azimuth lengths (compiled radix sequences): two of
    one length (one launch covers both), one each of the others."""
    specs = [synth.BurstSpec(n_lines=1480, n_samples=2400, rgdec=8, pri_code=21861, n_targets=20,
                             amp_mode="rayleigh", f_dc=30.0, seed=41),
             synth.BurstSpec(n_lines=1420, n_samples=2160, rgdec=6, pri_code=22276, n_targets=20,
                             amp_mode="rayleigh", seed=42),
             synth.BurstSpec(n_lines=1480, n_samples=2000, rgdec=8, pri_code=21861, n_targets=20,
                             amp_mode="rayleigh", seed=43),
             synth.BurstSpec(n_lines=1520, n_samples=2040, rgdec=11, pri_code=25868, n_targets=20,
                             amp_mode="rayleigh", seed=44)]
    return [pipeline.from_synth(synth.make_burst(s)) for s in specs]


@pytest.mark.parametrize("prec", [F32, F64])
@pytest.mark.parametrize("az_variant", [-1, 3, 4])
def test_batch_stages_match_burstwise(ctx, iw_class_bursts, prec, az_variant):
    """The stage-ordered batch (S1G_BATCH_STAGED: one launch per stage and
    shape class, the persistent kernels driven by per-burst descriptors)
    gives every burst the bits of its own chain, eager and as a replayed
    graph, as does the default burst-by-burst batch, with the azimuth
    variant chosen by the tuner or forced to either persistent form."""
    dt = torch.complex64 if prec == F32 else torch.complex128
    for st in ("az_fft", "az_ifft"):
        ctx.force_variant(st, az_variant)
    try:
        single = [db.focus(ctx, prec) for db in iw_class_bursts]
        torch.cuda.synchronize()
        outs = [torch.zeros((db.n_lines, db.n_samples), dtype=dt, device="cuda") for db in iw_class_bursts]
        descs = [db.burst_desc(o) for db, o in zip(iw_class_bursts, outs)]
        for staged, graph in ((True, False), (True, True), (True, True), (False, False)):
            for o in outs:
                o.zero_()
            ctx.focus_batch(descs, prec, graph=graph, staged=staged)
            torch.cuda.synchronize()
            for k, (o, e) in enumerate(zip(outs, single)):
                assert torch.equal(o, e), (staged, graph, k)
    finally:
        for st in ("az_fft", "az_ifft"):
            ctx.force_variant(st, -1)


def _af_problem(L=256, N=96, seed=3):
    rng = np.random.default_rng(seed)
    rd = (rng.standard_normal((L, N)) + 1j * rng.standard_normal((L, N)))
    rd[:, 10] += 40 * np.exp(1j * rng.uniform(-np.pi, np.pi, L))  # a bright column focuses under the right phase
    phase = rng.uniform(-1, 1, L)
    return rd, phase


def test_af_adjoint_f64_vs_oracle(ctx):
    rd, phase = _af_problem()
    e_ref, g_ref = oracle.af_entropy_grad(rd, phase)
    d_rd = torch.from_numpy(rd).cuda()
    d_ph = torch.from_numpy(phase).cuda()
    g = torch.zeros(rd.shape[0], dtype=torch.float64, device="cuda")
    e = ctx.af_adjoint(d_rd, d_ph, g)
    torch.cuda.synchronize()
    assert e == pytest.approx(e_ref, rel=1e-12)
    assert rel_l2(g.cpu().numpy(), g_ref) < 1e-9
    e2 = ctx.af_forward(d_rd, d_ph)
    assert e2 == pytest.approx(e_ref, rel=1e-12)


@pytest.mark.parametrize("prec,dt", [(F64, torch.complex128), (F32, torch.complex64)])
def test_af_finite_differences(ctx, prec, dt):
    """Central differences (eps = 1e-3, 32 seeded bins) of the FP64 entropy
    agree with the hand-written adjoint of either precision."""
    rd, phase = _af_problem(L=512, N=64, seed=5)
    d_rd = torch.from_numpy(rd).to(dt).cuda()
    d_rd64 = torch.from_numpy(rd).cuda()
    g = torch.zeros(rd.shape[0], dtype=torch.float64, device="cuda")
    ctx.af_adjoint(d_rd, torch.from_numpy(phase).cuda(), g)
    g = g.cpu().numpy()
    rng = np.random.default_rng(9)
    eps = 1e-3
    scale = np.abs(g).max()
    for i in rng.choice(rd.shape[0], 32, replace=False):
        p1 = phase.copy()
        p1[i] += eps
        p2 = phase.copy()
        p2[i] -= eps
        e1 = ctx.af_forward(d_rd64, torch.from_numpy(p1).cuda())
        e2 = ctx.af_forward(d_rd64, torch.from_numpy(p2).cuda())
        fd = (e1 - e2) / (2 * eps)
        tol = 1e-5 if prec == F64 else 1e-3
        assert abs(fd - g[i]) <= tol * scale, (i, fd, g[i])


@pytest.mark.parametrize("L,N,wide", [(256, 99, 0), (1024, 37, 5), (300, 40, 3)])
@pytest.mark.parametrize("prec,dt", [(F64, torch.complex128), (F32, torch.complex64)])
def test_af_shapes_vs_oracle(ctx, L, N, wide, prec, dt):
    """Lengths with compiled radices (256, 1024; a partial last column panel,
    a row stride wider than the row) and one without (300: the generic
    kernels that keep the image between the passes): entropy and gradient
    against the oracle, the image output against numpy."""
    rd, phase = _af_problem(L=L, N=N, seed=L + N)
    e_ref, g_ref = oracle.af_entropy_grad(rd, phase)
    full = torch.zeros((L, N + wide), dtype=dt, device="cuda")
    full[:, :N] = torch.from_numpy(rd).to(dt).cuda()
    d_rd = full[:, :N]
    d_ph = torch.from_numpy(phase).cuda()
    g = torch.zeros(L, dtype=torch.float64, device="cuda")
    e = ctx.af_adjoint(d_rd, d_ph, g)
    f64 = prec == F64
    assert e == pytest.approx(e_ref, rel=1e-12 if f64 else 1e-6)
    assert rel_l2(g.cpu().numpy(), g_ref) < (1e-9 if f64 else 1e-4)
    img = torch.empty((L, N), dtype=dt, device="cuda")
    e2 = ctx.af_forward(d_rd, d_ph, img)
    assert e2 == pytest.approx(e, rel=1e-12 if f64 else 1e-6)
    ref_img = np.fft.ifft(rd * np.exp(1j * phase)[:, None], axis=0)
    assert rel_l2(img.cpu().numpy(), ref_img) < (1e-12 if f64 else 1e-6)
