"""Every launch variant the autotuner can pick gives the oracle's result.

Each tunable stage is run with each of its variants forced
(s1g_force_variant) and compared with the CPU oracle at the stage's
tolerance; then the automatic choice is checked to be one of the
candidates, and a graph captured with a forced variant to replay the eager
result.
"""
import numpy as np
import pytest

import oracle
from chain_ref import oracle_chain
from conftest import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from sentinel1_decode_amd import capi, pipeline, synth  # noqa: E402

F32, F64 = capi.S1G_F32, capi.S1G_F64


@pytest.fixture
def vctx():
    c = capi.Context(0)
    yield c
    c.close()


def _variant_count(ctx, stage, run):
    ctx.force_variant(stage, 0)
    run()
    torch.cuda.synchronize()
    return ctx.last_variant(stage)[1]


@pytest.mark.parametrize("prec,tol", [(F64, 1e-11), (F32, 2e-6)])
def test_range_variants(vctx, prec, tol):
    rng = np.random.default_rng(5)
    N = 24000
    x = (rng.standard_normal((3, N)) + 1j * rng.standard_normal((3, N))) * 20
    txpsf, txprr, txpl, fs = -8.7e-1, 1.0789, 52.40481, 64.34523813
    ref = oracle.range_compress(x, txpsf, txprr, txpl, fs)
    dt = torch.complex128 if prec == F64 else torch.complex64
    d = torch.from_numpy(x).to(dt).cuda()
    desc = capi.RangeDesc(3, N, txpsf, txprr, txpl, fs)
    n = _variant_count(vctx, "range", lambda: vctx.range_compress(d, desc))
    assert n == (2 if prec == F32 else 1)
    for v in range(n):
        vctx.force_variant("range", v)
        y = vctx.range_compress(d, desc)
        torch.cuda.synchronize()
        assert vctx.last_variant("range") == (v, n)
        assert rel_l2(y.cpu().numpy(), ref) < tol, v


@pytest.mark.parametrize("L", [1480, 1420, 1520])
@pytest.mark.parametrize("inverse", [False, True])
def test_azimuth_variants(vctx, L, inverse):
    rng = np.random.default_rng(L + inverse)
    a = rng.standard_normal((L, 100)) + 1j * rng.standard_normal((L, 100))
    ref = oracle.azimuth_ifft(a) if inverse else oracle.azimuth_fft(a)
    stage = "az_ifft" if inverse else "az_fft"
    for prec, dt, tol in ((F64, torch.complex128, 1e-13), (F32, torch.complex64, 2e-6)):
        x0 = torch.from_numpy(a).to(dt).cuda()
        n = _variant_count(vctx, stage, lambda: vctx.azimuth_fft(x0.clone(), inverse=inverse))
        assert n == 5
        for v in range(n):
            vctx.force_variant(stage, v)
            x = x0.clone()
            vctx.azimuth_fft(x, inverse=inverse)
            torch.cuda.synchronize()
            assert vctx.last_variant(stage) == (v, n)
            assert rel_l2(x.cpu().numpy(), ref) < tol, (prec, v)


@pytest.mark.parametrize("inverse", [False, True])
def test_azimuth_persistent_many_panels(vctx, inverse):
    """The persistent variants (3; 4 with the first and last passes in
    registers) walk several panels per workgroup, the last one partial:
    4500 columns = 563 panels of 8 (FP32) / 1125 of 4 (FP64)."""
    L, N = 1480, 4500
    rng = np.random.default_rng(11 + inverse)
    a = rng.standard_normal((L, N)) + 1j * rng.standard_normal((L, N))
    ref = oracle.azimuth_ifft(a) if inverse else oracle.azimuth_fft(a)
    stage = "az_ifft" if inverse else "az_fft"
    for prec, dt, tol in ((F64, torch.complex128, 1e-13), (F32, torch.complex64, 2e-6)):
        for v in (3, 4):
            vctx.force_variant(stage, v)
            x = torch.from_numpy(a).to(dt).cuda()
            vctx.azimuth_fft(x, inverse=inverse)
            torch.cuda.synchronize()
            assert vctx.last_variant(stage) == (v, 5)
            assert rel_l2(x.cpu().numpy(), ref) < tol, (prec, v)
    vctx.force_variant(stage, -1)


@pytest.fixture(scope="module")
def rd_case():
    spec = synth.config0(n_lines=192, n_samples=4096, grid_spacing=(300, 40), seed=7)
    burst = synth.make_burst(spec)
    return burst, oracle_chain(burst)


@pytest.mark.parametrize("prec,tol", [(F64, 1e-11), (F32, 2e-6)])
def test_rcmc_variants(vctx, rd_case, prec, tol):
    burst, ch = rd_case
    db = pipeline.from_synth(burst)
    dt = torch.complex128 if prec == F64 else torch.complex64
    x = torch.from_numpy(ch["rd"]).to(dt).cuda()
    n = _variant_count(vctx, "rcmc_mf", lambda: vctx.rcmc_matched_filter(x, db.azimuth_desc()))
    assert n == (7 if prec == F32 else 3)
    for v in range(n):
        vctx.force_variant("rcmc_mf", v)
        y = vctx.rcmc_matched_filter(x, db.azimuth_desc())
        torch.cuda.synchronize()
        assert vctx.last_variant("rcmc_mf") == (v, n)
        assert rel_l2(y.cpu().numpy(), ch["mf"]) < tol, v


@pytest.mark.parametrize("lam_scale", [1.0, 4.0, 12.0, 60.0])
def test_rcmc_runs_against_fp64(vctx, rd_case, lam_scale):
    """FP32 run kernels (variants 3, 4; linear geometry 5, 6) against the FP64
    path on the same input.
    A longer wavelength raises u = lambda^2 f^2 / (4 v^2) (about 1e-5 at
    the band edge here): x4 spreads the fractional offsets of the outer
    rows' runs past the interpolation limit (per-sample weights), x12 and
    x60 take u past the series limit (square-root geometry; at x60 the
    migration also outgrows the LDS tile)."""
    burst, ch = rd_case
    db = pipeline.from_synth(burst)
    desc = db.azimuth_desc()
    desc.wavelength *= lam_scale
    x64 = torch.from_numpy(ch["rd"]).cuda()
    vctx.force_variant("rcmc_mf", 0)
    ref = vctx.rcmc_matched_filter(x64, desc).cpu().numpy()
    x32 = x64.to(torch.complex64)
    for v in (0, 3, 4, 5, 6):
        vctx.force_variant("rcmc_mf", v)
        y = vctx.rcmc_matched_filter(x32, desc)
        torch.cuda.synchronize()
        assert vctx.last_variant("rcmc_mf") == (v, 7)
        assert rel_l2(y.cpu().numpy(), ref) < 2e-6, (lam_scale, v)
    vctx.force_variant("rcmc_mf", -1)


@pytest.mark.parametrize("field", ["v_rel", "f_dc", "slant_ranges"])
def test_rcmc_linear_kernel_general_geometry(vctx, rd_case, field):
    """The linear-geometry kernels (variants 5, 6) on columns that break
    their assumptions (v_rel or f_dc varying with range, slant ranges off a
    straight line): the per-sample path against the FP64 path."""
    burst, ch = rd_case
    db = pipeline.from_synth(burst)
    desc = db.azimuth_desc()
    n = db.n_samples
    ramp = torch.linspace(0.0, 1.0, n, dtype=torch.float64, device="cuda")
    if field == "v_rel":
        arr = db.d_v + 25.0 * ramp
        desc.v_rel = arr.data_ptr()
    elif field == "f_dc":
        arr = db.d_fdc + 40.0 * ramp
        desc.f_dc = arr.data_ptr()
    else:
        arr = db.d_r + 3.0 * ramp * ramp
        desc.slant_ranges = arr.data_ptr()
    x64 = torch.from_numpy(ch["rd"]).cuda()
    vctx.force_variant("rcmc_mf", 0)
    ref = vctx.rcmc_matched_filter(x64, desc).cpu().numpy()
    x32 = x64.to(torch.complex64)
    for v in (5, 6):
        vctx.force_variant("rcmc_mf", v)
        y = vctx.rcmc_matched_filter(x32, desc)
        torch.cuda.synchronize()
        assert rel_l2(y.cpu().numpy(), ref) < 2e-6, (field, v)
    vctx.force_variant("rcmc_mf", -1)
    del arr


def test_autotuned_chain_and_forced_graph(vctx):
    """Automatic choice stays among the candidates and matches a forced
    run; a graph captured under forced variants replays the eager result."""
    spec = synth.config1(n_lines=256, n_samples=24000)
    burst = synth.make_burst(spec)
    db = pipeline.from_synth(burst)
    auto = db.focus(vctx, F32)
    torch.cuda.synchronize()
    for st in ("range", "az_fft", "rcmc_mf", "az_ifft"):
        v, n = vctx.last_variant(st)
        assert 0 <= v < n, st
    for st in ("range", "az_fft", "rcmc_mf", "az_ifft"):
        vctx.force_variant(st, 1)
    forced = db.focus(vctx, F32)
    torch.cuda.synchronize()
    assert rel_l2(forced.cpu().numpy(), auto.cpu().numpy()) < 1e-6
    out = torch.empty_like(forced)
    desc = db.burst_desc(out)
    for rep in range(3):
        out.zero_()
        vctx.focus_batch([desc], F32, graph=True)
        torch.cuda.synchronize()
        assert torch.equal(out, forced), rep


def test_tuning_cache_file(tmp_path):
    """Choices timed in one context, saved, and loaded into a context that
    may not time anything, give the same variants and the same bits; a file
    from another kind of device and a malformed one are refused."""
    burst = synth.make_burst(synth.config0(n_lines=256, n_samples=2048))
    db = pipeline.from_synth(burst)
    a = capi.Context(0)
    out_a = db.focus(a, F32).clone()
    torch.cuda.synchronize()
    stages = ("range", "az_fft", "rcmc_mf", "az_ifft")
    chosen = {s: a.last_variant(s) for s in stages}
    path = tmp_path / "tune.txt"
    a.tune_save(str(path))
    a.close()
    text = path.read_text().splitlines()
    assert text[0].startswith("s1gpu-tune 1 ") and len(text) == 1 + int(text[0].split()[-1])
    b = capi.Context(0)
    b.set_autotune(False)
    b.tune_load(str(path))
    out_b = db.focus(b, F32)
    torch.cuda.synchronize()
    assert {s: b.last_variant(s) for s in stages} == chosen
    assert torch.equal(out_a, out_b)
    head = text[0].split()
    foreign = tmp_path / "foreign.txt"
    foreign.write_text(" ".join(head[:2] + ["gfx000:1"] + head[3:]) + "\n" + "\n".join(text[1:]) + "\n")
    with pytest.raises(capi.S1GError):
        b.tune_load(str(foreign))
    bad = tmp_path / "bad.txt"
    bad.write_text(text[0] + "\n1 2 3\n")
    with pytest.raises(capi.S1GError):
        b.tune_load(str(bad))
    b.close()


@pytest.mark.parametrize("n", [1480, 4902])
@pytest.mark.parametrize("prec,tol", [(F64, 1e-12), (F32, 3e-6)])
def test_dft_radix_orders(vctx, n, prec, tol):
    """Every radix order the DFT tuner can pick for rows that fit LDS gives
    numpy's transform (1480 = 37 * 5 * 8: four orders; 4902 = 43 * 19 * 3 * 2:
    three distinct ones)."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))
    dt = torch.complex128 if prec == F64 else torch.complex64
    x0 = torch.from_numpy(x).to(dt).cuda()
    nv = _variant_count(vctx, "dft", lambda: vctx.dft(x0.clone(), axis=1))
    assert nv == (4 if n == 1480 else 3)
    for inverse in (False, True):
        ref = np.fft.ifft(x, axis=1) if inverse else np.fft.fft(x, axis=1)
        for v in range(nv):
            vctx.force_variant("dft", v)
            d = x0.clone()
            vctx.dft(d, axis=1, inverse=inverse)
            torch.cuda.synchronize()
            assert vctx.last_variant("dft") == (v, nv)
            assert rel_l2(d.cpu().numpy(), ref) < tol, (v, inverse)
    vctx.force_variant("dft", -1)
