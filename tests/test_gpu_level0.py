"""GPU: Level-0 files through the header-parse kernel and the Level-0 product
(S1_Decoder's bookkeeping and products), effective velocities, secondary
range compression, the RCMC phase ramp and the FFT plan handle, against the
CPU FP64 oracle."""
import types

import numpy as np
import pytest

import oracle
from chain_ref import oracle_chain
from conftest import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from sentinel1_decode_amd import capi, pipeline, synth  # noqa: E402

F32, F64 = capi.S1G_F32, capi.S1G_F64


@pytest.fixture(scope="module")
def l0():
    specs = []
    for k in range(2):
        for sw, rg, pri in ((10, 6, 22276), (11, 8, 21861), (12, 11, 25868)):
            specs.append(synth.BurstSpec(n_lines=1100, n_samples=1024, rgdec=rg, pri_code=pri, n_targets=8,
                                         swath_number=sw, seed=200 + 3 * k + sw, clutter_sigma=0.1))
    s = synth.level0_stream(specs, lead_packets=300)
    offs = oracle.index_level0(s.stream)
    return s, offs, oracle.header_fields(s.stream, offs)


def test_parse_headers_kernel(ctx, l0):
    s, offs, F = l0
    d_stream = torch.from_numpy(s.stream).cuda()
    d_offs = torch.from_numpy(np.asarray(offs, np.int64)).cuda()
    f, uo, ul, nq = ctx.parse_headers(d_stream, d_offs)
    torch.cuda.synchronize()
    assert np.array_equal(f.cpu().numpy(), F)
    assert np.array_equal(uo.cpu().numpy(), np.asarray(offs) + 68)
    assert np.array_equal(ul.cpu().numpy(), F[:, 7] + 1 - 62)
    assert np.array_equal(nq.cpu().numpy(), F[:, oracle.SEC["num_quadratures"]])


def _oracle_burst(s, idx, spec):
    """The packets idx of the stream as the oracle chain takes them."""
    offs = np.asarray(oracle.index_level0(s.stream), np.uint64)[idx] + 68
    F = oracle.header_fields(s.stream, [int(o) - 68 for o in offs])
    return types.SimpleNamespace(stream=s.stream, ud_offsets=offs, ud_lengths=(F[:, 7] + 1 - 62).astype(np.uint32),
                                 num_quads=F[:, oracle.SEC["num_quadratures"]].astype(np.int32), spec=spec)


def test_level0_product(ctx, l0):
    s, offs, F = l0
    prod = pipeline.Level0Product(s.stream, ctx)
    assert prod.swaths() == {"IW1": 2500, "IW2": 2200, "IW3": 2200}
    for sw, name in ((10, "IW1"), (11, "IW2"), (12, "IW3")):
        built = [list(range(first, first + n)) for swn, first, n, _ in s.bursts if swn == sw]
        if sw == 10:
            built = built[1:]
        assert prod.n_bursts(name) == len(built)
        for b, want in enumerate(built):
            assert prod.packet_indices(name, b).tolist() == want
    with pytest.raises(KeyError):
        prod.get_burst("EW1", 0)
    with pytest.raises(IndexError):
        prod.get_burst("IW2", 5)
    # decode of one burst: bit-exact against the oracle
    idx = prod.packet_indices("IW2", 1)
    spec = [sp for swn, first, n, sp in s.bursts if first == idx[0]][0]
    ob = _oracle_burst(s, idx, spec)
    dec = prod.get_burst("IW2", 1, prec=F64).cpu().numpy()
    ref_dec, st = oracle.decode_packets(ob.stream, ob.ud_offsets, ob.ud_lengths, ob.num_quads)
    assert (st == 0).all()
    assert np.array_equal(dec.view(np.uint64), ref_dec.view(np.uint64))
    # focusing with the orbit's effective velocities: FP64 against the oracle
    db = prod.burst("IW2", 1)
    v = db.geo.v_rel
    assert 6500 < v.mean() < 7800 and np.all(np.diff(v) < 0)
    ref = oracle_chain(ob, v_rel=v, f_dc=0.0)
    slc = prod.get_azimuth_compressed_burst("IW2", 1, prec=F64).cpu().numpy()
    assert rel_l2(slc, ref["slc"]) < 1e-10
    rc = prod.get_range_compressed_burst("IW2", 1, range_doppler=True, prec=F64).cpu().numpy()
    assert rel_l2(rc, ref["rd"]) < 1e-11
    # a swath: its bursts in one batch, 20 zero lines after each
    sw = prod.get_azimuth_compressed_swath("IW2", prec=F32).cpu().numpy()
    L = db.n_lines
    assert sw.shape[0] == sum(len(x) + 20 for x in (prod.packet_indices("IW2", b) for b in range(2)))
    one = prod.get_azimuth_compressed_burst("IW2", 1, prec=F32).cpu().numpy()
    start = len(prod.packet_indices("IW2", 0)) + 20
    assert np.array_equal(sw[start: start + L], one)
    assert not sw[start + L: start + L + 20].any()


def test_effective_velocities_vs_oracle(ctx):
    rng = np.random.default_rng(1)
    pos = np.stack([rng.uniform(-7.1e6, 7.1e6, 5), rng.uniform(-2e6, 2e6, 5), rng.uniform(-6e6, 6e6, 5)], 1)
    vn = rng.uniform(7400, 7600, 5)
    R = np.linspace(800e3, 950e3, 3000)
    got = ctx.effective_velocities(torch.from_numpy(pos).cuda(), torch.from_numpy(vn).cuda(),
                                   torch.from_numpy(R).cuda()).cpu().numpy()
    for i in range(5):
        ref = oracle.effective_velocities(pos[i], vn[i], R)
        assert np.max(np.abs(got[i] / ref - 1)) < 1e-14


@pytest.mark.parametrize("prec,dt,tol", [(F64, torch.complex128, 1e-13), (F32, torch.complex64, 1e-6)])
def test_src_and_rcmc_phase_vs_oracle(ctx, prec, dt, tol):
    rng = np.random.default_rng(2)
    rows, N = 6, 4000
    x = rng.standard_normal((rows, N)) + 1j * rng.standard_normal((rows, N))
    R = np.linspace(800e3, 880e3, N)
    v = rng.uniform(6900, 7200, (rows, N))
    fr = np.fft.fftfreq(N) * 64.34523813e6
    fdc = np.linspace(-20, 45, N)
    faz = np.linspace(-800, 800, rows)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    X = d(x).to(dt)
    fac = torch.zeros((rows, N), dtype=torch.float64, device="cuda")
    ctx.secondary_range_compression(X, d(v), d(R), d(fr), d(fdc), d(faz), fac)
    Y = X.clone()
    ctx.rcmc_phase(Y, d(R), d(fr), fac)
    X, Y, fac = X.cpu().numpy(), Y.cpu().numpy(), fac.cpu().numpy()
    for i in range(rows):
        rx, rf = oracle.secondary_range_compression(x[i], v[i], R, fr, fdc, faz[i])
        assert np.array_equal(fac[i], rf)
        assert rel_l2(X[i], rx) < tol
        ry = oracle.rcmc_phase(rx, R, fr, rf)
        assert rel_l2(Y[i], ry) < tol


@pytest.mark.parametrize("n", [8, 1480, 1867, 24000])
def test_fft_plan_unnormalised(ctx, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    for inverse in (False, True):
        p = ctx.plan(F64, n, inverse)
        d = torch.from_numpy(x).cuda()
        p.execute(d)
        want = np.fft.ifft(x, axis=1) * n if inverse else np.fft.fft(x, axis=1)
        assert rel_l2(d.cpu().numpy(), want) < 1e-13
        p.close()


@pytest.mark.parametrize("mode", ["norm_log", "norm", "mag", "real", "imag"])
@pytest.mark.parametrize("dt", [torch.complex128, torch.complex64])
def test_scale_vs_oracle(ctx, mode, dt):
    rng = np.random.default_rng(7)
    x = (rng.standard_normal((300, 1000)) + 1j * rng.standard_normal((300, 1000))) * 1e4
    x[17, 400] = 1e7 + 1e7j
    d = torch.from_numpy(x).to(dt).cuda()
    got = ctx.scale(d, mode).cpu().numpy()
    ref = oracle.scale_2d(d.cpu().numpy(), mode)
    # FP64 |x| (hypot) on both sides, rounded once to float: within an ulp
    assert np.all(np.abs(got - ref) <= np.spacing(np.abs(ref)))


def test_to_cf32(ctx):
    rng = np.random.default_rng(8)
    x = rng.standard_normal((50, 333)) + 1j * rng.standard_normal((50, 333))
    got = ctx.to_cf32(torch.from_numpy(x).cuda()).cpu().numpy()
    assert np.array_equal(got, x.astype(np.complex64))


def test_tops_burst_vs_oracle(ctx, l0):
    """Level0Product.get_tops_burst (the TOPS branch on the device) against
    the oracle's TOPS chain of the same burst, FP64 and FP32."""
    from chain_ref import oracle_tops
    s, offs, F = l0
    prod = pipeline.Level0Product(s.stream, ctx)
    for swath in ("IW1", "IW2"):
        idx = prod.packet_indices(swath, 1)
        uo = np.asarray(offs, np.uint64)[idx] + 68
        Fb = F[idx]
        ob = types.SimpleNamespace(stream=s.stream, ud_offsets=uo,
                                   ud_lengths=(Fb[:, 7] + 1 - 62).astype(np.uint32),
                                   num_quads=Fb[:, oracle.SEC["num_quadratures"]].astype(np.int32),
                                   spec=[sp for _, first, _, sp in s.bursts if first == idx[0]][0])
        want = oracle_tops(ob, Fb, oracle.state_vectors(F))["slc"]
        got = prod.get_tops_burst(swath, 1, prec=F64).cpu().numpy()
        assert got.shape == want.shape
        # stage by stage on the same inputs the GPU is within ~1e-15 of the
        # oracle; from packets, the range lines differ at ~5e-14 (two FFT
        # algorithms), the centroid estimated from them at ~5e-15, and the
        # line filter's phase 4 pi R D / lambda (~2e8 rad) turns the D values
        # whose last bit that flips into ~2e-8 rad: ~1e-10 end to end
        assert rel_l2(got, want) < 1e-9, swath
        # FP32: the centroid estimated from FP32 range lines moves by ~1e-7
        # relative (a few 1e-5 Hz); the time unfolding's ramps multiply that by
        # kt t ~ 1e4 (rad / Hz), so the end-to-end error is ~1e-4
        got32 = prod.get_tops_burst(swath, 1, prec=F32).cpu().numpy()
        e32 = rel_l2(got32, want)
        print(f"{swath}: FP64 {rel_l2(got, want):.3g} FP32 {e32:.3g}")
        assert e32 < 2e-3, swath
