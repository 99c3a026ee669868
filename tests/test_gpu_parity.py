"""GPU parity: the HIP path (through the C ABI) against the CPU FP64 oracle.

Bars: decode bit-exact (F64 doubles equal, F32 equals the doubles rounded
once); FP64 stages within 1e-11..1e-10 relative L2 of the oracle (the FFT
algorithms differ, so round-off differs at the 1e-15 level per pass); FP32
stages within 1e-5 relative L2 of FP64.
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


def _bits_equal(got, ref, prec):
    if prec == F64:
        return np.array_equal(np.ascontiguousarray(got).view(np.uint64), np.ascontiguousarray(ref).view(np.uint64))
    r32 = ref.astype(np.complex64)
    return np.array_equal(np.ascontiguousarray(got).view(np.uint32), r32.view(np.uint32))


@pytest.fixture(scope="module")
def small_burst():
    spec = synth.config0(n_lines=192, n_samples=4096, grid_spacing=(300, 40), seed=7)
    return synth.make_burst(spec)


@pytest.fixture(scope="module")
def small_chain(small_burst):
    return oracle_chain(small_burst)


@pytest.mark.parametrize("prec", [F64, F32])
def test_decode_bitexact_burst(ctx, small_burst, prec):
    db = pipeline.from_synth(small_burst)
    out, st = db.decode(ctx, prec)
    torch.cuda.synchronize()
    ref, rst = oracle.decode_packets(small_burst.stream, small_burst.ud_offsets, small_burst.ud_lengths,
                                     small_burst.num_quads)
    assert (rst == 0).all()
    assert (st.cpu().numpy() == 0).all()
    assert _bits_equal(out.cpu().numpy(), ref, prec)


def _edge_packets():
    """Packets that exercise the decoder's corner cases, with the oracle's
    verdict for each: odd num_quads (short last block), one quad, a packet
    truncated inside a code, one cut inside a header field, BRC 5-7, an
    empty packet, and user data at unaligned byte offsets."""
    rng = np.random.default_rng(11)
    packets, nqs = [], []
    for nq in (1, 2, 127, 128, 129, 130, 255, 300, 1000):
        nb = (nq + 127) // 128
        iq = (rng.standard_normal(2 * nq) + 1j * rng.standard_normal(2 * nq)) * rng.uniform(0.5, 60)
        brc = rng.integers(0, 5, nb).astype(np.uint8)
        th = rng.integers(0, 256, nb).astype(np.uint8)
        th[rng.random(nb) < 0.3] = rng.integers(0, 4)
        packets.append(synth.encode_packet(iq, brc, th))
        nqs.append(nq)
    base = packets[5]
    packets.append(base[: len(base) // 2])          # truncated inside the data
    nqs.append(130)
    packets.append(base[:1])                        # truncated inside the first BRC/ symbols
    nqs.append(130)
    for bad in (5, 6, 7):                           # BRC field > 4
        b = bytearray(packets[3])
        b[0] = (b[0] & 0x1F) | (bad << 5)
        packets.append(bytes(b))
        nqs.append(129)
    packets.append(b"\x00\x00\x00\x00")             # num_quads = 0
    nqs.append(0)
    return packets, nqs


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("misalign", [0, 1, 3])
def test_decode_edge_cases(ctx, prec, misalign):
    packets, nqs = _edge_packets()
    offs, blob = [], bytearray()
    for p in packets:
        blob += b"\xA5" * (misalign + 5)  # junk between packets
        offs.append(len(blob))
        blob += p
    blob += b"\x5A" * 16
    stream = np.frombuffer(bytes(blob), np.uint8)
    lens = np.array([len(p) for p in packets], np.uint32)
    offs = np.array(offs, np.uint64)
    nq = np.array(nqs, np.int32)
    maxq = int(nq.max())
    d_stream = torch.from_numpy(stream.copy()).cuda()
    out, st = ctx.decode(d_stream, torch.from_numpy(offs.view(np.int64)).cuda(),
                         torch.from_numpy(lens.view(np.int32)).cuda(), torch.from_numpy(nq).cuda(),
                         int(lens.max()), maxq, prec)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    st = st.cpu().numpy()
    for i, p in enumerate(packets):
        rst, sig, _, _ = oracle.decode_packet(p, int(nq[i]))
        assert st[i] == rst, (i, st[i], rst)
        n = 2 * int(nq[i])
        if rst == 0:
            assert _bits_equal(got[i, :n], sig, prec), i
        else:
            assert not np.any(got[i, :n]), i
    assert set(st.tolist()) >= {0, 1, 2}


@pytest.mark.parametrize("prec", [F64, F32])
def test_decode_bitexact_large_packets(ctx, prec):
    """RFI lines compress badly (BRC 4, ~44 KB packets, above the 20 KB the
    decoder stages in LDS): packets of both kinds in one launch decode
    bit-exactly."""
    spec = synth.config1(n_lines=48, n_samples=24000, seed=77)
    raw = spec.raw()
    mask = synth.add_rfi(raw, spec, seed=78, frac=0.4, run=(5, 10))
    burst = synth.encode_burst(spec, raw)
    lens = np.asarray(burst.ud_lengths)
    assert lens[mask].min() > 20 * 1024 > lens[~mask].max()
    db = pipeline.from_synth(burst)
    out, st = db.decode(ctx, prec)
    torch.cuda.synchronize()
    ref, rst = oracle.decode_packets(burst.stream, burst.ud_offsets, burst.ud_lengths, burst.num_quads)
    assert (rst == 0).all() and (st.cpu().numpy() == 0).all()
    assert _bits_equal(out.cpu().numpy(), ref, prec)


@pytest.mark.parametrize("prec", [F64, F32])
def test_decode_bitexact_more_than_128_blocks(ctx, prec):
    """36,000-sample lines: 141 BAQ blocks per channel, more than the 128
    lanes of the final value round, so every IE/IO/QE item stays with wave 1
    and QO takes two rounds; bit-exact against the oracle."""
    spec = synth.config1(n_lines=6, n_samples=36000, n_targets=20, seed=91)
    burst = synth.make_burst(spec)
    assert ((np.asarray(burst.num_quads) + 127) // 128 > 128).all()
    db = pipeline.from_synth(burst)
    out, st = db.decode(ctx, prec)
    torch.cuda.synchronize()
    ref, rst = oracle.decode_packets(burst.stream, burst.ud_offsets, burst.ud_lengths, burst.num_quads)
    assert (rst == 0).all() and (st.cpu().numpy() == 0).all()
    assert _bits_equal(out.cpu().numpy(), ref, prec)


@pytest.mark.parametrize("prec,tol", [(F64, 1e-11), (F32, 2e-6)])
def test_range_compress(ctx, small_burst, small_chain, prec, tol):
    db = pipeline.from_synth(small_burst)
    dt = torch.complex128 if prec == F64 else torch.complex64
    x = torch.from_numpy(small_chain["decoded"]).to(dt).cuda()
    y = ctx.range_compress(x, db.geo.range_desc())
    torch.cuda.synchronize()
    assert rel_l2(y.cpu().numpy(), small_chain["rc"]) < tol


@pytest.mark.parametrize("prec,tol", [(F64, 1e-11), (F32, 2e-6)])
@pytest.mark.parametrize("n_samples,fs", [(24000, 64.34523813), (21600, 50.04629632), (20400, 54.59595962)])
def test_range_compress_iw_lengths(ctx, prec, tol, n_samples, fs):
    """Full IW1/IW2/IW3 line lengths (two overlap-save blocks per line in
    FP32, five in FP64), a few lines, against the oracle's literal
    pad / FFT / multiply / IFFT / fftshift / crop."""
    rng = np.random.default_rng(n_samples)
    x = (rng.standard_normal((3, n_samples)) + 1j * rng.standard_normal((3, n_samples))) * 20
    txpsf, txprr, txpl = -8.7e-1, 1.0789, 52.40481
    ref = oracle.range_compress(x, txpsf, txprr, txpl, fs)
    dt = torch.complex128 if prec == F64 else torch.complex64
    d = torch.from_numpy(x).to(dt).cuda()
    y = ctx.range_compress(d, capi.RangeDesc(3, n_samples, txpsf, txprr, txpl, fs))
    torch.cuda.synchronize()
    assert rel_l2(y.cpu().numpy(), ref) < tol


@pytest.mark.parametrize("L", [192, 1480, 1420, 1520, 1024])
def test_azimuth_fft_lengths(ctx, L):
    rng = np.random.default_rng(L)
    a = rng.standard_normal((L, 100)) + 1j * rng.standard_normal((L, 100))
    ref_f = oracle.azimuth_fft(a)
    ref_i = oracle.azimuth_ifft(a)
    for prec, dt, tol in ((F64, torch.complex128, 1e-13), (F32, torch.complex64, 2e-6)):
        x = torch.from_numpy(a).to(dt).cuda()
        ctx.azimuth_fft(x, inverse=False)
        y = torch.from_numpy(a).to(dt).cuda()
        ctx.azimuth_fft(y, inverse=True)
        torch.cuda.synchronize()
        assert rel_l2(x.cpu().numpy(), ref_f) < tol, (prec, L)
        assert rel_l2(y.cpu().numpy(), ref_i) < tol, (prec, L)


@pytest.mark.parametrize("prec,tol", [(F64, 1e-11), (F32, 2e-6)])
def test_rcmc_matched_filter(ctx, small_burst, small_chain, prec, tol):
    db = pipeline.from_synth(small_burst)
    dt = torch.complex128 if prec == F64 else torch.complex64
    x = torch.from_numpy(small_chain["rd"]).to(dt).cuda()
    y = ctx.rcmc_matched_filter(x, db.azimuth_desc())
    torch.cuda.synchronize()
    assert rel_l2(y.cpu().numpy(), small_chain["mf"]) < tol


def test_focus_burst_f64_vs_oracle(ctx, small_burst, small_chain):
    db = pipeline.from_synth(small_burst)
    slc = db.focus(ctx, F64)
    torch.cuda.synchronize()
    assert (db.d_status.cpu().numpy() == 0).all()
    assert rel_l2(slc.cpu().numpy(), small_chain["slc"]) < 1e-10


def test_focus_burst_f32_vs_f64(ctx, small_burst):
    db = pipeline.from_synth(small_burst)
    a = db.focus(ctx, F64).cpu().numpy()
    b = db.focus(ctx, F32).cpu().numpy()
    torch.cuda.synchronize()
    assert rel_l2(b, a) < 1e-5


def _rfi_lines(seed=21, L=24, N=3000):
    """Random clutter lines with strong tones on a third of them, and ECCM
    thresholds placed far from every decision the data produces."""
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((L, N)) + 1j * rng.standard_normal((L, N))) * 10
    n = np.arange(N)
    for l in range(0, L, 3):
        for _ in range(3):
            x[l] += 300 * np.exp(2j * np.pi * (rng.uniform(-0.45, 0.45) * n + rng.uniform()))
    spec_max = np.abs(np.fft.fft(x, axis=1)).max(1)
    det = 0.5 * (spec_max[1::3].max() + spec_max[0::3].min())
    return x, det, 2000.0


# N = 3000: the detection spectrum in one sub-spectrum; 21600 and 20400 (IW1
# and IW3 lines): four (FP32) or eight (FP64) sub-spectra, 20400 with a
# radix-17 pass.
@pytest.mark.parametrize("N", [3000, 21600, 20400])
@pytest.mark.parametrize("prec", [F64, F32])
def test_eccm_vs_oracle(ctx, prec, N):
    x, det, mit = _rfi_lines(L=24 if N == 3000 else 12, N=N)
    ref, rflags = oracle.eccm(x, det, mit)
    dt = torch.complex128 if prec == F64 else torch.complex64
    d = torch.from_numpy(x).to(dt).cuda()
    flags = torch.zeros(x.shape[0], dtype=torch.uint8, device="cuda")
    desc = capi.EccmDesc(x.shape[0], x.shape[1], 128, 64, det, mit)
    ctx.eccm(d, desc, flags)
    torch.cuda.synchronize()
    assert np.array_equal(flags.cpu().numpy(), rflags)
    assert rflags.sum() == x.shape[0] // 3
    tol = 1e-12 if prec == F64 else 1e-5
    assert rel_l2(d.cpu().numpy(), ref) < tol


def test_focus_burst_with_eccm_f64(ctx):
    spec = synth.config0(n_lines=160, n_samples=4096, grid_spacing=(300, 40), seed=13)
    raw = spec.raw()
    mask = synth.add_rfi(raw, spec, seed=3, frac=0.2, run=(10, 20))
    burst = synth.encode_burst(spec, raw)
    dec, _ = oracle.decode_packets(burst.stream, burst.ud_offsets, burst.ud_lengths, burst.num_quads)
    smax = np.abs(np.fft.fft(dec, axis=1)).max(1)
    det = 0.5 * (smax[~mask].max() + smax[mask].min())
    assert smax[~mask].max() < det < smax[mask].min()
    mit = 4000.0
    ref = oracle_chain(burst, eccm=(det, mit))
    assert np.array_equal(ref["eccm_flags"].astype(bool), mask)
    db = pipeline.from_synth(burst)
    flags = torch.zeros(spec.n_lines, dtype=torch.uint8, device="cuda")
    slc = db.focus(ctx, F64, eccm=capi.EccmDesc(spec.n_lines, spec.n_samples, 128, 64, det, mit), eccm_flags=flags)
    torch.cuda.synchronize()
    assert np.array_equal(flags.cpu().numpy().astype(bool), mask)
    assert rel_l2(slc.cpu().numpy(), ref["slc"]) < 1e-10
