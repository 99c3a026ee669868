"""GPU regression tests for host-side faults that were fixed:

* a captured batch graph replayed after the context's workspace was
  reallocated (a larger or FP64 call) must not write the freed buffers: it
  is captured again and gives the eager result;
* a packet whose num_quads exceeds the stated maximum is rejected with
  S1G_PKT_ERR_ARG and writes nothing outside its own row;
* s1g_check_status reports the first failed packet (S1G_EDECODE);
* the FP32 RCMC interpolation weights at an offset whose FP64 fraction lies
  within 2^-25 below 1 (it rounds to 1.0f): on the IW1- and IW3-like burst
  geometries this happens at (132, 5776) and (83, 14526) and gave NaN
  (0 * inf on the k = 1 tap), which the azimuth IFFT spread over the column.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from sentinel1_decode_amd import capi, pipeline, synth  # noqa: E402

F32, F64 = capi.S1G_F32, capi.S1G_F64


def test_graph_replay_after_workspace_reallocation():
    ctx = capi.Context(0)  # a fresh context: its workspace starts at this batch's size
    try:
        small = [pipeline.from_synth(synth.make_burst(synth.config0(n_lines=96, n_samples=2048, seed=41 + k,
                                                                    grid_spacing=(300, 20))))
                 for k in range(2)]
        big = pipeline.from_synth(synth.make_burst(synth.config0(n_lines=256, n_samples=4096, seed=43,
                                                                 grid_spacing=(300, 40))))
        eager = [db.focus(ctx, F32) for db in small]
        torch.cuda.synchronize()
        outs = [torch.zeros_like(e) for e in eager]
        descs = [db.burst_desc(o) for db, o in zip(small, outs)]
        ctx.focus_batch(descs, F32, graph=True)
        torch.cuda.synchronize()
        for o, e in zip(outs, eager):
            assert torch.equal(o, e)
        # FP64 on a larger burst: the workspace (and column table) grow,
        # freeing the buffers the captured graph was built on
        ref64 = big.focus(ctx, F64)
        torch.cuda.synchronize()
        for rep in range(3):
            for o in outs:
                o.zero_()
            ctx.focus_batch(descs, F32, graph=True)
            torch.cuda.synchronize()
            for o, e in zip(outs, eager):
                assert torch.equal(o, e), rep
            # the FP64 result, computed before the replays, is untouched
            assert torch.equal(big.focus(ctx, F64), ref64), rep
    finally:
        ctx.close()


def test_decode_rejects_num_quads_above_maximum(ctx):
    """Three packets; the middle one claims 100 quads more than the
    max_num_quads the call states.  It fails with S1G_PKT_ERR_ARG, its row
    is zero-filled, and the rows around it (filled with a sentinel where
    nothing should be written) hold exactly the oracle's decode."""
    rng = np.random.default_rng(5)
    nqs = [600, 700, 600]
    packets = []
    for nq in nqs:
        nb = (nq + 127) // 128
        iq = (rng.standard_normal(2 * nq) + 1j * rng.standard_normal(2 * nq)) * 20
        packets.append(synth.encode_packet(iq, np.full(nb, 2, np.uint8), np.full(nb, 40, np.uint8)))
    offs, blob = [], bytearray()
    for p in packets:
        offs.append(len(blob))
        blob += p
    max_quads = 600
    ld = 2 * max_quads + 64  # spare room past the stated maximum in every row
    lens = np.array([len(p) for p in packets], np.uint32)
    d = lambda a: torch.from_numpy(np.array(a)).cuda()
    for prec, dt in ((F64, torch.complex128), (F32, torch.complex64)):
        out = torch.full((3, ld), 7 + 7j, dtype=dt, device="cuda")
        st = torch.full((3,), -9, dtype=torch.int32, device="cuda")
        ctx.decode(d(np.frombuffer(bytes(blob), np.uint8)), d(np.array(offs, np.int64)), d(lens.view(np.int32)),
                   d(np.array(nqs, np.int32)), int(lens.max()), max_quads, prec, out=out, status=st)
        torch.cuda.synchronize()
        s = st.cpu().numpy()
        got = out.cpu().numpy()
        assert s.tolist() == [0, 5, 0]
        for i in (0, 2):
            _, ref, _, _ = oracle.decode_packet(packets[i], nqs[i])
            want = ref.astype(got.dtype)
            assert np.array_equal(got[i, :1200], want)
            assert (got[i, 1200:] == 7 + 7j).all()
        # the rejected row: zeros where a row of 2 * num_quads would lie,
        # bounded by the row's leading dimension
        assert not got[1].any()
        with pytest.raises(capi.S1GError, match="packet 1"):
            ctx.check_status(st)


def test_check_status_clean_and_failed(ctx):
    st = torch.zeros(1000, dtype=torch.int32, device="cuda")
    ctx.check_status(st)
    st[617] = 2
    st[900] = 1
    with pytest.raises(capi.S1GError, match="packet 617: read past the end"):
        ctx.check_status(st)


@pytest.mark.parametrize("k,pixels", [(0, [(132, 5776), (1287, 5776)]), (2, [(83, 14526), (1436, 14526)])])
def test_rcmc_fp32_offset_rounding_to_one(ctx, k, pixels):
    spec = synth.config3_specs()[k]
    db = pipeline.from_synth(synth.make_burst(spec))
    L, N = db.n_lines, db.n_samples
    # the pixels' migration (FP64) sits just below an integer: it rounds to 1.0f
    R = db.geo.slant_ranges
    rdr = R / db.geo.range_spacing
    lam = pipeline.WAVELENGTH
    f = np.linspace(-db.geo.prf / 2, db.geo.prf / 2, L)
    for (i, j) in pixels:
        u = lam * lam * f[i] * f[i] / (4 * spec.v_ms ** 2)
        sh = rdr[j] * (1 / np.sqrt(1 - u) - 1)
        assert np.float32(sh - np.floor(sh)) == np.float32(1.0), (i, j, sh)
    rng = np.random.default_rng(k)
    x = rng.standard_normal((L, N)) + 1j * rng.standard_normal((L, N))
    x64 = torch.from_numpy(x).cuda()
    x32 = x64.to(torch.complex64)
    ref = ctx.rcmc_matched_filter(x64, db.azimuth_desc())
    for v in range(7):
        ctx.force_variant("rcmc_mf", v)
        try:
            y = ctx.rcmc_matched_filter(x32, db.azimuth_desc())
            torch.cuda.synchronize()
        finally:
            ctx.force_variant("rcmc_mf", -1)
        assert bool(torch.isfinite(torch.view_as_real(y)).all()), v
        for (i, j) in pixels:
            assert abs(complex(y[i, j]) - complex(ref[i, j])) < 1e-4 * abs(complex(ref[i, j])) + 1e-3, (v, i, j)
        err = float((y.to(torch.complex128) - ref).norm() / ref.norm())
        assert err < 2e-6, (v, err)


@pytest.fixture(scope="module")
def rfi_burst():
    """64 lines of 24,000 samples with strong narrowband RFI (30-40 dB INR)
    on about a fifth of them: the RFI lines' packets (36-48 KB) are longer
    than the decode's short-packet class (20 KB), the others (~19 KB)
    shorter."""
    spec = synth.config1(n_lines=64, n_targets=20, seed=2011)
    raw = spec.raw()
    synth.add_rfi(raw, spec, seed=2011, frac=0.2, run=(4, 8), inr_db=(30.0, 40.0))
    return synth.encode_burst(spec, raw)


def test_decode_size_classes_bit_exact(ctx, rfi_burst):
    """Packets of both length classes (decoded by two launches on two
    streams) decode bit-exactly, FP64 and FP32."""
    lens = np.asarray(rfi_burst.ud_lengths)
    assert lens.max() > 24 * 1024 and lens.min() <= 20 * 1024, (lens.min(), lens.max())
    ref, rst = oracle.decode_packets(rfi_burst.stream, rfi_burst.ud_offsets, rfi_burst.ud_lengths,
                                     rfi_burst.num_quads)
    assert (np.asarray(rst) == 0).all()
    db = pipeline.from_synth(rfi_burst)
    for prec in (F64, F32):
        dec, st = db.decode(ctx, prec)
        torch.cuda.synchronize()
        assert (st.cpu().numpy() == 0).all(), prec
        got = dec.cpu().numpy()
        want = ref if prec == F64 else ref.astype(np.complex64)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), prec


def test_decode_size_classes_in_batch_graph(ctx, rfi_burst):
    """The two-stream decode inside a captured batch of three copies of the
    burst, the first and last with ECCM (so the chains run in order, each
    decode forking its long packets onto the decode stream and joining
    back).  Replayed twice, every image and flag vector equals the burst's
    single eager call."""
    db = pipeline.from_synth(rfi_burst)
    L, N = rfi_burst.spec.n_lines, rfi_burst.spec.n_samples
    ed = capi.EccmDesc(L, N, 128, 64, 1.0e5, 4000.0)
    f_ref = torch.zeros(L, dtype=torch.uint8, device="cuda")
    e_eccm = db.focus(ctx, F32, eccm=ed, eccm_flags=f_ref)
    e_plain = db.focus(ctx, F32)
    torch.cuda.synchronize()
    assert 0 < int(f_ref.sum()) < L
    assert not torch.equal(e_eccm, e_plain)
    outs = [torch.zeros_like(e_plain) for _ in range(3)]
    flags = [torch.full_like(f_ref, 7), torch.full_like(f_ref, 7)]
    descs = [db.burst_desc(outs[0], eccm=ed, eccm_flags=flags[0]), db.burst_desc(outs[1]),
             db.burst_desc(outs[2], eccm=ed, eccm_flags=flags[1])]
    for rep in range(2):
        for o in outs:
            o.zero_()
        ctx.focus_batch(descs, F32, graph=True)
        torch.cuda.synchronize()
        assert torch.equal(outs[0], e_eccm), rep
        assert torch.equal(outs[1], e_plain), rep
        assert torch.equal(outs[2], e_eccm), rep
        for f in flags:
            assert torch.equal(f, f_ref), rep
