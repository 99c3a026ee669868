"""GPU parity on the IW2 burst geometry (24,000 samples per line, 30 Hz
Doppler centroid) against the CPU FP64 oracle: the burst chain in FP64
(within 1e-10 relative L2) and FP32 (within 1e-5), decode bit-exact, and the
FP32 ECCM chain on a burst with narrowband RFI (same flags as the oracle,
SLC within 1e-5 of the oracle's FP64 chain)."""
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
def iw2():
    # the IW2 burst's geometry and scene statistics on 512 of its 1,480 lines
    spec = synth.config1(n_lines=512, n_targets=140, seed=101)
    burst = synth.make_burst(spec)
    print(f"iw2 burst {spec.n_lines} x {spec.n_samples}, f_dc {spec.f_dc}", flush=True)
    return burst, oracle_chain(burst)


@pytest.mark.timeout(600)
def test_iw2_fdc30_chain_vs_oracle(ctx, iw2):
    burst, ref = iw2
    assert burst.spec.f_dc == 30.0
    db = pipeline.from_synth(burst)
    dec, st = db.decode(ctx, F64)
    torch.cuda.synchronize()
    assert (st.cpu().numpy() == 0).all()
    assert np.array_equal(dec.cpu().numpy().view(np.uint64), ref["decoded"].view(np.uint64))
    s64 = db.focus(ctx, F64).cpu().numpy()
    s32 = db.focus(ctx, F32).cpu().numpy()
    e64 = rel_l2(s64, ref["slc"])
    e32 = rel_l2(s32, ref["slc"])
    print(f"rel-L2 vs oracle: FP64 {e64:.2e}, FP32 {e32:.2e}", flush=True)
    assert e64 < 1e-10
    assert e32 < 1e-5


def _window_gradients(x, fsz=128, stride=64):
    """|central difference| along frequency of every ECCM window spectrum of
    each line (windows as signal_processing.cpp:929-1002 cut them: start
    k * stride, an overrunning window ending at N - 1, zero-padded)."""
    L, N = x.shape
    nw = int(np.floor((N / fsz) * (fsz / stride)))
    w = np.zeros((L, nw, fsz), np.complex128)
    for k in range(nw):
        s, e = k * stride, k * stride + fsz
        if e > N:
            e = N - 1
        w[:, k, : e - s] = x[:, s:e]
    W = np.fft.fft(w, axis=-1)
    g = np.empty(W.shape)
    g[..., 1:-1] = np.abs(W[..., 2:] - W[..., :-2]) / 2
    g[..., 0] = np.abs(W[..., 1] - W[..., 0])
    g[..., -1] = np.abs(W[..., -1] - W[..., -2])
    return g


@pytest.mark.timeout(600)
def test_iw2_rfi_eccm_chain_fp32_vs_oracle(ctx):
    spec = synth.config1(n_lines=192, n_targets=60, seed=2003)
    raw = spec.raw()
    mask = synth.add_rfi(raw, spec, seed=2003, frac=0.2, run=(10, 30))
    burst = synth.encode_burst(spec, raw)
    del raw
    dec, _ = oracle.decode_packets(burst.stream, burst.ud_offsets, burst.ud_lengths, burst.num_quads)
    smax = np.abs(np.fft.fft(dec, axis=1)).max(1)
    det = float(np.sqrt(smax[~mask].max() * smax[mask].min()))
    assert smax[~mask].max() < det < smax[mask].min()
    # mitigation threshold: in the widest gap (relative) between the RFI
    # lines' window gradients above twice the largest clean-line gradient,
    # at its geometric midpoint, so FP32 and FP64 zero the same bins (one bin
    # decided differently moves the SLC by about 5e-4 relative L2)
    g_rfi = _window_gradients(dec[mask]).ravel()
    base = 2.0 * _window_gradients(dec[~mask][:32]).max()
    sel = np.sort(g_rfi[(g_rfi > base) & (g_rfi < 2 * base)])
    gaps = sel[1:] / sel[:-1] - 1
    o = int(np.argmax(gaps))
    assert gaps[o] > 5e-5, gaps[o]
    mit = float(np.sqrt(sel[o] * sel[o + 1]))
    ref = oracle_chain(burst, eccm=(det, mit))
    assert np.array_equal(ref["eccm_flags"].astype(bool), mask)
    db = pipeline.from_synth(burst)
    ed = capi.EccmDesc(spec.n_lines, spec.n_samples, 128, 64, det, mit)
    for prec, tol in ((F32, 1e-5), (F64, 1e-10)):
        flags = torch.zeros(spec.n_lines, dtype=torch.uint8, device="cuda")
        slc = db.focus(ctx, prec, eccm=ed, eccm_flags=flags)
        torch.cuda.synchronize()
        assert np.array_equal(flags.cpu().numpy().astype(bool), mask), prec
        err = rel_l2(slc.cpu().numpy(), ref["slc"])
        print(f"ECCM chain prec {prec}: rel-L2 vs oracle {err:.2e}, {int(mask.sum())} RFI lines", flush=True)
        assert err < tol, prec
