"""The general signal-processing entry points against numpy and the oracle.

DFTs of every length kind (power of two, mixed radix, lengths with a prime
factor above the pass radices that go through Bluestein, including the
reference's padded range length 14936 = 2^3 * 1867), along rows and
columns; transpose; the pulse-compression product; the reference function;
the replica chirp; windowed-sinc interpolation.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from sentinel1_decode_amd import capi  # noqa: E402

F32, F64 = capi.S1G_F32, capi.S1G_F64
DT = {F32: torch.complex64, F64: torch.complex128}
TOL = {F32: 3e-6, F64: 1e-12}


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# 1..19608: whole rows in LDS (FP64 up to 10,240 points); 24000 and 30744
# (and FP64 19608): global passes; 1867 and 14936: Bluestein.
@pytest.mark.parametrize("n", [1, 7, 256, 1480, 1867, 4902, 14936, 19608, 24000, 30744])
@pytest.mark.parametrize("prec", [F64, F32])
def test_dft_rows(ctx, n, prec):
    x = _rand((3, n), n)
    for inverse in (False, True):
        ref = np.fft.ifft(x, axis=1) if inverse else np.fft.fft(x, axis=1)
        d = torch.from_numpy(x).to(DT[prec]).cuda()
        ctx.dft(d, axis=1, inverse=inverse)
        torch.cuda.synchronize()
        tol = TOL[prec] * (10 if n > 1000 else 1)
        assert rel_l2(d.cpu().numpy(), ref) < tol, (n, inverse)


@pytest.mark.parametrize("rows", [1480, 1024, 1867])
@pytest.mark.parametrize("prec", [F64, F32])
def test_dft_columns(ctx, rows, prec):
    x = _rand((rows, 37), rows)
    for inverse in (False, True):
        ref = np.fft.ifft(x, axis=0) if inverse else np.fft.fft(x, axis=0)
        d = torch.from_numpy(x).to(DT[prec]).cuda()
        ctx.dft(d, axis=0, inverse=inverse)
        torch.cuda.synchronize()
        assert rel_l2(d.cpu().numpy(), ref) < TOL[prec] * 10, (rows, inverse)


def test_dft_matches_oracle_fft(ctx):
    """Same convention as the oracle's FFTW restatement (forward)."""
    x = _rand(14936, 3)
    d = torch.from_numpy(np.ascontiguousarray(x.reshape(1, -1))).cuda()
    ctx.dft(d, axis=1)
    torch.cuda.synchronize()
    assert rel_l2(d.cpu().numpy()[0], oracle.fft(x)) < 1e-11


@pytest.mark.parametrize("prec", [F64, F32])
def test_transpose_and_multiply(ctx, prec):
    x = _rand((67, 301), 9)
    d = torch.from_numpy(x).to(DT[prec]).cuda()
    t = ctx.transpose(d)
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy(), x.T.astype(t.cpu().numpy().dtype))
    v = _rand(301, 10)
    dv = torch.from_numpy(v).to(DT[prec]).cuda()
    ctx.multiply(d, dv)
    ref = x * v[None, :]
    torch.cuda.synchronize()
    assert rel_l2(d.cpu().numpy(), ref) < TOL[prec]
    d2 = torch.from_numpy(x).to(DT[prec]).cuda()
    ctx.multiply(d2, dv, conj=True)
    torch.cuda.synchronize()
    assert rel_l2(d2.cpu().numpy(), x * np.conj(v)[None, :]) < TOL[prec]


@pytest.mark.parametrize("n_samples,fs", [(8192, 64.34523813), (512, 25.02314816)])
def test_replica_chirp_and_reference_function(ctx, n_samples, fs):
    txpsf, txprr, txpl = -8.7e-1, 1.0789, 52.40481
    ref_chirp = oracle.replica_chirp(txpsf, txprr, txpl, fs, n_samples)
    chirp = ctx.replica_chirp(n_samples, txpsf, txprr, txpl, fs)
    torch.cuda.synchronize()
    got = chirp.cpu().numpy()
    assert got.shape == ref_chirp.shape
    assert rel_l2(got, ref_chirp) < 1e-13
    rf = ctx.reference_function(chirp)
    torch.cuda.synchronize()
    assert rel_l2(rf.cpu().numpy(), oracle.reference_function(ref_chirp)) < 1e-11
    c32 = ctx.replica_chirp(n_samples, txpsf, txprr, txpl, fs, dtype=torch.complex64)
    rf32 = ctx.reference_function(c32)
    torch.cuda.synchronize()
    assert rel_l2(rf32.cpu().numpy(), oracle.reference_function(ref_chirp)) < 1e-5


@pytest.mark.parametrize("half_width", [8, 4])
def test_sinc_interpolate(ctx, half_width):
    x = _rand(200, 4)
    rng = np.random.default_rng(5)
    t = np.concatenate([rng.uniform(-5, 205, 300), np.arange(0, 200, 17.0), [0.0, 199.0, 3.5]])
    ref = np.array([oracle.sinc_interpolate(x, float(tq), half_width) for tq in t])
    d = torch.from_numpy(x).cuda()
    dt = torch.from_numpy(t).cuda()
    got = ctx.sinc_interpolate(d, dt, half_width).cpu().numpy()
    assert rel_l2(got, ref) < 1e-13
    # integer positions reproduce the samples (interior)
    k = np.arange(0, 200, 17)
    assert np.allclose(got[300:300 + k.size], x[k], rtol=0, atol=1e-12)
