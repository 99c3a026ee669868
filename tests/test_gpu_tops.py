"""GPU: the TOPS (IW) stages -- Doppler centroid, azimuth-frequency UFR, the
per-line filter, azimuth-time UFR -- and their chain, against the CPU FP64
oracle on the same inputs.

Tolerances: the Doppler centroid is the reference's arithmetic on the same
samples (serial lag sums without contraction, host fit), so it must agree
to 1e-12 relative; the FFT stages to rel-L2 1e-10 in FP64 and 1e-5 in FP32
(FP32 data, FP64 phases)."""
import math

import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from sentinel1_decode_amd import capi  # noqa: E402

PRF = 1717.128973878037
LAM = oracle.WAVELENGTH
C = 299792458.0


def _dt(prec):
    return torch.complex64 if prec == "f32" else torch.complex128


def _tone_burst(L, n, rate, seed, f_lo=-500.0, f_hi=400.0):
    """Range-compressed-like columns: a Doppler tone per sample under the
    TOPS ramp, plus noise."""
    rng = np.random.default_rng(seed)
    t = np.linspace(0.0, 1.06, L)
    f = np.linspace(f_lo, f_hi, n) + 30 * np.sin(np.arange(n) / 50.0)
    i = np.arange(L)[:, None]
    amp = 1.0 + 0.5 * np.cos(2 * np.pi * i / L)
    x = amp * np.exp(2j * np.pi * f[None, :] * i / PRF) * np.exp(-1j * math.pi * rate * (t - 0.53) ** 2)[:, None]
    x = x + 0.3 * (rng.standard_normal((L, n)) + 1j * rng.standard_normal((L, n)))
    return x


def _geometry(n, rows, seed=1):
    rng = np.random.default_rng(seed)
    R = np.linspace(800e3, 800e3 + n * 2.34, n)
    fs = 64.34523813e6
    fr = np.linspace(-fs / 2, fs / 2, n)
    tc = np.linspace(1.1e-7, 1.4e-7, n)
    # a platform near 700 km altitude, moving along the line
    t = np.arange(rows) * 1e-3
    pos = np.stack([7.07e6 + 10 * t, 1.2e5 + 7100 * t, 2.0e5 + 30 * t], axis=1)
    vn = 7100.0 + 0.01 * rng.standard_normal(rows)
    return R, fr, tc, pos, vn


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_doppler_centroid(ctx, prec):
    L, n, rate = 400, 1500, -6000.0
    x = _tone_burst(L, n, rate, 3).astype(np.complex64 if prec == "f32" else np.complex128)
    tn, tf = 5.32e-3, 5.32e-3 + n / 64.34523813e6
    want = oracle.doppler_centroid(x.astype(np.complex128), rate, PRF, tn, tf)
    got = ctx.doppler_centroid(torch.from_numpy(x).cuda(), rate, PRF, tn, tf).cpu().numpy()
    assert np.allclose(got, want, rtol=1e-12, atol=0), np.abs(got - want).max()
    # the estimate follows the tones: a quadratic through their block means
    assert np.abs(want - np.linspace(-500, 400, n)).max() < 80


def test_dc_rate_host():
    for a, b, n, v in ((200, 824, 1500, 7500.0), (900, 100, 1200, 7600.0)):
        assert capi.doppler_centroid_rate(582.37, a, b, n, v) == oracle.doppler_centroid_rate(582.37, a, b, n, v)
    assert capi.azimuth_beam_angle(0) == -0.018
    assert math.isnan(capi.azimuth_beam_angle(1024))


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("L,n,nr,bw", [(400, 512, 2.3, 600.0), (257, 300, 3.0, 675.0), (300, 64, 2.71, 600.0)])
def test_azimuth_frequency_ufr(ctx, prec, L, n, nr, bw):
    rng = np.random.default_rng(L + n)
    x = rng.standard_normal((L, n)) + 1j * rng.standard_normal((L, n))
    T = L / PRF
    rate = -nr * PRF / T
    want = oracle.az_freq_ufr(x, rate, T, PRF, bw)
    d = torch.from_numpy(x).to(_dt(prec)).cuda()
    got = ctx.azimuth_frequency_ufr(d, rate, T, PRF, bw).cpu().numpy()
    assert got.shape == want.shape
    tol = 1e-5 if prec == "f32" else 1e-10
    assert rel_l2(got, want) < tol


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_tops_filter_lines(ctx, prec):
    rows, n = 300, 480
    rng = np.random.default_rng(5)
    x = rng.standard_normal((rows, n)) + 1j * rng.standard_normal((rows, n))
    R, fr, tc, pos, vn = _geometry(n, rows)
    fdc = np.linspace(-40.0, 60.0, n)
    faz = np.linspace(-2000.0, 2000.0, rows)
    ref = float(R[n // 2])
    vrel = np.stack([oracle.effective_velocities(pos[i], vn[i], R) for i in range(rows)])
    want, fm_want = oracle.tops_rows(x, vrel, R, fr, fdc, faz, tc, ref)
    d = torch.from_numpy(x).to(_dt(prec)).cuda()
    g = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
    fm = ctx.tops_filter_lines(d, g(pos), g(vn), g(faz), g(R), g(fr), g(fdc), g(tc), ref)
    got = d.cpu().numpy()
    tol = 1e-5 if prec == "f32" else 1e-10
    assert rel_l2(got, want) < tol
    assert np.allclose(fm.cpu().numpy(), fm_want, rtol=1e-13, atol=0)


def _fm_plane(n_az, n, R, seed):
    rng = np.random.default_rng(seed)
    v = 7100.0 + 2 * rng.standard_normal((n_az, 1))
    D = np.sqrt(1 - (LAM * np.linspace(-2500, 2500, n_az)[:, None] / (2 * v)) ** 2)
    return -2 * v ** 2 * D ** 3 / (LAM * R[None, :])


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("n_az,n,nr,swath", [(920, 512, 2.3, 10), (700, 96, 2.9, 11), (1000, 300, 2.0, 12)])
def test_azimuth_time_ufr(ctx, prec, n_az, n, nr, swath):
    rng = np.random.default_rng(n_az)
    L = int(n_az / nr) + 1
    T = L / PRF
    rate = -nr * PRF / T
    bw = 675.0 if swath == 11 else 600.0
    R = np.linspace(800e3, 800e3 + n * 2.34, n)
    fm = _fm_plane(n_az, n, R, n)
    dc = np.linspace(-30.0, 50.0, n)
    x = rng.standard_normal((n_az, n)) + 1j * rng.standard_normal((n_az, n))
    want, p = oracle.az_time_ufr(x, dc, fm, rate, T, PRF, bw, swath)
    g = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
    d = torch.from_numpy(x).to(_dt(prec)).cuda()
    got, gp = ctx.azimuth_time_ufr(d, g(dc), g(fm), rate, T, PRF, bw, swath)
    for k in ("num_tiles", "shape", "lpf_start", "lpf_end", "downsample_shape", "output_shape", "crop_start",
              "crop_end"):
        assert gp[k] == p[k], k
    for k in ("bandwidth", "focused_time", "n_ref", "ka_mean"):
        assert gp[k] == p[k], k
    got = got.cpu().numpy()
    assert got.shape == want.shape
    tol = 1e-5 if prec == "f32" else 1e-10
    assert rel_l2(got, want) < tol


def test_time_ufr_rejects_a_window_outside(ctx):
    """A burst so short that the gate [lpf_start, lpf_end) leaves the tiled
    row is refused (the reference would read outside its arrays)."""
    n_az, n = 64, 32
    R = np.linspace(800e3, 800e3 + n * 2.34, n)
    fm = _fm_plane(n_az, n, R, 0)
    g = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
    x = torch.zeros((n_az, n), dtype=torch.complex128, device="cuda")
    # nr = 1: bandwidth nr prf - 600 = 1117 Hz of bins, but 2 tiles of 64
    T = 0.2
    with pytest.raises(capi.S1GError):
        ctx.azimuth_time_ufr(x, g(np.zeros(n)), g(fm), -PRF / T, T, PRF, 600.0, 10)
    p, _, _ = oracle.tops_time_params(fm, np.zeros(n), 1.0, PRF, T, 600.0, 10)
    assert not p["ok"] and p["lpf_start"] < 0


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_tops_chain(ctx, prec):
    """Doppler centroid -> frequency UFR -> line filter -> azimuth IFFT ->
    time UFR, each GPU stage fed by the previous GPU stage, against the oracle
    chain fed the same range-compressed burst."""
    L, n, nr, swath = 420, 384, 2.4, 10
    T = L / PRF
    rate = -nr * PRF / T
    bw = 600.0
    x = _tone_burst(L, n, rate, 9, -200.0, 250.0)
    if prec == "f32":
        x = x.astype(np.complex64).astype(np.complex128)
    tn, tf = 5.4e-3, 5.4e-3 + n / 64.34523813e6
    R, fr, tc, pos0, vn0 = _geometry(n, 1)
    # oracle chain
    dc = oracle.doppler_centroid(x, rate, PRF, tn, tf)
    u = oracle.az_freq_ufr(x, rate, T, PRF, bw)
    rows = u.shape[0]
    _, _, _, pos, vn = _geometry(n, rows, seed=4)
    faz = np.linspace(-nr * PRF / 2, nr * PRF / 2, rows)
    vrel = np.stack([oracle.effective_velocities(pos[i], vn[i], R) for i in range(rows)])
    ref = float(R[n // 2])
    u, fm = oracle.tops_rows(u, vrel, R, fr, dc, faz, tc, ref)
    u = np.fft.ifft(u, axis=0)
    want, p = oracle.az_time_ufr(u, dc, fm, rate, T, PRF, bw, swath)
    # GPU chain
    g = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
    d = torch.from_numpy(x).to(_dt(prec)).cuda()
    d_dc = ctx.doppler_centroid(d, rate, PRF, tn, tf)
    assert np.allclose(d_dc.cpu().numpy(), dc, rtol=1e-12, atol=0)
    du = ctx.azimuth_frequency_ufr(d, rate, T, PRF, bw)
    d_fm = ctx.tops_filter_lines(du, g(pos), g(vn), g(faz), g(R), g(fr), d_dc, g(tc), ref)
    ctx.dft(du, axis=0, inverse=True)
    got, gp = ctx.azimuth_time_ufr(du, d_dc, d_fm, rate, T, PRF, bw, swath)
    assert gp["crop_end"] - gp["crop_start"] == want.shape[0]
    tol = 2e-5 if prec == "f32" else 1e-10
    assert rel_l2(got.cpu().numpy(), want) < tol
