"""The oracle's TOPS (IW) chain against closed forms: Doppler-centroid
estimation (a deramped tone, a linear centroid profile, the unwrapping of
in-range estimates), the azimuth-frequency unfolding (no-ramp periodicity
and a chirp that the ramp cancels), the resamplers on polynomials, the
per-line filter on cases where it reduces to a phase, and the
azimuth-time unfolding on a tone built from the expanded deramp phase."""
import math

import numpy as np
import pytest

import oracle

PRF = 1717.128973878037


def test_azimuth_beam_angle_and_rate():
    a = np.linspace(-0.018, 0.018, 1024)
    for addr in (0, 1, 511, 512, 1023):
        assert oracle.azimuth_beam_angle(addr) == pytest.approx(a[addr], abs=1e-18)
    # the steering rate over a burst and the rate -2 v k / lambda
    pri_us, n, v = 582.37, 1500, 7500.0
    k = (a[824] - a[200]) / (pri_us * 1e-6 * n)
    assert oracle.doppler_centroid_rate(pri_us, 200, 824, n, v) == pytest.approx(-2 * v * k / oracle.WAVELENGTH,
                                                                                   rel=1e-13)


def _tone_columns(L, f, rate):
    """Columns whose deramped samples are tones exp(2 pi i f i / prf) under
    an amplitude that grows along the burst."""
    t = np.linspace(0.0, 1.06, L)
    ph = math.pi * rate * (t - 0.53) ** 2
    i = np.arange(L)[:, None]
    amp = 1.0 + 3.0 * i / L
    return amp * np.exp(2j * np.pi * f[None, :] * i / PRF) * np.exp(-1j * ph)[:, None]


def test_dc_wrapped_recovers_a_deramped_tone():
    f = np.linspace(-800.0, 800.0, 37)
    for rate in (0.0, -7000.0, 3000.0):
        w = oracle.dc_wrapped(_tone_columns(300, f, rate), rate, PRF)
        assert np.allclose(w, f, atol=1e-9), rate
    # with the opposite ramp the residual chirp, weighted towards the end of
    # the burst, moves the estimate
    w = oracle.dc_wrapped(_tone_columns(300, f, -7000.0), 7000.0, PRF)
    assert not np.allclose(w, f, atol=1.0)


def test_dc_unwrap_linear_profile():
    """Full blocks (n = 30 bs + 1): the block means are linear in the block
    times, the quadratic fit is exact and its value at sample j is
    alpha + beta ((bs - 1) / 2 + j 29 bs / (n - 1))."""
    bs = 11
    n = 30 * bs + 1
    alpha, beta = -120.0, 0.9
    w = alpha + beta * np.arange(n)
    dc = oracle.dc_unwrap(w, 5.3e-3, 5.5e-3, PRF)
    j = np.arange(n)
    want = alpha + beta * ((bs - 1) / 2 + j * 29 * bs / (n - 1))
    assert np.allclose(dc, want, rtol=0, atol=1e-7 * abs(want).max())


def test_dc_unwrap_is_the_fit_of_the_block_means():
    """Estimates inside (-prf/2, prf/2) come back from the unwrapping as they
    went in: the 0/1 sequence floor(|exp(i theta)|) peaks at bin 0 (a = 0,
    b = 0).  The output is then the least-squares quadratic through the block
    means, here checked against numpy's polyfit."""
    rng = np.random.default_rng(11)
    n = 1000
    w = rng.uniform(-PRF / 2 * 0.99, PRF / 2 * 0.99, n)
    tn, tf = 5.3e-3, 5.5e-3
    dc = oracle.dc_unwrap(w, tn, tf, PRF)
    bs = n // 30
    means = []
    for b in range(30):
        s, e = b * bs, (b * bs + bs if b * bs + bs < n else n - 1)
        means.append(w[s:e].mean())
    bt = np.linspace(tn, tf, 30) - tn
    c = np.polyfit(bt, means, 2)
    t = np.linspace(tn, tf, n) - tn
    assert np.allclose(dc, np.polyval(c, t), rtol=0, atol=1e-6 * PRF)


def _rot(x, L):
    """std::rotate(begin, end - L / 2, end) along axis 0."""
    return np.roll(x, L // 2, axis=0)


def test_freq_ufr_without_ramp_is_a_gated_reversal():
    """1 / rate = 0, integer nr: the tiled spectrum is periodic, so its FFT
    is nr L (-1)^m x[-m] on bins k = nr m (even L) and 0 elsewhere; the
    output is the IFFT of the window [start, end) of that."""
    rng = np.random.default_rng(2)
    L, n, nr, bw = 64, 5, 3, 100.0
    x = rng.standard_normal((L, n)) + 1j * rng.standard_normal((L, n))
    out = oracle.az_freq_ufr(x, 1.0, 1.0, PRF, bw, nr=nr, inv_rate=0.0)
    shape = nr * L
    assert out.shape == (shape, n)
    Y = np.zeros((shape, n), complex)
    m = np.arange(L)
    Y[nr * m] = nr * L * ((-1.0) ** m)[:, None] * x[(-m) % L]
    mid = shape // 2
    start, end = int(mid - math.floor(bw) / 2), int(mid + math.floor(bw) / 2)
    Z = np.zeros_like(Y)
    Z[: end - start] = Y[start:end]
    assert np.allclose(out, np.fft.ifft(Z, axis=0), atol=1e-12 * abs(x).max() * nr)


def test_freq_ufr_partial_tile():
    """Non-integer nr: floor(nr) whole copies and floor(frac L) samples of the
    shifted spectrum, then the gate (numpy FFTs)."""
    rng = np.random.default_rng(4)
    L, n, nr, bw = 40, 3, 2.5, 61.7
    x = rng.standard_normal((L, n)) + 1j * rng.standard_normal((L, n))
    out = oracle.az_freq_ufr(x, 1.0, 1.0, PRF, bw, nr=nr, inv_rate=0.0)
    S = _rot(np.fft.fft(x, axis=0), L)
    T = np.concatenate([S, S, S[: L // 2]])
    Y = np.fft.fft(T, axis=0)
    shape = 100
    mid = shape // 2
    start, end = int(mid - 61 / 2), int(mid + 61 / 2)
    assert (start, end) == (19, 80)
    Z = np.zeros_like(Y)
    Z[: end - start] = Y[start:end]
    assert np.allclose(out, np.fft.ifft(Z, axis=0), atol=1e-11 * abs(x).max())


def test_freq_ufr_ramp_cancels_a_chirp():
    """nr = 1: a spectrum exp(i pi f^2 / rate) exp(2 pi i a k0 / L) is a tone
    after the deramp, a single bin k0 after the FFT, exp(2 pi i a (k0 -
    start) / L) after the gate and IFFT, and that times the chirp again."""
    L, n, rate, bw = 256, 3, -3000.0, 150.0
    f = np.linspace(-PRF / 2, PRF / 2, L)
    a = np.arange(L)
    mid = L // 2
    start = int(mid - 150 / 2)
    for k0 in (start, start + 7, start + 149):
        S = np.exp(1j * np.pi * f ** 2 / rate) * np.exp(2j * np.pi * a * k0 / L)
        X = np.roll(S, -(L // 2))  # undo the rotation
        x = np.fft.ifft(X)[:, None] * np.ones(n)
        out = oracle.az_freq_ufr(x, rate, 1.0, PRF, bw, nr=1.0)
        want = np.exp(1j * np.pi * f ** 2 / rate) * np.exp(2j * np.pi * a * (k0 - start) / L)
        assert np.abs(out - want[:, None]).max() < 1e-10, k0
    # a bin outside the window is removed
    S = np.exp(1j * np.pi * f ** 2 / rate) * np.exp(2j * np.pi * a * (start - 3) / L)
    x = np.fft.ifft(np.roll(S, -(L // 2)))[:, None]
    assert np.abs(oracle.az_freq_ufr(x, rate, 1.0, PRF, bw, nr=1.0)).max() < 1e-10


def test_linear_resample_on_lines():
    for n_in, n_out in ((7, 23), (50, 13), (100, 1000), (1000, 1234)):
        y = 3.0 - 0.25 * np.arange(n_in) / (n_in - 1)
        out = oracle.linear_resample(y, n_out)
        want = 3.0 - 0.25 * np.arange(n_out) / (n_out - 1)
        assert np.allclose(out, want, rtol=0, atol=1e-12), (n_in, n_out)
        assert out[-1] == y[-1]
    y = np.random.default_rng(0).standard_normal(9)
    assert np.array_equal(oracle.linear_resample(y, 9), y)


def test_quadratic_resample_on_parabolas():
    """Three-point Lagrange is exact on a parabola; output 0 interpolates
    through a repeated point (0 / 0)."""
    for n_in, n_out in ((40, 17), (1000, 333), (12000, 1700)):
        u = np.arange(n_in) / (n_in - 1)
        y = (1 + 2j) - (3 - 1j) * u + 0.5j * u ** 2
        out = oracle.quadratic_resample(y, n_out)
        x = np.arange(n_out) / (n_out - 1)
        want = (1 + 2j) - (3 - 1j) * x + 0.5j * x ** 2
        assert np.isnan(out[0])
        assert np.allclose(out[1:], want[1:], rtol=0, atol=1e-10), (n_in, n_out)
    y = np.arange(5) + 0j
    assert np.array_equal(oracle.quadratic_resample(y, 5), y)
    with pytest.raises(ValueError):
        oracle.quadratic_resample(np.ones(2, complex), 5)


def test_tops_rows_reduce_to_a_phase():
    """f_a + f_dc = 0: D = 1, no range compression or RCMC shift, so a line
    comes out n x / sqrt((R_ref / R)^3) exp(4 pi i R / lambda) and the FM rate
    is -2 v^2 / (lambda R).  With v very large (D -> 1, K -> inf) and f != 0
    the time correction exp(2 pi i f t_c) is left on top."""
    rng = np.random.default_rng(8)
    rows, n = 4, 48
    x = rng.standard_normal((rows, n)) + 1j * rng.standard_normal((rows, n))
    R = np.linspace(800e3, 801e3, n)
    fr = np.linspace(-3e7, 3e7, n)
    tc = np.linspace(1e-7, 3e-7, n)
    lam = oracle.WAVELENGTH
    ref = R[n // 2]
    v = np.full((rows, n), 7100.0)
    faz = np.array([-30.0, -30.0, -30.0, -30.0])
    y, fm = oracle.tops_rows(x, v, R, fr, 30.0, faz, tc, ref)
    # 4 pi R / lambda is ~1.8e8 rad: its rounding alone (ulp 3e-8 rad) allows
    # a few 1e-8 between two evaluation orders
    want = n * x / np.sqrt((ref / R) ** 3) * np.exp(4j * np.pi * R / lam)
    assert np.allclose(y, want, rtol=1e-6, atol=0)
    assert np.allclose(fm, -2 * 7100.0 ** 2 / (lam * R), rtol=1e-14)
    faz = np.array([100.0, 250.0, -400.0, 0.0])
    v = np.full((rows, n), 1e12)
    y, fm = oracle.tops_rows(x, v, R, fr, 30.0, faz, tc, ref)
    want = n * x / np.sqrt((ref / R) ** 3) * np.exp(4j * np.pi * R / lam) * np.exp(
        2j * np.pi * (faz[:, None] + 30.0) * tc)
    assert np.allclose(y, want, rtol=1e-6, atol=0)


def test_time_params_worked_example():
    """ka = 2000 Hz/s on every line, prf = 2000 Hz, nr = 2 (nr prf = 4000),
    B_d = 600, T = 0.5 s: bandwidth 3400, focused time 0.9 (0.5 + 2200 /
    2000) = 1.44 s, n_ref = 100 / -2000 = -0.05 s, tiles ceil(0.67 / 0.5) -
    floor(-0.77 / 0.5) = 4, output 1000 lines, downsampled to 1156."""
    n_az, n, prf = 1200, 64, 2000.0
    fm = np.full((n_az, n), -2000.0)
    dc = np.full(n, 100.0)
    p, _, _ = oracle.tops_time_params(fm, dc, 2.0, prf, 0.5, 600.0, 10)
    assert p["bandwidth"] == 3400.0
    assert p["focused_time"] == pytest.approx(1.44, rel=1e-14)
    assert p["n_ref"] == -0.05
    assert p["num_tiles"] == 4 and p["shape"] == 4800
    assert (p["lpf_start"], p["lpf_end"]) == (2400 - 1700, 2400 + 1700)
    assert p["output_shape"] == 1000 and p["downsample_shape"] == 1156
    assert (p["crop_start"], p["crop_end"]) == (78, 1078)
    # IW2: x 1.05 and the wider extension floor(675 * 1.15) = 776
    p, _, _ = oracle.tops_time_params(fm, dc, 2.0, prf, 0.5, 675.0, 11)
    assert p["focused_time"] == pytest.approx(1.05 * (0.5 + (4000 - 675 - 1350) / 2000), rel=1e-14)
    assert p["downsample_shape"] == 1776
    # a truncation towards zero: bandwidth / 2 = 1662.5 -> start = mid - 1662.5
    assert (p["lpf_start"], p["lpf_end"]) == (int(p["shape"] // 2 - 1662.5), int(p["shape"] // 2 + 1662.5))


def test_fm_reference_index():
    fm = np.arange(12.0 * 20).reshape(12, 20)
    # in range (n_az / 2 < n and n / 2 < n_az): line n / 2, sample n_az / 2
    assert oracle.fm_reference(fm) == fm[10, 6]
    fm = np.arange(5.0 * 40).reshape(5, 40)
    assert oracle.fm_reference(fm) == fm[2, 20]


def test_time_ufr_on_an_expanded_tone():
    """One tile (n_ref = T / 2, focused time < T).  With dct = -f_dc / k_a the
    deramp phase expands to -pi kt tau^2 - 2 pi f_dc tau + pi dct^2 (kt + 2 ka) (tau
    the burst-centred time) and the reramp phase to pi kt tau^2 + 2 pi f_dc
    tau - pi dct^2 (kt + 2 ka).  An input that the deramp turns into the tone
    of bin lpf_start becomes 1 after the gate, stays 1 through the quadratic
    resampling, and leaves the reramp's phase alone."""
    n_az, n, T = 1000, 6, 0.5
    prf = 2400.0
    nr = 1500.0 / prf
    ka = np.linspace(-2100.0, -1900.0, n)
    fm = np.tile(ka, (n_az, 1))
    dc = ka * (T / 2) + np.linspace(-5.0, 5.0, n)
    dc[n // 2] = ka[n // 2] * T / 2
    for rate in (1e300, -8000.0):
        p, d, k = oracle.tops_time_params(fm, dc, nr, prf, T, 600.0, 10)
        assert p["ok"] and p["num_tiles"] == 1, p
        shape, start, ds = p["shape"], p["lpf_start"], p["downsample_shape"]
        ft, n_ref = p["focused_time"], p["n_ref"]
        tau1 = np.linspace(-ft / 2, ft / 2, shape)[:, None]
        tau2 = np.linspace(n_ref - ft / 2, n_ref + ft / 2, ds)[p["crop_start"]: p["crop_end"], None]
        kt = -ka * rate / (rate - ka)
        dct = -dc / ka
        th1 = -np.pi * kt * tau1 ** 2 - 2 * np.pi * dc * tau1 + np.pi * dct ** 2 * (kt + 2 * ka)
        th2 = np.pi * kt * tau2 ** 2 + 2 * np.pi * dc * tau2 - np.pi * dct ** 2 * (kt + 2 * ka)
        a = np.arange(shape)[:, None]
        x = np.exp(2j * np.pi * a * start / shape) * np.exp(-1j * th1)
        out, p2 = oracle.az_time_ufr(x, dc, fm, rate, T, prf, 600.0, 10, nr=nr)
        assert out.shape == (p["crop_end"] - p["crop_start"], n)
        assert np.abs(out - np.exp(1j * th2)).max() < 1e-8, rate
