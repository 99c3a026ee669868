"""Pins for the oracle's azimuth chain and burst geometry (runs without a GPU).

The checks here do not restate the oracle; each one compares it with
something fixed independently, and each is shown to fail on a planted
mistake (the oracle built with -DORC_MUTATION=k, see s1_oracle.c):

* The stationary-phase spectrum of the hyperbolic range history
  R(eta) = sqrt(r0^2 + v^2 eta^2) is checked against a numerical FFT of that
  history: Phi(f) = -4 pi r0 D(f) / lambda - 2 pi f eta0 - pi/4,
  D(f) = sqrt(1 - (lambda f / 2 v)^2).
* A range-Doppler point target built from that spectrum (Gaussian range
  pulse migrated to r0 / D - r0, spectral phase -4 pi r0 D / lambda) through
  orc_rcmc_mf and orc_azimuth_ifft must focus to a discrete delta.  With the
  reference's Doppler labels f_i = linspace(-PRF/2, PRF/2, L)
  (s1_decoder.cpp:604) on the rows of the fftshifted spectrum, whose true
  bin frequencies are g_i = (i - L/2) PRF / L, a target placed at
  eta0 = l0 / PRF + t_c L / (L - 1) focuses on line l0 with value
  A (r0 / R_ref)^1.5 exp(i (pi l0 + pi PRF t_c / (L - 1) + 2 pi f_dc t_c)):
  the i-dependent parts of -2 pi g_i eta0 + 2 pi (f_i + f_dc) t_c
  + 2 pi i l0 / L cancel, the constant parts give that phase, and the factor
  (-1)^l0 is the fftshift the inverse transform does not undo
  (s1_decoder.cpp:682).  f_dc = 0, t_c = 0 is the textbook RDA case.
* Band-limited to 0.8 PRF, the focused response is an unweighted sinc:
  PSLR -13.26 dB and 3-dB width 0.886 / 0.8 lines.
* A time-domain echo from the synthetic-scene simulator (its own hyperbolic
  model, chirp and sampling) through the whole oracle chain lands on the
  range cell of the target and on the azimuth line the stationary-phase
  prediction gives.
* Slant ranges and timing corrections (packet.cpp:284-342) from the header
  codes of the README's sample packet agree with the same quantities
  evaluated from the physical values the README prints for that packet.
"""
import math
import os

import numpy as np
import pytest

import oracle
from sentinel1_decode_amd import synth

C_LIGHT = 299792458.0
LAM = C_LIGHT / 5.405000454334350e9
PRF = 1.0 / (21861 / 37.53472224 * 1e-6)
V = 7100.0
FS = 64.34523813e6
DR = C_LIGHT / (2 * FS)


def D_of(f, v=V):
    return np.sqrt(1.0 - (LAM * f / (2.0 * v)) ** 2)


# --------------------------------------------------------- physics of D
def test_stationary_phase_spectrum_of_hyperbolic_history():
    """The azimuth spectrum of exp(-i 4 pi R(eta) / lambda) has phase
    -4 pi r0 D(f) / lambda - 2 pi f eta0 - pi/4 (stationary phase): the D
    that the point-target checks below build their spectra with."""
    r0, n = 850e3, 1 << 13
    eta0 = 0.37 / PRF
    eta = (np.arange(n) - n / 2) / PRF - eta0
    R = np.sqrt(r0 ** 2 + V ** 2 * eta ** 2)
    s = np.exp(-4j * np.pi * R / LAM)
    # the history is sampled at the PRF: keep the instantaneous Doppler
    # -2 v^2 eta / (lambda R) inside +-0.45 PRF with a smooth (raised-cosine)
    # taper from 0.38 PRF, so nothing aliases and the aperture ends stay out
    # of the band checked
    fd = np.abs(2 * V * V * eta / (LAM * R)) / PRF
    w = np.where(fd < 0.38, 1.0, np.where(fd > 0.45, 0.0, 0.5 * (1 + np.cos(np.pi * (fd - 0.38) / 0.07))))
    S = np.fft.fftshift(np.fft.fft(np.fft.ifftshift(s * w)))
    f = (np.arange(n) - n / 2) * PRF / n
    # positions relative to the centre sample: eta_l = (l - n/2) / PRF, so
    # the target's zero-Doppler time is eta0 and the transform's time origin
    # the centre sample
    keep = np.abs(f) < 0.3 * PRF
    ph_model = -4 * np.pi * r0 * D_of(f) / LAM - 2 * np.pi * f * eta0 - np.pi / 4
    d = np.angle(S[keep] * np.exp(-1j * ph_model[keep]))
    d -= np.angle(np.mean(np.exp(1j * d)))  # constant of the taper
    assert np.max(np.abs(d)) < 2e-2
    # a factor 2 inside D (or a dropped factor 4 pi) breaks the agreement
    ph_bad = -4 * np.pi * r0 * D_of(f * math.sqrt(2)) / LAM - 2 * np.pi * f * eta0
    d2 = np.angle(S[keep] * np.exp(-1j * ph_bad[keep]))
    assert np.std(np.unwrap(d2)) > 10


# --------------------------------------------------- RD point target
def rd_point_target(L, N, f_dc, t_c, jt, l0, ref_ratio=0.9, s=2.5, band=None, amp=1.0):
    """Range-Doppler spectrum (rows = fftshifted Doppler bins) of one point
    target at column jt focusing to line l0 (l0 may be fractional), and the
    geometry arrays the oracle takes.  Returns (X, geo, expected_peak)."""
    i = np.arange(L)
    f = np.linspace(-PRF / 2, PRF / 2, L)          # the reference's row labels
    g = (i - L // 2) * PRF / L                      # true frequency of row i
    R = 850e3 + (np.arange(N) - jt) * DR
    r0 = R[jt]
    D = D_of(f + f_dc)
    mig = (r0 / D - r0) / DR
    eta0 = l0 / PRF + t_c * L / (L - 1)
    j = np.arange(N)
    X = amp * np.exp(-((j[None, :] - jt - mig[:, None]) ** 2) / (2 * s * s))
    X = X * np.exp(-4j * np.pi * r0 * D / LAM)[:, None] * np.exp(-2j * np.pi * g * eta0)[:, None]
    if band is not None:
        X = X * (np.abs(g) < band * PRF / 2)[:, None]
    tc = t_c + (np.arange(N) - jt) * (-0.5 / FS)   # the reference's slope -1/(2 fs)
    geo = dict(prf=PRF, range_spacing=DR, ref_range=r0 / ref_ratio, slant_ranges_=R, v_rel=np.full(N, V),
               f_dc=np.full(N, f_dc), t_corr=tc)
    peak = amp * ref_ratio ** 1.5 * np.exp(1j * (np.pi * l0 + np.pi * PRF * t_c / (L - 1)
                                                   + 2 * np.pi * f_dc * t_c))
    return X, geo, peak


def focus(X, geo):
    return oracle.azimuth_ifft(oracle.rcmc_mf(X, **geo))


def closed_form_failures(f_dc, t_c, L=512, N=96, jt=40, l0=301):
    """Failures (text) of the closed-form checks for one configuration."""
    X, geo, peak = rd_point_target(L, N, f_dc, t_c, jt, l0)
    y = focus(X, geo)
    col = y[:, jt]
    out = []
    k = int(np.argmax(np.abs(col)))
    if k != l0:
        out.append(f"peak on line {k}, expected {l0}")
    amp_err = abs(abs(col[l0]) / abs(peak) - 1)
    if amp_err > 2e-4:
        out.append(f"peak amplitude off by {amp_err:.2e}")
    ph_err = abs(np.angle(col[l0] / peak))
    if ph_err > 1e-3:
        out.append(f"peak phase off by {ph_err:.2e} rad")
    side = np.max(np.abs(np.delete(col, l0))) / abs(col[l0])
    if side > 1e-3:
        out.append(f"energy off the focused line {side:.2e}")
    # RCMC: the column holding the most energy on the focused line is jt
    if int(np.argmax(np.abs(y[l0]))) != jt:
        out.append(f"focused column {int(np.argmax(np.abs(y[l0])))}, expected {jt}")
    return out


@pytest.mark.parametrize("f_dc,t_c", [(0.0, 0.0), (30.0, 3.26e-3), (-45.0, 2.9e-3)])
def test_rd_point_target_closed_form(f_dc, t_c):
    """f_dc = 0, t_c = 0 is the textbook range-Doppler case: the matched
    filter exp(i 4 pi R D / lambda) undoes the spectral phase and the IFFT
    returns a delta at the target's line."""
    assert closed_form_failures(f_dc, t_c) == []


def test_rd_point_target_band_limited_sinc():
    """A rectangular 0.8 PRF Doppler band focuses to an unweighted sinc:
    PSLR -13.26 dB and 3-dB width 0.886 / 0.8 lines, at a fractional line."""
    L, N, jt = 512, 64, 30
    l0 = 250.37
    X, geo, peak = rd_point_target(L, N, 30.0, 3.26e-3, jt, l0, band=0.8)
    col = focus(X, geo)[:, jt]
    # band-limited interpolation of the focused column from its spectrum
    Y = np.fft.fft(col)
    t = np.arange(l0 - 8, l0 + 8, 1 / 64)
    e = np.exp(2j * np.pi * np.outer(t, np.arange(L)) / L)
    yt = np.abs(e @ Y) / L
    k = int(np.argmax(yt))
    assert abs(t[k] - l0) < 1 / 32
    half = yt >= yt[k] / math.sqrt(2)
    lo = k
    while lo > 0 and half[lo - 1]:
        lo -= 1
    hi = k
    while hi < len(t) - 1 and half[hi + 1]:
        hi += 1
    width = (hi - lo + 1) / 64
    assert width == pytest.approx(0.886 / 0.8, rel=0.03)
    # first sidelobe: the largest value outside the main lobe (first nulls at
    # +-1.25 lines for a 0.8 band)
    main = np.abs(t - t[k]) < 1.2
    pslr = 20 * np.log10(yt[~main].max() / yt[k])
    assert pslr == pytest.approx(-13.26, abs=0.3)


# Planted mistakes each check set must catch (s1_oracle.c ORC_MUTATION list).
AZIMUTH_MUTATIONS = {
    1: "sign of 4 pi R D / lambda",
    2: "factor 2 inside D",
    3: "amplitude weight dropped",
    4: "sign of the t_corr phase",
    5: "sign of the RCMC shift",
    6: "f_dc left out of D",
    7: "az_freqs spacing PRF / L",
}


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("oracle_mutants"))
    return {k: oracle.build_mutant(k, d) for k in list(AZIMUTH_MUTATIONS) + [8, 9]}


@pytest.mark.parametrize("mutation", sorted(AZIMUTH_MUTATIONS))
def test_closed_form_catches_planted_mistake(mutants, mutation):
    with oracle.using(mutants[mutation]):
        fails = closed_form_failures(30.0, 3.26e-3)
    assert fails, f"planted mistake not caught: {AZIMUTH_MUTATIONS[mutation]}"


# ------------------------------------------------ time-domain full chain
def predicted_azimuth_line(spec, n_t, a_t, geo):
    """Azimuth line (fractional) where the chain focuses a target at range
    cell n_t, zero-Doppler line a_t.  Row i of the fftshifted spectrum holds
    Doppler g_i = (i - L/2) PRF / L; there the target's spectrum is
    exp(-i 4 pi r0 D(g) / lambda - i 2 pi g eta0) (stationary phase) within
    the band the simulator illuminates (|g - f_dc| <= band_frac PRF / 2),
    and the matched filter multiplies it by
    exp(i 4 pi R_j D(f_i + f_dc) / lambda + i 2 pi (f_i + f_dc) t_c) with the
    reference's labels f_i (s1_decoder.cpp:604,658-676).  The prediction is
    the peak of the band-limited inverse transform of that product."""
    L = spec.n_lines
    i = np.arange(L)
    g = (i - L // 2) * PRF / L
    f = np.linspace(-PRF / 2, PRF / 2, L)
    r0 = spec.slant_range_near() + n_t * DR
    Rj = geo["R"][n_t]
    tc = geo["tc"][n_t]
    fd = spec.f_dc
    band = np.abs(g - fd) <= spec.band_frac * PRF / 2
    phi = 4 * np.pi / LAM * (Rj * D_of(f + fd) - r0 * D_of(g)) + 2 * np.pi * (f + fd) * tc \
        - 2 * np.pi * g * (a_t / PRF)
    Y = np.where(band, np.exp(1j * phi), 0.0)
    # y(t) = sum_i Y_i e^{2 pi i i t / L} (the inverse transform of the rows)
    t0 = a_t - tc * PRF  # a first guess; the search spans +-40 lines
    t = t0 + np.arange(-40.0, 40.0, 1 / 64)
    y = np.abs(np.exp(2j * np.pi * np.outer(t, i) / L) @ Y)
    k = int(np.argmax(y))
    t = t[k] + np.arange(-1 / 64, 1 / 64, 1 / 4096)
    y = np.abs(np.exp(2j * np.pi * np.outer(t, i) / L) @ Y)
    return float(t[np.argmax(y)])


def interp_peak(x, near, k):
    """Position of |x|'s peak near sample `near`, from the band-limited
    interpolant sum_k X_k e^{2 pi i k t / n} / n over frequency indices k."""
    n = len(x)
    t = near + np.arange(-2.0, 2.0, 1 / 512)
    v = np.abs(np.exp(2j * np.pi * np.outer(t, k) / n) @ np.fft.fft(x))
    return float(t[np.argmax(v)])


def predicted_range_cell(spec, n_t):
    """Range sample where the reference's range compression puts an echo
    centred on sample n_t.  Its replica samples the pulse at
    t_u = u tau / (nc - 1) (linspace with the endpoint, packet.cpp:345-373)
    while the echo is sampled every 1/fs, a stretch of b = 1/fs - tau/(nc - 1)
    per sample; the correlation is stationary where the instantaneous
    frequencies meet at the pulse centre, which is half a sample late plus a
    term in the chirp's centre frequency f_c:
    j - n_t = 1/2 - f_c b fs / (K tau / (nc - 1))."""
    fs = spec.fs_mhz * 1e6
    tau = spec.txpl_us * 1e-6
    K = spec.txprr * 1e12
    nc = int(math.floor(spec.fs_mhz * spec.txpl_us))
    delta = tau / (nc - 1)
    b = 1 / fs - delta
    f_c = spec.txpsf * 1e6 + K * tau / 2
    return n_t + 0.5 - f_c * b * fs / (K * delta)


@pytest.mark.parametrize("f_dc", [0.0, 30.0])
def test_time_domain_point_target_lands_on_prediction(f_dc):
    from chain_ref import oracle_chain
    spec = synth.config0(n_lines=512, n_samples=4096, grid_spacing=(1, 1), clutter_sigma=0.0,
                         noise_sigma=0.0, f_dc=f_dc, seed=11)
    spec.target_grid = False
    spec.n_targets = 0
    n_t, a_t = 1500.0, 260.0
    spec.targets = lambda: np.array([[n_t, a_t, 1.0, 0.0]])
    b = synth.make_burst(spec)
    out = oracle_chain(b)
    slc = out["slc"]
    i, j = np.unravel_index(np.argmax(np.abs(slc)), slc.shape)
    # sub-sample peaks by exact band-limited interpolation: range lines are
    # baseband, azimuth columns hold the Doppler rows 0..L-1 (the inverse
    # transform leaves the forward fftshift in place)
    N, L = spec.n_samples, spec.n_lines
    rng_peak = interp_peak(slc[i], j, np.fft.fftfreq(N) * N)
    assert abs(rng_peak - predicted_range_cell(spec, n_t)) < 0.1, (rng_peak, predicted_range_cell(spec, n_t))
    az_peak = interp_peak(slc[:, j], i, np.arange(L))
    want = predicted_azimuth_line(spec, int(j), a_t, out["geo"])
    # 0.2 lines: with f_dc != 0 the reference's RCMC shift uses
    # D(f_i + f_dc) where the row holds Doppler g_i, leaving up to 0.12
    # samples of residual migration linear in Doppler; the column, off the
    # range peak by half a sample, then sees the range pulse's phase slope as
    # an azimuth shift (0.1 lines per column here), which the prediction
    # leaves out.  A flipped t_corr sign moves the peak by 11 lines, f_dc
    # left out of D by 24 lines.
    assert abs(az_peak - want) < 0.2, (az_peak, want)


# --------------------------------------------------- burst geometry
# The README's sample packet (README.md:94-165): header codes and the
# physical values the reference prints for them.
README_SEC = dict(swst=3681, pri=21859, rank=9, range_decimation=8, num_quadratures=11938, pulse_length=1967,
                  tx_ramp_rate_sign=1, tx_ramp_rate_mag=1605, pulse_start_frequency_sign=0,
                  pulse_start_frequency_mag=12335, swl=13979, baq_mode=12, swath_number=10)
README_PRINTED = dict(pri_us=582.367, swst_us=98.0692, fs_mhz=64.34523813)


def readme_sec():
    s = np.zeros(46, np.int64)
    for k, v in README_SEC.items():
        s[synth.SECONDARY_FIELDS.index(k)] = v
    return s


def geometry_failures():
    sec = readme_sec()
    n = 2 * README_SEC["num_quadratures"]
    fs = README_PRINTED["fs_mhz"] * 1e6
    pri = README_PRINTED["pri_us"] * 1e-6
    swst = README_PRINTED["swst_us"] * 1e-6
    dts = 320 / (8 * 37.53472224) * 1e-6  # suppressed-data time (structs.h:33)
    out = []
    R = oracle.slant_ranges(sec)
    near = C_LIGHT / 2 * (README_SEC["rank"] * pri + swst + dts)
    if abs(R[0] / near - 1) > 2e-6:
        out.append(f"near slant range {R[0]} vs {near}")
    # the window spans N / fs of two-way time, endpoint included
    span = C_LIGHT / 2 * n / fs
    if abs((R[-1] - R[0]) / span - 1) > 1e-9:
        out.append(f"slant-range span {R[-1] - R[0]} vs {span}")
    if np.max(np.abs(np.diff(R, 2))) > 1e-6:
        out.append("slant ranges not uniformly spaced")
    tc = oracle.timing_corrections(sec)
    j = np.arange(n)
    want = (swst + 1.439e-6) + (swst + README_SEC["rank"] * pri - j / fs) / 2
    if np.max(np.abs(tc / want - 1)) > 2e-6:
        out.append(f"timing corrections off by {np.max(np.abs(tc / want - 1)):.2e}")
    return out


def test_geometry_from_readme_packet():
    assert geometry_failures() == []
    p = oracle.packet_params(readme_sec())
    assert p["pri_us"] == pytest.approx(README_PRINTED["pri_us"], abs=5e-4)
    assert p["swst_us"] == pytest.approx(README_PRINTED["swst_us"], abs=5e-5)


@pytest.mark.parametrize("mutation", [8, 9])
def test_geometry_catches_planted_mistake(mutants, mutation):
    with oracle.using(mutants[mutation]):
        assert geometry_failures(), f"planted mistake {mutation} not caught"
