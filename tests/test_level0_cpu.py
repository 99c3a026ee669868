"""Level-0 files and burst geometry without a GPU: packet framing, burst
segmentation and state vectors (the C ABI's host functions against the
oracle, and both against how the synthetic stream was built), and the
oracle's effective velocity, secondary range compression and RCMC phase
against closed forms."""
import math

import numpy as np
import pytest

import oracle
from sentinel1_decode_amd import capi, synth


@pytest.fixture(scope="module")
def l0():
    """IW1, IW2, IW3 bursts twice over in acquisition order, a 400-packet
    partial burst of IW1 before them and a short (300-packet) last burst."""
    specs = []
    for k in range(2):
        for sw, rg, pri in ((10, 6, 22276), (11, 8, 21861), (12, 11, 25868)):
            specs.append(synth.BurstSpec(n_lines=1100, n_samples=512, rgdec=rg, pri_code=pri, n_targets=5,
                                         swath_number=sw, seed=100 + 3 * k + sw, clutter_sigma=0.1))
    specs[-1] = synth.BurstSpec(n_lines=300, n_samples=512, rgdec=11, pri_code=25868, n_targets=5, swath_number=12,
                                seed=999)
    s = synth.level0_stream(specs, lead_packets=400)
    offs = oracle.index_level0(s.stream)
    return s, offs, oracle.header_fields(s.stream, offs)


def test_index_level0(l0):
    s, offs, _ = l0
    assert len(offs) == s.n_packets
    got = capi.index_level0(s.stream)
    assert got.tolist() == offs
    # a stream cut inside its last packet loses that packet only
    cut = s.stream[: offs[-1] + 100]
    assert capi.index_level0(cut).tolist() == offs[:-1]
    assert oracle.index_level0(cut) == offs[:-1]
    assert capi.index_level0(np.zeros(0, np.uint8)).size == 0


def test_segment_bursts_as_built(l0):
    s, offs, F = l0
    built = {}
    for sw, first, n, _ in s.bursts:
        built.setdefault(sw, []).append(list(range(first, first + n)))
    for sw in (10, 11, 12):
        want = built[sw]
        if sw == 10:
            want = want[1:]  # the 400-packet lead burst is dropped
        ref = oracle.segment_bursts(F, sw)
        assert ref == want, sw
        of, nb = capi.segment_bursts(F, sw)
        assert nb == len(want)
        got = [np.nonzero(of == b)[0].tolist() for b in range(nb)]
        assert got == want, sw
        assert (of[[i for i in range(len(of)) if all(i not in b for b in want)]] == -1).all()
    # the short last burst of IW3 is kept; a calibration read is one burst
    assert len(oracle.segment_bursts(F, 12)[-1]) == 300
    of, nb = capi.segment_bursts(F, 11, cal=True)
    assert nb == 1 and (of >= 0).sum() == sum(len(b) for b in built[11])
    assert oracle.segment_bursts(F, 11, cal=True) == [sum(built[11], [])]


def test_state_vectors_round_trip(l0):
    s, offs, F = l0
    ref = oracle.state_vectors(F)
    # one vector per second of stream, each the orbit model at that second
    assert len(ref) >= 3
    t = [v["time"] for v in ref]
    assert all(b > a for a, b in zip(t, t[1:]))
    for v in ref:
        pos, vel = synth._orbit_state(float(v["time"]))
        assert v["position"] == [float(x) for x in pos]
        assert v["velocity"] == [float(np.float32(x)) for x in vel]
    got = capi.state_vectors(F)
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        assert g.time == r["time"]
        assert list(g.position) == r["position"] and list(g.velocity) == r["velocity"]
        assert list(g.quaternion) == r["quaternion"] and list(g.angular_rate) == r["angular_rate"]


def test_time_stamp_rule():
    """parse_time_stamp keeps bit 23 and bit 8 of the fine time only:
    coarse + 0.5 bit23 + 9 * 2^-16 bit8."""
    for coarse, fine in ((1406987562, 0), (7, 1 << 23), (7, 1 << 8), (7, 0xFFFFFF), (12, 0x123456)):
        raw = (coarse << 24) | fine
        want = coarse + 0.5 * ((fine >> 23) & 1) + 9 * 2.0 ** -16 * ((fine >> 8) & 1)
        assert oracle.parse_time_stamp(raw) == want
    # the C ABI through a cycle of ancillary words carrying that stamp
    for coarse, fine in ((1406987562, 0x800100), (99, 0x000100)):
        F = np.zeros((1 + 3 * 64, 54), np.int64)
        words = synth._subcomm_cycle(0)
        words[37:41] = synth._words((coarse << 24) | fine, 4)
        for k in range(1, F.shape[0]):
            idx = (k - 1) % 64 + 1
            F[k, oracle.SEC["sc_data_word_index"]] = idx
            F[k, oracle.SEC["sc_data_word"]] = words[idx]
        got = capi.state_vectors(F)
        assert got[0].time == coarse + 0.5 * ((fine >> 23) & 1) + 9 * 2.0 ** -16 * ((fine >> 8) & 1)
        assert len(got) == 1  # the second cycle repeats the time stamp


def test_interpolate_state(l0):
    _, _, F = l0
    sv = capi.state_vectors(F)
    t = np.array([v.time for v in sv])
    P = np.array([list(v.position) for v in sv])
    rng = np.random.default_rng(3)
    for q in rng.uniform(t[0], t[-1], 20):
        out = capi.interpolate_state(sv, q)
        for d in range(3):
            assert out.position[d] == pytest.approx(np.interp(q, t, P[:, d]), rel=1e-14, abs=1e-6)
    # outside the span: the first / last vector; at a vector's own time: it
    assert list(capi.interpolate_state(sv, t[0] - 5).position) == list(sv[0].position)
    assert list(capi.interpolate_state(sv, t[-1] + 5).position) == list(sv[-1].position)
    for k in range(len(sv)):
        assert list(capi.interpolate_state(sv, t[k]).position) == list(sv[k].position)


def test_effective_velocity_closed_forms():
    """Over the equator (z = 0, so the reference's ratio z / x is 0 and the
    Earth radius is the semi-major axis a), at the nadir slant range H - a
    the ground-projected speed is v a / H and v_rel = v sqrt(a / H); farther
    out it falls with the cosine of the Earth-centred angle."""
    a = 6378137.0
    pos = np.array([7071137.0, 0.0, 0.0])
    H = float(np.linalg.norm(pos))
    v = 7500.0
    R = np.array([H - a, H - a + 1e5, 850e3, 950e3])
    got = oracle.effective_velocities(pos, v, R)
    assert got[0] == pytest.approx(v * math.sqrt(a / H), rel=1e-15)
    beta = (a * a + H * H - R ** 2) / (2 * a * H)
    assert np.allclose(got, np.sqrt(v * a * v * beta / H), rtol=1e-14)
    assert np.all(np.diff(got) < 0)
    # over the pole (z / x large) the radius tends to the semi-minor axis
    b = 6356752.3142
    pol = np.array([1.0, 0.0, 1e12])
    lat = pol[2] / pol[0]
    re = math.sqrt(((a * a * math.cos(lat)) ** 2 + (b * b * math.sin(lat)) ** 2)
                   / ((a * math.cos(lat)) ** 2 + (b * math.sin(lat)) ** 2))
    Hp = float(np.linalg.norm(pol))
    assert oracle.effective_velocities(pol, v, np.array([Hp - re]))[0] == pytest.approx(v * math.sqrt(re / Hp),
                                                                                          rel=1e-12)


def test_rcmc_phase_is_a_range_shift():
    """exp(-4 i PI f_r dR / c) on a line's range spectrum moves the line by
    2 dR fs / c samples (shift theorem): a band-limited pulse lands there."""
    N, fs = 512, 64.34523813e6
    n = np.arange(N)
    pulse = np.exp(-((n - 200.0) ** 2) / (2 * 3.0 ** 2))
    spec = np.fft.fft(pulse)
    fr = np.fft.fftfreq(N) * fs
    R = np.full(N, 850e3)
    for shift_samples in (0.0, 2.05, -3.5):
        dR = shift_samples * 299792458.0 / (2 * fs)
        D = np.full(N, R[0] / (R[0] + dR))  # R / D - R = dR
        out = oracle.rcmc_phase(spec, R, fr, D)
        y = np.fft.ifft(out)
        want = np.exp(-((n - 200.0 - shift_samples) ** 2) / (2 * 3.0 ** 2))
        assert np.abs(y - want).max() < 1e-9, shift_samples
    # factors of 1: identity
    assert np.array_equal(oracle.rcmc_phase(spec, R, fr, np.ones(N)), spec)


def test_secondary_range_compression_identities():
    """Zero range frequency or zero Doppler (f_a + f_dc = 0) leaves the line
    unchanged; the returned factor is the range-Doppler D of the matched
    filter; the filter is unit modulus."""
    rng = np.random.default_rng(5)
    N = 64
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    R = np.linspace(800e3, 801e3, N)
    v = np.full(N, 7100.0)
    lam = 299792458.0 / 5.405000454334350e9
    y, D = oracle.secondary_range_compression(x, v, R, np.zeros(N), np.full(N, 30.0), 400.0)
    assert np.array_equal(y, x)
    assert np.allclose(D, np.sqrt(1 - (lam * 430.0 / (2 * 7100.0)) ** 2), rtol=1e-15)
    y, D = oracle.secondary_range_compression(x, v, R, np.linspace(-3e7, 3e7, N), np.full(N, 30.0), -30.0)
    assert np.array_equal(y, x) and (D == 1.0).all()
    y, _ = oracle.secondary_range_compression(x, v, R, np.linspace(-3e7, 3e7, N), np.zeros(N), 800.0)
    assert np.allclose(np.abs(y), np.abs(x), rtol=1e-14)
    assert not np.allclose(y, x)
