"""Synthetic Sentinel-1 raw scenes encoded as Level-0 FDBAQ space packets.

The simulator and encoder are C (``s1synth.c``, OpenMP); this module holds the
scene definitions of the five benchmark workloads and returns packet streams
in the form the decoder consumes: one byte stream of concatenated space
packets plus the offset and length of each packet's user data.

Conventions recorded here (they are choices, not reference behaviour):

* clutter is white circular complex Gaussian in the raw domain, CN(0, sigma^2);
* FDBAQ sample rates are the table rates of the range-decimation codes
  (code 6 -> 50.046 MHz, 8 -> 64.345 MHz, 11 -> 54.596 MHz);
* rank and SWST are chosen so that the slant range of sample N/2 is 850 km;
* the processed Doppler band is 0.8 PRF around the Doppler centroid.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

F_REF = 37.53472224
SPEED_OF_LIGHT = 299792458.0
CENTER_FREQ = 5.405000454334350e9
WAVELENGTH = SPEED_OF_LIGHT / CENTER_FREQ
DELTA_T_SUPPRESSED = (320 / (8 * F_REF)) * 1e-6
RANGE_DECIMATION = [112.6041667, 100.0925926, 0.0, 83.41049387, 66.72839509, 56.30208336, 50.04629632,
                    25.02314816, 64.34523813, 46.91840280, 17.32371796, 54.59595962]

SECONDARY_FIELDS = [
    "coarse_time", "fine_time", "sync_marker", "data_take_id", "ecc_number", "na_1", "test_mode",
    "rx_channel_id", "instrument_configuration_id", "sc_data_word_index", "sc_data_word",
    "space_packet_count", "pri_count", "error_flag", "na_2", "baq_mode", "baq_block_length", "na_3",
    "range_decimation", "rx_gain", "tx_ramp_rate_sign", "tx_ramp_rate_mag", "pulse_start_frequency_sign",
    "pulse_start_frequency_mag", "pulse_length", "na_4", "rank", "pri", "swst", "swl", "ssb_flag",
    "polarisation", "temperature_compensation", "na_5", "elevation_beam_address", "na_6",
    "azimuth_beam_address", "calibration_mode", "na_7", "tx_pulse_number", "signal_type", "na_8", "swap",
    "swath_number", "num_quadratures", "na_9"]
PRIMARY_FIELDS = ["packet_version_number", "packet_type", "secondary_header_flag", "process_id",
                  "process_category", "sequence_flags", "packet_sequence_count", "packet_data_length"]


def build(force: bool = False) -> str:
    so = os.path.join(_HERE, "libs1synth.so")
    src = os.path.join(_HERE, "s1synth.c")
    if force or not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-shared", "-std=gnu11", "-o", so, src, "-lm"],
                       check=True)
    return so


class _Scene(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("fs_hz", "k_hz_s", "tau_s", "f_start_hz", "prf_hz", "wavelength", "v_ms", "r_near", "f_dc",
                 "band_frac", "clutter_sigma", "noise_sigma", "gain")]


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "libs1synth.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        L.syn_raw_scene.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(_Scene), C.c_void_p, C.c_int,
                                    C.c_uint64]
        L.syn_pack_headers.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.syn_encode_packet.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_int64]
        L.syn_encode_packet.restype = C.c_int64
        L.syn_encode_burst.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_uint64,
                                       C.c_double, C.c_double, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        L.syn_encode_burst.restype = C.c_int64
        L.syn_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]
        L.syn_uniform.restype = C.c_double
        L.syn_block_codes.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_int, C.c_double, C.c_double,
                                      C.c_void_p, C.c_void_p, C.c_void_p]
        _LIB = L
    return _LIB


# ------------------------------------------------------------------ headers
def pack_headers(primary: dict, secondary: dict) -> bytes:
    p = np.array([int(primary.get(k, 0)) for k in PRIMARY_FIELDS], np.int64)
    s = np.array([int(secondary.get(k, 0)) for k in SECONDARY_FIELDS], np.int64)
    out = np.zeros(68, np.uint8)
    lib().syn_pack_headers(out.ctypes.data, p.ctypes.data, s.ctypes.data)
    return out.tobytes()


def encode_packet(iq: np.ndarray, brc, thidx, scale=None) -> bytes:
    """Encode 2*NQ complex samples with the given per-block BRC/THIDX."""
    iq = np.ascontiguousarray(iq, dtype=np.complex128)
    nq = iq.size // 2
    nb = (nq + 127) // 128
    brc = np.ascontiguousarray(brc, np.uint8)
    thidx = np.ascontiguousarray(thidx, np.uint8)
    sc = np.ones(nb) if scale is None else np.ascontiguousarray(scale, np.float64)
    cap = nq * 5 + 8 * nb + 64
    out = np.zeros(cap, np.uint8)
    n = lib().syn_encode_packet(iq.ctypes.data, nq, brc.ctypes.data, thidx.ctypes.data, sc.ctypes.data,
                                out.ctypes.data, cap)
    if n < 0:
        raise RuntimeError("encoder buffer overflow")
    return out[:n].tobytes()


# ------------------------------------------------------------------ scenes
@dataclasses.dataclass
class BurstSpec:
    """One burst: geometry, radar codes and scene content."""
    n_lines: int
    n_samples: int
    rgdec: int = 8
    pri_code: int = 21861
    pulse_length_code: int = 1967
    txprr_mag: int = 1606
    swath_number: int = 11
    f_dc: float = 0.0
    v_ms: float = 7100.0
    r0: float = 850e3
    rank: int = -1  # -1: the largest rank that keeps SWST >= 0
    n_targets: int = 0
    target_grid: bool = False
    grid_spacing: tuple = (600, 200)
    amp_mode: str = "unit"  # unit | rayleigh
    clutter_sigma: float = 0.1
    noise_sigma: float = 0.0
    gain: float = 5.0
    simple_frac: float = 0.03
    band_frac: float = 0.8
    seed: int = 0

    # ---- header-derived physical quantities (packet.cpp accessors) ----
    @property
    def fs_mhz(self):
        return RANGE_DECIMATION[self.rgdec]

    @property
    def pri_us(self):
        return self.pri_code / F_REF

    @property
    def prf(self):
        return 1.0 / (self.pri_us * 1e-6)

    @property
    def txpl_us(self):
        return self.pulse_length_code / F_REF

    @property
    def txprr(self):
        return self.txprr_mag * (F_REF ** 2 / 2097152)

    @property
    def txpsf_sign_mag(self):
        start = -0.5 * self.txprr * self.txpl_us  # MHz, chirp centred on 0
        mag = int(round(abs(start - self.txprr / (4 * F_REF)) * 16384 / F_REF))
        return (0 if start < 0 else 1), mag

    @property
    def txpsf(self):
        s, m = self.txpsf_sign_mag
        return ((-1 if s == 0 else 1) * m * (F_REF / 16384)) + (self.txprr / (4 * F_REF))

    def _near_delay(self):
        fs = self.fs_mhz * 1e6
        n = self.n_samples
        # R_{N/2} = c/2 (near + (N/2) (N/fs)/(N-1)) = r0
        return 2 * self.r0 / SPEED_OF_LIGHT - (n // 2) * (n / fs) / (n - 1)

    @property
    def rank_value(self):
        if self.rank >= 0:
            return self.rank
        return int(math.floor((self._near_delay() - DELTA_T_SUPPRESSED) / (self.pri_us * 1e-6)))

    @property
    def swst_code(self):
        swst = self._near_delay() - self.rank_value * self.pri_us * 1e-6 - DELTA_T_SUPPRESSED
        return int(round(swst * 1e6 * F_REF))

    def headers(self):
        s_sign, s_mag = self.txpsf_sign_mag
        prim = dict(packet_version_number=0, packet_type=0, secondary_header_flag=1, process_id=65,
                    process_category=12, sequence_flags=3, packet_sequence_count=100, packet_data_length=0)
        sec = dict(coarse_time=1406987562, fine_time=7757, sync_marker=0x352EF853, data_take_id=225252800,
                   ecc_number=8, test_mode=0, rx_channel_id=0, instrument_configuration_id=7,
                   sc_data_word_index=1, sc_data_word=0, space_packet_count=240533, pri_count=243384,
                   baq_mode=12, baq_block_length=31, range_decimation=self.rgdec, rx_gain=8,
                   tx_ramp_rate_sign=1, tx_ramp_rate_mag=self.txprr_mag, pulse_start_frequency_sign=s_sign,
                   pulse_start_frequency_mag=s_mag, pulse_length=self.pulse_length_code, rank=self.rank_value,
                   pri=self.pri_code, swst=self.swst_code,
                   swl=int(round(self.n_samples / self.fs_mhz * F_REF)), polarisation=7,
                   temperature_compensation=3, elevation_beam_address=6, azimuth_beam_address=385,
                   tx_pulse_number=6, signal_type=0, swap=1, swath_number=self.swath_number,
                   num_quadratures=self.n_samples // 2)
        return prim, sec

    def slant_range_near(self):
        """c/2 * (rank PRI + SWST + dt_suppressed) (packet.cpp:284-301)."""
        near = self.rank_value * self.pri_us * 1e-6 + (self.swst_code / F_REF) * 1e-6 + DELTA_T_SUPPRESSED
        return near * SPEED_OF_LIGHT / 2

    def targets(self):
        rng = np.random.default_rng(1000 + self.seed)
        L, N = self.n_lines, self.n_samples
        if self.target_grid:
            dn, da = self.grid_spacing
            rs = N / 2 + dn * np.array([-1, 0, 1])
            az = L / 2 + da * np.array([-1, 0, 1])
            pts = [(r, a, 1.0, 0.0) for a in az for r in rs]
            return np.array(pts, np.float64)
        if self.n_targets == 0:
            return np.zeros((0, 4))
        r = rng.uniform(0, N, self.n_targets)
        a = rng.uniform(0, L, self.n_targets)
        amp = rng.rayleigh(1.0, self.n_targets) if self.amp_mode == "rayleigh" else np.ones(self.n_targets)
        ph = rng.uniform(-np.pi, np.pi, self.n_targets)
        return np.stack([r, a, amp, ph], 1)

    def raw(self):
        """Raw echo [L, N] complex128 (already multiplied by `gain`)."""
        L, N = self.n_lines, self.n_samples
        fs = self.fs_mhz * 1e6
        sc = _Scene(fs_hz=fs, k_hz_s=self.txprr * 1e12, tau_s=self.txpl_us * 1e-6,
                    f_start_hz=self.txpsf * 1e6, prf_hz=self.prf, wavelength=WAVELENGTH, v_ms=self.v_ms,
                    r_near=self.slant_range_near(), f_dc=self.f_dc, band_frac=self.band_frac,
                    clutter_sigma=self.clutter_sigma, noise_sigma=self.noise_sigma, gain=self.gain)
        tg = np.ascontiguousarray(self.targets(), np.float64)
        out = np.empty((L, N), np.complex128)
        lib().syn_raw_scene(out.ctypes.data, L, N, C.byref(sc), tg.ctypes.data if tg.size else None,
                            tg.shape[0], self.seed)
        return out


@dataclasses.dataclass
class Burst:
    spec: BurstSpec
    stream: np.ndarray       # uint8, concatenated space packets
    ud_offsets: np.ndarray   # uint64, user-data offset of each packet in `stream`
    ud_lengths: np.ndarray   # uint32, user-data byte length of each packet
    num_quads: np.ndarray    # int32 per packet
    raw: np.ndarray | None = None

    @property
    def n_lines(self):
        return self.spec.n_lines

    @property
    def n_samples(self):
        return self.spec.n_samples

    @property
    def packet_bytes(self):
        return int(self.stream.size)

    def packet(self, i):
        start = int(self.ud_offsets[i]) - 68
        return self.stream[start: int(self.ud_offsets[i]) + int(self.ud_lengths[i])].tobytes()


def encode_burst(spec: BurstSpec, raw: np.ndarray, keep_raw=False) -> Burst:
    L, N = raw.shape
    nq = N // 2
    nb = (nq + 127) // 128
    # sigma reference for the BRC choice: median block RMS of the scene
    q = raw.reshape(L, -1)
    blk = np.abs(q[:, : (N // 256) * 256].reshape(L, -1, 256)) ** 2 if N >= 256 else np.abs(q) ** 2
    sig_ref = float(np.sqrt(np.median(blk.mean(-1)) / 2.0)) if blk.size else 1.0
    sig_ref = max(sig_ref, 1e-3)
    prim, sec = spec.headers()
    p = np.array([int(prim.get(k, 0)) for k in PRIMARY_FIELDS], np.int64)
    s = np.array([int(sec.get(k, 0)) for k in SECONDARY_FIELDS], np.int64)
    cap = L * (68 + nq * 5 + 8 * nb + 64)
    stream = np.zeros(cap, np.uint8)
    offs = np.zeros(L, np.uint64)
    lens = np.zeros(L, np.uint32)
    rawc = np.ascontiguousarray(raw, np.complex128)
    n = lib().syn_encode_burst(rawc.ctypes.data, L, N, p.ctypes.data, s.ctypes.data, spec.seed, sig_ref,
                               spec.simple_frac, stream.ctypes.data, cap, offs.ctypes.data, lens.ctypes.data)
    if n < 0:
        raise RuntimeError("burst encoder overflow")
    return Burst(spec, stream[:n].copy(), offs, lens, np.full(L, nq, np.int32), rawc if keep_raw else None)


def make_burst(spec: BurstSpec, keep_raw=False) -> Burst:
    return encode_burst(spec, spec.raw(), keep_raw=keep_raw)


# ------------------------------------------------------------------ workloads
def config0(**kw) -> BurstSpec:
    """1,024 x 8,192, nine unit targets on a 3x3 grid, clutter 0.05, noise 0.01, f_dc = 0."""
    d = dict(n_lines=1024, n_samples=8192, target_grid=True, clutter_sigma=0.05, noise_sigma=0.01,
             gain=20.0, f_dc=0.0, seed=0, swath_number=11)
    d.update(kw)
    return BurstSpec(**d)


def config1(**kw) -> BurstSpec:
    """One IW2-sized burst: 1,480 x 24,000, 400 Rayleigh targets, clutter 0.1, f_dc = 30 Hz."""
    d = dict(n_lines=1480, n_samples=24000, n_targets=400, amp_mode="rayleigh", clutter_sigma=0.1,
             gain=5.0, f_dc=30.0, seed=1, swath_number=11)
    d.update(kw)
    return BurstSpec(**d)


def config3_specs(seed=3):
    """27 uneven bursts: 9 IW1-like, 9 IW2-like, 9 IW3-like (200 targets each)."""
    out = []
    for k in range(9):
        out.append(BurstSpec(n_lines=1420, n_samples=21600, rgdec=6, pri_code=22276, n_targets=200,
                             amp_mode="rayleigh", clutter_sigma=0.1, gain=5.0, swath_number=10,
                             seed=seed * 1000 + 3 * k))
        out.append(BurstSpec(n_lines=1480, n_samples=24000, rgdec=8, pri_code=21861, n_targets=200,
                             amp_mode="rayleigh", clutter_sigma=0.1, gain=5.0, swath_number=11,
                             seed=seed * 1000 + 3 * k + 1))
        out.append(BurstSpec(n_lines=1520, n_samples=20400, rgdec=11, pri_code=25868, n_targets=200,
                             amp_mode="rayleigh", clutter_sigma=0.1, gain=5.0, swath_number=12,
                             seed=seed * 1000 + 3 * k + 2))
    return out


def config4(**kw) -> BurstSpec:
    """Autofocus burst: 4,096 x 8,192, 200 targets, clutter 0.1."""
    d = dict(n_lines=4096, n_samples=8192, n_targets=200, amp_mode="rayleigh", clutter_sigma=0.1, gain=5.0,
             f_dc=0.0, seed=4)
    d.update(kw)
    return BurstSpec(**d)


def add_rfi(raw: np.ndarray, spec: BurstSpec, seed: int, frac=0.2, run=(10, 50), tones=3,
            inr_db=(15.0, 25.0), max_offset_hz=28e6):
    """Narrowband RFI on `frac` of the lines in contiguous runs; each affected
    line carries `tones` tones at uniform offsets within +-max_offset_hz whose
    power relative to the line's mean power is drawn from inr_db.  Returns the
    boolean mask of affected lines."""
    rng = np.random.default_rng(seed)
    L, N = raw.shape
    mask = np.zeros(L, bool)
    target = int(round(frac * L))
    while mask.sum() < target:
        length = int(rng.integers(run[0], run[1] + 1))
        start = int(rng.integers(0, max(1, L - length)))
        mask[start: start + length] = True
    fs = spec.fs_mhz * 1e6
    n = np.arange(N)
    for l in np.nonzero(mask)[0]:
        p = float(np.mean(np.abs(raw[l]) ** 2))
        for _ in range(tones):
            f = rng.uniform(-max_offset_hz, max_offset_hz)
            inr = rng.uniform(*inr_db)
            a = math.sqrt(p * 10 ** (inr / 10))
            raw[l] += a * np.exp(1j * (2 * np.pi * f / fs * n + rng.uniform(-np.pi, np.pi)))
    return mask


# ------------------------------------------------------------ Level-0 streams
def _orbit_state(t):
    """A circular 693 km orbit (inclination 98.18 deg) in an Earth-fixed frame
    without Earth rotation: ECEF position (m) and velocity (m/s) at time t (s)."""
    r = 6378137.0 + 693e3
    mu = 3.986004418e14
    w = math.sqrt(mu / r ** 3)
    inc = math.radians(98.18)
    c, s_ = math.cos(w * t), math.sin(w * t)
    pos = np.array([r * c, r * s_ * math.cos(inc), r * s_ * math.sin(inc)])
    vel = np.array([-r * w * s_, r * w * c * math.cos(inc), r * w * c * math.sin(inc)])
    return pos, vel


def _words(value, n_words):
    return [(value >> (16 * (n_words - 1 - k))) & 0xFFFF for k in range(n_words)]


def _subcomm_cycle(t_coarse):
    """The 64 sub-commutated ancillary words (word indices 1..64) carrying the
    state vector at time t_coarse (integer seconds): positions as IEEE
    doubles (4 words each), velocities, quaternion and rates as IEEE floats
    (2 words), the time stamp as 32 coarse + 24 fine bits (4 words)."""
    import struct
    pos, vel = _orbit_state(float(t_coarse))
    d64 = lambda x: struct.unpack(">Q", struct.pack(">d", float(x)))[0]
    f32 = lambda x: struct.unpack(">I", struct.pack(">f", float(x)))[0]
    w = [0] * 65
    for k, idx in enumerate((1, 5, 9)):
        w[idx: idx + 4] = _words(d64(pos[k]), 4)
    for k, idx in enumerate((13, 15, 17)):
        w[idx: idx + 2] = _words(f32(vel[k]), 2)
    w[19:23] = _words(int(t_coarse) << 24, 4)                       # pod data stamp
    for k, (idx, q) in enumerate(zip((23, 25, 27, 29), (0.5, -0.5, 0.5, 0.5))):
        w[idx: idx + 2] = _words(f32(q), 2)
    for k, idx in enumerate((31, 33, 35)):
        w[idx: idx + 2] = _words(f32(1e-3 * (k + 1)), 2)
    w[37:41] = _words(int(t_coarse) << 24, 4)                       # data time stamp
    return w


@dataclasses.dataclass
class Level0:
    """A Level-0 byte stream of several bursts, with what built it: per burst
    its spec, swath number and the stream indices of its packets, and the
    orbit's state-vector times."""
    stream: np.ndarray
    bursts: list            # [(swath_number, first_packet, n_packets, BurstSpec)]
    n_packets: int


def level0_stream(specs, t0=1406987562, first_pri=243384, lead_packets=0, beam_addresses=(100, 900)):
    """Concatenate bursts (BurstSpec, encoded in turn) into one Level-0 stream
    in acquisition order.  Each packet's headers are rewritten as the
    instrument writes them: pri_count and space_packet_count count up over
    the stream (so a swath's consecutive bursts are separated by a PRI gap),
    azimuth_beam_address ramps within a burst and restarts with the next
    one, coarse/fine time follow the PRI, and sc_data_word_index cycles
    1..64 carrying the orbit's state vector (one per second of stream).
    `lead_packets` > 0 puts that many packets of the first spec's swath
    before everything (a partial burst the reader must drop).
    `beam_addresses` are the first and last azimuth beam addresses of a
    burst's ramp (IW sweeps about 213..810, +-0.0105 rad)."""
    chunks, bursts = [], []
    pkt = 0
    clock = []  # each packet's time since the stream start (s), one PRI apart
    items = []
    if lead_packets:
        sp0 = dataclasses.replace(specs[0], n_lines=lead_packets, seed=specs[0].seed + 7777)
        items.append(sp0)
    items.extend(specs)
    for sp in items:
        b = make_burst(sp)
        n = sp.n_lines
        bursts.append((sp.swath_number, pkt, n, sp))
        for l in range(n):
            raw = bytearray(b.packet(l))
            p_, s_ = _parse_header_fields(bytes(raw[:68]))
            clock.append(clock[-1] + 1.0 / sp.prf if clock else 0.0)
            t = clock[-1]
            cycle_pos = (pkt + l) % 64
            sc_idx = cycle_pos + 1
            # one cycle carries one state vector: the one at its first packet
            words = _subcomm_cycle(t0 + int(clock[pkt + l - cycle_pos]))
            s_["pri_count"] = first_pri + pkt + l
            s_["space_packet_count"] = 240533 + pkt + l
            a0, a1 = beam_addresses
            s_["azimuth_beam_address"] = a0 + (l * (a1 - a0)) // max(n - 1, 1)
            s_["coarse_time"] = t0 + int(t)
            s_["fine_time"] = max(1, int((t - int(t)) * 65536))
            s_["sc_data_word_index"] = sc_idx
            s_["sc_data_word"] = words[sc_idx]
            p_["packet_sequence_count"] = (pkt + l) & 0x3FFF
            raw[:68] = pack_headers(p_, s_)
            chunks.append(bytes(raw))
        pkt += n
    stream = np.frombuffer(b"".join(chunks), np.uint8).copy()
    return Level0(stream, bursts, pkt)


def _parse_header_fields(h):
    vals = []
    bit = 0
    widths = [3, 1, 1, 7, 4, 2, 14, 16, 32, 16, 32, 32, 8, 1, 3, 4, 32, 8, 16, 32, 32, 1, 2, 5, 8, 8, 8, 8, 1, 15, 1,
              15, 24, 3, 5, 24, 24, 24, 1, 3, 2, 2, 4, 2, 10, 2, 1, 5, 4, 3, 1, 8, 16, 8]
    v = int.from_bytes(h, "big")
    total = 8 * len(h)
    for w in widths:
        vals.append((v >> (total - bit - w)) & ((1 << w) - 1))
        bit += w
    return dict(zip(PRIMARY_FIELDS, vals[:8])), dict(zip(SECONDARY_FIELDS, vals[8:]))
