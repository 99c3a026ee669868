"""CPU FP64 oracle for the Sentinel-1 L0 -> SLC hot path.

TEST INFRASTRUCTURE ONLY.  This package is the parity checker and the timed
CPU baseline.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it; the product package never does.

Thin numpy/ctypes wrapper over ``libs1oracle.so`` (built from
``s1_oracle.c`` by ``oracle/Makefile``).  Each wrapper names the reference
function (AndrewPlayer3/sentinel1_decode ``src/`` file:line) it restates.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build(force: bool = False) -> None:
    """Compile the oracle library (and the reference decode shim where the
    reference sources exist)."""
    so = os.path.join(_HERE, "libs1oracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "libs1oracle.so")
        if not os.path.exists(so):
            build()
        _LIB = _bind(C.CDLL(so))
    return _LIB


def build_mutant(mutation: int, dest_dir: str) -> str:
    """Compile the oracle with planted mistake `mutation` (see the
    ORC_MUTATION list in s1_oracle.c) into dest_dir; returns the library path.
    Used by tests that show a check fails on the mistake it is meant to catch."""
    so = os.path.join(dest_dir, f"libs1oracle_m{mutation}.so")
    subprocess.run(["make", "-s", "-C", _HERE, f"MUTANT={so}", f"MUTATION={mutation}", "mutant"], check=True)
    return so


@contextlib.contextmanager
def using(so_path: str):
    """Route every wrapper in this module through the library at so_path
    (an oracle variant from build_mutant) for the duration of the block."""
    global _LIB
    saved = lib()
    _LIB = _bind(C.CDLL(so_path))
    try:
        yield
    finally:
        _LIB = saved


def _bind(L):
    """Declare the argument types of the oracle's C functions on library L."""
    L.orc_read_n_bits.argtypes = [_u8, C.c_int64, C.c_int64, C.c_int, C.POINTER(C.c_uint64)]
    L.orc_huffman_decode.argtypes = [_u8, C.c_int64, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int)]
    L.orc_s_value_d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]
    L.orc_s_value_d.restype = C.c_double
    L.orc_decode_packet_d.argtypes = [_u8, C.c_int64, C.c_int, _dp, _u8, _u8]
    L.orc_decode_packets.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                     C.c_void_p, C.c_int64, C.c_void_p]
    L.orc_parse_headers.argtypes = [_u8, C.c_int64, C.c_void_p, C.c_void_p]
    L.orc_packet_params.argtypes = [C.c_void_p, _dp]
    L.orc_slant_ranges.argtypes = [C.c_void_p, C.c_int, _dp]
    L.orc_timing_corrections.argtypes = [C.c_void_p, _dp]
    L.orc_linspace.argtypes = [C.c_double, C.c_double, C.c_int, _dp]
    L.orc_fft.argtypes = [_dp, C.c_int, C.c_int]
    L.orc_fft_many.argtypes = [_dp, C.c_int, C.c_int, C.c_int64, C.c_int]
    L.orc_replica_chirp_length.argtypes = [C.c_double, C.c_double]
    L.orc_replica_chirp.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp]
    L.orc_reference_function.argtypes = [_dp, C.c_int, _dp]
    L.orc_range_compress.argtypes = [_dp, C.c_int, C.c_int, C.c_int64, C.c_double, C.c_double,
                                     C.c_double, C.c_double]
    L.orc_azimuth_fft.argtypes = [_dp, C.c_int, C.c_int, C.c_int64]
    L.orc_azimuth_ifft.argtypes = [_dp, C.c_int, C.c_int, C.c_int64]
    L.orc_rcmc_mf.argtypes = [_dp, C.c_int, C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double,
                              _dp, _dp, _dp, _dp, C.c_int]
    L.orc_sinc_interpolate.argtypes = [_dp, C.c_int, C.c_double, C.c_int, _dp]
    L.orc_eccm.argtypes = [_dp, C.c_int, C.c_int, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_double, _u8]
    L.orc_af_entropy_grad.argtypes = [_dp, C.c_int, C.c_int, C.c_int64, _dp, _dp]
    L.orc_af_entropy_grad.restype = C.c_double
    L.orc_effective_velocities.argtypes = [_dp, C.c_double, _dp, C.c_int, _dp]
    L.orc_secondary_range_compression.argtypes = [_dp, C.c_int, _dp, _dp, _dp, _dp, C.c_double, _dp]
    L.orc_rcmc_phase.argtypes = [_dp, C.c_int, _dp, _dp, _dp]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set_threads.restype = C.c_int
    L.orc_dc_wrapped.argtypes = [_dp, C.c_int, C.c_int, C.c_int64, C.c_double, C.c_double, _dp]
    L.orc_dc_unwrap.argtypes = [_dp, C.c_int, C.c_int, _dp, _dp, C.c_double, _dp]
    L.orc_dc_unwrap.restype = C.c_int
    L.orc_tops_freq_shape.argtypes = [C.c_int, C.c_double]
    L.orc_tops_freq_shape.restype = C.c_int
    L.orc_az_freq_ufr.argtypes = [_dp, C.c_int, C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
                                  _dp]
    L.orc_tops_rows.argtypes = [_dp, C.c_int, C.c_int, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, _dp]
    L.orc_linear_resample_many.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp]
    L.orc_quadratic_resample.argtypes = [_dp, C.c_int, C.c_int, _dp]
    L.orc_tops_time_params.argtypes = [_dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_int, _dp, np.ctypeslib.ndpointer(np.int32)]
    L.orc_tops_time_params.restype = C.c_int
    L.orc_az_time_ufr.argtypes = [_dp, C.c_int, C.c_int, C.c_int64, _dp, _dp, C.c_double, _dp,
                                  np.ctypeslib.ndpointer(np.int32), _dp]
    return L


def set_threads(n: int) -> int:
    """OpenMP threads for the oracle's loops (n <= 0: query); returns the
    count in effect."""
    return lib().orc_set_threads(int(n))


def _table_call(fn):
    out = dict(thidx=np.zeros(5, np.int32), mflag=np.zeros(5, np.int32), simple=np.zeros(50), nrl=np.zeros(80),
               sf=np.zeros(256), rgdec=np.zeros(12), consts=np.zeros(8), huff=np.zeros(160, np.int32),
               huff_n=np.zeros(5, np.int32), huff_start=np.zeros(5, np.int32))
    fn(*[v.ctypes.data for v in out.values()])
    out["simple"] = out["simple"].reshape(5, 10)
    out["nrl"] = out["nrl"].reshape(5, 16)
    h = out["huff"].reshape(5, 16, 2)
    out["huff"] = {b: sorted((int(h[b, k, 0]), int(h[b, k, 1])) for k in range(out["huff_n"][b])) for b in range(5)}
    return out


def tables():
    """The oracle's decode tables and constants."""
    L = lib()
    L.orc_tables.argtypes = [C.c_void_p] * 10
    return _table_call(L.orc_tables)


def ref_tables():
    """The reference's own tables (structs.h) via the shim, or None."""
    R = ref_lib()
    if R is None:
        return None
    R.ref_tables.argtypes = [C.c_void_p] * 10
    return _table_call(R.ref_tables)


def ref_parse_headers(packet: bytes):
    """The reference's parse_packet_header over its own width tables, or None."""
    R = ref_lib()
    if R is None:
        return None
    R.ref_parse_headers.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    d = np.frombuffer(bytes(packet[:68]), np.uint8).copy()
    p = np.zeros(8, np.int64)
    s = np.zeros(46, np.int64)
    if R.ref_parse_headers(d.ctypes.data, p.ctypes.data, s.ctypes.data):
        raise ValueError("reference header parse failed")
    return p, s


def ref_lib():
    """The reference's own decoding_utils.cpp behind a C shim, or None where it
    was not built (the reference tree is absent)."""
    global _REF
    if _REF is None:
        so = os.path.join(_HERE, "_ref", "libref_decoding.so")
        if not os.path.exists(so):
            return None
        R = C.CDLL(so)
        R.ref_read_n_bits.argtypes = [_u8, C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_uint64)]
        R.ref_huffman_decode.argtypes = [_u8, C.c_int64, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        _REF = R
    return _REF


def _c128(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


def _as_f64(a):
    return a.view(np.float64).reshape(-1) if a.size else np.zeros(2)


# ---------------------------------------------------------------- decoding
def read_n_bits(data: bytes, start_bit: int, n: int):
    """decoding_utils.cpp:95-134; returns (status, value)."""
    d = np.frombuffer(bytes(data) or b"\0", dtype=np.uint8).copy()
    out = C.c_uint64(0)
    st = lib().orc_read_n_bits(d, len(data), start_bit, n, C.byref(out))
    return st, out.value


def huffman_decode(data: bytes, brc: int, bit_index: int):
    """decoding_utils.cpp:44-77; returns (status, m_code, new_bit_index)."""
    d = np.frombuffer(bytes(data) or b"\0", dtype=np.uint8).copy()
    bi = C.c_int64(bit_index)
    m = C.c_int(-1)
    st = lib().orc_huffman_decode(d, len(data), brc, C.byref(bi), C.byref(m))
    return st, m.value, bi.value


def s_value(brc: int, thidx: int, sign: int, m_code: int):
    """packet.cpp:678-693; returns (status, value)."""
    st = C.c_int(0)
    v = lib().orc_s_value_d(brc, thidx, sign, m_code, C.byref(st))
    return st.value, v


def decode_packet(user_data: bytes, num_quads: int):
    """packet.cpp:404-453,696-818 -> (status, complex128[2*NQ], brc[NB], thidx[NB])."""
    nb = -(-2 * num_quads // 256) if num_quads > 0 else 0
    d = np.frombuffer(bytes(user_data) or b"\0", dtype=np.uint8).copy()
    out = np.zeros(max(4 * num_quads, 2), dtype=np.float64)
    brc = np.zeros(max(nb, 1), np.uint8)
    thr = np.zeros(max(nb, 1), np.uint8)
    st = lib().orc_decode_packet_d(d, len(user_data), num_quads, out, brc, thr)
    sig = out[: 4 * num_quads].view(np.complex128).copy()
    return st, sig, brc[:nb].copy(), thr[:nb].copy()


def decode_packets(stream: np.ndarray, offsets, lengths, num_quads, n_samples=None):
    """Batch decode (s1_decoder.cpp:462-468) -> (complex128[L, N], status[L])."""
    stream = np.ascontiguousarray(stream, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    nq = np.ascontiguousarray(num_quads, dtype=np.int32)
    L = len(offsets)
    N = int(n_samples if n_samples is not None else 2 * int(nq.max(initial=0)))
    out = np.zeros((L, max(N, 1)), np.complex128)
    status = np.zeros(L, np.int32)
    lib().orc_decode_packets(stream.ctypes.data, offsets.ctypes.data, lengths.ctypes.data, nq.ctypes.data,
                             L, out.ctypes.data, out.shape[1], status.ctypes.data)
    return out[:, :N], status


# ---------------------------------------------------------------- headers
def parse_headers(packet: bytes):
    """parse_packet_header (packet.cpp:829-846) over the 6 + 62 header bytes."""
    d = np.frombuffer(bytes(packet), dtype=np.uint8).copy()
    p = np.zeros(8, np.int64)
    s = np.zeros(46, np.int64)
    st = lib().orc_parse_headers(d, len(d), p.ctypes.data, s.ctypes.data)
    if st:
        raise ValueError("packet shorter than its headers")
    return p, s


def packet_params(sec: np.ndarray):
    """packet.cpp:29-100,277-281,376-380 as a dict."""
    sec = np.ascontiguousarray(sec, dtype=np.int64)
    out = np.zeros(10)
    lib().orc_packet_params(sec.ctypes.data, out)
    keys = ["pri_us", "pulse_length_us", "swl_us", "swst_us", "rx_gain_db", "txprr", "txpsf", "fs_mhz",
            "replica_chirp_length", "baq_block_length"]
    return dict(zip(keys, out.tolist()))


def slant_ranges(sec, num_ranges=0):
    sec = np.ascontiguousarray(sec, dtype=np.int64)
    n = num_ranges if num_ranges > 0 else 2 * int(sec[44])
    out = np.zeros(n)
    lib().orc_slant_ranges(sec.ctypes.data, n, out)
    return out


def timing_corrections(sec):
    sec = np.ascontiguousarray(sec, dtype=np.int64)
    out = np.zeros(2 * int(sec[44]))
    lib().orc_timing_corrections(sec.ctypes.data, out)
    return out


def linspace(start, end, size):
    out = np.zeros(size)
    lib().orc_linspace(start, end, size, out)
    return out


# ---------------------------------------------------------------- DSP
def fft(x, inverse=False):
    """Unnormalised DFT with FFTW's sign conventions (signal_processing.cpp:516-578)."""
    y = _c128(x).copy()
    lib().orc_fft(_as_f64(y), y.size, int(inverse))
    return y


def fft_rows(x, inverse=False):
    y = np.ascontiguousarray(np.array(x, dtype=np.complex128, copy=True))
    lib().orc_fft_many(_as_f64(y), y.shape[1], y.shape[0], y.shape[1], int(inverse))
    return y


def replica_chirp_length(fs_mhz, txpl_us):
    return lib().orc_replica_chirp_length(fs_mhz, txpl_us)


def replica_chirp(txpsf, txprr, txpl, fs_mhz, num_range):
    nc = replica_chirp_length(fs_mhz, txpl)
    out = np.zeros(num_range + 2 * nc, np.complex128)
    lib().orc_replica_chirp(txpsf, txprr, txpl, fs_mhz, num_range, _as_f64(out))
    return out


def reference_function(chirp):
    chirp = _c128(chirp)
    out = np.zeros_like(chirp)
    lib().orc_reference_function(_as_f64(chirp), chirp.size, _as_f64(out))
    return out


def range_compress(data, txpsf, txprr, txpl, fs_mhz):
    """_range_compress without ECCM (s1_decoder.cpp:476-497), returns a copy."""
    y = np.ascontiguousarray(np.array(data, dtype=np.complex128, copy=True))
    L, N = y.shape
    lib().orc_range_compress(_as_f64(y), L, N, N, txpsf, txprr, txpl, fs_mhz)
    return y


def azimuth_fft(data):
    y = np.ascontiguousarray(np.array(data, dtype=np.complex128, copy=True))
    lib().orc_azimuth_fft(_as_f64(y), y.shape[0], y.shape[1], y.shape[1])
    return y


def azimuth_ifft(data):
    y = np.ascontiguousarray(np.array(data, dtype=np.complex128, copy=True))
    lib().orc_azimuth_ifft(_as_f64(y), y.shape[0], y.shape[1], y.shape[1])
    return y


def sinc_interpolate(row, t, half_width=8):
    row = _c128(row)
    out = np.zeros(2)
    lib().orc_sinc_interpolate(_as_f64(row), row.size, float(t), half_width, out)
    return complex(out[0], out[1])


def rcmc_mf(data, prf, range_spacing, ref_range, slant_ranges_, v_rel, f_dc, t_corr, apply_rcmc=True):
    y = np.ascontiguousarray(np.array(data, dtype=np.complex128, copy=True))
    L, N = y.shape
    f = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (N,)))
    lib().orc_rcmc_mf(_as_f64(y), L, N, N, prf, range_spacing, ref_range, f(slant_ranges_), f(v_rel), f(f_dc),
                      f(t_corr), int(apply_rcmc))
    return y


def eccm(data, det, mit, fft_size=128, stride=64):
    """signal_processing.cpp:913-1004 -> (cleaned copy, per-line detection flags)."""
    y = np.ascontiguousarray(np.array(data, dtype=np.complex128, copy=True))
    L, N = y.shape
    flags = np.zeros(L, np.uint8)
    lib().orc_eccm(_as_f64(y), L, N, N, fft_size, stride, float(det), float(mit), flags)
    return y, flags


def af_entropy_grad(rd, phase):
    rd = np.ascontiguousarray(rd, dtype=np.complex128)
    L, N = rd.shape
    ph = np.ascontiguousarray(phase, dtype=np.float64)
    g = np.zeros(L)
    e = lib().orc_af_entropy_grad(_as_f64(rd), L, N, N, ph, g)
    return e, g


# ------------------------------------------------------- Level-0 files
def index_level0(stream) -> list:
    """get_next_packet / get_packets (packet.cpp:850-905): packet k starts
    where packet k-1 ends; its length is 6 + packet_data_length + 1 (primary
    header bytes 4-5); a packet running past the end of the stream is not
    returned.  Returns the packets' start offsets."""
    b = bytes(stream)
    out, pos = [], 0
    while len(b) - pos >= 6:
        n = 6 + ((b[pos + 4] << 8) | b[pos + 5]) + 1
        if pos + n > len(b):
            break
        out.append(pos)
        pos += n
    return out


SEC = {name: 8 + i for i, name in enumerate([
    "coarse_time", "fine_time", "sync_marker", "data_take_id", "ecc_number", "na_1", "test_mode",
    "rx_channel_id", "instrument_configuration_id", "sc_data_word_index", "sc_data_word", "space_packet_count",
    "pri_count", "error_flag", "na_2", "baq_mode", "baq_block_length", "na_3", "range_decimation", "rx_gain",
    "tx_ramp_rate_sign", "tx_ramp_rate_mag", "pulse_start_frequency_sign", "pulse_start_frequency_mag",
    "pulse_length", "na_4", "rank", "pri", "swst", "swl", "ssb_flag", "polarisation", "temperature_compensation",
    "na_5", "elevation_beam_address", "na_6", "azimuth_beam_address", "calibration_mode", "na_7",
    "tx_pulse_number", "signal_type", "na_8", "swap", "swath_number", "num_quadratures", "na_9"])}


def header_fields(stream, offsets) -> np.ndarray:
    """[n][54] header fields of the packets at `offsets` (parse_packet_header,
    packet.cpp:829-846)."""
    b = np.frombuffer(bytes(stream), np.uint8)
    out = np.zeros((len(offsets), 54), np.int64)
    for k, o in enumerate(offsets):
        p, s_ = parse_headers(b[o: o + 68].tobytes())
        out[k, :8] = p
        out[k, 8:] = s_
    return out


def segment_bursts(fields, swath_number, cal=False) -> list:
    """get_packets_in_bursts (packet.cpp:1026-1076) over the flat packet list:
    the packet indices of each burst of one swath."""
    def i32(v):
        v = int(v) & 0xFFFFFFFF
        return v - (1 << 32) if v >= 1 << 31 else v
    bursts, cur = [], []
    az = pri = prev_az = prev_pri = 0
    n = len(fields)
    for i in range(n):
        f = fields[i]
        if i32(f[SEC["swath_number"]]) == swath_number:
            if not cal:
                az = i32(f[SEC["azimuth_beam_address"]])
                pri = i32(f[SEC["pri_count"]])
                if prev_az == 0:
                    prev_az, prev_pri = az, pri
                if az < prev_az or pri > prev_pri + 1:
                    if len(cur) < 1000:
                        cur = []
                    else:
                        bursts.append(cur)
                        cur = []
                prev_pri, prev_az = pri, az
            cur.append(i)
        if i == n - 1 and len(cur) > 0:
            bursts.append(cur)
    return bursts


# SUB_COMM_KEY_POS (structs.h:622-689): the quantity of each word index
_SUBCOMM = (["dummy"] + ["px"] * 4 + ["py"] * 4 + ["pz"] * 4 + ["vx"] * 2 + ["vy"] * 2 + ["vz"] * 2
            + ["pod"] * 4 + ["q0"] * 2 + ["q1"] * 2 + ["q2"] * 2 + ["q3"] * 2 + ["wx"] * 2 + ["wy"] * 2
            + ["wz"] * 2 + ["time"] * 4 + ["other"] * 24)


def int_to_ieee754(number, is_double):
    """decoding_utils.cpp:149-170."""
    msh = 52 if is_double else 23
    emask = 0x7FF0000000000000 if is_double else 0x7F800000
    bias = 1023 if is_double else 127
    ssh = 63 if is_double else 31
    sign = (number >> ssh) & 1
    e = ((number & emask) >> msh) - bias
    total, power = 0.0, -1
    for i in range(msh):
        total += ((number >> (msh - i - 1)) & 1) * 2.0 ** power
        power -= 1
    return (-1 if sign else 1) * 2.0 ** e * (total + 1.0)


def parse_time_stamp(raw):
    """state_vectors.cpp:5-16, loop as written (only bit 8 of the fine time
    survives its first step)."""
    coarse = (raw >> 24) & 0xFFFFFFFF
    fine_raw = raw & 0xFFFFFF
    fine = 0.0
    for i in range(23, -1, -1):
        fine += (fine_raw >> i) * 2.0 ** (i - 24)
        fine_raw &= 1 << 8
    return float(coarse) + fine


def state_vectors(fields) -> list:
    """build_data_word_dicts + STATE_VECTORS::parse_subcomm_dicts
    (state_vectors.cpp:19-91): dicts of time, position, velocity, quaternion
    and angular-rate words, one per complete 64-word cycle, kept while their
    time stamps increase."""
    n = len(fields)
    start = 1
    while start < n and int(fields[start][SEC["sc_data_word_index"]]) != 1:
        start += 1
    out, latest = [], 0.0
    raw = {k: 0 for k in set(_SUBCOMM)}
    for i in range(start, n):
        idx = int(fields[i][SEC["sc_data_word_index"]])
        if idx == 1 and i != start:
            d = dict(time=parse_time_stamp(raw["time"]),
                     position=[int_to_ieee754(raw[k], True) for k in ("px", "py", "pz")],
                     velocity=[int_to_ieee754(raw[k], False) for k in ("vx", "vy", "vz")],
                     quaternion=[int_to_ieee754(raw[k], False) for k in ("q0", "q1", "q2", "q3")],
                     angular_rate=[int_to_ieee754(raw[k], False) for k in ("wx", "wy", "wz")])
            if d["time"] > latest:
                latest = d["time"]
                out.append(d)
            raw = {k: 0 for k in raw}
        if 0 <= idx <= 64:
            k = _SUBCOMM[idx]
            raw[k] = ((raw[k] << 16) | (int(fields[i][SEC["sc_data_word"]]) & 0xFFFF)) & 0xFFFFFFFFFFFFFFFF
    return out


def effective_velocities(position, v, slant_ranges_):
    R = np.ascontiguousarray(slant_ranges_, np.float64)
    out = np.zeros(R.size)
    lib().orc_effective_velocities(np.ascontiguousarray(position, np.float64), float(v), R, R.size, out)
    return out


def secondary_range_compression(line, v_rel, slant_ranges_, range_freqs, f_dc, az_freq):
    y = np.array(line, np.complex128, copy=True)
    n = y.size
    f = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (n,)))
    fac = np.zeros(n)
    lib().orc_secondary_range_compression(_as_f64(y), n, f(v_rel), f(slant_ranges_), f(range_freqs), f(f_dc),
                                          float(az_freq), fac)
    return y, fac


def rcmc_phase(line, slant_ranges_, range_freqs, factors):
    y = np.array(line, np.complex128, copy=True)
    n = y.size
    f = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (n,)))
    lib().orc_rcmc_phase(_as_f64(y), n, f(slant_ranges_), f(range_freqs), f(factors))
    return y


# ----------------------------------------------------------- image output
SCALE_MODES = {"norm_log": 0, "norm": 1, "mag": 2, "real": 3, "imag": 4}


def scale_2d(img, mode):
    """scale / norm_2d / magnitude_2d (signal_processing.cpp:367-460,
    1007-1071) in FP64, stored as float32: norm adds 1e-12 to |x| before
    dividing by max |x|; mag is |x|^2."""
    x = np.asarray(img, np.complex128)
    if mode in ("real", "imag"):
        return (x.real if mode == "real" else x.imag).astype(np.float32)
    a = np.hypot(x.real, x.imag)
    if mode == "mag":
        return (a * a).astype(np.float32)
    mx = a.max()
    v = a + 1e-12
    return (20.0 * np.log10(v / mx) if mode == "norm_log" else v / mx).astype(np.float32)


# -------------------------------------------------------------- TOPS (IW)
WAVELENGTH = 299792458.0 / 5.405000454334350e9


def azimuth_beam_angle(address):
    """L0Packet::get_azimuth_beam_angle (packet.cpp:83-90):
    linspace(-0.018, 0.018, 1024)[address] in radians."""
    return float(linspace(-0.018, 0.018, 1024)[int(address)])


def doppler_centroid_rate(pri_us, first_address, last_address, n_packets, velocity):
    """get_doppler_centroid_rate (doppler.cpp:224-232): -2 v k_psi / lambda,
    k_psi = (last - first beam angle) / (pri n_packets)."""
    pri = pri_us * 1e-6
    steer = (azimuth_beam_angle(last_address) - azimuth_beam_angle(first_address)) / (pri * n_packets)
    return (-2 * velocity * steer) / WAVELENGTH


def dc_wrapped(rc, rate, prf):
    """get_wrapped_estimates over dce_preconditioning (doppler.cpp:4-143)
    with the burst duration fixed at 1.06 s as get_doppler_centroid passes it
    (:161): one estimate per range column."""
    x = _c128(rc)
    L, n = x.shape
    out = np.zeros(n)
    lib().orc_dc_wrapped(_as_f64(x), L, n, n, float(rate), float(prf), out)
    return out


def dc_unwrap(wrapped, t_near, t_far, prf, n_blocks=30):
    """The rest of get_doppler_centroid (doppler.cpp:146-221): block means,
    get_unwrapped_estimates (:42-95), polyfit (signal_processing.cpp:159-214)
    and its value at every range sample; the range times are
    L0Packet::get_slant_range_times (packet.cpp:304-318) = linspace(t_near,
    t_far, count)."""
    w = np.ascontiguousarray(wrapped, np.float64)
    n = w.size
    bt = linspace(t_near, t_far, n_blocks)
    t = linspace(t_near, t_far, n)
    out = np.zeros(n)
    if lib().orc_dc_unwrap(w, n, n_blocks, bt, t, float(prf), out) != 0:
        raise ValueError("need at least two range blocks")
    return out


def doppler_centroid(rc, rate, prf, t_near, t_far, n_blocks=30):
    """get_doppler_centroid (doppler.cpp:146-221)."""
    return dc_unwrap(dc_wrapped(rc, rate, prf), t_near, t_far, prf, n_blocks)


def tops_replicas(rate, burst_duration, prf):
    """num_replicas = |rate T / prf| (image_formation.cpp:142)."""
    return abs(rate * burst_duration / prf)


def az_freq_ufr(rc, rate, burst_duration, prf, bw, nr=None, inv_rate=None):
    """azimuth_frequency_ufr (image_formation.cpp:122-235): [floor(nr L)][n].
    nr / inv_rate override |rate T / prf| and 1 / rate (drives the chain
    without its ramp, for closed-form checks)."""
    x = _c128(rc)
    L, n = x.shape
    nr = tops_replicas(rate, burst_duration, prf) if nr is None else float(nr)
    ir = 1.0 / rate if inv_rate is None else float(inv_rate)
    shape = lib().orc_tops_freq_shape(L, nr)
    mid = shape // 2
    start, end = int(mid - np.floor(bw) / 2), int(mid + np.floor(bw) / 2)
    if shape < 1 or start < 0 or end > shape:
        raise ValueError("the low-pass window leaves the spectrum")
    out = np.zeros((shape, n), np.complex128)
    lib().orc_az_freq_ufr(_as_f64(x), L, n, n, nr, ir, float(prf), float(bw), _as_f64(out))
    return out


def tops_rows(data, v_rel, slant_ranges_, range_freqs, f_dc, az_freqs, t_corr, ref_range):
    """The TOPS branch of the Doppler-row loop of S1_Decoder::_azimuth_compress
    (s1_decoder.cpp:626-677) on every line: range FFT, secondary range
    compression, unnormalised range IFFT, amplitude weight, RCMC phase,
    azimuth filter.  Returns (lines, azimuth FM rate [lines][n])."""
    y = np.array(data, np.complex128, copy=True)
    rows, n = y.shape
    f = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (n,)))
    v = np.ascontiguousarray(np.broadcast_to(np.asarray(v_rel, np.float64), (rows, n)))
    fm = np.zeros((rows, n))
    lib().orc_tops_rows(_as_f64(y), rows, n, n, v, f(slant_ranges_), f(range_freqs), f(f_dc),
                        np.ascontiguousarray(az_freqs, np.float64), f(t_corr), float(ref_range), fm)
    return y, fm


def linear_resample(x, n_out):
    """linear_resample (signal_processing.cpp:32-63), real rows."""
    a = np.ascontiguousarray(np.atleast_2d(x), np.float64)
    out = np.zeros((a.shape[0], n_out))
    lib().orc_linear_resample_many(a, a.shape[1], int(n_out), a.shape[0], out)
    return out.reshape(-1) if np.ndim(x) == 1 else out


def quadratic_resample(x, n_out):
    """quadratic_resample (signal_processing.cpp:109-155), one complex row."""
    a = _c128(x)
    if a.size < 3 and a.size != n_out:
        raise ValueError("Need at least 3 points for quadratic interpolation.")
    out = np.zeros(int(n_out), np.complex128)
    lib().orc_quadratic_resample(_as_f64(a), a.size, int(n_out), _as_f64(out))
    return out


def fm_reference(fm):
    """The FM rate n_ref divides by (image_formation.cpp:252): the reference
    reads az_fm_rate[n_az / 2][n / 2] of the transposed [n][n_az] array, i.e.
    line n / 2, range sample n_az / 2; where that is out of range (as for
    IW bursts, n > 2 n_az) it reads past the vector, and the value taken
    here is the evident intent, line n_az / 2 at range sample n / 2."""
    n_az, n = fm.shape
    if n_az // 2 < n and n // 2 < n_az:
        return float(fm[n // 2, n_az // 2])
    return float(fm[n_az // 2, n // 2])


def tops_time_params(fm, dc, nr, prf, burst_duration, bw, swath):
    """The scalar set-up of azimuth_time_ufr (image_formation.cpp:238-292)."""
    fm = np.ascontiguousarray(fm, np.float64)
    n_az, n = fm.shape
    d = np.zeros(4)
    k = np.zeros(8, np.int32)
    rc = lib().orc_tops_time_params(np.ascontiguousarray(fm[:, 0]), n_az, n, float(dc[n // 2]), fm_reference(fm),
                                    float(nr), float(prf), float(burst_duration), float(bw), int(swath), d, k)
    keys_d = ("bandwidth", "focused_time", "n_ref", "ka_mean")
    keys_k = ("num_tiles", "shape", "lpf_start", "lpf_end", "downsample_shape", "output_shape", "crop_start",
              "crop_end")
    p = dict(zip(keys_d, d.tolist()))
    p.update(zip(keys_k, k.tolist()))
    p["ok"] = rc == 0
    return p, d, k


def az_time_ufr(ac, dc, fm, rate, burst_duration, prf, bw, swath, nr=None):
    """azimuth_time_ufr (image_formation.cpp:238-388): [crop rows][n].  nr
    overrides |rate T / prf| (for closed-form checks)."""
    x = _c128(ac)
    n_az, n = x.shape
    nr = tops_replicas(rate, burst_duration, prf) if nr is None else float(nr)
    p, d, k = tops_time_params(fm, dc, nr, prf, burst_duration, bw, swath)
    if not p["ok"]:
        raise ValueError(f"azimuth time UFR set-up out of range: {p}")
    out = np.zeros((int(k[7] - k[6]), n), np.complex128)
    lib().orc_az_time_ufr(_as_f64(x), n_az, n, n, np.ascontiguousarray(dc, np.float64),
                          np.ascontiguousarray(fm, np.float64), float(rate), d, k, _as_f64(out))
    return out, p


def _lerp(a, b, t):
    """std::lerp as libstdc++ writes it (the interpolation of
    interpolate_vector, state_vectors.cpp:142-153)."""
    if (a <= 0 and b >= 0) or (a >= 0 and b <= 0):
        return t * b + (1 - t) * a
    if t == 1:
        return b
    x = a + t * (b - a)
    if (t > 1) == (b > a):
        return x if b < x else b
    return x if b > x else b


def interpolate_state(times, positions, velocities, time):
    """STATE_VECTORS::interpolate (state_vectors.cpp:127-200): the first /
    last vector outside the span, else a bisection for the bracketing pair
    (a vector at exactly the time gives a == b and t = 0 / 0) and lerp of
    every component.  Returns (position, velocity).  A pivot equal to the
    time ends the bisection with the window's start twice (a == b, so t is
    (time - a) / 0)."""
    times = [float(t) for t in times]
    if time > times[-1]:
        return list(positions[-1]), list(velocities[-1])
    if time < times[0]:
        return list(positions[0]), list(velocities[0])
    lo, hi = 0, len(times)
    while True:
        size = hi - lo
        if size == 1:
            a_i, b_i = lo, hi
            break
        pivot = size // 2
        tp = times[lo + pivot]
        if tp == time:
            a_i = b_i = lo
            break
        if tp > time:
            hi = lo + pivot
        else:
            lo = lo + pivot
    a, b = times[a_i], times[b_i]
    with np.errstate(invalid="ignore", divide="ignore"):
        t = float(np.float64(time - a) / np.float64(b - a))
    pos = [_lerp(float(p), float(q), t) for p, q in zip(positions[a_i], positions[b_i])]
    vel = [_lerp(float(p), float(q), t) for p, q in zip(velocities[a_i], velocities[b_i])]
    return pos, vel
