"""ctypes binding of the C ABI in include/s1gpu.h (libs1gpu.so).

Device buffers are torch tensors (torch is used for device memory and the
stream only); every call runs on torch's current stream.  The library must be
built (``python -m sentinel1_decode_amd.build``); a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libs1gpu.so")

S1G_F64 = 0
S1G_F32 = 1
S1G_OK = 0
S1G_BATCH_GRAPH = 1
S1G_BATCH_STAGED = 2
PKT_STATUS = {0: "ok", 1: "brc", 2: "overrun", 3: "huffman", 4: "mcode", 5: "arg"}


class RangeDesc(C.Structure):
    _fields_ = [("n_lines", C.c_int32), ("n_samples", C.c_int32), ("txpsf_mhz", C.c_double),
                ("txprr_mhz_us", C.c_double), ("txpl_us", C.c_double), ("fs_mhz", C.c_double)]


class AzimuthDesc(C.Structure):
    _fields_ = [("n_lines", C.c_int32), ("n_samples", C.c_int32), ("apply_rcmc", C.c_int32),
                ("reserved", C.c_int32), ("prf", C.c_double), ("wavelength", C.c_double),
                ("range_spacing", C.c_double), ("ref_range", C.c_double), ("slant_ranges", C.c_void_p),
                ("v_rel", C.c_void_p), ("f_dc", C.c_void_p), ("t_corr", C.c_void_p)]


class EccmDesc(C.Structure):
    _fields_ = [("n_lines", C.c_int32), ("n_samples", C.c_int32), ("fft_size", C.c_int32), ("stride", C.c_int32),
                ("detection_threshold", C.c_double), ("mitigation_threshold", C.c_double)]


class BurstDesc(C.Structure):
    _fields_ = [("user_data", C.c_void_p), ("ud_offsets", C.c_void_p), ("ud_lengths", C.c_void_p),
                ("num_quads", C.c_void_p), ("status", C.c_void_p), ("max_user_data_bytes", C.c_int32),
                ("max_num_quads", C.c_int32), ("use_eccm", C.c_int32), ("reserved", C.c_int32),
                ("range", RangeDesc), ("azimuth", AzimuthDesc), ("eccm", EccmDesc), ("eccm_flags", C.c_void_p),
                ("out", C.c_void_p), ("ld_out", C.c_int64)]


class StateVector(C.Structure):
    _fields_ = [("time", C.c_double), ("position", C.c_double * 3), ("velocity", C.c_double * 3),
                ("quaternion", C.c_double * 4), ("angular_rate", C.c_double * 3)]


class PacketInfo(C.Structure):
    _fields_ = [("primary", C.c_int64 * 8), ("secondary", C.c_int64 * 46), ("num_quads", C.c_int32),
                ("baq_mode", C.c_int32), ("range_decimation", C.c_int32), ("rank", C.c_int32),
                ("swath_number", C.c_int32), ("user_data_length", C.c_int32),
                ("replica_chirp_length", C.c_int32), ("baq_block_length", C.c_int32), ("data_format", C.c_char),
                ("reserved", C.c_char * 7), ("pri_us", C.c_double), ("pulse_length_us", C.c_double),
                ("swl_us", C.c_double), ("swst_us", C.c_double), ("rx_gain_db", C.c_double),
                ("txprr_mhz_us", C.c_double), ("txpsf_mhz", C.c_double), ("fs_mhz", C.c_double)]


class TimeUfrParams(C.Structure):
    _fields_ = [("bandwidth", C.c_double), ("focused_time", C.c_double), ("n_ref", C.c_double),
                ("ka_mean", C.c_double), ("num_tiles", C.c_int32), ("shape", C.c_int32),
                ("lpf_start", C.c_int32), ("lpf_end", C.c_int32), ("downsample_shape", C.c_int32),
                ("output_shape", C.c_int32), ("crop_start", C.c_int32), ("crop_end", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, restype, argtypes) of every exported entry point
_SIGS = [
    ("s1g_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    ("s1g_destroy", C.c_int, [C.c_void_p]),
    ("s1g_last_error", C.c_char_p, [C.c_void_p]),
    ("s1g_set_stream", C.c_int, [C.c_void_p, C.c_void_p]),
    ("s1g_synchronize", C.c_int, [C.c_void_p]),
    ("s1g_decode_fdbaq", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                   C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p]),
    ("s1g_check_status", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    ("s1g_fdbaq_symbol", C.c_int, [C.c_int, C.c_uint32, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("s1g_parse_packet_header", C.c_int, [C.c_void_p, C.c_int64, C.POINTER(PacketInfo)]),
    ("s1g_slant_ranges", C.c_int, [C.POINTER(PacketInfo), C.c_int, C.c_void_p]),
    ("s1g_timing_corrections", C.c_int, [C.POINTER(PacketInfo), C.c_void_p]),
    ("s1g_range_compress", C.c_int, [C.c_void_p, C.POINTER(RangeDesc), C.c_int, C.c_void_p, C.c_int64, C.c_void_p,
                                     C.c_int64]),
    ("s1g_replica_chirp_length", C.c_int, [C.c_double, C.c_double]),
    ("s1g_azimuth_fft", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_int]),
    ("s1g_rcmc_matched_filter", C.c_int, [C.c_void_p, C.POINTER(AzimuthDesc), C.c_int, C.c_void_p, C.c_int64,
                                          C.c_void_p, C.c_int64]),
    ("s1g_rcmc_matched_filter_noalloc", C.c_int, [C.c_void_p, C.POINTER(AzimuthDesc), C.c_int, C.c_void_p,
                                                  C.c_int64, C.c_void_p, C.c_int64]),
    ("s1g_eccm", C.c_int, [C.c_void_p, C.POINTER(EccmDesc), C.c_int, C.c_void_p, C.c_int64, C.c_void_p]),
    ("s1g_focus_burst", C.c_int, [C.c_void_p, C.POINTER(BurstDesc), C.c_int]),
    ("s1g_focus_batch", C.c_int, [C.c_void_p, C.POINTER(BurstDesc), C.c_int, C.c_int, C.c_int]),
    ("s1g_af_forward", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p,
                                 C.c_void_p, C.POINTER(C.c_double)]),
    ("s1g_af_adjoint", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p,
                                 C.POINTER(C.c_double), C.c_void_p]),
    ("s1g_profile_enable", C.c_int, [C.c_void_p, C.c_int]),
    ("s1g_profile_read", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    ("s1g_dft", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_int, C.c_int]),
    ("s1g_transpose", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p,
                                C.c_int64]),
    ("s1g_multiply", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int]),
    ("s1g_reference_function", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]),
    ("s1g_replica_chirp", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_void_p]),
    ("s1g_sinc_interpolate", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                       C.c_void_p]),
    ("s1g_packet_info_from_fields", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("s1g_index_level0", C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    ("s1g_parse_headers", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p]),
    ("s1g_segment_bursts", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_int)]),
    ("s1g_state_vectors", C.c_int, [C.c_void_p, C.c_int, C.POINTER(StateVector), C.c_int, C.POINTER(C.c_int)]),
    ("s1g_interpolate_state", C.c_int, [C.POINTER(StateVector), C.c_int, C.c_double, C.POINTER(StateVector)]),
    ("s1g_line_states", C.c_int, [C.POINTER(StateVector), C.c_int, C.c_double, C.c_double, C.c_int, C.c_void_p,
                                   C.c_void_p]),
    ("s1g_effective_velocities", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                           C.c_void_p, C.c_int64]),
    ("s1g_secondary_range_compression", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64,
                                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                  C.c_void_p]),
    ("s1g_rcmc_phase", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p,
                                 C.c_void_p, C.c_void_p]),
    ("s1g_plan_create", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    ("s1g_plan_execute", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int]),
    ("s1g_plan_destroy", C.c_int, [C.c_void_p]),
    ("s1g_scale", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p,
                            C.c_int64]),
    ("s1g_to_cf32", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]),
    ("s1g_write_tiff", C.c_int, [C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_int]),
    ("s1g_azimuth_beam_angle", C.c_double, [C.c_int]),
    ("s1g_doppler_centroid_rate", C.c_double, [C.c_double, C.c_int, C.c_int, C.c_int, C.c_double]),
    ("s1g_doppler_centroid", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p]),
    ("s1g_tops_frequency_lines", C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double]),
    ("s1g_azimuth_frequency_ufr", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64,
                                            C.c_double, C.c_double, C.c_double, C.c_double, C.c_void_p,
                                            C.c_int64]),
    ("s1g_tops_filter_lines", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_double, C.c_void_p]),
    ("s1g_azimuth_time_ufr_params", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_double,
                                              C.c_double, C.c_double, C.c_double, C.c_int,
                                              C.POINTER(TimeUfrParams)]),
    ("s1g_azimuth_time_ufr", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p,
                                       C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                       C.c_void_p, C.c_int64]),
    ("s1g_tune_save", C.c_int, [C.c_void_p, C.c_char_p]),
    ("s1g_tune_load", C.c_int, [C.c_void_p, C.c_char_p]),
    ("s1g_set_autotune", C.c_int, [C.c_void_p, C.c_int]),
    ("s1g_force_variant", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("s1g_last_variant", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
]
EXPORTED = [s[0] for s in _SIGS]
STAGES = ["decode", "eccm", "range", "az_fft", "rcmc_mf", "az_ifft", "filter", "autofocus", "dft"]

_LIB = None


def lib():
    """Load libs1gpu.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m sentinel1_decode_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


class S1GError(RuntimeError):
    pass


# ------------------------------------------------------------ host helpers
def parse_packet_header(packet: bytes) -> PacketInfo:
    buf = np.frombuffer(bytes(packet), np.uint8)
    info = PacketInfo()
    rc = lib().s1g_parse_packet_header(buf.ctypes.data, buf.size, C.byref(info))
    if rc:
        raise S1GError(f"s1g_parse_packet_header failed ({rc})")
    return info


def linspace(start, end, size):
    """linspace (signal_processing.cpp:17-28): start + i (end - start) / (size - 1)."""
    if size == 1:
        return np.array([float(start)])
    step = (end - start) / (size - 1)
    return start + np.arange(size, dtype=np.float64) * step


def slant_ranges(info: PacketInfo, num_ranges: int = 0) -> np.ndarray:
    n = num_ranges if num_ranges > 0 else 2 * info.num_quads
    out = np.zeros(n)
    lib().s1g_slant_ranges(C.byref(info), n, out.ctypes.data)
    return out


def timing_corrections(info: PacketInfo) -> np.ndarray:
    out = np.zeros(2 * info.num_quads)
    lib().s1g_timing_corrections(C.byref(info), out.ctypes.data)
    return out


def fdbaq_symbol(brc: int, code10: int):
    ln, m = C.c_int(0), C.c_int(0)
    rc = lib().s1g_fdbaq_symbol(brc, code10, C.byref(ln), C.byref(m))
    if rc:
        raise S1GError("bad symbol query")
    return ln.value, m.value


def index_level0(stream: np.ndarray) -> np.ndarray:
    """Start offsets of the packets of a Level-0 byte stream (host)."""
    stream = np.ascontiguousarray(stream, np.uint8)
    n = C.c_int64(0)
    lib().s1g_index_level0(stream.ctypes.data, stream.size, None, 0, C.byref(n))
    out = np.zeros(n.value, np.uint64)
    if n.value:
        lib().s1g_index_level0(stream.ctypes.data, stream.size, out.ctypes.data, out.size, C.byref(n))
    return out


def segment_bursts(fields: np.ndarray, swath_number: int, cal: bool = False):
    """Per-packet burst index (-1: other swath or dropped) and the burst count."""
    fields = np.ascontiguousarray(fields, np.int64)
    out = np.zeros(fields.shape[0], np.int32)
    nb = C.c_int(0)
    rc = lib().s1g_segment_bursts(fields.ctypes.data, fields.shape[0], swath_number, int(cal), out.ctypes.data,
                                  C.byref(nb))
    if rc:
        raise S1GError(f"s1g_segment_bursts failed ({rc})")
    return out, nb.value


def state_vectors(fields: np.ndarray):
    fields = np.ascontiguousarray(fields, np.int64)
    n = C.c_int(0)
    lib().s1g_state_vectors(fields.ctypes.data, fields.shape[0], None, 0, C.byref(n))
    arr = (StateVector * max(n.value, 1))()
    lib().s1g_state_vectors(fields.ctypes.data, fields.shape[0], arr, n.value, C.byref(n))
    return list(arr)[: n.value]


def interpolate_state(vectors, time: float) -> StateVector:
    arr = (StateVector * len(vectors))(*vectors)
    out = StateVector()
    rc = lib().s1g_interpolate_state(arr, len(vectors), float(time), C.byref(out))
    if rc:
        raise S1GError(f"s1g_interpolate_state failed ({rc})")
    return out


def line_states(vectors, t0: float, dt: float, count: int):
    """Positions [count][3] and speeds [count] of the platform at t0 + dt i."""
    arr = (StateVector * len(vectors))(*vectors)
    pos = np.zeros((count, 3))
    vn = np.zeros(count)
    rc = lib().s1g_line_states(arr, len(vectors), float(t0), float(dt), int(count), pos.ctypes.data, vn.ctypes.data)
    if rc:
        raise S1GError(f"s1g_line_states failed ({rc})")
    return pos, vn


SCALE_MODES = {"norm_log": 0, "norm": 1, "mag": 2, "real": 3, "imag": 4}


def write_tiff(path, img: np.ndarray, flip_rows=True):
    """A float image as an uncompressed BigTIFF (last row first by default,
    as the reference writes it)."""
    img = np.ascontiguousarray(img, np.float32)
    rc = lib().s1g_write_tiff(str(path).encode(), img.ctypes.data, img.shape[0], img.shape[1], int(flip_rows))
    if rc:
        raise S1GError(f"s1g_write_tiff failed ({rc})")


def azimuth_beam_angle(address):
    """L0Packet::get_azimuth_beam_angle (host)."""
    return lib().s1g_azimuth_beam_angle(int(address))


def doppler_centroid_rate(pri_us, first_address, last_address, n_packets, velocity):
    """get_doppler_centroid_rate (host)."""
    return lib().s1g_doppler_centroid_rate(float(pri_us), int(first_address), int(last_address), int(n_packets),
                                           float(velocity))


# ------------------------------------------------------------- context
def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class Context:
    """One s1g context bound to a HIP device; calls follow torch's current stream."""

    def __init__(self, device: int = 0):
        import torch
        self.torch = torch
        self.device = device
        h = C.c_void_p()
        rc = lib().s1g_create(C.byref(h), device)
        if rc:
            raise S1GError(f"s1g_create failed ({rc})")
        self.h = h

    def close(self):
        if self.h:
            lib().s1g_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _bind_stream(self):
        s = self.torch.cuda.current_stream(self.device).cuda_stream
        lib().s1g_set_stream(self.h, C.c_void_p(s))

    def _check(self, rc, what):
        if rc:
            msg = lib().s1g_last_error(self.h)
            raise S1GError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # ------------------------------------------------ general signal processing
    def dft(self, x, axis=1, inverse=False):
        """In-place DFT of a 2-D complex tensor along `axis` (FFTW conventions,
        inverse scaled by 1/n)."""
        self._bind_stream()
        ld = x.stride(0) if x.shape[0] > 1 else x.shape[1]
        self._check(lib().s1g_dft(self.h, _prec_of(x), x.shape[0], x.shape[1], _ptr(x), ld, axis, int(inverse)),
                    "dft")
        return x

    def transpose(self, x):
        import torch
        out = torch.empty((x.shape[1], x.shape[0]), dtype=x.dtype, device=x.device)
        self._bind_stream()
        self._check(lib().s1g_transpose(self.h, _prec_of(x), x.shape[0], x.shape[1], _ptr(x), x.stride(0), _ptr(out),
                                        out.stride(0)), "transpose")
        return out

    def multiply(self, x, vec, conj=False):
        self._bind_stream()
        self._check(lib().s1g_multiply(self.h, _prec_of(x), x.shape[0], vec.shape[0], _ptr(x), x.stride(0),
                                       _ptr(vec), int(conj)), "multiply")
        return x

    def reference_function(self, chirp):
        import torch
        out = torch.empty_like(chirp)
        self._bind_stream()
        self._check(lib().s1g_reference_function(self.h, _prec_of(chirp), _ptr(chirp), chirp.shape[0], _ptr(out)),
                    "reference_function")
        return out

    def replica_chirp(self, n_samples, txpsf, txprr, txpl, fs, dtype=None, device="cuda"):
        import torch
        dtype = dtype or torch.complex128
        nc = int(np.floor(fs * txpl))
        out = torch.empty(n_samples + 2 * nc, dtype=dtype, device=device)
        self._bind_stream()
        self._check(lib().s1g_replica_chirp(self.h, _prec_of(out), n_samples, txpsf, txprr, txpl, fs, _ptr(out)),
                    "replica_chirp")
        return out

    def sinc_interpolate(self, signal, t, half_width=8):
        import torch
        out = torch.empty(t.shape[0], dtype=signal.dtype, device=signal.device)
        self._bind_stream()
        self._check(lib().s1g_sinc_interpolate(self.h, _prec_of(signal), _ptr(signal), signal.shape[0], _ptr(t),
                                               t.shape[0], half_width, _ptr(out)), "sinc_interpolate")
        return out

    # ------------------------------------------------------ Level-0 files
    def parse_headers(self, d_stream, d_offsets, fields=True):
        """Device stream + packet offsets -> (fields [n, 54] or None, user-data
        offsets, lengths, num_quads), all device tensors."""
        torch = self.torch
        n = d_offsets.numel()
        dev = d_stream.device
        f = torch.empty((n, 54), dtype=torch.int64, device=dev) if fields else None
        uo = torch.empty(n, dtype=torch.int64, device=dev)
        ul = torch.empty(n, dtype=torch.int32, device=dev)
        nq = torch.empty(n, dtype=torch.int32, device=dev)
        self._bind_stream()
        self._check(lib().s1g_parse_headers(self.h, _ptr(d_stream), _ptr(d_offsets), n, _ptr(f), _ptr(uo), _ptr(ul),
                                            _ptr(nq)), "parse_headers")
        return f, uo, ul, nq

    def effective_velocities(self, positions, velocity_norms, slant_ranges_):
        torch = self.torch
        out = torch.empty((positions.shape[0], slant_ranges_.numel()), dtype=torch.float64, device=positions.device)
        self._bind_stream()
        self._check(lib().s1g_effective_velocities(self.h, _ptr(positions), _ptr(velocity_norms), positions.shape[0],
                                                   _ptr(slant_ranges_), slant_ranges_.numel(), _ptr(out),
                                                   out.stride(0)), "effective_velocities")
        return out

    def secondary_range_compression(self, x, v_rel, slant_ranges_, range_freqs, f_dc, az_freqs, factors=None):
        self._bind_stream()
        self._check(lib().s1g_secondary_range_compression(self.h, _prec_of(x), x.shape[0], x.shape[1], _ptr(x),
                                                          x.stride(0), _ptr(v_rel), _ptr(slant_ranges_),
                                                          _ptr(range_freqs), _ptr(f_dc), _ptr(az_freqs),
                                                          _ptr(factors)), "secondary_range_compression")
        return x

    def rcmc_phase(self, x, slant_ranges_, range_freqs, factors):
        self._bind_stream()
        self._check(lib().s1g_rcmc_phase(self.h, _prec_of(x), x.shape[0], x.shape[1], _ptr(x), x.stride(0),
                                         _ptr(slant_ranges_), _ptr(range_freqs), _ptr(factors)), "rcmc_phase")
        return x

    # --------------------------------------------------------------- TOPS
    def doppler_centroid(self, rc, rate, prf, t_near, t_far, n_blocks=30):
        """get_doppler_centroid of a range-compressed burst -> float64[n]."""
        out = self.torch.empty(rc.shape[1], dtype=self.torch.float64, device=rc.device)
        self._bind_stream()
        self._check(lib().s1g_doppler_centroid(self.h, _prec_of(rc), rc.shape[0], rc.shape[1], _ptr(rc),
                                               rc.stride(0), float(rate), float(prf), float(t_near), float(t_far),
                                               int(n_blocks), _ptr(out)), "doppler_centroid")
        return out

    def azimuth_frequency_ufr(self, rc, rate, burst_duration, prf, doppler_bandwidth):
        lines = lib().s1g_tops_frequency_lines(rc.shape[0], float(rate), float(burst_duration), float(prf))
        if lines < 1:
            raise S1GError("azimuth_frequency_ufr: no output lines")
        out = self.torch.empty((lines, rc.shape[1]), dtype=rc.dtype, device=rc.device)
        self._bind_stream()
        self._check(lib().s1g_azimuth_frequency_ufr(self.h, _prec_of(rc), rc.shape[0], rc.shape[1], _ptr(rc),
                                                    rc.stride(0), float(rate), float(burst_duration), float(prf),
                                                    float(doppler_bandwidth), _ptr(out), out.stride(0)),
                    "azimuth_frequency_ufr")
        return out

    def tops_filter_lines(self, x, positions, velocity_norms, az_freqs, slant_ranges_, range_freqs, f_dc, t_corr,
                          ref_range):
        """In place on x; returns the FM rate float64[lines][n]."""
        fm = self.torch.empty(x.shape, dtype=self.torch.float64, device=x.device)
        self._bind_stream()
        self._check(lib().s1g_tops_filter_lines(self.h, _prec_of(x), x.shape[0], x.shape[1], _ptr(x), x.stride(0),
                                                _ptr(positions), _ptr(velocity_norms), _ptr(az_freqs),
                                                _ptr(slant_ranges_), _ptr(range_freqs), _ptr(f_dc), _ptr(t_corr),
                                                float(ref_range), _ptr(fm)), "tops_filter_lines")
        return fm

    def azimuth_time_ufr_params(self, f_dc, fm, rate, burst_duration, prf, doppler_bandwidth, swath_number):
        p = TimeUfrParams()
        self._bind_stream()
        self._check(lib().s1g_azimuth_time_ufr_params(self.h, fm.shape[0], fm.shape[1], _ptr(f_dc), _ptr(fm),
                                                      float(rate), float(burst_duration), float(prf),
                                                      float(doppler_bandwidth), int(swath_number), C.byref(p)),
                    "azimuth_time_ufr_params")
        return p.as_dict()

    def azimuth_time_ufr(self, x, f_dc, fm, rate, burst_duration, prf, doppler_bandwidth, swath_number):
        p = self.azimuth_time_ufr_params(f_dc, fm, rate, burst_duration, prf, doppler_bandwidth, swath_number)
        out = self.torch.empty((p["crop_end"] - p["crop_start"], x.shape[1]), dtype=x.dtype, device=x.device)
        self._check(lib().s1g_azimuth_time_ufr(self.h, _prec_of(x), x.shape[0], x.shape[1], _ptr(x), x.stride(0),
                                               _ptr(f_dc), _ptr(fm), float(rate), float(burst_duration), float(prf),
                                               float(doppler_bandwidth), int(swath_number), _ptr(out),
                                               out.stride(0)), "azimuth_time_ufr")
        return out, p

    def scale(self, x, mode):
        """Display scaling of a complex image to float (mode: norm_log, norm,
        mag, real, imag)."""
        torch = self.torch
        out = torch.empty(x.shape, dtype=torch.float32, device=x.device)
        self._bind_stream()
        self._check(lib().s1g_scale(self.h, _prec_of(x), SCALE_MODES[mode], x.shape[0], x.shape[1], _ptr(x),
                                    x.stride(0), _ptr(out), out.stride(0)), "scale")
        return out

    def to_cf32(self, x):
        torch = self.torch
        out = torch.empty(x.shape, dtype=torch.complex64, device=x.device)
        self._bind_stream()
        self._check(lib().s1g_to_cf32(self.h, x.shape[0], x.shape[1], _ptr(x), x.stride(0), _ptr(out),
                                      out.stride(0)), "to_cf32")
        return out

    def plan(self, prec, n, inverse=False):
        """An unnormalised FFT plan of length n (the opaque s1g_plan)."""
        return Plan(self, prec, n, inverse)

    # ------------------------------------------------------------- autotuner
    def set_autotune(self, on=True):
        self._check(lib().s1g_set_autotune(self.h, int(on)), "set_autotune")

    def force_variant(self, stage, variant):
        """Stage name (see STAGES) or index; variant -1 returns to automatic choice."""
        st = STAGES.index(stage) if isinstance(stage, str) else int(stage)
        self._check(lib().s1g_force_variant(self.h, st, int(variant)), "force_variant")

    def tune_save(self, path):
        """Write the context's autotuning choices to `path`."""
        self._check(lib().s1g_tune_save(self.h, os.fsencode(path)), "tune_save")

    def tune_load(self, path):
        """Merge the choices saved in `path` (same kind of device)."""
        self._check(lib().s1g_tune_load(self.h, os.fsencode(path)), "tune_load")

    def last_variant(self, stage):
        """(variant used by the stage's last call, number of variants it had)."""
        st = STAGES.index(stage) if isinstance(stage, str) else int(stage)
        v, n = C.c_int(0), C.c_int(0)
        self._check(lib().s1g_last_variant(self.h, st, C.byref(v), C.byref(n)), "last_variant")
        return v.value, n.value

    def profile(self, on=True):
        self._check(lib().s1g_profile_enable(self.h, int(on)), "profile_enable")

    def profile_read(self, reset=True):
        ms = np.zeros(len(STAGES))
        n = np.zeros(len(STAGES), np.int64)
        self._check(lib().s1g_profile_read(self.h, ms.ctypes.data, n.ctypes.data, int(reset)), "profile_read")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(STAGES)}

    def synchronize(self):
        self._check(lib().s1g_synchronize(self.h), "synchronize")

    def check_status(self, status):
        """Wait for the stream and raise S1GError naming the first failed
        packet of a device status array (s1g_check_status)."""
        first = C.c_int(-1)
        self._bind_stream()
        self._check(lib().s1g_check_status(self.h, _ptr(status), status.numel(), C.byref(first)), "check_status")

    # ---------------------------------------------------------------- decode
    def decode(self, user_data, ud_offsets, ud_lengths, num_quads, max_bytes, max_quads, prec=S1G_F32,
               out=None, status=None):
        """Device tensors in, complex [n_packets, 2*max_quads] + int32 status out."""
        torch = self.torch
        n = ud_offsets.numel()
        dt = torch.complex64 if prec == S1G_F32 else torch.complex128
        if out is None:
            out = torch.empty((n, 2 * max_quads), dtype=dt, device=user_data.device)
        if status is None:
            status = torch.empty(n, dtype=torch.int32, device=user_data.device)
        self._bind_stream()
        rc = lib().s1g_decode_fdbaq(self.h, _ptr(user_data), _ptr(ud_offsets), _ptr(ud_lengths), _ptr(num_quads), n,
                                    int(max_bytes), int(max_quads), prec, _ptr(out), out.stride(0), _ptr(status))
        self._check(rc, "decode")
        return out, status

    # ----------------------------------------------------------------- range
    def range_compress(self, x, desc: RangeDesc, prec=None, out=None):
        torch = self.torch
        prec = _prec_of(x) if prec is None else prec
        if out is None:
            out = torch.empty_like(x)
        self._bind_stream()
        rc = lib().s1g_range_compress(self.h, C.byref(desc), prec, _ptr(x), x.stride(0), _ptr(out), out.stride(0))
        self._check(rc, "range_compress")
        return out

    # --------------------------------------------------------------- azimuth
    def azimuth_fft(self, x, inverse=False):
        self._bind_stream()
        rc = lib().s1g_azimuth_fft(self.h, _prec_of(x), x.shape[0], x.shape[1], _ptr(x), x.stride(0), int(inverse))
        self._check(rc, "azimuth_fft")
        return x

    def rcmc_matched_filter(self, x, desc: AzimuthDesc, out=None):
        torch = self.torch
        if out is None:
            out = torch.empty_like(x)
        self._bind_stream()
        rc = lib().s1g_rcmc_matched_filter(self.h, C.byref(desc), _prec_of(x), _ptr(x), x.stride(0), _ptr(out),
                                           out.stride(0))
        self._check(rc, "rcmc_matched_filter")
        return out

    def eccm(self, x, desc: EccmDesc, flags=None):
        self._bind_stream()
        rc = lib().s1g_eccm(self.h, C.byref(desc), _prec_of(x), _ptr(x), x.stride(0), _ptr(flags))
        self._check(rc, "eccm")
        return x

    # ----------------------------------------------------------------- chain
    def focus_burst(self, desc: BurstDesc, prec):
        self._bind_stream()
        self._check(lib().s1g_focus_burst(self.h, C.byref(desc), prec), "focus_burst")

    def focus_batch(self, descs, prec, graph=False, staged=False):
        """s1g_focus_batch: the bursts' chains one after another (default),
        or with staged one launch per stage over all bursts (no ECCM,
        disjoint outputs); graph captures and replays."""
        arr = (BurstDesc * len(descs))(*descs)
        flags = (S1G_BATCH_GRAPH if graph else 0) | (S1G_BATCH_STAGED if staged else 0)
        self._bind_stream()
        self._check(lib().s1g_focus_batch(self.h, arr, len(descs), prec, flags), "focus_batch")

    # ------------------------------------------------------------- autofocus
    def af_forward(self, rd, phase, img=None):
        e = C.c_double(0)
        self._bind_stream()
        rc = lib().s1g_af_forward(self.h, _prec_of(rd), rd.shape[0], rd.shape[1], _ptr(rd), rd.stride(0),
                                  _ptr(phase), _ptr(img), C.byref(e))
        self._check(rc, "af_forward")
        return e.value

    def af_adjoint(self, rd, phase, grad):
        e = C.c_double(0)
        self._bind_stream()
        rc = lib().s1g_af_adjoint(self.h, _prec_of(rd), rd.shape[0], rd.shape[1], _ptr(rd), rd.stride(0),
                                  _ptr(phase), C.byref(e), _ptr(grad))
        self._check(rc, "af_adjoint")
        return e.value


class Plan:
    """s1g_plan_create / execute / destroy: in-place, unnormalised DFT of
    rows of length n (the handle that stands in for FFTW's per-row plans)."""

    def __init__(self, ctx, prec, n, inverse=False):
        self.ctx = ctx
        self.n = n
        h = C.c_void_p()
        ctx._check(lib().s1g_plan_create(ctx.h, prec, n, int(inverse), C.byref(h)), "plan_create")
        self.h = h

    def execute(self, x):
        self.ctx._bind_stream()
        self.ctx._check(lib().s1g_plan_execute(self.h, _ptr(x), x.stride(0) if x.dim() == 2 else self.n,
                                               x.shape[0] if x.dim() == 2 else 1), "plan_execute")
        return x

    def close(self):
        if self.h:
            lib().s1g_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _prec_of(x):
    import torch
    if x.dtype == torch.complex64:
        return S1G_F32
    if x.dtype == torch.complex128:
        return S1G_F64
    raise S1GError(f"unsupported dtype {x.dtype}")
