"""Burst-level host logic over the C ABI, mirroring S1_Decoder's burst path.

S1_Decoder::_azimuth_compress (s1_decoder.cpp:511-701) derives everything the
chain needs from the burst's first packet: replica-chirp parameters
(packet.cpp:345-380), slant ranges (packet.cpp:284-301), timing corrections
(packet.cpp:322-342), PRF = 1 / PRI, and the reference range
slant_ranges[N/2] (s1_decoder.cpp:614).  Effective velocity and Doppler
centroid come from orbit data and DC estimation in the reference; here they
are inputs (per range sample or scalars).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

from . import capi

SPEED_OF_LIGHT = 299792458.0
CENTER_FREQ = 5.405000454334350e9
WAVELENGTH = SPEED_OF_LIGHT / CENTER_FREQ


@dataclasses.dataclass
class BurstGeometry:
    n_lines: int
    n_samples: int
    info: capi.PacketInfo
    prf: float
    slant_ranges: np.ndarray
    t_corr: np.ndarray
    v_rel: np.ndarray
    f_dc: np.ndarray

    @property
    def ref_range(self):
        return float(self.slant_ranges[self.n_samples // 2])

    @property
    def range_spacing(self):
        r = self.slant_ranges
        return float((r[-1] - r[0]) / (len(r) - 1))

    def range_desc(self) -> capi.RangeDesc:
        i = self.info
        return capi.RangeDesc(self.n_lines, self.n_samples, i.txpsf_mhz, i.txprr_mhz_us, i.pulse_length_us, i.fs_mhz)


def geometry_from_packets(first_packet: bytes, n_lines: int, v_rel=7100.0, f_dc=0.0) -> BurstGeometry:
    return geometry_from_info(capi.parse_packet_header(first_packet), n_lines, v_rel, f_dc)


def geometry_from_info(info, n_lines: int, v_rel=7100.0, f_dc=0.0) -> BurstGeometry:
    n = 2 * info.num_quads
    r = capi.slant_ranges(info)
    tc = capi.timing_corrections(info)
    prf = 1 / (info.pri_us * 1e-6)
    v = np.broadcast_to(np.asarray(v_rel, np.float64), (n,)).copy()
    fd = np.broadcast_to(np.asarray(f_dc, np.float64), (n,)).copy()
    return BurstGeometry(n_lines, n, info, prf, r, tc, v, fd)


class DeviceBurst:
    """Packets of one burst resident on the device, with the descriptors the
    chain needs.  `stream` is the concatenated packet bytes (headers
    included), `ud_offsets`/`ud_lengths` locate each user-data field."""

    def __init__(self, stream, ud_offsets, ud_lengths, num_quads, v_rel=7100.0, f_dc=0.0, device="cuda"):
        import torch
        self.torch = torch
        stream = np.ascontiguousarray(stream, np.uint8)
        offs = np.ascontiguousarray(ud_offsets, np.uint64)
        lens = np.ascontiguousarray(ud_lengths, np.uint32)
        nq = np.ascontiguousarray(num_quads, np.int32)
        first = stream[int(offs[0]) - 68: int(offs[0])].tobytes()
        dev = torch.device(device)
        self._init(geometry_from_packets(first, len(offs), v_rel, f_dc), torch.from_numpy(stream).to(dev),
                   torch.from_numpy(offs.view(np.int64)).to(dev), torch.from_numpy(lens.view(np.int32)).to(dev),
                   torch.from_numpy(nq).to(dev), int(lens.max()), int(nq.max()), int(stream.size))

    @classmethod
    def from_device(cls, geo, d_stream, d_offs, d_lens, d_nq, max_bytes, max_quads, packet_bytes):
        """A burst over packets already on the device (user-data offsets into
        d_stream, lengths and num_quads as s1g_parse_headers writes them)."""
        import torch
        self = cls.__new__(cls)
        self.torch = torch
        self._init(geo, d_stream, d_offs, d_lens, d_nq, max_bytes, max_quads, packet_bytes)
        return self

    def _init(self, geo, d_stream, d_offs, d_lens, d_nq, max_bytes, max_quads, packet_bytes):
        torch = self.torch
        self.geo = geo
        self.n_lines, self.n_samples = int(d_offs.numel()), geo.n_samples
        self.max_bytes, self.max_quads, self.packet_bytes = max_bytes, max_quads, packet_bytes
        dev = d_stream.device
        self.d_stream, self.d_offs, self.d_lens, self.d_nq = d_stream, d_offs, d_lens, d_nq
        self.d_status = torch.zeros(self.n_lines, dtype=torch.int32, device=dev)
        g = self.geo
        self.d_r = torch.from_numpy(g.slant_ranges).to(dev)
        self.d_v = torch.from_numpy(g.v_rel).to(dev)
        self.d_fdc = torch.from_numpy(g.f_dc).to(dev)
        self.d_tc = torch.from_numpy(g.t_corr).to(dev)

    def azimuth_desc(self, apply_rcmc=True) -> capi.AzimuthDesc:
        g = self.geo
        return capi.AzimuthDesc(self.n_lines, self.n_samples, int(apply_rcmc), 0, g.prf, WAVELENGTH,
                                g.range_spacing, g.ref_range, self.d_r.data_ptr(), self.d_v.data_ptr(),
                                self.d_fdc.data_ptr(), self.d_tc.data_ptr())

    def burst_desc(self, out, eccm: capi.EccmDesc | None = None, eccm_flags=None,
                   apply_rcmc=True) -> capi.BurstDesc:
        d = capi.BurstDesc()
        d.user_data = self.d_stream.data_ptr()
        d.ud_offsets = self.d_offs.data_ptr()
        d.ud_lengths = self.d_lens.data_ptr()
        d.num_quads = self.d_nq.data_ptr()
        d.status = self.d_status.data_ptr()
        d.max_user_data_bytes = self.max_bytes
        d.max_num_quads = self.max_quads
        d.use_eccm = int(eccm is not None)
        d.range = self.geo.range_desc()
        d.azimuth = self.azimuth_desc(apply_rcmc)
        if eccm is not None:
            d.eccm = eccm
        d.eccm_flags = eccm_flags.data_ptr() if eccm_flags is not None else None
        d.out = out.data_ptr()
        d.ld_out = out.stride(0)
        return d

    def decode(self, ctx: capi.Context, prec):
        return ctx.decode(self.d_stream, self.d_offs, self.d_lens, self.d_nq, self.max_bytes, self.max_quads, prec,
                          status=self.d_status)

    def focus(self, ctx: capi.Context, prec, out=None, eccm=None, eccm_flags=None, apply_rcmc=True):
        torch = self.torch
        dt = torch.complex64 if prec == capi.S1G_F32 else torch.complex128
        if out is None:
            out = torch.empty((self.n_lines, self.n_samples), dtype=dt, device=self.d_stream.device)
        d = self.burst_desc(out, eccm, eccm_flags, apply_rcmc)
        ctx.focus_burst(d, prec)
        return out


def from_synth(burst, v_rel=None, f_dc=None, device="cuda") -> DeviceBurst:
    spec = burst.spec
    return DeviceBurst(burst.stream, burst.ud_offsets, burst.ud_lengths, burst.num_quads,
                       v_rel=spec.v_ms if v_rel is None else v_rel, f_dc=spec.f_dc if f_dc is None else f_dc,
                       device=device)


# ------------------------------------------------------------ Level-0 files
# Swath number -> name (SWATH_NUM_TO_STRING, structs.h:489-559).
SWATH_NAMES = {0: "S1", 1: "S2", 2: "S3", 3: "S4", 4: "S5_N", 5: "S5_S", 6: "S6", 10: "IW1", 11: "IW2", 12: "IW3",
               20: "EW1", 21: "EW2", 22: "EW3", 23: "EW4", 24: "EW5", 25: "EW1_100MHz_CAL", 26: "EW2_100MHz_CAL",
               27: "EW3_100MHz_CAL", 28: "EW4_100MHz_CAL", 29: "EW5_100MHz_CAL", 30: "WV1", 31: "WV2",
               32: "WV1_100MHz_CAL", 33: "WV2_100MHz_CAL", 36: "S1_100MHz_CAL", 37: "S2_100MHz_CAL",
               38: "S3_100MHz_CAL", 39: "S4_100MHz_CAL", 40: "S5_100MHz_N_CAL", 41: "S5_100MHz_S_CAL",
               42: "S6_100MHz_CAL", 43: "IW1_100MHz_CAL", 44: "IW2_100MHz_CAL", 45: "IW3_100MHz_CAL",
               50: "S1_TXCAL", 51: "S2_TXCAL", 52: "S3_TXCAL", 53: "S4_TXCAL", 54: "S5_N_TXCAL", 55: "S5_S_TXCAL",
               56: "S6_TXCAL", 60: "IW1_TXCAL", 61: "IW2_TXCAL", 62: "IW3_TXCAL", 70: "EW1_TXCAL", 71: "EW2_TXCAL",
               72: "EW3_TXCAL", 73: "EW4_TXCAL", 74: "EW5_TXCAL", 75: "EW1_100MHz_TXCAL", 76: "EW2_100MHz_TXCAL",
               77: "EW3_100MHz_TXCAL", 78: "EW4_100MHz_TXCAL", 79: "EW5_100MHz_TXCAL", 80: "WV1_TXCAL",
               81: "WV2_TXCAL", 82: "WV1_100MHz_TXCAL", 83: "WV2_100MHz_TXCAL", 86: "S1_100MHz_TXCAL",
               87: "S2_100MHz_TXCAL", 88: "S3_100MHz_TXCAL", 89: "S4_100MHz_TXCAL", 90: "S5_100MHz_N_TXCAL",
               91: "S5_100MHz_S_TXCAL", 92: "S6_100MHz_TXCAL", 93: "IW1_100MHz_TXCAL", 94: "IW2_100MHz_TXCAL",
               95: "IW3_100MHz_TXCAL"}
ECHO_SWATH_NUMBERS = {0, 1, 2, 3, 4, 5, 6, 10, 11, 12, 20, 21, 22, 23, 24, 30, 31}
F_SWATH, F_PRI_COUNT, F_COARSE, F_FINE = 8 + 43, 8 + 12, 8, 9


class Level0Product:
    """S1_Decoder (s1_decoder.h:13-77) over a Level-0 byte stream, on one GPU.

    The stream (or the file at `path`) is indexed on the host
    (s1g_index_level0), copied to the device once, and every packet header
    is parsed there (s1g_parse_headers); echo swaths are split into bursts
    and calibration swaths kept whole (s1g_segment_bursts); the state
    vectors come from the packets' ancillary words (s1g_state_vectors).
    Bursts reference the one device copy of the stream."""

    def __init__(self, source, ctx: capi.Context, device="cuda"):
        import torch
        self.torch = torch
        self.ctx = ctx
        if isinstance(source, (str, os.PathLike)):
            stream = np.fromfile(source, np.uint8)
        elif isinstance(source, (bytes, bytearray)):
            stream = np.frombuffer(bytes(source), np.uint8)
        else:
            stream = np.ascontiguousarray(source, np.uint8)
        self.stream = stream
        self.offsets = capi.index_level0(stream)
        dev = torch.device(device)
        self.d_stream = torch.from_numpy(stream).to(dev)
        d_offs = torch.from_numpy(self.offsets.view(np.int64)).to(dev)
        f, self.d_ud_offs, self.d_ud_lens, self.d_nq = ctx.parse_headers(self.d_stream, d_offs)
        self.fields = f.cpu().numpy()
        self.ud_lens = self.d_ud_lens.cpu().numpy()
        self.nq = self.d_nq.cpu().numpy()
        self.swath_counts = {}
        for num, cnt in zip(*np.unique(self.fields[:, F_SWATH], return_counts=True)):
            self.swath_counts[SWATH_NAMES.get(int(num), str(int(num)))] = int(cnt)
        self._bursts = {}
        for num in np.unique(self.fields[:, F_SWATH]):
            num = int(num)
            if num not in SWATH_NAMES:
                continue
            of, nb = capi.segment_bursts(self.fields, num, cal=num not in ECHO_SWATH_NUMBERS)
            self._bursts[SWATH_NAMES[num]] = [np.nonzero(of == b)[0] for b in range(nb)]
        self.state_vectors = capi.state_vectors(self.fields)
        self.eccm = None

    # -- S1_Decoder's bookkeeping (s1_decoder.cpp:13-93)
    def swaths(self):
        return dict(self.swath_counts)

    def n_bursts(self, swath):
        return len(self._validate(swath))

    def _validate(self, swath, burst=None):
        if swath not in self._bursts:
            raise KeyError(f"The requested swath, {swath}, does not exist in the data. "
                           f"The available swaths are: {sorted(self.swath_counts)}")
        b = self._bursts[swath]
        if burst is not None and not 0 <= burst < len(b):
            raise IndexError(f"The requested burst, {burst}, does not exist in the data. "
                             f"There are {len(b)} bursts in swath {swath}.")
        return b

    def apply_eccm(self, detection_threshold, mitigation_threshold):
        self.eccm = (float(detection_threshold), float(mitigation_threshold))

    def packet_indices(self, swath, burst):
        return self._validate(swath, burst)[burst]

    def packet_time(self, k):
        """L0Packet::get_time (packet.cpp:56-63): coarse + 1 / fine."""
        return float(self.fields[k, F_COARSE]) + 1.0 / float(self.fields[k, F_FINE])

    def burst_velocity(self, idx, geo):
        """Effective velocity per range sample of a burst from the orbit: the
        reference evaluates the platform state per Doppler row at
        t0 + i (L / PRF) / PRF (s1_decoder.cpp:537,579-591); the range-Doppler
        kernels take one value per range sample, the state at the burst's
        middle row (i = L/2)."""
        torch = self.torch
        if not self.state_vectors:
            return np.full(geo.n_samples, 7100.0)
        L = len(idx)
        t = self.packet_time(int(idx[0])) + (L / geo.prf) / geo.prf * (L // 2)
        st = capi.interpolate_state(self.state_vectors, t)
        pos = torch.tensor([list(st.position)], dtype=torch.float64, device=self.d_stream.device)
        v = float(np.sqrt(sum(x * x for x in st.velocity)))
        vn = torch.tensor([v], dtype=torch.float64, device=self.d_stream.device)
        R = torch.from_numpy(geo.slant_ranges).to(self.d_stream.device)
        return self.ctx.effective_velocities(pos, vn, R)[0].cpu().numpy()

    def burst(self, swath, burst, v_rel=None, f_dc=0.0) -> DeviceBurst:
        idx = self.packet_indices(swath, burst)
        torch = self.torch
        di = torch.from_numpy(idx.astype(np.int64)).to(self.d_stream.device)
        info = capi.PacketInfo()
        first = int(idx[0])
        row = np.ascontiguousarray(self.fields[first], np.int64)
        rc = capi.lib().s1g_packet_info_from_fields(row[:8].ctypes.data, row[8:].ctypes.data, C.byref(info))
        if rc:
            raise capi.S1GError(f"packet {first}: unknown BAQ mode")
        geo = geometry_from_info(info, len(idx), 7100.0, f_dc)
        geo.v_rel = np.ascontiguousarray(self.burst_velocity(idx, geo) if v_rel is None else
                                         np.broadcast_to(np.asarray(v_rel, np.float64), (geo.n_samples,)))
        return DeviceBurst.from_device(geo, self.d_stream, self.d_ud_offs.index_select(0, di),
                                       self.d_ud_lens.index_select(0, di), self.d_nq.index_select(0, di),
                                       int(self.ud_lens[idx].max()), int(self.nq[idx].max()),
                                       int(self.stream.size))

    # -- the reference's products (s1_decoder.cpp:96-349)
    def get_burst(self, swath, burst, prec=capi.S1G_F64):
        db = self.burst(swath, burst)
        out, st = db.decode(self.ctx, prec)
        self.ctx.check_status(st)
        return out[:, : db.n_samples]

    def _eccm_desc(self, db):
        if self.eccm is None:
            return None
        return capi.EccmDesc(db.n_lines, db.n_samples, 128, 64, self.eccm[0], self.eccm[1])

    def get_range_compressed_burst(self, swath, burst, range_doppler=False, prec=capi.S1G_F64):
        db = self.burst(swath, burst)
        dec, st = db.decode(self.ctx, prec)
        self.ctx.check_status(st)
        x = dec[:, : db.n_samples].contiguous()
        ed = self._eccm_desc(db)
        if ed is not None:
            self.ctx.eccm(x, ed)
        rc = self.ctx.range_compress(x, db.geo.range_desc())
        if range_doppler:
            self.ctx.azimuth_fft(rc)
        return rc

    def get_tops_burst(self, swath, burst, prec=capi.S1G_F64):
        """The TOPS branch of S1_Decoder::_azimuth_compress (s1_decoder.cpp:
        511-701): range compression, the Doppler-centroid rate from the first
        state vector's speed and the burst's beam angles, the Doppler centroid
        over 30 range blocks, azimuth-frequency unfolding, the per-line filter
        with the platform state at each unfolded line's time, azimuth IFFT and
        azimuth-time unfolding.  Returns the focused burst (device)."""
        torch = self.torch
        if not self.state_vectors:
            raise capi.S1GError("TOPS focusing needs the file's state vectors")
        idx = self.packet_indices(swath, burst)
        rc = self.get_range_compressed_burst(swath, burst, prec=prec)
        L, N = rc.shape
        first = int(idx[0])
        info = capi.PacketInfo()
        row = np.ascontiguousarray(self.fields[first], np.int64)
        capi.lib().s1g_packet_info_from_fields(row[:8].ctypes.data, row[8:].ctypes.data, C.byref(info))
        prf = 1 / (info.pri_us * 1e-6)
        T = L / prf
        swath_number = int(info.swath_number)
        bw = 675.0 if swath_number == 11 else 600.0
        v0 = self.state_vectors[0].velocity
        v_norm = float(np.sqrt(v0[0] ** 2.0 + v0[1] ** 2.0 + v0[2] ** 2.0))
        az_col = 8 + 36
        rate = capi.doppler_centroid_rate(info.pri_us, int(self.fields[first, az_col]),
                                          int(self.fields[int(idx[-1]), az_col]), L, v_norm)
        rank = float(info.secondary[26])
        near = rank * (info.pri_us * 1e-6) + info.swst_us * 1e-6 + (320 / (8 * 37.53472224)) * 1e-6
        far = near + (2.0 * info.num_quads) / (info.fs_mhz * 1e6)
        dc = self.ctx.doppler_centroid(rc, rate, prf, near, far, 30)
        u = self.ctx.azimuth_frequency_ufr(rc, rate, T, prf, bw)
        rows = u.shape[0]
        t0 = self.packet_time(first)
        dt = (T / prf) / (rows / L)
        pos, vn = capi.line_states(self.state_vectors, t0, dt, rows)
        geo = geometry_from_info(info, L)
        nr = abs(rate * T / prf)
        dev = self.d_stream.device
        g = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
        faz = capi.linspace(-nr * prf / 2, nr * prf / 2, rows)
        fs = info.fs_mhz * 1e6
        fr = capi.linspace(-fs / 2, fs / 2, N)
        fm = self.ctx.tops_filter_lines(u, g(pos), g(vn), g(faz), g(geo.slant_ranges), g(fr), dc, g(geo.t_corr),
                                        float(geo.slant_ranges[N // 2]))
        self.ctx.dft(u, axis=0, inverse=True)
        out, _ = self.ctx.azimuth_time_ufr(u, dc, fm, rate, T, prf, bw, swath_number)
        return out

    def get_azimuth_compressed_burst(self, swath, burst, prec=capi.S1G_F64):
        """The range-Doppler chain of one burst (the hot path; the C++
        S1_Decoder, like the reference, focuses bursts with the TOPS branch,
        which get_tops_burst gives here)."""
        db = self.burst(swath, burst)
        ed = self._eccm_desc(db)
        out = db.focus(self.ctx, prec, eccm=ed)
        self.ctx.check_status(db.d_status)
        return out

    def get_azimuth_compressed_swath(self, swath, prec=capi.S1G_F64, graph=True):
        """The swath's bursts of at least 1,000 packets focused in one batch,
        stacked with 20 zero lines after each (s1_decoder.cpp:307-349)."""
        torch = self.torch
        bursts = [self.burst(swath, b) for b in range(self.n_bursts(swath))]
        bursts = [db for db in bursts if db.n_lines >= 1000]
        if not bursts:
            return None
        N = max(db.n_samples for db in bursts)
        dt = torch.complex64 if prec == capi.S1G_F32 else torch.complex128
        rows = sum(db.n_lines + 20 for db in bursts)
        out = torch.zeros((rows, N), dtype=dt, device=self.d_stream.device)
        descs, r = [], 0
        for db in bursts:
            descs.append(db.burst_desc(out[r: r + db.n_lines], eccm=self._eccm_desc(db)))
            r += db.n_lines + 20
        self.ctx.focus_batch(descs, prec, graph=graph)
        for db in bursts:
            self.ctx.check_status(db.d_status)
        return out
