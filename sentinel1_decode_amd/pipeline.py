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
    info = capi.parse_packet_header(first_packet)
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
        self.geo = geometry_from_packets(first, len(offs), v_rel, f_dc)
        self.n_lines, self.n_samples = len(offs), self.geo.n_samples
        self.max_bytes = int(lens.max())
        self.max_quads = int(nq.max())
        self.packet_bytes = int(stream.size)
        dev = torch.device(device)
        self.d_stream = torch.from_numpy(stream).to(dev)
        self.d_offs = torch.from_numpy(offs.view(np.int64)).to(dev)
        self.d_lens = torch.from_numpy(lens.view(np.int32)).to(dev)
        self.d_nq = torch.from_numpy(nq).to(dev)
        self.d_status = torch.zeros(len(offs), dtype=torch.int32, device=dev)
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
