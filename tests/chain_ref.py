"""Oracle (CPU FP64) evaluation of the burst chain, stage by stage, from the
packet stream alone: the geometry comes from the oracle's own header parser."""
import numpy as np

import oracle

WAVELENGTH = 299792458.0 / 5.405000454334350e9


def oracle_geometry(burst, v_rel=None, f_dc=None):
    first = burst.stream[int(burst.ud_offsets[0]) - 68: int(burst.ud_offsets[0])].tobytes()
    _, sec = oracle.parse_headers(first)
    par = oracle.packet_params(sec)
    n = 2 * int(sec[44])
    r = oracle.slant_ranges(sec)
    tc = oracle.timing_corrections(sec)
    prf = 1 / (par["pri_us"] * 1e-6)
    v = np.broadcast_to(np.asarray(burst.spec.v_ms if v_rel is None else v_rel, np.float64), (n,)).copy()
    fd = np.broadcast_to(np.asarray(burst.spec.f_dc if f_dc is None else f_dc, np.float64), (n,)).copy()
    return dict(par=par, sec=sec, n=n, R=r, tc=tc, prf=prf, v=v, fdc=fd, ref_range=float(r[n // 2]),
                spacing=float((r[-1] - r[0]) / (n - 1)))


def oracle_chain(burst, apply_rcmc=True, eccm=None, v_rel=None, f_dc=None):
    g = oracle_geometry(burst, v_rel, f_dc)
    raw, st = oracle.decode_packets(burst.stream, burst.ud_offsets, burst.ud_lengths, burst.num_quads)
    out = {"decoded": raw, "status": st}
    x = raw
    if eccm is not None:
        x, flags = oracle.eccm(x, *eccm)
        out["eccm"] = x
        out["eccm_flags"] = flags
    p = g["par"]
    rc = oracle.range_compress(x, p["txpsf"], p["txprr"], p["pulse_length_us"], p["fs_mhz"])
    out["rc"] = rc
    rd = oracle.azimuth_fft(rc)
    out["rd"] = rd
    mf = oracle.rcmc_mf(rd, g["prf"], g["spacing"], g["ref_range"], g["R"], g["v"], g["fdc"], g["tc"], apply_rcmc)
    out["mf"] = mf
    out["slc"] = oracle.azimuth_ifft(mf)
    out["geo"] = g
    return out
