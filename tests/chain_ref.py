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


def oracle_tops(burst, fields, state_vectors, eccm=None):
    """The oracle's TOPS branch of S1_Decoder::_azimuth_compress
    (s1_decoder.cpp:511-701) for one burst: `fields` are the burst's header
    rows (oracle.header_fields), `state_vectors` the file's (oracle.state_vectors).
    Returns the focused burst and the intermediate products."""
    S = oracle.SEC
    g = oracle_geometry(burst)
    raw, st = oracle.decode_packets(burst.stream, burst.ud_offsets, burst.ud_lengths, burst.num_quads)
    x = raw if eccm is None else oracle.eccm(raw, *eccm)[0]
    p = g["par"]
    rc = oracle.range_compress(x, p["txpsf"], p["txprr"], p["pulse_length_us"], p["fs_mhz"])
    L, n = rc.shape
    sec = g["sec"]
    prf = g["prf"]
    T = L / prf
    swath = int(fields[0][S["swath_number"]])
    bw = 675.0 if swath == 11 else 600.0
    v0 = state_vectors[0]["velocity"]
    v_norm = float(np.sqrt(v0[0] ** 2.0 + v0[1] ** 2.0 + v0[2] ** 2.0))
    rate = oracle.doppler_centroid_rate(p["pri_us"], int(fields[0][S["azimuth_beam_address"]]),
                                        int(fields[-1][S["azimuth_beam_address"]]), L, v_norm)
    # L0Packet::get_slant_range_times span (packet.cpp:304-318)
    near = float(sec[26]) * (p["pri_us"] * 1e-6) + p["swst_us"] * 1e-6 + (320 / (8 * 37.53472224)) * 1e-6
    far = near + (2.0 * int(sec[44])) / (p["fs_mhz"] * 1e6)
    dc = oracle.doppler_centroid(rc, rate, prf, near, far, 30)
    u = oracle.az_freq_ufr(rc, rate, T, prf, bw)
    rows = u.shape[0]
    t0 = float(fields[0][S["coarse_time"]]) + 1.0 / float(fields[0][S["fine_time"]])
    dt = (T / prf) / (rows / L)
    times = [v["time"] for v in state_vectors]
    P = [v["position"] for v in state_vectors]
    V = [v["velocity"] for v in state_vectors]
    vrel = np.zeros((rows, n))
    for i in range(rows):
        pos, vel = oracle.interpolate_state(times, P, V, t0 + dt * i)
        vn = float(np.sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]))
        vrel[i] = oracle.effective_velocities(np.asarray(pos), vn, g["R"])
    nr = oracle.tops_replicas(rate, T, prf)
    faz = oracle.linspace(-nr * prf / 2, nr * prf / 2, rows)
    fs = p["fs_mhz"] * 1e6
    fr = oracle.linspace(-fs / 2, fs / 2, n)
    u, fm = oracle.tops_rows(u, vrel, g["R"], fr, dc, faz, g["tc"], g["ref_range"])
    u = oracle.azimuth_ifft(u)
    out, params = oracle.az_time_ufr(u, dc, fm, rate, T, prf, bw, swath)
    return dict(slc=out, rc=rc, dc=dc, rate=rate, freq_ufr_lines=rows, params=params, status=st)
