// test_dropin.cpp -- exercises the C++ drop-in surface (include/s1/*.h) the
// way the reference's own code calls it.  Building blocks are checked
// against direct evaluations of their definitions (long-double DFT, the
// reference's sinc and reference-function formulas); packet decoding and
// burst focusing against oracle vectors in the directory given as argv[1]
// (written by tests/test_gpu_cpp_dropin.py).  Prints PASS/FAIL per check;
// exit status 1 if any check fails.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <string>

#include "s1/image_formation.h"
#include "s1/image_write.h"
#include "s1/s1_decoder.h"
#include "s1/signal_processing.h"
#include "s1/state_vectors.h"

static int failures = 0;

static void report(const std::string& name, bool ok, double err) {
    std::printf("%s %s %.3e\n", ok ? "PASS" : "FAIL", name.c_str(), err);
    std::fflush(stdout);
    if (!ok) ++failures;
}

static double rel(const CF_VEC_1D& a, const CF_VEC_1D& b) {
    if (a.size() != b.size()) return 1e30;
    long double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += std::norm(a[i] - b[i]);
        den += std::norm(b[i]);
    }
    return den > 0 ? (double)std::sqrt(num / den) : (double)std::sqrt(num);
}

static double rel2(const CF_VEC_2D& a, const CF_VEC_2D& b) {
    CF_VEC_1D fa, fb;
    for (auto& r : a) fa.insert(fa.end(), r.begin(), r.end());
    for (auto& r : b) fb.insert(fb.end(), r.begin(), r.end());
    return rel(fa, fb);
}

static CF_VEC_1D direct_dft(const CF_VEC_1D& x, int n, bool inverse) {
    CF_VEC_1D y(x);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < n; ++k) {
        long double re = 0, im = 0;
        for (int j = 0; j < n; ++j) {
            const long long e = ((long long)j * k) % n;
            const long double a = (inverse ? 2 : -2) * pi * (long double)e / n;
            re += x[j].real() * cosl(a) - x[j].imag() * sinl(a);
            im += x[j].real() * sinl(a) + x[j].imag() * cosl(a);
        }
        y[k] = cdouble((double)re, (double)im);
    }
    if (inverse)
        for (auto& v : y) v /= (double)n;
    return y;
}

static CF_VEC_1D random_vec(size_t n, unsigned seed) {
    std::mt19937_64 g(seed);
    std::normal_distribution<double> d;
    CF_VEC_1D v(n);
    for (auto& x : v) x = cdouble(d(g), d(g));
    return v;
}

template <typename T>
static std::vector<T> read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    std::vector<T> v;
    if (!f) return v;
    f.seekg(0, std::ios::end);
    const size_t bytes = (size_t)f.tellg();
    f.seekg(0);
    v.resize(bytes / sizeof(T));
    f.read((char*)v.data(), (std::streamsize)(v.size() * sizeof(T)));
    return v;
}

static void building_blocks() {
    for (int n : {1, 12, 1480, 1867}) {
        const CF_VEC_1D x = random_vec(n, n);
        report("compute_1d_dft n=" + std::to_string(n), rel(compute_1d_dft(x), direct_dft(x, n, false)) < 1e-11,
               rel(compute_1d_dft(x), direct_dft(x, n, false)));
        CF_VEC_1D y = x;
        compute_1d_dft_in_place(y, 0, false);
        compute_1d_dft_in_place(y, 0, true);
        report("dft round trip n=" + std::to_string(n), rel(y, x) < 1e-12, rel(y, x));
    }
    {   // fft_size shorter than the signal: leading samples transformed, the
        // inverse scales the whole vector (signal_processing.cpp:550-578)
        const CF_VEC_1D x = random_vec(100, 7);
        CF_VEC_1D want = direct_dft(CF_VEC_1D(x.begin(), x.begin() + 64), 64, true);
        for (int i = 64; i < 100; ++i) want.push_back(x[i] / 64.0);
        const CF_VEC_1D got = compute_1d_dft(x, 64, true);
        report("compute_1d_dft fft_size < size", rel(got, want) < 1e-12, rel(got, want));
    }
    {   // axis transforms
        const int R = 37, C = 20;
        CF_VEC_2D a(R);
        for (int r = 0; r < R; ++r) a[r] = random_vec(C, 100 + r);
        CF_VEC_2D rows = compute_axis_dft(a, 0, 1, false), want_rows(R);
        for (int r = 0; r < R; ++r) want_rows[r] = direct_dft(a[r], C, false);
        report("compute_axis_dft axis 1", rel2(rows, want_rows) < 1e-12, rel2(rows, want_rows));
        CF_VEC_2D cols = compute_axis_dft(a, 0, 0, true);
        CF_VEC_2D at = transpose(a), want_t(C);
        for (int c = 0; c < C; ++c) want_t[c] = direct_dft(at[c], R, true);
        const CF_VEC_2D want_cols = transpose(want_t);
        report("compute_axis_dft axis 0 inverse", rel2(cols, want_cols) < 1e-12, rel2(cols, want_cols));
        bool tr = true;
        for (int r = 0; r < R; ++r)
            for (int c = 0; c < C; ++c) tr = tr && at[c][r] == a[r][c];
        report("transpose", tr, 0);
        CF_VEC_2D b = a;
        compute_axis_dft_in_place(b, 0, 1, false);
        report("compute_axis_dft_in_place axis 1", rel2(b, want_rows) < 1e-12, rel2(b, want_rows));
    }
    {
        CF_VEC_1D v = {1, 2, 3, 4, 5};
        fftshift_in_place(v);
        report("fftshift odd", v == CF_VEC_1D({4, 5, 1, 2, 3}), 0);
        CF_VEC_1D w = {1, 2, 3, 4};
        fftshift_in_place(w);
        report("fftshift even", w == CF_VEC_1D({3, 4, 1, 2}), 0);
    }
    {   // sinc_interpolate: the reference's formula, evaluated here in double
        const CF_VEC_1D x = random_vec(64, 3);
        const double PI = 3.14159265358979;
        double worst = 0;
        for (double t : {0.0, 5.0, 10.25, 31.5, 62.9, -3.0, 70.2}) {
            for (int L : {8, 3}) {
                const int n = (int)std::floor(t);
                cdouble want = 0.0;
                for (int k = -L; k <= L; ++k) {
                    const int idx = n + k;
                    if (idx < 0 || idx >= 64) continue;
                    const double xx = t - idx;
                    const double w = 0.5 * (1 + std::cos(PI * xx / L));
                    const double s = xx == 0.0 ? 1.0 : std::sin(PI * xx) / (PI * xx);
                    want += x[idx] * s * w;
                }
                const cdouble got = sinc_interpolate(x, t, L);
                worst = std::max(worst, std::abs(got - want) / std::max(1.0, std::abs(want)));
            }
        }
        report("sinc_interpolate", worst < 1e-13, worst);
    }
    {   // reference function of a random "chirp" and pulse compression
        const CF_VEC_1D c = random_vec(301, 11);
        long double e = 0;
        for (auto& v : c) {
            const long double p = std::norm(v);
            e += p * (p / 301.0L);
        }
        CF_VEC_1D want = direct_dft(c, 301, false);
        for (auto& v : want) v = std::conj(v) / (double)e;
        const CF_VEC_1D got = get_reference_function(c);
        report("get_reference_function", rel(got, want) < 1e-11, rel(got, want));
        const CF_VEC_1D s = random_vec(400, 12);
        const CF_VEC_1D pc = pulse_compression(s, got);
        CF_VEC_1D pw = s;
        for (int i = 0; i < 301; ++i) pw[i] = got[i] * s[i];
        report("pulse_compression", rel(pc, pw) < 1e-15, rel(pc, pw));
    }
    {   // ECCM: identity when no line reaches the detection threshold
        CF_VEC_2D lines(6);
        for (int r = 0; r < 6; ++r) lines[r] = random_vec(1000, 50 + r);
        CF_VEC_2D out = lines;
        eccm(out, 128, 64, 1e30, 1e30);
        report("eccm below threshold", rel2(out, lines) == 0.0, rel2(out, lines));
    }
}

static void packets_and_bursts(const std::string& dir) {
    const auto blob = read_file<uint8_t>(dir + "/packets.bin");
    const auto index = read_file<int64_t>(dir + "/index.bin");  // (offset, length) per packet
    const auto dec = read_file<cdouble>(dir + "/decoded.bin");
    const auto rc = read_file<cdouble>(dir + "/range_compressed.bin");
    const auto slc = read_file<cdouble>(dir + "/slc.bin");
    const auto geo = read_file<double>(dir + "/geometry.bin");  // v_rel, f_dc
    if (blob.empty() || index.empty() || geo.size() != 2) {
        report("packet fixtures present", false, 0);
        return;
    }
    const size_t L = index.size() / 2;
    PACKET_VEC_1D packets;
    for (size_t i = 0; i < L; ++i) {
        const UINT8_VEC_1D bytes(blob.begin() + index[2 * i], blob.begin() + index[2 * i] + index[2 * i + 1]);
        packets.push_back(L0Packet::from_bytes(bytes, (int)i));
    }
    const size_t N = 2 * (size_t)packets[0].get_num_quads();
    report("burst shape", dec.size() == L * N, 0);
    {   // the reference constructor from header maps gives the same packet
        std::unordered_map<std::string, int> prim, sec;
        for (auto k : {"packet_version_number", "packet_type", "secondary_header_flag", "process_id",
                       "process_category", "sequence_flags", "packet_sequence_count", "packet_data_length"})
            prim[k] = packets[0].primary_header(k);
        for (auto k : {"coarse_time", "fine_time", "sync_marker", "data_take_id", "ecc_number", "na_1",
                       "test_mode", "rx_channel_id", "instrument_configuration_id", "sc_data_word_index",
                       "sc_data_word", "space_packet_count", "pri_count", "error_flag", "na_2", "baq_mode",
                       "baq_block_length", "na_3", "range_decimation", "rx_gain", "tx_ramp_rate_sign",
                       "tx_ramp_rate_mag", "pulse_start_frequency_sign", "pulse_start_frequency_mag",
                       "pulse_length", "na_4", "rank", "pri", "swst", "swl", "ssb_flag", "polarisation",
                       "temperature_compensation", "na_5", "elevation_beam_address", "na_6",
                       "azimuth_beam_address", "calibration_mode", "na_7", "tx_pulse_number", "signal_type",
                       "na_8", "swap", "swath_number", "num_quadratures", "na_9"})
            sec[k] = packets[0].secondary_header(k);
        L0Packet q(prim, sec, packets[0].get_user_data(), 0);
        const bool same = q.get_num_quads() == packets[0].get_num_quads() &&
                          q.get_pulse_length() == packets[0].get_pulse_length() &&
                          q.get_tx_ramp_rate() == packets[0].get_tx_ramp_rate() &&
                          q.get_start_frequency() == packets[0].get_start_frequency() &&
                          q.get_range_sample_rate() == packets[0].get_range_sample_rate() &&
                          q.get_slant_ranges() == packets[0].get_slant_ranges();
        report("L0Packet from header maps", same, 0);
        const CF_VEC_1D s0(dec.begin(), dec.begin() + N);
        report("L0Packet::get_signal bit-exact", q.get_signal() == s0, rel(q.get_signal(), s0));
    }
    {
        PACKET_VEC_1D decoded = decode_packets(packets);
        size_t mism = 0;
        for (size_t i = 0; i < L; ++i) {
            const CF_VEC_1D s = decoded[i].get_signal();
            for (size_t j = 0; j < N; ++j) {
                const cdouble a = s[j], b = dec[i * N + j];
                if (std::memcmp(&a, &b, sizeof a)) ++mism;
            }
        }
        report("decode_packets bit-exact", mism == 0, (double)mism);
    }
    {
        const CF_VEC_1D chirp = packets[0].get_replica_chirp();
        report("get_replica_chirp length",
               (int)chirp.size() == 2 * packets[0].get_num_quads() + 2 * packets[0].get_replica_chirp_length(), 0);
    }
    {
        CF_VEC_2D r = range_compress_burst(packets);
        CF_VEC_1D flat;
        for (auto& row : r) flat.insert(flat.end(), row.begin(), row.end());
        report("range_compress_burst vs oracle", rel(flat, rc) < 1e-11, rel(flat, rc));
    }
    {
        CF_VEC_2D img = azimuth_compress_burst(packets, {geo[0]}, {geo[1]});
        CF_VEC_1D flat;
        for (auto& row : img) flat.insert(flat.end(), row.begin(), row.end());
        report("azimuth_compress_burst FP64 vs oracle", rel(flat, slc) < 1e-10, rel(flat, slc));
        s1::set_burst_precision(S1G_F32);
        CF_VEC_2D img32 = azimuth_compress_burst(packets, {geo[0]}, {geo[1]});
        s1::set_burst_precision(S1G_F64);
        CF_VEC_1D f32;
        for (auto& row : img32) f32.insert(f32.end(), row.begin(), row.end());
        report("azimuth_compress_burst FP32 vs oracle", rel(f32, slc) < 1e-5, rel(f32, slc));
    }
    {   // a packet whose user data ends early throws like the reference
        UINT8_VEC_1D bytes(blob.begin() + index[0], blob.begin() + index[0] + index[1]);
        L0Packet p = L0Packet::from_bytes(bytes, 0);
        bool threw = false;
        try {
            PACKET_VEC_1D one{p};
            const_cast<UINT8_VEC_1D&>(one[0].get_user_data()).resize(16);
            decode_packets_in_place(one);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        report("truncated user data throws", threw, 0);
    }
    {   // the same inside a burst: the reference's decode throws out of
        // _range_compress (s1_decoder.cpp:462-468), so the burst calls throw
        PACKET_VEC_1D bad = packets;
        const_cast<UINT8_VEC_1D&>(bad[L / 2].get_user_data()).resize(16);
        bool threw_az = false, threw_rc = false;
        try {
            (void)azimuth_compress_burst(bad, {geo[0]}, {geo[1]});
        } catch (const std::runtime_error& e) {
            threw_az = std::string(e.what()).find("packet " + std::to_string(L / 2)) != std::string::npos;
        }
        try {
            (void)range_compress_burst(bad);
        } catch (const std::runtime_error& e) {
            threw_rc = std::string(e.what()).find("packet " + std::to_string(L / 2)) != std::string::npos;
        }
        report("truncated packet in azimuth_compress_burst throws", threw_az, 0);
        report("truncated packet in range_compress_burst throws", threw_rc, 0);
    }
}

static CF_VEC_1D flat(const CF_VEC_2D& a) {
    CF_VEC_1D f;
    for (auto& r : a) f.insert(f.end(), r.begin(), r.end());
    return f;
}

// S1_Decoder over a Level-0 file written by the test driver, against the
// oracle's decode, range-Doppler data and SLC of IW2 burst 1.
static void level0_file(const std::string& dir) {
    const auto counts = read_file<int64_t>(dir + "/l0_counts.bin");
    const auto dec = read_file<cdouble>(dir + "/l0_decoded.bin");
    const auto rd = read_file<cdouble>(dir + "/l0_rd.bin");
    const auto slc = read_file<cdouble>(dir + "/l0_slc.bin");
    const auto vrel = read_file<double>(dir + "/l0_vrel.bin");
    if (counts.size() != 7 || dec.empty()) {
        report("level0 fixtures present", false, 0);
        return;
    }
    S1_Decoder d(dir + "/level0.dat");
    const auto& sc = d.swath_counts();
    report("S1_Decoder swath counts", sc.at("IW1") == counts[0] && sc.at("IW2") == counts[1] &&
                                          sc.at("IW3") == counts[2] && sc.size() == 3, 0);
    report("S1_Decoder bursts per swath", (int64_t)d.bursts("IW1").size() == counts[3] &&
                                              (int64_t)d.bursts("IW2").size() == counts[4] &&
                                              (int64_t)d.bursts("IW3").size() == counts[5], 0);
    report("S1_Decoder state vectors", (int64_t)d.get_state_vectors().size() == counts[6] && counts[6] > 0, 0);
    bool threw = false;
    try {
        (void)d.get_burst("EW1", 0);
    } catch (const std::out_of_range&) {
        threw = true;
    }
    report("S1_Decoder unknown swath throws", threw, 0);
    const CF_VEC_1D b = flat(d.get_burst("IW2", 1));
    size_t mism = b.size() == dec.size() ? 0 : 1;
    for (size_t i = 0; i < b.size() && i < dec.size(); ++i) mism += std::memcmp(&b[i], &dec[i], sizeof(cdouble)) != 0;
    report("S1_Decoder::get_burst bit-exact", mism == 0, (double)mism);
    PACKET_VEC_1D packets = d.bursts("IW2")[1];
    const D_VEC_1D v = d.burst_effective_velocities(packets);
    double verr = v.size() == vrel.size() ? 0 : 1;
    for (size_t j = 0; j < v.size() && j < vrel.size(); ++j) verr = std::max(verr, std::abs(v[j] / vrel[j] - 1));
    report("S1_Decoder effective velocities", verr < 1e-14, verr);
    const CF_VEC_1D r = flat(d.get_range_compressed_burst("IW2", 1, true));
    report("S1_Decoder::get_range_compressed_burst (range-Doppler)", rel(r, rd) < 1e-11, rel(r, rd));
    // the range-Doppler chain of a burst (the reference's non-TOPS branch)
    const CF_VEC_1D a = flat(azimuth_compress_burst(packets, v, {0.0}));
    report("azimuth_compress_burst of S1_Decoder's burst", rel(a, slc) < 1e-10, rel(a, slc));
    // bursts and IW swaths take the TOPS branch
    const auto tops = read_file<cdouble>(dir + "/l0_tops.bin");
    const auto tshape = read_file<int64_t>(dir + "/l0_tops_shape.bin");
    const CF_VEC_2D tb = d.get_azimuth_compressed_burst("IW2", 1);
    const CF_VEC_1D t = flat(tb);
    const bool shape_ok = tshape.size() == 2 && (int64_t)tb.size() == tshape[0] && !tb.empty() &&
                          (int64_t)tb[0].size() == tshape[1];
    report("S1_Decoder::get_azimuth_compressed_burst (TOPS) shape", shape_ok, (double)tb.size());
    // from packets the centroid carries the range lines' ~5e-14 FFT rounding,
    // which flips the last bit of a few D values under the ~2e8 rad filter
    // phase: ~1e-10 end to end (each TOPS stage alone agrees to ~1e-15)
    report("S1_Decoder::get_azimuth_compressed_burst (TOPS) vs oracle", shape_ok && rel(t, tops) < 1e-9,
           rel(t, tops));
    const CF_VEC_2D sw = d.get_azimuth_compressed_swath("IW2");
    const size_t m0 = (size_t)tshape[0];
    bool ok = shape_ok && sw.size() == 2 * (m0 + 20);
    if (ok) {
        const CF_VEC_2D part(sw.begin() + (long)(m0 + 20), sw.begin() + (long)(2 * m0 + 20));
        ok = rel(flat(part), tops) < 1e-9;
    }
    report("S1_Decoder::get_azimuth_compressed_swath (TOPS bursts + 20 zero lines)", ok, (double)sw.size());
}

// The reference's image-formation helpers and FFTW-style plans.
static void geometry_and_plans() {
    const D_VEC_1D pos = {7071137.0, 0.0, 0.0};
    const double a = 6378137.0, H = 7071137.0, vv = 7500.0;
    const D_VEC_1D R = {H - a, 850e3};
    const D_VEC_1D v = get_effective_velocities(pos, vv, R);
    report("get_effective_velocities nadir closed form", std::abs(v[0] / (vv * std::sqrt(a / H)) - 1) < 1e-14,
           std::abs(v[0] / (vv * std::sqrt(a / H)) - 1));
    CF_VEC_1D line = random_vec(256, 21), keep = line;
    const D_VEC_1D vr(256, 7100.0), rr(256, 850e3), fr(256, 0.0), dc(256, 0.0);
    const D_VEC_1D fac = apply_secondary_range_compression(line, vr, rr, fr, dc, 500.0);
    const double lam = 299792458.0 / 5.405000454334350e9;
    const double Dw = std::sqrt(1 - (lam * lam * 500.0 * 500.0) / (4 * 7100.0 * 7100.0));
    report("apply_secondary_range_compression at zero range frequency", rel(line, keep) == 0.0 &&
                                                                             std::abs(fac[0] - Dw) < 1e-16, 0);
    apply_range_cell_migration_correction(line, rr, fr, fac);
    report("apply_range_cell_migration_correction at zero range frequency", rel(line, keep) == 0.0, 0);
    CF_VEC_2D rows = {random_vec(1480, 31), random_vec(1480, 32)};
    const CF_VEC_2D orig = rows;
    std::vector<fftw_plan> fwd = get_fftw_plans(rows);
    std::vector<fftw_plan> inv = get_fftw_plans(rows, FFTW_BACKWARD);
    for (auto& p : fwd) fftw_execute(p);
    const double e1 = rel(rows[1], direct_dft(orig[1], 1480, false));
    for (auto& p : inv) fftw_execute(p);
    CF_VEC_1D back = rows[0];
    for (auto& x : back) x /= 1480.0;  // the plans are unnormalised, like FFTW's
    destroy_fftw_plans(fwd);
    destroy_fftw_plans(inv);
    report("get_fftw_plans forward row", e1 < 1e-11, e1);
    report("get_fftw_plans backward unnormalised", rel(back, orig[0]) < 1e-12, rel(back, orig[0]));
    // display scaling and writers
    CF_VEC_2D img = {random_vec(40, 41), random_vec(40, 42), random_vec(40, 43)};
    double mx = 0;
    for (auto& r : img)
        for (auto& v : r) mx = std::max(mx, std::abs(v));
    const std::vector<float> n = scale(img, "norm");
    double worst = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 40; ++j)
            worst = std::max(worst, std::abs(n[i * 40 + j] - (float)((std::abs(img[i][j]) + 1e-12) / mx)) /
                                        (double)n[i * 40 + j]);
    report("scale norm", worst < 1.2e-7, worst);
    const F_VEC_2D m = magnitude_2d(img);
    report("magnitude_2d", std::abs(m[1][5] - (float)std::norm(img[1][5])) <= 1e-6f * m[1][5], 0);
    bool threw = false;
    try {
        (void)scale(img, "loudness");
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    report("scale unknown mode throws", threw, 0);
    write_cf32(img, "/tmp/s1_dropin_test.cf32");
    const auto cf = read_file<std::complex<float>>("/tmp/s1_dropin_test.cf32");
    report("write_cf32", cf.size() == 120 && cf[41] == std::complex<float>(img[1][1]), 0);
    write_tif(img, "/tmp/s1_dropin_test.tif", "norm_log");
    const auto tif = read_file<uint8_t>("/tmp/s1_dropin_test.tif");
    report("write_tif BigTIFF", tif.size() > 16 + 480 && tif[0] == 'I' && tif[2] == 43, 0);
}

int main(int argc, char** argv) {
    try {
        building_blocks();
        geometry_and_plans();
        if (argc > 1) {
            packets_and_bursts(argv[1]);
            level0_file(argv[1]);
        }
    } catch (const std::exception& e) {
        std::printf("FAIL exception %s\n", e.what());
        return 1;
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "ALL PASSED", failures);
    return failures ? 1 : 0;
}
