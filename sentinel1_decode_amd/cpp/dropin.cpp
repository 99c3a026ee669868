// dropin.cpp -- the reference's C++ call surface for the L0 -> SLC path
// (include/s1/*.h), implemented by marshalling std::vector data to the GPU
// and calling the C ABI (include/s1gpu.h).  Host code only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/s1/image_formation.h"
#include "../../include/s1/s1_decoder.h"
#include "../../include/s1/signal_processing.h"

namespace {

// Header field names: the keys of the reference's header maps, in header
// order (structs.h:54-166).
const char* const kPrimaryFields[8] = {"packet_version_number", "packet_type", "secondary_header_flag",
                                       "process_id", "process_category", "sequence_flags",
                                       "packet_sequence_count", "packet_data_length"};
const char* const kSecondaryFields[46] = {
    "coarse_time", "fine_time", "sync_marker", "data_take_id", "ecc_number", "na_1", "test_mode",
    "rx_channel_id", "instrument_configuration_id", "sc_data_word_index", "sc_data_word", "space_packet_count",
    "pri_count", "error_flag", "na_2", "baq_mode", "baq_block_length", "na_3", "range_decimation", "rx_gain",
    "tx_ramp_rate_sign", "tx_ramp_rate_mag", "pulse_start_frequency_sign", "pulse_start_frequency_mag",
    "pulse_length", "na_4", "rank", "pri", "swst", "swl", "ssb_flag", "polarisation", "temperature_compensation",
    "na_5", "elevation_beam_address", "na_6", "azimuth_beam_address", "calibration_mode", "na_7",
    "tx_pulse_number", "signal_type", "na_8", "swap", "swath_number", "num_quadratures", "na_9"};

const char* const kDecodeErrors[6] = {"ok", "BRC value above 4", "read past the end of the user data",
                                      "Huffman code not found", "magnitude code above the BRC flag",
                                      "packet larger than the stated maxima"};

thread_local int t_device = 0;
thread_local int t_burst_prec = S1G_F64;
thread_local s1g_ctx* t_ctx = nullptr;

void check(int rc, const char* what) {
    if (rc == S1G_OK) return;
    std::string msg = std::string(what) + ": " + s1g_last_error(t_ctx);
    if (rc == S1G_EARG || rc == S1G_EUNSUPPORTED) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

void hipcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device buffer owned for the duration of one call.
struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) {
        if (bytes) hipcheck(hipMalloc(&p, bytes), "hipMalloc");
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    template <typename T>
    T* as() const {
        return (T*)p;
    }
};

void upload(void* dst, const void* src, size_t bytes) {
    if (bytes) hipcheck(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
}
void download(void* dst, const void* src, size_t bytes) {
    if (bytes) hipcheck(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
}

// Rows (equal length n or at least n) <-> a dense [rows][ld] device array.
void upload_rows(const CF_VEC_2D& rows, int64_t ld, cdouble* dev) {
    for (size_t r = 0; r < rows.size(); ++r)
        upload(dev + r * ld, rows[r].data(), sizeof(cdouble) * std::min<size_t>(rows[r].size(), (size_t)ld));
}
void download_rows(CF_VEC_2D& rows, int64_t ld, const cdouble* dev) {
    for (size_t r = 0; r < rows.size(); ++r)
        download(rows[r].data(), dev + r * ld, sizeof(cdouble) * std::min<size_t>(rows[r].size(), (size_t)ld));
}

s1g_ctx* ctx() {
    if (!t_ctx) {
        hipcheck(hipSetDevice(t_device), "hipSetDevice");
        const int rc = s1g_create(&t_ctx, t_device);
        if (rc != S1G_OK) throw std::runtime_error("s1g_create failed (no usable GPU?)");
        // calls synchronise on return, so the null stream of this thread suffices
        (void)s1g_set_stream(t_ctx, nullptr);
    }
    return t_ctx;
}

// Throws runtime_error naming the first packet whose status[] entry (device
// array, one per packet) is not S1G_PKT_OK; waits for the stream first.
void throw_failed_packets(const PACKET_VEC_1D& packets, const int32_t* d_status) {
    int first = -1;
    const int rc = s1g_check_status(ctx(), d_status, (int)packets.size(), &first);
    if (rc == S1G_EDECODE && first >= 0) {
        std::vector<int32_t> st(packets.size());
        download(st.data(), d_status, sizeof(int32_t) * st.size());
        throw std::runtime_error("packet " + std::to_string(packets[(size_t)first].get_packet_index()) + ": " +
                                 kDecodeErrors[std::min(std::max(st[(size_t)first], 0), 5)]);
    }
    check(rc, "decode status");
}

// Rows x cols DFT along `axis` on a dense device array, FP64, in place.
void dft_dev(cdouble* d, int rows, int cols, int64_t ld, int axis, bool inverse) {
    check(s1g_dft(ctx(), S1G_F64, rows, cols, d, ld, axis, inverse ? 1 : 0), "dft");
}

// inverse DFT of the leading fft_size samples of rows that are longer: the
// reference scales the whole row by 1/fft_size (signal_processing.cpp:574-577).
void scale_tail(cdouble* d, int rows, int64_t ld, int from, int len, double s) {
    if (len <= from) return;
    std::vector<cdouble> v((size_t)(len - from), cdouble(s, 0.0));
    DevBuf dv(sizeof(cdouble) * v.size());
    upload(dv.p, v.data(), sizeof(cdouble) * v.size());
    check(s1g_multiply(ctx(), S1G_F64, rows, len - from, d + from, ld, dv.p, 0), "multiply");
}

}  // namespace

namespace s1 {
void set_device(int device) {
    if (t_ctx && device != t_device) {
        s1g_destroy(t_ctx);
        t_ctx = nullptr;
    }
    t_device = device;
}
void set_burst_precision(int prec) {
    if (prec != S1G_F32 && prec != S1G_F64) throw std::invalid_argument("precision must be S1G_F32 or S1G_F64");
    t_burst_prec = prec;
}
s1g_ctx* context() { return ctx(); }
}  // namespace s1

// ------------------------------------------------------------------ L0Packet
L0Packet::L0Packet(std::unordered_map<std::string, int> primary_header,
                   std::unordered_map<std::string, int> secondary_header, UINT8_VEC_1D raw_user_data,
                   int packet_index) {
    int64_t prim[8], sec[46];
    for (int i = 0; i < 8; ++i) prim[i] = primary_header.at(kPrimaryFields[i]);
    for (int i = 0; i < 46; ++i) sec[i] = secondary_header.at(kSecondaryFields[i]);
    if (s1g_packet_info_from_fields(prim, sec, &_info) != S1G_OK)
        throw std::runtime_error("unknown BAQ mode in the secondary header");
    _raw_user_data = std::move(raw_user_data);
    _packet_index = packet_index;
    if (_info.user_data_length != (int)_raw_user_data.size())
        throw std::runtime_error("The lenght of the user data field is invalid.");
    _is_empty = false;
}

L0Packet L0Packet::from_bytes(const UINT8_VEC_1D& packet, int packet_index) {
    L0Packet p;
    if (packet.size() < 68 || s1g_parse_packet_header(packet.data(), (int64_t)packet.size(), &p._info) != S1G_OK)
        throw std::runtime_error("not a Sentinel-1 space packet");
    const size_t ud = (size_t)std::max(p._info.user_data_length, 0);
    if (packet.size() < 68 + ud) throw std::runtime_error("The lenght of the user data field is invalid.");
    p._raw_user_data.assign(packet.begin() + 68, packet.begin() + 68 + ud);
    p._packet_index = packet_index;
    p._is_empty = false;
    return p;
}

int L0Packet::primary_header(const std::string& key) const {
    for (int i = 0; i < 8; ++i)
        if (key == kPrimaryFields[i]) return (int)_info.primary[i];
    throw std::out_of_range("unknown primary header field " + key);
}

int L0Packet::secondary_header(const std::string& key) const {
    for (int i = 0; i < 46; ++i)
        if (key == kSecondaryFields[i]) return (int)_info.secondary[i];
    throw std::out_of_range("unknown secondary header field " + key);
}

D_VEC_1D L0Packet::get_slant_ranges(int num_ranges) const {
    D_VEC_1D out(num_ranges > 0 ? num_ranges : 2 * _info.num_quads);
    if (out.empty()) return out;
    check(s1g_slant_ranges(&_info, (int)out.size(), out.data()), "slant ranges");
    return out;
}

D_VEC_1D L0Packet::get_timing_corrections() const {
    D_VEC_1D out(2 * _info.num_quads);
    if (out.empty()) return out;
    check(s1g_timing_corrections(&_info, out.data()), "timing corrections");
    return out;
}

CF_VEC_1D L0Packet::get_replica_chirp() const {
    const int nc = _info.replica_chirp_length;
    CF_VEC_1D out((size_t)(2 * _info.num_quads + 2 * nc));
    DevBuf d(sizeof(cdouble) * out.size());
    check(s1g_replica_chirp(ctx(), S1G_F64, 2 * _info.num_quads, _info.txpsf_mhz, _info.txprr_mhz_us,
                            _info.pulse_length_us, _info.fs_mhz, d.p),
          "replica chirp");
    download(out.data(), d.p, sizeof(cdouble) * out.size());
    return out;
}

void L0Packet::set_signal(CF_VEC_1D signal) {
    _signal = std::move(signal);
    _signal_set = true;
}

CF_VEC_1D L0Packet::get_signal() {
    if (!_signal_set) {
        PACKET_VEC_1D one{*this};
        decode_packets_in_place(one);
        set_signal(one[0]._signal);
    }
    return _signal;
}

void decode_packets_in_place(PACKET_VEC_1D& packets) {
    const int n = (int)packets.size();
    if (n == 0) return;
    std::vector<uint64_t> offs(n);
    std::vector<uint32_t> lens(n);
    std::vector<int32_t> nq(n);
    size_t total = 0;
    int max_bytes = 0, max_quads = 0;
    for (int i = 0; i < n; ++i) {
        if (packets[i].get_data_format() != 'D')
            throw std::runtime_error("only FDBAQ (data format D) packets are decoded on this path");
        offs[i] = total;
        lens[i] = (uint32_t)packets[i].get_user_data().size();
        nq[i] = packets[i].get_num_quads();
        total += lens[i];
        max_bytes = std::max(max_bytes, (int)lens[i]);
        max_quads = std::max(max_quads, nq[i]);
    }
    std::vector<uint8_t> blob(total);
    for (int i = 0; i < n; ++i) memcpy(blob.data() + offs[i], packets[i].get_user_data().data(), lens[i]);
    const int64_t ld = std::max(2 * max_quads, 1);
    DevBuf d_blob(total), d_offs(8 * (size_t)n), d_lens(4 * (size_t)n), d_nq(4 * (size_t)n), d_st(4 * (size_t)n);
    DevBuf d_out(sizeof(cdouble) * (size_t)n * ld);
    upload(d_blob.p, blob.data(), total);
    upload(d_offs.p, offs.data(), 8 * (size_t)n);
    upload(d_lens.p, lens.data(), 4 * (size_t)n);
    upload(d_nq.p, nq.data(), 4 * (size_t)n);
    const int rc = s1g_decode_fdbaq(ctx(), d_blob.as<uint8_t>(), d_offs.as<uint64_t>(), d_lens.as<uint32_t>(),
                                    d_nq.as<int32_t>(), n, max_bytes, max_quads, S1G_F64, d_out.p, ld,
                                    d_st.as<int32_t>());
    check(rc, "decode");
    throw_failed_packets(packets, d_st.as<int32_t>());
    std::vector<cdouble> host((size_t)n * ld);
    download(host.data(), d_out.p, sizeof(cdouble) * host.size());
    for (int i = 0; i < n; ++i)
        packets[i].set_signal(CF_VEC_1D(host.begin() + (size_t)i * ld, host.begin() + (size_t)i * ld + 2 * nq[i]));
}

PACKET_VEC_1D decode_packets(const PACKET_VEC_1D& packets) {
    PACKET_VEC_1D out = packets;
    decode_packets_in_place(out);
    return out;
}

// ------------------------------------------------------- signal processing
D_VEC_1D linspace(const double& start, const double& end, const int& size) {
    D_VEC_1D out((size_t)std::max(size, 0));
    if (size == 1) {
        out[0] = start;
    } else if (size > 1) {
        const double step = (end - start) / (size - 1);
        for (int i = 0; i < size; ++i) out[i] = start + i * step;
    }
    return out;
}

std::complex<double> sinc_interpolate(const CF_VEC_1D& signal, double t, int L) {
    DevBuf ds(sizeof(cdouble) * signal.size()), dt(sizeof(double)), dout(sizeof(cdouble));
    upload(ds.p, signal.data(), sizeof(cdouble) * signal.size());
    upload(dt.p, &t, sizeof(double));
    check(s1g_sinc_interpolate(ctx(), S1G_F64, ds.p, (int)signal.size(), dt.as<double>(), 1, L, dout.p),
          "sinc interpolation");
    cdouble r;
    download(&r, dout.p, sizeof r);
    return r;
}

CF_VEC_2D transpose(const CF_VEC_2D& arr) {
    const int rows = (int)arr.size();
    if (rows == 0) return {};
    const int cols = (int)arr[0].size();
    for (const auto& r : arr)
        if ((int)r.size() != cols) throw std::invalid_argument("transpose needs rows of equal length");
    DevBuf a(sizeof(cdouble) * (size_t)rows * cols), b(sizeof(cdouble) * (size_t)rows * cols);
    upload_rows(arr, cols, a.as<cdouble>());
    check(s1g_transpose(ctx(), S1G_F64, rows, cols, a.p, cols, b.p, rows), "transpose");
    CF_VEC_2D out((size_t)cols, CF_VEC_1D((size_t)rows));
    download_rows(out, rows, b.as<cdouble>());
    return out;
}

void compute_1d_dft_in_place(CF_VEC_1D& signal, int fft_size, const bool& inverse) {
    const int len = (int)signal.size();
    if (fft_size <= 0) fft_size = len;
    if (fft_size > len) throw std::invalid_argument("fft_size exceeds the signal length");
    if (len == 0) return;
    DevBuf d(sizeof(cdouble) * len);
    upload(d.p, signal.data(), sizeof(cdouble) * len);
    dft_dev(d.as<cdouble>(), 1, fft_size, len, 1, inverse);
    if (inverse) scale_tail(d.as<cdouble>(), 1, len, fft_size, len, 1.0 / fft_size);
    download(signal.data(), d.p, sizeof(cdouble) * len);
}

CF_VEC_1D compute_1d_dft(const CF_VEC_1D& signal, int fft_size, const bool& inverse) {
    CF_VEC_1D out = signal;
    compute_1d_dft_in_place(out, fft_size, inverse);
    return out;
}

void _compute_axis_dft(CF_VEC_2D& signals, int& fft_size, const bool& inverse) {
    const int rows = (int)signals.size();
    if (rows == 0) return;
    size_t ld = 0;
    for (const auto& r : signals) ld = std::max(ld, r.size());
    for (const auto& r : signals)
        if ((int)r.size() < fft_size) throw std::invalid_argument("fft_size exceeds a row's length");
    if (fft_size <= 0 || ld == 0) return;
    DevBuf d(sizeof(cdouble) * rows * ld);
    upload_rows(signals, (int64_t)ld, d.as<cdouble>());
    dft_dev(d.as<cdouble>(), rows, fft_size, (int64_t)ld, 1, inverse);
    if (inverse) scale_tail(d.as<cdouble>(), rows, (int64_t)ld, fft_size, (int)ld, 1.0 / fft_size);
    download_rows(signals, (int64_t)ld, d.as<cdouble>());
}

void compute_axis_dft_in_place(CF_VEC_2D& signals, int fft_size, const int& axis, const bool& inverse) {
    if (axis != 0 && axis != 1) throw std::invalid_argument("FFT axis must be 0 (rows) or 1 (cols).");
    if (signals.empty()) return;
    if (fft_size == 0) fft_size = axis ? (int)signals[0].size() : (int)signals.size();
    if (axis == 1) {
        _compute_axis_dft(signals, fft_size, inverse);
        return;
    }
    // along columns: the leading fft_size rows of every column
    const int rows = (int)signals.size(), cols = (int)signals[0].size();
    for (const auto& r : signals)
        if ((int)r.size() != cols) throw std::invalid_argument("axis-0 DFT needs rows of equal length");
    if (fft_size > rows) throw std::invalid_argument("fft_size exceeds the column length");
    DevBuf d(sizeof(cdouble) * (size_t)rows * cols);
    upload_rows(signals, cols, d.as<cdouble>());
    dft_dev(d.as<cdouble>(), fft_size, cols, cols, 0, inverse);
    if (inverse && fft_size < rows)
        scale_tail(d.as<cdouble>() + (size_t)fft_size * cols, rows - fft_size, cols, 0, cols, 1.0 / fft_size);
    download_rows(signals, cols, d.as<cdouble>());
}

CF_VEC_2D compute_axis_dft(CF_VEC_2D& signals, int fft_size, const int& axis, const bool& inverse) {
    CF_VEC_2D out = signals;
    compute_axis_dft_in_place(out, fft_size, axis, inverse);
    return out;
}

void fftshift_in_place(CF_VEC_1D& signal) {
    std::rotate(signal.begin(), signal.end() - (signal.size() / 2), signal.end());
}

void fftshift_in_place(CF_VEC_2D& signals) {
    for (auto& row : signals) fftshift_in_place(row);
}

void eccm(CF_VEC_2D& signals, const int& fft_size, const int& stride, const double& detection_threshold,
          const double& mitigation_threshold, const char&) {
    const int rows = (int)signals.size();
    if (rows == 0) return;
    const int cols = (int)signals[0].size();
    for (const auto& r : signals)
        if ((int)r.size() != cols) throw std::invalid_argument("ECCM needs lines of equal length");
    DevBuf d(sizeof(cdouble) * (size_t)rows * cols);
    upload_rows(signals, cols, d.as<cdouble>());
    s1g_eccm_desc e{rows, cols, fft_size, stride, detection_threshold, mitigation_threshold};
    check(s1g_eccm(ctx(), &e, S1G_F64, d.p, cols, nullptr), "eccm");
    download_rows(signals, cols, d.as<cdouble>());
}

// ---------------------------------------------------------- image formation
CF_VEC_1D pulse_compression(const CF_VEC_1D& signal, const CF_VEC_1D& reference) {
    if (reference.size() > signal.size()) throw std::invalid_argument("reference longer than the signal");
    const int n = (int)signal.size(), m = (int)reference.size();
    DevBuf ds(sizeof(cdouble) * n), dr(sizeof(cdouble) * m);
    upload(ds.p, signal.data(), sizeof(cdouble) * n);
    upload(dr.p, reference.data(), sizeof(cdouble) * m);
    if (m) check(s1g_multiply(ctx(), S1G_F64, 1, m, ds.p, n, dr.p, 0), "multiply");
    CF_VEC_1D out((size_t)n);
    download(out.data(), ds.p, sizeof(cdouble) * n);
    return out;
}

CF_VEC_1D get_reference_function(const CF_VEC_1D& replica_chirp) {
    const int n = (int)replica_chirp.size();
    CF_VEC_1D out((size_t)n);
    if (n == 0) return out;
    DevBuf d(sizeof(cdouble) * n);
    upload(d.p, replica_chirp.data(), sizeof(cdouble) * n);
    check(s1g_reference_function(ctx(), S1G_F64, d.p, n, d.p), "reference function");
    download(out.data(), d.p, sizeof(cdouble) * n);
    return out;
}

// ---------------------------------------------------------------- bursts
namespace {

struct DeviceBurst {
    int L = 0, N = 0, max_bytes = 0, max_quads = 0;
    std::vector<uint64_t> offs;
    std::vector<uint32_t> lens;
    std::vector<int32_t> nq;
    std::vector<uint8_t> blob;
};

DeviceBurst pack(const PACKET_VEC_1D& packets) {
    DeviceBurst b;
    b.L = (int)packets.size();
    if (b.L < 2) throw std::invalid_argument("a burst needs at least 2 packets");
    b.N = 2 * packets[0].get_num_quads();
    size_t total = 0;
    for (const auto& p : packets) {
        if (2 * p.get_num_quads() != b.N) throw std::invalid_argument("packets of a burst differ in num_quads");
        b.offs.push_back(total);
        b.lens.push_back((uint32_t)p.get_user_data().size());
        b.nq.push_back(p.get_num_quads());
        total += p.get_user_data().size();
        b.max_bytes = std::max(b.max_bytes, (int)p.get_user_data().size());
    }
    b.max_quads = b.N / 2;
    b.blob.resize(total);
    for (int i = 0; i < b.L; ++i)
        memcpy(b.blob.data() + b.offs[i], packets[i].get_user_data().data(), b.lens[i]);
    return b;
}

// Runs the chain (decode [+ ECCM] -> range -> [az FFT -> RCMC/MF -> az IFFT])
// and returns the burst as rows.
CF_VEC_2D run_burst(PACKET_VEC_1D& packets, bool azimuth, bool do_azimuth_fft, const D_VEC_1D* v_rel,
                    const D_VEC_1D* f_dc, const S1_ECCM_THRESHOLDS* eccm_thr) {
    const DeviceBurst b = pack(packets);
    const int L = b.L, N = b.N;
    const int prec = azimuth ? t_burst_prec : S1G_F64;
    const size_t es = prec == S1G_F32 ? 8 : 16;
    const s1g_packet_info& in = packets[0].info();
    s1g_ctx* c = ctx();
    DevBuf d_blob(b.blob.size()), d_offs(8 * (size_t)L), d_lens(4 * (size_t)L), d_nq(4 * (size_t)L),
        d_st(4 * (size_t)L);
    upload(d_blob.p, b.blob.data(), b.blob.size());
    upload(d_offs.p, b.offs.data(), 8 * (size_t)L);
    upload(d_lens.p, b.lens.data(), 4 * (size_t)L);
    upload(d_nq.p, b.nq.data(), 4 * (size_t)L);
    DevBuf d_out(es * (size_t)L * N);
    s1g_range_desc rd{L, N, in.txpsf_mhz, in.txprr_mhz_us, in.pulse_length_us, in.fs_mhz};
    s1g_eccm_desc ed{L, N, 128, 64, eccm_thr ? eccm_thr->detection : 0.0, eccm_thr ? eccm_thr->mitigation : 0.0};
    int rc;
    if (azimuth) {
        const D_VEC_1D r = packets[0].get_slant_ranges();
        const D_VEC_1D tc = packets[0].get_timing_corrections();
        D_VEC_1D v((size_t)N), fd((size_t)N);
        for (int j = 0; j < N; ++j) {
            v[j] = v_rel->size() == 1 ? (*v_rel)[0] : v_rel->at(j);
            fd[j] = f_dc->size() == 1 ? (*f_dc)[0] : f_dc->at(j);
        }
        DevBuf d_r(8 * (size_t)N), d_v(8 * (size_t)N), d_fd(8 * (size_t)N), d_tc(8 * (size_t)N);
        upload(d_r.p, r.data(), 8 * (size_t)N);
        upload(d_v.p, v.data(), 8 * (size_t)N);
        upload(d_fd.p, fd.data(), 8 * (size_t)N);
        upload(d_tc.p, tc.data(), 8 * (size_t)N);
        s1g_burst_desc bd;
        memset(&bd, 0, sizeof bd);
        bd.user_data = d_blob.as<uint8_t>();
        bd.ud_offsets = d_offs.as<uint64_t>();
        bd.ud_lengths = d_lens.as<uint32_t>();
        bd.num_quads = d_nq.as<int32_t>();
        bd.status = d_st.as<int32_t>();
        bd.max_user_data_bytes = b.max_bytes;
        bd.max_num_quads = b.max_quads;
        bd.use_eccm = eccm_thr != nullptr;
        bd.range = rd;
        bd.azimuth.n_lines = L;
        bd.azimuth.n_samples = N;
        bd.azimuth.apply_rcmc = 1;
        bd.azimuth.prf = 1.0 / (in.pri_us * 1e-6);
        bd.azimuth.wavelength = 299792458.0 / 5.405000454334350e9;
        bd.azimuth.range_spacing = (r[N - 1] - r[0]) / (N - 1);
        bd.azimuth.ref_range = r[N / 2];
        bd.azimuth.slant_ranges = d_r.as<double>();
        bd.azimuth.v_rel = d_v.as<double>();
        bd.azimuth.f_dc = d_fd.as<double>();
        bd.azimuth.t_corr = d_tc.as<double>();
        bd.eccm = ed;
        bd.out = d_out.p;
        bd.ld_out = N;
        rc = s1g_focus_burst(c, &bd, prec);
        if (rc == S1G_OK) rc = s1g_synchronize(c);
    } else {
        DevBuf d_dec(es * (size_t)L * N);
        rc = s1g_decode_fdbaq(c, d_blob.as<uint8_t>(), d_offs.as<uint64_t>(), d_lens.as<uint32_t>(),
                              d_nq.as<int32_t>(), L, b.max_bytes, b.max_quads, prec, d_dec.p, N, d_st.as<int32_t>());
        if (rc == S1G_OK && eccm_thr) rc = s1g_eccm(c, &ed, prec, d_dec.p, N, nullptr);
        if (rc == S1G_OK) rc = s1g_range_compress(c, &rd, prec, d_dec.p, N, d_out.p, N);
        if (rc == S1G_OK && do_azimuth_fft) rc = s1g_azimuth_fft(c, prec, L, N, d_out.p, N, 0);
        if (rc == S1G_OK) rc = s1g_synchronize(c);
    }
    check(rc, azimuth ? "azimuth compression" : "range compression");
    // the reference throws out of the decode (packet.cpp:795-798,
    // decoding_utils.cpp:69-72,107-119); the chain reports per packet
    throw_failed_packets(packets, d_st.as<int32_t>());
    CF_VEC_2D out((size_t)L, CF_VEC_1D((size_t)N));
    if (prec == S1G_F64) {
        download_rows(out, N, d_out.as<cdouble>());
    } else {
        std::vector<std::complex<float>> h((size_t)L * N);
        download(h.data(), d_out.p, es * h.size());
        for (int i = 0; i < L; ++i)
            for (int j = 0; j < N; ++j) out[i][j] = cdouble(h[(size_t)i * N + j]);
    }
    return out;
}

}  // namespace

CF_VEC_2D range_compress_burst(PACKET_VEC_1D& packets, const bool& do_azimuth_fft, const S1_ECCM_THRESHOLDS* eccm) {
    return run_burst(packets, false, do_azimuth_fft, nullptr, nullptr, eccm);
}

CF_VEC_2D azimuth_compress_burst(PACKET_VEC_1D& packets, const D_VEC_1D& effective_velocities,
                                 const D_VEC_1D& doppler_centroids, const S1_ECCM_THRESHOLDS* eccm) {
    if (effective_velocities.empty() || doppler_centroids.empty())
        throw std::invalid_argument("effective velocities and Doppler centroids are required");
    return run_burst(packets, true, false, &effective_velocities, &doppler_centroids, eccm);
}
