// dropin.cpp -- the reference's C++ call surface for the L0 -> SLC path
// (include/s1/*.h), implemented by marshalling std::vector data to the GPU
// and calling the C ABI (include/s1gpu.h).  Host code only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>

#include "../../include/s1/doppler.h"
#include "../../include/s1/image_formation.h"
#include "../../include/s1/image_write.h"
#include "../../include/s1/s1_decoder.h"
#include "../../include/s1/signal_processing.h"
#include "../../include/s1/state_vectors.h"

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

// Swath number -> name (SWATH_NUM_TO_STRING, structs.h:489-559), and the
// echo swaths among them (ECHO_SWATHS, structs.h:427-432).
const std::map<int, std::string> kSwathNames = {
    {0, "S1"}, {1, "S2"}, {2, "S3"}, {3, "S4"}, {4, "S5_N"}, {5, "S5_S"}, {6, "S6"}, {10, "IW1"}, {11, "IW2"},
    {12, "IW3"}, {20, "EW1"}, {21, "EW2"}, {22, "EW3"}, {23, "EW4"}, {24, "EW5"}, {25, "EW1_100MHz_CAL"},
    {26, "EW2_100MHz_CAL"}, {27, "EW3_100MHz_CAL"}, {28, "EW4_100MHz_CAL"}, {29, "EW5_100MHz_CAL"}, {30, "WV1"},
    {31, "WV2"}, {32, "WV1_100MHz_CAL"}, {33, "WV2_100MHz_CAL"}, {36, "S1_100MHz_CAL"}, {37, "S2_100MHz_CAL"},
    {38, "S3_100MHz_CAL"}, {39, "S4_100MHz_CAL"}, {40, "S5_100MHz_N_CAL"}, {41, "S5_100MHz_S_CAL"},
    {42, "S6_100MHz_CAL"}, {43, "IW1_100MHz_CAL"}, {44, "IW2_100MHz_CAL"}, {45, "IW3_100MHz_CAL"},
    {50, "S1_TXCAL"}, {51, "S2_TXCAL"}, {52, "S3_TXCAL"}, {53, "S4_TXCAL"}, {54, "S5_N_TXCAL"},
    {55, "S5_S_TXCAL"}, {56, "S6_TXCAL"}, {60, "IW1_TXCAL"}, {61, "IW2_TXCAL"}, {62, "IW3_TXCAL"},
    {70, "EW1_TXCAL"}, {71, "EW2_TXCAL"}, {72, "EW3_TXCAL"}, {73, "EW4_TXCAL"}, {74, "EW5_TXCAL"},
    {75, "EW1_100MHz_TXCAL"}, {76, "EW2_100MHz_TXCAL"}, {77, "EW3_100MHz_TXCAL"}, {78, "EW4_100MHz_TXCAL"},
    {79, "EW5_100MHz_TXCAL"}, {80, "WV1_TXCAL"}, {81, "WV2_TXCAL"}, {82, "WV1_100MHz_TXCAL"},
    {83, "WV2_100MHz_TXCAL"}, {86, "S1_100MHz_TXCAL"}, {87, "S2_100MHz_TXCAL"}, {88, "S3_100MHz_TXCAL"},
    {89, "S4_100MHz_TXCAL"}, {90, "S5_100MHz_N_TXCAL"}, {91, "S5_100MHz_S_TXCAL"}, {92, "S6_100MHz_TXCAL"},
    {93, "IW1_100MHz_TXCAL"}, {94, "IW2_100MHz_TXCAL"}, {95, "IW3_100MHz_TXCAL"}};
const std::set<int> kEchoSwaths = {0, 1, 2, 3, 4, 5, 6, 10, 11, 12, 20, 21, 22, 23, 24, 30, 31};
const std::set<int> kStripmapSwaths = {0, 1, 2, 3, 4, 5, 6};
const std::set<int> kIwSwaths = {10, 11, 12};

int swath_number_of(const std::string& name) {
    for (const auto& kv : kSwathNames)
        if (kv.second == name) return kv.first;
    return -1;
}

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

double L0Packet::get_time() const {
    const double coarse = (double)_info.secondary[0], fine = (double)_info.secondary[1];
    return coarse + (1.0 / fine);
}

std::string L0Packet::get_swath() const {
    auto it = kSwathNames.find(_info.swath_number);
    if (it == kSwathNames.end()) throw std::out_of_range("unknown swath number " + std::to_string(_info.swath_number));
    return it->second;
}

char L0Packet::get_rx_polarization() const {
    switch (_info.secondary[7]) {  // rx_channel_id
        case 0: return 'V';
        case 1: return 'H';
        default: throw std::runtime_error("The rx_channel_id is invalid.");
    }
}

char L0Packet::get_tx_polarization() const {
    const int64_t code = _info.secondary[31];  // polarisation
    if (code >= 0 && code <= 3) return 'H';
    if (code >= 4 && code <= 7) return 'V';
    throw std::runtime_error("The polarization code is invalid.");
}

double L0Packet::get_azimuth_beam_angle() const {
    return s1g_azimuth_beam_angle((int)_info.secondary[36]);  // azimuth_beam_address
}

namespace {
// The two-way time span of a packet's range samples (packet.cpp:304-318).
void slant_range_span(const s1g_packet_info& in, double* near, double* far) {
    const double start = in.swst_us * 1e-6, pri = in.pri_us * 1e-6, rank = (double)in.secondary[26];
    const double fs = in.fs_mhz * 1e+6, n_rng = 2.0 * in.num_quads;
    *near = rank * pri + start + (320 / (8 * 37.53472224)) * 1e-6;
    *far = *near + (n_rng / fs);
}
}  // namespace

D_VEC_1D L0Packet::get_slant_range_times(int num_ranges) const {
    if (num_ranges <= 0) num_ranges = 2 * _info.num_quads;
    double near, far;
    slant_range_span(_info, &near, &far);
    return linspace(near, far, num_ranges);
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

namespace {
template <typename V>
std::vector<V> linear_resample_t(const std::vector<V>& arr, int n_out) {
    const int n_in = (int)arr.size();
    if (n_in == n_out) return arr;
    if (n_in < 2 || n_out < 2) throw std::invalid_argument("linear_resample needs at least 2 points");
    std::vector<V> out((size_t)n_out);
    double x = 0.0;
    const double dx_in = 1.0 / (n_in - 1), dx_out = 1.0 / (n_out - 1);
    int index = 0;
    for (int i = 0; i < n_in - 1; i++) {
        const double x0 = i * dx_in, x1 = (i + 1) * dx_in;
        const V y0 = arr[i], y1 = arr[i + 1];
        while (x < x1 && index < n_out) {
            out[index] = y0 + ((x - x0) * (y1 - y0) / (x1 - x0));
            x += dx_out;
            index += 1;
        }
    }
    out[n_out - 1] = arr[n_in - 1];
    return out;
}
}  // namespace

D_VEC_1D linear_resample(const D_VEC_1D& arr, const int& num_output_samples) {
    return linear_resample_t(arr, num_output_samples);
}
CF_VEC_1D linear_resample(const CF_VEC_1D& arr, const int& num_output_samples) {
    return linear_resample_t(arr, num_output_samples);
}

CF_VEC_1D quadratic_resample(const CF_VEC_1D& arr, const int& num_output_samples) {
    const int n_in = (int)arr.size(), n_out = num_output_samples;
    if (n_in == n_out) return arr;
    if (n_in < 3) throw std::invalid_argument("Need at least 3 points for quadratic interpolation.");
    CF_VEC_1D out((size_t)std::max(n_out, 0));
    const double dx_in = 1.0 / (n_in - 1), dx_out = 1.0 / (n_out - 1);
    for (int j = 0; j < n_out; ++j) {
        const double x = j * dx_out;
        const int i = std::min(int(x / dx_in), n_in - 2);
        const int i0 = std::max(i - 1, 0), i1 = i, i2 = std::min(i + 1, n_in - 1);
        const double x0 = i0 * dx_in, x1 = i1 * dx_in, x2 = i2 * dx_in;
        const double l0 = (x - x1) * (x - x2) / ((x0 - x1) * (x0 - x2));
        const double l1 = (x - x0) * (x - x2) / ((x1 - x0) * (x1 - x2));
        const double l2 = (x - x0) * (x - x1) / ((x2 - x0) * (x2 - x1));
        out[j] = l0 * arr[i0] + l1 * arr[i1] + l2 * arr[i2];
    }
    return out;
}

D_VEC_1D polyfit(const D_VEC_1D& x, const D_VEC_1D& y) {
    if (x.size() != y.size() || x.size() < 3) throw std::invalid_argument("polyfit needs 3 or more (x, y) pairs");
    double Sx[5] = {0, 0, 0, 0, 0}, Sy[3] = {0, 0, 0};
    for (size_t i = 0; i < x.size(); ++i) {
        double xi = 1.0;
        for (int j = 0; j <= 4; ++j) {
            if (j <= 2) Sy[j] += xi * y[i];
            Sx[j] += xi;
            xi *= x[i];
        }
    }
    double A[3][3] = {{Sx[0], Sx[1], Sx[2]}, {Sx[1], Sx[2], Sx[3]}, {Sx[2], Sx[3], Sx[4]}};
    double b[3] = {Sy[0], Sy[1], Sy[2]};
    for (int i = 0; i < 3; ++i) {
        for (int k = i + 1; k < 3; ++k) {
            if (std::abs(A[k][i]) > std::abs(A[i][i])) {
                for (int j = 0; j < 3; ++j) std::swap(A[i][j], A[k][j]);
                std::swap(b[i], b[k]);
            }
        }
        for (int k = i + 1; k < 3; ++k) {
            const double f = A[k][i] / A[i][i];
            for (int j = i; j < 3; ++j) A[k][j] -= f * A[i][j];
            b[k] -= f * b[i];
        }
    }
    D_VEC_1D c(3);
    for (int i = 2; i >= 0; --i) {
        c[i] = b[i];
        for (int j = i + 1; j < 3; ++j) c[i] -= A[i][j] * c[j];
        c[i] /= A[i][i];
    }
    return c;
}

double polyval(const D_VEC_1D& p, const double& x) {
    double r = 0.0;
    for (size_t k = p.size(); k-- > 0;) r = r * x + p[k];
    return r;
}

D_VEC_1D fftfreq(int n, double d) {
    D_VEC_1D f((size_t)std::max(n, 0));
    const double val = 1.0 / (n * d);
    const int N = (n - 1) / 2 + 1;
    for (int i = 0; i < N; ++i) f[i] = i * val;
    for (int i = N; i < n; ++i) f[i] = (i - n) * val;
    return f;
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
    if (eccm_thr) (void)packets[0].get_rx_polarization();  // throws as the reference's ECCM call does
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

// --------------------------------------------------------- Level-0 files
namespace {

std::vector<uint8_t> read_file(const std::string& filename) {
    std::ifstream f(filename, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + filename);
    f.seekg(0, std::ios::end);
    const std::streamoff size = f.tellg();
    f.seekg(0);
    std::vector<uint8_t> bytes((size_t)std::max<std::streamoff>(size, 0));
    if (!bytes.empty()) f.read((char*)bytes.data(), (std::streamsize)bytes.size());
    return bytes;
}

// Header fields of packets as rows of S1G_NFIELDS (for the C ABI's
// segmentation and state-vector functions).
std::vector<int64_t> field_rows(const PACKET_VEC_1D& packets) {
    std::vector<int64_t> rows(packets.size() * S1G_NFIELDS);
    for (size_t k = 0; k < packets.size(); ++k) {
        const s1g_packet_info& in = packets[k].info();
        memcpy(&rows[k * S1G_NFIELDS], in.primary, sizeof in.primary);
        memcpy(&rows[k * S1G_NFIELDS + 8], in.secondary, sizeof in.secondary);
    }
    return rows;
}

}  // namespace

PACKET_VEC_1D get_packets(const std::string& filename, int num_packets) {
    const std::vector<uint8_t> bytes = read_file(filename);
    int64_t n = 0;
    check(s1g_index_level0(bytes.data(), (int64_t)bytes.size(), nullptr, 0, &n), "index");
    std::vector<uint64_t> offs((size_t)n);
    check(s1g_index_level0(bytes.data(), (int64_t)bytes.size(), offs.data(), n, &n), "index");
    PACKET_VEC_1D packets;
    const int64_t want = num_packets > 0 ? std::min<int64_t>(num_packets, n) : n;
    packets.reserve((size_t)want);
    for (int64_t k = 0; k < n && (int64_t)packets.size() < want; ++k) {
        const uint64_t o = offs[(size_t)k];
        const size_t len = 7 + (((size_t)bytes[o + 4] << 8) | bytes[o + 5]);
        try {
            packets.push_back(L0Packet::from_bytes(UINT8_VEC_1D(bytes.begin() + o, bytes.begin() + o + len),
                                                   (int)packets.size()));
        } catch (const std::runtime_error& e) {
            // the reference reports the packet and reads on (packet.cpp:909-914)
            std::cout << "Caught a runtime error while decoding packet #" << packets.size() << " - " << e.what()
                      << std::endl;
        }
    }
    return packets;
}

PACKET_VEC_1D get_packets_in_swath(const std::string& filename, const std::string& swath) {
    PACKET_VEC_1D out;
    for (auto& p : get_packets(filename, 0))
        if (p.get_swath() == swath) out.push_back(std::move(p));
    if (out.empty()) throw std::runtime_error("No packets found for swath: " + swath);
    return out;
}

PACKET_VEC_2D get_packets_in_bursts(PACKET_VEC_1D& packets, const std::string& swath, bool get_cal_packets) {
    const int num = swath_number_of(swath);
    PACKET_VEC_2D bursts;
    if (num < 0 || packets.empty()) return bursts;
    const std::vector<int64_t> rows = field_rows(packets);
    std::vector<int32_t> which(packets.size());
    int nb = 0;
    check(s1g_segment_bursts(rows.data(), (int)packets.size(), num, get_cal_packets ? 1 : 0, which.data(), &nb),
          "burst segmentation");
    bursts.resize((size_t)nb);
    for (size_t k = 0; k < packets.size(); ++k)
        if (which[k] >= 0) bursts[(size_t)which[k]].push_back(packets[k]);
    return bursts;
}

PACKET_VEC_2D get_packets_in_bursts(const std::string& filename, const std::string& swath) {
    PACKET_VEC_1D packets = get_packets_in_swath(filename, swath);
    return get_packets_in_bursts(packets, swath, false);
}

// ----------------------------------------------------------- state vectors
namespace {
STATE_VECTOR to_state_vector(const s1g_state_vector& v) {
    STATE_VECTOR s;
    s.time = v.time;
    s.position.assign(v.position, v.position + 3);
    s.velocity.assign(v.velocity, v.velocity + 3);
    s.quaternions.assign(v.quaternion, v.quaternion + 4);
    s.angular_rate.assign(v.angular_rate, v.angular_rate + 3);
    return s;
}
s1g_state_vector from_rows(const STATE_VECTORS& s, size_t i) {
    s1g_state_vector v{};
    v.time = s.times[i];
    for (int d = 0; d < 3; ++d) {
        v.position[d] = s.positions[i][d];
        v.velocity[d] = s.velocities[i][d];
        v.angular_rate[d] = s.angular_rates[i][d];
    }
    for (int d = 0; d < 4; ++d) v.quaternion[d] = s.quaternions[i][d];
    return v;
}
}  // namespace

STATE_VECTORS::STATE_VECTORS(PACKET_VEC_1D& packets) {
    const std::vector<int64_t> rows = field_rows(packets);
    int n = 0;
    check(s1g_state_vectors(rows.data(), (int)packets.size(), nullptr, 0, &n), "state vectors");
    std::vector<s1g_state_vector> v((size_t)n);
    check(s1g_state_vectors(rows.data(), (int)packets.size(), v.data(), n, &n), "state vectors");
    for (const auto& x : v) {
        const STATE_VECTOR s = to_state_vector(x);
        times.push_back(s.time);
        positions.push_back(s.position);
        velocities.push_back(s.velocity);
        quaternions.push_back(s.quaternions);
        angular_rates.push_back(s.angular_rate);
    }
}

STATE_VECTOR STATE_VECTORS::interpolate(const double& time) const {
    if (times.empty()) throw std::runtime_error("no state vectors to interpolate");
    std::vector<s1g_state_vector> v(times.size());
    for (size_t i = 0; i < times.size(); ++i) v[i] = from_rows(*this, i);
    s1g_state_vector out{};
    check(s1g_interpolate_state(v.data(), (int)v.size(), time, &out), "interpolate");
    return to_state_vector(out);
}

// ------------------------------------------------ image formation (rest)
std::pair<PACKET_VEC_2D, int> get_azimuth_blocks(PACKET_VEC_1D& packets) {
    if (packets.empty()) return {PACKET_VEC_2D{}, 0};
    PACKET_VEC_2D blocks;
    PACKET_VEC_1D current;
    int previous = 2 * packets.front().get_num_quads(), max_size = previous;
    for (const L0Packet& p : packets) {
        const int size = 2 * p.get_num_quads();
        if (size != previous && !current.empty()) {
            blocks.emplace_back(std::move(current));
            current.clear();
            previous = size;
        }
        max_size = std::max(max_size, size);
        current.push_back(p);
    }
    if (!current.empty()) blocks.emplace_back(std::move(current));
    return {std::move(blocks), max_size};
}

D_VEC_1D get_effective_velocities(const D_VEC_1D& position, const double& velocity, const D_VEC_1D& slant_ranges) {
    if (position.size() != 3) throw std::invalid_argument("position must have 3 components");
    const int n = (int)slant_ranges.size();
    D_VEC_1D out((size_t)n);
    if (n == 0) return out;
    DevBuf dp(3 * sizeof(double)), dv(sizeof(double)), dr(sizeof(double) * n), dout(sizeof(double) * n);
    upload(dp.p, position.data(), 3 * sizeof(double));
    upload(dv.p, &velocity, sizeof(double));
    upload(dr.p, slant_ranges.data(), sizeof(double) * n);
    check(s1g_effective_velocities(ctx(), dp.as<double>(), dv.as<double>(), 1, dr.as<double>(), n, dout.as<double>(),
                                   n),
          "effective velocities");
    download(out.data(), dout.p, sizeof(double) * n);
    return out;
}

D_VEC_1D apply_secondary_range_compression(CF_VEC_1D& range_line, const D_VEC_1D& effective_velocities,
                                           const D_VEC_1D& slant_ranges, const D_VEC_1D& range_freqs,
                                           const D_VEC_1D& doppler_centroids, const double& az_freq) {
    const size_t n = range_line.size();
    if (effective_velocities.size() < n || slant_ranges.size() < n || range_freqs.size() < n ||
        doppler_centroids.size() < n)
        throw std::out_of_range("geometry vectors shorter than the range line");
    D_VEC_1D factors(n);
    if (!n) return factors;
    DevBuf dx(sizeof(cdouble) * n), dv(8 * n), dr(8 * n), df(8 * n), dc(8 * n), da(8), dfac(8 * n);
    upload(dx.p, range_line.data(), sizeof(cdouble) * n);
    upload(dv.p, effective_velocities.data(), 8 * n);
    upload(dr.p, slant_ranges.data(), 8 * n);
    upload(df.p, range_freqs.data(), 8 * n);
    upload(dc.p, doppler_centroids.data(), 8 * n);
    upload(da.p, &az_freq, 8);
    check(s1g_secondary_range_compression(ctx(), S1G_F64, 1, (int)n, dx.p, (int64_t)n, dv.as<double>(),
                                          dr.as<double>(), df.as<double>(), dc.as<double>(), da.as<double>(),
                                          dfac.as<double>()),
          "secondary range compression");
    download(range_line.data(), dx.p, sizeof(cdouble) * n);
    download(factors.data(), dfac.p, 8 * n);
    return factors;
}

void apply_range_cell_migration_correction(CF_VEC_1D& range_line, const D_VEC_1D& slant_ranges,
                                           const D_VEC_1D& range_freqs, const D_VEC_1D& rcmc_factors) {
    const size_t n = range_line.size();
    if (slant_ranges.size() < n || range_freqs.size() < n || rcmc_factors.size() < n)
        throw std::out_of_range("geometry vectors shorter than the range line");
    if (!n) return;
    DevBuf dx(sizeof(cdouble) * n), dr(8 * n), df(8 * n), dfac(8 * n);
    upload(dx.p, range_line.data(), sizeof(cdouble) * n);
    upload(dr.p, slant_ranges.data(), 8 * n);
    upload(df.p, range_freqs.data(), 8 * n);
    upload(dfac.p, rcmc_factors.data(), 8 * n);
    check(s1g_rcmc_phase(ctx(), S1G_F64, 1, (int)n, dx.p, (int64_t)n, dr.as<double>(), df.as<double>(),
                         dfac.as<double>()),
          "rcmc phase");
    download(range_line.data(), dx.p, sizeof(cdouble) * n);
}

// ------------------------------------------------------------ TOPS (IW)
namespace {
// [rows][cols] host rows of equal length -> dense device array.
int check_rect(const CF_VEC_2D& a, const char* what) {
    if (a.empty()) throw std::invalid_argument(std::string(what) + ": empty burst");
    for (const auto& r : a)
        if (r.size() != a[0].size()) throw std::invalid_argument(std::string(what) + ": rows differ in length");
    return (int)a[0].size();
}
CF_VEC_2D download_burst(const DevBuf& d, int rows, int cols) {
    CF_VEC_2D out((size_t)rows, CF_VEC_1D((size_t)cols));
    download_rows(out, cols, d.as<cdouble>());
    return out;
}
}  // namespace

double get_doppler_centroid_rate(PACKET_VEC_1D& burst_packets, const double& velocity) {
    if (burst_packets.empty()) throw std::invalid_argument("get_doppler_centroid_rate: no packets");
    const int first = (int)burst_packets.front().info().secondary[36];
    const int last = (int)burst_packets.back().info().secondary[36];
    return s1g_doppler_centroid_rate(burst_packets[0].get_pri(), first, last, (int)burst_packets.size(), velocity);
}

D_VEC_1D get_doppler_centroid(CF_VEC_2D& range_compressed, const double& doppler_centroid_rate,
                              const double& /*burst_duration*/, L0Packet& first_packet, const int& num_rng_blocks) {
    const int n = check_rect(range_compressed, "get_doppler_centroid");
    const int L = (int)range_compressed.size();
    double near, far;
    slant_range_span(first_packet.info(), &near, &far);
    const double prf = 1.0 / (first_packet.get_pri() * 1e-6);
    DevBuf d(sizeof(cdouble) * (size_t)L * n), dc(8 * (size_t)n);
    upload_rows(range_compressed, n, d.as<cdouble>());
    check(s1g_doppler_centroid(ctx(), S1G_F64, L, n, d.p, n, doppler_centroid_rate, prf, near, far, num_rng_blocks,
                               dc.as<double>()),
          "Doppler centroid");
    D_VEC_1D out((size_t)n);
    download(out.data(), dc.p, 8 * (size_t)n);
    return out;
}

CF_VEC_1D get_tiled_signal(CF_VEC_1D& signal, const double& num_replicas) {
    const int n = (int)signal.size();
    const int shape = (int)std::floor(num_replicas * n);
    const int offset = (int)std::floor(num_replicas);
    CF_VEC_1D out((size_t)std::max(shape, 0));
    size_t k = 0;
    for (int i = 0; i < offset; ++i)
        for (int j = 0; j < n && k < out.size(); ++j) out[k++] = signal[j];
    if (std::floor(num_replicas) != num_replicas) {
        const int rest = (int)std::floor((num_replicas - std::floor(num_replicas)) * n);
        for (int j = 0; j < rest && k < out.size(); ++j) out[k++] = signal[j];
    }
    return out;
}

CF_VEC_2D azimuth_frequency_ufr(CF_VEC_2D& range_compressed, D_VEC_1D& /*dc_estimates*/,
                                L0Packet& /*initial_packet*/, const double& dc_rate, const double& burst_duration,
                                const double& prf, const double& doppler_bandwidth) {
    const int n = check_rect(range_compressed, "azimuth_frequency_ufr");
    const int L = (int)range_compressed.size();
    const int rows = s1g_tops_frequency_lines(L, dc_rate, burst_duration, prf);
    if (rows < 1) throw std::invalid_argument("azimuth_frequency_ufr: no output lines");
    DevBuf d(sizeof(cdouble) * (size_t)L * n), o(sizeof(cdouble) * (size_t)rows * n);
    upload_rows(range_compressed, n, d.as<cdouble>());
    check(s1g_azimuth_frequency_ufr(ctx(), S1G_F64, L, n, d.p, n, dc_rate, burst_duration, prf, doppler_bandwidth,
                                    o.p, n),
          "azimuth frequency UFR");
    return download_burst(o, rows, n);
}

CF_VEC_2D azimuth_time_ufr(CF_VEC_2D& azimuth_compressed, D_VEC_1D& dc_estimates, D_VEC_2D& az_fm_rate,
                           L0Packet& /*initial_packet*/, const double& dc_rate, const double& burst_duration,
                           const double& prf, const double& doppler_bandwidth, const int& swath_number) {
    const int n = check_rect(azimuth_compressed, "azimuth_time_ufr");
    const int L = (int)azimuth_compressed.size();
    if ((int)dc_estimates.size() < n || (int)az_fm_rate.size() != L)
        throw std::invalid_argument("azimuth_time_ufr: Doppler centroids or FM rates do not match the burst");
    std::vector<double> fm((size_t)L * n);
    for (int i = 0; i < L; ++i) {
        if ((int)az_fm_rate[i].size() < n) throw std::invalid_argument("azimuth_time_ufr: short FM-rate line");
        std::copy(az_fm_rate[i].begin(), az_fm_rate[i].begin() + n, fm.begin() + (size_t)i * n);
    }
    s1g_ctx* c = ctx();
    DevBuf d(sizeof(cdouble) * (size_t)L * n), dfm(8 * fm.size()), ddc(8 * (size_t)n);
    upload_rows(azimuth_compressed, n, d.as<cdouble>());
    upload(dfm.p, fm.data(), 8 * fm.size());
    upload(ddc.p, dc_estimates.data(), 8 * (size_t)n);
    s1g_time_ufr_params p;
    check(s1g_azimuth_time_ufr_params(c, L, n, ddc.as<double>(), dfm.as<double>(), dc_rate, burst_duration, prf,
                                      doppler_bandwidth, swath_number, &p),
          "azimuth time UFR");
    const int rows = p.crop_end - p.crop_start;
    DevBuf o(sizeof(cdouble) * (size_t)rows * n);
    check(s1g_azimuth_time_ufr(c, S1G_F64, L, n, d.p, n, ddc.as<double>(), dfm.as<double>(), dc_rate,
                               burst_duration, prf, doppler_bandwidth, swath_number, o.p, n),
          "azimuth time UFR");
    return download_burst(o, rows, n);
}

CF_VEC_2D tops_compress_burst(PACKET_VEC_1D& packets, const STATE_VECTORS& state_vectors,
                              const S1_ECCM_THRESHOLDS* eccm_thr) {
    if (state_vectors.size() == 0) throw std::runtime_error("TOPS focusing needs the file's state vectors");
    const DeviceBurst b = pack(packets);
    const int L = b.L, N = b.N;
    if (eccm_thr) (void)packets[0].get_rx_polarization();
    const int prec = t_burst_prec;
    const size_t es = prec == S1G_F32 ? 8 : 16;
    const L0Packet& first = packets[0];
    const s1g_packet_info& in = first.info();
    s1g_ctx* c = ctx();
    // decode (+ ECCM) and range compression, kept on the device
    DevBuf d_blob(b.blob.size()), d_offs(8 * (size_t)L), d_lens(4 * (size_t)L), d_nq(4 * (size_t)L),
        d_st(4 * (size_t)L);
    upload(d_blob.p, b.blob.data(), b.blob.size());
    upload(d_offs.p, b.offs.data(), 8 * (size_t)L);
    upload(d_lens.p, b.lens.data(), 4 * (size_t)L);
    upload(d_nq.p, b.nq.data(), 4 * (size_t)L);
    DevBuf d_rc(es * (size_t)L * N);
    {
        DevBuf d_dec(es * (size_t)L * N);
        s1g_range_desc rd{L, N, in.txpsf_mhz, in.txprr_mhz_us, in.pulse_length_us, in.fs_mhz};
        s1g_eccm_desc ed{L, N, 128, 64, eccm_thr ? eccm_thr->detection : 0.0,
                         eccm_thr ? eccm_thr->mitigation : 0.0};
        int rc = s1g_decode_fdbaq(c, d_blob.as<uint8_t>(), d_offs.as<uint64_t>(), d_lens.as<uint32_t>(),
                                  d_nq.as<int32_t>(), L, b.max_bytes, b.max_quads, prec, d_dec.p, N,
                                  d_st.as<int32_t>());
        if (rc == S1G_OK && eccm_thr) rc = s1g_eccm(c, &ed, prec, d_dec.p, N, nullptr);
        if (rc == S1G_OK) rc = s1g_range_compress(c, &rd, prec, d_dec.p, N, d_rc.p, N);
        if (rc == S1G_OK) rc = s1g_synchronize(c);
        check(rc, "range compression");
        throw_failed_packets(packets, d_st.as<int32_t>());
    }
    // burst geometry (s1_decoder.cpp:525-570)
    const D_VEC_1D& v0 = state_vectors.velocities[0];
    const double v_norm = std::sqrt(std::pow(v0[0], 2.0) + std::pow(v0[1], 2.0) + std::pow(v0[2], 2.0));
    const double prf = 1 / (in.pri_us * 1e-6);
    const double T = double(L) / prf;
    const int swath = in.swath_number;
    const double bw = swath == 11 ? 675 : 600;
    const double t0 = first.get_time();
    double time_delta = T / prf;
    const double fs = in.fs_mhz * 1e+6;
    const double rate = get_doppler_centroid_rate(packets, v_norm);
    double near, far;
    slant_range_span(in, &near, &far);
    DevBuf d_dc(8 * (size_t)N);
    check(s1g_doppler_centroid(c, prec, L, N, d_rc.p, N, rate, prf, near, far, 30, d_dc.as<double>()),
          "Doppler centroid");
    const int rows = s1g_tops_frequency_lines(L, rate, T, prf);
    if (rows < 3) throw std::invalid_argument("TOPS focusing: the burst unfolds to fewer than 3 lines");
    DevBuf d_u(es * (size_t)rows * N);
    check(s1g_azimuth_frequency_ufr(c, prec, L, N, d_rc.p, N, rate, T, prf, bw, d_u.p, N), "azimuth frequency UFR");
    // a platform state per unfolded line (s1_decoder.cpp:573-591)
    time_delta /= double(rows) / double(L);
    std::vector<double> pos(3 * (size_t)rows), vn((size_t)rows);
    {
        std::vector<s1g_state_vector> sv(state_vectors.size());
        for (size_t i = 0; i < sv.size(); ++i) sv[i] = from_rows(state_vectors, i);
        check(s1g_line_states(sv.data(), (int)sv.size(), t0, time_delta, rows, pos.data(), vn.data()),
              "line states");
    }
    const double nr = std::abs(rate * T / prf);
    const D_VEC_1D az = linspace(-nr * prf / 2, nr * prf / 2, rows);
    const D_VEC_1D R = first.get_slant_ranges(), tc = first.get_timing_corrections();
    const D_VEC_1D fr = linspace(-fs / 2, fs / 2, N);
    DevBuf d_pos(8 * pos.size()), d_vn(8 * vn.size()), d_az(8 * (size_t)rows), d_R(8 * (size_t)N),
        d_fr(8 * (size_t)N), d_tc(8 * (size_t)N), d_fm(8 * (size_t)rows * N);
    upload(d_pos.p, pos.data(), 8 * pos.size());
    upload(d_vn.p, vn.data(), 8 * vn.size());
    upload(d_az.p, az.data(), 8 * (size_t)rows);
    upload(d_R.p, R.data(), 8 * (size_t)N);
    upload(d_fr.p, fr.data(), 8 * (size_t)N);
    upload(d_tc.p, tc.data(), 8 * (size_t)N);
    check(s1g_tops_filter_lines(c, prec, rows, N, d_u.p, N, d_pos.as<double>(), d_vn.as<double>(), d_az.as<double>(),
                                d_R.as<double>(), d_fr.as<double>(), d_dc.as<double>(), d_tc.as<double>(), R[N / 2],
                                d_fm.as<double>()),
          "TOPS line filter");
    check(s1g_dft(c, prec, rows, N, d_u.p, N, 0, 1), "azimuth IFFT");
    s1g_time_ufr_params p;
    check(s1g_azimuth_time_ufr_params(c, rows, N, d_dc.as<double>(), d_fm.as<double>(), rate, T, prf, bw, swath, &p),
          "azimuth time UFR");
    const int out_rows = p.crop_end - p.crop_start;
    DevBuf d_out(es * (size_t)out_rows * N);
    check(s1g_azimuth_time_ufr(c, prec, rows, N, d_u.p, N, d_dc.as<double>(), d_fm.as<double>(), rate, T, prf, bw,
                               swath, d_out.p, N),
          "azimuth time UFR");
    check(s1g_synchronize(c), "TOPS focusing");
    CF_VEC_2D out((size_t)out_rows, CF_VEC_1D((size_t)N));
    if (prec == S1G_F64) {
        download_rows(out, N, d_out.as<cdouble>());
    } else {
        std::vector<std::complex<float>> h((size_t)out_rows * N);
        download(h.data(), d_out.p, es * h.size());
        for (int i = 0; i < out_rows; ++i)
            for (int j = 0; j < N; ++j) out[i][j] = cdouble(h[(size_t)i * N + j]);
    }
    return out;
}

// ------------------------------------------------------------- FFT plans
namespace s1 {
struct fft_plan_rec {
    s1g_plan* plan;
    CF_VEC_1D* row;
};
}  // namespace s1

std::vector<fftw_plan> get_fftw_plans(CF_VEC_2D& signals, const int& direction) {
    if (direction != FFTW_FORWARD && direction != FFTW_BACKWARD)
        throw std::invalid_argument("direction must be FFTW_FORWARD or FFTW_BACKWARD");
    std::vector<fftw_plan> plans;
    plans.reserve(signals.size());
    try {
        // each handle binds its row; the context caches the tables per length
        for (auto& row : signals) {
            s1g_plan* p = nullptr;
            if (!row.empty())
                check(s1g_plan_create(ctx(), S1G_F64, (int)row.size(), direction == FFTW_BACKWARD ? 1 : 0, &p),
                      "plan");
            plans.push_back(new s1::fft_plan_rec{p, &row});
        }
    } catch (...) {
        destroy_fftw_plans(plans);
        throw;
    }
    return plans;
}

void fftw_execute(const fftw_plan plan) {
    if (!plan || !plan->plan || plan->row->empty()) return;
    CF_VEC_1D& row = *plan->row;
    DevBuf d(sizeof(cdouble) * row.size());
    upload(d.p, row.data(), sizeof(cdouble) * row.size());
    check(s1g_plan_execute(plan->plan, d.p, (int64_t)row.size(), 1), "plan execute");
    download(row.data(), d.p, sizeof(cdouble) * row.size());
}

void destroy_fftw_plans(std::vector<fftw_plan>& plans) {
    for (auto& p : plans) {
        if (p) s1g_plan_destroy(p->plan);
        delete p;
        p = nullptr;
    }
}

// ------------------------------------------------------------ S1_Decoder
S1_Decoder::S1_Decoder(const std::string& filename) : _filename(filename) {
    _flat_packets = get_packets(_filename, 0);
    for (const L0Packet& p : _flat_packets) _swath_counts[p.get_swath()]++;
    for (const auto& kv : _swath_counts) {
        const int num = swath_number_of(kv.first);
        if (kEchoSwaths.count(num))
            _echo_packets[kv.first] = get_packets_in_bursts(_flat_packets, kv.first, false);
        else
            _cal_packets[kv.first] = get_packets_in_bursts(_flat_packets, kv.first, true);
    }
    _state_vectors = STATE_VECTORS(_flat_packets);
}

void S1_Decoder::apply_eccm(const int& detection_threshold, const int& mitigation_threshold) {
    _use_eccm = true;
    _eccm_detection_threshold = detection_threshold;
    _eccm_mitigation_threshold = mitigation_threshold;
}

// The reference prints the choices and ends the process (std::exit(1),
// s1_decoder.cpp:47-93); a library throws std::out_of_range with the same
// text instead.  As there, burst 0 is not checked.
void S1_Decoder::_validate_request(const std::string& swath, const int& burst) {
    if (!_swath_counts.count(swath)) {
        std::string msg = "The requested swath, " + swath + ", does not exist in the data. The available swaths are:";
        for (const auto& kv : _swath_counts) msg += " " + kv.first;
        throw std::out_of_range(msg);
    }
    if (burst > 0) {
        const auto& m = _cal_packets.count(swath) ? _cal_packets : _echo_packets;
        const size_t nb = m.at(swath).size();
        if ((size_t)burst >= nb)
            throw std::out_of_range("The requested burst, " + std::to_string(burst) +
                                    ", does not exist in the data. There are " + std::to_string(nb) +
                                    " bursts in swath " + swath + ".");
    }
}

const PACKET_VEC_2D& S1_Decoder::bursts(const std::string& swath) {
    _validate_request(swath);
    return _cal_packets.count(swath) ? _cal_packets.at(swath) : _echo_packets.at(swath);
}

CF_VEC_2D S1_Decoder::get_burst(const std::string& swath, const int& burst) {
    _validate_request(swath, burst);
    PACKET_VEC_1D packets = _echo_packets.at(swath).at((size_t)burst);
    decode_packets_in_place(packets);
    CF_VEC_2D out;
    out.reserve(packets.size());
    for (auto& p : packets) out.push_back(p.get_signal());
    return out;
}

CF_VEC_2D S1_Decoder::get_swath(const std::string& swath) {
    _validate_request(swath);
    const int num = swath_number_of(swath);
    PACKET_VEC_2D groups;
    if (_cal_packets.count(swath)) {
        groups = {_cal_packets.at(swath).at(0)};
    } else if (kIwSwaths.count(num)) {
        groups = _echo_packets.at(swath);
    } else if (kStripmapSwaths.count(num)) {
        groups = get_azimuth_blocks(_echo_packets.at(swath).at(0)).first;
    } else {
        throw std::domain_error("Only IW and SM modes are currently supported.");
    }
    size_t max_samples = 0;
    for (const auto& g : groups)
        if (!g.empty()) max_samples = std::max(max_samples, (size_t)(2 * g[0].get_num_quads()));
    CF_VEC_2D out;
    for (auto& g : groups) {
        PACKET_VEC_1D packets = g;
        decode_packets_in_place(packets);
        for (auto& p : packets) {
            CF_VEC_1D row = p.get_signal();
            if (!_cal_packets.count(swath)) row.resize(std::max(row.size(), max_samples));
            out.push_back(std::move(row));
        }
    }
    return out;
}

CF_VEC_2D S1_Decoder::_range_compress(PACKET_VEC_1D& packets, const bool& do_azimuth_fft) {
    S1_ECCM_THRESHOLDS thr{(double)_eccm_detection_threshold, (double)_eccm_mitigation_threshold};
    return range_compress_burst(packets, do_azimuth_fft, _use_eccm ? &thr : nullptr);
}

D_VEC_1D S1_Decoder::burst_effective_velocities(PACKET_VEC_1D& packets) {
    const L0Packet& first = packets.at(0);
    const D_VEC_1D R = first.get_slant_ranges();
    if (_state_vectors.size() == 0) return D_VEC_1D(R.size(), 7100.0);
    const double prf = 1.0 / (first.get_pri() * 1e-6);
    const double L = (double)packets.size();
    // the reference's row time t0 + i (L / PRF) / PRF at the middle row
    const double t = first.get_time() + (L / prf) / prf * (double)(packets.size() / 2);
    const STATE_VECTOR s = _state_vectors.interpolate(t);
    const double v = std::sqrt(s.velocity[0] * s.velocity[0] + s.velocity[1] * s.velocity[1] +
                               s.velocity[2] * s.velocity[2]);
    return get_effective_velocities(s.position, v, R);
}

CF_VEC_2D S1_Decoder::_azimuth_compress(PACKET_VEC_1D& packets, const bool& tops_mode) {
    S1_ECCM_THRESHOLDS thr{(double)_eccm_detection_threshold, (double)_eccm_mitigation_threshold};
    if (tops_mode) return tops_compress_burst(packets, _state_vectors, _use_eccm ? &thr : nullptr);
    const D_VEC_1D v = burst_effective_velocities(packets);
    return azimuth_compress_burst(packets, v, D_VEC_1D{0.0}, _use_eccm ? &thr : nullptr);
}

CF_VEC_2D S1_Decoder::get_range_compressed_burst(const std::string& swath, const int& burst,
                                                 const bool& range_doppler) {
    _validate_request(swath, burst);
    PACKET_VEC_1D packets = _echo_packets.at(swath).at((size_t)burst);
    return _range_compress(packets, range_doppler);
}

CF_VEC_2D S1_Decoder::get_azimuth_compressed_burst(const std::string& swath, const int& burst) {
    _validate_request(swath, burst);
    PACKET_VEC_1D packets = _echo_packets.at(swath).at((size_t)burst);
    return _azimuth_compress(packets, true);  // as s1_decoder.cpp:303, whatever the mode
}

namespace {
// Rows of one group appended to `out`, left-padded with zeros to max_size
// (the stripmap stacking of s1_decoder.cpp:221-240).
void append_rows(CF_VEC_2D& out, CF_VEC_2D&& rows, size_t max_size) {
    for (auto& r : rows) {
        if (r.size() < max_size) r.insert(r.begin(), max_size - r.size(), cdouble(0.0));
        out.push_back(std::move(r));
    }
}
}  // namespace

CF_VEC_2D S1_Decoder::get_range_compressed_swath(const std::string& swath, const bool& range_doppler) {
    _validate_request(swath);
    const int num = swath_number_of(swath);
    CF_VEC_2D out;
    if (kIwSwaths.count(num)) {
        for (auto& burst : _echo_packets.at(swath)) {
            PACKET_VEC_1D packets = burst;
            append_rows(out, _range_compress(packets, range_doppler), 0);
        }
    } else if (kStripmapSwaths.count(num)) {
        auto blocks = get_azimuth_blocks(_echo_packets.at(swath).at(0));
        for (auto& b : blocks.first) append_rows(out, _range_compress(b, range_doppler), (size_t)blocks.second);
    } else {
        throw std::invalid_argument(swath +
                                    " is not a supported swath type. Only IW and SM imaging modes are supported.");
    }
    return out;
}

CF_VEC_2D S1_Decoder::get_azimuth_compressed_swath(const std::string& swath) {
    _validate_request(swath);
    const int num = swath_number_of(swath);
    CF_VEC_2D out;
    auto pad20 = [&]() {
        if (!out.empty()) out.insert(out.end(), 20, CF_VEC_1D(out[0].size()));
    };
    if (kIwSwaths.count(num)) {
        for (auto& burst : _echo_packets.at(swath)) {
            if (burst.size() < 1000) continue;  // s1_decoder.cpp:322-330
            PACKET_VEC_1D packets = burst;
            append_rows(out, _azimuth_compress(packets, true), 0);
            pad20();
        }
    } else if (kStripmapSwaths.count(num)) {
        auto blocks = get_azimuth_blocks(_echo_packets.at(swath).at(0));
        for (auto& b : blocks.first) {
            append_rows(out, _azimuth_compress(b), (size_t)blocks.second);
            if (blocks.first.size() > 1) pad20();
        }
    } else {
        throw std::invalid_argument(swath +
                                    " is not a supported swath type. Only IW and SM imaging modes are supported.");
    }
    return out;
}


// ------------------------------------------------------------ image output
namespace {
int scale_mode(const std::string& m) {
    if (m == "norm_log") return S1G_SCALE_NORM_LOG;
    if (m == "norm") return S1G_SCALE_NORM;
    if (m == "mag") return S1G_SCALE_MAG;
    if (m == "real") return S1G_SCALE_REAL;
    if (m == "imag") return S1G_SCALE_IMAG;
    throw std::invalid_argument(m + " is not a valid scaling mode.");
}

// The image on the GPU, scaled, back as row-major floats.
std::vector<float> scale_dev(const CF_VEC_2D& img, int mode, int* rows_out, int* cols_out) {
    const int rows = (int)img.size();
    const int cols = rows ? (int)img[0].size() : 0;
    for (const auto& r : img)
        if ((int)r.size() != cols) throw std::invalid_argument("scaling needs rows of equal length");
    *rows_out = rows;
    *cols_out = cols;
    std::vector<float> out((size_t)rows * cols);
    if (out.empty()) return out;
    DevBuf d(sizeof(cdouble) * out.size()), o(sizeof(float) * out.size());
    upload_rows(img, cols, d.as<cdouble>());
    check(s1g_scale(ctx(), S1G_F64, mode, rows, cols, d.p, cols, o.as<float>(), cols), "scale");
    download(out.data(), o.p, sizeof(float) * out.size());
    return out;
}

std::vector<std::vector<float>> unflatten(const std::vector<float>& v, int rows, int cols) {
    std::vector<std::vector<float>> out((size_t)rows);
    for (int r = 0; r < rows; ++r) out[(size_t)r].assign(v.begin() + (long)r * cols, v.begin() + (long)(r + 1) * cols);
    return out;
}
}  // namespace

std::vector<std::vector<float>> norm_2d(const CF_VEC_2D& complex_values, const bool& log_scale) {
    int r, c;
    const std::vector<float> v = scale_dev(complex_values, log_scale ? S1G_SCALE_NORM_LOG : S1G_SCALE_NORM, &r, &c);
    return unflatten(v, r, c);
}

F_VEC_2D magnitude_2d(const CF_VEC_2D& complex_values) {
    int r, c;
    const std::vector<float> v = scale_dev(complex_values, S1G_SCALE_MAG, &r, &c);
    return unflatten(v, r, c);
}

std::vector<float> scale(const CF_VEC_2D& signal, const std::string& scaling_mode) {
    int r, c;
    return scale_dev(signal, scale_mode(scaling_mode), &r, &c);
}

D_VEC_1D norm_1d(const CF_VEC_1D& complex_values, const bool& log_scale) {
    D_VEC_1D n(complex_values.size());
    double mx = 0;
    for (size_t i = 0; i < n.size(); ++i) {
        n[i] = std::abs(complex_values[i]);
        if (n[i] > mx) mx = n[i];
    }
    for (auto& v : n) v = log_scale ? 20 * std::log10(v / mx) : v / mx;
    return n;
}

D_VEC_1D magnitude_1d(const CF_VEC_1D& complex_values) {
    D_VEC_1D m(complex_values.size());
    for (size_t i = 0; i < m.size(); ++i) m[i] = std::pow(std::abs(complex_values[i]), 2.0);
    return m;
}

D_VEC_1D scale(const CF_VEC_1D& signal, const std::string& scaling_mode) {
    const int mode = scale_mode(scaling_mode);
    if (mode == S1G_SCALE_NORM_LOG) return norm_1d(signal, true);
    if (mode == S1G_SCALE_NORM) return norm_1d(signal, false);
    if (mode == S1G_SCALE_MAG) return magnitude_1d(signal);
    D_VEC_1D out(signal.size());
    for (size_t i = 0; i < out.size(); ++i) out[i] = mode == S1G_SCALE_REAL ? signal[i].real() : signal[i].imag();
    return out;
}

void _write_tif(std::vector<float>& img_data, const int& rows, const int& cols, const std::string out_filename) {
    if ((size_t)rows * (size_t)cols > img_data.size()) throw std::invalid_argument("image smaller than rows x cols");
    if (s1g_write_tiff(out_filename.c_str(), img_data.data(), rows, cols, 1) != S1G_OK)
        throw std::runtime_error("cannot write " + out_filename);
}

void write_tif(CF_VEC_2D& img_data, const std::string& out_filename, const std::string& scaling_mode) {
    const int rows = (int)img_data.size(), cols = rows ? (int)img_data[0].size() : 0;
    std::vector<float> scaled = scale(img_data, scaling_mode);
    _write_tif(scaled, rows, cols, out_filename);
}

void write_cf32(const CF_VEC_2D& data, const std::string& out_filename) {
    std::ofstream f(out_filename, std::ios::binary);
    if (!f) throw std::runtime_error("Unable to open file " + out_filename + " for writing.");
    const int rows = (int)data.size(), cols = rows ? (int)data[0].size() : 0;
    if (rows && cols) {
        DevBuf d(sizeof(cdouble) * (size_t)rows * cols), o(8 * (size_t)rows * cols);
        upload_rows(data, cols, d.as<cdouble>());
        check(s1g_to_cf32(ctx(), rows, cols, d.p, cols, o.p, cols), "to_cf32");
        std::vector<std::complex<float>> h((size_t)rows * cols);
        download(h.data(), o.p, 8 * h.size());
        f.write((const char*)h.data(), (std::streamsize)(8 * h.size()));
    }
    if (!f) throw std::runtime_error("Failed to write to file " + out_filename + ".");
}
