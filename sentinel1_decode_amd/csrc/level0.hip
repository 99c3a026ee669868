// level0.hip -- Level-0 files and burst geometry: packet framing, the
// header-parse kernel, burst segmentation, state vectors from the
// sub-commutated ancillary words, effective velocities, and the reference's
// secondary range compression and RCMC phase ramp.
//
// Framing is a serial walk (each packet's length gives the next one's
// start), 8 bytes of index per ~20 KB packet, so it stays on the host; the
// 68 header bytes of every packet are parsed on the GPU, one wavefront per
// packet with one field per lane, straight into the arrays the decoder
// takes.  Segmentation and the ancillary-word assembly are sequential rules
// over a few header fields per packet and run on the host.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "s1g_internal.h"

using namespace s1g;

namespace {

// Secondary-header field positions (structs.h:68-115) in a row of
// S1G_NFIELDS fields (primary fields first).
constexpr int kFScIndex = 8 + 9, kFScWord = 8 + 10, kFPriCount = 8 + 12,
              kFAzBeam = 8 + 36, kFSwath = 8 + 43, kFNumQuads = 8 + 44, kFDataLen = 7;

__constant__ int c_field_bits[S1G_NFIELDS] = {S1G_HEADER_FIELD_BITS};

// One wavefront per packet: lane f < 54 extracts field f from the packet's
// 68 header bytes (big-endian, MSB first); fields are at most 32 bits, so a
// lane reads at most 5 bytes.
__global__ __launch_bounds__(256) void parse_headers_kernel(const uint8_t* __restrict__ stream,
                                                            const uint64_t* __restrict__ offsets, int n,
                                                            int64_t* __restrict__ fields,
                                                            uint64_t* __restrict__ ud_offsets,
                                                            uint32_t* __restrict__ ud_lengths,
                                                            int32_t* __restrict__ num_quads) {
    __shared__ int s_bit[S1G_NFIELDS];
    const int lane = threadIdx.x & 63;
    if (threadIdx.x == 0) {
        int b = 0;
        for (int f = 0; f < S1G_NFIELDS; ++f) {
            s_bit[f] = b;
            b += c_field_bits[f];
        }
    }
    __syncthreads();
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= n || lane >= S1G_NFIELDS) return;
    const uint8_t* p = stream + offsets[k];
    const int bit = s_bit[lane], w = c_field_bits[lane];
    const int b0 = bit >> 3, b1 = (bit + w - 1) >> 3;
    uint64_t acc = 0;
    for (int b = b0; b <= b1; ++b) acc = (acc << 8) | p[b];
    const int tail = 8 * (b1 + 1) - (bit + w);  // bits after the field in the last byte
    const uint64_t v = (acc >> tail) & ((w == 64) ? ~0ull : ((1ull << w) - 1));
    if (fields) fields[(int64_t)k * S1G_NFIELDS + lane] = (int64_t)v;
    if (lane == kFDataLen) {
        if (ud_offsets) ud_offsets[k] = offsets[k] + 68;
        if (ud_lengths) ud_lengths[k] = (uint32_t)(v + 1 - 62);
    }
    if (lane == kFNumQuads && num_quads) num_quads[k] = (int32_t)v;
}

// ---------------------------------------------------------- ancillary words
// Position of each sub-commutated word index 0..64 (SUB_COMM_KEY_POS,
// structs.h:622-689): the quantity it belongs to.  Words of one quantity
// arrive most significant first.
enum Sub { kDummy, kPx, kPy, kPz, kVx, kVy, kVz, kPod, kQ0, kQ1, kQ2, kQ3, kWx, kWy, kWz, kTime, kOther, kNSub };
const Sub kSubOfIndex[65] = {kDummy, kPx, kPx, kPx, kPx, kPy, kPy, kPy, kPy, kPz, kPz, kPz, kPz, kVx, kVx, kVy,
                             kVy, kVz, kVz, kPod, kPod, kPod, kPod, kQ0, kQ0, kQ1, kQ1, kQ2, kQ2, kQ3, kQ3, kWx,
                             kWx, kWy, kWy, kWz, kWz, kTime, kTime, kTime, kTime, kOther, kOther, kOther, kOther,
                             kOther, kOther, kOther, kOther, kOther, kOther, kOther, kOther, kOther, kOther, kOther,
                             kOther, kOther, kOther, kOther, kOther, kOther, kOther, kOther, kOther};

// The reference's bit-pattern -> floating-point rule (int_to_ieee754,
// decoding_utils.cpp:149-170): sign, biased exponent and the mantissa bits
// as a sum of powers of two, always with the implicit leading one.
double ieee_from_bits(uint64_t v, bool is_double) {
    const int mshift = is_double ? 52 : 23;
    const uint64_t emask = is_double ? 0x7FF0000000000000ull : 0x7f800000ull;
    const int bias = is_double ? 1023 : 127;
    const int sshift = is_double ? 63 : 31;
    const int sign = (int)((v >> sshift) & 1);
    const int e = (int)((v & emask) >> mshift) - bias;
    double total = 0.0;
    int power = -1;
    for (int i = 0; i < mshift; ++i) {
        const int bitv = (int)((v >> (mshift - i - 1)) & 1);
        total += bitv * ldexp(1.0, power);
        --power;
    }
    return (sign ? -1 : 1) * ldexp(1.0, e) * (total + 1.0);
}

// parse_time_stamp (state_vectors.cpp:5-16), its fine-time loop included:
// after the first step only bit 8 of the fine time is kept, so the result
// is coarse + 0.5 * bit23 + 9 * 2^-16 * bit8.
double time_from_bits(uint64_t raw) {
    const unsigned coarse = (unsigned)((raw >> 24) & 0xffffffffull);
    unsigned fine_raw = (unsigned)(raw & 0xffffff);
    double fine = 0.0;
    for (int i = 23; i >= 0; --i) {
        fine += (double)(fine_raw >> i) * ldexp(1.0, i - 24);
        fine_raw &= 1u << 8;
    }
    return (double)coarse + fine;
}

s1g_state_vector vector_from_words(const uint64_t* w) {
    s1g_state_vector sv;
    sv.time = time_from_bits(w[kTime]);
    sv.position[0] = ieee_from_bits(w[kPx], true);
    sv.position[1] = ieee_from_bits(w[kPy], true);
    sv.position[2] = ieee_from_bits(w[kPz], true);
    sv.velocity[0] = ieee_from_bits(w[kVx], false);
    sv.velocity[1] = ieee_from_bits(w[kVy], false);
    sv.velocity[2] = ieee_from_bits(w[kVz], false);
    sv.quaternion[0] = ieee_from_bits(w[kQ0], false);
    sv.quaternion[1] = ieee_from_bits(w[kQ1], false);
    sv.quaternion[2] = ieee_from_bits(w[kQ2], false);
    sv.quaternion[3] = ieee_from_bits(w[kQ3], false);
    sv.angular_rate[0] = ieee_from_bits(w[kWx], false);
    sv.angular_rate[1] = ieee_from_bits(w[kWy], false);
    sv.angular_rate[2] = ieee_from_bits(w[kWz], false);
    return sv;
}

// find_time_bounds (state_vectors.cpp:127-138): bisection over [lo, hi).
// A pivot equal to `time` ends the search; the reference then returns the
// window's start twice (and interpolates 0 / 0), here the pivot itself.
void time_bounds(const s1g_state_vector* v, int lo, int hi, double time, int* a, int* b) {
    while (true) {
        const int size = hi - lo;
        if (size == 1) {
            *a = lo;
            *b = hi;
            return;
        }
        const int pivot = size / 2;
        const double tp = v[lo + pivot].time;
        if (tp == time) {
            *a = lo + pivot;
            *b = lo + pivot;
            return;
        }
        if (tp > time)
            hi = lo + pivot;
        else
            lo = lo + pivot;
    }
}

// ------------------------------------------------------------- geometry
constexpr double kSemiMajor = 6378137.0;     // WGS84 (structs.h:26)
constexpr double kSemiMinor = 6356752.3142;  // WGS84 (structs.h:27)
constexpr double kC = 299792458.0;
constexpr double kF0 = 5.405000454334350e9;
constexpr double kPI = 3.14159265358979;  // the reference's PI (structs.h:23)

// get_effective_velocities (image_formation.cpp:391-419) in its expression
// order; "lat" is the ratio z / x, as in the reference.  The per-line Earth
// radius and platform distance (cos / sin of lat) come from the host's libm,
// the per-sample part is IEEE arithmetic without contractions, so v_rel is
// the reference's to the last bit (the azimuth filter's phase, 4 pi R D /
// lambda ~ 2e8 rad, turns one ulp of D into ~2e-8 rad).
void line_radii(double px, double py, double pz, double* re, double* alt) {
#pragma clang fp contract(off)
    const double lat = pz / px;
    const double cl = cos(lat), sl = sin(lat);
    const double t1 = kSemiMajor * kSemiMajor * cl, t2 = kSemiMinor * kSemiMinor * sl;
    const double num = t1 * t1 + t2 * t2;
    const double t3 = kSemiMajor * cl, t4 = kSemiMinor * sl;
    const double den = t3 * t3 + t4 * t4;
    *re = sqrt(num / den);
    *alt = sqrt(px * px + py * py + pz * pz);
}

__global__ void effective_velocity_kernel(const double* __restrict__ radii, const double* __restrict__ vnorm,
                                          int n_lines, const double* __restrict__ R, int n, double* __restrict__ out,
                                          int64_t ld) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= n || i >= n_lines) return;
    const double v = vnorm[i];
    const double re = radii[2 * i], alt = radii[2 * i + 1];
    const double r = R[j];
    const double beta = (re * re + alt * alt - r * r) / (2 * re * alt);
    const double vg = re * v * beta / alt;
    out[(int64_t)i * ld + j] = sqrt(v * vg);
}

template <typename T>
__global__ void src_kernel(cpx<T>* __restrict__ data, int64_t ld, int n_rows, int n, const double* __restrict__ vrel,
                           const double* __restrict__ R, const double* __restrict__ fr, const double* __restrict__ fdc,
                           const double* __restrict__ faz, double* __restrict__ factors) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= n || i >= n_rows) return;
    const double lam = kC / kF0;
    const double v = vrel[(int64_t)i * n + j];
    const double f = faz[i] + fdc[j];
    const double D = sqrt(1 - ((lam * lam) * (f * f)) / (4 * v * v));
    if (factors) factors[(int64_t)i * n + j] = D;
    double k = 2.0 * (v * v) * pow(kF0, 3.0) * (D * D);
    k /= (kC * R[j] * (f * f));
    const double ph = -1.0 * kPI * (fr[j] * fr[j]) / k;
    double s, c;
    phase_sincos<T>(ph, &s, &c);
    cpx<T>& x = data[(int64_t)i * ld + j];
    const double xr = (double)x.x, xi = (double)x.y;
    x.x = (T)(xr * c - xi * s);
    x.y = (T)(xr * s + xi * c);
}

template <typename T>
__global__ void rcmc_phase_kernel(cpx<T>* __restrict__ data, int64_t ld, int n_rows, int n,
                                  const double* __restrict__ R, const double* __restrict__ fr,
                                  const double* __restrict__ factors) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= n || i >= n_rows) return;
    const double shift = (R[j] / factors[(int64_t)i * n + j]) - R[j];
    const double ph = -4.0 * kPI * fr[j] * shift / kC;
    double s, c;
    phase_sincos<T>(ph, &s, &c);
    cpx<T>& x = data[(int64_t)i * ld + j];
    const double xr = (double)x.x, xi = (double)x.y;
    x.x = (T)(xr * c - xi * s);
    x.y = (T)(xr * s + xi * c);
}

int launch_status(s1g_ctx* c, const char* what) {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return S1G_OK;
    char buf[160];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    return ctx_error(c, S1G_EHIP, buf);
}

}  // namespace

extern "C" {

int s1g_index_level0(const uint8_t* stream, int64_t nbytes, uint64_t* offsets, int64_t cap, int64_t* n_packets) {
    if (!n_packets || nbytes < 0 || (nbytes && !stream) || cap < 0 || (cap && !offsets)) return S1G_EARG;
    int64_t pos = 0, n = 0;
    while (nbytes - pos >= 6) {
        const int64_t pdl = ((int64_t)stream[pos + 4] << 8) | stream[pos + 5];
        const int64_t len = 6 + pdl + 1;
        if (pos + len > nbytes) break;  // truncated last packet
        if (n < cap) offsets[n] = (uint64_t)pos;
        ++n;
        pos += len;
    }
    *n_packets = n;
    return S1G_OK;
}

int s1g_parse_headers(s1g_ctx* c, const uint8_t* stream, const uint64_t* offsets, int n_packets, int64_t* fields,
                      uint64_t* ud_offsets, uint32_t* ud_lengths, int32_t* num_quads) {
    if (!c) return S1G_EARG;
    if (n_packets < 0) return ctx_error(c, S1G_EARG, "negative packet count");
    if (n_packets == 0) return S1G_OK;
    if (!stream || !offsets) return ctx_error(c, S1G_EARG, "null stream or offsets");
    dim3 grid((unsigned)((n_packets + 3) / 4)), block(256);
    hipLaunchKernelGGL(parse_headers_kernel, grid, block, 0, ctx_stream(c), stream, offsets, n_packets, fields,
                       ud_offsets, ud_lengths, num_quads);
    return launch_status(c, "header parse");
}

int s1g_segment_bursts(const int64_t* fields, int n_packets, int swath_number, int cal, int32_t* burst_of_packet,
                       int* n_bursts) {
    if (!n_bursts || n_packets < 0 || (n_packets && (!fields || !burst_of_packet))) return S1G_EARG;
    // the reference keeps header fields as int (packet.h:142-167)
    auto f = [&](int k, int idx) { return (int)(int32_t)fields[(int64_t)k * S1G_NFIELDS + idx]; };
    std::vector<int> current;
    int nb = 0;
    int prev_az = 0, prev_pri = 0;
    auto close = [&](bool keep) {
        for (int k : current) burst_of_packet[k] = keep ? nb : -1;
        if (keep) ++nb;
        current.clear();
    };
    for (int k = 0; k < n_packets; ++k) {
        burst_of_packet[k] = -1;
        if (f(k, kFSwath) == swath_number) {
            if (!cal) {
                const int az = f(k, kFAzBeam), pri = f(k, kFPriCount);
                if (prev_az == 0) {
                    prev_az = az;
                    prev_pri = pri;
                }
                if (az < prev_az || (int64_t)pri > (int64_t)prev_pri + 1) close(current.size() >= 1000);
                prev_pri = pri;
                prev_az = az;
            }
            current.push_back(k);
        }
        if (k == n_packets - 1 && !current.empty()) close(true);
    }
    *n_bursts = nb;
    return S1G_OK;
}

int s1g_state_vectors(const int64_t* fields, int n_packets, s1g_state_vector* out, int cap, int* n_out) {
    if (!n_out || n_packets < 0 || (n_packets && !fields) || cap < 0 || (cap && !out)) return S1G_EARG;
    *n_out = 0;
    auto f = [&](int k, int idx) { return fields[(int64_t)k * S1G_NFIELDS + idx]; };
    // the first packet after packet 0 that opens a cycle (index 1)
    int start = 1;
    while (start < n_packets && f(start, kFScIndex) != 1) ++start;
    if (start >= n_packets) return S1G_OK;
    uint64_t words[kNSub] = {0};
    double latest = 0.0;  // the reference seeds its set of times with 0
    int n = 0;
    for (int k = start; k < n_packets; ++k) {
        const int idx = (int)f(k, kFScIndex);
        if (idx == 1 && k != start) {
            const s1g_state_vector sv = vector_from_words(words);
            if (sv.time > latest) {
                latest = sv.time;
                if (n < cap) out[n] = sv;
                ++n;
            }
            memset(words, 0, sizeof words);
        }
        if (idx < 0 || idx > 64) continue;
        const Sub q = kSubOfIndex[idx];
        words[q] = (words[q] << 16) | (uint64_t)(f(k, kFScWord) & 0xffff);
    }
    *n_out = n;
    return S1G_OK;
}

int s1g_interpolate_state(const s1g_state_vector* v, int n, double time, s1g_state_vector* out) {
    if (!v || n <= 0 || !out) return S1G_EARG;
    if (time > v[n - 1].time) {
        *out = v[n - 1];
        out->time = time;
        return S1G_OK;
    }
    if (time < v[0].time) {
        *out = v[0];
        out->time = time;
        return S1G_OK;
    }
    int a, b;
    time_bounds(v, 0, n, time, &a, &b);
    // b reaches n when the last bracket is the final vector alone, and a == b
    // when `time` equals a vector's time: the reference then reads past its
    // vectors or divides 0 / 0; here the bracket's own vector is returned
    if (b >= n) b = n - 1;
    if (a == b) {
        *out = v[a];
        out->time = time;
        return S1G_OK;
    }
    const double t = (time - v[a].time) / (v[b].time - v[a].time);
    // std::lerp (libstdc++): exact endpoints, monotone in t
    auto lerp = [&](double x, double y) {
        if ((x <= 0 && y >= 0) || (x >= 0 && y <= 0)) return t * y + (1 - t) * x;
        if (t == 1) return y;
        const double r = x + t * (y - x);
        return (t > 1) == (y > x) ? (y < r ? r : y) : (y > r ? r : y);
    };
    out->time = time;
    for (int d = 0; d < 3; ++d) {
        out->position[d] = lerp(v[a].position[d], v[b].position[d]);
        out->velocity[d] = lerp(v[a].velocity[d], v[b].velocity[d]);
        out->angular_rate[d] = lerp(v[a].angular_rate[d], v[b].angular_rate[d]);
    }
    for (int d = 0; d < 4; ++d) out->quaternion[d] = lerp(v[a].quaternion[d], v[b].quaternion[d]);
    return S1G_OK;
}

int s1g_line_states(const s1g_state_vector* v, int n, double t0, double dt, int count, double* positions,
                    double* velocity_norms) {
    if (!v || n <= 0 || count < 0 || (count && (!positions || !velocity_norms))) return S1G_EARG;
    for (int i = 0; i < count; ++i) {
        s1g_state_vector st;
        const int rc = s1g_interpolate_state(v, n, t0 + dt * i, &st);
        if (rc) return rc;
        for (int k = 0; k < 3; ++k) positions[3 * (int64_t)i + k] = st.position[k];
        const double* w = st.velocity;
        velocity_norms[i] = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    }
    return S1G_OK;
}

int s1g_effective_velocities(s1g_ctx* c, const double* positions, const double* velocity_norms, int n_lines,
                             const double* slant_ranges, int n_samples, double* out, int64_t ld_out) {
    if (!c) return S1G_EARG;
    if (n_lines < 0 || n_samples < 0 || ld_out < n_samples) return ctx_error(c, S1G_EARG, "bad shape");
    if (!n_lines || !n_samples) return S1G_OK;
    if (!positions || !velocity_norms || !slant_ranges || !out) return ctx_error(c, S1G_EARG, "null pointer");
    hipStream_t s = ctx_stream(c);
    std::vector<double> pos(3 * (size_t)n_lines), radii(2 * (size_t)n_lines);
    if (hipMemcpyAsync(pos.data(), positions, sizeof(double) * pos.size(), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return launch_status(c, "effective velocities");
    for (int i = 0; i < n_lines; ++i) line_radii(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], &radii[2 * i], &radii[2 * i + 1]);
    void* d_radii = nullptr;
    int rc = ctx_aux(c, 8, sizeof(double) * radii.size(), &d_radii);
    if (rc) return rc;
    if (hipMemcpyAsync(d_radii, radii.data(), sizeof(double) * radii.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return launch_status(c, "effective velocities");
    dim3 grid((unsigned)((n_samples + 255) / 256), (unsigned)n_lines), block(256);
    hipLaunchKernelGGL(effective_velocity_kernel, grid, block, 0, s, (const double*)d_radii, velocity_norms, n_lines,
                       slant_ranges, n_samples, out, ld_out);
    return launch_status(c, "effective velocities");
}

int s1g_secondary_range_compression(s1g_ctx* c, int prec, int n_rows, int n_samples, void* data, int64_t ld,
                                    const double* v_rel, const double* slant_ranges, const double* range_freqs,
                                    const double* f_dc, const double* az_freqs, double* rcmc_factors) {
    if (!c) return S1G_EARG;
    if (prec != S1G_F32 && prec != S1G_F64) return ctx_error(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64");
    if (n_rows < 0 || n_samples < 0 || ld < n_samples) return ctx_error(c, S1G_EARG, "bad shape");
    if (!n_rows || !n_samples) return S1G_OK;
    if (!data || !v_rel || !slant_ranges || !range_freqs || !f_dc || !az_freqs)
        return ctx_error(c, S1G_EARG, "null pointer");
    dim3 grid((unsigned)((n_samples + 255) / 256), (unsigned)n_rows), block(256);
    if (prec == S1G_F32)
        hipLaunchKernelGGL(src_kernel<float>, grid, block, 0, ctx_stream(c), (cpx<float>*)data, ld, n_rows,
                           n_samples, v_rel, slant_ranges, range_freqs, f_dc, az_freqs, rcmc_factors);
    else
        hipLaunchKernelGGL(src_kernel<double>, grid, block, 0, ctx_stream(c), (cpx<double>*)data, ld, n_rows,
                           n_samples, v_rel, slant_ranges, range_freqs, f_dc, az_freqs, rcmc_factors);
    return launch_status(c, "secondary range compression");
}

int s1g_rcmc_phase(s1g_ctx* c, int prec, int n_rows, int n_samples, void* data, int64_t ld,
                   const double* slant_ranges, const double* range_freqs, const double* rcmc_factors) {
    if (!c) return S1G_EARG;
    if (prec != S1G_F32 && prec != S1G_F64) return ctx_error(c, S1G_EARG, "precision must be S1G_F32 or S1G_F64");
    if (n_rows < 0 || n_samples < 0 || ld < n_samples) return ctx_error(c, S1G_EARG, "bad shape");
    if (!n_rows || !n_samples) return S1G_OK;
    if (!data || !slant_ranges || !range_freqs || !rcmc_factors) return ctx_error(c, S1G_EARG, "null pointer");
    dim3 grid((unsigned)((n_samples + 255) / 256), (unsigned)n_rows), block(256);
    if (prec == S1G_F32)
        hipLaunchKernelGGL(rcmc_phase_kernel<float>, grid, block, 0, ctx_stream(c), (cpx<float>*)data, ld, n_rows,
                           n_samples, slant_ranges, range_freqs, rcmc_factors);
    else
        hipLaunchKernelGGL(rcmc_phase_kernel<double>, grid, block, 0, ctx_stream(c), (cpx<double>*)data, ld,
                           n_rows, n_samples, slant_ranges, range_freqs, rcmc_factors);
    return launch_status(c, "rcmc phase");
}

}  // extern "C"
