/*
 * s1_decoder.h -- drop-in S1_Decoder (s1_decoder.h:13-77 of the reference):
 * a Level-0 file's packets split into swaths and bursts, its state vectors,
 * and the decoded, range-compressed and focused bursts and swaths, computed
 * on the GPU.  Also the burst-level bodies of S1_Decoder::_range_compress
 * and _azimuth_compress (s1_decoder.cpp:448-701) as free functions over one
 * burst's packets.
 *
 * Focusing follows the reference's two branches: bursts
 * (get_azimuth_compressed_burst) and IW swaths take the TOPS branch --
 * Doppler-centroid estimation, azimuth-frequency unfolding, the per-line
 * filter with a platform state per line, azimuth IFFT, azimuth-time
 * unfolding (tops_compress_burst); stripmap swaths take the range-Doppler
 * chain (decode, ECCM, range compression, azimuth FFT, RCMC + matched
 * filter, azimuth IFFT) with a zero Doppler centroid and the effective
 * velocity per range sample from the state vectors at the burst's middle
 * row.
 */
#pragma once

#include <set>

#include "packet.h"
#include "state_vectors.h"

struct S1_ECCM_THRESHOLDS {
    double detection;
    double mitigation;
};

/* Decode (+ ECCM) + range compression, optionally followed by the azimuth
 * FFT with fftshift (S1_Decoder::_range_compress, s1_decoder.cpp:448-508). */
CF_VEC_2D range_compress_burst(PACKET_VEC_1D& packets, const bool& do_azimuth_fft = false,
                               const S1_ECCM_THRESHOLDS* eccm = nullptr);

/* The whole chain to the focused burst: decode (+ ECCM), range compression,
 * azimuth FFT, RCMC + matched filter, azimuth IFFT
 * (S1_Decoder::_azimuth_compress, s1_decoder.cpp:511-701). */
CF_VEC_2D azimuth_compress_burst(PACKET_VEC_1D& packets, const D_VEC_1D& effective_velocities,
                                 const D_VEC_1D& doppler_centroids, const S1_ECCM_THRESHOLDS* eccm = nullptr);

/* The TOPS branch of S1_Decoder::_azimuth_compress (s1_decoder.cpp:511-701)
 * on the GPU: decode (+ ECCM) and range compression, the Doppler-centroid
 * rate from the first state vector's speed and the burst's beam angles, the
 * Doppler centroid over 30 range blocks, azimuth-frequency unfolding, the
 * per-line filter with the platform state interpolated at each unfolded
 * line's time, azimuth IFFT and azimuth-time unfolding.  std::runtime_error
 * without state vectors; std::invalid_argument when the burst's geometry
 * leaves the unfolding's windows (e.g. a burst without beam steering). */
CF_VEC_2D tops_compress_burst(PACKET_VEC_1D& packets, const STATE_VECTORS& state_vectors,
                              const S1_ECCM_THRESHOLDS* eccm = nullptr);

class S1_Decoder {
   public:
    S1_Decoder() {}
    explicit S1_Decoder(const std::string& filename);

    /* ECCM before range compression with the given thresholds
     * (s1_decoder.cpp:37-44); apply_eccm() without thresholds uses the
     * reference's defaults of zero. */
    void apply_eccm() { _use_eccm = true; }
    void apply_eccm(const int& detection_threshold, const int& mitigation_threshold);

    [[nodiscard]] STATE_VECTORS get_state_vectors() { return _state_vectors; }

    [[nodiscard]] CF_VEC_2D get_burst(const std::string& swath, const int& burst);
    [[nodiscard]] CF_VEC_2D get_swath(const std::string& swath);
    [[nodiscard]] CF_VEC_2D get_range_compressed_burst(const std::string& swath, const int& burst,
                                                       const bool& range_doppler = false);
    [[nodiscard]] CF_VEC_2D get_range_compressed_swath(const std::string& swath, const bool& range_doppler = false);
    [[nodiscard]] CF_VEC_2D get_azimuth_compressed_burst(const std::string& swath, const int& burst);
    [[nodiscard]] CF_VEC_2D get_azimuth_compressed_swath(const std::string& swath);

    /* Not in the reference's interface: the swaths present with their packet
     * counts, the bursts of a swath, and the effective velocities the
     * focusing of a burst uses. */
    [[nodiscard]] const std::unordered_map<std::string, int>& swath_counts() const { return _swath_counts; }
    [[nodiscard]] const PACKET_VEC_2D& bursts(const std::string& swath);
    [[nodiscard]] D_VEC_1D burst_effective_velocities(PACKET_VEC_1D& packets);

   private:
    std::string _filename;
    bool _use_eccm = false;
    int _eccm_detection_threshold = 0;
    int _eccm_mitigation_threshold = 0;
    PACKET_VEC_1D _flat_packets;
    std::unordered_map<std::string, int> _swath_counts;
    std::unordered_map<std::string, PACKET_VEC_2D> _echo_packets;
    std::unordered_map<std::string, PACKET_VEC_2D> _cal_packets;
    STATE_VECTORS _state_vectors;

    void _validate_request(const std::string& swath, const int& burst = -1);
    CF_VEC_2D _range_compress(PACKET_VEC_1D& packets, const bool& do_azimuth_fft = false);
    CF_VEC_2D _azimuth_compress(PACKET_VEC_1D& packets, const bool& tops_mode = false);
};
