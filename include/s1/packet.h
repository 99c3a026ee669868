/*
 * packet.h -- drop-in L0Packet (packet.h:63-215 of the reference) for the
 * Level-0 -> SLC path: header fields and the physical quantities derived
 * from them, the replica chirp, and the FDBAQ-decoded signal computed on
 * the GPU (s1g_decode_fdbaq), and reading a Level-0 file into packets and
 * bursts.  Members outside the path (printing, mode strings) are not
 * provided.
 *
 * Errors are C++ exceptions as in the reference: a packet whose user data
 * fails to decode throws std::runtime_error from get_signal() and
 * decode_packets*() (the C ABI reports them per packet).
 */
#pragma once

#include <stdexcept>

#include "misc_types.h"
#include "../s1gpu.h"

class L0Packet {
   public:
    L0Packet() {}

    /* Same signature and checks as the reference constructor
     * (packet.h:142-167): header maps keyed by the reference's field names
     * (structs.h:54-166) and the raw user data field. */
    L0Packet(std::unordered_map<std::string, int> primary_header,
             std::unordered_map<std::string, int> secondary_header, UINT8_VEC_1D raw_user_data, int packet_index);

    /* A packet from its bytes: 6-byte primary and 62-byte secondary header
     * followed by the user data field. */
    static L0Packet from_bytes(const UINT8_VEC_1D& packet, int packet_index = 0);

    [[nodiscard]] int primary_header(const std::string& key) const;
    [[nodiscard]] int secondary_header(const std::string& key) const;

    [[nodiscard]] bool is_empty() const { return _is_empty; }
    [[nodiscard]] int get_num_quads() const { return _info.num_quads; }
    [[nodiscard]] int get_num_baq_blocks() const { return (2 * _info.num_quads + 255) / 256; }
    [[nodiscard]] int get_user_data_length() const { return _info.user_data_length; }
    [[nodiscard]] char get_data_format() const { return _info.data_format; }
    [[nodiscard]] int get_packet_index() const { return _packet_index; }
    [[nodiscard]] int get_baq_block_length() const { return _info.baq_block_length; }
    [[nodiscard]] int get_replica_chirp_length() const { return _info.replica_chirp_length; }
    [[nodiscard]] double get_pulse_length() const { return _info.pulse_length_us; }
    [[nodiscard]] double get_tx_ramp_rate() const { return _info.txprr_mhz_us; }
    [[nodiscard]] double get_start_frequency() const { return _info.txpsf_mhz; }
    [[nodiscard]] double get_pri() const { return _info.pri_us; }
    [[nodiscard]] double get_swl() const { return _info.swl_us; }
    [[nodiscard]] double get_swst() const { return _info.swst_us; }
    [[nodiscard]] double get_rx_gain() const { return _info.rx_gain_db; }
    [[nodiscard]] double get_range_sample_rate() const { return _info.fs_mhz; }

    /* coarse_time + 1 / fine_time, as the reference computes it
     * (packet.cpp:56-63). */
    [[nodiscard]] double get_time() const;
    /* The swath's name from swath_number (SWATH_NUM_TO_STRING, structs.h:489-559). */
    [[nodiscard]] std::string get_swath() const;
    /* 'V' / 'H' from rx_channel_id 0 / 1, and from the polarisation code
     * 0-3 / 4-7 (packet.cpp:104-122); other codes throw std::runtime_error. */
    [[nodiscard]] char get_rx_polarization() const;
    [[nodiscard]] char get_tx_polarization() const;
    /* linspace(-0.018, 0.018, 1024)[azimuth_beam_address], radians
     * (packet.cpp:83-90). */
    [[nodiscard]] double get_azimuth_beam_angle() const;
    /* Two-way times of the range samples: linspace from rank * PRI + SWST +
     * the suppressed-data time over 2 * num_quads / fs (packet.cpp:304-318). */
    [[nodiscard]] D_VEC_1D get_slant_range_times(int num_ranges = 0) const;
    [[nodiscard]] D_VEC_1D get_slant_ranges(int num_ranges = 0) const;
    [[nodiscard]] D_VEC_1D get_timing_corrections() const;
    [[nodiscard]] CF_VEC_1D get_replica_chirp() const;
    [[nodiscard]] CF_VEC_1D get_signal();

    /* The raw user data field and the parsed header (not in the reference's
     * public interface; used to batch packets for the GPU). */
    [[nodiscard]] const UINT8_VEC_1D& get_user_data() const { return _raw_user_data; }
    [[nodiscard]] const s1g_packet_info& info() const { return _info; }
    [[nodiscard]] bool signal_is_set() const { return _signal_set; }
    void set_signal(CF_VEC_1D signal);

    typedef std::vector<L0Packet> PACKET_VEC_1D;
    typedef std::vector<std::vector<L0Packet>> PACKET_VEC_2D;

   private:
    s1g_packet_info _info{};
    UINT8_VEC_1D _raw_user_data;
    CF_VEC_1D _signal;
    int _packet_index = 0;
    bool _is_empty = true;
    bool _signal_set = false;
};

typedef L0Packet::PACKET_VEC_1D PACKET_VEC_1D;
typedef L0Packet::PACKET_VEC_2D PACKET_VEC_2D;

/* The packets of a Level-0 file (get_packets, packet.cpp:880-905): all of
 * them when num_packets is 0.  The file is read whole and indexed with
 * s1g_index_level0; a packet that cannot be constructed is reported and
 * skipped, as the reference skips it. */
PACKET_VEC_1D get_packets(const std::string& filename, int num_packets = 0);
/* The packets of one swath (get_packets_in_swath, packet.cpp:917-951);
 * std::runtime_error when there are none. */
PACKET_VEC_1D get_packets_in_swath(const std::string& filename, const std::string& swath);
/* The swath's packets split into bursts (get_packets_in_bursts,
 * packet.cpp:1026-1076): a burst ends where azimuth_beam_address decreases
 * or pri_count jumps; bursts of fewer than 1000 packets are dropped except
 * the last; calibration swaths stay whole (get_cal_packets). */
PACKET_VEC_2D get_packets_in_bursts(PACKET_VEC_1D& packets, const std::string& swath, bool get_cal_packets = false);
PACKET_VEC_2D get_packets_in_bursts(const std::string& filename, const std::string& swath);

/* decode_packets_in_place (packet.h:228): every packet's signal, decoded in
 * one GPU launch.  decode_packets (packet.h:238): the same on a copy. */
void decode_packets_in_place(PACKET_VEC_1D& packets);
PACKET_VEC_1D decode_packets(const PACKET_VEC_1D& packets);
