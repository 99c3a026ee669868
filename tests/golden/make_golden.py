"""Write the golden vectors under tests/golden/ from the reference's own code.

Runs only where the reference sources exist: `make -C oracle ref` compiles the
reference's decoding_utils.cpp (with structs.h) behind oracle/ref_shim.cpp.
Outputs (data only):
  ref_tables.json   structs.h decode tables and constants
  ref_huffman.npz   huffman_decode of every 10-bit pattern per BRC, plus
                    truncated inputs (status, m_code, bits consumed)
  ref_read_bits.npz read_n_bits on seeded random bytes/offsets/lengths
  ref_header_kat.json  the README's packet-0 header fields (README.md:94-142),
                    packed into 68 bytes and parsed back by the reference's
                    parse_packet_header loop over its own width tables
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from sentinel1_decode_amd import synth  # noqa: E402

README_PRIMARY = dict(packet_version_number=0, packet_type=0, secondary_header_flag=1, process_id=65,
                      process_category=12, sequence_flags=3, packet_sequence_count=11157, packet_data_length=18621)
README_SECONDARY = dict(
    coarse_time=1406987562, fine_time=7757, sync_marker=892270675, data_take_id=225252800, ecc_number=8,
    test_mode=0, rx_channel_id=0, instrument_configuration_id=7, sc_data_word_index=27, sc_data_word=48835,
    space_packet_count=240533, pri_count=243384, error_flag=0, baq_mode=12, baq_block_length=31,
    range_decimation=8, rx_gain=8, tx_ramp_rate_sign=1, tx_ramp_rate_mag=1605, pulse_start_frequency_sign=0,
    pulse_start_frequency_mag=12335, pulse_length=1967, rank=9, pri=21859, swst=3681, swl=13979, ssb_flag=0,
    polarisation=7, temperature_compensation=3, elevation_beam_address=6, azimuth_beam_address=385,
    calibration_mode=0, tx_pulse_number=6, signal_type=0, swap=1, swath_number=10, num_quadratures=11938)
# physical values printed by the README for that packet (README.md:150-162)
README_PHYSICAL = dict(pulse_length_us=52.4048, txprr=1.07823, txpsf=-28.2515, pri_us=582.367, swl_us=372.428,
                       swst_us=98.0692, rx_gain_db=-4.0, baq_block_length=256.0)


def main():
    oracle.build()
    R = oracle.ref_lib()
    if R is None:
        sys.exit("reference shim not built (needs the reference sources)")
    t = oracle.ref_tables()
    out = {k: (v.tolist() if hasattr(v, "tolist") else {str(b): p for b, p in v.items()}) for k, v in t.items()}
    with open(os.path.join(HERE, "ref_tables.json"), "w") as f:
        json.dump(out, f, indent=1)

    # huffman_decode over every 10-bit pattern, with and without room to finish
    st = np.zeros((5, 1024, 3), np.int32)
    trunc = np.zeros((5, 1024, 3), np.int32)
    for brc in range(5):
        for code in range(1024):
            for arr, nb in ((st, 4), (trunc, 1)):
                data = np.zeros(4, np.uint8)
                data[0] = (code >> 2) & 0xFF
                data[1] = (code & 3) << 6
                bi = C.c_int(0)
                m = C.c_int(-1)
                s = R.ref_huffman_decode(data, nb, brc, C.byref(bi), C.byref(m))
                arr[brc, code] = (s, m.value if s == 0 else -1, bi.value)
    np.savez_compressed(os.path.join(HERE, "ref_huffman.npz"), full=st, one_byte=trunc)

    rng = np.random.default_rng(123)
    data = rng.integers(0, 256, 64).astype(np.uint8)
    cases = []
    for _ in range(2000):
        start = int(rng.integers(0, 64 * 8 + 8))
        n = int(rng.integers(1, 65))
        v = C.c_uint64(0)
        s = R.ref_read_n_bits(data, data.size, start, n, C.byref(v))
        cases.append((start, n, s, v.value if s == 0 else 0))
    np.savez_compressed(os.path.join(HERE, "ref_read_bits.npz"), data=data,
                        cases=np.array(cases, dtype=np.uint64))

    hdr = synth.pack_headers(README_PRIMARY, README_SECONDARY)
    p, s = oracle.ref_parse_headers(hdr)
    with open(os.path.join(HERE, "ref_header_kat.json"), "w") as f:
        json.dump({"header_hex": hdr.hex(), "primary": p.tolist(), "secondary": s.tolist(),
                   "readme_primary": README_PRIMARY, "readme_secondary": README_SECONDARY,
                   "readme_physical": README_PHYSICAL}, f, indent=1)
    print("golden vectors written")


if __name__ == "__main__":
    main()
