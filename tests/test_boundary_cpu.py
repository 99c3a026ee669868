"""The C ABI without a GPU: the library loads, exports every function that
include/s1gpu.h declares, and its host-side functions (symbol rule, header
parse, slant ranges, timing corrections) agree with the oracle."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

import oracle
from sentinel1_decode_amd import capi, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_functions():
    text = open(os.path.join(ROOT, "include", "s1gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(s1g_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the binding covers the same set
    assert sorted(capi.EXPORTED) == names


def test_fdbaq_symbol_rule_exhaustive():
    """The decoder's clo-table rule equals the reference's prefix walk for all
    10-bit code windows of every BRC (golden vectors from the reference)."""
    g = np.load(os.path.join(GOLD, "ref_huffman.npz"))["full"]
    for brc in range(5):
        for code in range(1024):
            ln, m = capi.fdbaq_symbol(brc, code)
            st, rm, rbits = g[brc, code]
            assert st == 0
            assert (m, ln) == (rm, rbits), (brc, code)


def test_parse_header_readme_kat():
    kat = json.load(open(os.path.join(GOLD, "ref_header_kat.json")))
    info = capi.parse_packet_header(bytes.fromhex(kat["header_hex"]))
    assert list(info.primary) == kat["primary"]
    assert list(info.secondary) == kat["secondary"]
    assert info.num_quads == 11938 and info.baq_mode == 12 and info.data_format == b"D"
    assert info.user_data_length == 18621 + 1 - 62
    assert info.baq_block_length == 256
    assert info.replica_chirp_length == int(np.floor(64.34523813 * info.pulse_length_us))
    phys = kat["readme_physical"]
    got = dict(pulse_length_us=info.pulse_length_us, txprr=info.txprr_mhz_us, txpsf=info.txpsf_mhz,
               pri_us=info.pri_us, swl_us=info.swl_us, swst_us=info.swst_us, rx_gain_db=info.rx_gain_db,
               baq_block_length=info.baq_block_length)
    for k, v in phys.items():
        assert float(f"{got[k]:.6g}") == pytest.approx(v, rel=1e-12, abs=1e-12), k


def test_header_geometry_matches_oracle():
    for spec in (synth.config1(n_lines=3), synth.config0(n_lines=3),
                 synth.BurstSpec(n_lines=3, n_samples=21600, rgdec=6, pri_code=22276),
                 synth.BurstSpec(n_lines=3, n_samples=20400, rgdec=11, pri_code=25868)):
        b = synth.make_burst(spec)
        hdr = b.packet(0)[:68]
        info = capi.parse_packet_header(hdr)
        _, sec = oracle.parse_headers(hdr)
        par = oracle.packet_params(sec)
        assert info.txpsf_mhz == par["txpsf"] and info.txprr_mhz_us == par["txprr"]
        assert info.pulse_length_us == par["pulse_length_us"] and info.fs_mhz == par["fs_mhz"]
        assert info.replica_chirp_length == par["replica_chirp_length"]
        assert np.array_equal(capi.slant_ranges(info), oracle.slant_ranges(sec))
        assert np.array_equal(capi.timing_corrections(info), oracle.timing_corrections(sec))
        # the scene places range sample N/2 at 850 km
        assert abs(capi.slant_ranges(info)[spec.n_samples // 2] - 850e3) < 5.0


def test_bad_arguments_fail_cleanly():
    lib = capi.lib()
    assert lib.s1g_fdbaq_symbol(5, 0, None, None) == -1
    assert lib.s1g_fdbaq_symbol(0, 1024, None, None) == -1
    assert lib.s1g_parse_packet_header(None, 0, None) == -1
    info = capi.PacketInfo()
    short = np.zeros(10, np.uint8)
    assert lib.s1g_parse_packet_header(short.ctypes.data, 10, C.byref(info)) == -1
    # null context everywhere returns S1G_EARG, never crashes
    assert lib.s1g_decode_fdbaq(None, None, None, None, None, 1, 0, 0, 1, None, 0, None) == -1
    assert lib.s1g_range_compress(None, None, 1, None, 0, None, 0) == -1
    assert lib.s1g_focus_burst(None, None, 1) == -1
    assert lib.s1g_set_stream(None, None) == -1
    assert lib.s1g_destroy(None) == 0
    assert lib.s1g_replica_chirp_length(64.34523813, 1967 / 37.53472224) == 3372


def test_create_without_gpu_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = capi.lib().s1g_create(C.byref(h), 0)
    assert rc == -2 and not h.value


def test_cpp_dropin_headers_compile_standalone(tmp_path):
    """include/s1/*.h are plain C++17 (no HIP headers): a reference caller
    compiles against them with -I include/s1."""
    src = tmp_path / "caller.cpp"
    src.write_text(
        '#include "packet.h"\n#include "signal_processing.h"\n#include "image_formation.h"\n'
        '#include "s1_decoder.h"\n'
        "int main() { CF_VEC_1D v(8); PACKET_VEC_1D p; (void)v; (void)p;\n"
        "  CF_VEC_1D (*f)(const CF_VEC_1D&, int, const bool&) = compute_1d_dft; (void)f; return 0; }\n")
    import subprocess
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I", os.path.join(ROOT, "include", "s1"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_dropin_library_exports_reference_names():
    import subprocess
    lib = os.path.join(ROOT, "sentinel1_decode_amd", "libs1dropin.so")
    assert os.path.exists(lib), "build the drop-in: python -m sentinel1_decode_amd.build"
    syms = subprocess.run(["nm", "-DC", "--defined-only", lib], capture_output=True, text=True).stdout
    for name in ("decode_packets_in_place(", "decode_packets(", "L0Packet::get_signal()",
                 "L0Packet::get_replica_chirp() const", "compute_1d_dft(", "compute_1d_dft_in_place(",
                 "_compute_axis_dft(", "compute_axis_dft(", "compute_axis_dft_in_place(", "transpose(",
                 "fftshift_in_place(", "sinc_interpolate(", "eccm(", "pulse_compression(",
                 "get_reference_function(", "linspace(", "range_compress_burst(", "azimuth_compress_burst("):
        assert name in syms, name


def test_cpp_dropin_fails_loudly_without_gpu():
    """No CPU fallback: the drop-in raises when no GPU is present."""
    import subprocess
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    exe = os.path.join(ROOT, "tests", "cpp", "test_dropin")
    assert os.path.exists(exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "FAIL exception" in r.stdout
