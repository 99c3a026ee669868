"""The C++ drop-in surface (include/s1/*.h) called the reference's way.

tests/cpp/test_dropin.cpp checks the building blocks against direct
evaluations of their definitions, and packet decoding / burst focusing
against the oracle vectors written here for a small burst (192 lines x
4096 samples): decode bit-exact, range compression 1e-11, FP64 SLC 1e-10,
FP32 SLC 1e-5 (relative L2).
"""
import os
import subprocess

import numpy as np
import pytest

from chain_ref import oracle_chain

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_dropin")


def test_cpp_dropin(tmp_path):
    from sentinel1_decode_amd import synth
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/test_dropin is not built (python -m sentinel1_decode_amd.build)")
    spec = synth.config0(n_lines=192, n_samples=4096, grid_spacing=(300, 40), seed=7)
    burst = synth.make_burst(spec)
    ch = oracle_chain(burst)
    offs = np.asarray(burst.ud_offsets, np.int64) - 68
    lens = np.asarray(burst.ud_lengths, np.int64) + 68
    np.asarray(burst.stream, np.uint8).tofile(tmp_path / "packets.bin")
    np.stack([offs, lens], axis=1).astype(np.int64).tofile(tmp_path / "index.bin")
    np.ascontiguousarray(ch["decoded"], np.complex128).tofile(tmp_path / "decoded.bin")
    np.ascontiguousarray(ch["rc"], np.complex128).tofile(tmp_path / "range_compressed.bin")
    np.ascontiguousarray(ch["slc"], np.complex128).tofile(tmp_path / "slc.bin")
    np.array([spec.v_ms, spec.f_dc], np.float64).tofile(tmp_path / "geometry.bin")
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
    assert r.stdout.count("PASS ") >= 25
