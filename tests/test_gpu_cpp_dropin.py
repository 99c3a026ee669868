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
    _write_level0_case(tmp_path)
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
    assert r.stdout.count("PASS ") >= 40


def _write_level0_case(d):
    """A Level-0 file (IW1/IW2/IW3 bursts, a partial lead burst) for
    S1_Decoder, and the oracle's decode, range-Doppler data and SLC of IW2
    burst 1 with the effective velocities the GPU product derives from the
    file's state vectors."""
    import types

    import torch
    import oracle
    from chain_ref import oracle_chain, oracle_tops
    from sentinel1_decode_amd import capi, pipeline, synth
    specs = []
    for k in range(2):
        for sw, rg, pri in ((10, 6, 22276), (11, 8, 21861), (12, 11, 25868)):
            specs.append(synth.BurstSpec(n_lines=1050, n_samples=768, rgdec=rg, pri_code=pri, n_targets=6,
                                         swath_number=sw, seed=300 + 3 * k + sw, clutter_sigma=0.1))
    s = synth.level0_stream(specs, lead_packets=200)
    s.stream.tofile(d / "level0.dat")
    ctx = capi.Context(0)
    prod = pipeline.Level0Product(s.stream, ctx)
    idx = prod.packet_indices("IW2", 1)
    v = prod.burst("IW2", 1).geo.v_rel
    offs = np.asarray(oracle.index_level0(s.stream), np.uint64)[idx] + 68
    F = oracle.header_fields(s.stream, [int(o) - 68 for o in offs])
    spec = [sp for swn, first, n, sp in s.bursts if first == idx[0]][0]
    ob = types.SimpleNamespace(stream=s.stream, ud_offsets=offs, ud_lengths=(F[:, 7] + 1 - 62).astype(np.uint32),
                               num_quads=F[:, oracle.SEC["num_quadratures"]].astype(np.int32), spec=spec)
    ref = oracle_chain(ob, v_rel=v, f_dc=0.0)
    np.ascontiguousarray(ref["decoded"], np.complex128).tofile(d / "l0_decoded.bin")
    np.ascontiguousarray(ref["rd"], np.complex128).tofile(d / "l0_rd.bin")
    np.ascontiguousarray(ref["slc"], np.complex128).tofile(d / "l0_slc.bin")
    np.ascontiguousarray(v, np.float64).tofile(d / "l0_vrel.bin")
    # the TOPS branch that get_azimuth_compressed_burst takes
    tops = oracle_tops(ob, F, oracle.state_vectors(oracle.header_fields(s.stream, oracle.index_level0(s.stream))))
    np.ascontiguousarray(tops["slc"], np.complex128).tofile(d / "l0_tops.bin")
    np.asarray(tops["slc"].shape, np.int64).tofile(d / "l0_tops_shape.bin")
    counts = [prod.swaths()[n] for n in ("IW1", "IW2", "IW3")] + [prod.n_bursts(n) for n in ("IW1", "IW2", "IW3")]
    np.asarray(counts + [len(prod.state_vectors)], np.int64).tofile(d / "l0_counts.bin")
    ctx.close()
