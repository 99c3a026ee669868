"""The output side without a GPU: the BigTIFF writer (read back with a
minimal BigTIFF parser written here from the format's layout) and the
oracle's display scaling against its identities."""
import struct

import numpy as np
import pytest

import oracle
from sentinel1_decode_amd import capi


def read_bigtiff(path):
    """(tags, image) of a single-strip little-endian BigTIFF of floats."""
    b = open(path, "rb").read()
    assert b[:2] == b"II"
    version, offsize, _, ifd = struct.unpack_from("<HHHQ", b, 2)
    assert (version, offsize) == (43, 8)
    (n,) = struct.unpack_from("<Q", b, ifd)
    tags = {}
    for k in range(n):
        tag, typ, count, value = struct.unpack_from("<HHQQ", b, ifd + 8 + 20 * k)
        tags[tag] = (value & 0xFFFF) if typ == 3 else value
    (nxt,) = struct.unpack_from("<Q", b, ifd + 8 + 20 * n)
    assert nxt == 0
    w, h = tags[256], tags[257]
    img = np.frombuffer(b, np.float32, count=w * h, offset=tags[273]).reshape(h, w)
    return tags, img


def test_write_tiff_layout(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.standard_normal((7, 13)).astype(np.float32)
    p = tmp_path / "x.tif"
    capi.write_tiff(p, img)
    tags, back = read_bigtiff(p)
    assert tags[258] == 32 and tags[339] == 3 and tags[277] == 1 and tags[259] == 1 and tags[262] == 1
    assert tags[279] == img.nbytes and tags[278] == 7
    # the reference writes the last data row first (image_write.cpp:19-24)
    assert np.array_equal(back, img[::-1])
    capi.write_tiff(p, img, flip_rows=False)
    assert np.array_equal(read_bigtiff(p)[1], img)


def test_oracle_scaling_identities():
    rng = np.random.default_rng(1)
    x = (rng.standard_normal((5, 40)) + 1j * rng.standard_normal((5, 40))) * 3
    a = np.abs(x)
    k = np.unravel_index(np.argmax(a), a.shape)
    n = oracle.scale_2d(x, "norm")
    assert n[k] == np.float32(1 + 1e-12 / a[k]) and n.max() == n[k] and (n > 0).all()
    lg = oracle.scale_2d(x, "norm_log")
    assert abs(lg[k]) < 1e-9 and (lg <= 1e-9).all()
    assert np.allclose(10 ** (lg / 20), n, rtol=1e-6)
    assert np.allclose(oracle.scale_2d(x, "mag"), x.real ** 2 + x.imag ** 2, rtol=1e-7)
    assert np.array_equal(oracle.scale_2d(x, "real"), x.real.astype(np.float32))
    assert np.array_equal(oracle.scale_2d(x, "imag"), x.imag.astype(np.float32))
