"""s1g_rcmc_matched_filter_noalloc (no device allocation; column table in a
fixed device array) against s1g_rcmc_matched_filter, through the C ABI:
bit-equal outputs for every launch variant of both precisions, on two calls
in a row, the second with a smaller burst; sizes above the stated maximum
are rejected."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from sentinel1_decode_amd import capi, pipeline, synth  # noqa: E402

F32, F64 = capi.S1G_F32, capi.S1G_F64


def noalloc(ctx, x, desc, out):
    L = capi.lib()
    f = L.s1g_rcmc_matched_filter_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.POINTER(capi.AzimuthDesc), C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]
    ctx._bind_stream()
    rc = f(ctx.h, C.byref(desc), capi._prec_of(x), C.c_void_p(x.data_ptr()), x.stride(0),
           C.c_void_p(out.data_ptr()), out.stride(0))
    return rc


@pytest.fixture(scope="module")
def two_bursts():
    big = synth.config1(n_lines=160, n_samples=6000, n_targets=30, seed=61)
    small = synth.config0(n_lines=96, n_samples=2048, grid_spacing=(300, 20), f_dc=12.0, seed=62)
    return [pipeline.from_synth(synth.make_burst(s)) for s in (big, small)]


@pytest.mark.parametrize("prec", [F32, F64])
def test_noalloc_matches_allocating_entry_point(ctx, two_bursts, prec):
    dt = torch.complex64 if prec == F32 else torch.complex128
    nvar = capi.lib().s1g_rcmc_matched_filter  # noqa: F841 (loaded)
    rng = np.random.default_rng(3)
    data = []
    for db in two_bursts:
        x = rng.standard_normal((db.n_lines, db.n_samples)) + 1j * rng.standard_normal((db.n_lines, db.n_samples))
        data.append(torch.from_numpy(x).to(dt).cuda())
    n_variants = 7 if prec == F32 else 3
    for v in range(n_variants):
        ctx.force_variant("rcmc_mf", v)
        try:
            # two calls in a row, the second on the smaller burst
            for db, x in zip(two_bursts, data):
                desc = db.azimuth_desc()
                want = ctx.rcmc_matched_filter(x, desc)
                got = torch.full_like(x, complex(np.nan, np.nan))
                assert noalloc(ctx, x, desc, got) == 0, capi.lib().s1g_last_error(ctx.h)
                torch.cuda.synchronize()
                assert torch.equal(got, want), (v, db.n_lines, db.n_samples)
                assert ctx.last_variant("rcmc_mf")[0] == v
        finally:
            ctx.force_variant("rcmc_mf", -1)


def test_noalloc_rejects_above_maximum(ctx, two_bursts):
    db = two_bursts[1]
    desc = db.azimuth_desc()
    desc.n_samples = 65537
    x = torch.zeros((db.n_lines, 65537), dtype=torch.complex64, device="cuda")
    out = torch.zeros_like(x)
    assert noalloc(ctx, x, desc, out) == -1
    assert b"no-alloc maximum" in capi.lib().s1g_last_error(ctx.h)
