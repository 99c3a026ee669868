"""Entropy-minimising autofocus through the hand-written adjoint.

The adjoint (s1g_af_adjoint) is wrapped in a torch.autograd.Function here,
in the test harness only, and driven by torch.optim.LBFGS.  A synthetic
burst is focused up to the matched filter (the range-Doppler spectrum is the
checkpoint: it is computed once and every entropy / gradient evaluation
recomputes the image from it); a known azimuth phase error -- quadratic,
2 rad at the aperture edge, plus a seeded Gaussian random walk of
0.05 rad per Doppler bin -- is applied per Doppler row.  Entropy is blind to
a constant phase and to a linear phase (an azimuth shift), so the residual
is measured after removing its best-fit constant and linear terms.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from sentinel1_decode_amd import capi, pipeline, synth  # noqa: E402

F32 = capi.S1G_F32


class ImageEntropy(torch.autograd.Function):
    """E(phase) of the image IFFT_az(rd e^{i phase}) / L; the gradient is the
    library's hand-written adjoint."""

    @staticmethod
    def forward(fctx, phase, rd, s1ctx):
        g = torch.zeros_like(phase)
        e = s1ctx.af_adjoint(rd, phase.detach().contiguous(), g)
        fctx.save_for_backward(g)
        return torch.tensor(e, dtype=phase.dtype, device=phase.device)

    @staticmethod
    def backward(fctx, grad_out):
        (g,) = fctx.saved_tensors
        return grad_out * g, None, None


def phase_error(L, seed, edge_rad=2.0, walk_rad=0.05):
    rng = np.random.default_rng(seed)
    f = np.linspace(-1.0, 1.0, L)
    return edge_rad * f ** 2 + np.cumsum(rng.normal(0.0, walk_rad, L))


def detrended_rms(x, band=0.75):
    """RMS after removing the best-fit constant + linear phase, over the
    Doppler bins |f| <= band * PRF/2 (the scene's Doppler band holds the
    energy that determines the phase)."""
    i = np.arange(x.size, dtype=np.float64)
    keep = np.abs(np.linspace(-1.0, 1.0, x.size)) <= band
    a = np.vstack([np.ones_like(i), i]).T[keep]
    coef, *_ = np.linalg.lstsq(a, x[keep], rcond=None)
    r = x[keep] - a @ coef
    return float(np.sqrt(np.mean(r ** 2)))


def range_doppler(ctx, db, prec):
    """decode -> range compression -> azimuth FFT -> RCMC + matched filter."""
    dec, _ = db.decode(ctx, prec)
    rc = ctx.range_compress(dec, db.geo.range_desc())
    ctx.azimuth_fft(rc, inverse=False)
    return ctx.rcmc_matched_filter(rc, db.azimuth_desc())


def autofocus(ctx, rd_err, iters, history=10):
    L = rd_err.shape[0]
    phase = torch.zeros(L, dtype=torch.float64, device=rd_err.device, requires_grad=True)
    opt = torch.optim.LBFGS([phase], lr=1.0, max_iter=iters, history_size=history,
                            line_search_fn="strong_wolfe")

    def closure():
        opt.zero_grad()
        e = ImageEntropy.apply(phase, rd_err, ctx)
        e.backward()
        return e

    opt.step(closure)
    return phase.detach()


def test_autofocus_recovers_phase_error(ctx):
    """The error-free burst is autofocused too: in Doppler bins with little
    scene energy, per-bin phases are fitted to speckle, so the recovered
    error is judged against that solution (the differential residual), and
    against the injected error itself over the central, well-illuminated
    quarter of the band."""
    spec = synth.config4(n_lines=1024, n_samples=2048, n_targets=60, seed=41)
    db = pipeline.from_synth(synth.make_burst(spec))
    rd = range_doppler(ctx, db, F32)
    L = rd.shape[0]
    err = phase_error(L, seed=7)
    rot = torch.from_numpy(np.exp(1j * err)).to(rd.dtype).cuda()
    rd_err = (rd * rot[:, None]).contiguous()
    zero = torch.zeros(L, dtype=torch.float64, device="cuda")
    e0 = ctx.af_forward(rd_err, zero)
    est_clean = autofocus(ctx, rd, iters=30).cpu().numpy()
    est = autofocus(ctx, rd_err, iters=30)
    e1 = ctx.af_forward(rd_err, est)
    est = est.cpu().numpy()
    torch.cuda.synchronize()
    diff = detrended_rms(err + est - est_clean)
    print(f"entropy {e0:.5f} -> {e1:.5f}; residual phase RMS (band 0.75) {detrended_rms(err):.3f} -> "
          f"{detrended_rms(err + est):.3f}, differential {diff:.4f}; central quarter "
          f"{detrended_rms(err, 0.25):.3f} -> {detrended_rms(err + est, 0.25):.3f} rad")
    assert e1 < e0
    assert diff < 0.1 * detrended_rms(err)
    assert detrended_rms(err + est, 0.25) < 0.5 * detrended_rms(err, 0.25)
