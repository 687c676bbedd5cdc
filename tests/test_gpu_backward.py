"""Gradients of CrossCorr_Norm and SqrDistance (ocipCrossCorr_NormBackward,
ocipSqrDistanceBackward) through the C ABI.

The library's forward and backward entry points are wrapped in a torch.autograd.Function
here, in the test harness only.  Checks:
* small images: against torch float64 autograd of an independent float64 forward
  (torch conv2d on the CPU), for several template shapes including even sizes;
* the 1024x1024 / 31x31 workload: against FP64 central finite differences (h = 1e-3) of an
  FP64 forward (scipy FFT correlation) at 64 seeded coordinates (32 source pixels, 32
  template taps), with L = sum(w1 * NCC) + sum(w2 * SSD), w ~ N(0, 1); and all 16 images'
  full gradients against an FP64 reference gradient (torch float64 autograd through an FFT
  forward), one call per image and op, and in one ocipProximityBackwardBatch call;
* ocipProximityBackwardBatch against the single-image, single-op calls added.
"""
import ctypes

import numpy as np
import pytest

from tests.gpu_util import say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ocip():
    from openclipp_amd import ocip as mod
    ctx = mod.initialize()
    yield mod
    mod.lib.ocipUninitialize(ctx)


def _run(ocip, fwd_name, bwd_name):
    class Op(torch.autograd.Function):
        @staticmethod
        def forward(ctx, s, t):
            S = ocip.Image(np.ascontiguousarray(s.detach().numpy(), dtype=np.float32))
            T = ocip.Image(np.ascontiguousarray(t.detach().numpy(), dtype=np.float32))
            D = ocip.Image(shape=s.shape, dtype=np.float32)
            ocip.check(getattr(ocip.lib, fwd_name)(S.handle, T.handle, D.handle), fwd_name)
            out = torch.from_numpy(D.read().copy())
            ctx.save_for_backward(s, t)
            return out

        @staticmethod
        def backward(ctx, g):
            s, t = ctx.saved_tensors
            S = ocip.Image(np.ascontiguousarray(s.detach().numpy(), dtype=np.float32))
            T = ocip.Image(np.ascontiguousarray(t.detach().numpy(), dtype=np.float32))
            G = ocip.Image(np.ascontiguousarray(g.detach().numpy(), dtype=np.float32))
            dS = ocip.Image(shape=s.shape, dtype=np.float32)
            dT = ocip.Image(shape=t.shape, dtype=np.float32)
            ocip.check(getattr(ocip.lib, bwd_name)(S.handle, T.handle, G.handle, dS.handle,
                                                   dT.handle), bwd_name)
            return torch.from_numpy(dS.read().copy()), torch.from_numpy(dT.read().copy())
    return Op


def ref_maps_torch(s, t):
    """float64 forward: zero padding, template origin at (-tw/2, -th/2)."""
    th, tw = t.shape
    cy, cx = th // 2, tw // 2
    x = torch.nn.functional.pad(s[None, None], (cx, tw - 1 - cx, cy, th - 1 - cy))
    cc = torch.nn.functional.conv2d(x, t[None, None])[0, 0]
    ss = torch.nn.functional.conv2d(x * x, torch.ones_like(t)[None, None])[0, 0]
    tt = (t * t).sum()
    ssd = ss - 2 * cc + tt
    return cc / torch.sqrt(ss * tt), ssd


@pytest.mark.parametrize("tw,th", [(31, 31), (8, 6), (9, 9), (16, 5)])
def test_gradients_match_float64_autograd(ocip, tw, th):
    rng = np.random.default_rng(tw * 10 + th)
    H, W = 70, 90
    s_np = rng.random((H, W)).astype(np.float32)
    t_np = rng.normal(0, 1, (th, tw)).astype(np.float32)
    w1 = rng.normal(0, 1, (H, W))
    w2 = rng.normal(0, 1, (H, W))
    say("autograd", tw, th)
    # library path
    NCC = _run(ocip, "ocipCrossCorr_Norm", "ocipCrossCorr_NormBackward")
    SSD = _run(ocip, "ocipSqrDistance", "ocipSqrDistanceBackward")
    s = torch.tensor(s_np, requires_grad=True)
    t = torch.tensor(t_np, requires_grad=True)
    loss = (torch.from_numpy(w1).float() * NCC.apply(s, t)).sum() + \
           (torch.from_numpy(w2).float() * SSD.apply(s, t)).sum()
    loss.backward()
    # float64 reference
    s64 = torch.tensor(s_np.astype(np.float64), requires_grad=True)
    t64 = torch.tensor(t_np.astype(np.float64), requires_grad=True)
    ncc, ssd = ref_maps_torch(s64, t64)
    (torch.from_numpy(w1) * ncc + torch.from_numpy(w2) * ssd).sum().backward()
    for got, want in [(s.grad, s64.grad), (t.grad, t64.grad)]:
        got = got.double().numpy()
        want = want.numpy()
        scale = np.abs(want).max()
        assert np.max(np.abs(got - want)) <= 2e-4 * scale, (np.max(np.abs(got - want)) / scale)


def test_gradient_of_each_op_separately(ocip):
    """Each backward alone (the other gradient output NULL) equals its share."""
    rng = np.random.default_rng(1)
    s_np = rng.random((40, 50)).astype(np.float32)
    t_np = rng.random((7, 7)).astype(np.float32)
    g = rng.normal(0, 1, (40, 50)).astype(np.float32)
    S, T, G = ocip.Image(s_np), ocip.Image(t_np), ocip.Image(g)
    dS = ocip.Image(shape=(40, 50), dtype=np.float32)
    dT = ocip.Image(shape=(7, 7), dtype=np.float32)
    ocip.check(ocip.lib.ocipSqrDistanceBackward(S.handle, T.handle, G.handle, dS.handle, dT.handle))
    a_s, a_t = dS.read().copy(), dT.read().copy()
    dS2 = ocip.Image(shape=(40, 50), dtype=np.float32)
    ocip.check(ocip.lib.ocipSqrDistanceBackward(S.handle, T.handle, G.handle, dS2.handle, None))
    dT2 = ocip.Image(shape=(7, 7), dtype=np.float32)
    ocip.check(ocip.lib.ocipSqrDistanceBackward(S.handle, T.handle, G.handle, None, dT2.handle))
    np.testing.assert_array_equal(dS2.read(), a_s)
    np.testing.assert_array_equal(dT2.read(), a_t)
    bad = ocip.Image(shape=(7, 8), dtype=np.float32)
    assert ocip.lib.ocipSqrDistanceBackward(S.handle, T.handle, G.handle, None, bad.handle) == \
        ocip.CL_INVALID_IMAGE_SIZE


def fp64_loss(s, t, w1, w2):
    """L = sum(w1 * NCC) + sum(w2 * SSD) in float64 via FFT correlation."""
    import scipy.signal
    th, tw = t.shape
    cy, cx = th // 2, tw // 2
    P = np.pad(s, [(cy, th - 1 - cy), (cx, tw - 1 - cx)])
    cc = scipy.signal.fftconvolve(P, t[::-1, ::-1], mode="valid")
    ss = scipy.signal.fftconvolve(P * P, np.ones_like(t), mode="valid")
    tt = (t * t).sum()
    ncc = cc / np.sqrt(ss * tt)
    ssd = ss - 2 * cc + tt
    return (w1 * ncc).sum() + (w2 * ssd).sum()


def test_diff_1024_finite_differences(ocip):
    from openclipp_amd import workloads
    srcs, tpl, w_ncc, w_ssd = workloads.diff_1024x16()
    # two of the sixteen images: the template gradient is a sum over images
    imgs = [0, 7]
    dT_total = np.zeros(tpl.shape, np.float64)
    dS_all = {}
    for k in imgs:
        say("diff_1024 image", k)
        S, T = ocip.Image(srcs[k]), ocip.Image(tpl)
        dS = ocip.Image(shape=srcs[k].shape, dtype=np.float32)
        dT = ocip.Image(shape=tpl.shape, dtype=np.float32)
        G1, G2 = ocip.Image(w_ncc[k]), ocip.Image(w_ssd[k])
        ocip.check(ocip.lib.ocipCrossCorr_NormBackward(S.handle, T.handle, G1.handle, dS.handle, dT.handle))
        ds1, dt1 = dS.read().copy(), dT.read().copy()
        ocip.check(ocip.lib.ocipSqrDistanceBackward(S.handle, T.handle, G2.handle, dS.handle, dT.handle))
        ds2, dt2 = dS.read().copy(), dT.read().copy()
        dS_all[k] = ds1.astype(np.float64) + ds2
        dT_total += dt1.astype(np.float64) + dt2

    rng = np.random.default_rng(64)
    h = 1e-3
    s64 = {k: srcs[k].astype(np.float64) for k in imgs}
    t64 = tpl.astype(np.float64)

    def total_loss(tt, ss=None):
        return sum(fp64_loss(ss if (ss is not None and k == imgs[0]) else s64[k], tt,
                             w_ncc[k].astype(np.float64), w_ssd[k].astype(np.float64))
                   for k in imgs)

    # 32 template taps: central differences of the summed loss
    for _ in range(32):
        v, u = rng.integers(0, tpl.shape[0]), rng.integers(0, tpl.shape[1])
        tp, tm = t64.copy(), t64.copy()
        tp[v, u] += h
        tm[v, u] -= h
        fd = (total_loss(tp) - total_loss(tm)) / (2 * h)
        got = dT_total[v, u]
        assert abs(got - fd) <= 1e-3 * abs(fd) + 2e-2, (v, u, got, fd)
    # 32 source pixels of the first image (the loss of the other image does not depend on them)
    k = imgs[0]
    for _ in range(32):
        i, j = rng.integers(0, 1024), rng.integers(0, 1024)
        # only outputs within the template footprint change: evaluate a local window
        th, tw = tpl.shape
        y0, y1 = max(0, i - th), min(1024, i + th + 1)
        x0, x1 = max(0, j - tw), min(1024, j + tw + 1)
        r0, r1 = max(0, y0 - th), min(1024, y1 + th)
        c0, c1 = max(0, x0 - tw), min(1024, x1 + tw)

        def local(ds):
            s = s64[k][r0:r1, c0:c1].copy()
            s[i - r0, j - c0] += ds
            full = np.zeros((r1 - r0, c1 - c0))
            w1 = np.zeros_like(full)
            w2 = np.zeros_like(full)
            w1[y0 - r0:y1 - r0, x0 - c0:x1 - c0] = w_ncc[k][y0:y1, x0:x1]
            w2[y0 - r0:y1 - r0, x0 - c0:x1 - c0] = w_ssd[k][y0:y1, x0:x1]
            # windows of outputs inside [y0,y1)x[x0,x1) lie inside the crop unless they touch
            # the image border, where the crop border is the image border too
            return fp64_loss(s, t64, w1, w2)

        fd = (local(h) - local(-h)) / (2 * h)
        got = dS_all[k][i, j]
        assert abs(got - fd) <= 1e-3 * abs(fd) + 2e-3, (i, j, got, fd)


def ref_grad_fft(s_np, t_np, w1, w2):
    """FP64 reference gradient of L = sum(w1 * NCC) + sum(w2 * SSD): torch float64 autograd
    (CPU) through a forward whose correlations are taken with FFTs (zero padding, template
    origin at (-tw/2, -th/2)); independent of the library's gradient formulas."""
    s = torch.tensor(s_np.astype(np.float64), requires_grad=True)
    t = torch.tensor(t_np.astype(np.float64), requires_grad=True)
    th, tw = t.shape
    H, W = s.shape
    cy, cx = th // 2, tw // 2
    P = torch.nn.functional.pad(s, (cx, tw - 1 - cx, cy, th - 1 - cy))
    Hp, Wp = P.shape

    def corr(A, B):
        # circular cross-correlation; rows < H and columns < W never wrap
        FA = torch.fft.rfft2(A)
        FB = torch.fft.rfft2(B, s=(Hp, Wp))
        return torch.fft.irfft2(FA * FB.conj(), s=(Hp, Wp))[:H, :W]

    cc = corr(P, t)
    ss = corr(P * P, torch.ones_like(t))
    tt = (t * t).sum()
    ncc = cc / torch.sqrt(ss * tt)
    ssd = ss - 2 * cc + tt
    ((torch.from_numpy(w1.astype(np.float64)) * ncc).sum() +
     (torch.from_numpy(w2.astype(np.float64)) * ssd).sum()).backward()
    return s.grad.numpy(), t.grad.numpy()


def test_diff_1024_all_images_against_fp64_reference(ocip):
    """All 16 images of the 1024x1024 / 31x31 workload: every full source-gradient map, and
    the template gradient summed over the batch, within 2e-4 of the largest component of the
    FP64 reference gradient."""
    from openclipp_amd import workloads
    srcs, tpl, w_ncc, w_ssd = workloads.diff_1024x16()
    T = ocip.Image(tpl)
    dS = ocip.Image(shape=srcs[0].shape, dtype=np.float32)
    dT = ocip.Image(shape=tpl.shape, dtype=np.float32)
    dT_got = np.zeros(tpl.shape, np.float64)
    dT_want = np.zeros(tpl.shape, np.float64)
    worst_s = 0.0
    for k in range(len(srcs)):
        say("diff_1024 fp64 reference, image", k)
        S, G1, G2 = ocip.Image(srcs[k]), ocip.Image(w_ncc[k]), ocip.Image(w_ssd[k])
        ocip.check(ocip.lib.ocipCrossCorr_NormBackward(S.handle, T.handle, G1.handle, dS.handle,
                                                       dT.handle))
        ds, dt = dS.read().astype(np.float64), dT.read().astype(np.float64)
        ocip.check(ocip.lib.ocipSqrDistanceBackward(S.handle, T.handle, G2.handle, dS.handle,
                                                    dT.handle))
        ds += dS.read()
        dT_got += dt + dT.read()
        want_s, want_t = ref_grad_fft(srcs[k], tpl, w_ncc[k], w_ssd[k])
        dT_want += want_t
        err = np.max(np.abs(ds - want_s)) / np.abs(want_s).max()
        worst_s = max(worst_s, float(err))
        assert err <= 2e-4, (k, err)
        for i in (S, G1, G2):
            i.release()
    err_t = np.max(np.abs(dT_got - dT_want)) / np.abs(dT_want).max()
    say("diff_1024 relative error: dS", worst_s, "dT", float(err_t))
    assert err_t <= 2e-4


def _batch_backward(ocip, srcs, tpl, g_ncc, g_ssd, want_ds=True, want_dt=True):
    n = len(srcs)
    Ss = [ocip.Image(s) for s in srcs]
    T = ocip.Image(tpl)
    G1 = [ocip.Image(g) for g in g_ncc] if g_ncc is not None else None
    G2 = [ocip.Image(g) for g in g_ssd] if g_ssd is not None else None
    dS = [ocip.Image(shape=srcs[0].shape, dtype=np.float32) for _ in range(n)] if want_ds else None
    dT = ocip.Image(shape=tpl.shape, dtype=np.float32) if want_dt else None
    ocip.check(ocip.lib.ocipProximityBackwardBatch(
        ocip.handles(Ss), T.handle, ocip.handles(G1) if G1 else None,
        ocip.handles(G2) if G2 else None, ocip.handles(dS) if dS else None,
        dT.handle if dT else None, n), "ocipProximityBackwardBatch")
    ds = [d.read().astype(np.float64) for d in dS] if dS else None
    dt = dT.read().astype(np.float64) if dT else None
    return ds, dt


def _single_backward(ocip, src, tpl, g, fn):
    S, T, G = ocip.Image(src), ocip.Image(tpl), ocip.Image(g)
    dS = ocip.Image(shape=src.shape, dtype=np.float32)
    dT = ocip.Image(shape=tpl.shape, dtype=np.float32)
    ocip.check(fn(S.handle, T.handle, G.handle, dS.handle, dT.handle))
    return dS.read().astype(np.float64), dT.read().astype(np.float64)


@pytest.mark.parametrize("tw,th", [(9, 7), (31, 31), (40, 36)])
def test_batch_backward_equals_single_calls(ocip, tw, th):
    """ocipProximityBackwardBatch over 3 images, both ops: dS_i and the batch-summed dT equal
    the single-image, single-op calls added, within the rounding of one combined map against
    two: 5e-5 of the largest component for dS (split-BF16 correlations, whose products carry
    up to 2^-14 relative error), 1e-5 for dT on the FP32 contraction (40 x 36: FP32 partial
    sums of 4096 products per tile), 5e-5 for dT on the split-BF16 contraction (up to
    32 x 32); with one op absent (GradNcc or GradSsd NULL) the same."""
    rng = np.random.default_rng(tw * 100 + th)
    n, H, W = 3, 130, 170
    srcs = [rng.random((H, W)).astype(np.float32) for _ in range(n)]
    tpl = rng.normal(0, 1, (th, tw)).astype(np.float32)
    g1 = [rng.normal(0, 1, (H, W)).astype(np.float32) for _ in range(n)]
    g2 = [rng.normal(0, 1, (H, W)).astype(np.float32) for _ in range(n)]
    singles = [(_single_backward(ocip, srcs[i], tpl, g1[i], ocip.lib.ocipCrossCorr_NormBackward),
                _single_backward(ocip, srcs[i], tpl, g2[i], ocip.lib.ocipSqrDistanceBackward))
               for i in range(n)]
    for which in ("both", "ncc", "ssd"):
        say("batch backward", tw, th, which)
        ds, dt = _batch_backward(ocip, srcs, tpl, g1 if which != "ssd" else None,
                                 g2 if which != "ncc" else None)
        want_t = sum((a[1] if which != "ssd" else 0) + (b[1] if which != "ncc" else 0)
                     for a, b in singles)
        tol_t = 5e-5 if tw <= 32 and th <= 32 else 1e-5
        assert np.max(np.abs(dt - want_t)) <= tol_t * np.abs(want_t).max(), which
        for i in range(n):
            a, b = singles[i]
            want_s = (a[0] if which != "ssd" else 0) + (b[0] if which != "ncc" else 0)
            assert np.max(np.abs(ds[i] - want_s)) <= 5e-5 * np.abs(want_s).max(), (which, i)


def test_batch_backward_chunks_and_missing_outputs(ocip):
    """19 images (two chunks of at most 16 images per correlation launch), source gradients
    requested for every image but the multiples of 3: each requested dS_i and the summed dT
    equal the single-image calls added (tolerances as above), the other dS images are left
    untouched."""
    rng = np.random.default_rng(19)
    n, H, W, th, tw = 19, 61, 75, 7, 10
    srcs = [rng.random((H, W)).astype(np.float32) for _ in range(n)]
    tpl = rng.normal(0, 1, (th, tw)).astype(np.float32)
    g1 = [rng.normal(0, 1, (H, W)).astype(np.float32) for _ in range(n)]
    g2 = [rng.normal(0, 1, (H, W)).astype(np.float32) for _ in range(n)]
    Ss = [ocip.Image(x) for x in srcs]
    T = ocip.Image(tpl)
    G1 = [ocip.Image(x) for x in g1]
    G2 = [ocip.Image(x) for x in g2]
    fill = np.full((H, W), 7.0, np.float32)
    dS = [ocip.Image(fill.copy()).send() for _ in range(n)]
    dT = ocip.Image(shape=tpl.shape, dtype=np.float32)
    hds = (ctypes.c_void_p * n)(*[None if i % 3 == 0 else dS[i].handle.value for i in range(n)])
    say("batch backward chunks", n)
    ocip.check(ocip.lib.ocipProximityBackwardBatch(ocip.handles(Ss), T.handle, ocip.handles(G1),
                                                   ocip.handles(G2), hds, dT.handle, n))
    want_t = np.zeros(tpl.shape)
    for i in range(n):
        a = _single_backward(ocip, srcs[i], tpl, g1[i], ocip.lib.ocipCrossCorr_NormBackward)
        b = _single_backward(ocip, srcs[i], tpl, g2[i], ocip.lib.ocipSqrDistanceBackward)
        want_t += a[1] + b[1]
        got = dS[i].read()
        if i % 3 == 0:
            np.testing.assert_array_equal(got, fill)
        else:
            want_s = a[0] + b[0]
            assert np.max(np.abs(got - want_s)) <= 5e-5 * np.abs(want_s).max(), i
    assert np.max(np.abs(dT.read() - want_t)) <= 5e-5 * np.abs(want_t).max()


def test_batch_backward_large_template_against_fp64_reference(ocip):
    """A 130 x 120 template, whose window sums come from FP64 integrals (too large for the
    box-sum tile) and whose template gradient from the FP32 contraction: two images in one
    ocipProximityBackwardBatch call, every dS_i and the summed dT within 2e-4 of the largest
    component of the FP64 reference gradient."""
    rng = np.random.default_rng(130)
    n, H, W, th, tw = 2, 200, 260, 120, 130
    srcs = [rng.random((H, W)).astype(np.float32) for _ in range(n)]
    tpl = rng.normal(0, 1, (th, tw)).astype(np.float32)
    g1 = [rng.normal(0, 1, (H, W)).astype(np.float32) for _ in range(n)]
    g2 = [rng.normal(0, 1, (H, W)).astype(np.float32) for _ in range(n)]
    say("batch backward large template", tw, th)
    ds, dt = _batch_backward(ocip, srcs, tpl, g1, g2)
    want_t = np.zeros(tpl.shape)
    for k in range(n):
        want_s, wt = ref_grad_fft(srcs[k], tpl, g1[k], g2[k])
        want_t += wt
        assert np.max(np.abs(ds[k] - want_s)) / np.abs(want_s).max() <= 2e-4, k
    assert np.max(np.abs(dt - want_t)) / np.abs(want_t).max() <= 2e-4


def test_batch_backward_diff_1024_against_fp64_reference(ocip):
    """The whole 16 x 1024x1024 / 31x31 workload in one ocipProximityBackwardBatch call (both
    ops): every dS_i and the summed dT within 2e-4 of the largest component of the FP64
    reference gradient of L = sum w1 NCC + sum w2 SSD."""
    from openclipp_amd import workloads
    srcs, tpl, w_ncc, w_ssd = workloads.diff_1024x16()
    say("batch backward diff_1024")
    ds, dt = _batch_backward(ocip, list(srcs), tpl, list(w_ncc), list(w_ssd))
    want_t = np.zeros(tpl.shape)
    for k in range(len(srcs)):
        want_s, wt = ref_grad_fft(srcs[k], tpl, w_ncc[k], w_ssd[k])
        want_t += wt
        err = np.max(np.abs(ds[k] - want_s)) / np.abs(want_s).max()
        assert err <= 2e-4, (k, err)
    assert np.max(np.abs(dt - want_t)) / np.abs(want_t).max() <= 2e-4


def test_batch_backward_errors(ocip):
    src = np.zeros((40, 50), np.float32)
    S, T = ocip.Image(src), ocip.Image(np.ones((5, 5), np.float32))
    G = ocip.Image(src)
    Gbad = ocip.Image(np.zeros((40, 51), np.float32))
    L = ocip.lib.ocipProximityBackwardBatch
    assert L(ocip.handles([S]), T.handle, None, None, None, None, 1) == ocip.CL_INVALID_VALUE
    assert L(ocip.handles([S]), T.handle, ocip.handles([Gbad]), None, None, None, 1) == \
        ocip.CL_INVALID_IMAGE_SIZE
    assert L(None, T.handle, None, None, None, None, 0) == ocip.CL_SUCCESS
    assert L(ocip.handles([S]), T.handle, ocip.handles([G]), None, None, None, 1) == ocip.CL_SUCCESS
