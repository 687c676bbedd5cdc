"""Template-matching throughput on one MI355X, through the C ABI of libOpenCLIPP.so.

Workload (one step = one ocipCrossCorr_NormEx call on inputs already resident in HBM):
4096x4096 F32 image, i.i.d. N(0.5, 0.2^2) clamped to [0,1]; 64x64 F32 template cropped at a
seeded offset and scaled by 1.1; normalised cross-correlation map (4096x4096 F32).  A step
includes everything the call does: template staging, the FP64 sum-of-squares integral that
feeds the denominators, and the correlation itself.

Metric: Gpix*tpix/s = W*H*tw*th / seconds / 1e9.  The headline uses BF16 MFMA products with
FP32 accumulation; the plain FP32 VALU path is measured beside it on the same inputs, and the
maximum / mean absolute NCC error of both against an FP64 evaluation at sampled pixels is
reported.

Roofline: the dominant kernel's algorithmic FLOPs (2 per template pixel per output pixel)
divided by its average duration, timed with HIP events on the library's own stream
(ocipEnableKernelTiming) over a second run of the timed steps.  Its `traffic` is the HBM bytes per launch
of that kernel from two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) on this command,
calibrated on a device copy of known size (--copy-calibration), kept in
profiles/crosscorr_norm_4096_t64_traffic.md.

CPU baseline: the reference's CrossCorr_Norm on this host's cores (oracle/baseline.c: the
reference's per-output order and F32 arithmetic, adjacent outputs in SIMD lanes as a CPU OpenCL
runtime runs work-items, OpenMP over rows), on the whole map of the same workload.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"bf16": 2500.0, "f16": 2500.0, "bf16x3": 2500.0, "f32": 157.3}
PREC_CODE = {"f32": 0, "bf16": 1, "f16": 2, "bf16x3": 3}
DOMINANT = {"f32": "prox_tile_f32", "bf16": "crosscorr_mfma", "f16": "crosscorr_mfma",
            "bf16x3": "crosscorr_mfma"}


def ncc_f64_at(src, tpl, xs, ys):
    """FP64 normalised cross-correlation at the given pixels (zero-padded, centred)."""
    th, tw = tpl.shape
    cy, cx = th // 2, tw // 2
    P = np.pad(src.astype(np.float64), [(cy, th - 1 - cy), (cx, tw - 1 - cx)])
    t = tpl.astype(np.float64)
    tt = (t * t).sum()
    out = np.empty(len(xs))
    for k in range(0, len(xs), 256):
        xb, yb = xs[k:k + 256], ys[k:k + 256]
        win = np.stack([P[y:y + th, x:x + tw] for x, y in zip(xb, yb)])
        out[k:k + 256] = (win * t).sum(axis=(1, 2)) / np.sqrt((win * win).sum(axis=(1, 2)) * tt)
    return out


def time_steps(ocip, fn, steps, warmup, kernel):
    """ms per step over `steps` calls (wall clock between two ocipFinish), then the average
    duration of `kernel` from HIP events recorded on the library's stream around each launch
    over a second run of the same steps (kept out of the first run: the event records sit
    between the kernels of a step)."""
    for _ in range(warmup):
        ocip.check(fn())
    ocip.check(ocip.lib.ocipFinish())
    t0 = time.perf_counter()
    for _ in range(steps):
        ocip.check(fn())
    ocip.check(ocip.lib.ocipFinish())
    t1 = time.perf_counter()
    ocip.check(ocip.lib.ocipEnableKernelTiming(b"\x01"))
    for _ in range(steps):
        ocip.check(fn())
    ocip.check(ocip.lib.ocipFinish())
    kms, kn = ocip.kernel_time(kernel)
    ocip.check(ocip.lib.ocipEnableKernelTiming(b"\x00"))
    return (t1 - t0) * 1e3 / steps, (kms / kn if kn else None)


def batch_bench(ocip, steps):
    """256 templates with sides 8..96 against one 2048x2048 F32 image, CrossCorr_Norm, one
    ocipCrossCorr_NormBatch call per step (BF16 MFMA, one recorded graph), next to the plain
    FP32 path on the same batch."""
    from openclipp_amd import workloads
    src, tpls, _ = workloads.batch_2048()
    H, W = src.shape
    tpix = sum(W * H * t.shape[0] * t.shape[1] for t in tpls)
    S = ocip.Image(src).send()
    Ts = [ocip.Image(t).send() for t in tpls]
    Ds = [ocip.Image(shape=(H, W), dtype=np.float32, device_only=True) for _ in tpls]
    th, dh = ocip.handles(Ts), ocip.handles(Ds)
    out = {"workload": f"CrossCorr_Norm, {W}x{H} F32 U[0,1), {len(tpls)} templates 8..96 px",
           "tpix_per_step": tpix}
    for name, prec, n in (("bf16", 1, steps), ("f32", 0, 1)):
        fn = lambda: ocip.lib.ocipCrossCorr_NormBatch(S.handle, th, dh, len(Ts), prec)
        ms, kms = time_steps(ocip, fn, n, 1, "crosscorr_mfma_batch" if prec else "prox_tile_f32")
        out[name] = {"value": round(tpix / (ms * 1e-3) / 1e9, 3), "ms_per_step": round(ms, 3)}
        if prec:
            ach = 2.0 * tpix / (ms * 1e-3) / 1e12
            out[name]["achieved_tflops"] = round(ach, 1)
            out[name]["frac_of_bf16_peak"] = round(ach / PEAK_TFLOPS["bf16"], 4)
    out["speedup_bf16_vs_f32"] = round(out["bf16"]["value"] / out["f32"]["value"], 3)
    # the same batch, SqrDistance_Norm (BF16 MFMA, SS - 2 CC + TT)
    fn = lambda: ocip.lib.ocipSqrDistance_NormBatch(S.handle, th, dh, len(Ts), 1)
    ms, _ = time_steps(ocip, fn, steps, 1, "crosscorr_mfma_batch")
    out["ssd_norm_bf16"] = {"value": round(tpix / (ms * 1e-3) / 1e9, 3), "ms_per_step": round(ms, 3)}
    for i in [S] + Ts + Ds:
        i.release()
    return out


def fft_bench(ocip, src, steps):
    """The FFT path (ocipCrossCorrFFT_Norm, FP64 transforms) on the 4096x4096 image with the
    64x64 template, and a 256x256 crop of the same image against the BF16 MFMA direct path
    (the FFT's cost does not depend on the template size, the direct path's grows with it)."""
    import ctypes
    H, W = src.shape
    out = {}
    S = ocip.Image(src).send()
    D = ocip.Image(shape=(H, W), dtype=np.float32, device_only=True)
    for name, (th, tw) in (("t64", (64, 64)), ("t256", (256, 256))):
        tpl = np.ascontiguousarray(src[1000:1000 + th, 2000:2000 + tw])
        T = ocip.Image(tpl).send()
        prog = ctypes.c_void_p()
        ocip.check(ocip.lib.ocipPrepareImageProximityFFT(ctypes.byref(prog), S.handle, T.handle))
        tpix = W * H * tw * th
        fn = lambda: ocip.lib.ocipCrossCorrFFT_Norm(prog, S.handle, T.handle, D.handle)
        ms, _ = time_steps(ocip, fn, steps, 1, "fft_cols_pair")
        row = {"fft_value": round(tpix / (ms * 1e-3) / 1e9, 3), "fft_ms_per_step": round(ms, 3)}
        if name == "t256":
            fn2 = lambda: ocip.lib.ocipCrossCorr_NormEx(S.handle, T.handle, D.handle, 1)
            ms2, _ = time_steps(ocip, fn2, max(2, steps // 2), 1, "crosscorr_mfma")
            row.update({"mfma_bf16_value": round(tpix / (ms2 * 1e-3) / 1e9, 3),
                        "mfma_bf16_ms_per_step": round(ms2, 3)})
        out[name] = row
        ocip.check(ocip.lib.ocipReleaseProgram(prog))
        T.release()
    S.release()
    D.release()
    return out


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "crosscorr_norm_4096_t64_traffic.md")


def configs_bench(ocip, steps):
    """The other template-matching workloads, each timed like the headline (wall clock over
    `steps` calls between two ocipFinish, kernel averages from HIP events in a second run):
    the 1024^2 F32 / 16x16 plain path (SqrDistance and CrossCorr_Norm; for a template this
    small the kernel is bounded by HBM, so its achieved GB/s are given with the algorithmic
    bytes source + map + template), the 4096^2 U8 / 32x32 integer workload (all five ops,
    exact integer sums), and the gradients of the 16 x 1024^2 / 31x31 workload (one
    ocipProximityBackwardBatch call per step: both ops' gradients for the 16 images)."""
    from openclipp_amd import workloads
    out = {}

    src, tpl, _ = workloads.plain_1024()
    H, W = src.shape
    th, tw = tpl.shape
    S, T = ocip.Image(src).send(), ocip.Image(tpl).send()
    D = ocip.Image(shape=(H, W), dtype=np.float32)
    plain = {"workload": f"{W}x{H} F32 U[0,1), {tw}x{th} crop + N(0, 0.01^2), plain F32 path"}
    for op in ("SqrDistance", "CrossCorr_Norm"):
        ms, kms = time_steps(ocip, lambda: ocip.proximity(op, S, T, D), steps, 2, "prox_tile_f32")
        alg = W * H * 4 * 2 + tw * th * 4
        plain[op] = {"ms_per_step": round(ms, 4), "kernel_ms": round(kms, 4) if kms else None,
                     "gpix_tpix_per_s": round(W * H * tw * th / (ms * 1e-3) / 1e9, 1),
                     "kernel_hbm_gbps": round(alg / (kms * 1e-3) / 1e9, 1) if kms else None,
                     "hbm_peak_gbps": 8000.0}
    out["plain_1024"] = plain
    for i in (S, T, D):
        i.release()

    src, tpl, _ = workloads.integer_4096()
    H, W = src.shape
    th, tw = tpl.shape
    S, T = ocip.Image(src).send(), ocip.Image(tpl).send()
    D = ocip.Image(shape=(H, W), dtype=np.float32)
    integer = {"workload": f"{W}x{H} U8 blurred field, {tw}x{th} crop + U{{-3..3}}, exact integer sums"}
    for op in ocip.PROXIMITY_OPS:
        # the exact-integer MFMA kernel (BF16 products) for the cross-term ops, the VALU kernel
        # for AbsDistance
        kname = "prox_tile_int" if op == "AbsDistance" else "crosscorr_mfma_int"
        ms, kms = time_steps(ocip, lambda: ocip.proximity(op, S, T, D), max(3, steps // 2), 1,
                             kname)
        row = {"ms_per_step": round(ms, 4), "kernel": kname,
               "kernel_ms": round(kms, 4) if kms else None,
               "gpix_tpix_per_s": round(W * H * tw * th / (ms * 1e-3) / 1e9, 1)}
        if kms and op != "AbsDistance":
            # the kernel's algorithmic work is the cross term, 2 ops per tpix, on BF16 MFMA
            ach = 2.0 * W * H * tw * th / (kms * 1e-3) / 1e12
            row["roofline"] = {"bound": "mfma", "achieved": round(ach, 1),
                               "peak": PEAK_TFLOPS["bf16"], "unit": "TFLOP/s",
                               "frac": round(ach / PEAK_TFLOPS["bf16"], 4)}
        elif kms:
            # v_sad_u8: 4 taps (3 ops each) per lane and instruction at the non-packed VALU
            # instruction rate (half the 157.3 TFLOP/s packed-FMA figure)
            ach = 3.0 * W * H * tw * th / (kms * 1e-3) / 1e12
            peak_sad = 3 * 4 * PEAK_TFLOPS["f32"] / 2
            row["roofline"] = {"bound": "valu", "achieved": round(ach, 1),
                               "peak": round(peak_sad, 1), "unit": "TOP/s",
                               "frac": round(ach / peak_sad, 4)}
        integer[op] = row
    out["integer_4096"] = integer
    for i in (S, T, D):
        i.release()

    srcs, tpl, w_ncc, w_ssd = workloads.diff_1024x16()
    n, H, W = srcs.shape
    th, tw = tpl.shape
    T = ocip.Image(tpl).send()
    Ss = [ocip.Image(srcs[k]).send() for k in range(n)]
    G1 = [ocip.Image(w_ncc[k]).send() for k in range(n)]
    G2 = [ocip.Image(w_ssd[k]).send() for k in range(n)]
    dSs = [ocip.Image(shape=(H, W), dtype=np.float32, device_only=True) for _ in range(n)]
    dT = ocip.Image(shape=(th, tw), dtype=np.float32)
    hS, h1, h2, hdS = ocip.handles(Ss), ocip.handles(G1), ocip.handles(G2), ocip.handles(dSs)

    def grads():
        # one call: both ops' gradients for the 16 images, dT summed over the batch
        return ocip.lib.ocipProximityBackwardBatch(hS, T.handle, h1, h2, hdS, dT.handle, n)

    ms, kms = time_steps(ocip, grads, max(2, steps // 8), 1, "backward")
    # algorithmic work: 2 FLOP per tpix per gradient map, two maps (source and template) per
    # op, two ops per image
    gflop = 2.0 * 2 * 2 * n * W * H * tw * th / 1e9
    out["diff_1024x16"] = {
        "workload": f"{n} x {W}x{H} F32 U[0,1), {tw}x{th} N(0,1) template: CrossCorr_Norm and "
                    f"SqrDistance gradients wrt images and template, {n} images per step "
                    f"(ocipProximityBackwardBatch)",
        "ms_per_step": round(ms, 4),
        "gradient_gpix_tpix_per_s": round(2 * 2 * n * W * H * tw * th / (ms * 1e-3) / 1e9, 1),
        "roofline": {"bound": "mfma", "achieved": round(gflop / ms, 2),
                     "peak": PEAK_TFLOPS["f32"], "unit": "TFLOP/s",
                     "frac": round(gflop / ms / PEAK_TFLOPS["f32"], 4),
                     "note": "whole step (every kernel of the batched backward call): "
                             "F32-equivalent algorithmic FLOP, computed as split-BF16 MFMA "
                             "correlations and contraction with FP64 window sums, against "
                             "the FP32 peak"}}
    for i in Ss + G1 + G2 + dSs + [T, dT]:
        i.release()
    return out


def measured_traffic():
    """HBM bytes per launch of the headline kernel (CrossCorr_Norm, BF16, this workload) from
    the committed summary of two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE), each scaled
    by what the counter reads for a device copy of known size; None if absent."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"HBM bytes per launch, crosscorr_mfma[^:]*: (\d+)", text)
    return int(m.group(1)) if m else None


def copy_calibration():
    """One device-to-device copy of 256 MiB by a torch elementwise kernel: its FETCH_SIZE and
    WRITE_SIZE counter readings (KiB per dispatch) against the 262144 KiB moved give the
    scale that turns those counters into HBM bytes."""
    import torch
    n = 64 * 1024 * 1024  # float32 elements = 256 MiB
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4}), flush=True)
    return 0


def cpu_baseline(src, tpl, target_s=10.0):
    """The reference's CrossCorr_Norm on this host's CPU (oracle/baseline.c: each output in the
    reference's order and precision, adjacent outputs in SIMD lanes as a CPU OpenCL runtime
    runs work-items, OpenMP over rows), timed as the reference's Bench times an op: one warm
    run, then timed runs of the whole map (wall clock), as many as fit in about target_s."""
    from oracle import oracle as orc
    # the threads this process may use (the GPU box grants 16 per GPU through OMP_NUM_THREADS)
    nthreads = max(1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    H, W = src.shape
    th, tw = tpl.shape
    out = np.zeros((H, W), np.float32)
    t0 = time.perf_counter()
    _, isa = orc.baseline_cc_norm(src, tpl, nthreads=nthreads, out=out)   # warm run
    warm = time.perf_counter() - t0
    runs = int(max(1, min(20, target_s / max(warm, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(runs):
        orc.baseline_cc_norm(src, tpl, nthreads=nthreads, out=out)
    dt = (time.perf_counter() - t0) / runs
    return {"value": round(W * H * tw * th / dt / 1e9, 3), "unit": "Gpix*tpix/s",
            "cores": nthreads, "kind": "port",
            "sample": f"CrossCorr_Norm, the whole {W}x{H} map with the {tw}x{th} template, "
                      f"reference order and F32 arithmetic, {isa} lanes across adjacent outputs, "
                      f"{nthreads} threads of {os.cpu_count()} logical CPUs; {runs} timed runs "
                      f"after a warm run, {dt:.2f} s each"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="bf16", choices=list(PREC_CODE))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-fft", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the plain 1024^2, integer 4096^2 and gradient workloads")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 256 MiB device buffer (scale for rocprofv3 FETCH_SIZE/"
                         "WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the template-"
              "matching path has no multi-GPU mode in this build.", file=sys.stderr)
        return 2

    from openclipp_amd import ocip, workloads
    src, tpl, (ox, oy) = workloads.precision_4096()
    H, W = src.shape
    th, tw = tpl.shape
    tpix = W * H * tw * th

    ctx = ocip.initialize()
    S, T = ocip.Image(src).send(), ocip.Image(tpl).send()
    D = ocip.Image(shape=(H, W), dtype=np.float32)
    D32 = ocip.Image(shape=(H, W), dtype=np.float32)
    ocip.check(ocip.lib.ocipFinish())

    def step(prec):
        return lambda: ocip.lib.ocipCrossCorr_NormEx(S.handle, T.handle, D.handle if prec else D32.handle, prec)

    prec = args.precision
    ms, kms = time_steps(ocip, step(PREC_CODE[prec]), args.steps, args.warmup, DOMINANT[prec])
    value = tpix / (ms * 1e-3) / 1e9
    # the same contraction without the Norm epilogue (ocipCrossCorrEx)
    ms_cc, kms_cc = time_steps(ocip, lambda: ocip.lib.ocipCrossCorrEx(S.handle, T.handle, D32.handle,
                                                                      PREC_CODE[prec]),
                               max(3, args.steps // 2), 1, DOMINANT[prec])
    # plain FP32 VALU path on the same inputs
    ms32, kms32 = time_steps(ocip, step(0), max(3, args.steps // 2), 1, DOMINANT["f32"])
    value32 = tpix / (ms32 * 1e-3) / 1e9

    # accuracy against FP64 at sampled pixels (and the match location)
    rng = np.random.default_rng(123)
    xs = np.concatenate([rng.integers(0, W, 4095), [ox + tw // 2]])
    ys = np.concatenate([rng.integers(0, H, 4095), [oy + th // 2]])
    want = ncc_f64_at(src, tpl, xs, ys)
    got = D.read()[ys, xs].astype(np.float64) if prec != "f32" else None
    got32 = D32.read()[ys, xs].astype(np.float64)
    err = np.abs(got - want) if got is not None else np.abs(got32 - want)
    err32 = np.abs(got32 - want)

    # the other MFMA precisions on the same workload: FP16 and split BF16 (hi + lo, three
    # products: the F32-accuracy figure)
    others = {}
    for p in ("f16", "bf16x3"):
        if p == prec:
            continue
        ms_p, kms_p = time_steps(ocip, step(PREC_CODE[p]), max(3, args.steps // 2), 1, DOMINANT[p])
        e = np.abs(D.read()[ys, xs].astype(np.float64) - want)
        ach_p = 2.0 * tpix / (kms_p * 1e-3) / 1e12 if kms_p else None
        others[p] = {"value": round(tpix / (ms_p * 1e-3) / 1e9, 3), "ms_per_step": round(ms_p, 4),
                     "kernel_ms": round(kms_p, 4) if kms_p else None,
                     "achieved_tflops": round(ach_p, 1) if ach_p else None,
                     "frac": round(ach_p / PEAK_TFLOPS[p], 4) if ach_p else None,
                     "max_abs_err": float(e.max()), "mean_abs_err": float(e.mean())}

    flops = 2.0 * tpix
    achieved = flops / (kms * 1e-3) / 1e12 if kms else None
    peak = PEAK_TFLOPS[prec]
    line = {
        "metric": "NCC/SSD Gpix·tpix/s, 4096² F32, BF16-MFMA vs FP32 HIP, err≤1e-3",
        "value": round(value, 3), "unit": "Gpix·tpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": prec, "data": "synthetic",
        "config": {"workload": f"CrossCorr_Norm, {W}x{H} F32 N(0.5,0.2) clamped, {tw}x{th} "
                               f"crop x1.1, precision {prec}",
                   "image": [W, H], "template": [tw, th], "precision": prec},
        "roofline": {"bound": "mfma" if prec != "f32" else "valu",
                     "kernel": DOMINANT[prec],
                     "achieved": round(achieved, 2) if achieved else None, "peak": peak,
                     "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4) if achieved else None,
                     "kernel_ms": round(kms, 4) if kms else None,
                     "traffic": measured_traffic() if prec == "bf16" else None},
        "max_abs_err": float(err.max()), "mean_abs_err": float(err.mean()),
        "plain_f32": {"value": round(value32, 3), "ms_per_step": round(ms32, 4),
                      "kernel_ms": round(kms32, 4) if kms32 else None,
                      "achieved_tflops": round(flops / (kms32 * 1e-3) / 1e12, 2) if kms32 else None,
                      "max_abs_err": float(err32.max()), "mean_abs_err": float(err32.mean())},
        "speedup_vs_plain_f32": round(value / value32, 3),
        "other_precisions": others,
        "crosscorr_same_precision": {"value": round(tpix / (ms_cc * 1e-3) / 1e9, 3),
                                     "ms_per_step": round(ms_cc, 4),
                                     "kernel_ms": round(kms_cc, 4) if kms_cc else None},
    }
    if not args.no_batch:
        line["batch_2048"] = batch_bench(ocip, max(2, args.steps // 4))
    if not args.no_fft:
        line["fft_4096"] = fft_bench(ocip, src, max(2, args.steps // 4))
    if not args.no_configs:
        line["configs"] = configs_bench(ocip, max(4, args.steps // 2))
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(src, tpl)
    for i in (S, T, D, D32):
        i.release()
    ocip.lib.ocipUninitialize(ctx)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
