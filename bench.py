#!/usr/bin/env python
"""Benchmark of the EFF/OLA operator on one MI355X.

Workload (one "step"): forward y = B(x, f) plus both exact transposes
(x-adjoint F'g and f-adjoint X'g of one residual g) for a batch of frames of
1024 x 1024 float32 images, 10 x 10 patches at 50 % overlap (window 188,
hop 94, FFT tile 256 x 256), 41 x 41 motion-streak PSFs.  Inputs are
synthetic (seeded piecewise-smooth images, random-walk PSFs) and already
resident in HBM when the timed region starts.

Prints ONE JSON line.  value = Mpix/s = frames * 1024^2 / time per step.

    python bench.py [--steps K] [--warmup W] [--frames B] [--cpu-seconds S]
"""
import argparse
import json
import os
import platform
import sys
import time

H = W = 1024
CSF = (10, 10)
SF = (41, 41)
OVERLAP = 0.5
HBM_PEAK_GBS = 8000.0
VALU_FP32_PEAK = 78.6     # TFLOP/s, non-packed FP32 FMA (256 CU x 64 lanes x 2 x 2.4 GHz)
VALU_FP32_PACKED = 157.3  # TFLOP/s with packed FP32 / FP32 MFMA


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--frames', type=int, default=1, help='frames per step (each its own plan call)')
    ap.add_argument('--cpu-seconds', type=float, default=12.0, help='bounded CPU-baseline sample')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--calib', action='store_true',
                    help='run one 64 MiB elementwise copy first (counter calibration)')
    return ap.parse_args()


def fft_flops(n):
    """Nominal flops of one complex FFT of length n (5 n log2 n)."""
    import math
    return 5.0 * n * math.log2(n)


def kernel_models(plan, frames):
    """Algorithmic bytes (unique inputs read once + outputs written once) and
    nominal flops of each kernel launch of one forward + adjoint_xf call."""
    P = plan.P
    Th, Tw = plan.tile
    NB = Tw // 2 + 1
    sw0, sw1 = plan.sw
    sf0, sf1 = plan.sf
    ext0 = sw0 + sf0 - 1
    sy0, sy1 = plan.sy
    c8 = 8.0
    rowfft = fft_flops(Tw)
    colfft = fft_flops(Th)

    def r2c(rows, src_bytes):
        return dict(bytes=src_bytes + P * NB * rows * c8, flops=P * ((rows + 1) // 2) * rowfft)

    m = {}
    # forward: rows_r2c(x) + rows_r2c(f) + cols_fwd + rows_c2r_ola
    m['rows_r2c'] = [r2c(sw0, H * W * 4.0), r2c(sf0, P * sf0 * sf1 * 4.0),
                     r2c(ext0, sy0 * sy1 * 4.0), r2c(sw0, H * W * 4.0), r2c(sf0, P * sf0 * sf1 * 4.0)]
    m['cols_fwd'] = [dict(bytes=P * NB * (sw0 + sf0 + ext0) * c8, flops=P * NB * (3 * colfft + 6 * Th))]
    m['rows_c2r_ola'] = [dict(bytes=P * ext0 * NB * c8 + sy0 * sy1 * 4.0, flops=sy0 * 3 * ((CSF[1] + 1) // 2) * rowfft)]
    m['cols_adj2'] = [dict(bytes=P * NB * (ext0 + sw0 + sf0) * c8 + P * NB * (sw0 + sf0) * c8,
                           flops=P * NB * (5 * colfft + 12 * Th))]
    m['rows_c2r_ola_win'] = [dict(bytes=P * sw0 * NB * c8 + H * W * 4.0, flops=H * 2 * ((CSF[1] + 1) // 2) * rowfft)]
    m['rows_c2r_stack'] = [dict(bytes=P * sf0 * NB * c8 + P * sf0 * sf1 * 4.0, flops=P * ((sf0 + 1) // 2) * rowfft)]
    return m


def operator_bytes():
    """Algorithmic bytes of fwd + x-adj + f-adj (inputs and outputs once)."""
    Sx = H * W
    Sy = (H - SF[0] + 1) * (W - SF[1] + 1)
    Pf = CSF[0] * CSF[1] * SF[0] * SF[1]
    return 4.0 * 3 * (Sx + Sy + Pf)


def main():
    a = parse()
    if a.gpus != 1:
        print(f'bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU sharding of this '
              'operator is a design note, not a measured path.', file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from vmbd_amd import _lib, synth
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan

    dev = torch.device('cuda', 0)
    torch.cuda.set_device(dev)
    wa = win2winaux((H, W), CSF, OVERLAP)
    plan = OlaPlan((H, W), SF, wa, mode='valid', device=dev)
    B = a.frames
    rng = np.random.default_rng(1234)
    xs_np = [synth.piecewise_smooth((H, W), rng) for _ in range(B)]
    fs_np = [synth.motion_psfs(plan.P, SF, rng, steps=40) for _ in range(B)]
    gs_np = [rng.standard_normal(plan.sy) * 0.01 for _ in range(B)]
    to = lambda v: torch.as_tensor(np.ascontiguousarray(v, dtype=np.float32), device=dev)
    xs = [to(v) for v in xs_np]
    fs = [to(v) for v in fs_np]
    gs = [to(v) for v in gs_np]
    ys = [torch.empty(plan.sy, device=dev) for _ in range(B)]
    xbs = [torch.empty((H, W), device=dev) for _ in range(B)]
    fbs = [torch.empty((plan.P,) + SF, device=dev) for _ in range(B)]

    if a.calib:
        src = torch.ones(16 * 1024 * 1024, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    def step():
        for b in range(B):
            plan.forward(xs[b], fs[b], out=ys[b])
            plan.adjoint_xf(gs[b], xs[b], fs[b], exact=True, out_x=xbs[b], out_f=fbs[b])

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    mpix = B * H * W / (ms * 1e-3) / 1e6

    # per-kernel durations over a second timed pass, events recorded by the
    # library around each launch on the launch stream
    _lib.prof_reset()
    _lib.prof_enable(True)
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    rep = _lib.prof_report()
    _lib.prof_enable(False)
    _lib.prof_reset()
    models = kernel_models(plan, B)
    kernels = {}
    for name, (cnt, tot, avg) in rep.items():
        mm = models.get(name)
        entry = dict(launches_per_step=cnt / a.steps, avg_us=avg, total_us_per_step=tot / a.steps)
        if mm:
            per_launch_b = sum(d['bytes'] for d in mm) / len(mm)
            per_launch_f = sum(d['flops'] for d in mm) / len(mm)
            entry['alg_bytes'] = per_launch_b
            entry['hbm_gbs'] = per_launch_b / (avg * 1e-6) / 1e9
            entry['alg_flops'] = per_launch_f
            entry['tflops'] = per_launch_f / (avg * 1e-6) / 1e12
        kernels[name] = entry
    dom = max(kernels, key=lambda k: kernels[k]['total_us_per_step'])
    d = kernels[dom]
    roofline = dict(kernel=dom, bound='hbm', achieved=round(d.get('hbm_gbs', 0.0), 1), peak=HBM_PEAK_GBS,
                    unit='GB/s', frac=round(d.get('hbm_gbs', 0.0) / HBM_PEAK_GBS, 4), traffic=None,
                    avg_us=round(d['avg_us'], 2),
                    valu=dict(achieved_tflops=round(d.get('tflops', 0.0), 2), peak_nonpacked=VALU_FP32_PEAK,
                              frac=round(d.get('tflops', 0.0) / VALU_FP32_PEAK, 4)))
    op_bytes = operator_bytes() * B
    step_hbm = op_bytes / (ms * 1e-3) / 1e9

    # accuracy of this exact workload against the FP64 oracle + CPU baseline
    cpu = None
    acc = None
    if not a.no_cpu:
        from oracle import ola_oracle as O
        owa = O.Win2WinAux((H, W), CSF, OVERLAP)
        x64 = np.float32(xs_np[0]).astype(np.float64)
        f64 = np.float32(fs_np[0]).astype(np.float64)
        g64 = np.float32(gs_np[0]).astype(np.float64)
        n = 0
        t0 = time.perf_counter()
        while True:
            y_ref = O.forward(x64, f64, owa)
            xb_ref = O.adjoint_x(g64, f64, owa, (H, W), exact=True)
            fb_ref = O.adjoint_f(g64, x64, owa, SF, exact=True)
            n += 1
            if time.perf_counter() - t0 >= a.cpu_seconds:
                break
        dt = (time.perf_counter() - t0) / n
        cpu = dict(value=round(H * W / dt / 1e6, 4), unit='Mpix/s', cores=1, kind='port',
                   sample=f'{n} x (fwd + x-adj + f-adj) of one 1024^2 frame, FP64 NumPy oracle, single thread, '
                          f'{dt:.2f} s each on {platform.processor() or platform.machine()} '
                          f'({os.cpu_count()} host cores visible)')

        def rl2(u, v):
            return float(np.linalg.norm((u - v).ravel()) / np.linalg.norm(v.ravel()))

        acc = dict(y=rl2(ys[0].cpu().numpy().astype(np.float64), y_ref),
                   xbar=rl2(xbs[0].cpu().numpy().astype(np.float64), xb_ref),
                   fbar=rl2(fbs[0].cpu().numpy().astype(np.float64), fb_ref))

    out = {
        'metric': 'OLA fwd+adj Mpix/s at 1024^2 FP32 (rel err <= 1e-5); %HBM peak',
        'value': round(mpix, 2),
        'unit': 'Mpix/s',
        'n_gpus': 1,
        'steps': a.steps,
        'warmup': a.warmup,
        'ms_per_step': round(ms, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'fp32',
        'data': 'synthetic (seeded piecewise-smooth 1024^2 images, 41x41 random-walk motion PSFs, N(0,0.01^2) residual)',
        'config': {'workload': 'EFF/OLA forward + exact x-adjoint + exact f-adjoint, 1024x1024 fp32, '
                               '10x10 patches 50% overlap (sw 188, hop 94), 41x41 PSFs, FFT tile 256x256',
                   'frames_per_step': B, 'mode': 'valid'},
        'hbm_frac_step': round(step_hbm / HBM_PEAK_GBS, 5),
        'step_alg_bytes': op_bytes,
        'roofline': roofline,
        'kernels': kernels,
        'cpu_baseline': cpu,
        'rel_l2_vs_fp64': acc,
    }
    print(json.dumps(out))


if __name__ == '__main__':
    main()
