#!/usr/bin/env python
"""Benchmark of the EFF/OLA operator on one MI355X.

Workload (one "step"): forward y = B(x, f) plus both exact transposes
(x-adjoint F'g and f-adjoint X'g of a residual g) for a batch of frames of
1024 x 1024 float32 images, 10 x 10 patches at 50 % overlap (window 188,
hop 94, FFT tile 256 x 256), 41 x 41 motion-streak PSFs.  The step holds
`streams` batches of `frames` frames (default 2 x 32), each a plan of its
own whose launches carry all patches of its frames, the batches forked over
that many streams; the step is captured once in a HIP graph and replayed.
Inputs are synthetic (seeded piecewise-smooth images, random-walk PSFs,
Gaussian residuals), distinct per frame, and already resident in HBM when
the timed region starts.

Prints ONE JSON line.  value = Mpix/s = frames * 1024^2 / time per step.
Beside it: the same step in the plain per-operator GPU form (the
reference's launch sequence, one kernel per operation, full complex spectra)
and the online MFBD loop in frames/s (1024^2 latent, 20 projected-BB
f-iterations + 20 x-iterations per frame, graph-captured).

    python bench.py [--steps K] [--warmup W] [--frames B] [--streams S] [--no-graph] [--cpu-seconds S]
                    [--no-plain] [--no-mfbd] [--mfbd-frames N] [--no-mixed] [--no-deconv]
"""
import argparse
import json
import math
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--frames', type=int, default=32, help='frames per batch, batched into the same launches')
    ap.add_argument('--streams', type=int, default=2,
                    help='batches per step, each on a stream of its own inside the graph')
    ap.add_argument('--no-graph', action='store_true', help='launch eagerly instead of replaying a HIP graph')
    ap.add_argument('--cpu-seconds', type=float, default=18.0,
                    help='bounded CPU-baseline sample (split over its three variants)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-plain', action='store_true', help='skip the plain per-operator comparator')
    ap.add_argument('--no-mfbd', action='store_true', help='skip the MFBD frames/s measurement')
    ap.add_argument('--mfbd-frames', type=int, default=10, help='timed MFBD frames')
    ap.add_argument('--no-mixed', action='store_true', help='skip the 64-frame mixed-geometry batch')
    ap.add_argument('--no-deconv', action='store_true', help='skip the closed-form deconvolution timings')
    ap.add_argument('--calib', action='store_true',
                    help='run one 64 MiB elementwise copy first (counter calibration)')
    return ap.parse_args()


def fft_flops(n):
    """Nominal flops of one complex FFT of length n (5 n log2 n)."""
    return 5.0 * n * math.log2(n)


def kernel_models(plan, B):
    """Per-launch algorithmic bytes (unique inputs read once + outputs written
    once) and nominal flops of the launches of one forward + adjoint_xf step
    (row transforms of x and f reused by the adjoint)."""
    P = plan.P
    Th, Tw = plan.tile
    # complex row bins carried between the passes: 0..Tw/2-1, bin 0 holding
    # the real bins 0 and Tw/2
    NB = Tw // 2
    sw0 = plan.sw[0]
    sf0, sf1 = plan.sf
    ext0 = sw0 + sf0 - 1
    nv = (plan.csf[0] - 1) * plan.nhop[0] + ext0
    sy0, sy1 = plan.sy
    rf, cf = fft_flops(Tw), fft_flops(Th)
    half = lambda r: (r + 1) // 2
    # separable windows: image rows shared by the patches of a patch column,
    # (csf0-1)*hop0 + sw0 rows per patch column (per-row patch factors applied
    # by the column pass)
    sx0 = (plan.csf[0] - 1) * plan.nhop[0] + sw0
    xrows = plan.csf[1] * sx0 if plan.separable else P * sw0
    m = {
        'rows_r2c_multi': dict(bytes=B * (H * W * 4 + P * NB * sw0 * 8 + P * sf0 * sf1 * 4 + P * NB * sf0 * 8),
                               flops=B * P * (half(sw0) + half(sf0)) * rf),
        'rows_r2c_x': dict(bytes=B * (H * W * 4 + plan.csf[1] * NB * sx0 * 8),
                           flops=B * plan.csf[1] * half(sx0) * rf),
        'rows_r2c_psf': dict(bytes=B * (P * sf0 * sf1 * 4 + P * NB * sf0 * 8), flops=B * P * half(sf0) * rf),
        # the shared image rows and the PSF rows in one launch
        'rows_r2c_x_psf': dict(bytes=B * (H * W * 4 + plan.csf[1] * NB * sx0 * 8 + P * sf0 * sf1 * 4
                                          + P * NB * sf0 * 8),
                               flops=B * (plan.csf[1] * half(sx0) + P * half(sf0)) * rf),
        # residual rows shared by the patches of a patch column: one transform
        # per (patch column, virtual row pair) over (csf0-1)*hop0 + ext0 rows
        'rows_r2c': dict(bytes=B * (sy0 * sy1 * 4 + plan.csf[1] * NB * nv * 8),
                         flops=B * plan.csf[1] * half(nv) * rf),
        'cols_fwd': dict(bytes=B * NB * (xrows + P * (sf0 + ext0)) * 8, flops=B * P * NB * (3 * cf + 6 * Th)),
        # summed overlap-add: one inverse row transform per output row and
        # pair of patch columns (the covering rows added before it)
        'rows_c2r_ola': dict(bytes=B * (P * ext0 * NB * 8 + sy0 * sy1 * 4), flops=B * sy0 * half(plan.csf[1]) * rf),
        'cols_adj2': dict(bytes=B * NB * (xrows + P * (ext0 + sw0 + 2 * sf0)) * 8,
                          flops=B * P * NB * (5 * cf + 12 * Th)),
        'rows_c2r_ola_win': dict(bytes=B * (P * sw0 * NB * 8 + H * W * 4),
                                 flops=B * (H * half(plan.csf[1]) if plan.separable else P * sw0 / 2) * rf),
        'rows_c2r_stack': dict(bytes=B * (P * sf0 * NB * 8 + P * sf0 * sf1 * 4), flops=B * P * half(sf0) * rf),
    }
    return m


def operator_bytes():
    """Algorithmic bytes of fwd + x-adj + f-adj of one frame (inputs and
    outputs once)."""
    Sx = H * W
    Sy = (H - SF[0] + 1) * (W - SF[1] + 1)
    Pf = CSF[0] * CSF[1] * SF[0] * SF[1]
    return 4.0 * 3 * (Sx + Sy + Pf)


def operator_flops():
    """Algorithmic flops of fwd + x-adj + f-adj of one frame: P (5 real
    T^2 FFTs + 9 T^2), real FFT = half of 5 N log2 N for N = T^2 points."""
    P = CSF[0] * CSF[1]
    T = 256
    rfft = 2.5 * T * T * math.log2(T * T)
    return P * (5 * rfft + 9 * T * T)


def attainable_hbm_frac():
    """HBM fraction the step can reach when it runs at the non-packed FP32
    VALU peak: t_HBM / t_VALU of the algorithmic bytes and flops."""
    return (operator_bytes() / (HBM_PEAK_GBS * 1e9)) / (operator_flops() / (VALU_FP32_PEAK * 1e12))


def operator_roofline(kernels, B, ms, nstep):
    """Roofline on the operator's algorithmic bytes and flops (one fwd +
    x-adj + f-adj per frame, inputs and outputs once, windows generated in
    the kernels): `achieved`/`frac` are the B frames' bytes over the dominant
    kernel's average launch time (a launch carries B frames; the intermediate
    spectra it moves are not algorithmic bytes, they are under `kernels`),
    `step` the nstep frames' bytes and flops over the whole step's time.  The step is VALU-issue bound (52
    flop per byte against a ridge of 10 for non-packed FP32), so `valu`
    gives its flop rate against the non-packed FP32 peak, and the HBM
    fraction it can reach is capped at attainable_hbm_frac()."""
    dom = max(kernels, key=lambda k: kernels[k]['total_us_per_step'])
    d = kernels[dom]
    by = operator_bytes() * B
    gbs = by / (d['avg_us'] * 1e-6) / 1e9
    step_gbs = operator_bytes() * nstep / (ms * 1e-3) / 1e9
    fl = operator_flops() * nstep
    step_tf = fl / (ms * 1e-3) / 1e12
    return dict(kernel=dom, bound='valu', achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit='GB/s',
                frac=round(gbs / HBM_PEAK_GBS, 4), traffic=measured_traffic(dom, B), avg_us=d['avg_us'],
                alg_bytes_per_launch=by,
                # the bound: algorithmic flops of the step over the step's time
                valu=dict(basis='step', alg_flops=fl, achieved_tflops=round(step_tf, 2),
                          peak_nonpacked=VALU_FP32_PEAK, frac=round(step_tf / VALU_FP32_PEAK, 4)),
                step=dict(hbm_gbs=round(step_gbs, 1), hbm_frac=round(step_gbs / HBM_PEAK_GBS, 5),
                          valu_tflops=round(step_tf, 2), valu_frac=round(step_tf / VALU_FP32_PEAK, 4)),
                attainable_hbm_frac=round(attainable_hbm_frac(), 4))


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                            'ola_1024_10x10_sf41_b32_hbm_traffic.md')


def measured_traffic(kernel, frames):
    """HBM bytes per launch of `kernel` at this workload from the committed
    rocprofv3 --pmc summary (FETCH_SIZE x its calibration factor +
    WRITE_SIZE), or None when the summary has no line for it or was measured
    at another number of frames per launch."""
    try:
        with open(TRAFFIC_FILE) as fh:
            lines = [ln.split() for ln in fh]
    except OSError:
        return None
    if ['frames_per_launch', str(frames)] not in lines:
        return None
    for parts in lines:
        if len(parts) == 3 and parts[0] == 'traffic_bytes_per_launch' and parts[1] == kernel:
            return int(parts[2])
    return None


def host_threads():
    """Threads the CPU baseline may use: OMP_NUM_THREADS when set (the host
    share of the GPU's node), else the CPUs this process may run on."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    env = os.environ.get('OMP_NUM_THREADS', '')
    return max(1, min(avail, int(env))) if env.isdigit() and int(env) > 0 else avail


def cpu_model():
    try:
        with open('/proc/cpuinfo') as fh:
            for ln in fh:
                if ln.startswith('model name'):
                    return ln.split(':', 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def plain_form(wa, x, f, g, timed, a):
    """fwd + both transposes of one frame in the plain per-operator form."""
    from vmbd_amd.plain import PlainOla
    p = PlainOla((H, W), SF, wa)
    steps = max(3, min(a.steps, 10))
    for _ in range(2):
        p.step(x, f, g)
    ms = timed(lambda: p.step(x, f, g), steps)
    return {'ms_per_frame': round(ms, 4), 'mpix_s': round(H * W / (ms * 1e-3) / 1e6, 2), 'steps': steps,
            'form': 'reference launch sequence (F ctor, X ctor, F.cnv, F.cnvtp, X.cnvtp), one HIP kernel per '
                    'operation, complex (T x T) spectra, patch-local transposes'}


def mfbd_rate(sx, csf, sf, walk, nframes, timed, iters=20, prior=None, tune=True):
    """Online MFBD frames/s: latent sx, PSF grid csf of sf PSFs, per frame
    `iters` PBB f-iterations + `iters` PBB x-iterations (exact adjoints) and
    the optional prior pass, captured in a HIP graph and replayed per frame."""
    import numpy as np
    import torch

    from vmbd_amd import synth
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.mfbd import OnlineMFBD
    from vmbd_amd.ola import OlaPlan

    dev = torch.device('cuda', 0)
    rng = np.random.default_rng(99)
    wa = win2winaux(sx, csf, OVERLAP)
    plan = OlaPlan(sx, sf, wa, 'valid', device=dev)
    xt = torch.as_tensor(synth.piecewise_smooth(sx, rng).astype(np.float32), device=dev)
    obs = []
    for _ in range(4):
        ft = torch.as_tensor(synth.motion_psfs(plan.P, sf, rng, steps=walk).astype(np.float32), device=dev)
        yt = plan.forward(xt, ft)
        noise = torch.as_tensor((0.01 * rng.standard_normal(plan.sy)).astype(np.float32), device=dev)
        obs.append(yt + noise)
    m = OnlineMFBD(obs[0], sf, csf, f_solver='pbb', x_solver='pbb', f_iters=iters, x_iters=iters, f_beta=1e-3,
                   exact=True, capture=True, prior=prior, device=dev, tune=tune)
    for i in range(2):
        m.step(obs[i % 4])
    k = [0]

    def one():
        m.step(obs[k[0] % 4])
        k[0] += 1

    ms = timed(one, nframes)
    return {'frames_per_s': round(1e3 / ms, 2), 'ms_per_frame': round(ms, 3), 'frames': nframes,
            'latent': list(sx), 'csf': list(csf), 'sf': list(sf), 'tile': list(m.plan.tile),
            'launch_shape': m.plan.tuning, 'autotuned': bool(tune), 'f_iters': iters, 'x_iters': iters,
            'prior': prior, 'graph': True}


C4_GEOMS = [((512, 512), (12, 12), (15, 15)), ((512, 512), (6, 6), (41, 41)), ((768, 1024), (8, 8), (25, 25)),
            ((768, 1024), (6, 6), (15, 15)), ((1024, 1024), (10, 10), (41, 41)), ((1024, 1024), (12, 12), (25, 25))]


def deconv_rates(wa, x, f, timed, reps=5):
    """ms per OlaGPU.deconv call at the benchmark geometry (1024^2, 10 x 10
    patches, 41 x 41 PSFs) for the closed-form modes, fused into the
    operator's passes and as the reference's plain launch sequence, on an
    observation y = B(x, f) + N(0, 0.01^2).  Each call includes its host-side
    set-up (tapering, gradients, output allocation)."""
    import numpy as np
    import torch

    from vmbd_amd.olaGPU import OlaGPU

    y = OlaGPU(f, (H, W), 'valid', wa).cnv(x)
    y = y + 0.01 * torch.randn(y.shape, generator=torch.Generator(device=y.device).manual_seed(5), device=y.device)
    out = {}
    for mode, inst, kw in (('xdirect', 'F', dict(alpha=0.05, beta=0.01)), ('direct', 'F', dict(alpha=0.05, beta=0.01)),
                           ('gdirect', 'X', dict(alpha=0.01, beta=0.01)), ('sparse', 'F', dict(alpha=1e-4))):
        e = {}
        for fused in (True, False):
            Z = OlaGPU(f, (H, W), 'valid', wa, fused=fused) if inst == 'F' else OlaGPU(x, SF, 'valid', wa, fused=fused)
            Z.deconv(y, mode=mode, **kw)
            e['fused_ms' if fused else 'plain_ms'] = round(timed(lambda: Z.deconv(y, mode=mode, **kw), reps), 3)
            del Z
        e['speedup'] = round(e['plain_ms'] / e['fused_ms'], 2)
        out[mode] = e
    return out


def mixed_rate(nframes, timed, steps):
    """64 frames of six geometries (512^2 / 768x1024 / 1024^2, 6x6..12x12
    patches, 15/25/41 PSFs): forward + both exact adjoints of every frame,
    one graph replay per step, in three launch forms: each pass of all groups
    as one launch per distinct tile (grouped, one stream), one such batch per
    tile forked over three streams, and each group's own launches forked
    over three streams."""
    import numpy as np
    import torch

    from vmbd_amd import synth
    from vmbd_amd.batch import MixedBatch

    dev = torch.device('cuda', 0)
    rng = np.random.default_rng(7)
    specs = [C4_GEOMS[i % len(C4_GEOMS)] for i in range(nframes)]
    to = lambda v: torch.as_tensor(np.ascontiguousarray(v, dtype=np.float32), device=dev)
    xs, fs, gs = [], [], []
    for sx, csf, sf in specs:
        xs.append(to(synth.piecewise_smooth(sx, rng)))
        fs.append(to(synth.motion_psfs(csf[0] * csf[1], sf, rng, steps=sf[0])))
        gs.append(to(0.01 * rng.standard_normal((sx[0] - sf[0] + 1, sx[1] - sf[1] + 1))))
    out = {'frames': nframes, 'geometries': [[list(a), list(b), list(c)] for a, b, c in C4_GEOMS], 'graph': True}
    outs = {}
    for name, kw in (('grouped', dict(grouped=True, streams=1)), ('grouped_streams3', dict(grouped=True, streams=3)),
                     ('streams3', dict(grouped=False, streams=3))):
        mb = MixedBatch(specs, device=dev, **kw)
        mb.load(xs, fs, gs)
        replay = mb.capture()
        for _ in range(2):
            replay()
        ms = timed(replay, steps)
        out[name] = {'ms_per_step': round(ms, 3), 'frames_per_s': round(nframes / (ms * 1e-3), 1),
                     'mpix_s': round(mb.pixels / (ms * 1e-3) / 1e6, 1), 'launches_per_step': mb.launches,
                     'groups': len(mb.groups)}
        outs[name] = [tuple(t.clone() for t in mb.outputs(i)) for i in range(nframes)]
        del mb
    out['same_bits'] = all(torch.equal(a, b) for k in ('grouped', 'grouped_streams3') for i in range(nframes)
                           for a, b in zip(outs[k][i], outs['streams3'][i]))
    best = min(('grouped', 'grouped_streams3', 'streams3'), key=lambda k: out[k]['ms_per_step'])
    out.update(best=best, mpix_s=out[best]['mpix_s'], frames_per_s=out[best]['frames_per_s'])
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f'bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU sharding of this '
              'operator is a design note, not a measured path.', file=sys.stderr)
        sys.exit(2)
    if a.frames < 1 or a.streams < 1:
        print('bench.py: --frames and --streams take positive counts', file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from vmbd_amd import _lib, synth
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan

    dev = torch.device('cuda', 0)
    torch.cuda.set_device(dev)
    B, S = a.frames, a.streams
    N = S * B  # frames per step
    wa = win2winaux((H, W), CSF, OVERLAP)
    plans = [OlaPlan((H, W), SF, wa, mode='valid', device=dev, max_frames=B) for _ in range(S)]
    plan = plans[0]
    rng = np.random.default_rng(1234)
    x_np = np.stack([synth.piecewise_smooth((H, W), rng) for _ in range(N)])
    f_np = np.stack([synth.motion_psfs(plan.P, SF, rng, steps=40) for _ in range(N)])
    g_np = rng.standard_normal((N,) + plan.sy) * 0.01
    to = lambda v: torch.as_tensor(np.ascontiguousarray(v, dtype=np.float32), device=dev)
    x, f, g = to(x_np), to(f_np), to(g_np)
    y = torch.empty((N,) + plan.sy, device=dev)
    xb = torch.empty((N, H, W), device=dev)
    fb = torch.empty((N, plan.P) + SF, device=dev)
    # batch k's views, made once: the adjoint reuses the forward's row
    # transforms only for the same tensors
    views = [[t[k * B:(k + 1) * B] for t in (x, f, g, y, xb, fb)] for k in range(S)]

    if a.calib:
        src = torch.ones(16 * 1024 * 1024, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    def batch(k):
        xk, fk, gk, yk, xbk, fbk = views[k]
        plans[k].forward(xk, fk, out=yk)
        plans[k].adjoint_xf(gk, xk, fk, exact=True, out_x=xbk, out_f=fbk)

    def step():  # eager: the batches one after the other on the current stream
        for k in range(S):
            batch(k)

    def forked_step(side):
        """The batches over the current stream and the streams in side,
        forked from and joined back into the current one."""
        cur = torch.cuda.current_stream()
        for ss in side:
            ss.wait_stream(cur)
        for k in range(S):
            with torch.cuda.stream(cur if k == 0 else side[k - 1]):
                batch(k)
        for ss in side:
            cur.wait_stream(ss)

    def timed(fn, steps):
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    eager_ms = timed(step, a.steps)
    ms = eager_ms
    graph = None
    if not a.no_graph:
        s = torch.cuda.Stream()
        side = [torch.cuda.Stream() for _ in range(S - 1)]
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            step()  # the adjoint's reuse decision is taken outside the capture
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, stream=s):
                forked_step(side)
        torch.cuda.current_stream().wait_stream(s)
        for _ in range(a.warmup):
            graph.replay()
        ms = timed(graph.replay, a.steps)
    mpix = N * H * W / (ms * 1e-3) / 1e6

    # one batch alone on one stream, graph-replayed: the step without the
    # overlap of the batches
    one_batch = None
    if graph is not None and S > 1:
        s1 = torch.cuda.Stream()
        s1.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s1):
            batch(0)
            graph1 = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph1, stream=s1):
                batch(0)
        torch.cuda.current_stream().wait_stream(s1)
        for _ in range(a.warmup):
            graph1.replay()
        ms_b = timed(graph1.replay, a.steps)
        one_batch = {'frames': B, 'ms': round(ms_b, 4), 'mpix_s': round(B * H * W / (ms_b * 1e-3) / 1e6, 2)}
        del graph1

    # single-frame eager step, for reference
    plan1 = OlaPlan((H, W), SF, wa, mode='valid', device=dev, max_frames=1)

    def step1():
        plan1.forward(x[0], f[0], out=y[0])
        plan1.adjoint_xf(g[0], x[0], f[0], exact=True, out_x=xb[0], out_f=fb[0])

    for _ in range(a.warmup):
        step1()
    ms1 = timed(step1, a.steps)

    # per-kernel durations over an eager pass, events recorded by the library
    # around each launch on the launch stream
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
        e = dict(launches_per_step=cnt / a.steps, avg_us=round(avg, 3), total_us_per_step=round(tot / a.steps, 3))
        mm = models.get(name)
        if mm:
            # the kernel's own intermediate operands (row/column spectra)
            e['kernel_bytes'] = mm['bytes']
            e['kernel_bytes_gbs'] = round(mm['bytes'] / (avg * 1e-6) / 1e9, 1)
            e['nominal_flops'] = mm['flops']
            e['tflops'] = round(mm['flops'] / (avg * 1e-6) / 1e12, 3)
        kernels[name] = e
    roofline = operator_roofline(kernels, B, ms, N)
    op_bytes = operator_bytes() * N

    plain = None
    if not a.no_plain:
        plain = plain_form(wa, x[0], f[0], g[0], timed, a)
        plain['fused_speedup'] = round(mpix / plain['mpix_s'], 2)
        plain['fused_speedup_single_frame_eager'] = round(H * W / (ms1 * 1e-3) / 1e6 / plain['mpix_s'], 2)

    mixed = None
    if not a.no_mixed:
        mixed = mixed_rate(64, timed, max(3, min(a.steps, 10)))

    deconv = None
    if not a.no_deconv:
        deconv = deconv_rates(wa, x[0], f[0], timed)

    mfbd = None
    if not a.no_mfbd:
        mfbd = {'1024': mfbd_rate((H, W), CSF, SF, 40, a.mfbd_frames, timed),
                '512': mfbd_rate((512, 512), (8, 8), (31, 31), 15, a.mfbd_frames, timed),
                '2048_prior': mfbd_rate((2048, 2048), (16, 16), (63, 63), 60, a.mfbd_frames, timed, iters=10,
                                        prior='bilateral_shock')}

    cpu = None
    acc = None
    if not a.no_cpu:
        from oracle import ola_oracle as O
        owa = O.Win2WinAux((H, W), CSF, OVERLAP)
        x64 = np.float32(x_np[0]).astype(np.float64)
        f64 = np.float32(f_np[0]).astype(np.float64)
        g64 = np.float32(g_np[0]).astype(np.float64)
        threads = host_threads()

        def cpu_rate(workers, single, seconds):
            """Mpix/s of the oracle's fwd + x-adj + f-adj on one frame, timed
            for about `seconds` (at least one step); returns (rate, steps,
            s/step, outputs of the last step)."""
            old = O.set_fft(workers, single)
            try:
                n, t0 = 0, time.perf_counter()
                while True:
                    outs = (O.forward(x64, f64, owa), O.adjoint_x(g64, f64, owa, (H, W), exact=True),
                            O.adjoint_f(g64, x64, owa, SF, exact=True))
                    n += 1
                    if time.perf_counter() - t0 >= seconds:
                        break
                dt = (time.perf_counter() - t0) / n
            finally:
                O.set_fft(*old)
            return H * W / dt / 1e6, n, dt, outs

        where = f'{cpu_model()}, {os.cpu_count()} logical cores visible'
        r1, n1, d1, (y_ref, xb_ref, fb_ref) = cpu_rate(1, False, a.cpu_seconds / 3)
        rA, nA, dA, _ = cpu_rate(threads, False, a.cpu_seconds / 3)
        rS, nS, dS, _ = cpu_rate(threads, True, a.cpu_seconds / 3)
        cpu = dict(value=round(rA, 4), unit='Mpix/s', cores=threads, kind='port',
                   sample=f'{nA} x (fwd + x-adj + f-adj) of one 1024^2 frame, FP64 NumPy oracle with scipy.fft '
                          f'on {threads} threads, {dA:.2f} s each ({where})',
                   variants=[dict(value=round(r1, 4), cores=1, precision='fp64',
                                  sample=f'{n1} steps, numpy.fft single thread, {d1:.2f} s each'),
                             dict(value=round(rA, 4), cores=threads, precision='fp64',
                                  sample=f'{nA} steps, scipy.fft workers={threads}, {dA:.2f} s each'),
                             dict(value=round(rS, 4), cores=threads, precision='fp32 FFTs and products',
                                  sample=f'{nS} steps, scipy.fft complex64 workers={threads}, {dS:.2f} s each')])
        if graph is not None:
            graph.replay()
            torch.cuda.synchronize()

        def rl2(u, v):
            return float(np.linalg.norm((u - v).ravel()) / np.linalg.norm(v.ravel()))

        acc = dict(y=rl2(y[0].cpu().numpy().astype(np.float64), y_ref),
                   xbar=rl2(xb[0].cpu().numpy().astype(np.float64), xb_ref),
                   fbar=rl2(fb[0].cpu().numpy().astype(np.float64), fb_ref))

    out = {
        'metric': 'OLA fwd+adj Mpix/s & MFBD frames/s at 1024\u00b2 FP32 (rel err \u22641e-5); %HBM peak',
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
                   'frames_per_step': N, 'frames_per_launch': B, 'streams': S, 'graph': graph is not None,
                   'mode': 'valid'},
        'ms_per_frame': round(ms / N, 4),
        'eager_batched_ms_per_step': round(eager_ms, 4),
        'single_frame_eager': {'ms': round(ms1, 4), 'mpix_s': round(H * W / (ms1 * 1e-3) / 1e6, 2)},
        'one_batch_one_stream': one_batch,
        'step_alg_bytes': op_bytes,
        'roofline': roofline,
        'kernels': kernels,
        'cpu_baseline': cpu,
        'rel_l2_vs_fp64': acc,
        'plain_form': plain,
        'mfbd_frames_per_s': mfbd['1024']['frames_per_s'] if mfbd else None,
        'mfbd': mfbd,
        'mixed_batch': mixed,
        'deconv': deconv,
    }
    print(json.dumps(out))


if __name__ == '__main__':
    main()
