"""Frames of uneven size, patch grid and PSF support in one step.

    mb = MixedBatch([(sx_i, csf_i, sf_i) for each frame], overlap=0.5)
    mb.load(xs, fs, gs)            # per-frame device tensors -> resident stacks
    mb.step()                      # forward + both exact adjoints of every frame
    y_i, xbar_i, fbar_i = mb.outputs(i)

Frames are grouped by geometry; each group is one plan whose launches carry
all patches of all its frames (the kernels take the frame count as a grid
dimension).  ``capture()`` records the launches of every group in one HIP
graph, so a step of the whole mixed set is one graph replay; with
``streams=k > 1`` (default 3) the graph forks the groups over k streams
(largest groups first, each to the stream with the fewest pixels so far)
and joins them at its end, so one group's launches fill the tail of
another's (64 frames of six geometries: 2.48 ms per step with 3 streams,
2.50 with 2, 2.60 with 1; 4 streams no faster than 3).  The eager
``step()`` runs the groups in order on the current stream.  With
``tune=True`` each group's FFT tile comes from the autotuner.
"""
from collections import OrderedDict

import torch

from .imagetools import win2winaux
from .ola import OlaPlan


class MixedBatch:
    def __init__(self, specs, overlap=0.5, tune=False, exact=True, device=None, streams=3):
        if int(streams) < 1:
            raise ValueError(f'streams must be >= 1, got {streams}')
        self.streams = int(streams)
        self.device = torch.device(device) if device is not None else torch.device('cuda', torch.cuda.current_device())
        self.exact = bool(exact)
        self.specs = [(tuple(int(v) for v in sx), (int(c[0]), int(c[1])), (int(s[0]), int(s[1])))
                      for sx, c, s in specs]
        groups = OrderedDict()
        for i, key in enumerate(self.specs):
            groups.setdefault(key, []).append(i)
        self.groups = []
        self.where = {}
        for (sx, csf, sf), idx in groups.items():
            wa = win2winaux(sx, csf, overlap, device=self.device)
            tile = None
            if tune:
                from .autotune import best_tile
                tile, _ = best_tile(sx, sf, wa, frames=len(idx), overlap=overlap, device=self.device)
            plan = OlaPlan(sx, sf, wa, 'valid', tile=tile, device=self.device, max_frames=len(idx))
            n = len(idx)
            d, f32 = self.device, torch.float32
            bufs = dict(x=torch.zeros((n,) + sx, dtype=f32, device=d),
                        f=torch.zeros((n, plan.P) + sf, dtype=f32, device=d),
                        g=torch.zeros((n,) + plan.sy, dtype=f32, device=d),
                        y=torch.zeros((n,) + plan.sy, dtype=f32, device=d),
                        xb=torch.zeros((n,) + sx, dtype=f32, device=d),
                        fb=torch.zeros((n, plan.P) + sf, dtype=f32, device=d))
            gi = len(self.groups)
            self.groups.append(dict(plan=plan, idx=idx, **bufs))
            for j, i in enumerate(idx):
                self.where[i] = (gi, j)
        self._graph = None

    @property
    def pixels(self):
        return sum(sx[0] * sx[1] for sx, _, _ in self.specs)

    def load(self, xs, fs, gs):
        """Copy frame i's image, PSF stack and residual into the group stacks."""
        for i in range(len(self.specs)):
            gi, j = self.where[i]
            G = self.groups[gi]
            G['x'][j].copy_(xs[i])
            G['f'][j].copy_(fs[i])
            G['g'][j].copy_(gs[i])

    def _group_step(self, G):
        p = G['plan']
        p.forward(G['x'], G['f'], out=G['y'])
        p.adjoint_xf(G['g'], G['x'], G['f'], exact=self.exact, out_x=G['xb'], out_f=G['fb'])

    def step(self):
        for G in self.groups:
            self._group_step(G)

    def _lanes(self):
        """Group indices per stream: largest groups first, each to the stream
        with the fewest pixels so far."""
        k = min(self.streams, len(self.groups))
        load = [0] * k
        lanes = [[] for _ in range(k)]
        work = lambda gi: len(self.groups[gi]['idx']) * self.groups[gi]['x'][0].numel()
        for gi in sorted(range(len(self.groups)), key=work, reverse=True):
            q = load.index(min(load))
            lanes[q].append(gi)
            load[q] += work(gi)
        return lanes

    def _forked_step(self, side):
        """One step with the groups spread over the current stream and the
        streams in side, which fork from and join back into the current one."""
        cur = torch.cuda.current_stream()
        lanes = self._lanes()
        for ss in side[:len(lanes) - 1]:
            ss.wait_stream(cur)
        for q, lane in enumerate(lanes):
            with torch.cuda.stream(cur if q == 0 else side[q - 1]):
                for gi in lane:
                    self._group_step(self.groups[gi])
        for ss in side[:len(lanes) - 1]:
            cur.wait_stream(ss)

    def capture(self):
        """Record one step of every group in a HIP graph; returns replay()."""
        s = torch.cuda.Stream(device=self.device)
        side = [torch.cuda.Stream(device=self.device) for _ in range(self.streams - 1)]
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.step()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=s):
                if side:
                    self._forked_step(side)
                else:
                    self.step()
        torch.cuda.current_stream().wait_stream(s)
        self._graph = g
        return g.replay

    def outputs(self, i):
        gi, j = self.where[i]
        G = self.groups[gi]
        return G['y'][j], G['xb'][j], G['fb'][j]
