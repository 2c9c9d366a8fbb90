"""Frames of uneven size, patch grid and PSF support in one step.

    mb = MixedBatch([(sx_i, csf_i, sf_i) for each frame], overlap=0.5)
    mb.load(xs, fs, gs)            # per-frame device tensors -> resident stacks
    mb.step()                      # forward + both exact adjoints of every frame
    y_i, xbar_i, fbar_i = mb.outputs(i)

Frames are grouped by geometry; each group is one plan whose launches carry
all patches of all its frames (the kernels take the frame count as a grid
dimension).  ``capture()`` records the launches of every group in one HIP
graph, so a step of the whole mixed set is one graph replay.  With
``tune=True`` each group's FFT tile comes from the autotuner.
"""
from collections import OrderedDict

import torch

from .imagetools import win2winaux
from .ola import OlaPlan


class MixedBatch:
    def __init__(self, specs, overlap=0.5, tune=False, exact=True, device=None):
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

    def step(self):
        for G in self.groups:
            p = G['plan']
            p.forward(G['x'], G['f'], out=G['y'])
            p.adjoint_xf(G['g'], G['x'], G['f'], exact=self.exact, out_x=G['xb'], out_f=G['fb'])

    def capture(self):
        """Record one step of every group in a HIP graph; returns replay()."""
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.step()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=s):
                self.step()
        torch.cuda.current_stream().wait_stream(s)
        self._graph = g
        return g.replay

    def outputs(self, i):
        gi, j = self.where[i]
        G = self.groups[gi]
        return G['y'][j], G['xb'][j], G['fb'][j]
