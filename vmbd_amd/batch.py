"""Frames of uneven size, patch grid and PSF support in one step.

    mb = MixedBatch([(sx_i, csf_i, sf_i) for each frame], overlap=0.5)
    mb.load(xs, fs, gs)            # per-frame device tensors -> resident stacks
    mb.step()                      # forward + both exact adjoints of every frame
    y_i, xbar_i, fbar_i = mb.outputs(i)

Frames are grouped by geometry; each group is one plan over the resident
stacks of its frames (the frame is a grid dimension of every pass).  Two
launch forms:

* ``grouped=True`` (default): the library's batch (vmbd_batch_*) runs each
  pass of several groups as one launch over a device table of groups, one
  launch per pass and distinct tile (the row, overlap-add and stack passes
  are specialised to the tile width, the column passes to its height).  With
  ``streams=1`` one batch holds every group (64 frames of six geometries on
  three tiles: 22 launches per step instead of 42); with ``streams=k > 1``
  the groups are split by tile into one batch per tile and ``capture()``
  forks those batches over k streams (24 launches on three streams).
  Groups whose image rows are shared per patch column (separable windows,
  the default, at launches that fill the GPU) transform them in a row pass
  of their own.
* ``grouped=False``: the groups' own launches, seven per group; with
  ``streams=k > 1`` (default 3) ``capture()`` forks the groups over k
  streams inside the graph (largest groups first, each to the stream with
  the fewest pixels so far) so that one group's launches fill the tail of
  another's.

``capture()`` records one step in a HIP graph and returns its replay.  With
``tune=True`` each group's FFT tile and its plan's launch shapes come from
the autotuner (the grouped launches take the tile; their launch shapes are
chosen per launch).
"""
import ctypes
from collections import OrderedDict

import torch

from . import _lib
from .imagetools import win2winaux
from .ola import OlaPlan


class MixedBatch:
    def __init__(self, specs, overlap=0.5, tune=False, exact=True, device=None, streams=3, grouped=True):
        if int(streams) < 1:
            raise ValueError(f'streams must be >= 1, got {streams}')
        self.streams = int(streams)
        self.grouped = bool(grouped)
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
            cfg = dict(tile=None, ola_threads=0, col_threads=0)
            if tune:
                from .autotune import best_config
                cfg, _ = best_config(sx, sf, wa, frames=len(idx), overlap=overlap, device=self.device)
            plan = OlaPlan(sx, sf, wa, 'valid', tile=cfg['tile'], device=self.device, max_frames=len(idx))
            plan.set_tuning(cfg['ola_threads'], cfg['col_threads'])
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
        self._vb = None

    def __del__(self):
        vbs = getattr(self, '_vb', None)
        if vbs:
            try:
                torch.cuda.synchronize(self.device)
                for vb in vbs.values():
                    _lib.lib().vmbd_batch_destroy(vb)
            except Exception:
                pass
            self._vb = None

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

    # -- grouped launches -----------------------------------------------------
    def _parts(self):
        """Group indices per library batch: all groups in one batch on one
        stream; split by tile when the step forks over several streams."""
        if self.streams == 1:
            return {None: list(range(len(self.groups)))}
        parts = OrderedDict()
        for gi, G in enumerate(self.groups):
            parts.setdefault(G['plan'].tile, []).append(gi)
        return parts

    def _batches(self):
        """The library batches bound to the group stacks (created once)."""
        if self._vb is None:
            L = _lib.lib()
            self._vb = OrderedDict()
            for key, gis in self._parts().items():
                gs = [self.groups[gi] for gi in gis]
                n = len(gs)
                arr = lambda vals: (ctypes.c_void_p * n)(*vals)
                ptrs = lambda k: arr([G[k].data_ptr() for G in gs])
                nfr = (ctypes.c_int32 * n)(*[len(G['idx']) for G in gs])
                with torch.cuda.device(self.device):
                    h = L.vmbd_batch_create(arr([G['plan']._h.value for G in gs]), n, nfr, ptrs('x'), ptrs('f'),
                                            ptrs('g'), ptrs('y'), ptrs('xb'), ptrs('fb'), int(self.exact))
                if not h:
                    raise _lib.VmbdError(f'vmbd_batch_create: {_lib.last_error()}', L.vmbd_last_status())
                self._vb[key] = ctypes.c_void_p(h)
        return self._vb

    @property
    def launches(self):
        """Kernel launches per step."""
        if self.grouped:
            return sum(int(_lib.lib().vmbd_batch_launches(vb)) for vb in self._batches().values())
        return 7 * len(self.groups)

    def _batch_step(self, vb):
        _lib.check(_lib.lib().vmbd_batch_step(vb, _lib.stream_ptr()), 'vmbd_batch_step')

    def _mark_written(self):
        for G in self.groups:
            _lib.wrote(G['y'], G['xb'], G['fb'])
            G['plan']._rows_key = None

    # -- steps ----------------------------------------------------------------
    def _group_step(self, G):
        p = G['plan']
        p.forward(G['x'], G['f'], out=G['y'])
        p.adjoint_xf(G['g'], G['x'], G['f'], exact=self.exact, out_x=G['xb'], out_f=G['fb'])

    def step(self):
        if self.grouped:
            for vb in self._batches().values():
                self._batch_step(vb)
            self._mark_written()
            return
        for G in self.groups:
            self._group_step(G)

    def _lanes(self, units, work):
        """Units per stream: largest first, each to the stream with the
        fewest pixels so far."""
        k = min(self.streams, len(units))
        load = [0] * k
        lanes = [[] for _ in range(k)]
        for u in sorted(units, key=work, reverse=True):
            q = load.index(min(load))
            lanes[q].append(u)
            load[q] += work(u)
        return lanes

    def _forked_step(self, side):
        """One step with the groups (or, grouped, the per-tile batches) spread
        over the current stream and the streams in side, which fork from and
        join back into the current one."""
        cur = torch.cuda.current_stream()
        gwork = lambda gi: len(self.groups[gi]['idx']) * self.groups[gi]['x'][0].numel()
        if self.grouped:
            parts = self._parts()
            vbs = self._batches()
            lanes = self._lanes(list(parts), lambda key: sum(gwork(gi) for gi in parts[key]))
            run = lambda key: self._batch_step(vbs[key])
        else:
            lanes = self._lanes(list(range(len(self.groups))), gwork)
            run = lambda gi: self._group_step(self.groups[gi])
        for ss in side[:len(lanes) - 1]:
            ss.wait_stream(cur)
        for q, lane in enumerate(lanes):
            with torch.cuda.stream(cur if q == 0 else side[q - 1]):
                for u in lane:
                    run(u)
        for ss in side[:len(lanes) - 1]:
            cur.wait_stream(ss)
        if self.grouped:
            self._mark_written()

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
