"""Device-resident OSEM drivers over the C-ABI.

``OSEMGroup`` runs several OSEM reconstructions that share the scanner,
the sinogram geometry, the image geometry and the subset count (noise
realisations or dynamic frames of one scan): every sub-iteration is one
batched projector launch (fir_osem_project_batch: one walk per LOR serves
up to 8 images) followed by the per-image update, Gaussian and cut. After a
first eager iteration, one whole iteration is captured in a HIP graph and
replayed.

With a ``PathCache`` the projector pass streams the cached system matrix
(fir_cache_osem_project) instead of walking every LOR.

``BatchedOSEM`` spreads reconstructions of uneven size (different image
grids, subset counts and count levels) over such groups and captures one
iteration of all of them in a single HIP graph.

Per sub-iteration each image follows reconAlgos::OSEM
(reconAlgos.cc:119-186): projector pass, bp /= sens, x *= bp, Gaussian every
``conv_interval`` sub-iterations (with edge restore), circular cut.
"""
from __future__ import annotations

import time
from typing import List, Optional, Sequence

import torch

from . import autotune
from . import lib as firlib


class OSEMGroup:
    def __init__(self, ctx: firlib.Context, vol: firlib.VolDesc,
                 n_subsets: int, ys: Sequence[torch.Tensor],
                 biases: Optional[Sequence[Optional[torch.Tensor]]] = None,
                 sens: Optional[torch.Tensor] = None,
                 fwhm=(0.0, 0.0, 0.0), conv_interval: int = 0,
                 cut: float = 0.0, x0: Optional[Sequence[torch.Tensor]] = None,
                 cache: Optional[firlib.PathCache] = None):
        self.ctx, self.vol, self.n_subsets = ctx, vol, n_subsets
        if cache is not None and cache.n_subsets != n_subsets:
            raise ValueError("the path cache was built for "
                             f"{cache.n_subsets} subsets, not {n_subsets}")
        self.cache = cache
        nv = vol.n_voxels
        self.ys = [y.contiguous() for y in ys]
        self.biases = list(biases) if biases is not None else None
        self.fwhm = tuple(float(f) for f in fwhm)
        self.conv_interval = conv_interval
        self.cut = float(cut)
        self.convolve = conv_interval > 0 and all(f > 0 for f in self.fwhm)
        if self.convolve and n_subsets % conv_interval != 0:
            raise ValueError("the convolution schedule must repeat every "
                             "iteration (conv_interval must divide the "
                             "number of subsets)")
        dev = self.ys[0].device
        if sens is None:
            sens = torch.zeros(n_subsets * nv, device=dev)
            if cache is not None:
                cache.sensitivity(sens)
            else:
                ctx.sensitivity(vol, sens, n_subsets)
        self.sens = sens
        self.x = ([t.clone() for t in x0] if x0 is not None else
                  [torch.ones(nv, device=dev) for _ in self.ys])
        self.bp = [torch.zeros(nv, device=dev) for _ in self.ys]
        for x in self.x:
            ctx.cut_circle(vol, x, self.cut)
        self.graph = None
        self.iterations_done = 0

    @property
    def n_images(self) -> int:
        return len(self.ys)

    def _subiteration(self, subset: int):
        nv = self.vol.n_voxels
        ctx, vol = self.ctx, self.vol
        if self.cache is not None:
            self.cache.osem_project(self.x, self.ys, self.biases, self.bp,
                                    subset)
        else:
            ctx.osem_project_batch(vol, self.x, self.ys, self.biases, self.bp,
                                   subset, self.n_subsets)
        sens_s = self.sens[subset * nv:(subset + 1) * nv]
        conv = self.convolve and (subset + 1) % self.conv_interval == 0
        for x, bp in zip(self.x, self.bp):
            ctx.osem_update(vol, x, bp, sens_s, self.cut,
                            apply_cut=not conv, reset=True)
            if conv:
                ctx.convolve(vol, x, self.fwhm, self.cut)
                ctx.cut_circle(vol, x, self.cut)

    def iteration_eager(self):
        for s in range(self.n_subsets):
            self._subiteration(s)
        self.iterations_done += 1

    def capture(self):
        """Capture one iteration in a HIP graph (after one eager iteration,
        which also sizes the library's Gaussian scratch and taps, so that
        nothing allocates or synchronizes during capture)."""
        if self.iterations_done == 0:
            self.iteration_eager()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.ctx.set_stream(torch.cuda.current_stream())
            for s in range(self.n_subsets):
                self._subiteration(s)
        self.ctx.set_stream(torch.cuda.current_stream())
        self.graph = g

    def iteration(self):
        if self.graph is None:
            self.iteration_eager()
        else:
            self.graph.replay()
            self.iterations_done += 1


class BatchedOSEM:
    """Reconstructions of uneven size grouped by (image grid, subset count).

    ``solves`` is a list of dicts with keys ``vol`` (VolDesc), ``n_subsets``
    and ``y`` (device sinogram); every solve shares one scanner/sinogram
    geometry, given by ``workload``.

    Without a cache format every group gets its own context and the plain
    projector. With ``cache_format`` (a FIR_LEN_* value) all groups share one
    context and each group gets a path cache; caches of one image grid share
    the LOR-major path matrix, so only the per-subset-count transposes are
    extra. ``cache_zslices`` is their layout's z-slice interleave
    (Context.set_cache_zslices) and ``cache_lanes`` the lanes per LOR and
    per voxel their quads are laid out for (Context.set_cache_lanes): with
    4 lanes per LOR, batches of 4 or 8 images take their line integrals
    from slab-staged sheets (FIR_FORM_SLABS).
    """

    def __init__(self, workload, solves: List[dict], fwhm=(0.0, 0.0, 0.0),
                 conv_interval: int = 0, cut: float = 0.0, device: int = 0,
                 cache_format: Optional[int] = None,
                 cache_row_order: int = firlib.CACHE_ROW_ORDER,
                 cache_zslices: int = 2, cache_lanes=(4, 16)):
        keys = []
        for s in solves:
            v = s["vol"]
            k = (v.nx, v.ny, v.nz, v.dx, v.dy, v.dz, v.ox, v.oy, v.oz,
                 s["n_subsets"])
            if k not in keys:
                keys.append(k)
        self.groups = []
        self.members = []
        self.caches = []
        self.graph = None
        shared = None
        if cache_format is not None:
            shared = firlib.context_for(workload, device)
            shared.set_uneven_subsets(True)
            # groups project several images per launch: 4 or 8 take their
            # line integrals from slab-staged sheets (4 lanes per LOR suit
            # them best), 2 gather from interleaved image copies (2
            # interleaved z-slices suit that best); the sheet form's family
            # lists serve the backprojections (and make the voxel-major
            # transposes unnecessary)
            shared.set_cache_zslices(cache_zslices)
            shared.set_cache_lanes(*cache_lanes)
            shared.set_cache_sheets(True)
        self.ctx = shared
        for k in keys:
            idx = [i for i, s in enumerate(solves)
                   if (s["vol"].nx, s["vol"].ny, s["vol"].nz, s["vol"].dx,
                       s["vol"].dy, s["vol"].dz, s["vol"].ox, s["vol"].oy,
                       s["vol"].oz, s["n_subsets"]) == k]
            vol = solves[idx[0]]["vol"]
            cache = None
            if shared is not None:
                ctx = shared
                cache = ctx.path_cache(vol, k[-1], cache_format,
                                       row_order=cache_row_order)
                self.caches.append(cache)
            else:
                ctx = firlib.context_for(workload, device)
                ctx.set_uneven_subsets(True)
            g = OSEMGroup(ctx, vol, k[-1], [solves[i]["y"] for i in idx],
                          fwhm=fwhm, conv_interval=conv_interval, cut=cut,
                          cache=cache)
            self.groups.append(g)
            self.members.append(idx)

    def capture(self):
        """Capture one iteration of every group in a single HIP graph: one
        launch replays the sub-iterations of all solves. Every group first
        runs one eager iteration, so that the contexts' Gaussian scratch and
        taps are sized for all of them before capture."""
        for g in self.groups:
            if g.iterations_done == 0:
                g.iteration_eager()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            st = torch.cuda.current_stream()
            for g in self.groups:
                g.ctx.set_stream(st)
                for s in range(g.n_subsets):
                    g._subiteration(s)
        for g in self.groups:
            g.ctx.set_stream(torch.cuda.current_stream())
        self.graph = graph

    def iteration(self):
        if self.graph is None:
            for g in self.groups:
                g.iteration()
        else:
            self.graph.replay()
            for g in self.groups:
                g.iterations_done += 1

    def images(self) -> List[torch.Tensor]:
        out = [None] * sum(len(m) for m in self.members)
        for g, m in zip(self.groups, self.members):
            for j, i in enumerate(m):
                out[i] = g.x[j]
        return out

    def close(self):
        self.graph = None
        for g in self.groups:
            g.graph = None
        for c in self.caches:
            c.close()
        self.caches = []


class MultiOSEM:
    """Reconstructions of uneven size advanced together through
    fir_osem_step: every phase of a sub-iteration -- the projector (pack,
    line integrals, family backprojection, per-voxel sums), the update and
    the Gaussian -- runs for all solves scheduled at that step in one
    launch, the projector phases as persistent kernels over one table of
    (job, family) work items.

    Solves are grouped by (image grid, subset count), each group on its own
    path cache (caches of one grid share the path matrix), its solves split
    into jobs of up to 8 images. One iteration of every solve takes K =
    max(subset counts) steps: a group with S subsets runs its sub-iteration
    s at step floor(s * K / S), so each solve still visits its subsets in
    order (reconAlgos.cc:101-199) and the steps' work stays spread out. The
    iteration's K steps are captured in one HIP graph and replayed.

    ``solves`` as for BatchedOSEM; conv_interval 0 (no Gaussian) or 1
    (Gaussian every sub-iteration). With ``autotune_layout`` the caches'
    LOR sort key (row order) and lanes per LOR are the fastest of
    autotune.MULTI_LAYOUTS for one step of the largest group's first job
    (``self.tuning``); otherwise ``cache_row_order`` and ``cache_lanes``.
    Each image gets what BatchedOSEM gives it on caches of the same layout
    (``self.cache_row_order``, ``self.cache_lanes``)."""

    def __init__(self, workload, solves: List[dict], fwhm=(0.0, 0.0, 0.0),
                 conv_interval: int = 0, cut: float = 0.0, device: int = 0,
                 cache_format: int = firlib.FIR_LEN_SNORM16,
                 cache_row_order: int = firlib.CACHE_ROW_ORDER,
                 cache_zslices: int = 2, cache_lanes: int = 4,
                 autotune_layout: bool = False):
        if conv_interval not in (0, 1):
            raise ValueError("MultiOSEM convolves every sub-iteration or "
                             "never (conv_interval 1 or 0)")
        self.fwhm = tuple(float(f) for f in fwhm)
        conv = conv_interval == 1 and all(f > 0 for f in self.fwhm)
        self.cut = float(cut)
        ctx = firlib.context_for(workload, device)
        ctx.set_uneven_subsets(True)
        ctx.set_cache_zslices(cache_zslices)
        ctx.set_cache_lanes(cache_lanes, 16)  # the slab form's lanes per LOR
        ctx.set_cache_sheets(True)
        self.ctx = ctx
        dev = solves[0]["y"].device
        keys, self.members, self.caches, self.sens = [], [], [], []
        for i, s in enumerate(solves):
            v = s["vol"]
            k = (v.nx, v.ny, v.nz, v.dx, v.dy, v.dz, v.ox, v.oy, v.oz,
                 s["n_subsets"])
            if k not in keys:
                keys.append(k)
                self.members.append([])
            self.members[keys.index(k)].append(i)
        self.x, self.bp, self.y = [], [], []
        for s in solves:
            nv = s["vol"].n_voxels
            x = torch.ones(nv, device=dev)
            ctx.cut_circle(s["vol"], x, self.cut)
            self.x.append(x)
            self.bp.append(torch.zeros(nv, device=dev))
            self.y.append(s["y"].contiguous())
        self.tuning, self.tuning_s = None, 0.0
        if autotune_layout:
            t0 = time.perf_counter()
            # the group with the most solves (then the largest grid) and the
            # images of its first job
            gi = max(range(len(keys)), key=lambda g: (
                len(self.members[g]), solves[self.members[g][0]]["vol"].n_voxels))
            mem = self.members[gi][:firlib.JOB_MAX_IMAGES]
            self.tuning = autotune.tune_multi(
                ctx, solves[mem[0]]["vol"], keys[gi][-1],
                [self.y[i] for i in mem],
                self.fwhm if conv else (0.0, 0.0, 0.0), cache_format)
            cache_row_order = self.tuning.best["row_order"]
            cache_lanes = self.tuning.best["fwd_lanes"]
            self.tuning_s = time.perf_counter() - t0
        self.cache_row_order, self.cache_lanes = cache_row_order, cache_lanes
        groups = []
        for k, mem in zip(keys, self.members):
            vol = solves[mem[0]]["vol"]
            S = k[-1]
            cache = ctx.path_cache(vol, S, cache_format,
                                   row_order=cache_row_order)
            self.caches.append(cache)
            sens = torch.zeros(S * vol.n_voxels, device=dev)
            cache.sensitivity(sens)
            self.sens.append(sens)
            # jobs of 8 images and one of the rest (up to 4 images take the
            # cheaper 4-image slot)
            chunks = [mem[a:a + firlib.JOB_MAX_IMAGES]
                      for a in range(0, len(mem), firlib.JOB_MAX_IMAGES)]
            groups.append((cache, sens, S, vol.n_voxels, chunks))
        K = max(g[2] for g in groups)
        plan = [[] for _ in range(K)]
        for cache, sens, S, nv, chunks in groups:
            for sub in range(S):
                k = sub * K // S
                for ch in chunks:
                    plan[k].append((cache, sens, sub, nv, ch))
        n_ws = max(len(p) for p in plan)
        ws_bytes = max(c.job_workspace() for c in self.caches)
        self._ws_raw = []
        self.workspaces = []
        for _ in range(n_ws):
            raw = torch.empty(ws_bytes + 256, dtype=torch.uint8, device=dev)
            off = (-raw.data_ptr()) % 256
            self._ws_raw.append(raw)
            self.workspaces.append(raw[off:off + ws_bytes])
        self.steps = []
        for p in plan:
            jobs = []
            for slot, (cache, sens, sub, nv, ch) in enumerate(p):
                jobs.append(dict(
                    cache=cache, subset=sub,
                    x=[self.x[i] for i in ch], y=[self.y[i] for i in ch],
                    bp=[self.bp[i] for i in ch],
                    sens=sens[sub * nv:(sub + 1) * nv],
                    workspace=self.workspaces[slot]))
            self.steps.append(firlib.OsemStep(
                jobs, self.fwhm if conv else (0.0, 0.0, 0.0), self.cut))
        self.n_steps = K
        self.graph = None
        self.iterations_done = 0

    def iteration_eager(self):
        for st in self.steps:
            st.run()
        self.iterations_done += 1

    def capture(self):
        """One eager iteration, then the iteration's steps captured in one
        HIP graph."""
        if self.iterations_done == 0:
            self.iteration_eager()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.ctx.set_stream(torch.cuda.current_stream())
            for st in self.steps:
                st.run()
        self.ctx.set_stream(torch.cuda.current_stream())
        self.graph = g

    def iteration(self):
        if self.graph is None:
            self.iteration_eager()
        else:
            self.graph.replay()
            self.iterations_done += 1

    def images(self) -> List[torch.Tensor]:
        return self.x

    def close(self):
        self.graph = None
        for st in self.steps:
            st.close()
        self.steps = []
        for c in self.caches:
            c.close()
        self.caches = []
        self.ctx.close()
