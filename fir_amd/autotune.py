"""Autotuner of the projector kernels: picks, by timing the real workload on
the device, the LOR sort key, the lanes per LOR / per voxel and the z-slice
interleave of a path cache's layout, the LOR sort key and lanes per LOR of
the multi-solve steps, and the row order and per-XCD scatter copies of the
plain projector.

Every candidate changes only the FP32 summation order of the results (the
tests check that), so the choice is purely a matter of speed. Timings are
HIP events around the OSEM projector pass of one subset on the caller's
stream, best of ``reps`` repetitions after one warm-up launch.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import torch

from . import lib as firlib

# Row orders a cache is tried with: per view (tangential bin, pair by centre
# slice) and the sinogram's own (pair, tangential bin).
CACHE_ROW_ORDERS = (firlib.CACHE_ROW_ORDER, firlib.FIR_ORDER_VIEW_MAJOR)
CACHE_LANES = ((8, 16), (16, 16), (8, 8))
# Layouts tried: (row order, lanes per LOR, lanes per voxel, z-slices).
CACHE_LAYOUTS = ((firlib.CACHE_ROW_ORDER, 8, 16, 1),
                 (firlib.CACHE_ROW_ORDER, 8, 16, 2),
                 (firlib.CACHE_ROW_ORDER, 8, 16, 4),
                 (firlib.CACHE_ROW_ORDER, 8, 16, 8),
                 (firlib.CACHE_ROW_ORDER, 16, 16, 4),
                 (firlib.CACHE_ROW_ORDER, 8, 8, 4),
                 (firlib.FIR_ORDER_VIEW_MAJOR, 8, 16, 4))
# Layouts the multi-solve steps are tried with: (row order, lanes per LOR).
# The steps take their line integrals from slab-staged sheets, which need a
# family's rows together (tangential-major rows per view); the order of the
# ring pairs within a family (by centre slice or by segment) and the lanes
# per LOR are free.
MULTI_LAYOUTS = ((firlib.CACHE_ROW_ORDER, 4),
                 (firlib.CACHE_ROW_ORDER, 2),
                 (firlib.CACHE_ROW_ORDER, 8),
                 (firlib.FIR_ORDER_VIEW_MAJOR | firlib.FIR_ORDER_TANG_MAJOR, 4),
                 (firlib.FIR_ORDER_VIEW_MAJOR | firlib.FIR_ORDER_TANG_MAJOR, 2))
PLAIN_ORDERS = (firlib.FIR_ORDER_SEGMENT,
                firlib.FIR_ORDER_SEGMENT | firlib.FIR_ORDER_TANG_MAJOR,
                firlib.FIR_ORDER_VIEW_MAJOR | firlib.FIR_ORDER_ZCENTER)
PLAIN_COPIES = (1, 8)


@dataclass
class Trial:
    config: dict
    ms: float


@dataclass
class Result:
    best: dict
    best_ms: float
    trials: List[Trial] = field(default_factory=list)

    def table(self) -> List[dict]:
        return [dict(t.config, ms=round(t.ms, 4)) for t in self.trials]


def _time(fn, reps: int) -> float:
    stream = torch.cuda.current_stream()
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def tune_cache(ctx: firlib.Context, vol: firlib.VolDesc, n_subsets: int,
               xs: Sequence[torch.Tensor], ys: Sequence[torch.Tensor],
               len_format: int = firlib.FIR_LEN_SNORM16,
               layouts: Sequence[Tuple[int, int, int, int]] = CACHE_LAYOUTS,
               subset: int = 0, reps: int = 3
               ) -> Tuple[firlib.PathCache, Result]:
    """Build a path cache per layout (row order, lanes per LOR and per voxel,
    z-slice interleave: Context.set_cache_lanes / set_cache_zslices), time
    the OSEM projector pass of ``len(xs)`` images on each, and return the
    fastest cache with the table of trials. Only the best cache so far is
    kept while the next is built."""
    bps = [torch.zeros(vol.n_voxels, device=xs[0].device) for _ in xs]
    trials: List[Trial] = []
    best = None
    best_cache: Optional[firlib.PathCache] = None
    old_lanes = getattr(ctx, "_cache_lanes", (8, 16))
    old_zq = getattr(ctx, "_cache_zq", 4)
    try:
        for order, fl, bl, zq in layouts:
            ctx.set_cache_lanes(fl, bl)
            ctx.set_cache_zslices(zq)
            cache = ctx.path_cache(vol, n_subsets, len_format, row_order=order)
            ms = _time(lambda: cache.osem_project(list(xs), list(ys), None,
                                                  bps, subset), reps)
            cfg = {"row_order": order, "fwd_lanes": fl, "bp_lanes": bl,
                   "zslices": zq}
            trials.append(Trial(cfg, ms))
            if best is None or ms < best[1]:
                if best_cache is not None:
                    best_cache.close()
                best, best_cache = (cfg, ms), cache
            else:
                cache.close()
            torch.cuda.synchronize()
    finally:
        ctx.set_cache_lanes(*old_lanes)
        ctx.set_cache_zslices(old_zq)
    return best_cache, Result(best[0], best[1], trials)


def tune_plain(ctx: firlib.Context, vol: firlib.VolDesc, n_subsets: int,
               x: torch.Tensor, y: torch.Tensor,
               orders: Sequence[int] = PLAIN_ORDERS,
               copies: Sequence[int] = PLAIN_COPIES,
               subset: int = 0, reps: int = 3) -> Result:
    """Time the plain (on-the-fly walk) OSEM projector pass for every row
    order and scatter-copy count, leave the fastest set on the context and
    return the table."""
    bp = torch.zeros(vol.n_voxels, device=x.device)
    trials: List[Trial] = []
    for order in orders:
        ctx.set_lor_order(order)
        for cp in copies:
            ctx.set_scatter_copies(cp)
            ms = _time(lambda: ctx.osem_project(vol, x, y, None, bp, subset,
                                                n_subsets), reps)
            trials.append(Trial({"lor_order": order, "scatter_copies": cp},
                                ms))
    best = min(trials, key=lambda t: t.ms)
    ctx.set_lor_order(best.config["lor_order"])
    ctx.set_scatter_copies(best.config["scatter_copies"])
    return Result(best.config, best.ms, trials)


def tune_multi(ctx: firlib.Context, vol: firlib.VolDesc, n_subsets: int,
               ys: Sequence[torch.Tensor], fwhm=(0.0, 0.0, 0.0),
               len_format: int = firlib.FIR_LEN_SNORM16,
               layouts: Sequence[Tuple[int, int]] = MULTI_LAYOUTS,
               subset: int = 0, reps: int = 3) -> Result:
    """Time one multi-solve step (fir_osem_step: the projector phases,
    update and Gaussian) of one job of ``len(ys)`` images (up to 8, from
    x = 1) on a cache of each layout (row order, lanes per LOR), and leave
    the fastest lanes on the context. Layouts whose cache lacks the slab
    form or the family lists, which the steps need, are skipped; every cache
    is freed. The chosen row order is ``best["row_order"]``."""
    dev = ys[0].device
    nv = vol.n_voxels
    bps = [torch.zeros(nv, device=dev) for _ in ys]
    trials: List[Trial] = []
    old_lanes = getattr(ctx, "_cache_lanes", (4, 16))
    best = None
    try:
        for order, lanes in layouts:
            ctx.set_cache_lanes(lanes, old_lanes[1])
            cache = ctx.path_cache(vol, n_subsets, len_format, row_order=order)
            try:
                forms = cache.forms()
                if not (forms["slabs"] and forms["family_lists"]):
                    continue
                sens = torch.zeros(n_subsets * nv, device=dev)
                cache.sensitivity(sens)
                xs = [torch.ones(nv, device=dev) for _ in ys]
                nb = cache.job_workspace()
                raw = torch.empty(nb + 256, dtype=torch.uint8, device=dev)
                off = (-raw.data_ptr()) % 256
                st = firlib.OsemStep([dict(
                    cache=cache, subset=subset, x=xs, y=list(ys), bp=bps,
                    sens=sens[subset * nv:(subset + 1) * nv],
                    workspace=raw[off:off + nb])], fwhm)
                try:
                    ms = _time(st.run, reps)
                finally:
                    st.close()
                torch.cuda.synchronize()
            finally:
                cache.close()
            cfg = {"row_order": order, "fwd_lanes": lanes}
            trials.append(Trial(cfg, ms))
            if best is None or ms < best.ms:
                best = trials[-1]
    finally:
        if best is None:
            ctx.set_cache_lanes(*old_lanes)
    if best is None:
        raise RuntimeError("no layout gives the multi-solve steps' cache "
                           "forms (slab form and family lists)")
    ctx.set_cache_lanes(best.config["fwd_lanes"], old_lanes[1])
    return Result(best.config, best.ms, trials)
