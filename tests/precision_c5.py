"""Precision report for the 64-ring scanner (64 rings x 672 crystals,
309 657 600 LORs, 256x256x128 image at 1.6 mm, geometric walk): the GPU's
plain path (FP64 walk, FP32 values) against the all-FP64 oracle.

Not collected by pytest (a long measurement, not a test). Two steps:

    python tests/precision_c5.py reference --out DIR [--threads N]
        CPU, FP64 oracle: forward projection of the phantom (every 997th bin
        kept), backprojection of the count sinogram, 24-subset sensitivity
        and OSEM 2 x 24 from a uniform image. Writes DIR/*.npy.
    python tests/precision_c5.py gpu --ref DIR [--partial-cache]
        GPU: the same four results through the C-ABI; prints one JSON line
        with their errors against DIR and the GPU timings. With
        --partial-cache, also a partial path cache (snorm16 lengths with a
        per-LOR FP32 scale; the subsets that fit in device memory, the
        others walked on the fly): its forward projection over the cached
        subsets' sampled bins, the backprojection and the OSEM 2 x 24 run.

The count sinogram is y[b] = (b * 2654435761 mod 2^32) >> 27 (integers 0-31):
deterministic and identical on both sides without shipping 1.2 GB of data.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from fir_amd import synth  # noqa: E402

SAMPLE = 997          # forward projection: every 997th bin is compared
N_ITER, N_SUB = 2, 24


def counts(n_bins):
    y = np.empty(n_bins, np.float32)
    step = 1 << 26
    for a in range(0, n_bins, step):
        b = np.arange(a, min(a + step, n_bins), dtype=np.uint64)
        y[a:a + b.size] = (((b * np.uint64(2654435761)) & np.uint64(0xffffffff))
                           >> np.uint64(27)).astype(np.float32)
    return y


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def reference(out, threads):
    import pyoracle as O
    w = synth.c5()
    os.makedirs(out, exist_ok=True)
    xy, z = O.scanner_tables(w.scanner)
    act, _ = synth.phantom(w)
    t = time.time()
    f = O.forward(w.vol, w.proj, xy, z, act.ravel().astype(np.float64),
                  geometric=True, nthreads=threads)
    np.save(os.path.join(out, "fwd_samples.npy"), f[::SAMPLE])
    log(f"forward {time.time() - t:.0f} s")
    del f
    y = counts(w.n_lors).astype(np.float64)
    t = time.time()
    bp = O.backward(w.vol, w.proj, xy, z, y, 1, geometric=True,
                    nthreads=threads)
    np.save(os.path.join(out, "bp.npy"), bp.astype(np.float32))
    log(f"backprojection {time.time() - t:.0f} s")
    t = time.time()
    sens = O.sensitivity(w.vol, w.proj, xy, z, N_SUB, dtype=np.float64,
                         geometric=True, nthreads=threads)
    log(f"sensitivity {time.time() - t:.0f} s")
    t = time.time()
    x = O.osem(w.vol, w.proj, xy, z, y, sens,
               np.ones(w.vol.n_voxels, np.float64), N_ITER, N_SUB,
               geometric=True, nthreads=threads)
    np.save(os.path.join(out, "osem.npy"), x.astype(np.float32))
    np.save(os.path.join(out, "sens_sum.npy"),
            sens.reshape(N_SUB, -1).sum(0).astype(np.float32))
    log(f"OSEM {time.time() - t:.0f} s")


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rel_max(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.abs(a - b).max() / np.abs(b).max())


def sample_views(spec, n_bins):
    """View of each sampled bin b = 0, SAMPLE, 2 SAMPLE, ...: segments in
    order, each [view][axial][tang] (ProjData.h:16-21)."""
    import pyoracle as O
    p = O.proj_desc(spec)
    na = np.asarray(p.n_axial, np.int64)
    base = np.concatenate([[0], np.cumsum(na * p.n_views * p.n_tang)])
    b = np.arange(0, n_bins, SAMPLE, dtype=np.int64)
    k = np.searchsorted(base, b, side="right") - 1
    return (b - base[k]) // p.n_tang // na[k]


def hybrid(w, ctx, v, x, y, sens, ref, out):
    """Partial path cache + on-the-fly walk for the other subsets."""
    import torch
    from fir_amd import lib as firlib
    nb, nv = ctx.n_bins, v.n_voxels
    st = torch.cuda.current_stream()
    sino = torch.zeros(nb, device="cuda")
    frames = torch.zeros(N_SUB * nv, device="cuda")
    xo = torch.ones(nv, device="cuda")
    bpo = torch.zeros(nv, device="cuda")
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    t = time.perf_counter()
    try:
        cache = ctx.path_cache(v, N_SUB, firlib.FIR_LEN_SNORM16,
                               row_order=firlib.CACHE_ROW_ORDER, max_bytes=0)
    except firlib.FirError as e:
        out["partial_cache"] = {"error": str(e)}
        return
    torch.cuda.synchronize()
    k = cache.n_cached
    info = cache.info()
    h = out["partial_cache"] = {
        "len_format": "snorm16 (per-LOR FP32 scale)", "n_cached": k,
        "n_subsets": N_SUB, "elements": info["elements"],
        "device_gb": round(info["device_bytes"] / 1e9, 2),
        "build_s": round(time.perf_counter() - t, 2)}
    log(f"partial cache: {k} of {N_SUB} subsets, {h['device_gb']} GB")

    def timed(fn):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        fn()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    h["fwd_cached_subsets_ms"] = timed(lambda: cache.forward(x, sino))
    f = sino[::SAMPLE].cpu().numpy()
    del sino

    def bp_all():
        frames.zero_()
        for s in range(k, N_SUB):
            ctx.backward_subset(v, y, frames[s * nv:(s + 1) * nv], s, N_SUB)
        cache.backward(y, frames)
    h["bp_ms"] = timed(bp_all)
    bp = frames.view(N_SUB, -1).sum(0).cpu().numpy()

    def osem():
        total = N_ITER * N_SUB
        for it in range(N_ITER):
            for s in range(N_SUB):
                sub = it * N_SUB + s + 1
                sens_s = sens[s * nv:(s + 1) * nv]
                if s < k:
                    cache.osem_project([xo], [y], None, [bpo], s)
                    ctx.osem_update(v, xo, bpo, sens_s, reset=sub != total)
                else:
                    ctx.osem_subiter(v, xo, y, None, sens_s, bpo, s, N_SUB,
                                     reset=sub != total)
    h["osem_2x24_ms"] = timed(osem)
    h["cached_subiter_ms"] = timed(
        lambda: cache.osem_project([xo], [y], None, [bpo], 0))
    if ref:
        rf = np.load(os.path.join(ref, "fwd_samples.npy"))
        cached = sample_views(w.proj, nb) % N_SUB < k
        fc, rc = f[cached], rf[cached]
        big = rc > 1e-3 * rc.max()
        h["fwd_vs_fp64"] = {
            "rel_l2": rel_l2(fc, rc),
            "rel_max_over_significant_bins":
                float((np.abs(fc - rc)[big] / rc[big]).max()),
            "bins_compared": int(cached.sum())}
        rb = np.load(os.path.join(ref, "bp.npy"))
        h["bp_vs_fp64"] = {"rel_l2": rel_l2(bp, rb), "rel_max": rel_max(bp, rb)}
        h["osem_2x24_vs_fp64"] = {"rel_l2": rel_l2(
            xo.cpu().numpy(), np.load(os.path.join(ref, "osem.npy")))}
    cache.close()


def gpu(ref, partial_cache=False):
    import torch
    from fir_amd import lib as firlib
    w = synth.c5()
    ctx = firlib.context_for(w, geometric=True)
    v = firlib.VolDesc.from_spec(w.vol)
    nb, nv = ctx.n_bins, v.n_voxels
    act, _ = synth.phantom(w)
    x = torch.from_numpy(act.ravel()).cuda()
    st = torch.cuda.current_stream()

    def timed(fn):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        fn()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    out = {"workload": w.name, "n_lors": nb, "walk": "geometric"}
    sino = torch.zeros(nb, device="cuda")
    ctx.forward(v, x, sino)
    out["fwd_ms"] = timed(lambda: ctx.forward(v, x, sino))
    f = sino[::SAMPLE].cpu().numpy()
    log("forward done")
    del sino
    y = torch.from_numpy(counts(nb)).cuda()
    bp = torch.zeros(nv, device="cuda")
    out["bp_ms"] = timed(lambda: (bp.zero_(), ctx.backward(v, y, bp, 1)))
    log("backprojection done")
    sens = torch.zeros(N_SUB * nv, device="cuda")
    out["sens_ms"] = timed(lambda: ctx.sensitivity(v, sens, N_SUB))
    xo = torch.ones(nv, device="cuda")
    bpo = torch.zeros(nv, device="cuda")

    def osem():
        total = N_ITER * N_SUB
        for it in range(N_ITER):
            for s in range(N_SUB):
                sub = it * N_SUB + s + 1
                ctx.osem_subiter(v, xo, y, None, sens[s * nv:(s + 1) * nv],
                                 bpo, s, N_SUB, reset=sub != total)
    out["osem_2x24_ms"] = timed(osem)
    log("OSEM done")
    if ref:
        rf = np.load(os.path.join(ref, "fwd_samples.npy"))
        big = rf > 1e-3 * rf.max()
        out["fwd_vs_fp64"] = {
            "rel_l2": rel_l2(f, rf),
            "rel_max_over_significant_bins":
                float((np.abs(f - rf)[big] / rf[big]).max()),
            "bins_compared": int(rf.size)}
        out["bp_vs_fp64"] = {"rel_l2": rel_l2(bp.cpu().numpy(),
                                               np.load(os.path.join(ref, "bp.npy"))),
                             "rel_max": rel_max(bp.cpu().numpy(),
                                                np.load(os.path.join(ref, "bp.npy")))}
        out["sens_vs_fp64"] = {"rel_l2": rel_l2(
            sens.view(N_SUB, -1).sum(0).cpu().numpy(),
            np.load(os.path.join(ref, "sens_sum.npy")))}
        out["osem_2x24_vs_fp64"] = {"rel_l2": rel_l2(
            xo.cpu().numpy(), np.load(os.path.join(ref, "osem.npy")))}
    if partial_cache:
        del bp, bpo, xo
        hybrid(w, ctx, v, x, y, sens, ref, out)
    print(json.dumps(out), flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["reference", "gpu"])
    ap.add_argument("--out", default="precision_c5_ref")
    ap.add_argument("--ref", default=None)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--partial-cache", action="store_true")
    a = ap.parse_args()
    if a.mode == "reference":
        reference(a.out, a.threads)
    else:
        gpu(a.ref, a.partial_cache)


if __name__ == "__main__":
    main()
