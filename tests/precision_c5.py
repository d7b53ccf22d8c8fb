"""Precision report for the 64-ring scanner (64 rings x 672 crystals,
309 657 600 LORs, 256x256x128 image at 1.6 mm, geometric walk): the GPU's
plain path (FP64 walk, FP32 values) against the all-FP64 oracle.

Not collected by pytest (a long measurement, not a test). Two steps:

    python tests/precision_c5.py reference --out DIR [--threads N]
        CPU, FP64 oracle: forward projection of the phantom (every 997th bin
        kept), backprojection of the count sinogram, 24-subset sensitivity
        and OSEM 2 x 24 from a uniform image. Writes DIR/*.npy.
    python tests/precision_c5.py gpu --ref DIR
        GPU: the same four results through the C-ABI; prints one JSON line
        with their errors against DIR and the GPU timings.

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


def gpu(ref):
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
    print(json.dumps(out), flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["reference", "gpu"])
    ap.add_argument("--out", default="precision_c5_ref")
    ap.add_argument("--ref", default=None)
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    if a.mode == "reference":
        reference(a.out, a.threads)
    else:
        gpu(a.ref)


if __name__ == "__main__":
    main()
