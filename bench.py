"""Benchmark: FIR's projector + OSEM hot path on one MI355X.

Workload: the clinical-scale plain path (24 rings x 448 crystals, radius
330 mm, span 1, 47 segments, 325 tangential bins = 41 932 800 LORs;
160x160x47 image at 2x2x3.27 mm), a seeded 12-ellipsoid phantom in a 200 mm
water cylinder, attenuation-corrected Poisson data at 5e8 counts, OSEM with
16 subsets and a 4 mm FWHM Gaussian every sub-iteration.

One step = one full 16-subset OSEM iteration (every LOR forward- and
back-projected once per iteration, plus the per-sub-iteration update and
Gaussian) with the projector streaming the path cache (the Siddon system
matrix kept in HBM with snorm16 lengths; OSEM images within 1e-4 relative L2
of the reference). Inputs and the cache are resident in HBM before timing
starts; the cache build is timed separately. The counts are those of the
parity fixture (synth.parity_counts), so before timing, the golden schedule
(3 iterations) is run on the timed path and on the plain path and compared
with the reference's FIR_OSEM image of the same counts and with the FP64
oracle's (``osem_parity``).

    value = LORs per step / step time  [M LORs/s, fwd+backproj]

The JSON line also carries: the same iteration on the plain path (FP64
Siddon walk per LOR, bit-exact forward), 8 reconstructions batched over one
cache read per sub-iteration in a HIP graph, the separately timed forward
and backprojection passes, the roofline of the dominant kernel pair (the
cached OSEM projector: line integrals + transposed backprojection) and the
reference CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def traffic_from_profiles():
    """Measured HBM bytes per launch of the dominant kernel pair (FETCH_SIZE
    and WRITE_SIZE, calibrated against a copy of known size), from the
    committed rocprofv3 summary profiles/osem_cached_c2_traffic.md."""
    p = os.path.join(ROOT, "profiles", "osem_cached_c2_traffic.md")
    if not os.path.exists(p):
        return None, None
    val, commit = None, None
    for line in open(p):
        if line.startswith("bytes_per_launch:"):
            try:
                val = int(line.split(":", 1)[1])
            except ValueError:
                val = None
        elif line.startswith("commit:"):
            commit = line.split(":", 1)[1].strip()
    return val, commit


FP64_PEAK_TFLOPS = 78.6


def fp64_flop_from_profiles():
    """FP64 operations per walked projector launch (SQ_INSTS_VALU_*_F64
    counters x 64 lanes, FMA counted twice), from the committed rocprofv3
    summary profiles/plain_osem_project_c2_fp64_counters.md."""
    p = os.path.join(ROOT, "profiles", "plain_osem_project_c2_fp64_counters.md")
    if not os.path.exists(p):
        return None, None
    val, commit = None, None
    for line in open(p):
        if line.startswith("fp64_flop_per_launch:"):
            try:
                val = int(line.split(":", 1)[1])
            except ValueError:
                val = None
        elif line.startswith("commit:"):
            commit = line.split(":", 1)[1].strip()
    return val, commit


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def parity_golden(w):
    """The reference's OSEM image of this workload's counts, with the FP64
    oracle's (tests/golden/c2_osem_golden.npz), or None for workloads it
    does not cover."""
    p = os.path.join(ROOT, "tests", "golden", "c2_osem_golden.npz")
    if w.name != "c2_24x448_160x160x47" or not os.path.exists(p):
        return None
    return dict(np.load(p))


def copy_bandwidth(reps=10):
    """Achievable HBM bandwidth: a 1 GiB device-to-device copy (1 GiB read
    and 1 GiB written per copy), mean of reps copies timed with HIP
    events on torch's current stream."""
    import torch
    src = torch.ones(1 << 28, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2 * (1 << 30) / (ms * 1e-3) / 1e9


def host_cpu():
    """CPU model, sockets and logical CPUs of this host (lscpu)."""
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True,
                             timeout=30).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    for line in out.splitlines():
        k, _, v = line.partition(":")
        k, v = k.strip(), v.strip()
        if k == "Model name":
            info["model"] = v
        elif k == "Socket(s)":
            info["sockets"] = v
        elif k == "Core(s) per socket":
            info["cores_per_socket"] = v
        elif k == "Thread(s) per core":
            info["threads_per_core"] = v
    return info


def cpu_baseline(w, n_subsets):
    """Reference FIR (compiled from its own sources into oracle/_ref) timed
    in-process on this host: one 16-subset OSEM iteration at the workload's
    geometry with a uniform image and sinogram. Falls back to the oracle
    restatement if the reference build is absent."""
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    host = host_cpu()
    reps = 3
    if os.path.exists(drv):
        with tempfile.TemporaryDirectory() as d:
            w.scanner.write(os.path.join(d, "s.scan"))
            w.proj.write(os.path.join(d, "p.hs"))
            w.vol.write(os.path.join(d, "v.h33"))
            env = dict(os.environ, OMP_NUM_THREADS=str(threads),
                       OMP_PROC_BIND="close", OMP_PLACES="cores")
            r = subprocess.run([drv, "time", "subiter", "v.h33", "s.scan",
                                "p.hs", str(n_subsets), str(reps)], cwd=d,
                               env=env, capture_output=True, text=True,
                               timeout=900)
            if r.returncode != 0:
                log("reference driver failed:", r.stderr[-500:])
                return None
            res = json.loads(r.stdout.strip().splitlines()[-1])
        t = res["median_s"]
        return {"value": w.n_lors / t / 1e6, "unit": "M LORs/s",
                "cores": res["threads"], "kind": "reference",
                "osem_iter_s": t, "min_s": res["min_s"], "reps": reps,
                "omp": "OMP_PROC_BIND=close OMP_PLACES=cores",
                "host": host,
                "sample": (f"one {n_subsets}-subset OSEM iteration of the "
                           "reference FIR built from source (oracle/_ref), "
                           "same geometry, uniform image and sinogram, "
                           f"median of {reps} timed runs after 1 warm-up")}
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    xy, z = O.scanner_tables(w.scanner)
    sens = np.ones(w.vol.n_voxels * n_subsets, np.float32)
    y = np.ones(w.n_lors, np.float32)
    x0 = np.ones(w.vol.n_voxels, np.float32)
    t0 = time.perf_counter()
    O.osem(w.vol, w.proj, xy, z, y, sens, x0, 1, n_subsets,
           nthreads=threads)
    t = time.perf_counter() - t0
    return {"value": w.n_lors / t / 1e6, "unit": "M LORs/s",
            "cores": threads, "kind": "port", "osem_iter_s": t,
            "host": host,
            "sample": f"one {n_subsets}-subset OSEM iteration of the CPU "
                      "oracle, uniform image and sinogram"}


def run_uneven(args):
    """Batched uneven reconstructions: 64 OSEM solves of the 24-ring data
    (noise realisations / dynamic frames with total counts log-uniform in
    [1e5, 1e8], seeds 100-163) on mixed 128x128x47 and 160x160x47 grids with
    8, 16 or 24 subsets (24 does not divide the 224 views: interleaved
    uneven subsets), 4 mm Gaussian every sub-iteration. Groups of equal
    (grid, subset count) share each read of their path cache (one shared
    context; caches of one grid share the path matrix). One step = one
    iteration of all 64 solves: 24 fir_osem_step launches sets (every phase
    of every solve scheduled at a step in one launch, the projector phases
    persistent over one work queue) captured in one HIP graph, on caches
    whose LOR sort key and lanes per LOR the autotuner picks; the same
    solves through per-group graphs (BatchedOSEM) on caches of the same
    layout are timed beside it.
    """
    import torch
    from fir_amd import interfile as io
    from fir_amd import lib as firlib
    from fir_amd import recon
    from fir_amd import synth

    w = synth.c2()
    ctx = firlib.context_for(w)
    v160 = firlib.VolDesc.from_spec(w.vol)
    v128 = firlib.VolDesc.from_spec(io.VolSpec((128, 128, 47),
                                               (2.0, 2.0, 3.27),
                                               (-127.0, -127.0, 0.0)))
    act, _ = synth.phantom(w)
    y0 = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v160, torch.from_numpy(act.ravel()).cuda(), y0)
    y0 = y0 / y0.sum()
    solves = []
    for i in range(64):
        seed = 100 + i
        total = 10 ** np.random.default_rng(seed).uniform(5.0, 8.0)
        g = torch.Generator(device="cuda").manual_seed(seed)
        y = torch.poisson(y0 * float(total), generator=g)
        solves.append(dict(vol=v128 if i % 2 else v160,
                           n_subsets=(8, 16, 24)[i % 3], y=y))
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)

    def timed(obj):
        obj.capture()
        for _ in range(args.warmup):
            obj.iteration()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(args.steps):
            obj.iteration()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    t0 = time.perf_counter()
    m = recon.MultiOSEM(w, solves, fwhm=w.fwhm, conv_interval=1,
                        autotune_layout=True)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    used = free0 - torch.cuda.mem_get_info()[0]
    tuning, m_tuning_s = m.tuning, m.tuning_s
    log(f"multi-solve steps: {m.n_steps} per iteration, layout tuned "
        f"({tuning.best}), caches and workspaces set up in {setup_s:.1f} s")
    ms = timed(m)
    multi_imgs = [x.cpu() for x in m.images()]
    layout = dict(cache_row_order=m.cache_row_order,
                  cache_lanes=(m.cache_lanes, 16))
    m.close()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b = recon.BatchedOSEM(w, solves, fwhm=w.fwhm, conv_interval=1,
                          cache_format=firlib.FIR_LEN_SNORM16, **layout)
    torch.cuda.synchronize()
    group_setup_s = time.perf_counter() - t0
    log(f"{len(b.groups)} groups, caches built in {group_setup_s:.1f} s")
    n_groups = len(b.groups)
    group_ms = timed(b)
    # both paths ran capture + warmup + steps iterations from x = 1 on
    # caches of the same layout
    same = all(torch.equal(a, x.cpu()) for a, x in zip(multi_imgs,
                                                        b.images()))
    b.close()
    del multi_imgs
    # the plain path, one solve of the same data alone (160 grid, 16 subsets)
    single = recon.OSEMGroup(ctx, v160, 16, [solves[4]["y"]],
                             fwhm=w.fwhm, conv_interval=1)
    single.iteration_eager()
    torch.cuda.synchronize()
    e0.record(stream)
    single.iteration_eager()
    e1.record(stream)
    torch.cuda.synchronize()
    plain_ms = e0.elapsed_time(e1)
    out = {
        "metric": "M LORs/s fwd+backproj & OSEM iter ms: plain vs batched "
                  "GPU at rel err<=1e-4",
        "value": round(64 * w.n_lors / (ms * 1e-3) / 1e6, 3),
        "unit": "M LORs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded phantom, Poisson counts log-uniform "
                "1e5..1e8)",
        "config": {"workload": "uneven_64_solves_c2_data",
                   "grids": "128x128x47 and 160x160x47 @ 2x2x3.27 mm",
                   "subsets": [8, 16, 24], "n_solves": 64,
                   "groups": n_groups,
                   "path": "path caches (snorm16, LOR sort key and lanes "
                           "per LOR autotuned); 24 multi-solve steps per "
                           "iteration, one launch per phase for all solves "
                           "(persistent projector kernels over one work "
                           "queue), captured in one HIP graph"},
        "iteration_ms_all_solves": round(ms, 3),
        "iteration_ms_per_solve": round(ms / 64, 4),
        "per_group_graphs_iteration_ms": round(group_ms, 3),
        "images_identical_to_per_group_graphs": bool(same),
        "plain_single_solve_iteration_ms": round(plain_ms, 3),
        "speedup_per_solve_vs_plain": round(plain_ms * 64 / ms, 3),
        "layout_autotune": {"best": tuning.best,
                            "trials_ms": tuning.table(),
                            "tuning_s": round(m_tuning_s, 2)},
        "cache_setup_s": round(setup_s - m_tuning_s, 2),
        "per_group_cache_setup_s": round(group_setup_s, 2),
        "cache_device_gb": round(used / 1e9, 2),
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def cli_end_to_end(w, counts, hu, n_iter=3):
    """The drop-in FIR_OSEM executable run on the workload's files, end to
    end: process start, Interfile reads, ACF from the HU map, sensitivity,
    n_iter iterations of the workload's subsets with its Gaussian, image
    write. Wall time with the on-the-fly walk and with the path cache."""
    from fir_amd import interfile as io
    exe = os.path.join(ROOT, "fir_amd", "bin", "FIR_OSEM")
    fw = ",".join(f"{f:g}" for f in w.fwhm)
    res = {"iterations": n_iter, "subsets": w.n_subsets,
           "run": "FIR_OSEM osem.params: counts + HU map read from "
                  "Interfile, ACF, sensitivity, Gaussian every "
                  "sub-iteration, image written"}
    with tempfile.TemporaryDirectory() as d:
        w.scanner.write(os.path.join(d, "scanner.scan"))
        w.vol.write(os.path.join(d, "tmpl.h33"))
        w.vol.write(os.path.join(d, "hu.h33"), hu)
        io.write_proj(os.path.join(d, "y.hs"), w.proj, counts)
        io.write_params(os.path.join(d, "osem.params"), "OSEM PARAMETERS", [
            ("input projection file", "y.hs"),
            ("scanner file", "scanner.scan"),
            ("output volume header", "tmpl.h33"),
            ("output volume file name", "x"),
            ("number of iterations", str(n_iter)),
            ("number of subsets", str(w.n_subsets)),
            ("attenuation volume in HU", "hu.h33"),
            ("convolution interval", "1"),
            ("convolution FHWM XYZ in mm", "{" + fw + "}")])
        for name, env in (("plain_s", {"FIR_PATH_CACHE": "off"}),
                          ("path_cache_s", {"FIR_PATH_CACHE": "snorm16"})):
            t0 = time.perf_counter()
            r = subprocess.run([exe, "osem.params"], cwd=d,
                               env=dict(os.environ, **env),
                               capture_output=True, text=True, timeout=300)
            dt = time.perf_counter() - t0
            if r.returncode != 0:
                log("FIR_OSEM failed:", r.stderr[-500:])
                return None
            res[name] = round(dt, 3)
            log(f"FIR_OSEM end to end ({name[:-2]}): {dt:.2f} s")
    return res


def host_io_rates(w, sino):
    """Upload/download rates of the C2 sinogram through the host layer
    (fir_host_selftest upload), page-locked (default) and ordinary memory."""
    import subprocess
    import tempfile
    from fir_amd import interfile as io
    sel = os.path.join(ROOT, "fir_amd", "bin", "fir_host_selftest")
    if not os.path.exists(sel):
        return None
    out = {}
    with tempfile.TemporaryDirectory() as d:
        io.write_proj(os.path.join(d, "y.hs"), w.proj, sino.cpu().numpy())
        for flag, key in (("1", "pinned"), ("0", "pageable")):
            r = subprocess.run([sel, "upload", "y.hs", "5"], cwd=d,
                               capture_output=True, text=True, timeout=300,
                               env=dict(os.environ, FIR_PINNED=flag))
            if r.returncode != 0:
                return {"error": r.stderr[-300:]}
            out[key] = json.loads(r.stdout.strip().splitlines()[-1])
    log(f"host I/O (sinogram {out['pinned']['bytes'] / 1e6:.0f} MB): "
        f"page-locked {out['pinned']['h2d_gb_s']} / "
        f"{out['pinned']['d2h_gb_s']} GB/s up/down, ordinary "
        f"{out['pageable']['h2d_gb_s']} / {out['pageable']['d2h_gb_s']} GB/s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2",
                    help="c2 (default) or uneven (64 batched solves)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cli", action="store_true",
                    help="skip the end-to-end FIR_OSEM runs")
    ap.add_argument("--batch", type=int, default=8,
                    help="reconstructions in the batched measurement (0: skip)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run only the cached projector launch (twice), "
                         "one walked projector launch and a 1 GiB device "
                         "copy (for rocprofv3 --pmc runs)")
    args = ap.parse_args()
    if args.workload == "uneven":
        if args.gpus != 1:
            print("bench.py measures one GPU", file=sys.stderr)
            sys.exit(2)
        run_uneven(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not "
              "supported (multi-GPU sharding is a design note only)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from fir_amd import lib as firlib
    from fir_amd import synth

    w = synth.WORKLOADS[args.workload]()
    n_sub = w.n_subsets
    fwhm = w.fwhm
    log("workload", w.name, "LORs", w.n_lors)
    ctx = firlib.context_for(w)
    v = firlib.VolDesc.from_spec(w.vol)
    nvox = v.n_voxels
    nb = ctx.n_bins

    # --- synthetic data, resident in HBM ---------------------------------
    # The counts are those of the parity fixture: Poisson draws from the
    # forward projections of the phantom and of its mu map, reproducible bit
    # for bit on any host (synth.parity_counts).
    act, hu = synth.phantom(w)
    act_d = torch.from_numpy(act.ravel()).cuda()
    mu = torch.from_numpy(hu.ravel()).cuda()
    y0 = torch.zeros(nb, device="cuda")
    ctx.forward(v, act_d, y0)
    ctx.hu_to_mu(mu)
    acf = torch.zeros(nb, device="cuda")
    ctx.forward(v, mu, acf)
    y0h, mh = y0.cpu().numpy(), acf.cpu().numpy()
    counts_h, scale = synth.parity_counts(w, y0h, mh)
    mean_f = torch.from_numpy(
        (y0h * synth.attenuation(mh) * scale).astype(np.float32)).cuda()
    del y0h, mh
    ctx.exp_clamped(acf)
    y = torch.from_numpy(counts_h).cuda()
    counts = None if args.no_cli or args.traffic_probe else counts_h
    del counts_h
    ctx.mul_clamped(y, acf)  # FIR_OSEM: inputProj *= ACF (OSEM.cc:336-337)
    sens = torch.zeros(n_sub * nvox, device="cuda")
    ctx.sensitivity(v, sens, n_sub)
    x = torch.ones(nvox, device="cuda")
    bp = torch.zeros(nvox, device="cuda")
    torch.cuda.synchronize()
    del act_d, y0
    golden = parity_golden(w)

    stream = torch.cuda.current_stream()

    # --- the drop-in executable, end to end (before this process takes the
    # device memory the caches need) -------------------------------------
    cli = None
    if counts is not None:
        cli = cli_end_to_end(w, counts, hu)
        del counts

    # --- path cache, row order and lanes picked by the autotuner ---------
    from fir_amd import autotune
    cache_t0 = time.perf_counter()
    cache, tres = autotune.tune_cache(ctx, v, n_sub, [x], [y], reps=2)
    torch.cuda.synchronize()
    tune_s = time.perf_counter() - cache_t0
    cache.close()
    cache_t0 = time.perf_counter()
    ctx.set_cache_lanes(tres.best["fwd_lanes"], tres.best["bp_lanes"])
    ctx.set_cache_zslices(tres.best["zslices"])
    cache = ctx.path_cache(v, n_sub, firlib.FIR_LEN_SNORM16,
                           row_order=tres.best["row_order"])
    torch.cuda.synchronize()
    cache_build_s = time.perf_counter() - cache_t0
    cinfo = cache.info()
    log(f"path cache: {cinfo['elements']} elements, "
        f"{cinfo['device_bytes'] / 1e9:.2f} GB, built in {cache_build_s:.2f} s;"
        f" autotuned {tres.best} in {tune_s:.1f} s")
    sens_c = torch.zeros(n_sub * nvox, device="cuda")
    cache.sensitivity(sens_c)

    if args.traffic_probe:
        # one launch of each cached projector kernel pair at steady state,
        # then a plain copy of known size whose FETCH_SIZE/WRITE_SIZE
        # calibrate the counters (1 GiB read + 1 GiB written)
        cache.osem_project([x], [y], None, [bp], 3)
        torch.cuda.synchronize()
        cache.osem_project([x], [y], None, [bp], 3)
        # the walked projector's launch of the same subset (FP64 counters)
        bp.zero_()
        ctx.osem_project(v, x, y, None, bp, 3, n_sub)
        src = torch.ones(1 << 28, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("traffic probe done")
        return

    def osem_iteration(project, sens_v, evs=None):
        for s in range(n_sub):
            if evs is not None:
                evs[s][0].record(stream)
            project(s)
            if evs is not None:
                evs[s][1].record(stream)
            ctx.osem_update(v, x, bp, sens_v[s * nvox:(s + 1) * nvox], 0.0,
                            apply_cut=False, reset=True)
            ctx.convolve(v, x, fwhm, 0.0)

    def cached_project(s):
        cache.osem_project([x], [y], None, [bp], s)

    def plain_project(s):
        ctx.osem_project(v, x, y, None, bp, s, n_sub)

    def timed_iterations(project, sens_v, steps, warmup):
        x.fill_(1.0)
        bp.zero_()
        for _ in range(warmup):
            osem_iteration(project, sens_v)
        torch.cuda.synchronize()
        evs = [[(torch.cuda.Event(enable_timing=True),
                 torch.cuda.Event(enable_timing=True)) for _ in range(n_sub)]
               for _ in range(steps)]
        t_start = torch.cuda.Event(enable_timing=True)
        t_end = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t_start.record(stream)
        for k in range(steps):
            osem_iteration(project, sens_v, evs[k])
        t_end.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = t_start.elapsed_time(t_end) / steps
        proj = float(np.mean([a.elapsed_time(b) for row in evs
                              for (a, b) in row]))
        return ms, proj, wall / steps * 1e3

    def parity(project, sens_v):
        """The golden schedule from x = 1 on this path, against the
        reference's FIR_OSEM image and the FP64 oracle's (untimed)."""
        if golden is None:
            return None
        x.fill_(1.0)
        bp.zero_()
        for _ in range(int(golden["schedule"][0])):
            osem_iteration(project, sens_v)
        torch.cuda.synchronize()
        xh = x.cpu().numpy()
        return {"rel_l2_vs_reference": float(f"{rel_l2(xh, golden['x_ref']):.3e}"),
                "rel_l2_vs_fp64_oracle":
                    float(f"{rel_l2(xh, golden['x_f64']):.3e}"),
                "schedule": (f"{int(golden['schedule'][0])} iterations x "
                             f"{int(golden['schedule'][1])} subsets, ACF, "
                             "4 mm Gaussian every sub-iteration, from x = 1"),
                "reference": "FIR_OSEM of the reference built from source, "
                             "same counts (tests/golden/c2_osem_golden.npz)"}

    cached_parity = parity(cached_project, sens_c)
    if cached_parity:
        log(f"cached path vs reference: rel L2 "
            f"{cached_parity['rel_l2_vs_reference']:.2e}")

    # --- the measured step: cached OSEM iterations ------------------------
    ms_step, proj_avg_ms, wall_ms = timed_iterations(
        cached_project, sens_c, args.steps, args.warmup)
    log(f"cached OSEM iteration {ms_step:.3f} ms (wall {wall_ms:.3f} ms), "
        f"projector launch avg {proj_avg_ms:.3f} ms")

    # --- the plain path (FP64 walk per LOR), same schedule ----------------
    pres = autotune.tune_plain(ctx, v, n_sub, x, y, reps=2)
    log(f"plain projector autotuned: {pres.best}")
    plain_parity = parity(plain_project, sens)
    plain_ms, plain_proj_ms, _ = timed_iterations(
        plain_project, sens, max(1, min(args.steps, 2)), 1)
    log(f"plain OSEM iteration {plain_ms:.3f} ms, projector launch avg "
        f"{plain_proj_ms:.3f} ms")

    # --- separately timed full forward and backprojection passes ---------
    sino = torch.zeros(nb, device="cuda")
    img = torch.zeros(nvox, device="cuda")
    frames = torch.zeros(n_sub * nvox, device="cuda")
    e0, e1 = (torch.cuda.Event(enable_timing=True) for _ in range(2))

    def ev_time(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    fwd_ms = ev_time(lambda: ctx.forward(v, x, sino))
    bp_ms = ev_time(lambda: ctx.backward_subset(v, y, img, 0, 1))
    cfwd_ms = ev_time(lambda: cache.forward(x, sino))
    cbp_ms = ev_time(lambda: cache.backward(y, frames))
    log(f"plain forward {fwd_ms:.3f} ms, backprojection {bp_ms:.3f} ms; "
        f"cached forward {cfwd_ms:.3f} ms, backprojection {cbp_ms:.3f} ms")

    # --- batched: reconstructions sharing each read of the cache ---------
    batched = None
    if args.batch > 0:
        from fir_amd import recon
        g2 = torch.Generator(device="cuda").manual_seed(w.seed + 1)
        scales = np.logspace(-2, 1, args.batch)
        ys = []
        for sc in scales:
            yi = torch.poisson((mean_f * float(sc)), generator=g2)
            ctx.mul_clamped(yi, acf)
            ys.append(yi)
        # batches of 8 images take their line integrals from slab-staged
        # sheets, best with 4 lanes per LOR, and backproject family by
        # family from the sheet form's lists: build that cache in place of
        # the single-image one (same row order)
        bz = 2
        blanes = (4, 16)
        cache.close()
        ctx.set_cache_zslices(bz)
        ctx.set_cache_lanes(*blanes)
        cache = ctx.path_cache(v, n_sub, firlib.FIR_LEN_SNORM16,
                               row_order=tres.best["row_order"])
        grp = recon.OSEMGroup(ctx, v, n_sub, ys, sens=sens_c, fwhm=fwhm,
                              conv_interval=1, cache=cache)
        grp.capture()
        grp.iteration()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(args.steps):
            grp.iteration()
        e1.record(stream)
        torch.cuda.synchronize()
        bms = e0.elapsed_time(e1) / args.steps
        batched = {
            "n_images": args.batch,
            "path": "cached: line integrals from each family's sheet of "
                    "all images staged in LDS slab by slab, family "
                    "backprojection of all images from the sheet form's "
                    "lists; one HIP graph per iteration",
            "cache_forms": cache.forms(),
            "cache_lanes": list(blanes),
            "osem_iter_ms": round(bms, 4),
            "osem_iter_ms_per_recon": round(bms / args.batch, 4),
            "mlors_s_per_recon_aggregate": round(
                args.batch * w.n_lors / (bms * 1e-3) / 1e6, 3),
            "speedup_per_recon_vs_plain": round(plain_ms * args.batch / bms,
                                                3),
            "speedup_per_recon_vs_cached": round(ms_step * args.batch / bms,
                                                 3),
        }
        log(f"batched x{args.batch}: {bms:.3f} ms per iteration "
            f"({bms / args.batch:.3f} ms per reconstruction)")
        del grp, ys
        ctx.set_stream(stream)

    # --- list-mode events into the sinogram (crystal pairs -> bins) ------
    n_ev = 100_000_000
    gen = torch.Generator(device="cuda").manual_seed(w.seed + 2)
    ncr, nr = w.proj.n_crystals_per_ring, w.proj.n_rings
    c1 = torch.randint(0, ncr, (n_ev,), generator=gen, device="cuda",
                       dtype=torch.int32)
    sep = torch.randint(-(w.proj.n_tang_coords // 2),
                        w.proj.n_tang_coords // 2 + 1, (n_ev,),
                        generator=gen, device="cuda", dtype=torch.int32)
    events = torch.stack([
        torch.randint(0, nr, (n_ev,), generator=gen, device="cuda",
                      dtype=torch.int32), c1,
        torch.randint(0, nr, (n_ev,), generator=gen, device="cuda",
                      dtype=torch.int32),
        (c1 + ncr // 2 + sep) % ncr], 1).contiguous()
    del c1, sep
    hist = torch.zeros(ctx.n_bins, device="cuda")
    kept = ctx.histogram_events(events, hist, count=True)
    lm_ms = ev_time(lambda: ctx.histogram_events(events, hist))
    listmode = {
        "events": n_ev, "kept": kept, "ms": round(lm_ms, 4),
        "m_events_s": round(n_ev / (lm_ms * 1e-3) / 1e6, 1),
        "events_gb_s": round(16 * n_ev / (lm_ms * 1e-3) / 1e9, 1),
        "source": "uniform ring pairs, crystal pairs within the stored "
                  "tangential range"}
    log(f"list-mode histogram: {n_ev} events in {lm_ms:.3f} ms "
        f"({listmode['m_events_s']} M events/s)")
    del events, hist

    # --- host I/O: sinogram transfers from page-locked vs ordinary memory
    host_io = host_io_rates(w, sino)

    # --- roofline of the dominant kernel pair (cached OSEM projector) ----
    # Algorithmic bytes per launch (one sub-iteration's LORs): forward +
    # backprojection in the cached-path format, 12 B per LOR and 6 B per
    # path element each way, plus 4 B (x read) + 8 B (bp write, sens read)
    # per voxel.
    n_e_total = cinfo["elements"]
    lors_per_sub = w.n_lors / n_sub
    n_e_per_sub = n_e_total / n_sub
    bytes_launch = 2 * (12 * lors_per_sub + 6 * n_e_per_sub) + 12 * nvox
    achieved = bytes_launch / (proj_avg_ms * 1e-3) / 1e9
    if batched is not None:
        # The batch reads the matrix once for all its images: summed over
        # its reconstructions, the algorithmic bytes of one single-image
        # launch each per sub-iteration. A throughput figure per
        # reconstruction, not a fraction of HBM's bandwidth.
        per_sub_ms = batched["osem_iter_ms"] / n_sub
        eq = args.batch * bytes_launch / (per_sub_ms * 1e-3) / 1e9
        batched["single_image_bytes_rate_gb_s"] = round(eq, 1)
    traffic, traffic_commit = traffic_from_profiles()
    # the walked projector against the FP64 vector peak (its FP64
    # operations per launch counted by rocprofv3, the launch timed here)
    flop, flop_commit = fp64_flop_from_profiles()
    plain_fp64 = None
    if flop and plain_proj_ms > 0:
        tflops = flop / (plain_proj_ms * 1e-3) / 1e12
        plain_fp64 = {
            "kernel": "k_project_fam<true> (walked OSEM projector launch, "
                      "FP64 Siddon walk per LOR)",
            "bound": "latency of the FP64 walk",
            "achieved": round(tflops, 3), "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tflops / FP64_PEAK_TFLOPS, 4),
            "flop_per_launch": flop,
            "flop_source": ("rocprofv3 --pmc SQ_INSTS_VALU_*_F64, "
                            "profiles/plain_osem_project_c2_fp64_counters.md"
                            f" at commit {flop_commit}")}
    copy_gbs = copy_bandwidth()
    log(f"device copy: {copy_gbs:.0f} GB/s (1 GiB read + 1 GiB written)")

    cpu = None
    if not args.no_cpu:
        log("timing the reference CPU path (one OSEM iteration)...")
        cpu = cpu_baseline(w, n_sub)

    value = w.n_lors / (ms_step * 1e-3) / 1e6
    out = {
        "metric": "M LORs/s fwd+backproj & OSEM iter ms: plain vs batched "
                  "GPU at rel err<=1e-4",
        "value": round(value, 3),
        "unit": "M LORs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, Poisson counts)",
        "config": {
            "workload": w.name,
            "scanner": "24 rings x 448 crystals, r=330 mm, span 1, 47 seg",
            "n_lors": w.n_lors,
            "image": "160x160x47 @ 2x2x3.27 mm",
            "osem": f"{n_sub} subsets, 4 mm FWHM Gaussian every subiter",
            "path": "path cache (Siddon matrix in HBM, snorm16 lengths)",
        },
        "osem_iter_ms": round(ms_step, 4),
        "osem_subiter_ms": round(ms_step / n_sub, 4),
        "cache": {"build_s": round(cache_build_s, 3),
                  "autotune": {"best": tres.best,
                               "trials_ms": tres.table(),
                               "seconds": round(tune_s, 2)},
                  "elements": int(n_e_total),
                  "rows": cinfo["rows"],
                  "device_gb": round(cinfo["device_bytes"] / 1e9, 3),
                  "fwd_ms": round(cfwd_ms, 4),
                  "fwd_mlors_s": round(w.n_lors / (cfwd_ms * 1e-3) / 1e6, 3),
                  "bp_all_subsets_ms": round(cbp_ms, 4),
                  "bp_mlors_s": round(w.n_lors / (cbp_ms * 1e-3) / 1e6, 3)},
        "plain": {"osem_iter_ms": round(plain_ms, 4),
                  "autotune": {"best": pres.best, "trials_ms": pres.table()},
                  "projector_launch_ms": round(plain_proj_ms, 4),
                  "fp64_roofline": plain_fp64,
                  "mlors_s": round(w.n_lors / (plain_ms * 1e-3) / 1e6, 3),
                  "fwd_ms": round(fwd_ms, 4),
                  "bp_ms": round(bp_ms, 4),
                  "osem_parity": plain_parity,
                  "fwd_mlors_s": round(w.n_lors / (fwd_ms * 1e-3) / 1e6, 3),
                  "bp_mlors_s": round(w.n_lors / (bp_ms * 1e-3) / 1e6, 3)},
        "speedup_cached_vs_plain": round(plain_ms / ms_step, 3),
        "osem_parity": cached_parity,
        "path_elements_per_iter": int(n_e_total),
        "roofline": {
            "kernel": "k_pc_osem_fam + k_pc_fam_reduce (one cached OSEM "
                      "projector launch: line integrals from LDS-staged "
                      "sheets, each family's voxel lists summed against its "
                      "weights in LDS, then the families' partials summed "
                      "per voxel)",
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, "
                               "profiles/osem_cached_c2_traffic.md at "
                               f"commit {traffic_commit}"),
            "achievable_copy_gb_s": round(copy_gbs, 1),
            "frac_of_copy": round(achieved / copy_gbs, 4),
            "avg_launch_ms": round(proj_avg_ms, 4),
            "algorithmic_bytes_per_launch": int(bytes_launch),
        },
        "batched": batched,
        "listmode": listmode,
        "host_io": host_io,
        "fir_osem_end_to_end": cli,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu_iter"] = round(cpu["osem_iter_s"] * 1e3 / ms_step, 2)
        out["speedup_plain_vs_cpu_iter"] = round(
            cpu["osem_iter_s"] * 1e3 / plain_ms, 2)
    print(json.dumps(out), flush=True)
    cache.close()
    ctx.close()


if __name__ == "__main__":
    main()
