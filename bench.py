"""Benchmark of the exact squared EDT + local thickness (Tb.Th and Tb.Sp) path.

One JSON line.  The headline (metric/value) is the plain one-volume form: one
step = one bm_analyze_thickness_spacing_device() call on the 512^3 workload
(plain_512: sigma 4, BV/TV 0.15, z-cylinder ROI of diameter 460, 10.5 um,
Standard mode): both phases through EDT, ridges, sphere fill, clean-up, mask,
statistics and quantiles, with bone and ROI already resident in HBM.
value = voxels / step time (Gvoxel/s, the volume counted once).

Beside it:
  batched  96 volumes of uneven shapes (axes U{64..384}) analyzed as one plan
           replayed from a hipGraph, against 96 sequential one-volume calls on
           the same resident inputs (Gvoxel/s of each, speed-up, equality).
  serving  a seeded Poisson stream (150 requests/s for 60 s, lognormal sizes
           of median 192^3) through the request loop with a 200 GB HBM budget:
           p50/p99 latency (completion - submission) and Gvoxel/s.
  analyze  the whole analyze (every Metrics field, topology and surface included)
           on the plain workload: time per call, the topology passes' share,
           exact agreement of the integer measures with the oracle.
  tiff     the input step: the same bone volume as a 512-page TIFF (LZW and
           uncompressed) read into HBM on all host threads and on one.
  sdf      the signed distance field of the same bone volume smoothed with the
           30 um Gaussian (signed_distance_field_um + gaussian_smooth_nearest_in_place),
           per-pass HBM fractions, the oracle's time and bit-exact parity.

The roofline object describes the kernel with the largest share of the plain
step, from per-kernel HIP events recorded on the launching stream; the
algorithmic bytes per voxel of each kernel are listed in KERNEL_BYTES.  The CPU
baseline runs the oracle (a C restatement of the reference algorithm, OpenMP)
on a bounded sample of the plain workload on this host.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per voxel of one launch over one item (one phase of a
# volume); every launch of the plain step serves the bone and marrow items.
# The EDT figures are the 22 B/voxel/phase of the algorithm (bone + ROI in,
# an int32 distance out per pass); this build moves 18 (its X pass writes u16).
KERNEL_BYTES = {
    "edt_x": 6.0,           # bone + ROI u8 in, int32 row distance out
    "edt_y": 8.0,           # int32 in, int32 out
    "edt_z": 8.0,           # int32 in, int32 out
    "ridges": 4.0,          # squared EDT in (ridge lists are per ridge, excluded)
    "fill_gather": 4.0,     # propagated map out (ridge lists are per ridge, excluded)
    "cleanup_stats": 12.0,  # propagated map and squared EDT in, thickness map out
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_plain_512.txt")
COUNTERS_FILE = os.path.join(ROOT, "profiles", "counters_tile_kernels_512.txt")
# The kernel that takes most of the plain step (timed inside the timed steps).
DOMINANT_KERNEL = "fill_gather"
# 32-bit VALU lane operations per second: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
# (one wave64 instruction per SIMD every 4 cycles).
VALU_LANE_OPS_PEAK = 256 * 4 * 16 * 2.4e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="plain_512", choices=["plain_512", "plain_small", "large"])
    ap.add_argument("--mode", default="standard", choices=["standard", "classic"])
    ap.add_argument("--forms", default="plain,batched,serving,sdf,analyze,tiff",
                    help="comma list of plain, batched, serving, sdf, analyze, tiff")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-host-call", action="store_true",
                    help="skip timing the host-pointer entry point")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-fill-work", action="store_true",
                    help="skip counting the fill work (two extra stage calls; profiling runs)")
    ap.add_argument("--cpu-runs", type=int, default=3,
                    help="runs of the CPU baseline (the median is reported)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--serve-rate", type=float, default=150.0)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--serve-max-batch", type=int, default=32)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-pool", type=int, default=96,
                    help="generated request volumes of the serving stream")
    ap.add_argument("--serve-saturated-rate", type=float, default=1200.0)
    ap.add_argument("--serve-saturated-seconds", type=float, default=8.0)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a device-to-device copy of this size (torch), the "
                         "reference for scaling FETCH_SIZE/WRITE_SIZE counters")
    return ap.parse_args()


def calibrate_copy(mib):
    """A plain copy kernel of known size: src (mib MiB) read once, dst written
    once, three times."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    ok = bool(torch.equal(dst, src))
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3,
                      "equal": ok}))
    return 0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def same_result(a, b):
    import math
    if isinstance(a, Exception) or isinstance(b, Exception):
        return type(a) is type(b) and str(a) == str(b)
    for x, y in ((a.thickness, b.thickness), (a.spacing, b.spacing)):
        for f in ("mean", "standard_deviation", "maximum", "median", "percentile_10",
                  "percentile_25", "percentile_75", "percentile_90", "count"):
            u, v = getattr(x, f), getattr(y, f)
            if not (u == v or (isinstance(u, float) and math.isnan(u) and math.isnan(v))):
                return False
    return a.bone_volume_fraction == b.bone_volume_fraction


def read_traffic():
    """HBM bytes per launch of the dominant kernel from the committed counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    try:
        vals = {}
        for line in open(TRAFFIC_FILE):
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
        return {k: float(vals[k]) for k in ("fill_gather_bytes_per_launch",
                                            "edt_bytes_per_step") if k in vals}
    except (OSError, ValueError):
        return {}


def read_busy(kernel):
    """VALU and SALU busy fractions of a kernel from the committed counter
    summary (rocprofv3 --pmc, SQ_INSTS_VALU/SALU against GRBM_GUI_ACTIVE)."""
    try:
        for line in open(COUNTERS_FILE):
            f = line.split()
            if f and f[0] == kernel and f[-1] == "%":
                return {"valu_busy": float(f[-4]) / 100.0, "salu_busy": float(f[-2]) / 100.0,
                        "source": os.path.relpath(COUNTERS_FILE, ROOT)}
    except (OSError, ValueError, IndexError):
        pass
    return None


def host_cpu():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return model, affinity


def fill_work(bm, wl):
    """W_fill of both analysis phases: the sum over ridges of the ball voxels
    within the array (the reference propagation's visits,
    src/local_thickness.rs:182-201), counted on the device outside the timed
    region through bm_local_thickness_stages."""
    import numpy as np
    total = 0
    per = {}
    for name, want in (("bone", True), ("marrow", False)):
        sel = ((wl.roi != 0) & ((wl.bone != 0) == want)).astype(np.uint8)
        st = bm.local_thickness_stages(bm.BinaryVolume(sel, wl.width, wl.height, wl.depth),
                                       bm.Phase.Foreground)
        per[name] = {"fill_visits": st["fill_visits"], "ridges": st["ridge_count"]}
        total += st["fill_visits"]
    return total, per


def plain_form(args, bm, native, hip, workloads, cpu_threads):
    t0 = time.time()
    wl = getattr(workloads, args.workload)(threads=cpu_threads)
    log(f"generated {wl.name} in {time.time() - t0:.1f}s")
    mode = bm.FeatureMode.Standard if args.mode == "standard" else bm.FeatureMode.Classic
    d_bone = hip.DeviceBuffer(wl.n)
    d_roi = hip.DeviceBuffer(wl.n)
    d_bone.upload(wl.bone)
    d_roi.upload(wl.roi)
    stream = hip.stream()

    def step():
        return bm.api.analyze_thickness_spacing_device(
            d_bone.value, d_roi.value, wl.width, wl.height, wl.depth, wl.spacing, mode,
            stream.value)

    for _ in range(args.warmup):
        res = step()
    hip.synchronize()
    native.profile_read()  # drop warm-up records
    # The timed steps bracket only the dominant kernel's launches with events
    # (two per step: events around every launch cost about 0.1 ms per step);
    # every kernel's time comes from a profiled repeat of the same steps.
    native.profile_only([DOMINANT_KERNEL])
    native.profile_enable(True)
    e0, e1 = hip.Event(), hip.Event()
    hip.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        res = step()
    e1.record(stream)
    hip.synchronize()
    total_ms = e0.elapsed_ms(e1)
    native.profile_enable(False)
    timed_kernels = native.profile_read()
    native.profile_only(None)
    native.profile_enable(True)
    for _ in range(args.steps):
        step()
    hip.synchronize()
    native.profile_enable(False)
    kernels = native.profile_read()
    log(f"plain: {args.steps} steps in {total_ms:.2f} ms")
    ms_per_step = total_ms / args.steps
    gvox = wl.n / (ms_per_step * 1e-3) / 1e9

    # Roofline of the kernel with the largest time share (one launch serves
    # both phases: 2 items of wl.n voxels), timed inside the timed steps.
    name, (launches, kms) = max(kernels.items(), key=lambda kv: kv[1][1])
    if name in timed_kernels:
        launches, kms = timed_kernels[name]
    avg_ms = kms / launches
    per_vox = KERNEL_BYTES.get(name, 0.0)
    alg_bytes = per_vox * wl.n * 2
    hbm_gbs = alg_bytes / (avg_ms * 1e-3) / 1e9 if per_vox else None
    traffic = read_traffic()
    if name == "fill_gather" and not args.no_fill_work:
        # The sphere fill is bound by vector/scalar instruction issue: its
        # algorithmic work is W_fill, the reference propagation's voxel visits,
        # against the vector lanes' rate at one 32-bit operation per visit.
        w_fill, w_per = fill_work(bm, wl)
        visits_per_s = w_fill / (avg_ms * 1e-3)
        roofline = {"kernel": name, "bound": "valu", "achieved": visits_per_s / 1e9,
                    "peak": VALU_LANE_OPS_PEAK / 1e9, "unit": "Gvisit/s",
                    "frac": visits_per_s / VALU_LANE_OPS_PEAK,
                    "traffic": traffic.get("fill_gather_bytes_per_launch"),
                    "avg_launch_ms": avg_ms, "fill_visits_per_launch": w_fill,
                    "fill_work_per_phase": w_per,
                    "hbm": {"algorithmic_bytes_per_launch": alg_bytes, "achieved_gbs": hbm_gbs,
                            "frac": hbm_gbs / HBM_PEAK_GBS},
                    "issue": read_busy("k_fill_gather"),
                    "note": "W_fill visits per second against 256 CUs x 4 SIMDs x 16 lanes "
                            "x 2.4 GHz 32-bit VALU lane operations (one per visit); the "
                            "issue object gives the kernel's VALU/SALU busy fractions"}
    else:
        roofline = {"kernel": name, "bound": "hbm", "achieved": hbm_gbs, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": (hbm_gbs / HBM_PEAK_GBS) if hbm_gbs else None,
                    "traffic": None, "avg_launch_ms": avg_ms,
                    "algorithmic_bytes_per_launch": alg_bytes}
    edt = {k: kernels[k][1] / kernels[k][0] for k in ("edt_x", "edt_y", "edt_z") if k in kernels}
    edt_ms = sum(edt.values())
    edt_gbs = 22.0 * 2 * wl.n / (edt_ms * 1e-3) / 1e9 if edt_ms else None

    # The host-pointer entry point (bm_analyze_thickness_spacing): inputs copied
    # over PCIe each call, for the PCIe-inclusive rate.
    host_ms = None
    if not args.no_host_call:
        bvol = bm.BinaryVolume(wl.bone, wl.width, wl.height, wl.depth)
        rvol = bm.BinaryVolume(wl.roi, wl.width, wl.height, wl.depth)
        bm.analyze_thickness_spacing(bvol, rvol, wl.spacing, mode)
        t1 = time.perf_counter()
        for _ in range(3):
            bm.analyze_thickness_spacing(bvol, rvol, wl.spacing, mode)
        host_ms = (time.perf_counter() - t1) / 3 * 1e3

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as orc
        from _stats import relative_differences
        quant = mode == bm.FeatureMode.Standard
        model, affinity = host_cpu()
        # The sample is the whole volume: about 20 s of CPU work at 16 threads,
        # run `cpu_runs` times (median).
        orc.set_threads(cpu_threads)
        times = []
        for _ in range(max(1, args.cpu_runs)):
            t1 = time.time()
            ref = orc.analyze_thickness_spacing(wl.bone, wl.roi, wl.width, wl.height, wl.depth,
                                                wl.spacing, include_quantiles=quant)
            times.append(time.time() - t1)
        cpu_s = sorted(times)[len(times) // 2]
        # The same call on one thread and on all threads, on the 128^3 workload
        # (a sample that one thread finishes in seconds).
        small = workloads.plain_small(threads=cpu_threads)
        one = {}
        for label, th in (("one_thread", 1), ("all_threads", cpu_threads)):
            orc.set_threads(th)
            t1 = time.time()
            orc.analyze_thickness_spacing(small.bone, small.roi, small.width, small.height,
                                          small.depth, small.spacing, include_quantiles=quant)
            one[label] = small.n / (time.time() - t1) / 1e9
        orc.set_threads(cpu_threads)
        cpu = {"value": wl.n / cpu_s / 1e9, "unit": "Gvoxel/s", "cores": cpu_threads,
               "kind": "port",
               "sample": f"the full {wl.name} volume ({wl.width}x{wl.height}x{wl.depth}), both "
                         f"phases with quantiles, oracle (C/OpenMP restatement), median of "
                         f"{len(times)} runs: {', '.join(f'{t:.1f}' for t in times)} s",
               "host_cpu": model, "host_cpus_in_affinity": affinity,
               "threads_note": "the OpenMP pool is the box's CPU share (16 threads for one "
                               "GPU), not every CPU in the affinity mask",
               "plain_small_128": {"gvox_per_s_one_thread": one["one_thread"],
                                   "gvox_per_s_all_threads": one["all_threads"],
                                   "threads": cpu_threads}}
        ok = True
        diffs = {}
        for name_, g, r in (("thickness", res.thickness, ref.thickness),
                            ("spacing", res.spacing, ref.spacing)):
            ok &= g.count == r.count and g.maximum == r.maximum
            if quant:
                ok &= all(getattr(g, f) == getattr(r, f) for f in (
                    "median", "percentile_10", "percentile_25", "percentile_75", "percentile_90"))
            diffs[name_] = relative_differences(g, r)
        parity = {"counts_maxima_quantiles_equal": bool(ok), "relative_differences": diffs}
    return dict(wl=wl, res=res, ms_per_step=ms_per_step, gvox=gvox, roofline=roofline,
                edt=edt, edt_gbs=edt_gbs, kernels=kernels, cpu=cpu, parity=parity,
                host_ms=host_ms)


# Algorithmic HBM bytes per voxel of the smoothing passes (f64): the x pass
# reads the bone byte and the squared EDT of the voxel's phase and writes the
# x-smoothed field; the y and z passes read and write one f64 each.
SMOOTH_BYTES = {"smooth_x": 1.0 + 4.0 + 8.0, "smooth_y": 16.0, "smooth_z": 16.0}


def sdf_form(args, bm, native, hip, wl, cpu_threads):
    """signed_distance_field_um + gaussian_smooth_nearest_in_place on the bone
    of the plain workload, 30 um bandwidth (the callers' default) at its
    spacing: one bm_smoothed_signed_distance_field_um_device call per step."""
    import numpy as np
    hs = wl.spacing[0]
    sigma_vox = 30.0 / hs
    d_bone = hip.DeviceBuffer(wl.n)
    d_bone.upload(wl.bone)
    d_out = hip.DeviceBuffer(wl.n * 8)
    stream = hip.stream()

    def step():
        bm.api.smoothed_signed_distance_field_um_device(
            d_bone.value, wl.width, wl.height, wl.depth, hs, sigma_vox, d_out.value,
            stream.value)

    for _ in range(max(1, args.warmup)):
        step()
    hip.synchronize()
    native.profile_read()
    native.profile_enable(True)
    e0, e1 = hip.Event(), hip.Event()
    steps = max(1, args.steps // 2)
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    hip.synchronize()
    total_ms = e0.elapsed_ms(e1)
    native.profile_enable(False)
    kernels = native.profile_read()
    ms = total_ms / steps
    passes = {}
    for k, b in SMOOTH_BYTES.items():
        if k in kernels:
            avg = kernels[k][1] / kernels[k][0]
            gbs = b * wl.n / (avg * 1e-3) / 1e9
            passes[k] = {"avg_launch_ms": avg, "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS,
                         "algorithmic_bytes_per_voxel": b}
    got = np.empty(wl.n, np.float64)
    d_out.download(got)
    log(f"sdf: {steps} calls in {total_ms:.2f} ms")
    cpu = parity = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as orc
        orc.set_threads(cpu_threads)
        t1 = time.time()
        ref = orc.gaussian_smooth_nearest(
            orc.signed_distance_field_um(wl.bone, wl.width, wl.height, wl.depth, hs),
            wl.width, wl.height, wl.depth, sigma_vox)
        cpu_s = time.time() - t1
        cpu = {"value": wl.n / cpu_s / 1e9, "unit": "Gvoxel/s", "cores": cpu_threads,
               "kind": "port",
               "sample": f"the full {wl.name} bone volume, oracle (C/OpenMP restatement), "
                         f"{cpu_s:.1f} s"}
        parity = bool(np.array_equal(got.view(np.uint64), ref.view(np.uint64)))
    return {"workload": f"{wl.name} bone, sigma 30 um at {hs} um (sigma_vox {sigma_vox:.4f}, "
                        f"radius {len(bm.gaussian_kernel(sigma_vox)) // 2})",
            "ms_per_call": ms, "gvox_per_s": wl.n / (ms * 1e-3) / 1e9,
            "kernels_ms_per_call": {k: v[1] / steps for k, v in kernels.items()},
            "smoothing_passes": passes, "cpu_baseline": cpu, "bit_exact_vs_oracle": parity}


def analyze_form(args, bm, native, hip, wl, cpu_threads, plain_cpu):
    """The whole analyze (every Metrics field: thickness/spacing, cubical
    measures, Betti numbers, BoneJ connectivity, surface, ROI-boundary
    fraction): one bm_analyze_device call per step on the plain workload."""
    d_bone, d_roi = hip.DeviceBuffer(wl.n), hip.DeviceBuffer(wl.n)
    d_bone.upload(wl.bone)
    d_roi.upload(wl.roi)
    stream = hip.stream()

    def step():
        return bm.api.analyze_device(d_bone.value, d_roi.value, wl.width, wl.height, wl.depth,
                                     wl.spacing, bm.FeatureMode.Standard, stream.value)

    for _ in range(max(1, args.warmup)):
        m = step()
    hip.synchronize()
    native.profile_read()
    native.profile_enable(True)
    steps = max(1, args.steps // 2)
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(steps):
        m = step()
    e1.record(stream)
    hip.synchronize()
    total_ms = e0.elapsed_ms(e1)
    native.profile_enable(False)
    kernels = native.profile_read()
    ms = total_ms / steps
    topo = {k: v[1] / steps for k, v in kernels.items() if k.startswith(("topo_", "ccl_"))}
    log(f"analyze: {steps} calls in {total_ms:.2f} ms")
    cpu = parity = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as orc
        orc.set_threads(cpu_threads)
        t1 = time.time()
        rest = orc.analyze_rest(wl.bone, wl.roi, wl.width, wl.height, wl.depth, wl.spacing)
        rest_s = time.time() - t1
        t = m.topology
        parity = bool((t.raw_euler, t.beta0, t.beta1, t.beta2) == (
            rest.topology.raw_euler, rest.topology.beta0, rest.topology.beta1,
            rest.topology.beta2)
            and m.bonej_connectivity.connectivity == rest.bonej.connectivity
            and m.geometry.bone_surface == rest.bone_surface
            and m.geometry.mean_breadth == rest.mean_breadth
            and m.roi_boundary_bone_fraction == rest.roi_boundary_bone_fraction)
        thick_s = wl.n / (plain_cpu["value"] * 1e9) if plain_cpu else None
        cpu = {"rest_s": rest_s, "thickness_spacing_s": thick_s, "cores": cpu_threads,
               "kind": "port",
               "value": (wl.n / ((thick_s or 0.0) + rest_s) / 1e9), "unit": "Gvoxel/s",
               "sample": f"the full {wl.name} volume: the oracle's thickness/spacing time from "
                         f"the plain form plus its topology/surface restatement "
                         f"({rest_s:.1f} s; flood fills serial as in the reference)"}
    return {"ms_per_call": ms, "gvox_per_s": wl.n / (ms * 1e-3) / 1e9,
            "topology_kernels_ms_per_call": topo, "topology_ms_per_call": sum(topo.values()),
            "beta": [m.topology.beta0, m.topology.beta1, m.topology.beta2],
            "cpu_baseline": cpu, "topology_exact_vs_oracle": parity}


def tiff_form(args, bm, hip, wl, cpu_threads):
    """The input step: the plain workload's bone volume as a 512-page 8-bit
    TIFF (LZW and uncompressed, in the page cache) read into HBM by
    bm_read_binary_tiff_device, on all host threads and on one (the
    reference reads pages one after another)."""
    import tempfile
    import numpy as np
    from PIL import Image
    out = {}
    d_out = hip.DeviceBuffer(wl.n)
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        pages = [Image.fromarray(p) for p in wl.bone.reshape(wl.depth, wl.height, wl.width)]
        for comp in ("tiff_lzw", "raw"):
            path = os.path.join(tmp, comp + ".tif")
            pages[0].save(path, save_all=True, append_images=pages[1:], compression=comp)
            size = os.path.getsize(path)
            row = {"file_mb": size / 1e6}
            for label, threads in (("all_threads", cpu_threads), ("one_thread", 1)):
                bm.read_binary_tiff_device(path, d_out.value, wl.n, strict_binary=True,
                                           threads=threads)
                reps = 3 if threads > 1 else 1
                t0 = time.perf_counter()
                for _ in range(reps):
                    bm.read_binary_tiff_device(path, d_out.value, wl.n, strict_binary=True,
                                               threads=threads)
                s = (time.perf_counter() - t0) / reps
                row[label] = {"ms": s * 1e3, "gvox_per_s": wl.n / s / 1e9, "threads": threads}
            got = np.empty(wl.n, np.uint8)
            d_out.download(got)
            row["exact"] = bool(np.array_equal(got, (wl.bone > 0).astype(np.uint8)))
            out[comp] = row
            log(f"tiff {comp}: {row['all_threads']['ms']:.1f} ms")
    return out


def batched_form(args, bm, hip, workloads, cpu_threads):
    t0 = time.time()
    vols = workloads.batched(threads=cpu_threads)
    total = sum(v.n for v in vols)
    log(f"generated 96 volumes ({total / 1e9:.3f} Gvoxel) in {time.time() - t0:.1f}s")
    bufs = []
    for v in vols:
        b, r = hip.DeviceBuffer(v.n), hip.DeviceBuffer(v.n)
        b.upload(v.bone)
        r.upload(v.roi)
        bufs.append((b, r))
    mode = bm.FeatureMode.Standard
    items = [(b.value, r.value, v.width, v.height, v.depth, v.spacing, mode)
             for v, (b, r) in zip(vols, bufs)]
    stream = hip.stream()

    def plain_all():
        out = []
        for it in items:
            out.append(bm.api.analyze_thickness_spacing_device(*it, stream=stream.value))
        return out

    ref = plain_all()
    hip.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.batch_steps):
        ref = plain_all()
    hip.synchronize()
    plain_s = (time.perf_counter() - t1) / args.batch_steps
    t1 = time.perf_counter()
    batch = bm.Batch(items, use_graph=True)
    create_s = time.perf_counter() - t1
    got = batch.run()
    t1 = time.perf_counter()
    for _ in range(args.batch_steps):
        got = batch.run()
    batch_s = (time.perf_counter() - t1) / args.batch_steps
    equal = all(same_result(a, b) for a, b in zip(got, ref))
    out = {"workload": "96 volumes, axes U{64..384}, sigma U[2,6], BV/TV U[0.08,0.35], "
                       "cylinder ROI 70-95% of min(w,h), spacing U[6,20] um, seed 4",
           "voxels": total, "plain_ms": plain_s * 1e3, "batched_ms": batch_s * 1e3,
           "plain_gvox_per_s": total / plain_s / 1e9, "batched_gvox_per_s": total / batch_s / 1e9,
           "speedup": plain_s / batch_s, "graph": batch.device_bytes() > 0,
           "create_ms": create_s * 1e3, "device_gb": batch.device_bytes() / 1e9,
           "tuning_cache": os.environ.get("BM_TUNE_CACHE"),
           "results_identical_to_plain": bool(equal)}
    batch.close()
    log(f"batched: plain {plain_s * 1e3:.1f} ms, batched {batch_s * 1e3:.1f} ms")
    return out


def serving_pool(args, workloads, cpu_threads):
    import numpy as np
    t0 = time.time()
    pool = workloads.serving_pool(args.serve_pool, 5, threads=cpu_threads)
    log(f"serving: pool of {len(pool)} volumes (median "
        f"{np.median([r.n for r in pool]) ** (1 / 3):.0f}^3) in {time.time() - t0:.1f}s")
    return pool


def serving_form(args, bm, workloads, pool, rate, seconds):
    """Requests of the pool arriving as a Poisson stream at `rate` per second
    for `seconds`: arrival i carries pool[i % len(pool)]."""
    import numpy as np
    arrivals = workloads.serving_arrivals(rate=rate, duration=seconds, seed=5)
    srv = bm.Server(budget_bytes=int(args.serve_budget_gb * 1e9), max_batch=args.serve_max_batch,
                    slots=args.serve_slots)
    # warm-up: one request of each of a few sizes
    for r in pool[:4]:
        srv.poll(srv.submit(r.bone, r.roi, r.width, r.height, r.depth, r.spacing), wait=True)
    warm = srv.stats()  # subtracted below: the stream's own batches and busy time
    tickets = []
    done = []
    lock = threading.Condition()
    finished = [False]

    def poller():
        i = 0
        while True:
            with lock:
                while i >= len(tickets) and not finished[0]:
                    lock.wait(0.05)
                if i >= len(tickets) and finished[0]:
                    return
                t, req = tickets[i]
            res, tim = srv.poll(t, wait=True)
            done.append((req, res, tim))
            i += 1

    th = threading.Thread(target=poller)
    th.start()
    start = time.perf_counter()
    max_lag = 0.0
    last_progress = start
    for k, a in enumerate(arrivals):
        now = time.perf_counter() - start
        if a > now:
            time.sleep(a - now)
        # the request's time from its arrival to the server's submit stamp
        lag = time.perf_counter() - start - a
        max_lag = max(max_lag, lag)
        r = pool[k % len(pool)]
        t = srv.submit(r.bone, r.roi, r.width, r.height, r.depth, r.spacing)
        with lock:
            tickets.append((t, (r, lag)))
            lock.notify()
        if time.perf_counter() - last_progress > 10:
            last_progress = time.perf_counter()
            log(f"serving: {k + 1} submitted, {len(done)} done")
    with lock:
        finished[0] = True
        lock.notify()
    th.join()
    stats = srv.stats()
    srv.close()
    for k in ("batches", "sum_batch", "busy_s"):
        stats[k] -= warm[k]
    # latency = completion - arrival: the server's submit-to-done time plus
    # the time from the scheduled arrival to the submit call
    lat = np.array([tim["done_s"] - tim["submit_s"] + rq[1] for rq, _, tim in done]) * 1e3
    errors = sum(1 for _, res, _ in done if isinstance(res, Exception))
    vox = sum(rq[0].n for rq, res, _ in done if not isinstance(res, Exception))
    t_first = min(tim["submit_s"] for _, _, tim in done)
    t_last = max(tim["done_s"] for _, _, tim in done)
    out = {"workload": f"Poisson {rate:g}/s for {seconds:g} s, seed 5; pool of {len(pool)} "
                       "generated volumes (each its own sigma U[2,6], BV/TV U[0.08,0.35], "
                       "cylinder ROI 70-95% of min(w,h), spacing U[6,20] um), lognormal voxels "
                       "(median 192^3, sigma_ln 0.7) clipped to 64^3..512^3, axes jittered "
                       "+-20%",
           "requests": len(done), "errors": errors, "latency": "completion - arrival",
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
           "gvox_per_s": vox / (t_last - t_first) / 1e9,
           "offered_gvox_per_s": vox / seconds / 1e9,
           "budget_gb": args.serve_budget_gb, "slots": args.serve_slots,
           "max_batch": args.serve_max_batch, "batches": stats["batches"],
           "mean_batch": stats["sum_batch"] / max(stats["batches"], 1),
           "largest_batch": stats["max_batch"], "max_inflight_batches": stats["max_inflight"],
           "peak_device_gb": stats["peak_device_bytes"] / 1e9,
           "gpu_busy_frac": stats["busy_s"] / (t_last - t_first),
           "max_submit_lag_ms": max_lag * 1e3}
    log(f"serving: p50 {out['p50_ms']:.1f} ms, p99 {out['p99_ms']:.1f} ms, "
        f"{out['gvox_per_s']:.3f} Gvoxel/s")
    return out


def main():
    args = parse()
    if args.calibrate_copy_mib:
        return calibrate_copy(args.calibrate_copy_mib)
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (volumes shard "
              "across GPUs by whole volumes; see DESIGN.md). Run with --gpus 1.",
              file=sys.stderr)
        return 2
    import bone_morphometry_amd as bm
    from bone_morphometry_amd import _native, hip, workloads

    if _native.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    cpu_threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    if "plain" not in forms and "analyze" in forms:  # the whole analyze alone (profiling runs)
        wl = getattr(workloads, args.workload)(threads=cpu_threads)
        full = analyze_form(args, bm, _native, hip, wl, cpu_threads, None)
        print(json.dumps({
            "metric": "Gvoxel/s whole analyze", "value": full["gvox_per_s"], "unit": "Gvoxel/s",
            "n_gpus": 1, "steps": max(1, args.steps // 2), "warmup": max(1, args.warmup),
            "ms_per_step": full["ms_per_call"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u32", "data": "synthetic",
            "config": {"workload": wl.name, "shape": [wl.width, wl.height, wl.depth]},
            "full_analyze": full}))
        return 0
    if "plain" not in forms:  # the smoothed field alone (profiling runs)
        wl = getattr(workloads, args.workload)(threads=cpu_threads)
        sdf = sdf_form(args, bm, _native, hip, wl, cpu_threads)
        print(json.dumps({
            "metric": "Gvoxel/s smoothed signed distance field", "value": sdf["gvox_per_s"],
            "unit": "Gvoxel/s", "n_gpus": 1, "steps": max(1, args.steps // 2),
            "warmup": max(1, args.warmup), "ms_per_step": sdf["ms_per_call"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic", "config": {"workload": wl.name,
                                            "shape": [wl.width, wl.height, wl.depth]},
            "smoothed_sdf": sdf}))
        return 0
    p = plain_form(args, bm, _native, hip, workloads, cpu_threads)
    batched = batched_form(args, bm, hip, workloads, cpu_threads) if "batched" in forms else None
    serving = saturated = None
    if "serving" in forms:
        pool = serving_pool(args, workloads, cpu_threads)
        serving = serving_form(args, bm, workloads, pool, args.serve_rate, args.serve_seconds)
        # the same pool offered faster than one slot serves it: continuous
        # batching under load
        saturated = serving_form(args, bm, workloads, pool, args.serve_saturated_rate,
                                 args.serve_saturated_seconds)
    wl = p["wl"]
    sdf = sdf_form(args, bm, _native, hip, wl, cpu_threads) if "sdf" in forms else None
    full = (analyze_form(args, bm, _native, hip, wl, cpu_threads, p["cpu"])
            if "analyze" in forms else None)
    tiff = tiff_form(args, bm, hip, wl, cpu_threads) if "tiff" in forms else None
    line = {
        "metric": "Gvoxel/s EDT+local thickness at 512^3",
        "value": p["gvox"],
        "unit": "Gvoxel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": p["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": wl.name, "shape": [wl.width, wl.height, wl.depth],
                   "roi": {"plain_512": "z-cylinder d=460",
                           "large_1024x1024x1536": "ellipsoid 500x480x740"}.get(wl.name, "full"),
                   "spacing_um": wl.spacing[0], "mode": args.mode, "phases": 2},
        "roofline": p["roofline"],
        "cpu_baseline": p["cpu"],
        "edt": {"ms_per_launch": p["edt"], "gbs_22B_per_voxel_phase": p["edt_gbs"],
                "frac": (p["edt_gbs"] / HBM_PEAK_GBS) if p["edt_gbs"] else None},
        "kernels_ms_per_step": {k: v[1] / args.steps for k, v in p["kernels"].items()},
        "kernel_timing": ("the timed steps bracket only %s's launches with HIP events; every "
                          "kernel's time (kernels_ms_per_step, edt) comes from a profiled repeat "
                          "of the same steps" % DOMINANT_KERNEL),
        "parity_vs_oracle": p["parity"],
        "host_pointer_call": None if p["host_ms"] is None else {
            "ms": p["host_ms"], "gvox_per_s_pcie_inclusive": wl.n / (p["host_ms"] * 1e-3) / 1e9},
        "tb_th_mean_um": p["res"].thickness.mean,
        "tb_sp_mean_um": p["res"].spacing.mean,
        "batched": batched,
        "serving": serving,
        "serving_saturated": saturated,
        "smoothed_sdf": sdf,
        "full_analyze": full,
        "tiff_input": tiff,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
