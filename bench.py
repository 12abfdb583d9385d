"""Benchmark of the exact squared EDT + local thickness (Tb.Th and Tb.Sp) path.

One step = one bm_analyze_thickness_spacing_device() call on the 512^3
workload (plain_512: sigma 4, BV/TV 0.15, z-cylinder ROI of diameter 460,
10.5 um, Standard mode): both phases through EDT, ridges, sphere fill,
clean-up, mask and statistics, with bone and ROI already resident in HBM.
value = voxels / step time (Gvoxel/s, the volume counted once).

The roofline object describes the kernel with the largest share of the timed
region, from per-kernel HIP events recorded on the launching stream;
algorithmic bytes per voxel of each kernel are listed in KERNEL_BYTES.  The
CPU baseline runs the oracle (a C restatement of the reference algorithm,
OpenMP) on the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per voxel of one launch (one phase), plus bytes per ridge.
KERNEL_BYTES = {
    "edt_x": (6.0, 0.0),        # bone + ROI u8 in, int32 row distance out
    "edt_y": (8.0, 0.0),        # int32 in, int32 out
    "edt_z": (8.0, 0.0),        # int32 in, int32 out
    "ridges": (4.0, 8.0),       # squared EDT in; (index, r^2) per ridge out
    "fill_scatter": (4.0, 8.0),  # propagated map out; ridge list in
    "fill_gather": (4.0, 8.0),
    "cleanup_stats": (12.0, 0.0),  # propagated map + squared EDT in, map out
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="plain_512", choices=["plain_512", "plain_small"])
    ap.add_argument("--mode", default="standard", choices=["standard", "classic"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (volumes shard "
              "across GPUs by whole volumes; see DESIGN.md). Run with --gpus 1.",
              file=sys.stderr)
        return 2
    import numpy as np
    import bone_morphometry_amd as bm
    from bone_morphometry_amd import _native, hip, workloads

    if _native.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    cpu_threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    t0 = time.time()
    wl = getattr(workloads, args.workload)(threads=cpu_threads)
    print(f"[bench] generated {wl.name} in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
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
    _native.profile_read()  # drop warm-up records
    _native.profile_enable(True)
    e0, e1 = hip.Event(), hip.Event()
    hip.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        res = step()
    e1.record(stream)
    hip.synchronize()
    total_ms = e0.elapsed_ms(e1)
    _native.profile_enable(False)
    kernels = _native.profile_read()
    print(f"[bench] {args.steps} steps in {total_ms:.2f} ms", file=sys.stderr, flush=True)

    ms_per_step = total_ms / args.steps
    gvox = wl.n / (ms_per_step * 1e-3) / 1e9

    # Roofline of the kernel with the largest time share.
    phases = 2
    dom = max(kernels.items(), key=lambda kv: kv[1][1])
    name, (launches, kms) = dom
    avg_ms = kms / launches
    per_vox, per_ridge = KERNEL_BYTES.get(name, (0.0, 0.0))
    alg_bytes = per_vox * wl.n
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9 if per_vox else None
    roofline = {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                "traffic": None, "avg_launch_ms": avg_ms,
                "algorithmic_bytes_per_launch": alg_bytes}
    edt = {k: kernels[k][1] / kernels[k][0] for k in ("edt_x", "edt_y", "edt_z") if k in kernels}
    edt_ms = sum(edt.values())
    edt_gbs = 22.0 * wl.n / (edt_ms * 1e-3) / 1e9 if edt_ms else None

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as orc
        orc.set_threads(cpu_threads)
        t1 = time.time()
        ref = orc.analyze_thickness_spacing(wl.bone, wl.roi, wl.width, wl.height, wl.depth,
                                            wl.spacing, include_quantiles=(mode ==
                                                                           bm.FeatureMode.Standard))
        cpu_s = time.time() - t1
        cpu = {"value": wl.n / cpu_s / 1e9, "unit": "Gvoxel/s", "cores": cpu_threads,
               "kind": "port",
               "sample": f"the full {wl.name} workload, both phases, oracle (C/OpenMP), "
                         f"{cpu_s:.1f} s"}
        ok = True
        for g, r in ((res.thickness, ref.thickness), (res.spacing, ref.spacing)):
            ok &= g.count == r.count and g.maximum == r.maximum
            ok &= abs(g.mean - r.mean) <= 1e-12 * abs(r.mean)
            tol = max(1e-12, (r.count - 1) * 2.0 ** -53 *
                      (1 + 1 / r.coefficient_of_variation ** 2) / 2)
            ok &= abs(g.standard_deviation - r.standard_deviation) <= tol * r.standard_deviation
            if mode == bm.FeatureMode.Standard:
                ok &= g.median == r.median and g.percentile_90 == r.percentile_90
        parity = bool(ok)

    line = {
        "metric": "Gvoxel/s EDT+local thickness at 512^3",
        "value": gvox,
        "unit": "Gvoxel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": wl.name, "shape": [wl.width, wl.height, wl.depth],
                   "roi": "z-cylinder d=460" if wl.name == "plain_512" else "full",
                   "spacing_um": wl.spacing[0], "mode": args.mode, "phases": phases},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "edt": {"ms_per_phase": edt, "gbs_22B_per_voxel": edt_gbs,
                "frac": (edt_gbs / HBM_PEAK_GBS) if edt_gbs else None},
        "kernels_ms_per_step": {k: v[1] / args.steps for k, v in kernels.items()},
        "parity_vs_oracle": parity,
        "tb_th_mean_um": res.thickness.mean,
        "tb_sp_mean_um": res.spacing.mean,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
