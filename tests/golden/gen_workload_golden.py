"""Writes the oracle's results on the benchmark workloads as golden fixtures.

The oracle (oracle/liboracle.so, the CPU restatement of the reference) needs
minutes to an hour on these sizes, so its results are computed once here and
committed; the GPU tests compare the HIP path against them.  The volumes are
regenerated by the seeded generator (libbm_synth.so), whose output depends
only on its arguments.

  large      1024x1024x1536, sigma (5, 5, 9), BV/TV 0.12, ellipsoid ROI
             (500, 480, 740), 8 um: per phase (bone and marrow inside the ROI)
             the xxh3-128 digests of the squared EDT (u32), ridge flags (u8),
             propagated r^2 (u32) and thickness map (f32), the ridge count,
             the fill work, and nonzero_stats with quantiles at 8 um.  Each
             phase runs in its own process (about 50 GB of host memory).
  batched    the 96 volumes of the batched workload: analyze's thickness and
             spacing statistics (Standard mode) and volumes.
  serving    the request pool of the serving test (24 volumes, seed 6): the
             same fields.

usage: python tests/golden/gen_workload_golden.py large|batched|serving [threads]
"""
import json
import os
import subprocess
import sys
import time

import numpy as np
import xxhash

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
ROOT = os.path.dirname(TESTS)
sys.path.insert(0, ROOT)
sys.path.insert(0, TESTS)

import _oracle as orc  # noqa: E402
from bone_morphometry_amd import workloads  # noqa: E402

STATS_FIELDS = ("mean", "standard_deviation", "maximum", "median", "percentile_10",
                "percentile_25", "percentile_75", "percentile_90", "interquartile_range",
                "coefficient_of_variation", "count")

SERVING_POOL = dict(count=24, seed=6)


def digest(a):
    return xxhash.xxh3_128_hexdigest(memoryview(np.ascontiguousarray(a)).cast("B"))


def stats_dict(st):
    return {f: (int(getattr(st, f)) if f == "count" else float(getattr(st, f)))
            for f in STATS_FIELDS}


def large_phase(phase_name):
    """One phase of the large workload (run in a child process)."""
    wl = workloads.large()
    sel = orc.phase_inside(wl.bone, wl.roi, phase_name == "bone")
    del wl.bone, wl.roi
    t = time.time()
    st = orc.local_thickness_stages(sel, 1024, 1024, 1536, 0)
    out = dict(seconds=time.time() - t,
               squared_edt=digest(st["squared_edt"]), ridge_flags=digest(st["ridge_flags"]),
               propagated=digest(st["propagated"]), thickness=digest(st["thickness"]),
               ridge_count=int(st["ridge_count"]), fill_visits=int(st["fill_visits"]),
               max_d2=int(st["squared_edt"].max()), selected=int(np.count_nonzero(sel)))
    del st["squared_edt"], st["ridge_flags"], st["propagated"]
    out["stats"] = stats_dict(orc.nonzero_stats(st["thickness"], 8.0, True))
    print(json.dumps(out))


def large():
    # resumable: a phase already in the output file is kept
    path = os.path.join(HERE, "large_oracle.json")
    res = {"workload": "large_1024x1024x1536", "spacing": 8.0}
    if os.path.exists(path):
        with open(path) as f:
            res.update(json.load(f))
    for ph in ("bone", "marrow"):
        if ph in res:
            continue
        r = subprocess.run([sys.executable, __file__, "large-phase", ph], check=True,
                           stdout=subprocess.PIPE, text=True)
        res[ph] = json.loads(r.stdout.strip().splitlines()[-1])
        print(ph, res[ph]["seconds"], flush=True)
        with open(path, "w") as f:
            json.dump(res, f, indent=1)
    return res


def analyze_record(wl):
    m = orc.analyze_thickness_spacing(wl.bone, wl.roi, wl.width, wl.height, wl.depth,
                                      wl.spacing)
    return dict(shape=[wl.width, wl.height, wl.depth], spacing=wl.spacing[0],
                thickness=stats_dict(m.thickness), spacing_stats=stats_dict(m.spacing),
                bone_volume=m.bone_volume, total_volume=m.total_volume,
                bone_volume_fraction=m.bone_volume_fraction)


def batched():
    out = []
    for i, p in enumerate(workloads.batched_params()):
        out.append(analyze_record(workloads.make_volume(p)))
        print(i, out[-1]["shape"], flush=True)
    return {"workload": "batched_96", "volumes": out}


def serving():
    out = []
    for i, p in enumerate(workloads.serving_pool_params(**SERVING_POOL)):
        out.append(analyze_record(workloads.make_volume(p)))
        print(i, out[-1]["shape"], flush=True)
    return {"workload": "serving_pool", "pool": SERVING_POOL, "volumes": out}


def main():
    what = sys.argv[1]
    if what == "large-phase":
        large_phase(sys.argv[2])
        return
    if len(sys.argv) > 2:
        orc.set_threads(int(sys.argv[2]))
    res = {"large": large, "batched": batched, "serving": serving}[what]()
    path = os.path.join(HERE, f"{what}_oracle.json")
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
