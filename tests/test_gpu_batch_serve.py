"""GPU tests of the batched and serving forms: every volume's result must be
bit-identical to the one-volume call (bm_analyze_thickness_spacing_device) on
the same input, with a hipGraph or without, alone or mixed with other
requests, and whatever the tunables; per-volume errors are the one-volume
call's errors."""
import json
import math
import os
import time

import numpy as np
import pytest

import _oracle as orc
import bone_morphometry_amd as bm
from _stats import check_stats
from bone_morphometry_amd import _native, hip, synth, workloads

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def _check_golden(got, g, sequential=False):
    """A ThicknessSpacing against an oracle record of gen_workload_golden.py."""
    assert not isinstance(got, bm.BmError), str(got)
    check_stats(got.thickness, g["thickness"], sequential, g["spacing"])
    check_stats(got.spacing, g["spacing_stats"], sequential, g["spacing"])
    assert (got.bone_volume, got.total_volume, got.bone_volume_fraction) == \
        (g["bone_volume"], g["total_volume"], g["bone_volume_fraction"])


def _volumes(seed, count, lo=24, hi=72):
    """Small volumes of uneven shapes from the synthetic generator."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        w, h, d = (int(v) for v in rng.integers(lo, hi + 1, size=3))
        bone = synth.volume(w, h, d, float(rng.uniform(1.5, 3.0)), float(rng.uniform(0.1, 0.35)),
                            int(rng.integers(1, 2**31)), 0)
        roi = synth.roi_cylinder(w, h, d, float(rng.uniform(0.7, 0.95)) * min(w, h))
        out.append(dict(w=w, h=h, d=d, bone=bone, roi=roi,
                        spacing=(float(rng.uniform(6, 20)),) * 3))
    return out


def _upload(vols):
    bufs = []
    for v in vols:
        b = hip.DeviceBuffer(v["bone"].size)
        r = hip.DeviceBuffer(v["roi"].size)
        b.upload(v["bone"])
        r.upload(v["roi"])
        bufs.append((b, r))
    return bufs


def _items(vols, bufs, mode=bm.FeatureMode.Standard):
    return [(b.value, r.value, v["w"], v["h"], v["d"], v["spacing"], mode)
            for v, (b, r) in zip(vols, bufs)]


def _plain(item):
    b, r, w, h, d, sp, mode = item
    try:
        return bm.api.analyze_thickness_spacing_device(b, r, w, h, d, sp, mode)
    except bm.BmError as e:
        return e


def _same(a, b):
    """Bit-identical results (NaN quantiles compare equal), or equal errors."""
    if isinstance(a, bm.BmError) or isinstance(b, bm.BmError):
        return (isinstance(a, bm.BmError) and isinstance(b, bm.BmError) and str(a) == str(b)
                and a.code == b.code)
    for x, y in ((a.thickness, b.thickness), (a.spacing, b.spacing)):
        for f in ("mean", "standard_deviation", "maximum", "median", "percentile_10",
                  "percentile_25", "percentile_75", "percentile_90", "interquartile_range",
                  "coefficient_of_variation", "count"):
            u, v = getattr(x, f), getattr(y, f)
            if not (u == v or (isinstance(u, float) and math.isnan(u) and math.isnan(v))):
                return False
    return (a.bone_volume, a.total_volume, a.bone_volume_fraction) == \
        (b.bone_volume, b.total_volume, b.bone_volume_fraction)


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_one_volume_calls(use_graph):
    vols = _volumes(11, 7)
    bufs = _upload(vols)
    items = _items(vols, bufs)
    items[3] = items[3][:6] + (bm.FeatureMode.Classic,)  # quantiles off for one volume
    ref = [_plain(it) for it in items]
    batch = bm.Batch(items, use_graph=use_graph)
    assert batch.device_bytes() > 0
    for _ in range(2):  # the graph is replayed: every run starts from a clean state
        got = batch.run()
        assert all(_same(g, r) for g, r in zip(got, ref))
    batch.close()
    # and each against the oracle
    for v, it, g in zip(vols, items, got):
        o = orc.analyze_thickness_spacing(v["bone"], v["roi"], v["w"], v["h"], v["d"],
                                          v["spacing"],
                                          include_quantiles=it[6] == bm.FeatureMode.Standard)
        check_stats(g.thickness, o.thickness, scale=v["spacing"][0])
        check_stats(g.spacing, o.spacing, scale=v["spacing"][0])


def test_batched_workload_matches_oracle():
    """The 96-volume batched workload (axes 64..384, sigma 2..6, BV/TV
    0.08..0.35) as one graph-replayed batch: every volume's statistics against
    the oracle's (tests/golden/batched_oracle.json), in both summation
    orders."""
    golden = _golden("batched_oracle.json")["volumes"]
    params = workloads.batched_params()
    assert len(golden) == len(params) == 96
    vols, bufs, items = [], [], []
    for p in params:
        wl = workloads.make_volume(p)
        b, r = hip.DeviceBuffer(wl.n), hip.DeviceBuffer(wl.n)
        b.upload(wl.bone)
        r.upload(wl.roi)
        bufs.append((b, r))
        items.append((b.value, r.value, wl.width, wl.height, wl.depth, wl.spacing,
                      bm.FeatureMode.Standard))
    batch = bm.Batch(items, use_graph=True)
    for _ in range(2):
        got = batch.run()
        for g, gold in zip(got, golden):
            _check_golden(g, gold)
    batch.close()
    with bm.stats_order(bm.StatsOrder.Sequential):
        batch = bm.Batch(items, use_graph=True)
    got = batch.run()
    batch.close()
    for g, gold in zip(got, golden):
        _check_golden(g, gold, sequential=True)


def test_serving_stream_matches_oracle():
    """A time-compressed serving stream: Poisson arrivals at 150/s for 4 s
    (about 600 requests) over the 24-volume request pool (lognormal sizes of
    median 192^3, each volume its own sigma and BV/TV), submitted at their
    arrival times: every result equals the oracle's for its volume
    (tests/golden/serving_oracle.json)."""
    g = _golden("serving_oracle.json")
    pool = workloads.serving_pool(**g["pool"])
    golden = g["volumes"]
    arrivals = workloads.serving_arrivals(rate=150.0, duration=4.0, seed=7)
    srv = bm.Server(budget_bytes=48 << 30, max_batch=32, slots=2)
    t0 = time.perf_counter()
    tickets = []
    for i, at in enumerate(arrivals):
        delay = at - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        q = pool[i % len(pool)]
        tickets.append(srv.submit(q.bone, q.roi, q.width, q.height, q.depth, q.spacing))
    sizes = []
    for i, t in enumerate(tickets):
        res, tim = srv.poll(t, wait=True)
        sizes.append(tim["batch_size"])
        _check_golden(res, golden[i % len(pool)])
    st = srv.stats()
    srv.close()
    assert st["requests_done"] == len(tickets) > 400
    assert max(sizes) > 1  # requests that arrived together shared a batch


def test_batch_reports_per_volume_errors():
    vols = _volumes(12, 3)
    bufs = _upload(vols)
    items = _items(vols, bufs)
    # an empty ROI, a bad spacing and a bad shape, among valid volumes
    empty = hip.DeviceBuffer(vols[0]["roi"].size)
    empty.upload(np.zeros_like(vols[0]["roi"]))
    items.append((bufs[0][0].value, empty.value, vols[0]["w"], vols[0]["h"], vols[0]["d"],
                  vols[0]["spacing"], bm.FeatureMode.Standard))
    items.append(items[1][:5] + ((1.0, float("nan"), 1.0), bm.FeatureMode.Standard))
    items.append((items[1][0], items[1][1], 1, 4, 4, (1.0, 1.0, 1.0), bm.FeatureMode.Standard))
    got = bm.Batch(items).run()
    ref = [_plain(it) for it in items]
    assert isinstance(got[3], bm.BmError) and "ROI mask contains no included voxels" in str(got[3])
    assert isinstance(got[4], bm.BmError) and "[1.0, NaN, 1.0]" in str(got[4])
    assert isinstance(got[5], bm.BmError)
    assert all(_same(g, r) for g, r in zip(got, ref))


def test_server_result_alone_equals_result_in_mix():
    vols = _volumes(13, 10)
    srv = bm.Server(budget_bytes=4 << 30, max_batch=4, slots=2)
    a = vols[0]
    t = srv.submit(a["bone"], a["roi"], a["w"], a["h"], a["d"], a["spacing"])
    alone, tim = srv.poll(t, wait=True)
    assert tim["batch_size"] == 1 and tim["done_s"] >= tim["start_s"] >= tim["submit_s"]
    tickets = [srv.submit(v["bone"], v["roi"], v["w"], v["h"], v["d"], v["spacing"])
               for v in vols[1:5] + [a] + vols[5:]]
    res = [srv.poll(t, wait=True) for t in tickets]
    mixed = res[4][0]
    assert _same(alone, mixed)
    assert max(r[1]["batch_size"] for r in res) > 1  # they were batched together
    bufs = _upload(vols)
    ref = [_plain(it) for it in _items(vols, bufs)]
    order = [1, 2, 3, 4, 0, 5, 6, 7, 8, 9]
    assert all(_same(res[k][0], ref[order[k]]) for k in range(len(res)))
    st = srv.stats()
    assert st["requests_done"] == 11 and st["batches"] >= 3 and st["max_batch"] <= 4
    srv.close()


def test_server_rejects_requests_beyond_the_budget():
    vols = _volumes(14, 1, 64, 64)
    v = vols[0]
    srv = bm.Server(budget_bytes=1 << 20, max_batch=4, slots=2)  # 512 KiB per slot
    t = srv.submit(v["bone"], v["roi"], v["w"], v["h"], v["d"], v["spacing"])
    res, _ = srv.poll(t, wait=True)
    assert isinstance(res, bm.BmError) and res.code == _native.BM_ERR_UNSUPPORTED
    t = srv.submit(v["bone"], v["roi"], v["w"], v["h"], v["d"], (0.0, 1.0, 1.0))
    res, _ = srv.poll(t, wait=True)
    assert isinstance(res, bm.BmError) and "finite and positive" in str(res)
    srv.close()


def test_autotune_writes_cache_and_keeps_results(tmp_path):
    vols = _volumes(15, 3, 40, 140)
    bufs = _upload(vols)
    items = _items(vols, bufs)
    before = [_plain(it) for it in items]
    path = tmp_path / "tune.json"
    bm.autotune(items, path)
    cache = json.loads(path.read_text())
    (arch, classes), = cache.items()
    assert arch.startswith("gfx")
    assert len(classes) >= 1
    for entry in classes.values():
        assert entry["y_rows"] in (64, 128, 256, 512) and entry["z_rows"] in (64, 128, 256, 512)
        assert entry["fill_cap"] in (512, 1024) and entry["fill_tiles"] in (1, 2)
    bm.tuning_load(path)
    after = [_plain(it) for it in items]
    assert all(_same(a, b) for a, b in zip(after, before))
    batch = bm.Batch(items).run()
    assert all(_same(a, b) for a, b in zip(batch, before))
    bm.tuning_load(tmp_path / "absent.json")  # back to the defaults


@pytest.mark.parametrize("tile_order,fill_tiles,fill_cap", [(7, 1, 1024), (7, 2, 512), (2, 2, 1024),
                                                            (4, 1, 512)])
def test_tile_orders_and_fill_shapes_keep_results(tmp_path, tile_order, fill_tiles, fill_cap):
    """The grouped (4x4x4) tile order of the ridge, fill and clean-up kernels,
    two-tile fill blocks and the smaller fill capacity (more rounds) on
    volumes whose sides are not multiples of the tile groups: every result
    bit-identical to the default kernels'."""
    vols = _volumes(23, 3, 40, 140)
    bufs = _upload(vols)
    items = _items(vols, bufs)
    bm.tuning_load(tmp_path / "absent.json")  # the defaults
    before = [_plain(it) for it in items]
    classes = {f"h{(v['h'] + 63) // 64 * 64}_d{(v['d'] + 63) // 64 * 64}" for v in vols}
    entry = {"y_rows": 128, "y_halo": 16, "z_rows": 128, "z_halo": 16, "fill_cap": fill_cap,
             "tile_order": tile_order, "fill_tiles": fill_tiles}
    path = tmp_path / "tune.json"
    path.write_text(json.dumps({"gfx950": {c: entry for c in classes}}))
    bm.tuning_load(path)
    try:
        after = [_plain(it) for it in items]
    finally:
        bm.tuning_load(tmp_path / "absent.json")
    assert all(_same(a, b) for a, b in zip(after, before))


def test_server_failures_and_cancellation_leave_other_requests_unchanged(monkeypatch):
    """Per-request isolation in the server (the reference analyzes file after
    file and reports each failure on its own, src/main.rs:254-268): a batch
    whose device allocation fails (forced on the third batch by the test
    hook), a request with an empty ROI and a cancelled request each fail
    alone; every other request equals the one-volume call, and requests after
    the failures are served."""
    monkeypatch.setenv("BM_SERVER_FAIL_BATCH", "2")
    big = _volumes(17, 3, 150, 170)
    vols = _volumes(16, 7)
    empty = dict(vols[3])
    empty["roi"] = np.zeros_like(empty["roi"])
    reqs = big + vols[:3] + [empty] + vols[3:]
    srv = bm.Server(budget_bytes=8 << 30, max_batch=1, slots=1)
    tickets = [srv.submit(v["bone"], v["roi"], v["w"], v["h"], v["d"], v["spacing"])
               for v in reqs]
    assert srv.cancel(tickets[-1])  # queued behind the large volumes
    res = [srv.poll(t, wait=True)[0] for t in tickets]
    st = srv.stats()
    srv.close()
    bufs = _upload(reqs)
    ref = [_plain(it) for it in _items(reqs, bufs)]
    for k, (g, r) in enumerate(zip(res, ref)):
        if k == 2:
            assert isinstance(g, bm.BmError) and g.code == _native.BM_ERR_DEVICE
            assert "injected device allocation failure" in str(g)
        elif k == len(reqs) - 1:
            assert isinstance(g, bm.BmError) and str(g) == "request cancelled"
        else:
            assert _same(g, r), k
    assert isinstance(res[6], bm.BmError) and "ROI mask contains no included voxels" in str(res[6])
    assert st["requests_done"] == len(reqs) - 1  # the cancelled one never ran
    with pytest.raises(bm.BmError, match="unknown ticket"):
        s2 = bm.Server(budget_bytes=1 << 30, max_batch=1, slots=1)
        try:
            s2.cancel(12345)
        finally:
            s2.close()


def test_server_concurrent_clients_match_one_volume_calls():
    """Four client threads submit, poll and cancel on one server at once (the
    C calls release the GIL, so the server's queue, worker thread and result
    table are driven concurrently): every request that was not cancelled
    equals the one-volume call on its volume."""
    import threading
    vols = _volumes(18, 12)
    srv = bm.Server(budget_bytes=4 << 30, max_batch=4, slots=2)
    results = {}
    errors = []

    def client(k):
        try:
            mine = list(range(k, len(vols), 4))
            tickets = [(i, srv.submit(vols[i]["bone"], vols[i]["roi"], vols[i]["w"],
                                      vols[i]["h"], vols[i]["d"], vols[i]["spacing"]))
                       for i in mine]
            for i, t in tickets:
                results[i] = srv.poll(t, wait=True)[0]
            # a request cancelled right after its submission, and one after its result
            v = vols[mine[0]]
            t = srv.submit(v["bone"], v["roi"], v["w"], v["h"], v["d"], v["spacing"])
            srv.cancel(t)
            res = srv.poll(t, wait=True)[0]
            if isinstance(res, bm.BmError):
                assert str(res) == "request cancelled"
            else:
                assert _same(res, results[mine[0]])
            srv.stats()
        except BaseException as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    srv.close()
    assert not errors, errors
    bufs = _upload(vols)
    ref = [_plain(it) for it in _items(vols, bufs)]
    assert all(_same(results[i], ref[i]) for i in range(len(vols)))
