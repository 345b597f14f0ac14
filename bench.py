"""Benchmark of the MI355X spatial-query path (one JSON line on stdout).

Headline workload (the plain quest::SignedDistance FP64 configuration):
icosphere level 6 (40,962 vertices, 81,920 triangles, r = 1, each vertex
scaled by 1 + U[-1e-3, 1e-3]) and 4,000,000 queries U[-1.5,1.5]^3.
One step = signed-distance forward (distance, closest point, winning
element) + adjoint (query and vertex gradients for an N(0,1) FP32 upstream
gradient) over all queries, inputs resident in HBM.

  value            FP64-plain forward+backward queries/s
  fp32_traversal   the same step with FP32 traversal (FP64 finish), and its
                   error against the FP64 result
  roofline         dominant kernel (FP64 SDF forward) timed with HIP events on
                   its stream; algorithmic bytes per query
                   = 24 (query) + 8 (distance) + 24 (closest point) + 4 (element)
                   + V*104 (inner-node visits: 2 x 48 B boxes + 2 x 4 B links)
                   + L*84 (leaf checks: 3 x 24 B vertices + 3 x 4 B indices),
                   V, L = mean visits / leaf checks of the reference traversal
                   (the FP64 kernel follows it exactly; counted on the GPU);
                   traffic = HBM bytes per launch from the committed counter runs
  bvh_build        BVH build Mprims/s over the 81,920 triangles and over the
                   100,000 boxes of the findPoints workload
  find_points      100,000 boxes / 1,000,000 points: findPoints queries/s (FP64
                   and FP32 traversal), the count pass's roofline (24 B point +
                   8 B offset and count + 4 B per candidate + V x 104 B)
  configs          the three larger workloads at their full sizes: torus
                   (2,097,152 triangles, 16M queries, forward), wavy L7
                   icosphere (1M queries, forward + adjoint), L8 shape fit (8M
                   queries, 20 iterations): ms, queries/s, the forward's
                   roofline, the FP32 form's error against FP64, per-iteration
                   build / forward / backward times
  cpu_baseline     the CPU oracle (restatement of the reference's SEQ path),
                   forward+backward on the first 1,000,000 queries, timed on
                   every CPU this process may use (and on one thread for the
                   first 100,000); each configuration carries its own on a
                   bounded sample; host gives the thread count, nproc, the
                   affinity set, the cgroup CPU quota and the CPU model
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("SDF queries/s fwd+bwd, FP32-trav vs FP64 plain @≤1e-6 rel err; "
          "BVH build Mprims/s")
HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "sdf_eval_fp64_icosphere6_4M.traffic.json")


def parse():
    p = argparse.ArgumentParser(description=__doc__,
                                formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--queries", type=int, default=4_000_000)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-extras", action="store_true",
                   help="only the headline step (for profiler runs)")
    p.add_argument("--calib-copy", action="store_true",
                   help="also run one 1 GiB device copy (counter calibration)")
    p.add_argument("--cpu-sample", type=int, default=1_000_000)
    p.add_argument("--workload", default="sdf_l6",
                   choices=["sdf_l6", "find_points", "torus_mixed", "diff_l7", "shapefit_l8",
                            "self_isect_l8", "find_points_2d", "sdf_quads", "hex_candidates"],
                   help="sdf_l6 is the headline; the others print their own JSON line")
    return p.parse_args()


def host_cpus():
    """Threads for the CPU legs and a description of the host: every CPU this
    process may run on (its affinity set, capped by a cgroup CPU quota when
    one is set), with nproc and the CPU model."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(-(-int(q) // int(per))))
    except (OSError, ValueError):
        pass
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    threads = min(aff, quota) if quota else aff
    return threads, {"threads_used": threads, "nproc": os.cpu_count(), "affinity_cpus": aff,
                     "cgroup_quota_cpus": quota, "model": model}


def sync_timer(torch):
    def sync_time(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps
    return sync_time


def prof_reader(L, _lib):
    def get(name):
        n, ms = ctypes.c_int64(), ctypes.c_double()
        _lib.check(L.axsq_prof_get(name.encode(), ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def prof(name):  # ms per launch of the named range
        n, ms = get(name)
        return ms / max(n, 1)
    prof.launches = lambda name: get(name)[0]
    return prof


def sdf_roofline(kernel, ms, vbar, lbar, node_bytes, nq):
    """Roofline of one signed-distance forward launch over nq queries: 24 B
    query + 8 B distance + 24 B closest point + 4 B element + V node visits x
    node_bytes (104 FP64 / 56 FP32) + L leaf checks x 84 B."""
    bpq = 24 + 8 + 24 + 4 + node_bytes * vbar + 84.0 * lbar
    ach = bpq * nq / (ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "kernel": kernel, "avg_launch_ms": ms,
            "bytes_per_query": bpq, "mean_inner_visits": vbar, "mean_leaf_checks": lbar}


def sdf_config(torch, dev, L, _lib, v, t, q, g, steps, want_bwd, cpu_threads, cpu_sample,
               label):
    """Signed distance over one mesh and query set: FP64 and FP32-traversal
    forward (+ adjoint when want_bwd), ms per step, queries/s, the forward's
    roofline with the reference traversal's V and L (counted by the FP64
    kernel), the FP32 form's max / RMS relative error and sign mismatches
    against FP64; with cpu_threads, the oracle on a bounded sample."""
    from axom_amd import quest
    vd, td, qd = (torch.as_tensor(a, device=dev) for a in (v, t, q))
    gd = torch.as_tensor(g, device=dev) if g is not None else None
    nq = q.shape[0]
    sdf = quest.SignedDistance(vd, td, isWatertight=True, computeSign=True, fp_mode="fp32")
    dist = torch.empty(nq, dtype=torch.float64, device=dev)
    elem = torch.empty(nq, dtype=torch.int32, device=dev)
    prof = prof_reader(L, _lib)
    out = {"workload": label, "queries": nq, "triangles": int(t.shape[0])}
    res = {}
    for mode in ("fp64", "fp32"):
        def step():
            sdf.computeDistances(nq, qd, outSgnDist=dist, outElem=elem, fp_mode=mode)
            if want_bwd:
                sdf.backward(qd, dist, elem, gd)
        step()
        torch.cuda.synchronize()
        L.axsq_prof_reset()
        L.axsq_prof_enable(1)
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / steps
        L.axsq_prof_enable(0)
        res[mode] = dist.clone()
        key = "fwd_bwd" if want_bwd else "fwd"
        out[mode] = {"ms_per_step": el * 1e3, f"{key}_queries_per_s": nq / el,
                     "sdf_eval_ms": prof(f"sdf_eval_{mode}"),
                     "second_pass_walks": sdf.lastSuspended()}
        if want_bwd:
            out[mode]["sdf_backward_ms"] = prof("sdf_backward")
    sdf.enableStats(True)
    sdf.computeDistances(nq, qd, outSgnDist=dist, outElem=elem, fp_mode="fp64")
    visits, leaves = sdf.lastStats()
    sdf.enableStats(False)
    vbar, lbar = visits / nq, leaves / nq
    out["fp64"]["roofline"] = sdf_roofline("k_sdf_eval<FP64>", out["fp64"]["sdf_eval_ms"], vbar,
                                           lbar, 104.0, nq)
    out["fp32"]["roofline"] = sdf_roofline("k_sdf_eval<FP32 traversal>",
                                           out["fp32"]["sdf_eval_ms"], vbar, lbar, 56.0, nq)
    rel = (res["fp32"] - res["fp64"]).abs() / res["fp64"].abs().clamp_min(1e-300)
    out["fp32"]["max_rel_err_vs_fp64"] = float(rel.max())
    out["fp32"]["rms_rel_err_vs_fp64"] = float(torch.sqrt((rel * rel).mean()))
    out["fp32"]["sign_mismatches"] = int((torch.signbit(res["fp32"]) !=
                                          torch.signbit(res["fp64"])).sum())
    if cpu_threads:
        from oracle import oracle as orc
        # an even sample over the whole query set (the sets are ordered in parts)
        idx = np.linspace(0, nq - 1, min(cpu_sample, nq)).astype(np.int64)
        qs = np.ascontiguousarray(q[idx])
        t0 = time.perf_counter()
        o = orc.OracleSDF(v, t)
        tb = time.perf_counter() - t0
        t0 = time.perf_counter()
        r = o.eval(qs, nthreads=cpu_threads, want_normal=False)
        if want_bwd:
            o.backward(qs, r["dist"], r["elem"], np.ascontiguousarray(g[idx]))
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "queries_per_s": len(idx) / tc, "cores": cpu_threads, "kind": "port",
            "sample": f"{len(idx):,} queries evenly spaced over the {nq:,}, FP64 forward" +
                      (" + adjoint" if want_bwd else ""),
            "seconds": tc, "build_seconds_1_thread": tb,
            "matches_gpu_bitwise": bool(np.array_equal(r["dist"],
                                                       res["fp64"][idx].cpu().numpy()))}
    return out


def shapefit_config(torch, dev, v, t, q, target, cpu_threads, label, iterations=20):
    """The shape-fitting loop: per-iteration BVH rebuild, FP32-traversal
    forward and adjoint times, loss; with cpu_threads, the oracle's rebuild
    (one thread) and forward + adjoint over a bounded sample of the queries."""
    from axom_amd import shapefit
    vd, td, qd, tg = (torch.as_tensor(a, device=dev) for a in (v, t, q, target))
    vf, hist = shapefit.shape_fit(vd, td, qd, tg, iterations=iterations, lr=1e-3,
                                  fp_mode="fp32")
    n = len(hist)
    out = {"workload": label, "queries": int(q.shape[0]), "triangles": int(t.shape[0]),
           "iterations": n,
           "mean_build_ms": sum(h["build_ms"] for h in hist[1:]) / max(n - 1, 1),
           "mean_forward_ms": sum(h["forward_ms"] for h in hist) / n,
           "mean_backward_ms": sum(h["backward_ms"] for h in hist) / n,
           "loss_first_last": [hist[0]["loss"], hist[-1]["loss"]],
           "per_iteration": [{k: (round(x, 4) if isinstance(x, float) and k != "loss" else x)
                              for k, x in h.items()} for h in hist]}
    it_ms = out["mean_build_ms"] + out["mean_forward_ms"] + out["mean_backward_ms"]
    out["ms_per_iteration"] = it_ms
    out["fwd_bwd_queries_per_s"] = q.shape[0] / (it_ms / 1e3)
    out["build_mprims_per_s"] = t.shape[0] / (out["mean_build_ms"] / 1e3) / 1e6
    if cpu_threads:
        from oracle import oracle as orc
        m = min(200_000, q.shape[0])
        idx = np.linspace(0, q.shape[0] - 1, m).astype(np.int64)
        qs = np.ascontiguousarray(q[idx])
        t0 = time.perf_counter()
        o = orc.OracleSDF(v, t)
        tb = time.perf_counter() - t0
        t0 = time.perf_counter()
        r = o.eval(qs, nthreads=cpu_threads, want_cp=False, want_normal=False)
        gs = ((r["dist"] - target[idx]) * (2.0 / q.shape[0])).astype(np.float32)
        o.backward(qs, r["dist"], r["elem"], gs)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"build_ms_1_thread": tb * 1e3,
                               "fwd_bwd_queries_per_s": m / tc, "cores": cpu_threads,
                               "kind": "port",
                               "sample": f"first iteration: {m:,} of the queries, evenly "
                                         "spaced, FP64 forward + adjoint"}
    return out


def find_points_config(torch, dev, L, _lib, boxes, pts, cpu_threads, reps=10):
    """BVH build over the boxes and findPoints over the points (FP64 and FP32
    traversal), with the count pass's roofline: 24 B point + 8 B offset and
    count + 4 B per candidate + V inner-node visits x 104 B, V counted by the
    FP64 count pass (the reference walk's own visits)."""
    from axom_amd import spin
    prof = prof_reader(L, _lib)
    sync_time = sync_timer(torch)
    bx, p = torch.as_tensor(boxes, device=dev), torch.as_tensor(pts, device=dev)
    n = p.shape[0]
    out = {"workload": f"{boxes.shape[0]:,} boxes, {n:,} points; BVH build + findPoints",
           "queries": n, "boxes": int(boxes.shape[0])}
    bvh = spin.BVH(bx)
    off = torch.empty(n, dtype=torch.int32, device=dev)
    cnt = torch.empty(n, dtype=torch.int32, device=dev)
    tb = sync_time(lambda: bvh.initialize(bx), 20)
    out["build_ms"] = tb * 1e3
    out["build_mprims_per_s"] = boxes.shape[0] / tb / 1e6
    res = {}
    # a first call keeps no candidate ids yet (count, scan, fill); the timed
    # calls and their kernel ranges are those of calls after it
    bvh.findPoints(off, cnt, n, p)
    torch.cuda.synchronize()
    L.axsq_prof_reset()
    L.axsq_prof_enable(1)
    tq = sync_time(lambda: res.__setitem__("c", bvh.findPoints(off, cnt, n, p)), reps)
    L.axsq_prof_enable(0)
    calls = reps + 1  # sync_time's own warm-up call included
    cand = res["c"]
    out["ms"] = tq * 1e3
    out["queries_per_s"] = n / tq
    out["candidates"] = int(cand.numel())
    t_count = prof("bvh_find_count")
    bvh.enableStats(True)
    bvh.findPoints(off, cnt, n, p)
    vbar = bvh.lastStats() / n
    bvh.enableStats(False)
    bpq = 24 + 8 + 4.0 * cand.numel() / n + 104.0 * vbar
    ach = bpq * n / (t_count / 1e3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": ach / HBM_PEAK_GBS, "kernel": "k_find<points, count>",
                       "avg_launch_ms": t_count, "bytes_per_query": bpq,
                       "mean_inner_visits": vbar,
                       "other_kernels_ms": {"scan": prof("bvh_find_scan"),
                                            "copy": prof("bvh_find_copy"),
                                            "fill": prof("bvh_find_fill")},
                       "fill_launches_per_call": prof.launches("bvh_find_fill") / calls}
    bvh32 = spin.BVH(bx, fp_mode="fp32")
    off32, cnt32 = torch.empty_like(off), torch.empty_like(cnt)
    tq32 = sync_time(lambda: res.__setitem__("c32", bvh32.findPoints(off32, cnt32, n, p)), reps)
    out["fp32_traversal"] = {"ms": tq32 * 1e3, "queries_per_s": n / tq32,
                             "lists_identical_to_fp64": bool(
                                 torch.equal(res["c32"], cand) and torch.equal(off32, off) and
                                 torch.equal(cnt32, cnt))}
    if cpu_threads:
        from oracle import oracle as orc
        t0 = time.perf_counter()
        ob = orc.OracleBVH(boxes)
        el_b = time.perf_counter() - t0
        same_tree = bool(np.array_equal(ob.export()["nodes"], bvh.export()["nodes"].cpu().numpy()))
        t0 = time.perf_counter()
        ooff, ocnt, ocand = ob.find("points", pts, nthreads=cpu_threads)
        el_f = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "build_ms_1_thread": el_b * 1e3, "build_mprims_per_s": boxes.shape[0] / el_b / 1e6,
            "tree_matches_gpu_bitwise": same_tree,
            "ms": el_f * 1e3, "queries_per_s": n / el_f, "cores": cpu_threads, "kind": "port",
            "sample": f"the same {boxes.shape[0]:,} boxes (build, one thread) and {n:,} points",
            "lists_match_gpu_bitwise": bool(np.array_equal(ocnt, cnt.cpu().numpy()) and
                                            np.array_equal(ocand, cand.cpu().numpy()))}
    return out


LABEL3 = ("torus R=1, r=0.3+0.02 sin(7u) cos(5v), 1024x1024 quads = 2,097,152 tris; 16M "
          "queries (half uniform in the box padded by 0.2, half |d|<0.01 of the surface); "
          "SignedDistance forward, FP64 and FP32 traversal")
LABEL4 = ("icosphere L7 (327,680 tris), radius 1+0.1 sin(4 theta) cos(3 phi); 1M queries at "
          "normal offsets U[-0.1,0.1], upstream gradient N(0,1) FP32; forward + adjoint")
LABEL5 = ("icosphere L8 (1,310,720 tris) fitted to a rounded box (b=(0.8,0.6,0.5), r=0.1); 8M "
          "queries U[-1.2,1.2]^3; per iteration: BVH rebuild + FP32-traversal forward + "
          "L=mean((d-t)^2) + adjoint + V -= 1e-3 dL/dV")


def other_workload(args, torch, dev):
    """One workload on its own line (see axom_amd/synth.py for the inputs)."""
    from axom_amd import _lib, quest, spin, synth
    L = _lib.lib()
    sync_time = sync_timer(torch)
    cpu = 0 if args.no_cpu else host_cpus()[0]

    if args.workload == "find_points":
        boxes, pts = synth.workload1()
        return find_points_config(torch, dev, L, _lib, boxes, pts, cpu, reps=args.steps)
    if args.workload == "torus_mixed":
        v, t, q = synth.workload3()
        return sdf_config(torch, dev, L, _lib, v, t, q, None, args.steps, False, cpu, 100_000,
                          LABEL3)
    if args.workload == "diff_l7":
        v, t, q, g = synth.workload4()
        return sdf_config(torch, dev, L, _lib, v, t, q, g, args.steps, True, cpu, 100_000, LABEL4)
    if args.workload == "shapefit_l8":
        v, t, q, target = synth.workload5()
        return shapefit_config(torch, dev, v, t, q, target, cpu, LABEL5)
    if args.workload == "self_isect_l8":
        return self_isect_workload(args, torch, dev, sync_time)
    if args.workload == "find_points_2d":
        rng = np.random.default_rng(20)
        c = rng.random((100_000, 2))
        h = rng.uniform(0.001, 0.01, (100_000, 2))
        boxes = np.concatenate([c - h, c + h], 1)
        pts = rng.random((1_000_000, 2))
        bx, p = torch.as_tensor(boxes, device=dev), torch.as_tensor(pts, device=dev)
        bvh = spin.BVH(bx, NDIMS=2)
        off = torch.empty(p.shape[0], dtype=torch.int32, device=dev)
        cnt = torch.empty(p.shape[0], dtype=torch.int32, device=dev)
        tb = sync_time(lambda: bvh.initialize(bx), args.steps)
        tq = sync_time(lambda: bvh.findPoints(off, cnt, p.shape[0], p), args.steps)
        return {"workload": "2-D: 100,000 boxes U[0,1]^2 (half-extents U[0.001,0.01]), "
                            "1,000,000 points, FP64 build + findPoints (spin::BVH<2>)",
                "build_mprims_per_s": boxes.shape[0] / tb / 1e6, "build_ms": tb * 1e3,
                "find_points_queries_per_s": p.shape[0] / tq, "find_ms": tq * 1e3}
    if args.workload == "hex_candidates":
        iv, ih = synth.hex_grid(64, jitter=0.15, seed=61)
        qv, qh = synth.hex_grid(60, origin=(0.7, 0.3, 0.5), h=64 / 60 * 1.01, jitter=0.1,
                                seed=62, angle=0.5)
        ivd, ihd, qvd, qhd = (torch.as_tensor(a, device=dev) for a in (iv, ih, qv, qh))
        pairs = quest.findCandidatesBVH(ivd, ihd, qvd, qhd)
        tq = sync_time(lambda: quest.findCandidatesBVH(ivd, ihd, qvd, qhd), args.steps)
        out = {"workload": f"hex candidate search: insert 64^3 = {ih.shape[0]:,} hexes "
                           f"(jittered grid), query 60^3 = {qh.shape[0]:,} hexes (jittered, "
                           "turned 0.5 rad); cell boxes + BVH + findBoundingBoxes + pairs",
               "pairs": int(pairs.shape[0]), "ms": tq * 1e3,
               "query_cells_per_s": qh.shape[0] / tq}
        if not args.no_cpu:
            from oracle import oracle as orc
            nthreads = cpu
            m = 20_000
            t0 = time.perf_counter()
            want = orc.mesh_candidates(iv, ih, qv, qh[:m], nthreads=nthreads)
            tc = time.perf_counter() - t0
            got = pairs.cpu().numpy()
            out["cpu_baseline"] = {"query_cells_per_s": m / tc, "cores": nthreads, "kind": "port",
                                   "sample": f"first {m:,} query cells (with the full insert "
                                             "mesh's BVH build)",
                                   "matches_gpu_bitwise": bool(np.array_equal(
                                       got[got[:, 0] < m], want))}
        return out
    if args.workload == "sdf_quads":
        v, quads = synth.cube_sphere(128)
        conn = quads.reshape(-1)
        off = np.arange(0, conn.size + 1, 4, dtype=np.int32)
        q = synth.uniform(synth.SEED0 + 7, 0, 3 * args.queries, -1.5, 1.5).reshape(-1, 3)
        vd, cd, od, qd = (torch.as_tensor(a, device=dev) for a in (v, conn, off, q))
        sdf = quest.SignedDistance(vd, cd, cell_offsets=od)
        n = q.shape[0]
        d = torch.empty(n, dtype=torch.float64, device=dev)
        tq = sync_time(lambda: sdf.computeDistances(n, qd, outSgnDist=d), max(1, args.steps // 2))
        out = {"workload": f"quad sphere (cube subdivided 128 x 128 per face, projected: "
                           f"{quads.shape[0]:,} quads), {n:,} queries U[-1.5,1.5]^3; "
                           "SignedDistance forward over quad cells",
               "quads": int(quads.shape[0]), "queries_per_s": n / tq, "ms": tq * 1e3}
        if not args.no_cpu:
            from oracle import oracle as orc
            m = min(n, 200_000)
            nthreads = cpu
            o = orc.OracleSDF(v, conn, cell_offsets=off)
            t0 = time.perf_counter()
            r = o.eval(q[:m], nthreads=nthreads)
            tc = time.perf_counter() - t0
            out["cpu_baseline"] = {"queries_per_s": m / tc, "cores": nthreads, "kind": "port",
                                   "sample": f"first {m:,} queries",
                                   "matches_gpu_bitwise": bool(np.array_equal(
                                       r["dist"], d[:m].cpu().numpy()))}
        return out
    raise ValueError(f"unknown workload {args.workload}")


def self_isect_workload(args, torch, dev, sync_time):
    """quest::findTriMeshIntersectionsBVH on two crossing noisy L8 icospheres
    (2,621,440 triangles): corner boxes, BVH build, findBoundingBoxes of every
    box, triangle-triangle tests on the pairs i < j, ordered output. The
    oracle runs the same test on this host for the CPU figure and the
    pair-by-pair comparison."""
    from axom_amd import _lib, quest, synth
    v, t = synth.workload6()
    vd, td = torch.as_tensor(v, device=dev), torch.as_tensor(t, device=dev)
    L = _lib.lib()
    res = {}

    def run():
        res["out"] = quest.findTriMeshIntersectionsBVH(vd, td)
    L.axsq_prof_reset()
    L.axsq_prof_enable(1)
    reps = max(1, args.steps)
    ts = sync_time(run, reps)
    L.axsq_prof_enable(0)

    def prof(name):
        n, ms = ctypes.c_int64(), ctypes.c_double()
        _lib.check(L.axsq_prof_get(name.encode(), ctypes.byref(n), ctypes.byref(ms)))
        return ms.value / max(n.value, 1)
    pairs, degen = res["out"]
    out = {"workload": "two noisy icospheres L8 (2 x 1,310,720 tris, vertex noise N(0, (0.25 "
                       "edge)^2), second copy shifted by (0.5, 0.1, 0)); "
                       "findTriMeshIntersectionsBVH, threshold 1e-8",
           "triangles": int(t.shape[0]), "pairs": int(pairs.shape[0]),
           "degenerate": int(degen.shape[0]), "ms": ts * 1e3,
           "triangles_per_s": t.shape[0] / ts,
           "kernels_ms": {k: prof(k) for k in ("mesh_isect_setup", "bvh_build", "mesh_isect_walk",
                                                "mesh_isect_emit")}}
    if not args.no_cpu:
        from oracle import oracle as orc
        nthreads = host_cpus()[0]
        t0 = time.perf_counter()
        want_p, want_d = orc.mesh_self_intersections(v, t, 1e-8, nthreads=nthreads)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"seconds": tc, "triangles_per_s": t.shape[0] / tc,
                               "cores": nthreads, "kind": "port",
                               "matches_gpu_pairwise": bool(
                                   np.array_equal(pairs.cpu().numpy(), want_p) and
                                   np.array_equal(degen.cpu().numpy(), want_d))}
    return out


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported: "
              "splitting the queries across GPUs (tree and mesh replicated on each) is "
              "not implemented", file=sys.stderr)
        return 2

    import torch

    from axom_amd import _lib, quest, spin, synth

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    L = _lib.lib()
    if args.workload != "sdf_l6":
        print(json.dumps(other_workload(args, torch, dev)), flush=True)
        return 0
    cpu, host = host_cpus()
    if args.no_cpu:
        cpu = 0

    log("generating inputs")
    v, t, q = synth.workload2(nq=args.queries)
    g = synth.normal_f32(synth.SEED0 + 2, 40, q.shape[0])
    nq = q.shape[0]
    vd = torch.as_tensor(v, device=dev)
    td = torch.as_tensor(t, device=dev)
    qd = torch.as_tensor(q, device=dev)
    gd = torch.as_tensor(g, device=dev)
    dist = torch.empty(nq, dtype=torch.float64, device=dev)
    cp = torch.empty((nq, 3), dtype=torch.float64, device=dev)
    elem = torch.empty(nq, dtype=torch.int32, device=dev)

    sdf = quest.SignedDistance(vd, td, isWatertight=True, computeSign=True, fp_mode="fp32")

    def step(mode):
        sdf.computeDistances(nq, qd, outSgnDist=dist, outClosestPts=cp, outElem=elem,
                             fp_mode=mode)
        return sdf.backward(qd, dist, elem, gd)

    def timed(mode, steps, warmup):
        for _ in range(warmup):
            step(mode)
        torch.cuda.synchronize()
        L.axsq_prof_reset()
        L.axsq_prof_enable(1)
        t0 = time.perf_counter()
        for _ in range(steps):
            step(mode)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        L.axsq_prof_enable(0)
        return el

    def prof(name):
        n, ms = ctypes.c_int64(), ctypes.c_double()
        _lib.check(L.axsq_prof_get(name.encode(), ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    log("timing FP64 forward+backward")
    el64 = timed("fp64", args.steps, args.warmup)
    n_eval, ms_eval = prof("sdf_eval_fp64")
    n_p1, ms_p1 = prof("sdf_eval_fp64_pass1")
    n_p2, ms_p2 = prof("sdf_eval_fp64_pass2")
    n_bwd, ms_bwd = prof("sdf_backward")
    n_ord, ms_ord = prof("sdf_query_order")
    d64 = dist.clone()
    qps = nq * args.steps / el64
    out = {
        "metric": METRIC, "value": qps, "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1000.0 * el64 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "icosphere level 6 (40,962 verts, 81,920 tris, r=1, +-1e-3 "
                               "radial noise), %d queries U[-1.5,1.5]^3; plain FP64 "
                               "SignedDistance forward (dist, cp, elem) + adjoint" % nq,
                   "queries": nq, "triangles": int(t.shape[0]), "parallelism": "single GPU"},
        "kernels_ms_per_step": {"sdf_eval_fp64": ms_eval / max(n_eval, 1),
                                "sdf_eval_fp64_pass1": ms_p1 / max(n_p1, 1),
                                "sdf_eval_fp64_pass2": ms_p2 / max(n_p2, 1),
                                "sdf_query_order": ms_ord / max(n_ord, 1),
                                "sdf_backward": ms_bwd / max(n_bwd, 1)},
        "second_pass_walks": sdf.lastSuspended(),
        "host": host,
    }

    # roofline of the dominant kernel: visits / leaf checks of the traversal
    sdf.enableStats(True)
    sdf.computeDistances(nq, qd, outSgnDist=dist, outClosestPts=cp, outElem=elem, fp_mode="fp64")
    visits, leaves = sdf.lastStats()
    sdf.enableStats(False)
    vbar, lbar = visits / nq, leaves / nq
    bytes_per_query = 24 + 8 + 24 + 4 + 104.0 * vbar + 84.0 * lbar
    t_eval = ms_eval / max(n_eval, 1) / 1000.0
    achieved = bytes_per_query * nq / t_eval / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        traffic = json.load(open(PROFILE_TRAFFIC)).get("hbm_bytes_per_launch")
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                       "kernel": "k_sdf_eval<FP64>", "avg_launch_ms": t_eval * 1e3,
                       "bytes_per_query": bytes_per_query, "mean_inner_visits": vbar,
                       "mean_leaf_checks": lbar}

    if not args.no_extras:
        log("timing FP32-traversal forward+backward")
        el32 = timed("fp32", args.steps, args.warmup)
        n32, ms32 = prof("sdf_eval_fp32")
        d32 = dist.clone()
        rel = (d32 - d64).abs() / d64.abs().clamp_min(1e-300)
        out["fp32_traversal"] = {
            "value": nq * args.steps / el32, "unit": "queries/s",
            "ms_per_step": 1000.0 * el32 / args.steps,
            "sdf_eval_fp32_ms": ms32 / max(n32, 1),
            "max_rel_err_vs_fp64": float(rel.max()),
            "rms_rel_err_vs_fp64": float(torch.sqrt((rel * rel).mean())),
            "sign_mismatches": int((torch.signbit(d32) != torch.signbit(d64)).sum())}

        log("timing BVH build over the 81,920 triangles' boxes")
        tb = torch.empty((t.shape[0], 6), dtype=torch.float64, device=dev)
        a, b, c = vd[td[:, 0].long()], vd[td[:, 1].long()], vd[td[:, 2].long()]
        tb[:, :3] = torch.minimum(torch.minimum(a, b), c)
        tb[:, 3:] = torch.maximum(torch.maximum(a, b), c)
        bvh = spin.BVH(tb)
        el = sync_timer(torch)(lambda: bvh.initialize(tb), 20)
        out["bvh_build"] = {"icosphere6_tris": {"prims": int(tb.shape[0]), "ms": el * 1e3,
                                                "mprims_per_s": tb.shape[0] / el / 1e6}}
        log("timing BVH build and findPoints (100,000 boxes, 1,000,000 points)")
        boxes1, pts1 = synth.workload1()
        fp = find_points_config(torch, dev, L, _lib, boxes1, pts1, cpu)
        out["bvh_build"]["boxes_100k"] = {"prims": int(boxes1.shape[0]), "ms": fp["build_ms"],
                                          "mprims_per_s": fp["build_mprims_per_s"]}
        out["find_points"] = fp
        # the remaining configurations at their full sizes
        cfg = {}
        log("torus: 2,097,152 triangles, 16M queries")
        v3, t3, q3 = synth.workload3()
        cfg["torus_mixed_2M_tris_16M_queries"] = sdf_config(
            torch, dev, L, _lib, v3, t3, q3, None, 3, False, cpu, 100_000, LABEL3)
        del v3, t3, q3
        log("icosphere L7: 327,680 triangles, 1M queries, forward + adjoint")
        v4, t4, q4, g4 = synth.workload4()
        cfg["differentiable_l7_1M_queries"] = sdf_config(
            torch, dev, L, _lib, v4, t4, q4, g4, 5, True, cpu, 100_000, LABEL4)
        del v4, t4, q4, g4
        log("shape fit: icosphere L8, 8M queries, 20 iterations")
        v5, t5, q5, tg5 = synth.workload5()
        cfg["shape_fit_l8_8M_queries_20_iterations"] = shapefit_config(
            torch, dev, v5, t5, q5, tg5, cpu, LABEL5)
        del v5, t5, q5, tg5
        out["configs"] = cfg
        torch.cuda.empty_cache()

    if args.calib_copy:
        src = torch.empty(1 << 27, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["calib_copy_bytes"] = {"read": src.numel() * 8, "write": dst.numel() * 8}

    if not args.no_cpu:
        from oracle import oracle as orc
        ncpu = cpu
        ns = min(args.cpu_sample, nq)
        log(f"CPU baseline: oracle on {ns} queries, {ncpu} threads")
        o = orc.OracleSDF(v, t)
        t0 = time.perf_counter()
        r = o.eval(q[:ns], nthreads=ncpu, want_normal=False)
        o.backward(q[:ns], r["dist"], r["elem"], g[:ns])
        el = time.perf_counter() - t0
        same = bool(np.array_equal(r["dist"], d64[:ns].cpu().numpy()))
        out["cpu_baseline"] = {"value": ns / el, "unit": "queries/s", "cores": ncpu,
                               "kind": "port",
                               "sample": f"first {ns} of the {nq} queries, forward (dist, cp, "
                                         "elem) + adjoint, FP64",
                               "seconds": el, "matches_gpu_bitwise": same}
        # the same on one thread (the reference's SEQ execution), on a smaller sample
        n1 = min(100_000, ns)
        t0 = time.perf_counter()
        r1 = o.eval(q[:n1], nthreads=1, want_normal=False)
        o.backward(q[:n1], r1["dist"], r1["elem"], g[:n1])
        el1 = time.perf_counter() - t0
        out["cpu_baseline"]["single_thread"] = {
            "value": n1 / el1, "unit": "queries/s", "cores": 1,
            "sample": f"first {n1} queries, forward + adjoint, FP64", "seconds": el1}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
