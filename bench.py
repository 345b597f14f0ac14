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
  bvh_build        BVH build Mprims/s over the 81,920 triangles and over the
                   100,000-box findPoints workload
  find_points      findPoints queries/s (100,000 boxes, 1,000,000 points)
  roofline         dominant kernel (FP64 SDF forward) timed with HIP events on
                   its stream; algorithmic bytes per query
                   = 24 (query) + 8 (distance) + 24 (closest point) + 4 (element)
                   + V*104 (inner-node visits: 2 x 48 B boxes + 2 x 4 B links)
                   + L*84 (leaf checks: 3 x 24 B vertices + 3 x 4 B indices),
                   V, L = mean visits / leaf checks of the reference traversal
                   (the FP64 kernel follows it exactly; counted on the GPU)
  cpu_baseline     the CPU oracle (restatement of the reference's SEQ path),
                   forward+backward on the first 1,000,000 queries, timed on
                   this host's cores (and on one thread for the first 100,000)
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


def other_workload(args, torch, dev):
    """The remaining synthetic workloads (see axom_amd/synth.py for their inputs)."""
    from axom_amd import autograd, quest, shapefit, spin, synth

    def sync_time(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    if args.workload == "find_points":
        boxes, pts = synth.workload1()
        bx, p = torch.as_tensor(boxes, device=dev), torch.as_tensor(pts, device=dev)
        bvh = spin.BVH(bx)
        off = torch.empty(p.shape[0], dtype=torch.int32, device=dev)
        cnt = torch.empty(p.shape[0], dtype=torch.int32, device=dev)
        tb = sync_time(lambda: bvh.initialize(bx), args.steps)
        tq = sync_time(lambda: bvh.findPoints(off, cnt, p.shape[0], p), args.steps)
        c64 = bvh.findPoints(off, cnt, p.shape[0], p)
        bvh32 = spin.BVH(bx, fp_mode="fp32")
        off32, cnt32 = torch.empty_like(off), torch.empty_like(cnt)
        tb32 = sync_time(lambda: bvh32.initialize(bx), args.steps)
        tq32 = sync_time(lambda: bvh32.findPoints(off32, cnt32, p.shape[0], p), args.steps)
        c32 = bvh32.findPoints(off32, cnt32, p.shape[0], p)
        same = bool(torch.equal(c32, c64) and torch.equal(off32, off) and torch.equal(cnt32, cnt))
        return {"workload": "100,000 boxes U[0,1]^3 (half-extents U[0.001,0.01]), 1,000,000 "
                            "points, FP64 build + findPoints",
                "build_mprims_per_s": boxes.shape[0] / tb / 1e6, "build_ms": tb * 1e3,
                "find_points_queries_per_s": p.shape[0] / tq, "find_ms": tq * 1e3,
                "fp32_traversal": {"build_ms": tb32 * 1e3, "find_ms": tq32 * 1e3,
                                   "lists_identical_to_fp64": same}}
    if args.workload == "torus_mixed":
        v, t, q = synth.workload3()
        vd, td, qd = (torch.as_tensor(a, device=dev) for a in (v, t, q))
        sdf = quest.SignedDistance(vd, td, fp_mode="fp32")
        n = q.shape[0]
        d32 = torch.empty(n, dtype=torch.float64, device=dev)
        d64 = torch.empty(n, dtype=torch.float64, device=dev)
        t32 = sync_time(lambda: sdf.computeDistances(n, qd, outSgnDist=d32, fp_mode="fp32"),
                        args.steps)
        t64 = sync_time(lambda: sdf.computeDistances(n, qd, outSgnDist=d64, fp_mode="fp64"),
                        max(1, args.steps // 2))
        rel = (d32 - d64).abs() / d64.abs().clamp_min(1e-300)
        return {"workload": "torus R=1, r=0.3+0.02 sin(7u) cos(5v), 2,097,152 tris; 16M queries "
                            "(half box, half |d|<0.01); SDF forward",
                "fp32_traversal_queries_per_s": n / t32, "fp64_queries_per_s": n / t64,
                "max_rel_err": float(rel.max()), "rms_rel_err": float(torch.sqrt((rel * rel).mean())),
                "sign_mismatches": int((torch.signbit(d32) != torch.signbit(d64)).sum())}
    if args.workload == "diff_l7":
        v, t, q, g = synth.workload4()
        vd = torch.as_tensor(v, device=dev).requires_grad_(True)
        qd = torch.as_tensor(q, device=dev).requires_grad_(True)
        td = torch.as_tensor(t, device=dev)
        gd = torch.as_tensor(g, device=dev, dtype=torch.float64)
        out = {"workload": "icosphere L7 (327,680 tris), radius 1+0.1 sin(4 theta) cos(3 phi); 1M "
                           "queries at normal offsets U[-0.1,0.1]; SDF forward + adjoint"}
        for mode in ("fp64", "fp32"):
            sdf = quest.SignedDistance(vd.detach(), td, fp_mode="fp32")

            def fb():
                qd.grad = None
                vd.grad = None
                autograd.signed_distance(qd, vd, sdf, mode).backward(gd)
            tt = sync_time(fb, args.steps)
            out[f"{mode}_fwd_bwd_queries_per_s"] = q.shape[0] / tt
        return out
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
            nthreads = min(16, os.cpu_count() or 1)
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
            nthreads = min(16, os.cpu_count() or 1)
            o = orc.OracleSDF(v, conn, cell_offsets=off)
            t0 = time.perf_counter()
            r = o.eval(q[:m], nthreads=nthreads)
            tc = time.perf_counter() - t0
            out["cpu_baseline"] = {"queries_per_s": m / tc, "cores": nthreads, "kind": "port",
                                   "sample": f"first {m:,} queries",
                                   "matches_gpu_bitwise": bool(np.array_equal(
                                       r["dist"], d[:m].cpu().numpy()))}
        return out
    v, t, q, target = synth.workload5()
    vd, td, qd, tg = (torch.as_tensor(a, device=dev) for a in (v, t, q, target))
    _, hist = shapefit.shape_fit(vd, td, qd, tg, iterations=20, lr=1e-3, fp_mode="fp32")
    return {"workload": "icosphere L8 (1,310,720 tris) fitted to a rounded box (b=(0.8,0.6,0.5), "
                        "r=0.1); 8M queries U[-1.2,1.2]^3; 20 iterations of rebuild + FP32-traversal "
                        "forward + loss ||d - target||^2 + adjoint + vertex step of length lr=1e-3",
            "iterations": hist,
            "rms_residual_first_last": [hist[0]["rms_residual"], hist[-1]["rms_residual"]],
            "mean_build_ms": sum(h["build_ms"] for h in hist[1:]) / 19,
            "mean_forward_ms": sum(h["forward_ms"] for h in hist) / 20,
            "mean_backward_ms": sum(h["backward_ms"] for h in hist) / 20}


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
        nthreads = min(16, os.cpu_count() or 1)
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

        log("timing BVH builds and findPoints")
        tb = torch.empty((t.shape[0], 6), dtype=torch.float64, device=dev)
        a, b, c = vd[td[:, 0].long()], vd[td[:, 1].long()], vd[td[:, 2].long()]
        tb[:, :3] = torch.minimum(torch.minimum(a, b), c)
        tb[:, 3:] = torch.maximum(torch.maximum(a, b), c)
        boxes1, pts1 = synth.workload1()
        bx1 = torch.as_tensor(boxes1, device=dev)
        builds = {}
        for name, bx in (("icosphere6_tris", tb), ("boxes_100k", bx1)):
            bvh = spin.BVH(bx)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = 20
            for _ in range(reps):
                bvh.initialize(bx)
            torch.cuda.synchronize()
            el = (time.perf_counter() - t0) / reps
            builds[name] = {"prims": int(bx.shape[0]), "ms": el * 1e3,
                            "mprims_per_s": bx.shape[0] / el / 1e6}
        out["bvh_build"] = builds
        bvh = spin.BVH(bx1)
        p1 = torch.as_tensor(pts1, device=dev)
        off = torch.empty(p1.shape[0], dtype=torch.int32, device=dev)
        cnt = torch.empty(p1.shape[0], dtype=torch.int32, device=dev)
        bvh.findPoints(off, cnt, p1.shape[0], p1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            cand = bvh.findPoints(off, cnt, p1.shape[0], p1)
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / 10
        out["find_points"] = {"queries": int(p1.shape[0]), "boxes": int(bx1.shape[0]),
                              "ms": el * 1e3, "queries_per_s": p1.shape[0] / el,
                              "candidates": int(cand.numel())}
        if not args.no_cpu:
            # the oracle's build, one thread (the reference's SEQ build), and
            # its findPoints over all host threads, on the same boxes and points
            from oracle import oracle as orc
            ncpu = min(os.cpu_count() or 1, 16)
            t0 = time.perf_counter()
            ob = orc.OracleBVH(boxes1)
            el_b = time.perf_counter() - t0
            ex = bvh.export()
            same_tree = bool(np.array_equal(ob.export()["nodes"], ex["nodes"].cpu().numpy()))
            t0 = time.perf_counter()
            ooff, ocnt, ocand = ob.find("points", pts1, nthreads=ncpu)
            el_f = time.perf_counter() - t0
            same_find = bool(np.array_equal(ocnt, cnt.cpu().numpy()) and
                             np.array_equal(ocand, cand.cpu().numpy()))
            out["bvh_build"]["boxes_100k"]["cpu_baseline"] = {
                "ms": el_b * 1e3, "mprims_per_s": boxes1.shape[0] / el_b / 1e6, "cores": 1,
                "kind": "port", "sample": "the same 100,000 boxes",
                "matches_gpu_bitwise": same_tree}
            out["find_points"]["cpu_baseline"] = {
                "ms": el_f * 1e3, "queries_per_s": pts1.shape[0] / el_f, "cores": ncpu,
                "kind": "port", "sample": "the same 1,000,000 points",
                "matches_gpu_bitwise": same_find}

    if args.calib_copy:
        src = torch.empty(1 << 27, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["calib_copy_bytes"] = {"read": src.numel() * 8, "write": dst.numel() * 8}

    if not args.no_cpu:
        from oracle import oracle as orc
        ncpu = min(os.cpu_count() or 1, 16)
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
