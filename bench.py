"""VF-Ray throughput benchmark on one MI355X.

Default workload (`batch64`): 128^3 hex grid x 6 = 12,582,912 tets, 64 seeded
perspective views on the r = 2.5 sphere with sides drawn from
{256, 384, 512, 768, 1024}, front-to-back compositing with early termination
at alpha 0.99.  One step = rendering all 64 views.  `value` is the batched
renderer's Mrays/s (sum of W*H over the views / wall time of the step, inputs
resident on the device); the plain per-view renderer (one launch set and one
frame sync per view) is timed on the same views and reported beside it, with
its per-frame p50/p99.  The same run then measures the request stream
(`served`, see serve_measure): plain and batched capacity, the highest
Poisson rate whose p99 latency meets one 60 Hz frame, and a 10,000-request
stream's p50/p99.  `roofline` is the walk kernel's HBM traffic (rocprofv3
counters of the same launch, committed under profiles/) over its launch time
against the 8 TB/s peak.

`plain_gpu`: 64^3 grid (1,572,864 tets), one 1024^2 perspective view.

`concave`: 96^3 grid with a graded radial map and 12 voids (rays exit and
re-enter), 32 views of 768^2.

`serve`: the request-stream measurement alone; `value` is the served Mrays/s
at the latency bound.

`serve_capacity`: the serving workload's saturated replay alone (for kernel
traces); prints its own JSON line (Mrays/s, batches, makespan).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
B_STEP, B_RAY = 96, 48  # model bytes per tet step / per active ray (reference layout, cache-served included)
METRIC = "Mrays/s + p50/p99 frame latency, batched/served vs plain GPU at PSNR\u226550dB"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(x, t, tfa, views, seconds: float) -> dict:
    """The CPU oracle's front-to-back port (OpenMP) on the first views of the
    same workload until `seconds` of work; Mrays/s over the rays it did."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import vfray_amd as V
    xn = O.normalize(x)
    con, _ = O.connectivity(t, x.shape[0])
    ext = O.extf(con)
    lmt = O.lmt(xn, t)
    psi = V.psi_generate()

    def run(budget):
        rays, el, done = 0, 0.0, 0
        for v in views:
            proj, M, k, kx = V.view_matrix(v)
            t0 = time.perf_counter()
            O.render_ftb(xn, t, con, ext, tfa, psi, lmt, proj, M, k, v.width, v.height, kx=kx)
            el += time.perf_counter() - t0
            rays += v.width * v.height
            done += 1
            if el >= budget:
                break
        return rays, el, done

    threads = O.set_threads(0)
    rays, el, done = run(seconds)
    O.set_threads(1)
    rays1, el1, done1 = run(seconds)
    O.set_threads(threads)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": rays / el / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"first {done} of {len(views)} views ({rays} rays), CPU oracle front-to-back "
                      f"restatement, FP32, OpenMP {threads} threads, {el:.1f} s",
            "single_core": {"value": rays1 / el1 / 1e6, "unit": "Mrays/s",
                            "sample": f"first {done1} views ({rays1} rays), 1 thread, {el1:.1f} s"},
            "host": {"cpu_model": model, "nproc": os.cpu_count()}}


def measured_traffic(kernel: str, workload: str):
    """HBM/fabric bytes per traversal launch from the committed rocprofv3
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by what the counter reads
    for a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None
    txt = open(path).read()
    i = txt.rfind("```json")
    if i < 0:
        return None
    blob = json.loads(txt[i + 7: txt.index("```", i + 7)])
    return blob


SLO_P99_MS = 16.7   # served latency bound: one 60 Hz frame
SLO_LOADS = (1.0, 0.95, 0.9, 0.8, 0.7, 0.6, 0.5)


def serve_measure(ctx, args) -> dict:
    """The request-stream workload: 8 resident meshes (32^3 .. 160^3 grids),
    4 TFs each, requests with a random mesh, TF, viewpoint on the r = 2.5
    sphere and side in {256..1024}.

    * plain: the same server with one slot and one request per batch (one
      launch set and one completion per request, back to back), all requests
      queued at once: its throughput is the plain renderer's capacity;
    * capacity: the batched server saturated (all requests queued at once);
    * at_slo: the highest offered Poisson rate (fractions of the capacity,
      descending) whose p99 latency (arrival -> result on the device) meets
      SLO_P99_MS, with the measured Mrays/s over that stream;
    * stream: `requests` requests at `load` x capacity, p50/p99.
    Mrays/s is always measured: rays of the requests / (first arrival -> last
    completion)."""
    import vfray_amd as V
    from vfray_amd import workloads as WL

    t0 = time.time()
    meshes, tf_ids, tets = [], [], 0
    for (x, t, tfs) in WL.serve_meshes():
        meshes.append(ctx.mesh(x, t))
        tf_ids.append([ctx.tf(tf) for tf in tfs])
        tets += t.shape[0]
    log(f"8 resident meshes, {tets} tets, built in {time.time() - t0:.1f}s")
    sched = dict(persistent=args.persistent, wave_queue=args.wave_queue, occupancy=args.occupancy,
                 waves_per_cu=args.wpc, refill=args.refill or 32)
    reqs = WL.serve_requests(args.requests, len(meshes), tf_ids)
    rays_of = lambda rs: sum(r.view.width * r.view.height for r in rs)
    n_cap = min(len(reqs), max(200, args.requests // 5))
    sample = reqs[:n_cap]
    sp = V.Server(ctx, meshes, hbm_budget=8 << 30, streams=1, max_batch=1, **sched)
    sp.replay(sample[:20], np.zeros(20))
    plain = sp.replay(sample, np.zeros(len(sample)))
    sp.close()
    log(f"plain capacity {plain['mrays_per_s']:.1f} Mrays/s over {len(sample)} requests")
    sb = V.Server(ctx, meshes, hbm_budget=32 << 30, streams=args.streams, max_batch=256,
                  max_batch_pixels=args.batch_pixels, schedule=args.serve_schedule, **sched)
    sb.replay(sample[:64], np.zeros(64))
    cap = sb.replay(sample, np.zeros(len(sample)))
    log(f"batched capacity {cap['mrays_per_s']:.1f} Mrays/s, mean batch {cap['mean_batch']:.1f}")
    mean_rays = rays_of(reqs) / len(reqs)
    n_slo = min(len(reqs), 2000)
    at_slo, tried = None, []
    for load in SLO_LOADS:
        rate = load * cap["mrays_per_s"] * 1e6 / mean_rays
        r = sb.replay(reqs[:n_slo], WL.poisson_offsets(n_slo, rate))
        tried.append({"load": load, "offered_req_per_s": round(rate, 1), "mrays_per_s": round(r["mrays_per_s"], 1),
                      "p99_ms": round(r["latency_p99_ms"], 3)})
        log(f"load {load:.2f}: {r['mrays_per_s']:.1f} Mrays/s, p99 {r['latency_p99_ms']:.2f} ms")
        if r["latency_p99_ms"] <= SLO_P99_MS:
            at_slo = dict(r, load=load, offered_req_per_s=rate,
                          offered_mrays_per_s=rate * mean_rays / 1e6)
            break
    rate = args.load * cap["mrays_per_s"] * 1e6 / mean_rays
    stream = sb.replay(reqs, WL.poisson_offsets(len(reqs), rate))
    sb.close()
    log(f"stream of {len(reqs)} at {rate:.0f} req/s: {stream['mrays_per_s']:.1f} Mrays/s, "
        f"p50 {stream['latency_p50_ms']:.2f} ms p99 {stream['latency_p99_ms']:.2f} ms")
    for m in meshes:
        m.close()
    keep = ("mrays_per_s", "latency_p50_ms", "latency_p90_ms", "latency_p99_ms", "latency_max_ms", "mean_batch",
            "requests", "batches", "makespan_s", "anomalies")
    small = lambda d: {k: (round(d[k], 3) if isinstance(d[k], float) else d[k]) for k in keep if k in d}
    out = {
        "workload": "serve: 8 resident meshes 32^3..160^3 x 6 tets, 4 TFs each, sides 256..1024, "
                    "perspective, front to back, alpha 0.99",
        "tets": int(tets),
        "plain": small(plain),
        "capacity": small(cap),
        "capacity_vs_plain": round(cap["mrays_per_s"] / plain["mrays_per_s"], 3),
        "slo_p99_ms": SLO_P99_MS,
        "at_slo": None if at_slo is None else dict(small(at_slo), load=at_slo["load"],
                                                    offered_mrays_per_s=round(at_slo["offered_mrays_per_s"], 1)),
        "at_slo_vs_plain_capacity": None if at_slo is None else round(at_slo["mrays_per_s"] / plain["mrays_per_s"], 3),
        "slo_search": tried,
        "stream": dict(small(stream), load=args.load, offered_req_per_s=round(rate, 1)),
        "streams": args.streams, "max_batch_pixels": args.batch_pixels, "schedule": sched,
        "serve_schedule": ["slots", "pipeline"][args.serve_schedule],
    }
    return out


def serve_capacity(args) -> int:
    """The serving workload's saturated replay alone (all requests queued at
    once, the batched server of serve_measure), for kernel traces: one JSON
    line with its Mrays/s, batches and makespan."""
    import vfray_amd as V
    from vfray_amd import workloads as WL
    ctx = V.Context(0)
    meshes, tf_ids = [], []
    for (x, t, tfs) in WL.serve_meshes():
        meshes.append(ctx.mesh(x, t))
        tf_ids.append([ctx.tf(tf) for tf in tfs])
    reqs = WL.serve_requests(args.requests, len(meshes), tf_ids)
    sb = V.Server(ctx, meshes, hbm_budget=32 << 30, streams=args.streams, max_batch=256,
                  max_batch_pixels=args.batch_pixels, schedule=args.serve_schedule)
    sb.replay(reqs[:64], np.zeros(64))
    log("warm-up replay done")
    cap = sb.replay(reqs, np.zeros(len(reqs)))
    sb.close()
    for m in meshes:
        m.close()
    ctx.close()
    print(json.dumps({"workload": "serve_capacity", "requests": len(reqs), "streams": args.streams,
                      "max_batch_pixels": args.batch_pixels,
                      "serve_schedule": ["slots", "pipeline"][args.serve_schedule],
                      **{k: cap[k] for k in ("mrays_per_s", "batches", "mean_batch", "makespan_s", "rays",
                                             "active_rays", "steps")}}), flush=True)
    return 0


def serve_bench(args) -> int:
    import vfray_amd as V
    ctx = V.Context(0)
    sv = serve_measure(ctx, args)
    ctx.close()
    val = sv["at_slo"]["mrays_per_s"] if sv["at_slo"] else sv["stream"]["mrays_per_s"]
    out = {
        "metric": METRIC,
        "value": round(val, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(sv["stream"]["makespan_s"] * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "serve", "tets": sv["tets"], "requests": args.requests},
        "served": sv,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="batch64", choices=["batch64", "plain_gpu", "concave", "serve", "serve_capacity"])
    ap.add_argument("--requests", type=int, default=10000, help="serve: requests in the stream")
    ap.add_argument("--load", type=float, default=0.8, help="serve: offered load / batched capacity")
    ap.add_argument("--streams", type=int, default=2, help="serve: concurrent batches")
    ap.add_argument("--batch-pixels", type=int, default=4 << 20, help="serve: pixels per batch (0 = no cap)")
    ap.add_argument("--serve-schedule", type=int, default=0, help="serve: 0 slots, 1 pipeline")
    ap.add_argument("--cpu-seconds", type=float, default=6.0,
                    help="CPU baseline: seconds of work for each of its two legs (all threads, one thread)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="batch64: skip the request-stream measurement")
    ap.add_argument("--graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched: skip the schedule autotuner (one ray per thread unless --persistent)")
    ap.add_argument("--persistent", action="store_true", help="persistent traversal kernel with lane refill")
    ap.add_argument("--wave-queue", action="store_true", help="wave work-queue traversal kernel")
    ap.add_argument("--lds-records", action="store_true", help="walk with LDS-staged tet records")
    ap.add_argument("--pipelined", action="store_true",
                    help="batched: view chunks whose preprocessing overlaps the walk of earlier chunks")
    ap.add_argument("--occupancy", type=int, default=0)
    ap.add_argument("--wpc", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--probe", type=int, default=0,
                    help="only run the copy-kernel bandwidth probe over this many bytes")
    ap.add_argument("--diag", action="store_true", help="per-view diagnostics on stderr")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the renderer is "
              "single-GPU; views would shard across GPUs, one mesh replica each", file=sys.stderr)
        return 2

    import vfray_amd as V
    from vfray_amd import workloads as WL

    if args.probe:
        with V.Context(0) as c:
            g = c.bandwidth_probe(args.probe, 20)
        print(json.dumps({"probe_bytes": args.probe, "copy_gbps": round(g, 1)}), flush=True)
        return 0

    if args.workload == "serve":
        return serve_bench(args)
    if args.workload == "serve_capacity":
        return serve_capacity(args)

    t0 = time.time()
    if args.workload == "batch64":
        x, t, tfa, seed = WL.batch64()
        nviews, sides = 64, list(WL.SIDES_BATCH)
    elif args.workload == "concave":
        x, t, tfa, seed = WL.concave()
        nviews, sides = 32, [768]
    else:
        x, t, tfa, seed = WL.plain_gpu()
        nviews, sides = 1, [1024]
    log(f"mesh {x.shape[0]} verts {t.shape[0]} tets generated in {time.time() - t0:.1f}s")
    ctx = V.Context(0)
    t0 = time.time()
    mesh = ctx.mesh(x, t)
    log(f"device mesh built in {time.time() - t0:.1f}s: {mesh.info}")
    tf = ctx.tf(tfa)
    views = WL.sphere_views(seed, nviews, sides, tf)
    total_rays = sum(v.width * v.height for v in views)

    tune = None
    manual = (args.persistent or args.wave_queue or args.pipelined or args.occupancy or args.wpc or args.refill
              or args.lds_records)
    if not args.no_autotune and not manual:
        pt = ctx.plan(mesh, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED)
        tune = pt.autotune()
        pt.close()
        log(f"autotune: {tune}")
    kw = {} if tune is None else {"persistent": tune["persistent"], "wave_queue": tune["wave_queue"],
                                  "pipelined": tune["pipelined"], "waves_per_cu": tune["waves_per_cu"],
                                  "refill": tune["refill"], "occupancy": tune["occupancy"],
                                  "lds_records": tune["lds_records"]}
    if tune is None and manual:
        kw = {"persistent": args.persistent, "wave_queue": args.wave_queue, "pipelined": args.pipelined,
              "waves_per_cu": args.wpc, "refill": args.refill, "occupancy": args.occupancy,
              "lds_records": args.lds_records}
        tune = dict(kw)
    pb = ctx.plan(mesh, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, use_graph=args.graph, **kw)
    for _ in range(args.warmup):
        pb.execute()
    log("batched warm-up done")
    st_list = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st_list.append(pb.execute())
    wall = (time.perf_counter() - t0) / args.steps
    log(f"batched: {wall * 1e3:.2f} ms/step")
    st = st_list[-1]
    ms_evt = float(np.mean([s["ms_total"] for s in st_list]))
    # the walk kernel alone (one launch over every view) for the roofline
    if not kw.get("pipelined") and not args.graph:
        ms_trav = float(np.mean([s["ms_traversal"] for s in st_list]))
    else:
        kw1 = dict(kw, pipelined=False)
        ps = ctx.plan(mesh, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, **kw1)
        ps.execute()
        ms_trav = float(np.mean([ps.execute()["ms_traversal"] for _ in range(3)]))
        ps.close()
    log(f"walk launch alone: {ms_trav:.3f} ms")

    plain = None
    if not args.no_plain:
        pp = ctx.plan(mesh, views, mode=V.MODE_FTB, schedule=V.SCHED_PER_VIEW)
        for _ in range(args.warmup):
            pp.execute()
        sp = []
        t1 = time.perf_counter()
        for _ in range(args.steps):
            sp.append(pp.execute())
        pwall = (time.perf_counter() - t1) / args.steps
        plain = {"mrays_per_s": total_rays / pwall / 1e6, "ms_per_step": pwall * 1e3,
                 "frame_p50_ms": float(np.median([s["ms_frame_p50"] for s in sp])),
                 "frame_p99_ms": float(np.median([s["ms_frame_p99"] for s in sp]))}
        same = all(np.array_equal(pb.read(i)["rgba"], pp.read(i)["rgba"]) for i in range(min(3, nviews)))
        plain["batched_equals_plain"] = bool(same)
        if args.diag:
            ft = pp.frame_times()
            for i in range(nviews):
                r = pp.read(i, rgba=False, entry=False)
                act = r["steps"] > 0
                log(f"view {i} {views[i].width}^2 frame {ft[i]:.2f} ms rays {int(act.sum())} "
                    f"steps max {int(r['steps'].max())} mean {float(r['steps'][act].mean()):.1f} "
                    f"sum {int(r['steps'].sum())}")
            pd = ctx.plan(mesh, views, mode=V.MODE_FTB, schedule=V.SCHED_BATCHED, persistent=True)
            pd.execute()
            sd = pd.execute()
            log(f"batched, persistent kernel (defaults): traversal {sd['ms_traversal']:.2f} ms")
            pd.close()
        log(f"plain: {pwall * 1e3:.2f} ms/step")
        pp.close()

    steps_per_launch = st["steps"]
    rays_active = st["active_rays"]
    model_bytes = B_STEP * steps_per_launch + B_RAY * rays_active
    roof = None
    if ms_trav:
        kname = ("k_trav_persist" if kw.get("persistent") else "k_trav_wq" if kw.get("wave_queue")
                 else "k_trav_ldsr" if kw.get("lds_records") else "k_trav")
        tr = measured_traffic(kname, args.workload)
        # achieved: bytes the walk moves through HBM per launch (rocprofv3
        # FETCH_SIZE + WRITE_SIZE of the same launch, copy-kernel calibrated,
        # profiles/<kernel>_<workload>.md) over the launch time timed here
        ach = tr["traffic_bytes"] / (ms_trav * 1e-3) / 1e9 if tr else None
        roof = {"bound": "hbm", "achieved": round(ach, 1) if ach else None, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / PEAK_HBM_GBS, 4) if ach else None,
                "traffic": tr["traffic_bytes"] if tr else None,
                "traffic_source": f"profiles/{kname}_{args.workload}.md" if tr else None,
                "l2_hit": tr["l2_hit"] if tr else None,
                # what bounds the walk: the vector memory pipeline between the
                # waves and L2 (same profile; per CU, shares of the launch's cycles)
                "memory_pipeline": None if not tr or "l1_lookups_per_cu_cycle" not in tr else {
                    "l1_line_lookups_per_cu_cycle": tr["l1_lookups_per_cu_cycle"], "l1_hit": tr["l1_hit"],
                    "texture_data_busy": tr["td_busy"], "l1_waiting_on_l2": tr["l1_pending_stall"]},
                "kernel": kname, "kernel_ms": round(ms_trav, 3),
                "model_bytes": int(model_bytes),
                "model_gbs": round(model_bytes / (ms_trav * 1e-3) / 1e9, 1),
                "model_note": "96 B per tet step + 48 B per active ray in the reference's layout, every gather "
                              "counted, most of them served by L2 and the Infinity Cache: not HBM traffic"}
    served = None
    if args.workload == "batch64" and not args.no_serve:
        pb.close()
        pb = None
        served = serve_measure(ctx, args)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(x, t, tfa, views, args.cpu_seconds)
        log(f"cpu baseline {cpu['value']:.3f} Mrays/s")
    value = total_rays / wall / 1e6
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": args.workload, "tets": int(t.shape[0]), "views": nviews,
                   "rays_per_step": int(total_rays), "active_rays": int(rays_active),
                   "tet_steps": int(steps_per_launch), "term_alpha": 0.99, "graph": args.graph,
                   "pipelined": bool(kw.get("pipelined")) and not args.graph,
                   "schedule": tune or {"persistent": False}},
        "value_is": "batched renderer, all views of the workload in one launch set, inputs resident",
        "anomalous_rays": int(st["anomalies"]),
        "event_ms_per_step": round(ms_evt, 3),
        "event_ms_median": round(float(np.median([s["ms_total"] for s in st_list])), 3),
        "active_mrays_per_s": round(rays_active / wall / 1e6, 2),
        "gsteps_per_s": round(steps_per_launch / wall / 1e9, 3),
        "plain": plain,
        "speedup_vs_plain": round(value / plain["mrays_per_s"], 3) if plain else None,
        "target_vs_plain": 3.0,
        "served": served,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if pb is not None:
        pb.close()
    mesh.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
