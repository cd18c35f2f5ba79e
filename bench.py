"""Benchmark: incoherent ray casting on one MI355X.

Workload (one step): 16,777,216 closest-hit rays plus 16,777,216 shadow
segments against a displaced icosphere of 1,310,720 triangles (20 * 4^8, the
icosphere level closest to 2^20), SAH BVH built with the reference's builder.
Rays: origins uniform in the mesh bbox scaled 1.2x, directions uniform on S^2;
segment endpoints uniform in the same box.  Inputs are resident in HBM before
timing starts.

Prints one JSON line: Grays/s over both ray kinds, the closest-hit kernel's
roofline entry (algorithmic bytes from the reference-order traversal's own
node/primitive test counts, over its event-timed duration), and a CPU baseline
(the oracle's restatement of Bvh::intersect / Bvh::visibility on host cores,
timed on a bounded sample of the same rays).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
COPY_PROBE_BYTES = 1 << 29  # 512 MiB, read once and written once
# HBM bytes per launch of the traversal kernels on the default workload
# (icosphere level 8, 16,777,216 rays / segments, the library's default
# setting: rays in input order, one range per XCD, 8 LDS stack entries, refill
# at 32 idle lanes (40 for visibility), 128-ray chunks), from two
# `rocprofv3 --pmc` runs of `bench.py --traffic-probe`: FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0 (KiB), the factors being what the counters read for the
# probe's 512 MiB device copy.  Counters cannot be read inside this process, so
# the figure is reported only when the timed launches run that exact setting.
MEASURED_TRAFFIC = {
    "closest": (8302469.1 * 2.0 + 555521.4) * 1024,
    "shadow": (10929519.4 * 2.0 + 652414.1) * 1024,
}
# TCC_HIT_sum / (TCC_HIT_sum + TCC_MISS_sum) of the same launches (a third counter run)
MEASURED_L2_HIT = {
    "closest": 148800515.0 / (148800515.0 + 145830664.0),
    "shadow": 228117297.0 / (228117297.0 + 183153193.0),
}
MEASURED_TRAFFIC_SETTING = dict(level=8, rays=1 << 24, sort_bits=0, lds_stack=8, refill=32, waves_per_cu=0,
                                xcd_ranges=1, ray_chunk=128, vis_refill=40)
RAY_BYTES, HIT_BYTES, VIS_BYTES = 32, 16, 4
NODE_BYTES, TRI_BYTES, SPHERE_BYTES = 32, 36, 16


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def algorithmic_bytes(n, counts, out_bytes):
    c = counts.sum(axis=0, dtype=np.float64)
    return n * (RAY_BYTES + out_bytes) + NODE_BYTES * c[0] + TRI_BYTES * c[1] + SPHERE_BYTES * c[2]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cpus():
    """What the host offers the CPU baseline: nproc, the CPUs this process may
    run on, and the cgroup CPU quota (CPUs' worth of time, None if unlimited)."""
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return dict(nproc=os.cpu_count() or 1, affinity=affinity, cgroup_quota_cpus=quota)


# One GPU's share of the GPU box's CPUs: the CPU baseline uses at most this many
# threads, whatever nproc shows.
CPU_SHARE = 16


def baseline_threads(hc):
    n = min(hc["affinity"], CPU_SHARE)
    if hc["cgroup_quota_cpus"]:
        n = min(n, max(1, int(hc["cgroup_quota_cpus"])))
    return max(1, n)


_ORACLE_BUILD = "x86-64-v2 (prebuilt)"


def use_native_oracle():
    """Rebuild the oracle with -march=native on this host for the CPU-baseline
    legs (its results are the same; only its speed changes)."""
    global _ORACLE_BUILD
    import tempfile
    from oracle import oracle as O
    path = os.path.join(tempfile.mkdtemp(prefix="oracle_native_"), "liboracle_native.so")
    t0 = time.time()
    if O.build_native(path):
        O.use_library(path)
        _ORACLE_BUILD = "-O3 -march=native (built on this host in %.1f s)" % (time.time() - t0)
    log("oracle for the CPU baseline: " + _ORACLE_BUILD)


def cpu_record(threads):
    hc = host_cpus()
    return dict(cores=threads, cpu_model=cpu_model(), nproc=hc["nproc"], affinity=hc["affinity"],
                cgroup_quota_cpus=hc["cgroup_quota_cpus"], oracle_build=_ORACLE_BUILD,
                threads_rule="min(CPUs this process may use, %d = one GPU's CPU share, cgroup quota)" % CPU_SHARE)


def cpu_baseline(spec, rays8, segs8, sample, threads, repeats=5, single_sample=1 << 20):
    """The oracle's Bvh::intersect / Bvh::visibility restatement on `sample`
    rays and segments of the same batches, median of `repeats` passes after a
    warm-up, on `threads` host threads; plus one pass on one thread over the
    first `single_sample` rays and segments."""
    from gammaray_amd import scenes
    from oracle.oracle import OracleScene
    t0 = time.time()
    o = OracleScene(spec)
    log("oracle BVH built in %.1f s" % (time.time() - t0))
    r6 = scenes.rays6(rays8[:sample])
    s6 = scenes.rays6(segs8[:sample])
    o.intersect(r6[:4096], threads=threads)  # warm
    tc, ts = [], []
    for _ in range(repeats):
        t0 = time.perf_counter()
        o.intersect(r6, threads=threads)
        t1 = time.perf_counter()
        o.visibility(s6, threads=threads)
        t2 = time.perf_counter()
        tc.append(t1 - t0)
        ts.append(t2 - t1)
    c, sh = float(np.median(tc)), float(np.median(ts))
    m = min(single_sample, sample)
    t0 = time.perf_counter()
    o.intersect(r6[:m], threads=1)
    o.visibility(s6[:m], threads=1)
    t1 = 2 * m / (time.perf_counter() - t0) / 1e9
    return dict(value=2 * sample / (c + sh) / 1e9, unit="Grays/s", kind="port", **cpu_record(threads),
                single_thread_value=round(t1, 6),
                single_thread_sample="%d rays + %d segments on 1 thread" % (m, m),
                sample="%d closest-hit rays + %d shadow segments of the timed batches "
                       "(oracle Bvh::intersect / Bvh::visibility restatement, %d threads; median of %d passes: "
                       "closest %.2f s, shadow %.2f s; %.1f s of CPU work in all)" % (
                           sample, sample, threads, repeats, c, sh, sum(tc) + sum(ts)))


PATH_WORKLOADS = [
    # name, scene factory, width, height, spp, max_depth, integrator, light_depth, cpu spp
    ("box256_pt", "box_scene", 256, 256, 16, 6, "pt", 0, 4),
    ("showcase1080_pt_nee", "showcase_scene", 1920, 1080, 64, 8, "pt_nee", 0, 1),
    ("showcase1080_bdpt", "showcase_scene", 1920, 1080, 64, 8, "bdpt", 5, 1),
    ("showcase1080_bdpt_mis", "showcase_scene", 1920, 1080, 64, 8, "bdpt_mis", 5, 1),
    ("showcase1080_bdpt_lt", "showcase_scene", 1920, 1080, 64, 8, "bdpt_lt", 5, 1),
]


def bench_paths(ctx, stream, cpu, threads, only=None, tune=False):
    """Path-tracing throughput (Mpaths/s, one camera sample = one path) of the
    render workloads, GPU time by stream events around gr_trace, plus the
    oracle's throughput on a bounded number of iterations of the same frame.
    With tune, each scene's traversal setting is first chosen by the autotuner
    on 2^22 bounce rays and 2^22 light segments of that scene."""
    import torch
    from gammaray_amd import scenes
    from gammaray_amd.api import Film, Stage, trace
    out = {}
    tuned = {}
    for name, factory, W, H, spp, depth, integ, ldepth, cpu_spp in PATH_WORKLOADS:
        if only and name not in only:
            continue
        spec = getattr(scenes, factory)()
        stage = Stage(spec, ctx=ctx)
        if tune:
            if factory not in tuned:
                t0 = time.time()
                rays = torch.from_numpy(scenes.bounce_batch(spec, 1 << 22, seed=11)).cuda()
                segs = torch.from_numpy(scenes.light_segment_batch(spec, 1 << 22, seed=12)).cuda()
                tuned[factory] = stage.autotune(rays, segs)
                del rays, segs
                log("%s: autotuned %s (%.1f s)" % (factory, tuned[factory], time.time() - t0))
            stage.set_tuning(**tuned[factory])
        film = Film(ctx, W, H)
        trace(stage, film, integ, spp=1, max_depth=depth, light_depth=ldepth, seed=1)  # warm
        torch.cuda.synchronize()
        film.clear()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        trace(stage, film, integ, spp=spp, max_depth=depth, light_depth=ldepth, seed=7)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        rec = {"mpaths_s": round(W * H * spp / (ms * 1e-3) / 1e6, 2), "ms": round(ms, 3), "width": W,
               "height": H, "spp": spp, "max_depth": depth, "integrator": integ,
               "primitives": spec.num_components,
               "tuning": tuned.get(factory) or "the library's integrator default (refill 16, vis_refill 16, ray_chunk 64)"}
        if ldepth:
            rec["light_depth"] = ldepth
        if cpu:
            from oracle import oracle as O
            o = O.OracleScene(spec)
            oi = {"pt": O.PT, "pt_nee": O.PT_NEE, "bdpt": O.BDPT, "bdpt_mis": O.BDPT_MIS, "bdpt_lt": O.BDPT_LT}[integ]
            O.render(o, spec, W, H, oi, spp=1, max_depth=depth, light_depth=ldepth, seed=7, threads=threads)
            times = []
            for it in range(5):  # median of 5 iterations after a warm-up
                t0 = time.perf_counter()
                O.render(o, spec, W, H, oi, spp=cpu_spp, max_depth=depth, light_depth=ldepth, seed=7,
                         sample_offset=1 + it * cpu_spp, threads=threads)
                times.append(time.perf_counter() - t0)
            dt = float(np.median(times))
            rec["cpu_baseline"] = {"mpaths_s": round(W * H * cpu_spp / dt / 1e6, 4), "kind": "port",
                                   **cpu_record(threads),
                                   "sample": "median of 5 timed passes of %d iteration(s) of the same frame "
                                             "after a warm-up" % cpu_spp}
            rec["gpu_over_cpu"] = round(rec["mpaths_s"] / rec["cpu_baseline"]["mpaths_s"], 1)
        log("%s: %.1f Mpaths/s (%.2f ms)%s" % (name, rec["mpaths_s"], ms,
                                              ", cpu %.3f Mpaths/s" % rec["cpu_baseline"]["mpaths_s"] if cpu else ""))
        out[name] = rec
        film.close()
        stage.close()
    return out


def bench_batch(ctx, threads, pool=1 << 23, steps=16):
    """64 jobs (side log-uniform 128..1024, spp in {4,16,64,256}) over 8
    displaced-mesh scenes (10k..2M triangles), PT+NEE depth 8, through one
    path pool with the bounce step as a hipGraph; aggregate Mpaths/s and job
    latency, next to the same jobs rendered one after another with gr_trace."""
    import torch
    from gammaray_amd import scenes
    from gammaray_amd.api import Film, Stage, autotune_jobs, render_jobs, trace
    t0 = time.time()
    specs, jobs = scenes.batch_workload(seed=4)
    stages = [Stage(sp, ctx=ctx) for sp in specs]
    log("batch: 8 scenes %s triangles built (%.1f s)" % ([sp.num_components for sp in specs], time.time() - t0))
    # the batch traces bounce rays and NEE shadow segments in every scene at once: tune its
    # multi-scene traversal on rays of both kinds in each scene (leaving its surfaces, and from
    # them to its light), each naming its scene as the batch's own do
    per = 1 << 18
    rays = np.concatenate([scenes.bounce_batch(sp, per, seed=9 + k) for k, sp in enumerate(specs)])
    segs = np.concatenate([scenes.light_segment_batch(sp, per, seed=100 + k) for k, sp in enumerate(specs)])
    tag = np.repeat(np.arange(len(specs), dtype=np.uint32), per).view(np.float32)
    rays[:, 3] = tag
    segs[:, 3] = tag
    rays, segs = torch.from_numpy(rays).cuda(), torch.from_numpy(segs).cuda()
    tuning = autotune_jobs(ctx, stages, rays, segs)
    del rays, segs
    for st in stages:
        st.set_tuning(**tuning)
    films = [Film(ctx, j["side"], j["side"]) for j in jobs]
    jl = [dict(stage=stages[j["scene"]], film=films[k], spp=j["spp"], max_depth=8, seed=j["seed"])
          for k, j in enumerate(jobs)]
    total = sum(j["side"] ** 2 * j["spp"] for j in jobs)
    torch.cuda.synchronize()
    t = time.perf_counter()
    lat = render_jobs(ctx, jl, "pt_nee", path_pool=pool, steps_per_poll=steps, use_graph=True, tuning=tuning)
    wall = time.perf_counter() - t
    log("batch: %d samples in %.3f s -> %.1f Mpaths/s" % (total, wall, total / wall / 1e6))
    for f in films:
        f.clear()
    ctx.synchronize()
    t = time.perf_counter()
    lat_ff = render_jobs(ctx, jl, "pt_nee", path_pool=pool, steps_per_poll=steps, use_graph=True, tuning=tuning,
                         job_order="fewest_first")
    wall_ff = time.perf_counter() - t
    log("batch, fewest samples first: %.3f s, p50 %.0f ms p99 %.0f ms" % (
        wall_ff, np.percentile(lat_ff, 50), np.percentile(lat_ff, 99)))
    for f in films:
        f.clear()
    ctx.synchronize()
    seq_lat = []
    t = time.perf_counter()
    for job in jl:
        trace(job["stage"], job["film"], "pt_nee", spp=job["spp"], max_depth=8, seed=job["seed"])
        ctx.synchronize()
        seq_lat.append((time.perf_counter() - t) * 1e3)
    seq = time.perf_counter() - t
    log("batch: sequential gr_trace %.3f s -> %.1f Mpaths/s" % (seq, total / seq / 1e6))
    for f in films:
        f.close()
    return {"jobs": len(jobs), "scenes": len(specs), "triangles": [sp.num_components for sp in specs],
            "samples": total, "mpaths_s": round(total / wall / 1e6, 2), "wall_s": round(wall, 4),
            "latency_ms_p50": round(float(np.percentile(lat, 50)), 2),
            "latency_ms_p99": round(float(np.percentile(lat, 99)), 2),
            "job_order": "as given",
            "fewest_first": {"mpaths_s": round(total / wall_ff / 1e6, 2), "wall_s": round(wall_ff, 4),
                             "latency_ms_p50": round(float(np.percentile(lat_ff, 50)), 2),
                             "latency_ms_p99": round(float(np.percentile(lat_ff, 99)), 2)},
            "sample_ring": min(16 * pool, total), "sample_ring_bytes": min(16 * pool, total) * 24,
            "path_pool": pool, "graph": True, "tuning": tuning, "max_depth": 8, "integrator": "pt_nee",
            "sequential": {"mpaths_s": round(total / seq / 1e6, 2), "wall_s": round(seq, 4),
                           "latency_ms_p50": round(float(np.percentile(seq_lat, 50)), 2),
                           "latency_ms_p99": round(float(np.percentile(seq_lat, 99)), 2)}}


def _rmse(img, ref):
    """Root mean square error over the pixels' channels (non-finite pixels of
    either image excluded)."""
    d = ((img - ref) ** 2).ravel()
    return float(np.sqrt(d[np.isfinite(d)].mean()))


def _rel_mse_trimmed(img, ref, trim=1e-3):
    """Mean of (x - ref)^2 / (ref^2 + 0.01) over the pixel channels, dropping the
    largest `trim` fraction (the reference's own fireflies)."""
    d = ((img - ref) ** 2 / (ref ** 2 + 1e-2)).ravel()
    d = d[np.isfinite(d)]
    return float(np.sort(d)[: int(len(d) * (1 - trim))].mean())


GUIDE_FLOPS_PER_QUERY = 2 * (208 * 64 + 3 * 64 * 64 + 64 * 32)  # the five layers' multiply-adds
FP16_DENSE_PEAK_TFLOPS = 2500.0


def guide_inference_rate(ctx, guide, n=1 << 22, reps=5):
    """Throughput of the guide's MFMA inference kernel on n queries (HIP events
    on the context's stream around each launch)."""
    import torch
    raw = torch.rand((n, 8), dtype=torch.float32, device="cuda")
    guide.infer(raw)
    torch.cuda.synchronize()
    st = torch.cuda.ExternalStream(ctx.stream()) if ctx.stream() else torch.cuda.current_stream()
    times = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        guide.infer(raw)
        e1.record(st)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    ms = float(np.median(times))
    tflops = GUIDE_FLOPS_PER_QUERY * n / (ms * 1e-3) / 1e12
    log("guide inference: %d queries in %.3f ms -> %.1f TFLOP/s" % (n, ms, tflops))
    return {"queries": n, "ms": round(ms, 4), "flops_per_query": GUIDE_FLOPS_PER_QUERY,
            "tflops": round(tflops, 2), "peak_tflops": FP16_DENSE_PEAK_TFLOPS,
            "frac": round(tflops / FP16_DENSE_PEAK_TFLOPS, 4), "bound": "mfma", "dtype": "fp16 x fp16 -> fp32"}


GUIDED_CONFIG = dict(alpha=0.3, guided_depth=1, steps_per_pass=4, batch=16384, train_paths=1 << 19, warmup_steps=8,
                     learning_rate=5e-3, lobe_select="learned", select_mix=0.4, rr_threshold=0.1)


def _parse_overrides(text):
    """'k=v,k=v' -> dict with int, float or string values."""
    out = {}
    for item in filter(None, text.split(",")):
        k, v = item.split("=")
        for cast in (int, float, str):
            try:
                out[k] = cast(v)
                break
            except ValueError:
                pass
    return out


GUIDED_SCENES = {
    # name: (scene builder, width, height, description)
    "showcase": ("showcase_scene", 1920, 1080, "showcase 1920x1080, PT+NEE depth 8"),
    "room": ("indirect_room_scene", 1440, 1440, "indirect-light room 1440x1440 (light facing the ceiling), PT+NEE depth 8"),
}


def bench_guided(ctx, ref_spp=4096, spp=64, overrides=None, scene="showcase"):
    """Learned path guide: wall time for guided rendering (online training
    included) to reach the relative MSE that plain PT+NEE reaches with 64 spp,
    against a ref_spp PT+NEE render, on the 1080p showcase scene or on the
    indirectly lit room; and the same with the network switched off."""
    from gammaray_amd import scenes
    from gammaray_amd.api import Film, Guide, Stage, trace
    builder, W, H, label = GUIDED_SCENES[scene]
    spec = getattr(scenes, builder)()
    st = Stage(spec, ctx=ctx)

    def image(f):
        px = f.readback()
        return px[..., :3] / px[..., 3:4]

    def timed(fn):
        ctx.synchronize()
        t = time.perf_counter()
        fn()
        ctx.synchronize()
        return time.perf_counter() - t

    ref = Film(ctx, W, H)
    t_ref = timed(lambda: trace(st, ref, "pt_nee", spp=ref_spp, max_depth=8, seed=1000))
    R = image(ref)
    ref.close()
    plain = Film(ctx, W, H)
    trace(st, plain, "pt_nee", spp=1, max_depth=8, seed=5)
    plain.clear()
    t_plain = timed(lambda: trace(st, plain, "pt_nee", spp=spp, max_depth=8, seed=7))
    img = image(plain)
    e_plain = _rmse(img, R)
    r_plain = _rel_mse_trimmed(img, R)
    plain.close()
    # batches of 8 spp like the plain render; the guide's mixture at the first vertex, the learned
    # lobe choice and the roulette at every vertex (GUIDED_CONFIG)
    cfg = dict(GUIDED_CONFIG)
    cfg.update(overrides or {})

    def run(c):
        """8-spp batches to twice the plain spp: (curve rows, wall time to plain's RMSE, to its relMSE)."""
        g = Guide(ctx, **c)
        f = Film(ctx, W, H)
        tot, rows = 0.0, []
        for k in range(0, 2 * spp, 8):
            tot += timed(lambda: trace(st, f, "guided", spp=8, max_depth=8, seed=9, sample_offset=k, batch=8,
                                       guide=g))
            img = image(f)
            rows.append((k + 8, round(tot, 4), round(_rmse(img, R), 6), round(_rel_mse_trimmed(img, R), 6)))
        f.close()
        g.close()

        def reach(col, target):
            # wall time at which the curve first reaches `target` (linear between batches)
            prev = (0.0, float("inf"))
            for row in rows:
                if row[col] <= target:
                    t0, e0 = prev
                    return t0 + (row[1] - t0) * ((e0 - target) / (e0 - row[col])
                                                  if np.isfinite(e0) and e0 > row[col] else 1.0)
                prev = (row[1], row[col])
            return None
        return rows, reach(2, e_plain), reach(3, r_plain)

    g = Guide(ctx, **cfg)
    infer = guide_inference_rate(ctx, g)
    g.close()
    curve, reached, reached_rel = run(cfg)
    log("guided: plain %d spp %.3f s RMSE %.5f relMSE %.5f; guided reaches the RMSE at %s s, the relMSE at %s s" % (
        spp, t_plain, e_plain, r_plain, "%.3f" % reached if reached else "--", "%.3f" % reached_rel if reached_rel else "--"))
    # the same integrator with the network off: never sampled (alpha 0) and never trained, so it is
    # never queried either; lobe choice and roulette as above
    off = dict(cfg, alpha=0.0, steps_per_pass=0)
    curve_off, _, reached_off = run(off)
    log("guided without the network: reaches the relMSE at %s s" % ("%.3f" % reached_off if reached_off else "--"))
    return {"scene": label, "reference_spp": ref_spp,
            "reference_s": round(t_ref, 3),
            "error": "relMSE with the 0.1% largest pixel errors dropped (the time ratio).  The untrimmed RMSE is "
                     "reported beside it; single-pixel fireflies of one render or the other dominate it, so its "
                     "time ratio measures those few pixels, not the estimator",
            "plain_spp": spp, "plain_s": round(t_plain, 4), "plain_relmse_trimmed": round(r_plain, 6),
            "plain_rmse_untrimmed": round(e_plain, 6),
            "guided_config": cfg, "guided_curve": curve,
            "guided_curve_columns": ["spp", "s", "rmse_untrimmed", "relmse_trimmed"],
            "inference": infer,
            "time_to_equal_error_s": round(reached_rel, 4) if reached_rel else None,
            "time_ratio": round(reached_rel / t_plain, 3) if reached_rel else None,
            "time_ratio_rmse_untrimmed": round(reached / t_plain, 3) if reached else None,
            "without_network": {"config": off, "curve": curve_off,
                                "time_ratio": round(reached_off / t_plain, 3) if reached_off else None}}


def _without_inconsistent_lobes(spec):
    """The scene with clearcoat and specular transmission set to zero on every
    Disney material: all remaining lobes report their sampling densities."""
    for prim in spec.prims:
        m = prim.material
        if m[0] == 0:
            m[8] = 0.0   # specular_trans
            m[12] = 0.0  # clearcoat
    return spec


GUIDE_SWEEP = [("alpha 0.3, first vertex", dict(alpha=0.3, guided_depth=1)),
               ("alpha 0.5, first vertex", dict(alpha=0.5, guided_depth=1)),
               ("alpha 0.7, first vertex", dict(alpha=0.7, guided_depth=1)),
               ("alpha 0.3, two vertices", dict(alpha=0.3, guided_depth=2)),
               ("alpha 0.5, three vertices", dict(alpha=0.5, guided_depth=3)),
               ("alpha 0.3, first vertex, 16 Adam steps per batch", dict(alpha=0.3, guided_depth=1, steps_per_pass=16)),
               ("all lobes, alpha 0.3, first vertex", dict(alpha=0.3, guided_depth=1, lobes="all")),
               ("all lobes, alpha 0.5, first vertex", dict(alpha=0.5, guided_depth=1, lobes="all")),
               ("all lobes, alpha 0.3, two vertices", dict(alpha=0.3, guided_depth=2, lobes="all")),
               ("all lobes, alpha 0.3, all vertices", dict(alpha=0.3, guided_depth=0, lobes="all")),
               ("albedo lobe choice (mix 0.25), alpha 0.3, first vertex",
                dict(alpha=0.3, guided_depth=1, lobe_select="albedo", select_mix=0.25)),
               ("albedo lobe choice (mix 0.1), alpha 0.3, first vertex",
                dict(alpha=0.3, guided_depth=1, lobe_select="albedo", select_mix=0.1)),
               ("albedo lobe choice (mix 0.25), all lobes, alpha 0.3, first vertex",
                dict(alpha=0.3, guided_depth=1, lobes="all", lobe_select="albedo", select_mix=0.25)),
               ("albedo lobe choice (mix 0.25), alpha 0 (no guide)",
                dict(alpha=0.0, guided_depth=1, lobe_select="albedo", select_mix=0.25)),
               ("learned lobe choice (mix 0.25), alpha 0.3, first vertex",
                dict(alpha=0.3, guided_depth=1, lobe_select="learned", select_mix=0.25)),
               ("learned lobe choice (mix 0.1), alpha 0.3, first vertex",
                dict(alpha=0.3, guided_depth=1, lobe_select="learned", select_mix=0.1)),
               ("learned lobe choice (mix 0.25), alpha 0 (no guide)",
                dict(alpha=0.0, guided_depth=1, lobe_select="learned", select_mix=0.25)),
               ("learned lobe choice (mix 0.4), alpha 0.3, first vertex",
                dict(alpha=0.3, guided_depth=1, lobe_select="learned", select_mix=0.4)),
               ("learned lobe choice (mix 0.25), all lobes, alpha 0.3, first vertex",
                dict(alpha=0.3, guided_depth=1, lobes="all", lobe_select="learned", select_mix=0.25)),
               ("learned lobe choice (mix 0.5), alpha 0.3, first vertex",
                dict(alpha=0.3, guided_depth=1, lobe_select="learned", select_mix=0.5)),
               ("learned lobe choice (mix 0.4), alpha 0.3, first vertex, 4 Adam steps per batch",
                dict(alpha=0.3, guided_depth=1, lobe_select="learned", select_mix=0.4, steps_per_pass=4)),
               ("learned lobe choice (mix 0.4), alpha 0.3, first vertex, 262144 training paths",
                dict(alpha=0.3, guided_depth=1, lobe_select="learned", select_mix=0.4, train_paths=1 << 18)),
               ("learned lobe choice (mix 0.4), alpha 0.3, two vertices",
                dict(alpha=0.3, guided_depth=2, lobe_select="learned", select_mix=0.4)),
               ("learned lobe choice (mix 0.4), roulette 0.1, alpha 0.3, first vertex",
                dict(alpha=0.3, guided_depth=1, lobe_select="learned", select_mix=0.4, rr_threshold=0.1)),
               ("learned lobe choice (mix 0.4), roulette 0.25, alpha 0.3, first vertex",
                dict(alpha=0.3, guided_depth=1, lobe_select="learned", select_mix=0.4, rr_threshold=0.25)),
               ("diffuse lobes guided, learned lobe choice (mix 0.4), roulette 0.1, alpha 0.3, first vertex",
                dict(alpha=0.3, guided_depth=1, lobes="diffuse", lobe_select="learned", select_mix=0.4,
                     rr_threshold=0.1, steps_per_pass=4, train_paths=1 << 19)),
               ("diffuse lobes guided, learned lobe choice (mix 0.4), roulette 0.1, alpha 0.5, first vertex",
                dict(alpha=0.5, guided_depth=1, lobes="diffuse", lobe_select="learned", select_mix=0.4,
                     rr_threshold=0.1, steps_per_pass=4, train_paths=1 << 19)),
               ("diffuse lobes guided, learned lobe choice (mix 0.4), roulette 0.1, alpha 0.3, two vertices",
                dict(alpha=0.3, guided_depth=2, lobes="diffuse", lobe_select="learned", select_mix=0.4,
                     rr_threshold=0.1, steps_per_pass=4, train_paths=1 << 19)),
               ("network off, learned lobe choice (mix 0.4), roulette 0.1",
                dict(alpha=0.0, guided_depth=1, lobe_select="learned", select_mix=0.4, rr_threshold=0.1,
                     steps_per_pass=0, train_paths=1 << 19)),
               ("consistent lobes guided, learned lobe choice (mix 0.4), roulette 0.1, alpha 0.3, first vertex",
                dict(alpha=0.3, guided_depth=1, lobe_select="learned", select_mix=0.4, rr_threshold=0.1,
                     steps_per_pass=4, train_paths=1 << 19)),
               ("learned lobe choice (mix 0.4), roulette 0.5, alpha 0.3, first vertex",
                dict(alpha=0.3, guided_depth=1, lobe_select="learned", select_mix=0.4, rr_threshold=0.5)),
               ("learned lobe choice (mix 0.4), roulette 0.05, alpha 0.3, first vertex, 4 Adam steps per batch",
                dict(alpha=0.3, guided_depth=1, lobe_select="learned", select_mix=0.4, rr_threshold=0.05,
                     steps_per_pass=4)),
               ("learned lobe choice (mix 0.4), roulette 0.1, alpha 0.3, first vertex, 4 Adam steps per batch",
                dict(alpha=0.3, guided_depth=1, lobe_select="learned", select_mix=0.4, rr_threshold=0.1,
                     steps_per_pass=4)),
               ("learned lobe choice (mix 0.4), roulette 0.15, alpha 0.3, first vertex, 4 Adam steps per batch",
                dict(alpha=0.3, guided_depth=1, lobe_select="learned", select_mix=0.4, rr_threshold=0.15,
                     steps_per_pass=4)),
               ("learned lobe choice (mix 0.25), roulette 0.1, alpha 0.3, first vertex, 4 Adam steps per batch",
                dict(alpha=0.3, guided_depth=1, lobe_select="learned", select_mix=0.25, rr_threshold=0.1,
                     steps_per_pass=4))]


def bench_guide_sweep(ctx, ref_spp=2048, W=1920, H=1080, spp=128, only=None):
    """Guide settings on the 1080p showcase scene and on the same scene without
    its clearcoat and transmission lobes: trimmed relMSE against a ref_spp
    PT+NEE render and wall time after every 8-spp batch, for plain PT+NEE and
    for each setting in GUIDE_SWEEP (online training included)."""
    from gammaray_amd import scenes
    from gammaray_amd.api import Film, Guide, Stage, trace

    def image(f):
        px = f.readback()
        return px[..., :3] / px[..., 3:4]

    def timed(fn):
        ctx.synchronize()
        t = time.perf_counter()
        fn()
        ctx.synchronize()
        return time.perf_counter() - t

    def curve(st, R, integrator, guide=None):
        f = Film(ctx, W, H)
        if guide is None:
            trace(st, f, "pt_nee", spp=1, max_depth=8, seed=5)
            f.clear()
        tot, rows = 0.0, []
        for k in range(0, spp, 8):
            tot += timed(lambda: trace(st, f, integrator, spp=8, max_depth=8, seed=9, sample_offset=k, batch=8,
                                       guide=guide))
            rows.append((k + 8, round(tot, 4), round(_rel_mse_trimmed(image(f), R), 6)))
        f.close()
        return rows

    out = {}
    for name, spec in (("showcase", scenes.showcase_scene()),
                       ("showcase without clearcoat and transmission",
                        _without_inconsistent_lobes(scenes.showcase_scene()))):
        st = Stage(spec, ctx=ctx)
        ref = Film(ctx, W, H)
        trace(st, ref, "pt_nee", spp=ref_spp, max_depth=8, seed=1000)
        R = image(ref)
        ref.close()
        res = {"plain PT+NEE": curve(st, R, "pt_nee")}
        log("guide sweep, %s: plain %s" % (name, res["plain PT+NEE"][-1]))
        for label, cfg in GUIDE_SWEEP:
            if only and not any(w in label for w in only):
                continue
            c = dict(alpha=0.3, guided_depth=1, steps_per_pass=8, batch=16384, train_paths=1 << 16,
                     warmup_steps=8, learning_rate=5e-3)
            c.update(cfg)
            g = Guide(ctx, **c)
            res[label] = curve(st, R, "guided", g)
            g.close()
            log("guide sweep, %s: %s %s" % (name, label, res[label][-1]))
        out[name] = res
        st.close()
    return {"reference_spp": ref_spp, "columns": ["spp", "s", "relmse_trimmed"], "curves": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rays", type=int, default=1 << 24)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--occlusion", choices=["exact", "anyhit"], default="exact")
    ap.add_argument("--cpu-sample", type=int, default=1 << 24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-paths", action="store_true", help="skip the path-tracing workloads")
    ap.add_argument("--paths", default="",
                    help="comma-separated subset of the path workloads (%s)" % ", ".join(w[0] for w in PATH_WORKLOADS))
    ap.add_argument("--path-tuning", choices=["default", "auto"], default="default",
                    help="traversal setting of the path workloads: the library's default, or autotuned per scene "
                         "on its bounce rays and light segments")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-job batched workload")
    ap.add_argument("--no-guided", action="store_true", help="skip the learned-guide comparison")
    ap.add_argument("--ref-spp", type=int, default=4096, help="reference spp of the guided comparison")
    ap.add_argument("--guided-only", action="store_true",
                    help="only the guided-vs-plain comparison on the 1080p scene, printed as JSON")
    ap.add_argument("--guided-scene", choices=sorted(GUIDED_SCENES), default="showcase",
                    help="with --guided-only: the scene of the guided comparison")
    ap.add_argument("--guided-config", default="",
                    help="k=v overrides of the guide settings of the guided comparison, e.g. train_paths=262144")
    ap.add_argument("--sweep-only", default="",
                    help="with --guide-sweep: comma-separated substrings; only settings whose label has one")
    ap.add_argument("--guide-sweep", action="store_true",
                    help="only the guide-settings sweep on the 1080p scene (with and without its clearcoat and "
                         "transmission lobes), printed as JSON")
    ap.add_argument("--guide-inference-only", action="store_true",
                    help="for profiles: only the guide inference timing (4M queries), printed as JSON")
    ap.add_argument("--tuning", default="default",
                    help="'default' (the library's setting), 'auto' (autotune on the step's own mix "
                         "before timing), or "
                         "k=v pairs, e.g. sort_bits=18,lds_stack=16,waves_per_cu=0,refill=16")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for counter runs: one %d MiB device copy (calibration), then one launch "
                         "of each traversal kernel; no timing, no JSON" % (COPY_PROBE_BYTES >> 20))
    args = ap.parse_args()
    only = [w for w in args.paths.split(",") if w]
    unknown = set(only) - {w[0] for w in PATH_WORKLOADS}
    if unknown:
        sys.exit("unknown path workload(s): %s" % ", ".join(sorted(unknown)))
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (a frame would shard "
              "by pixel tiles or iteration ranges, ray batches by ranges). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if not args.no_cpu_baseline and not args.guide_inference_only and not args.traffic_probe \
            and not args.guide_sweep and not args.guided_only:
        use_native_oracle()

    import torch
    from gammaray_amd import scenes
    from gammaray_amd.api import GR_OCCLUSION_ANYHIT, GR_OCCLUSION_EXACT, Context, Stage

    mode = GR_OCCLUSION_EXACT if args.occlusion == "exact" else GR_OCCLUSION_ANYHIT
    torch.cuda.init()
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    if args.guide_inference_only:
        from gammaray_amd.api import Guide
        ctx = Context(0)
        ctx.set_stream(stream.cuda_stream)
        print(json.dumps(guide_inference_rate(ctx, Guide(ctx))), flush=True)
        return
    if args.guided_only:
        ctx = Context(0)
        ctx.set_stream(stream.cuda_stream)
        print(json.dumps(bench_guided(ctx, ref_spp=args.ref_spp, overrides=_parse_overrides(args.guided_config),
                                      scene=args.guided_scene)),
              flush=True)
        return
    if args.guide_sweep:
        ctx = Context(0)
        ctx.set_stream(stream.cuda_stream)
        only = [w for w in args.sweep_only.split(",") if w] if args.sweep_only else None
        print(json.dumps(bench_guide_sweep(ctx, only=only)), flush=True)
        return
    t0 = time.time()
    spec = scenes.icosphere_scene(level=args.level, seed=2)
    log("icosphere L%d: %d triangles (%.1f s)" % (args.level, spec.num_components, time.time() - t0))
    ctx = Context(0)
    ctx.set_stream(stream.cuda_stream)
    stage = Stage(spec, ctx=ctx)
    info = stage.info()
    log("BVH: %d nodes, depth %d, built in %.2f s" % (info["num_nodes"], info["max_depth"],
                                                       info["build_seconds"]))
    n = args.rays
    t0 = time.time()
    rays8 = scenes.ray_batch(spec, n, seed=20)
    segs8 = scenes.segment_batch(spec, n, seed=21)
    log("ray batches generated (%.1f s)" % (time.time() - t0))
    rays = torch.from_numpy(rays8).to(dev)
    segs = torch.from_numpy(segs8).to(dev)
    hits = torch.empty((n, 4), dtype=torch.int32, device=dev)
    vis = torch.empty((n,), dtype=torch.int32, device=dev)

    # Traversal settings: chosen before the timed region.
    if args.tuning == "auto":
        t0 = time.time()
        tuning = stage.autotune(rays, segs, mode)  # the step's own mix: closest hit + visibility
        log("autotuned %s (%.1f s)" % (tuning, time.time() - t0))
    elif args.tuning == "default":
        tuning = stage.get_tuning()
    else:
        kv = dict(item.split("=") for item in args.tuning.split(","))
        tuning = stage.set_tuning(**{k: int(v) for k, v in kv.items()})

    if args.traffic_probe:
        src = torch.ones(COPY_PROBE_BYTES // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        stage.intersect_ptr(rays.data_ptr(), n, hits.data_ptr())
        stage.occluded_ptr(segs.data_ptr(), n, vis.data_ptr(), mode)
        torch.cuda.synchronize()
        log("traffic probe done")
        return

    # Work counts of the reference-order traversal (untimed pass).
    _, cnt_c = stage.intersect(rays, counts=True)
    _, cnt_s = stage.visibility(segs, mode=mode, counts=True)
    torch.cuda.synchronize()
    cnt_c = cnt_c.cpu().numpy().view(np.uint32)
    cnt_s = cnt_s.cpu().numpy().view(np.uint32)
    bytes_closest = algorithmic_bytes(n, cnt_c, HIT_BYTES)
    bytes_shadow = algorithmic_bytes(n, cnt_s, VIS_BYTES)
    log("per ray: closest %.1f box %.2f tri; shadow %.1f box %.2f tri" % (
        cnt_c[:, 0].mean(), cnt_c[:, 1].mean(), cnt_s[:, 0].mean(), cnt_s[:, 1].mean()))
    del cnt_c, cnt_s

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        stage.intersect_ptr(rays.data_ptr(), n, hits.data_ptr())
        if ev is not None:
            ev[1].record(stream)
        stage.occluded_ptr(segs.data_ptr(), n, vis.data_ptr(), mode)
        if ev is not None:
            ev[2].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    ctx.set_kernel_timing(True)
    ctx.take_kernel_times()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        step(events[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ktimes = ctx.take_kernel_times()
    ctx.set_kernel_timing(False)
    if stage.take_overflow_count() != 0:
        log("warning: traversal stack overflow reported")
    ms_closest = float(np.mean([e[0].elapsed_time(e[1]) for e in events]))
    ms_shadow = float(np.mean([e[1].elapsed_time(e[2]) for e in events]))
    ms_step = wall * 1e3 / args.steps
    value = 2.0 * n / (ms_step * 1e-3) / 1e9
    log("closest %.3f ms, shadow %.3f ms, step %.3f ms" % (ms_closest, ms_shadow, ms_step))

    # Per-launch durations of the traversal kernels alone (HIP events on the
    # stream they run on, recorded by the library around each launch).
    def avg(kind):
        ms, cnt = ktimes[kind]
        return ms / max(cnt, 1)
    k_closest, k_shadow = avg("trace_closest"), avg("trace_shadow")
    sort_ms = ktimes["ray_sort"][0] / args.steps
    log("kernels: closest %.3f ms, shadow %.3f ms, sort %.3f ms per step" % (k_closest, k_shadow, sort_ms))
    dominant = "closest" if k_closest >= k_shadow else "shadow"
    ms_dom = k_closest if dominant == "closest" else k_shadow
    bytes_dom = bytes_closest if dominant == "closest" else bytes_shadow
    achieved = bytes_dom / (ms_dom * 1e-3) / 1e9
    setting = dict(level=args.level, rays=n, **{k: tuning.get(k) for k in
                                                ("sort_bits", "lds_stack", "refill", "waves_per_cu", "xcd_ranges",
                                                 "ray_chunk", "vis_refill")})
    traffic_dom = l2_hit = None
    if setting == MEASURED_TRAFFIC_SETTING and args.occlusion == "exact":
        traffic_dom = int(MEASURED_TRAFFIC[dominant])
        l2_hit = round(MEASURED_L2_HIT[dominant], 4)
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic_dom, l2_hit_rate=l2_hit,
                    kernel="k_trace<%s>" % ("closest" if dominant == "closest" else "occlusion"),
                    algorithmic_bytes_per_launch=int(bytes_dom),
                    kernel_ms=round(ms_dom, 4),
                    # both traversal kernels, whichever of them is the dominant one above
                    closest_kernel_gbs=round(bytes_closest / (k_closest * 1e-3) / 1e9, 1),
                    closest_kernel_ms=round(k_closest, 4),
                    shadow_kernel_gbs=round(bytes_shadow / (k_shadow * 1e-3) / 1e9, 1),
                    shadow_kernel_ms=round(k_shadow, 4),
                    sort_ms_per_step=round(sort_ms, 4),
                    closest_call_ms=round(ms_closest, 4), shadow_call_ms=round(ms_shadow, 4))

    cpu = None
    threads = baseline_threads(host_cpus())
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(spec, rays8, segs8, min(args.cpu_sample, n), threads)
        log("cpu baseline %.4f Grays/s on %d threads" % (cpu["value"], threads))
    del rays, segs, hits, vis, rays8, segs8
    paths = None
    if not args.no_paths:
        paths = bench_paths(ctx, stream, not args.no_cpu_baseline, threads, only=only or None,
                            tune=args.path_tuning == "auto")
    batch = None if args.no_batch else bench_batch(ctx, threads)
    guided = None if args.no_guided else bench_guided(ctx, ref_spp=args.ref_spp,
                                                      overrides=_parse_overrides(args.guided_config))

    out = {
        "metric": "Grays/s (closest-hit + shadow rays, incoherent, 1 GPU)",
        "value": round(value, 4),
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "incoherent ray cast, displaced icosphere L%d" % args.level,
                   "triangles": spec.num_components, "closest_rays": n, "shadow_segments": n,
                   "occlusion_mode": args.occlusion, "tuning": tuning, "bvh_nodes": info["num_nodes"],
                   "bvh_depth": info["max_depth"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "paths": paths,
        "batch": batch,
        "guided": guided,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
