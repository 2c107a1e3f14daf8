#!/usr/bin/env python3
"""Benchmark of the MI355X render path; prints ONE JSON line.

Workloads (--workload):
  path : (default) plain path tracer, 512x512, 64 spp, max depth 8: Lambert
         ground, 327,680-triangle displaced icosphere as a GGX rough conductor
         (alpha 0.3, ior 0.27+2.78i), 1x1 area light of radiance 10, 1024x512
         sky + sun env map importance-sampled through an FP64-built 2D CDF.
         One step = one launch of all 64 samples of every pixel (16.8 M paths).
  bsdf : rough-conductor BSDF operator, 2^22 independent FP32 queries
         (PCG32 seed 42 / stream 54), complex IOR 0.27+2.78i.
  batch: 64 uneven renders of the scene family, autotuned wavefront vs one
         megakernel launch per render.
  diff : inverse rendering step (primal + L2 loss + path-replay adjoint +
         texture update) over 8 ring views at 256x256, 16 + 16 spp.
  sss  : subsurface + multiple scattering, 512x512, 128 spp: whole-milk
         displaced icosphere (1,310,720 triangles, standard dipole, 2^20
         Poisson-disk probes, 4 gather samples per hit) beside a GGX rough
         glass sphere with the microsurface multiple-scattering walk.  One step
         = the probe pass plus one launch of all 128 samples (33.6 M paths).

A step is one pass of the hot path over one batch of synthetic input that is
already resident in HBM.  The timed region is bracketed by
torch.cuda.synchronize() on both sides; per-launch kernel time comes from HIP
events recorded on the stream the kernel is launched on.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
# HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE passes of these exact
# workloads, scaled by a 1 GiB copy kernel (profiles/hbm_traffic_pmc.md)
TRAFFIC_PATH_512_64SPP = 16.17e9
TRAFFIC_SSS_512_128SPP = 133.3e9
TRAFFIC_BSDF_2E22 = 218.4e6


def _cpu_threads():
    """Threads for the CPU baseline: the host cores this process may run on,
    or the machine's CPU share when OMP_NUM_THREADS states it."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def _host_cores():
    return {"nproc": os.cpu_count() or 0, "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
            else None, "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# --------------------------------------------------------------------- bsdf
def bench_bsdf(args):
    import torch

    from open_rendering_framework_amd import rt

    n = 1 << 22
    host = rt.gen_bsdf_queries(n, 42, 54)
    dev = [t.cuda() for t in host]
    out = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(7)]
    ior_re, ior_im = [0.27] * 3, [2.78] * 3
    for _ in range(args.warmup):
        rt.bsdf_ggx_conductor_batch(*dev, ior_re, ior_im, out=out)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for s in range(args.steps):
        ev[s][0].record()
        rt.bsdf_ggx_conductor_batch(*dev, ior_re, ior_im, out=out)
        ev[s][1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = sum(a.elapsed_time(b) for a, b in ev) / args.steps
    bytes_per_query = 52  # 24 B in (wo, alpha, u1, u2) + 28 B out (wi, pdf, RGB weight)
    achieved = bytes_per_query * n / (kms * 1e-3) / 1e9
    value = n * args.steps / wall / 1e6
    line = {
        "metric": "Mqueries/s, rough-conductor BSDF operator (2^22 FP32 queries)",
        "value": value,
        "unit": "Mqueries/s",
        "ms_per_step": wall * 1e3 / args.steps,
        "dtype": "f32",
        "config": {"workload": "ggx_conductor_operator", "queries": n, "ior": "0.27+2.78i",
                   "inputs": "PCG32 seed 42 seq 54"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_BSDF_2E22,
                     "kernel": "k_ggx_conductor", "kernel_ms": kms, "bytes_per_unit": bytes_per_query,
                     "algorithmic_bytes_per_launch": bytes_per_query * n},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_bsdf(host)
    return line


def cpu_baseline_bsdf(host):
    from tests.oracle_api import Oracle

    orc = Oracle()
    threads = _cpu_threads()
    ins = [t.numpy() for t in host]
    done, t0 = 0, time.perf_counter()
    while True:
        orc.bsdf_ggx_conductor(*ins, [0.27] * 3, [2.78] * 3, threads=threads)
        done += len(ins[0])
        el = time.perf_counter() - t0
        if el > 10.0:
            break
    return {"value": done / el / 1e6, "unit": "Mqueries/s", "cores": threads, "kind": "port", "host": _host_cores(),
            "sample": "%d x 2^22 operator queries (oracle FP32 restatement, OpenMP %d threads, %s)"
                      % (done >> 22, threads, _cpu_model())}


# --------------------------------------------------------------------- path
PATH_W = PATH_H = 512
PATH_SPP = 64


def bench_path(args):
    import torch

    from open_rendering_framework_amd import rt, scenes

    desc = scenes.path_scene(width=PATH_W, height=PATH_H)
    ctx = scenes.to_context(desc)
    stats = ctx.accel_build()
    stream = torch.cuda.current_stream()
    for w in range(max(1, args.warmup)):
        ctx.launch(PATH_W, PATH_H, 0, PATH_SPP, stream=stream)
    torch.cuda.synchronize()
    ctx.film_clear()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        ev[s][0].record(stream)
        ctx.launch(PATH_W, PATH_H, 0, PATH_SPP, stream=stream)
        ev[s][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = sum(a.elapsed_time(b) for a, b in ev) / args.steps
    c = ctx.counters()
    paths = PATH_W * PATH_H * PATH_SPP
    assert c["paths"] == paths * args.steps, c
    n_ext = c["ext_rays"] / c["paths"]
    n_shd = c["shadow_rays"] / c["paths"]
    # algorithmic bytes per path: extension segment 160 B (ray 32 B w+r, hit 16 B
    # w+r, path state 32 B r+w), shadow segment 72 B (ray w+r + result), FP64 RGBA
    # film 32 B per pixel shared by its spp paths
    b_path = n_ext * 160.0 + n_shd * 72.0 + 32.0 / PATH_SPP
    achieved = b_path * paths / (kms * 1e-3) / 1e9
    value = paths * args.steps / wall / 1e6
    line = {
        "metric": "Mpaths/s at 512^2 64spp (plain megakernel path tracer)",
        "value": value,
        "unit": "Mpaths/s",
        "ms_per_step": wall * 1e3 / args.steps,
        "dtype": "f32",
        "config": {"workload": "plain_path_tracer_512x512_64spp", "width": PATH_W, "height": PATH_H,
                   "spp": PATH_SPP, "max_depth": 8, "triangles": int(stats["tris"]),
                   "env": "1024x512 sky+sun, importance sampled", "form": "megakernel"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_PATH_512_64SPP,
                     "algorithmic_bytes_per_launch": b_path * paths, "kernel": "k_render_mega",
                     "kernel_ms": kms, "bytes_per_unit": b_path, "n_ext_per_path": n_ext,
                     "n_shadow_per_path": n_shd},
        "bvh": stats,
        "nonfinite_samples": c["nonfinite"],
    }
    if not args.no_batch:
        line["batched"] = batch_measure(1, 1)
    if not args.no_diff:
        line["differentiable"] = diff_measure(5, 1)
    if not args.no_sss:
        line["subsurface"] = sss_measure(2, 1)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_path(desc)
    return line


def cpu_baseline_path(desc):
    from tests.oracle_scene import OracleScene

    threads = _cpu_threads()
    orc = OracleScene(desc)
    frames, paths, t0 = 0, 0, time.perf_counter()
    while True:
        orc.render(PATH_W, PATH_H, frames, 1, threads=threads)
        frames += 1
        paths += PATH_W * PATH_H
        el = time.perf_counter() - t0
        if el > 10.0:
            break
    return {"value": paths / el / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port", "host": _host_cores(),
            "sample": "%d frames of 512x512 at 1 spp of the same scene and seeds (oracle FP32 restatement, "
                      "OpenMP %d threads, %s)" % (frames, threads, _cpu_model())}


# --------------------------------------------------------------------- batch
def batch_measure(steps, warmup, spp_scale=1.0, size_scale=1.0):
    """64 uneven renders of the scene family: autotuned wavefront vs one
    megakernel launch per render (same scene, seeds and material tables)."""
    import torch

    from open_rendering_framework_amd import rt, scenes

    desc, items, remap = scenes.batch_workload(spp_scale=spp_scale, size_scale=size_scale)
    ctx = scenes.to_context(desc)
    ctx.accel_build()
    paths = sum(it["width"] * it["height"] * it["spp"] for it in items)
    best, report = ctx.autotune(items, remap, spp_divisor=16)
    for _ in range(max(1, warmup)):
        ctx.render_batch(items, remap, mode=rt.BATCH_WAVEFRONT, config=best)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.render_batch(items, remap, mode=rt.BATCH_WAVEFRONT, config=best)
    torch.cuda.synchronize()
    wf = (time.perf_counter() - t0) / steps
    c = ctx.batch_counters()
    mk = ctx.render_batch(items, remap, mode=rt.BATCH_MEGAKERNEL)
    torch.cuda.synchronize()
    ctx.render_batch(items, remap, mode=rt.BATCH_PERSISTENT)  # warm
    pe = min(ctx.render_batch(items, remap, mode=rt.BATCH_PERSISTENT) for _ in range(max(1, steps)))
    torch.cuda.synchronize()
    return {
        "renders": len(items), "paths": paths,
        "wavefront_mpaths_s": paths / wf / 1e6, "wavefront_s": wf,
        "per_render_megakernel_mpaths_s": paths / mk / 1e6, "per_render_megakernel_s": mk,
        "speedup_vs_per_render_megakernel": mk / wf,
        "persistent_mpaths_s": paths / pe / 1e6, "persistent_s": pe,
        "persistent_speedup_vs_per_render_megakernel": mk / pe,
        "config": best.as_dict(), "autotune": report,
        "ext_rays_per_path": c["ext_rays"] / paths, "shadow_rays_per_path": c["shadow_rays"] / paths,
    }


def bench_batch(args):
    b = batch_measure(args.steps, args.warmup)
    line = {
        "metric": "Mpaths/s, 64 uneven renders in one batched wavefront pass",
        "value": b["wavefront_mpaths_s"],
        "unit": "Mpaths/s",
        "ms_per_step": b["wavefront_s"] * 1e3,
        "dtype": "f32",
        "config": {"workload": "batched_wavefront_64_renders", "renders": b["renders"], "paths": b["paths"]},
        "batched": b,
    }
    return line


# --------------------------------------------------------------------- diff
DIFF_SIZE, DIFF_VIEWS, DIFF_SPP, DIFF_TEX = 256, 8, 16, 64


def diff_setup(size=DIFF_SIZE, views_n=DIFF_VIEWS, spp=DIFF_SPP, tex=DIFF_TEX, level=7):
    """Inverse-rendering workload: the path-tracer scene with a roughness texture on
    the conductor sphere and an albedo texture on the ground, 8 ring views, targets
    rendered from the seed-99 ground-truth textures; the optimisation starts from
    constant textures (albedo 0.5, alpha 0.3)."""
    import torch

    from open_rendering_framework_amd import diff, scenes

    desc = scenes.path_scene(width=size, height=size, level=level)
    ctx = scenes.to_context(desc)
    ctx.accel_build()
    views = scenes.ring_views(n=views_n, size=size, spp=spp)
    R = diff.DiffRenderer(ctx, views, scenes.DIFF_ALBEDO_MATERIAL, scenes.DIFF_ROUGHNESS_MATERIAL, (tex, tex),
                          scenes.DIFF_ALBEDO_UV, scenes.DIFF_ROUGHNESS_UV)
    a_gt, r_gt = scenes.diff_textures(tex, seed=99, rough_range=(0.05, 0.8))
    with torch.no_grad():
        target = R(torch.from_numpy(a_gt).cuda(), torch.from_numpy(r_gt).cuda())
    alb = torch.full((tex, tex, 3), 0.5, device="cuda", requires_grad=True)
    rough = torch.full((tex, tex), 0.3, device="cuda", requires_grad=True)
    return ctx, R, target, alb, rough


def diff_step(R, target, alb, rough, lr_a=0.5, lr_r=0.5):
    """One gradient step: primal render, L2 image loss, path-replay adjoint, update."""
    import torch

    img = R(alb, rough)
    loss = ((img - target) ** 2).mean()
    loss.backward()
    with torch.no_grad():
        alb -= lr_a * alb.grad
        rough -= lr_r * rough.grad
        alb.clamp_(0.02, 0.98)
        rough.clamp_(0.05, 0.8)
        alb.grad = None
        rough.grad = None
    return loss


def diff_measure(steps, warmup):
    import torch

    ctx, R, target, alb, rough = diff_setup()
    losses = []
    for _ in range(max(1, warmup)):
        losses.append(float(diff_step(R, target, alb, rough)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = diff_step(R, target, alb, rough)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    losses.append(float(loss))
    checks = ctx.diff_counters()
    R.close()
    paths = DIFF_SIZE * DIFF_SIZE * DIFF_SPP * DIFF_VIEWS
    return {"grad_steps_per_s": steps / wall, "ms_per_step": wall * 1e3 / steps, "views": DIFF_VIEWS,
            "size": DIFF_SIZE, "spp_primal": DIFF_SPP, "spp_adjoint": DIFF_SPP, "texture": DIFF_TEX,
            "primal_paths_per_step": paths, "adjoint_paths_per_step": paths,
            "adjoint_replay_mismatches": checks["replay_mismatch"], "adjoint_nonfinite": checks["nonfinite"],
            "loss_first": losses[0], "loss_last": losses[-1], "steps": steps}


def bench_diff(args):
    d = diff_measure(args.steps, args.warmup)
    return {
        "metric": "adjoint grad-steps/s (8 views 256^2, 16+16 spp, 64x64 albedo + roughness textures)",
        "value": d["grad_steps_per_s"],
        "unit": "grad-steps/s",
        "ms_per_step": d["ms_per_step"],
        "dtype": "f32",
        "config": {"workload": "differentiable_inverse_rendering", "views": DIFF_VIEWS, "size": DIFF_SIZE,
                   "spp_primal": DIFF_SPP, "spp_adjoint": DIFF_SPP, "texture": DIFF_TEX},
        "differentiable": d,
    }


# --------------------------------------------------------------------- sss
SSS_W, SSS_H, SSS_SPP = 512, 512, 128


def sss_measure(steps, warmup):
    """Subsurface scene at 512x512, 128 spp; per step one probe pass + one launch."""
    import torch

    from open_rendering_framework_amd import scenes

    desc = scenes.subsurface_scene(width=SSS_W, height=SSS_H)
    ctx = scenes.to_context(desc)
    stats = ctx.accel_build()
    pr = ctx.probes()
    stream = torch.cuda.current_stream()
    for _ in range(max(1, warmup)):
        ctx.launch(SSS_W, SSS_H, 0, SSS_SPP, stream=stream)
    torch.cuda.synchronize()
    ctx.film_clear()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        ev[s][0].record(stream)
        ctx.launch(SSS_W, SSS_H, 0, SSS_SPP, stream=stream)
        ev[s][1].record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    kms = sum(a.elapsed_time(b) for a, b in ev) / steps
    c = ctx.counters()
    paths = SSS_W * SSS_H * SSS_SPP
    assert c["paths"] == paths * steps, c
    return {"mpaths_s": paths / wall / 1e6, "s_per_launch": wall, "launch_ms_events": kms, "paths": paths,
            "hbm_traffic_per_launch": TRAFFIC_SSS_512_128SPP, "hbm_gbs": TRAFFIC_SSS_512_128SPP / (kms * 1e-3) / 1e9,
            "spp": SSS_SPP, "triangles": int(stats["tris"]), "probes": pr["count"],
            "ext_rays_per_path": c["ext_rays"] / c["paths"], "shadow_rays_per_path": c["shadow_rays"] / c["paths"],
            "nonfinite_samples": c["nonfinite"]}


def bench_sss(args):
    d = sss_measure(args.steps, args.warmup)
    return {
        "metric": "Mpaths/s at 512^2 128spp (subsurface + multiple scattering)",
        "value": d["mpaths_s"],
        "unit": "Mpaths/s",
        "ms_per_step": d["s_per_launch"] * 1e3,
        "dtype": "f32",
        "config": {"workload": "subsurface_multiscatter_512x512_128spp", "width": SSS_W, "height": SSS_H,
                   "spp": SSS_SPP, "probes": d["probes"], "triangles": d["triangles"]},
        "subsurface": d,
    }


# --------------------------------------------------------------------- copy
def bench_copy(args):
    """Plain 1 GiB device copy (torch copy_): the known-size kernel whose
    FETCH_SIZE / WRITE_SIZE readings scale the counters of the render kernels."""
    import torch

    n = 1 << 28
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    return {"metric": "device copy GB/s (1 GiB, counter calibration)", "value": 2 * 4 * n / dt / 1e9, "unit": "GB/s",
            "ms_per_step": dt * 1e3, "dtype": "f32", "config": {"workload": "copy_1GiB", "bytes_each_way": 4 * n}}


WORKLOADS = {"copy": bench_copy, "path": bench_path, "bsdf": bench_bsdf, "batch": bench_batch, "diff": bench_diff,
             "sss": bench_sss}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="path", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-wavefront measurement")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable-render measurement")
    ap.add_argument("--no-sss", action="store_true", help="skip the subsurface-scene measurement")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only; multi-GPU runs are not supported "
              "(the render path would shard by image tile or sample range)", file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    line = WORKLOADS[args.workload](args)
    base = {"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "data": "synthetic"}
    base.update(line)
    print(json.dumps(base))


if __name__ == "__main__":
    main()
