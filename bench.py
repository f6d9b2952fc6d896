"""Benchmark: naive MIS path tracer with environment importance sampling at 1080p.

Workload (the headline configuration): 819,400 triangles = 40 displaced icospheres
(subdivision 5) + a 200-triangle ground, per-object Disney materials, a 2048 x 1024
FP32 sky environment (stored RGB16F), 1920 x 1080, 64 samples per pixel, depth 8,
Sobol sampler, NEE on, Russian roulette off.  One step = one 64-spp frame (64
progressive passes in one launch of the path-tracing kernel) with the scene already
resident in HBM.  Metric: Msamples/s = 1920 * 1080 * 64 / t_step / 1e6.

Also reported: the HBM roofline of the path-tracing kernel (algorithmic bytes per
launch counted by the kernel's counting variant, in the reference's data layout:
36 B per MTBVH row visited, 48 B per leaf triangle test, 108 B per surface
interaction, ...; divided by the kernel's average duration from HIP events on its
stream) and the CPU oracle's throughput on the host cores for a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


class Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")

    def event(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0


def profile_value(name, key):
    """A `key: value` line of a counter summary under profiles/, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def workload(name):
    from zillumgl_amd import scene
    if name == "c1":
        d = scene.config1(256, 256)
        return d, dict(spp=16, depth=5), ("Cornell box, 36 triangles, 256x256, 16 spp, depth 5, naive MIS PT")
    d = scene.config2()
    return d, dict(spp=64, depth=8), ("40 displaced icospheres (subdiv 5) + 200-tri ground = 819,400 tris, "
                                      "per-object Disney, 2048x1024 env map IS, 1920x1080, 64 spp, depth 8, "
                                      "naive MIS path tracer, Sobol")


def cpu_baseline(ps, depth, budget_s=20.0):
    """The CPU oracle (FP32 restatement) on the host cores: pass 0 (1 spp) over whole
    image rows, strided so the sample covers the full frame, until ~budget_s."""
    from oracle import oracle as orc
    # every core this process may run on, up to the 16-core share a one-GPU machine
    # gives each GPU (the rest belong to other GPUs' jobs)
    nproc = len(os.sched_getaffinity(0))
    threads = min(16, nproc)
    prm = orc.params(0, 0, 1, 1, depth)
    W, H = ps.width, ps.height
    t0 = time.perf_counter()
    done = 0
    rows = 0
    order = [y for start in range(8) for y in range(start, H, 8)]  # strided, so a cut sample spans the frame
    for y in order:
        orc.render_pixels(ps, prm, 0, y, W, 1, nthreads=threads)
        done += W
        rows += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": round(done / dt / 1e6, 5), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "nproc": nproc, "cpu_count": os.cpu_count(),
            "sample": f"{rows} strided full-width rows x 1 spp (pass 0) of the same frame = {done} samples, "
                      f"{dt:.1f} s, FP32 oracle (oracle/zgl_oracle.cpp, g++ -O3), {threads} threads"}


def adjoint_bench(hip, steps):
    """Path-replay adjoint on the differentiable-rendering workload: the tessellated
    Cornell box (201,588 triangles, 32 Disney materials, 256 x 128 sky environment),
    512 x 512, 32 spp, depth 6, gradients of an L2 image loss against a target rendered
    at parameters perturbed by N(0, 0.1) (clamped to [0, 1]) w.r.t. 160 material
    parameters and 98,304 environment texel channels.  t_forward is zgl_render of the
    same passes; t_adjoint is zgl_adjoint of passes not rendered before (its own logged
    forward pass, the adjoint replay, the upload of dLoss/dframe and the download of the
    gradients); t_image is zgl_render_replay at the parameters (logged forward pass and
    value replay); t_step is one optimisation step, the image and then the gradients of
    the same passes (the adjoint reuses the image's traversal log).  Each timed call takes
    passes of its own, as an optimisation with fresh samples per step would."""
    import numpy as np
    from zillumgl_amd import scene, render, diff, _lib
    ps = scene.pack(scene.config5())
    ctx = render.Context(0)
    ctx.upload(ps)
    stream = ctx.stream()
    W, H, spp, depth = ps.width, ps.height, 32, 6
    rng = np.random.default_rng(5)
    p0, e0 = diff.scene_parameters(ps)
    pt = np.clip(p0 + rng.normal(0, 0.1, p0.shape), 0, 1).astype(np.float32)

    def timed(fn):
        ms = []
        for _ in range(steps + 1):
            a, b = hip.event(), hip.event()
            hip.record(a, stream)
            fn()
            hip.record(b, stream)
            ms.append(hip.elapsed_ms(a, b))
        return sorted(ms[1:])[len(ms[1:]) // 2]

    t_fwd = timed(lambda: ctx.render(_lib.INTEGRATOR_NAIVE, 0, spp, 1, depth))
    rd = diff.Renderer(ctx, ps, spp, depth)
    target = rd.image(pt, e0)
    img = rd.image(p0, e0)
    d_image = 2.0 * (img - target)
    gm = ge = None
    fresh = iter(range(1, 10000))
    n_mat, env_shape = p0.shape[0], e0.shape[:2]
    d_frame = (np.asarray(d_image, np.float32) / np.float32(spp))

    def adj():
        nonlocal gm, ge
        k = next(fresh) * spp
        gm, ge = ctx.adjoint(d_frame, n_mat, env_shape, k, spp, 1 + k, depth)
    ctx.set_eval_params(p0, e0)
    t_adj = timed(adj)

    def image():
        k = next(fresh) * spp
        ctx.clear()
        ctx.render_replay(k, spp, 1 + k, depth)
    t_img = timed(image)

    def step():
        k = next(fresh) * spp
        ctx.clear()
        ctx.render_replay(k, spp, 1 + k, depth)
        ctx.adjoint(d_frame, n_mat, env_shape, k, spp, 1 + k, depth)
    t_step = timed(step)
    ctx.close()
    samples = W * H * spp
    return {"metric": "adjoint-gradient Msamples/s", "value": round(samples / (t_adj / 1e3) / 1e6, 3),
            "unit": "Msamples/s", "t_adjoint_ms": round(t_adj, 3), "t_forward_ms": round(t_fwd, 3),
            "adjoint_over_forward": round(t_adj / t_fwd, 3), "t_image_ms": round(t_img, 3),
            "t_step_ms": round(t_step, 3), "step_over_forward": round(t_step / t_fwd, 3),
            "gradient_norms": {"materials": float(np.linalg.norm(gm)), "env": float(np.linalg.norm(ge))},
            "workload": "tessellated Cornell box, 201,588 tris, 32 Disney materials, 256x128 env, 512x512, "
                        "32 spp, depth 6, L2 loss vs a target at parameters + N(0, 0.1); 160 material + "
                        "98,304 env parameters"}


def batch_bench(reps):
    """64 scenes (Cornell boxes tessellated to 1k-2M triangles, 128^2-1024^2, 4-64 spp,
    depth 3-10) rendered three ways: (a) per scene and per pass, one zgl_render launch
    each -- the reference's dispatch pattern (NaivePathIntegrator::renderOnePass,
    src/integrator/NaivePath.cpp:77-104, one dispatch per sample pass); (b) per scene, one
    zgl_render launch holding all its passes; (c) one batched persistent launch replayed
    from a HIP graph, after the autotuner picked its scene order, regeneration threshold
    and occupancy.  Frames are bit-identical between the three forms
    (tests/test_gpu_batch.py), so RMSE is equal."""
    from concurrent.futures import ThreadPoolExecutor
    from zillumgl_amd import scene, render, _lib
    descs = scene.config4()
    t0 = time.time()
    with ThreadPoolExecutor(16) as ex:
        packed = list(ex.map(lambda d: scene.pack(d[0]), descs))
    print(f"[bench] batch: 64 scenes packed in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    ctxs = []
    for ps in packed:
        c = render.Context(0)
        c.upload(ps)
        ctxs.append(c)
    passes = [(0, spp, 1, depth) for (_, spp, depth) in descs]
    samples = sum(d.width * d.height * spp for (d, spp, _) in descs)
    tris = sum(int(ps.indices.shape[0]) for ps in packed)

    def sequential():
        for c, (first, n, fc, depth) in zip(ctxs, passes):
            c.render(_lib.INTEGRATOR_NAIVE, first, n, fc, depth)
            c.synchronize()

    def per_pass():
        for c, (first, n, fc, depth) in zip(ctxs, passes):
            for k in range(n):
                c.render(_lib.INTEGRATOR_NAIVE, first + k, 1, fc + k, depth)
            c.synchronize()

    def timed(fn, reps=reps):
        fn()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        return sorted(ts)[len(ts) // 2]
    t_pass = timed(per_pass, 1)
    t_seq = timed(sequential)
    b = render.Batch(0)
    b.set(ctxs, passes)
    t = time.perf_counter()
    b.tune(reps=1)
    t_tune = time.perf_counter() - t
    cfg = b.config()
    b.launch()
    b.synchronize()
    bat = []
    for _ in range(reps):
        t = time.perf_counter()
        b.launch()
        b.synchronize()
        bat.append(time.perf_counter() - t)
    t_bat = sorted(bat)[len(bat) // 2]
    b.close()
    for c in ctxs:
        c.close()
    n_launches = sum(n for (_, n, _, _) in passes)
    return {"metric": "batched Msamples/s", "value": round(samples / t_bat / 1e6, 3), "unit": "Msamples/s",
            "t_batch_ms": round(t_bat * 1e3, 3),
            "t_per_pass_ms": round(t_pass * 1e3, 3), "per_pass_launches": n_launches,
            "per_pass_value": round(samples / t_pass / 1e6, 3),
            "speedup_vs_per_pass": round(t_pass / t_bat, 3),
            "t_sequential_ms": round(t_seq * 1e3, 3),
            "sequential_value": round(samples / t_seq / 1e6, 3), "speedup_vs_sequential": round(t_seq / t_bat, 3),
            "tuned": {"regen_threshold": cfg["regen_threshold"], "blocks_per_cu": cfg["blocks_per_cu"],
                      "order_is_work_descending": bool(list(cfg["order"]) != list(range(len(ctxs)))),
                      "tuning_s": round(t_tune, 1)},
            "workload": f"64 Cornell-box scenes, {tris:,} triangles in total (1k-2M each, log-uniform), "
                        f"128^2-1024^2, spp 4-64, depth 3-10, {samples / 1e6:.1f} M samples per launch; "
                        "frames bit-identical between the three forms; per_pass = one persistent launch per "
                        "scene and pass "
                        "(the reference's dispatch pattern), sequential = one persistent launch per scene"}


def integrators_bench(hip, reps):
    """The four camera/light integrators on the caustics workload (Cornell box + 4 small
    emissive triangles + a 10k-triangle IOR-1.5 sphere, 1024 x 1024, depth 8): one call
    of 16 passes each, median of reps.  Naive and global-queue tracers count camera
    samples; the light tracer counts light paths (64 blocks x 1536 per pass, the
    reference's LightPathIntegrator dispatch shape); the triple tracer counts camera
    samples, and runs 64 x 1536 light paths per pass besides."""
    from zillumgl_amd import scene, render, _lib
    ps = scene.pack(scene.config3())
    ctx = render.Context(0)
    ctx.upload(ps)
    st = ctx.stream()
    W, H, passes, depth, blocks = ps.width, ps.height, 16, 8, 64
    out = {"workload": f"Cornell box + 4 emissive triangles + 10k-triangle dielectric sphere, "
                       f"{int(ps.indices.shape[0]):,} tris, {W}x{H}, {passes} passes per call, depth {depth}"}
    for name, integ, unit_count in (("naive", _lib.INTEGRATOR_NAIVE, W * H),
                                    ("global_queue", _lib.INTEGRATOR_GLOBAL_QUEUE, W * H),
                                    ("light", _lib.INTEGRATOR_LIGHT, blocks * 1536),
                                    ("triple", _lib.INTEGRATOR_TRIPLE, W * H)):
        ms = []
        for r in range(reps + 1):
            ctx.clear()
            a, b = hip.event(), hip.event()
            hip.record(a, st)
            ctx.render(integ, 0, passes, 1, depth, blocks_one_pass=blocks, loops_per_pass=1)
            hip.record(b, st)
            t = hip.elapsed_ms(a, b)
            if r > 0:
                ms.append(t)
        t = sorted(ms)[len(ms) // 2]
        key = "light_Mpaths_s" if integ == _lib.INTEGRATOR_LIGHT else f"{name}_Msamples_s"
        out[key] = round(unit_count * passes / t / 1e3, 2)
        out[f"{name}_ms"] = round(t, 3)
        if integ != _lib.INTEGRATOR_GLOBAL_QUEUE:
            # algorithmic bytes of the call (reference layout) from the counting variants
            ctx.clear()
            ctx.reset_counters()
            ctx.render(integ, 0, passes, 1, depth, blocks_one_pass=blocks, loops_per_pass=1, count_bytes=True)
            ctx.synchronize()
            nbytes = float(ctx.counters()[0])
            out[f"{name}_bytes_per_unit"] = round(nbytes / (unit_count * passes), 1)
            out[f"{name}_algorithmic_GBps"] = round(nbytes / (t / 1e3) / 1e9, 1)
    ctx.close()
    return out


def bvh_bench(reps):
    """GPU MTBVH build (zgl_bvh_builder_build) of the largest scene of the batched workload
    (Cornell box tessellated to 1,319,508 triangles), wall clock per build including its
    per-level synchronisations, against the library's single-threaded host build of the
    same mesh; the two outputs are compared bit for bit."""
    import numpy as np
    from zillumgl_amd import scene, render, _lib
    descs = scene.config4()
    sizes = [sum(m.indices.shape[0] for m, _ in d.objects) + sum(m.indices.shape[0] for m, _ in d.lights)
             for d, _, _ in descs]
    ps = scene.pack(descs[int(np.argmax(sizes))][0], bvh="device")
    v, idx = ps.vertices, ps.indices.astype(np.uint32)
    n = idx.shape[0]
    b = render.BvhBuilder(0)
    b.upload(v, idx)
    b.build(0)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        b.build(0)
        ts.append(time.perf_counter() - t0)
    db, dt = b.download()
    st = b.stats()
    b.close()
    t0 = time.perf_counter()
    hb, ht = _lib.build_mtbvh(v, idx)
    host_s = time.perf_counter() - t0
    ms = float(np.median(ts)) * 1e3
    return {"metric": "MTBVH build throughput", "value": round(n / ms / 1e3, 1), "unit": "Mtriangles/s",
            "workload": f"largest batched-workload scene, {n} triangles", "triangles": n,
            "device_ms": round(ms, 3), "host_ms": round(host_s * 1e3, 1), "host_threads": 1,
            "speedup_vs_host": round(host_s * 1e3 / ms, 1),
            "identical": bool(np.array_equal(hb.view(np.uint32), db.view(np.uint32)) and np.array_equal(ht, dt)),
            "levels": st["levels"], "one_wave_subtrees": st["small_subtrees"], "depth": st["depth"]}


def rmse_vs_converged(ctx, W, H, spp, depth, ref_spp, precision="f64"):
    """RMSE of the benchmark image (mean of passes 0..spp-1) against a converged image of
    ref_spp further passes (spp..spp+ref_spp-1, independent of the image's samples) of
    the same estimator: by default rendered in FP64 (zgl_render_reference_f64, the FP64
    instance of the device library, equal to the oracle's FP64 mode per pixel,
    tests/test_gpu_reference64.py), or with precision "f32" by the FP32 kernel, summed
    per launch of 64 passes into FP64 on the host.  The reference's own noise adds
    spp / ref_spp of the image's variance to the squared RMSE; the value is corrected
    for it."""
    import numpy as np
    from zillumgl_amd import _lib
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_NAIVE, 0, spp, 1, depth)
    img = ctx.read_frame()[..., :3].astype(np.float64) / spp
    acc = np.zeros((H, W, 3), np.float64)
    chunk = 64
    t0 = time.time()
    for first in range(spp, spp + ref_spp, chunk):  # passes disjoint from the image's
        if precision == "f64":
            ctx.render_reference_f64(first, chunk, 1 + first, depth, rgb_sum=acc)
        else:
            ctx.clear()
            ctx.render(_lib.INTEGRATOR_NAIVE, first, chunk, 1 + first, depth)
            acc += ctx.read_frame()[..., :3]
        done = first + chunk - spp
        if done % (chunk * (4 if precision == "f64" else 128)) == 0:
            print(f"[bench] {precision} reference passes {done}/{ref_spp} ({time.time() - t0:.0f} s)",
                  file=sys.stderr, flush=True)
    ref = acc / ref_spp
    mse = float(((img - ref) ** 2).mean())
    mse_corrected = mse / (1.0 + spp / ref_spp)
    return {"reference_spp": ref_spp, "reference": "FP64 (zgl_render_reference_f64)" if precision == "f64"
            else "FP32 kernel, FP64 sums", "reference_s": round(time.time() - t0, 1),
            "rmse": round(mse_corrected ** 0.5, 6), "rmse_raw": round(mse ** 0.5, 6),
            "mean_radiance": round(float(ref.mean()), 6),
            "relative_rmse": round(mse_corrected ** 0.5 / float(ref.mean()), 6),
            "note": "same estimator and Sobol indices on the CPU oracle (frames equal bit for bit), so "
                    "the CPU baseline's RMSE is this one; the FP64 reference equals the oracle's FP64 mode per "
                    "pixel (tests/test_gpu_reference64.py)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-bvh", action="store_true")
    ap.add_argument("--no-integrators", action="store_true")
    ap.add_argument("--bvh-only", action="store_true", help="only the MTBVH build line (profiling)")
    ap.add_argument("--reference-spp", type=int, default=1024, metavar="N",
                    help="RMSE against an N-spp converged image (N a multiple of 64; 0 = skip)")
    ap.add_argument("--reference-precision", choices=["f64", "f32"], default="f64",
                    help="the converged image's arithmetic: the FP64 reference renderer (default) or the FP32 kernel")
    ap.add_argument("--threshold", type=int, default=0, help="persistent schedule regeneration threshold (0 = default)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES (PMC calibration) and exit")
    ap.add_argument("--only", choices=["adjoint", "batch", "integrators"],
                    help="only run that secondary line and print it (profiling)")
    args = ap.parse_args()
    if args.bvh_only:
        print(json.dumps(bvh_bench(max(args.steps, 5))))
        return
    if args.only:
        line = {"adjoint": lambda: adjoint_bench(Hip(), max(args.steps, 3)),
                "batch": lambda: batch_bench(max(args.steps, 3)),
                "integrators": lambda: integrators_bench(Hip(), max(args.steps, 3))}[args.only]()
        print(json.dumps(line))
        return
    if args.copy_probe:
        import torch
        src = torch.ones(args.copy_probe // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the library is measured on a single MI355X); --gpus {args.gpus} "
              f"is not supported", file=sys.stderr)
        sys.exit(2)

    from zillumgl_amd import scene, render, _lib
    desc, cfg, text = workload(args.workload)
    t0 = time.time()
    ps = scene.pack(desc)
    print(f"[bench] scene packed in {time.time() - t0:.1f}s: {ps.indices.shape[0]} triangles", file=sys.stderr,
          flush=True)
    ctx = render.Context(0)
    ctx.upload(ps)
    hip = Hip()
    stream = ctx.stream()
    W, H, spp, depth = ps.width, ps.height, cfg["spp"], cfg["depth"]

    def step():
        ctx.render(_lib.INTEGRATOR_NAIVE, 0, spp, 1, depth, regen_threshold=args.threshold)

    for i in range(args.warmup):
        step()
    ctx.synchronize()
    print(f"[bench] warmup done", file=sys.stderr, flush=True)

    evs = [(hip.event(), hip.event()) for _ in range(args.steps)]
    hip.sync()
    wall0 = time.perf_counter()
    for (a, b) in evs:
        hip.record(a, stream)
        step()
        hip.record(b, stream)
    hip.sync()
    wall = time.perf_counter() - wall0
    kernel_ms = [hip.elapsed_ms(a, b) for (a, b) in evs]
    ms_per_step = wall * 1000.0 / args.steps
    samples = W * H * spp
    value = samples * args.steps / wall / 1e6
    print(f"[bench] {args.steps} steps: {ms_per_step:.1f} ms/step (events {sum(kernel_ms) / len(kernel_ms):.1f} ms)",
          file=sys.stderr, flush=True)

    rmse = None
    if args.reference_spp:
        rmse = rmse_vs_converged(ctx, W, H, spp, depth, args.reference_spp, args.reference_precision)
        print(f"[bench] RMSE vs {args.reference_spp} spp: {rmse['rmse']}", file=sys.stderr, flush=True)

    roofline = None
    if not args.no_count:
        ctx.reset_counters()
        ctx.render(_lib.INTEGRATOR_NAIVE, 0, spp, 1, depth, count_bytes=True, regen_threshold=args.threshold)
        ctx.synchronize()
        cnt = ctx.counters()
        algo_bytes = float(cnt[0])
        avg_s = (sum(kernel_ms) / len(kernel_ms)) / 1000.0
        achieved = algo_bytes / avg_s / 1e9
        traffic = profile_value(f"naive_{args.workload}_traffic.md", "hbm_bytes_per_launch")
        lanes = profile_value(f"naive_{args.workload}_valu.md", "active_lanes_per_valu")
        busy = profile_value(f"naive_{args.workload}_valu.md", "valu_issue_busy")
        fabric = traffic / avg_s / 1e9 if traffic else None
        # achieved/frac: the algorithmic bytes of the reference's data layout; L2 serves
        # much of them, so the measured L2-to-HBM traffic is reported beside them, and
        # the measured limiter (VALU issue, lane use) from the counter profile
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "achieved_is": "algorithmic bytes per launch / event-timed launch",
                    "fabric_GBps": round(fabric, 1) if fabric else None,
                    "fabric_frac": round(fabric / HBM_PEAK_GBS, 4) if fabric else None,
                    "limiter": (f"VALU issue ({busy:.0%} busy, {lanes:.1f} of 64 lanes active per VALU instruction)"
                                if lanes and busy else None),
                    "algorithmic_bytes_per_launch": algo_bytes,
                    "bytes_per_sample": round(algo_bytes / samples, 1),
                    "rows_per_traversal": round(float(cnt[1]) / max(float(cnt[3]), 1.0), 2),
                    "leaf_tests_per_traversal": round(float(cnt[2]) / max(float(cnt[3]), 1.0), 2),
                    "traversals_per_sample": round(float(cnt[3]) / samples, 3),
                    "kernel": "k_naive_persistent<false, false> (zillumgl_amd/csrc/zgl_persistent.hip)"}
    ctx.close()

    adjoint = None
    if not args.no_adjoint and args.workload == "c2":
        adjoint = adjoint_bench(hip, max(args.steps, 3))
        print(f"[bench] adjoint {adjoint['value']} Msamples/s, {adjoint['adjoint_over_forward']}x forward; "
              f"image {adjoint['t_image_ms']} ms, step {adjoint['t_step_ms']} ms ({adjoint['step_over_forward']}x forward)",
              file=sys.stderr, flush=True)

    batch = None
    if not args.no_batch and args.workload == "c2":
        batch = batch_bench(max(args.steps, 3))
        print(f"[bench] batch {batch['value']} Msamples/s, "
              f"{batch['speedup_vs_per_pass']}x per-pass launches, "
              f"{batch['speedup_vs_sequential']}x per-scene persistent launches",
              file=sys.stderr, flush=True)

    bvh = None
    if not args.no_bvh and args.workload == "c2":
        bvh = bvh_bench(max(args.steps, 5))
        print(f"[bench] MTBVH build {bvh['device_ms']} ms ({bvh['value']} Mtriangles/s), host {bvh['host_ms']} ms, "
              f"identical {bvh['identical']}", file=sys.stderr, flush=True)

    integ = None
    if not args.no_integrators and args.workload == "c2":
        integ = integrators_bench(hip, max(args.steps, 3))
        print(f"[bench] integrators (caustics workload): {integ}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(ps, depth)
        print(f"[bench] cpu baseline {cpu['value']} Msamples/s", file=sys.stderr, flush=True)

    out = {
        "metric": "Msamples/s (1080p, depth 8, equal RMSE)" if args.workload == "c2" else "Msamples/s",
        "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded generator, zillumgl_amd/scene.py)",
        "config": {"workload": text, "width": W, "height": H, "spp": spp, "depth": depth,
                   "triangles": int(ps.indices.shape[0]),
                   "equal_rmse": "GPU and CPU run the same estimator on the same Sobol indices; frames are "
                                 "equal bit for bit (tests/test_gpu_parity.py)"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "adjoint": adjoint,
        "batch": batch,
        "bvh_build": bvh,
        "integrators": integ,
    }
    if rmse:
        out["rmse"] = rmse
    print(json.dumps(out))


if __name__ == "__main__":
    main()
