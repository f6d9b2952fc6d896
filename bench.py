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
    threads = min(16, os.cpu_count() or 1)
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
            "sample": f"{rows} strided full-width rows x 1 spp (pass 0) of the same frame = {done} samples, "
                      f"{dt:.1f} s, FP32 oracle (oracle/zgl_oracle.cpp), {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    ap.add_argument("--threshold", type=int, default=0, help="persistent schedule regeneration threshold (0 = default)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES (PMC calibration) and exit")
    args = ap.parse_args()
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
        print(f"bench.py measures one GPU only (this build runs on a single MI355X); --gpus {args.gpus} "
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

    roofline = None
    if not args.no_count:
        ctx.reset_counters()
        ctx.render(_lib.INTEGRATOR_NAIVE, 0, spp, 1, depth, count_bytes=True, regen_threshold=args.threshold)
        ctx.synchronize()
        cnt = ctx.counters()
        algo_bytes = float(cnt[0])
        avg_s = (sum(kernel_ms) / len(kernel_ms)) / 1000.0
        achieved = algo_bytes / avg_s / 1e9
        traffic = None
        tf = os.path.join(ROOT, "profiles", f"naive_{args.workload}_traffic.md")
        if os.path.exists(tf):
            for line in open(tf):
                if line.startswith("hbm_bytes_per_launch:"):
                    traffic = float(line.split(":")[1])
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "algorithmic_bytes_per_launch": algo_bytes,
                    "bytes_per_sample": round(algo_bytes / samples, 1),
                    "rows_per_traversal": round(float(cnt[1]) / max(float(cnt[3]), 1.0), 2),
                    "leaf_tests_per_traversal": round(float(cnt[2]) / max(float(cnt[3]), 1.0), 2),
                    "traversals_per_sample": round(float(cnt[3]) / samples, 3),
                    "kernel": "k_naive_persistent<false> (zillumgl_amd/csrc/zgl_persistent.hip)"}
    ctx.close()

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
                   "equal_rmse": "GPU and CPU run the same estimator on the same Sobol indices; frames agree "
                                 "per pixel within 1e-4 (tests/test_gpu_parity.py)"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
