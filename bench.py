#!/usr/bin/env python3
"""Narrow-band voxel updates/s of Felt's sparse-field level-set update on one MI355X.

Workload (one GPU): a single large surface on the plain GPU path — Surface<3,3>, float32,
512^3 isogrid in 32^3 partitions, a sphere of radius 150 at the centre radially perturbed by
8 real spherical harmonics (degree <= 6, amplitudes U[0,6) voxels, seed 0xFE1702), speed
clamp(-0.3 * (1 + 0.5 * n(x)) + 0.1 * curvature, -1, 1) with n seeded value noise on a
64-voxel lattice; incremental marching cubes of the changed partitions every 10 steps.

A step is one update(fn) over the whole narrow band (plus notify(), and march() every 10th
step).  Voxel updates per step = band size at the start of the step (all 2L+1 layers).  The
timed region starts with the surface resident in HBM.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "narrow-band voxel updates/s, batched+served vs plain on GPU; p99 step latency"
HBM_PEAK = 8.0e12
# algorithmic bytes per list entry of each timed pass
PASSES = {
    0: ("speed (curvature + noise functor)", 12),   # index, phi(x) read, delta write
    1: ("distance pass 1", 12),                     # index, phi(x) read, d write
    2: ("band compaction", 1),                      # one layer byte per pool voxel scanned
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(size=512, part=32, L=3):
    from felt_amd import workloads
    t = time.time()
    phi = workloads.harmonic_sphere_phi((size,) * 3, L, 150.0 * size / 512.0, 0xFE1702,
                                        amp=6.0 * size / 512.0)
    log(f"[bench] workload field {size}^3 built in {time.time() - t:.1f}s")
    return phi


def speed():
    from felt_amd import speed_noise_curv
    return speed_noise_curv(-0.3, 0.1, 0.5, 0xFE1702, 64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--part", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--pass-reps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs a surface on a single MI355X and has "
              "no multi-GPU path (see DESIGN.md for how it would shard).", file=sys.stderr)
        sys.exit(2)
    L = 3
    phi = make_workload(args.size, args.part, L)
    import felt_amd
    felt_amd.lib()
    s = felt_amd.Surface(3, L, (args.size,) * 3, (args.part,) * 3)
    t = time.time()
    s.load_dense(phi)
    s.sync()
    log(f"[bench] loaded onto GPU in {time.time() - t:.1f}s, band {s.stats().band_voxels}")
    polys = felt_amd.Polys(s)
    polys.invalidate()
    polys.march()
    sp = speed()
    mc_every = 10
    for i in range(args.warmup):
        s.update(sp)
        polys.notify()
        if (i + 1) % mc_every == 0:
            polys.march()
    s.sync()
    log(f"[bench] {args.warmup} warmup steps done")

    def run(steps, first, lat, tally):
        for i in range(steps):
            t0 = time.perf_counter()
            s.update(sp)
            polys.notify()
            tally["updates"] += s.step_info()[0]
            if (first + i + 1) % mc_every == 0:
                tm = time.perf_counter()
                polys.march()
                tally["march_s"] += time.perf_counter() - tm
                tally["spx"] += sum(polys.child_info(int(p))[2] for p in polys.changes())
            lat.append(time.perf_counter() - t0)
            if (i + 1) % 50 == 0:
                log(f"[bench] step {i + 1}/{steps}")

    # timed region: each update is one replay of the captured step graph
    lat, tally = [], {"updates": 0, "march_s": 0.0, "spx": 0}
    s.sync()
    t_all = time.perf_counter()
    run(args.steps, args.warmup, lat, tally)
    s.sync()
    total = time.perf_counter() - t_all
    updates, march_s, spx = tally["updates"], tally["march_s"], tally["spx"]
    # kernel durations: each pass's launches of one step, captured as a graph on the surface's
    # stream and replayed between two HIP events on the final state of the timed run
    ktab = {}
    for k, (name, bpi) in PASSES.items():
        us, items, n = s.time_pass(k, args.pass_reps)
        ktab[k] = (name, bpi, n, us, items)
    dom = max((k for k in ktab if k != 2), key=lambda k: ktab[k][2] * ktab[k][3])
    name, bpi, n, us, items = ktab[dom]
    avg_s = us / 1e6
    bytes_per_launch = bpi * items
    achieved = bytes_per_launch / avg_s
    lat_ms = np.array(lat) * 1e3
    final = s.stats()
    result = {
        "metric": METRIC,
        "value": updates / total,
        "unit": "voxel updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"single large surface, plain GPU path: Surface<3,3> {args.size}^3 isogrid "
                        f"in {args.part}^3 partitions, harmonic-perturbed sphere r="
                        f"{150.0 * args.size / 512:g}, speed clamp(-0.3(1+0.5n)+0.1curv), "
                        "incremental marching cubes every 10 steps",
            "size": args.size, "partition": args.part, "layers": L,
            "band_voxels_final": int(final.band_voxels),
        },
        "p50_step_ms": float(np.percentile(lat_ms, 50)),
        "p99_step_ms": float(np.percentile(lat_ms, 99)),
        "triangles_per_s": spx / march_s if march_s > 0 else None,
        "roofline": {
            "bound": "hbm",
            "kernel": name,
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": None,
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_us": avg_s * 1e6,
            "timing": f"the pass's launches of one step on the final state, captured as a "
                      f"HIP graph and replayed {args.pass_reps} times between two HIP events "
                      "on the surface's stream",
        },
        "kernels": {v[0]: {"launches_per_step": v[2], "us_per_launch": v[3],
                           "items_per_launch": v[4]} for v in ktab.values()},
    }
    s.close()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(phi, args.size, args.part, L, args.cpu_seconds)
    print(json.dumps(result), flush=True)


def cpu_baseline(phi, size, part, L, budget_s):
    """The CPU restatement (oracle/, OpenMP over partitions like the reference) on the same
    initial state, for a bounded number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    b = oracle_lib.binding()
    b.lib.felt_oracle_set_threads(cores)
    c = oracle_lib.surface(3, L, (size,) * 3, (part,) * 3)
    c.b.lib.felt_oracle_surface_set_omp(c.h, 1)
    t = time.time()
    c.load_dense(phi)
    log(f"[bench] oracle loaded in {time.time() - t:.1f}s")
    sp = speed()
    done, ups, el = 0, 0, 0.0
    while el < budget_s and done < 50:
        t0 = time.perf_counter()
        c.update(sp)
        el += time.perf_counter() - t0
        ups += int(c.stats().last_voxel_updates)
        done += 1
        log(f"[bench] cpu step {done}: {el:.1f}s")
    return {"value": ups / el, "unit": "voxel updates/s", "cores": cores, "kind": "port",
            "sample": f"first {done} update steps of the same {size}^3 workload from the same "
                      f"initial state ({el:.1f}s of CPU work), CPU restatement with OpenMP "
                      "over partitions"}


if __name__ == "__main__":
    main()
