#!/usr/bin/env python3
"""Narrow-band voxel updates/s of Felt's sparse-field level-set update on one MI355X:
batched vs plain on the mixed-size workload, plus the single large surface.

Headline workload (mixed sizes): 256 independent Surface<3,3> in float32.  Grid edge
log-uniform in [32, 256] rounded to a multiple of the partition size (8^3 partitions below an
edge of 96, else 16^3), 1-8 spheres of radius U[0.05, 0.3] * edge with uniform centres (phi =
min over spheres), per-surface speed drawn from {-0.4, 0.15 curv, -0.3 + 0.1 curv}, seed
0xFE1703.  A step advances every surface by one update(fn).
  - batched: the surfaces of each partition size form one batch; a step queues one captured
    HIP graph replay per batch (felt_batch_update_async, one stream each, so the two overlap)
    and then waits for both;
  - plain: one felt_surface per surface, update(fn) called surface after surface.
Voxel updates per step = band size at the start of the step (all 2L+1 layers), summed over
surfaces.  `value` is the batched rate; the plain rate and their ratio sit beside it.

Second workload (single large surface): Surface<3,3> 512^3 in 32^3 partitions, sphere r=150
perturbed by 8 spherical harmonics (seed 0xFE1702), speed clamp(-0.3(1+0.5n)+0.1curv, -1, 1),
incremental marching cubes every 10 steps.

The timed regions start with the surfaces resident in HBM.  Prints one JSON line on stdout;
progress goes to stderr.
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
MIXED_SEED = 0xFE1703
# algorithmic bytes per list entry of each timed pass
PASSES = {
    0: ("speed (curvature functor)", 12),   # index, phi(x) read, delta write
    1: ("distance pass 1", 12),             # index, phi(x) read, d write
    2: ("band compaction", 1),              # one layer byte per pool voxel scanned
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pctl(x, q):
    return float(np.percentile(np.asarray(x) * 1e3, q)) if len(x) else None


def mixed_workload(n):
    from felt_amd import workloads
    return workloads.mixed_surfaces(n, MIXED_SEED, 32, 256)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "distance_pass_traffic_batched_16.md")


def dist_traffic_per_entry():
    """(fetch, write) HBM bytes per entry of the first-round distance pass, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counters scaled by a copy of known size), or None
    when the summary is absent."""
    import re
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"Per-entry HBM bytes.*?fetch ([0-9.]+) write ([0-9.]+)", text, re.S)
    return (float(m.group(1)), float(m.group(2))) if m else None


def time_passes(obj, reps, traffic=False):
    """Per-pass timing on the current state; returns (table, roofline of the dominant pass).
    With `traffic` (the mixed workload's 16^3 batch, the one the counter summary measured) the
    roofline carries the HBM bytes per launch."""
    ktab = {}
    for k, (name, bpi) in PASSES.items():
        us, items, n = obj.time_pass(k, reps)
        ktab[k] = (name, bpi, n, us, items)
    dom = max((k for k in ktab if k != 2), key=lambda k: ktab[k][2] * ktab[k][3])
    name, bpi, n, us, items = ktab[dom]
    achieved = bpi * items / (us / 1e6)
    per = dist_traffic_per_entry() if traffic and name == "distance pass 1" else None
    roof = {"bound": "hbm", "kernel": name, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": achieved / HBM_PEAK,
            "traffic": (per[0] + per[1]) * items if per else None,
            "bytes_per_launch": bpi * items, "avg_launch_us": us,
            "timing": f"the pass's launches of one step on the final state, {reps} copies "
                      "captured in one HIP graph and replayed between two HIP events on the "
                      "handle's stream"}
    if per:
        roof["traffic_source"] = (
            f"HBM bytes per launch = entries x ({per[0]:.2f} fetched + {per[1]:.2f} written) B, "
            "per-entry figures from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of bench.py "
            "--traffic-probe scaled by a 256 MiB copy (profiles/"
            "distance_pass_traffic_batched_16.md)")
    table = {v[0]: {"launches_per_step": v[2], "us_per_launch": v[3], "items_per_launch": v[4]}
             for v in ktab.values()}
    return table, roof


def run_mixed(args):
    import felt_amd
    from felt_amd import workloads
    from felt_amd.batch import Batch, Group
    specs = mixed_workload(args.surfaces)
    groups = {}
    for i, sp in enumerate(specs):
        groups.setdefault(sp.part[0], []).append(i)
    lib = felt_amd.lib().lib
    # ---- batched ----
    t = time.time()
    batches = []
    for p, ids in sorted(groups.items()):
        bt = Batch(lib, 3, 3, [specs[i].size for i in ids], (p,) * 3)
        for j, i in enumerate(ids):
            bt.load_spheres(j, specs[i].centres, specs[i].radii)
        speeds = [workloads.speed_of(specs[i].speed) for i in ids]
        batches.append((bt, speeds))
    log(f"[bench] mixed: {len(specs)} surfaces in {len(batches)} batches loaded in "
        f"{time.time() - t:.1f}s")
    for _ in range(args.warmup):
        for bt, speeds in batches:
            bt.update(speeds)
    # launch shapes tuned on the warmed-up state (felt_batch_autotune), then one more untimed
    # step so the step graphs are captured with them
    tuned = {}
    if args.tune_reps > 0:
        t = time.time()
        for (p, _), (bt, speeds) in zip(sorted(groups.items()), batches):
            tuned[f"{p}^3"] = bt.autotune(args.tune_reps)
            bt.update(speeds)
        log(f"[bench] tuned in {time.time() - t:.1f}s: {tuned}")
    # Timed: one graph launch per batch, the batches on their own streams so they overlap.
    # Beside it, the same batches as one group (felt_group_update_async: one graph launch per
    # step); the runtime overlaps two streams better than two independent chains of one graph.
    group = Group([bt for bt, _ in batches]) if args.group_steps > 0 else None
    all_speeds = [speeds for _, speeds in batches]

    def one_step(grouped):
        if grouped:
            group.update_async(all_speeds)
            group.wait()
        else:
            # every batch's step queued (one stream each) before waiting, so they overlap
            for bt, speeds in batches:
                bt.update_async(speeds)
            for bt, _ in batches:
                bt.wait()
        return sum(bt.step_info()[0] for bt, _ in batches)

    lat, ups = [], 0
    t_all = time.perf_counter()
    for step in range(args.steps):
        t0 = time.perf_counter()
        ups += one_step(False)
        lat.append(time.perf_counter() - t0)
        if (step + 1) % 25 == 0:
            log(f"[bench] batched step {step + 1}/{args.steps}")
    total = time.perf_counter() - t_all
    batched = {"value": ups / total, "ms_per_step": total / args.steps * 1e3,
               "p50_step_ms": pctl(lat, 50), "p99_step_ms": pctl(lat, 99),
               "voxel_updates_per_step": ups / args.steps,
               "launch": "one graph launch per batch, the two batches on two streams"}
    if group is not None:
        one_step(True)  # captures the group's graph
        t0 = time.perf_counter()
        ups_g = sum(one_step(True) for _ in range(args.group_steps))
        t_g = time.perf_counter() - t0
        batched["group_launch"] = {"steps": args.group_steps,
                                   "ms_per_step": t_g / args.group_steps * 1e3,
                                   "value": ups_g / t_g,
                                   "launch": "one graph launch per step for both batches"}
        group.close()
    big = max(batches, key=lambda b: b[0].memory()[1])[0]
    ktab, roof = time_passes(big, args.pass_reps, traffic=True)
    mem = [b.memory() for b, _ in batches]
    for bt, _ in batches:
        bt.close()
    cfg = {"workload": f"mixed-size batch: {len(specs)} independent Surface<3,3> float32, edge "
                       "log-uniform 32-256 (8^3 partitions below 96, else 16^3), 1-8 spheres "
                       "r=U[0.05,0.3]*edge, speeds {-0.4, 0.15curv, -0.3+0.1curv}, seed "
                       f"0x{MIXED_SEED:X}; one update(fn) of every surface per step",
           "surfaces": len(specs),
           "batches": {f"{p}^3": len(ids) for p, ids in sorted(groups.items())},
           "launch_config": tuned or "default (256 threads per block, grid cap 4096, "
                                     "band compaction one wavefront per 512 voxels)",
           "device_bytes": int(sum(m[0] for m in mem)),
           "bricks_in_use": int(sum(m[1] for m in mem))}
    if args.plain_steps <= 0:
        return specs, batched, None, ktab, roof, cfg
    # ---- plain: one surface per launch sequence ----
    t = time.time()
    plain = []
    for sp in specs:
        s = felt_amd.Surface(3, 3, sp.size, sp.part)
        s.load_spheres(sp.centres, sp.radii)
        if args.plain_launches:
            s.set_graph(False)
        plain.append((s, workloads.speed_of(sp.speed)))
    log(f"[bench] plain: {len(plain)} surfaces loaded in {time.time() - t:.1f}s")
    for _ in range(args.warmup):
        for s, f in plain:
            s.update(f)
    lat_p, ups_p = [], 0
    t_all = time.perf_counter()
    for step in range(args.plain_steps):
        t0 = time.perf_counter()
        for s, f in plain:
            s.update(f)
            ups_p += s.step_info()[0]
        lat_p.append(time.perf_counter() - t0)
        if (step + 1) % 10 == 0:
            log(f"[bench] plain step {step + 1}/{args.plain_steps}")
    total_p = time.perf_counter() - t_all
    for s, _ in plain:
        s.close()
    plain_res = {"value": ups_p / total_p, "ms_per_step": total_p / args.plain_steps * 1e3,
                 "p50_step_ms": pctl(lat_p, 50), "p99_step_ms": pctl(lat_p, 99),
                 "steps": args.plain_steps, "voxel_updates_per_step": ups_p / args.plain_steps}
    return specs, batched, plain_res, ktab, roof, cfg


def run_single(args):
    import felt_amd
    from felt_amd import speed_noise_curv, workloads
    size, part, L = args.size, args.part, 3
    t = time.time()
    phi = workloads.harmonic_sphere_phi((size,) * 3, L, 150.0 * size / 512.0, 0xFE1702,
                                        amp=6.0 * size / 512.0)
    log(f"[bench] single: field {size}^3 built in {time.time() - t:.1f}s")
    s = felt_amd.Surface(3, L, (size,) * 3, (part,) * 3)
    if args.plain_launches:
        s.set_graph(False)
    s.load_dense(phi)
    del phi
    polys = felt_amd.Polys(s)
    polys.invalidate()
    polys.march()
    sp = speed_noise_curv(-0.3, 0.1, 0.5, 0xFE1702, 64)
    for i in range(args.warmup):
        s.update(sp)
        polys.notify()
    s.sync()
    lat, ups, march_s, spx = [], 0, 0.0, 0
    t_all = time.perf_counter()
    for i in range(args.single_steps):
        t0 = time.perf_counter()
        s.update(sp)
        polys.notify()
        ups += s.step_info()[0]
        if (i + 1) % 10 == 0:
            tm = time.perf_counter()
            polys.march()
            march_s += time.perf_counter() - tm
            spx += sum(polys.child_info(int(p))[2] for p in polys.changes())
        lat.append(time.perf_counter() - t0)
    s.sync()
    total = time.perf_counter() - t_all
    ktab, roof = time_passes(s, args.pass_reps)
    ray = ray_bench(s, size, args.rays) if args.rays > 0 else None
    # checkpoint / resume in the reference's archive format: save the final state, load it into a
    # new surface (host formatting and device gather / scatter included)
    t0 = time.perf_counter()
    data = s.save()
    t_save = time.perf_counter() - t0
    t0 = time.perf_counter()
    s2 = felt_amd.Surface.load(felt_amd.lib(), data)
    s2.sync()
    t_load = time.perf_counter() - t0
    same = s2.layer_size(0) == s.layer_size(0)
    s2.close()
    ckpt = {"archive_bytes": len(data), "save_s": t_save, "load_s": t_load,
            "save_MB_per_s": len(data) / t_save / 1e6, "load_MB_per_s": len(data) / t_load / 1e6,
            "zero_layer_matches": bool(same)}
    del data
    res = {"workload": f"single large surface, plain GPU path: Surface<3,3> {size}^3 isogrid in "
                       f"{part}^3 partitions, harmonic-perturbed sphere r={150.0 * size / 512:g}, "
                       "speed clamp(-0.3(1+0.5n)+0.1curv), incremental marching cubes every 10 "
                       "steps",
           "value": ups / total, "steps": args.single_steps,
           "ms_per_step": total / args.single_steps * 1e3,
           "p50_step_ms": pctl(lat, 50), "p99_step_ms": pctl(lat, 99),
           "triangles_per_s": spx / march_s if march_s > 0 else None,
           "band_voxels_final": int(s.stats().band_voxels),
           "roofline": roof, "kernels": ktab, "ray": ray, "checkpoint": ckpt}
    polys.close()
    s.close()
    log(f"[bench] single: {res['value'] / 1e9:.2f} G voxel updates/s")
    return res


def inside_merged(surface, pool=4):
    """Whether the inside region (phi < 0) is one face-connected component.  A 4x-pooled mask
    (OR over 4^3 blocks) can only join components, so while it has several the fronts have not
    merged; only when it has one is the full-resolution mask labelled."""
    from scipy import ndimage
    inside = surface.snapshot() < 0
    n = inside.shape
    if all(x % pool == 0 for x in n):
        shape = []
        for x in n:
            shape += [x // pool, pool]
        coarse = inside.reshape(shape).any(axis=tuple(range(1, 2 * len(n), 2)))
        if ndimage.label(coarse)[1] != 1:
            return False
    return ndimage.label(inside)[1] == 1


def run_large_3d(args):
    """The large topology-changing workload at full size: Surface<3,3> 1024^3 in 32^3
    partitions, 64 spheres r U[20,60] (seed 0xFE1705), -0.5 until the inside region is one
    face-connected component, then 0.2*curvature, 300 steps of update(fn) + 16 random 64^3
    update(lo,hi,fn) + notify after each + march.  The merge check (host labelling of the
    snapshot) runs every `check` steps outside the timed region.  A step whose round loop
    cycles (the reference's would never end) stops the run; the steps before it are reported."""
    import felt_amd
    from felt_amd import workloads
    from felt_amd._binding import FeltError, speed_constant, speed_curvature
    N, part, L, steps, check = args.large_size, 32, 3, args.large_steps, args.large_check
    rng = workloads.SplitMix64(0xFE1705)
    cs = [tuple(rng.uniform(-N / 2, N / 2) for _ in range(3)) for _ in range(64)]
    rs = [rng.uniform(20, 60) * N / 1024 for _ in range(64)]
    box = 64 * N // 1024
    t = time.time()
    s = felt_amd.Surface(3, L, (N,) * 3, (part,) * 3)
    s.load_spheres(cs, rs)
    # a step whose round loop cycles (the reference's would not end) fails after 1000 rounds
    s.set_round_limit(1000)
    polys = felt_amd.Polys(s)
    polys.invalidate()
    polys.march()
    s.sync()
    log(f"[bench] large 3d: {N}^3 loaded in {time.time() - t:.1f}s, band "
        f"{s.stats().band_voxels}")
    switch, lat, ups, spx, march_s, check_s = None, [], 0, 0, 0.0, 0.0
    stopped, done = None, 0
    t_all = time.perf_counter()
    for step in range(steps):
        sp = speed_constant(-0.5) if switch is None else speed_curvature(0.2)
        boxes = []
        for _ in range(16):
            lo = [rng.randint(-N // 2, N // 2 - box) for _ in range(3)]
            boxes.append((lo, [x + box - 1 for x in lo]))
        t0 = time.perf_counter()
        try:
            s.update(sp)
            ups += s.step_info()[0]
            polys.notify()
            for lo, hi in boxes:
                s.update_box(lo, hi, sp)
                ups += s.step_info()[0]
                polys.notify()
        except FeltError as e:
            stopped = {"step": step + 1, "error": str(e)}
            log(f"[bench] large 3d: step {step + 1} stopped: {e}")
            t_stop = time.perf_counter() - t0
            break
        done = step + 1
        tm = time.perf_counter()
        polys.march()
        march_s += time.perf_counter() - tm
        lat.append(time.perf_counter() - t0)
        spx += sum(polys.child_info(int(p))[2] for p in polys.changes())
        if switch is None and (step + 1) % check == 0:
            tc = time.perf_counter()
            if inside_merged(s):
                switch = step
            check_s += time.perf_counter() - tc
        if (step + 1) % 25 == 0:
            log(f"[bench] large 3d step {step + 1}/{steps}, switch {switch}, band "
                f"{s.stats().band_voxels}")
    total = time.perf_counter() - t_all - check_s - (t_stop if stopped else 0.0)
    ktab, roof = time_passes(s, args.pass_reps)
    res = {"workload": f"Surface<3,3> {N}^3 in {part}^3 partitions, 64 spheres r=U[20,60]*{N}/1024"
                       ", seed 0xFE1705; -0.5 until the inside region is one face-connected "
                       f"component (checked every {check} steps), then 0.2*curvature; per step "
                       f"update(fn), 16 random {box}^3 update(lo,hi,fn), notify after each, "
                       "incremental marching cubes",
           "steps": done, "stopped": stopped, "first_step_ms": lat[0] * 1e3 if lat else None,
           "switch_after_step": None if switch is None else switch + 1,
           "value": ups / total, "ms_per_step": total / max(done, 1) * 1e3,
           "p50_step_ms": pctl(lat, 50), "p99_step_ms": pctl(lat, 99),
           "voxel_updates_per_step": ups / max(done, 1),
           "triangles_per_s": spx / march_s if march_s > 0 else None,
           "band_voxels_final": int(s.stats().band_voxels),
           "roofline": roof, "kernels": ktab}
    polys.close()
    s.close()
    log(f"[bench] large 3d: {res['value'] / 1e9:.2f} G voxel updates/s, "
        f"{res['ms_per_step']:.2f} ms/step, switch after step {res['switch_after_step']}")
    return res


def run_large_2d(args):
    """Surface<2,2> 16384^2 in 64^2 partitions, 512 circles r U[4,64] (seed 0xFE1706), 300
    steps of update(-0.5) with notify and incremental marching squares."""
    import felt_amd
    from felt_amd import workloads
    from felt_amd._binding import speed_constant
    N, part, steps = args.large2d_size, 64, args.large_steps
    rng = workloads.SplitMix64(0xFE1706)
    cs = [tuple(rng.uniform(-N / 2, N / 2) for _ in range(2)) for _ in range(512)]
    rs = [rng.uniform(4, 64) * N / 16384 for _ in range(512)]
    s = felt_amd.Surface(2, 2, (N,) * 2, (part,) * 2)
    s.load_spheres(cs, rs)
    polys = felt_amd.Polys(s)
    polys.invalidate()
    polys.march()
    s.sync()
    sp = speed_constant(-0.5)
    lat, ups, spx, march_s = [], 0, 0, 0.0
    t_all = time.perf_counter()
    for step in range(steps):
        t0 = time.perf_counter()
        s.update(sp)
        ups += s.step_info()[0]
        polys.notify()
        tm = time.perf_counter()
        polys.march()
        march_s += time.perf_counter() - tm
        lat.append(time.perf_counter() - t0)
        spx += sum(polys.child_info(int(p))[2] for p in polys.changes())
        if (step + 1) % 50 == 0:
            log(f"[bench] large 2d step {step + 1}/{steps}, band {s.stats().band_voxels}")
    total = time.perf_counter() - t_all
    ktab, roof = time_passes(s, args.pass_reps)
    res = {"workload": f"Surface<2,2> {N}^2 in {part}^2 partitions, 512 circles "
                       f"r=U[4,64]*{N}/16384, seed 0xFE1706, {steps} steps of update(-0.5), "
                       "notify, incremental marching squares",
           "steps": steps, "first_step_ms": lat[0] * 1e3, "value": ups / total,
           "ms_per_step": total / steps * 1e3,
           "p50_step_ms": pctl(lat, 50), "p99_step_ms": pctl(lat, 99),
           "voxel_updates_per_step": ups / steps,
           "segments_per_s": spx / march_s if march_s > 0 else None,
           "band_voxels_final": int(s.stats().band_voxels), "roofline": roof, "kernels": ktab}
    polys.close()
    s.close()
    log(f"[bench] large 2d: {res['value'] / 1e9:.2f} G voxel updates/s, "
        f"{res['ms_per_step']:.2f} ms/step")
    return res


def ray_bench(s, size, n):
    """Surface::ray throughput on the final state: n rays from a shell of radius 0.45*size about
    the centre, aimed at points within 0.1*size of it; rays resident on the device."""
    import numpy as np
    rng = np.random.default_rng(0xFE1705)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    o = 0.45 * size * u
    tgt = rng.uniform(-0.1 * size, 0.1 * size, size=(n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o, d = o.astype(np.float32), d.astype(np.float32)
    us = s.time_ray(o, d, reps=5)
    k = min(n, 1 << 16)
    hits = s.ray(o[:k], d[:k])
    frac = float(np.mean(~np.all(hits == np.float32(3.4028235e38), axis=1)))
    log(f"[bench] ray: {n} rays in {us / 1e3:.2f} ms, {n / us:.1f} M rays/s, hit fraction {frac:.3f}")
    return {"rays": n, "us_per_launch": us, "rays_per_s": n / (us * 1e-6), "hit_fraction": frac,
            "workload": "Surface::ray from a shell of radius 0.45*size aimed within 0.1*size of "
                        "the centre, on the final state; one thread per ray, timed with HIP events "
                        "over 5 launches"}


SERVED_SEED = 0xFE1704


def run_served(args):
    """Served stream: resident surfaces, Poisson request arrivals, continuous batching."""
    import felt_amd
    from felt_amd import workloads
    from felt_amd.server import Server, request
    lib = felt_amd.lib().lib
    specs = workloads.mixed_surfaces(args.served_surfaces, SERVED_SEED, 64, 384, fixed_part=16)
    sizes = [sp.size for sp in specs]
    budget = int(args.budget_frac * 288e9)
    arrivals = workloads.served_requests(specs, args.served_window, args.rate, SERVED_SEED ^ 0x5EED)

    def make(max_active, target_ms=0.0):
        t = time.time()
        srv = Server(lib, 3, 3, sizes, (16,) * 3, budget, max_active, polygonise=True)
        srv.set_target(target_ms)
        for i, sp in enumerate(specs):
            srv.batch.load_spheres(i, sp.centres, sp.radii)
        srv.start()
        log(f"[bench] served: {len(specs)} surfaces resident in {time.time() - t:.1f}s, "
            f"{srv.stats()['device_bytes'] / 2**30:.1f} GiB")
        return srv

    def reqs(sel):
        return [request(i, lo, hi, workloads.speed_of(k), n, rid)
                for rid, (_, i, lo, hi, k, n) in sel]

    def stream(target_ms):
        """The arrival stream in real time (arrivals are submitted when their time has come)
        through a server with the given tick-time target."""
        srv = make(args.max_active, target_ms)
        todo = list(enumerate(arrivals))
        total = len(todo)
        active_sum, ticks = 0, 0
        t0 = time.perf_counter()
        while True:
            now = time.perf_counter() - t0
            k = 0
            while k < len(todo) and todo[k][1][0] <= now:
                k += 1
            if k:
                srv.submit(reqs(todo[:k]))
                todo = todo[k:]
            st = srv.tick()
            if st == 0:
                if not todo and srv.stats()["requests_done"] >= total:
                    break
                if todo:
                    time.sleep(max(0.0, min(0.001, todo[0][1][0] - (time.perf_counter() - t0))))
                continue
            active_sum += st
            ticks += 1
            srv.poll()
            if ticks % 200 == 0:
                log(f"[bench] served tick {ticks}: {srv.stats()['requests_done']}/{total} done")
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        return {"tick_target_ms": target_ms or None,
                "requests": total, "ticks": st["ticks"], "wall_s": wall,
                "value": st["voxel_updates"] / wall,
                "capacity": st["voxel_updates"] / st["busy_s"] if st["busy_s"] else None,
                "step_p50_ms": st["step_p50_ms"], "step_p99_ms": st["step_p99_ms"],
                "request_p50_ms": st["request_p50_ms"], "request_p99_ms": st["request_p99_ms"],
                "mean_in_flight": active_sum / max(1, ticks),
                "phase_ms": dict(zip(["bounded update", "notify", "march", "band refresh"],
                                     st["phase_ms"])),
                "admissions_deferred": st["admissions_deferred"],
                "budget_bytes": budget, "device_bytes": st["device_bytes"],
                "bricks_used": st["bricks_used"], "bricks_total": st["bricks_total"]}

    res = stream(args.tick_target)
    res["workload"] = (f"served stream: {len(specs)} resident Surface<3,3> float32, edge "
                       "log-uniform 64-384, 16^3 partitions, 1-8 spheres, seed "
                       f"0x{SERVED_SEED:X}; Poisson arrivals at {args.rate:g}/s for "
                       f"{args.served_window:g}s; request = update(lo,hi,fn) over a box of "
                       "U[0.05,1]*edge per axis, speed from the three, 1-50 steps, each step "
                       "followed by repolygonisation of the changed partitions")
    if args.tick_target:
        # the same stream with no tick-time target: largest batches, highest throughput
        res["throughput_mode"] = stream(0.0)
    # plain: the same stream one request step at a time (no batching), for a bounded sample
    srv = make(1)
    srv.submit(reqs(list(enumerate(arrivals))))
    for _ in range(args.plain_ticks):
        if srv.tick() == 0:
            break
    pst = srv.stats()
    srv.close()
    res["plain"] = {"capacity": pst["voxel_updates"] / pst["busy_s"] if pst["busy_s"] else None,
                    "ticks": pst["ticks"], "step_p50_ms": pst["step_p50_ms"],
                    "step_p99_ms": pst["step_p99_ms"]}
    if res["capacity"] and res["plain"]["capacity"]:
        res["served_vs_plain"] = res["capacity"] / res["plain"]["capacity"]
    tm = res.get("throughput_mode")
    if tm and tm["capacity"] and res["plain"]["capacity"]:
        tm["served_vs_plain"] = tm["capacity"] / res["plain"]["capacity"]
    log(f"[bench] served: {res['value'] / 1e9:.2f} G/s wall, capacity "
        f"{(res['capacity'] or 0) / 1e9:.2f} G/s, p99 step {res['step_p99_ms']:.2f} ms")
    return res


def cpu_baseline(specs, budget_s):
    """The CPU restatement (oracle/, OpenMP over partitions like the reference) on the mixed
    workload from the same initial state: every surface is loaded, then passes step each surface
    once in turn until the time budget is spent (only the update calls are timed)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from felt_amd import workloads
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    b = oracle_lib.binding()
    b.lib.felt_oracle_set_threads(cores)
    surfaces = []
    for sp in specs:
        c = oracle_lib.surface(3, 3, sp.size, sp.part)
        c.b.lib.felt_oracle_surface_set_omp(c.h, 1)
        c.load_dense(sp.phi())
        surfaces.append((c, workloads.speed_of(sp.speed)))
    passes, steps, ups, el = 0, 0, 0, 0.0
    while el < budget_s and passes < 50:
        for c, f in surfaces:
            t0 = time.perf_counter()
            c.update(f)
            el += time.perf_counter() - t0
            ups += int(c.stats().last_voxel_updates)
            steps += 1
            if el >= budget_s:
                break
        passes += 1
        log(f"[bench] cpu pass {passes}: {el:.1f}s")
    for c, _ in surfaces:
        c.close()
    return {"value": ups / el, "unit": "voxel updates/s", "cores": cores, "kind": "port",
            "sample": f"{steps} update steps ({passes} passes over the {len(specs)} surfaces) of "
                      f"the same mixed workload from the same initial state ({el:.1f}s of CPU "
                      "work), CPU restatement with OpenMP over partitions"}


def run_traffic_probe(args):
    """For rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs: a device-to-device copy of 256 MiB (the
    counters' scale), then the mixed workload's 16^3 batch stepped `warmup` times and its six
    first-round distance passes replayed `pass_reps` times."""
    import ctypes
    import felt_amd
    from felt_amd import workloads
    from felt_amd.batch import Batch
    hip = ctypes.CDLL("libamdhip64.so")
    n = 256 << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(n)) or \
            hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, ctypes.c_size_t(n))
    for _ in range(3):
        if hip.hipMemcpy(dst, src, ctypes.c_size_t(n), 3):  # device to device
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log("[bench] traffic probe: 3 copies of 256 MiB")
    specs = mixed_workload(args.surfaces)
    ids = [i for i, sp in enumerate(specs) if sp.part[0] == 16]
    bt = Batch(felt_amd.lib().lib, 3, 3, [specs[i].size for i in ids], (16,) * 3)
    for j, i in enumerate(ids):
        bt.load_spheres(j, specs[i].centres, specs[i].radii)
    speeds = [workloads.speed_of(specs[i].speed) for i in ids]
    for _ in range(args.warmup):
        bt.update(speeds)
    bt.sync()
    us, items, n_launch = bt.time_pass(1, args.pass_reps)
    bt.close()
    res = {"surfaces": len(ids), "distance_pass_us": us, "entries_per_launch": items,
           "launches_per_copy": n_launch, "copies": args.pass_reps}
    log(f"[bench] traffic probe: distance pass {us:.1f} us per launch, {items:.0f} entries")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--surfaces", type=int, default=256)
    ap.add_argument("--single-steps", type=int, default=200)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--part", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--pass-reps", type=int, default=20)
    ap.add_argument("--large", action="store_true",
                    help="also run the full-size topology-changing workloads (1024^3 3D, "
                         "16384^2 2D); minutes of host-side merge checks")
    ap.add_argument("--large-only", action="store_true")
    ap.add_argument("--large-size", type=int, default=1024)
    ap.add_argument("--large2d-size", type=int, default=16384)
    ap.add_argument("--large-steps", type=int, default=300)
    ap.add_argument("--large-check", type=int, default=5,
                    help="steps between merge checks of the 3D workload")
    ap.add_argument("--plain-launches", action="store_true",
                    help="plain surfaces (the mixed workload's plain path and the single "
                         "surface) launch their kernels one by one instead of replaying a "
                         "captured graph")
    ap.add_argument("--group-steps", type=int, default=25,
                    help="steps timed with the batches as one group (one launch per step)")
    ap.add_argument("--tune-reps", type=int, default=5,
                    help="felt_batch_autotune replays per candidate (0: default launch shapes)")
    ap.add_argument("--rays", type=int, default=1 << 20, help="rays cast on the single surface "
                    "(0 skips)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--served-only", action="store_true")
    ap.add_argument("--single-only", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="copy of known size + the 16^3 batch's distance passes, for PMC runs")
    ap.add_argument("--served-surfaces", type=int, default=1000)
    ap.add_argument("--served-window", type=float, default=2.0)
    ap.add_argument("--rate", type=float, default=2000.0)
    ap.add_argument("--budget-frac", type=float, default=0.25)
    ap.add_argument("--max-active", type=int, default=1000)
    # admission aims at 4.5 ms so that the tick p99 stays under the 5 ms latency target with
    # margin for a noisier host
    ap.add_argument("--tick-target", type=float, default=4.5)
    ap.add_argument("--plain-ticks", type=int, default=300)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the surfaces run on a single MI355X and there is no "
              "multi-GPU path; independent surfaces would shard by surface id, one pool per GPU.",
              file=sys.stderr)
        sys.exit(2)
    import felt_amd
    felt_amd.lib()
    if args.served_only:
        print(json.dumps({"served": run_served(args)}), flush=True)
        return
    if args.single_only:
        print(json.dumps({"single_surface": run_single(args)}), flush=True)
        return
    if args.traffic_probe:
        print(json.dumps({"traffic_probe": run_traffic_probe(args)}), flush=True)
        return
    if args.large_only:
        print(json.dumps({"large_3d": run_large_3d(args), "large_2d": run_large_2d(args)}),
              flush=True)
        return
    specs, batched, plain, ktab, roof, cfg = run_mixed(args)
    log(f"[bench] batched {batched['value'] / 1e9:.2f} G/s" +
        (f", plain {plain['value'] / 1e9:.2f} G/s" if plain else ""))
    result = {
        "metric": METRIC,
        "value": batched["value"],
        "unit": "voxel updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": batched["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": cfg,
        "p50_step_ms": batched["p50_step_ms"],
        "p99_step_ms": batched["p99_step_ms"],
        "voxel_updates_per_step": batched["voxel_updates_per_step"],
        "launch": batched["launch"],
        "group_launch": batched.get("group_launch"),
        "plain": plain,
        "batched_vs_plain": batched["value"] / plain["value"] if plain else None,
        "roofline": roof,
        "kernels": ktab,
    }
    if not args.no_served:
        result["served"] = run_served(args)
    if not args.no_single:
        result["single_surface"] = run_single(args)
    if args.large:
        result["large_3d"] = run_large_3d(args)
        result["large_2d"] = run_large_2d(args)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(specs, args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
