"""Oracle runs of many independent surfaces in worker processes, for the large workload tests.

The oracle is serial; independent surfaces (batch members, served surfaces) are stepped in
parallel by processes started with `spawn` (they never touch the GPU).  Results come back as
digests: per layer the SHA-1 of the sorted row-major keys of its voxels, and the SHA-1 of the
snapshot's float32 bytes, which the GPU side computes the same way."""
import hashlib
import os
from concurrent.futures import ProcessPoolExecutor
import multiprocessing as mp

import numpy as np

import parity


def layer_digests(layer_positions, size, L):
    """(count, sha1) of layer k for k = -L..L; layer_positions(k) -> (n, D) world positions."""
    out = []
    for k in range(-L, L + 1):
        keys = parity.position_keys(layer_positions(k), size)
        out.append((len(keys), hashlib.sha1(keys.tobytes()).hexdigest()))
    return tuple(out)


def phi_digest(snapshot: np.ndarray) -> str:
    return hashlib.sha1(np.ascontiguousarray(snapshot, dtype=np.float32).tobytes()).hexdigest()


def workers() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def run(fn, jobs):
    """fn(job) for every job in worker processes; results in job order."""
    with ProcessPoolExecutor(max_workers=min(workers(), len(jobs)),
                             mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(fn, jobs))


# ---- jobs (module level, so spawned workers can unpickle them) ----------------------------------

def mixed_member_job(args):
    """Oracle of members `ids` of mixed_surfaces(n, seed, lo, hi), each stepped `steps` global
    updates with its own speed: {(i, step): layer digests} at the checkpoints and {(i, "phi"):
    snapshot digest} after the last step."""
    import oracle_lib
    from felt_amd import workloads
    n, seed, lo, hi, ids, steps, checkpoints = args
    specs = workloads.mixed_surfaces(n, seed, lo, hi)
    out = {}
    for i in ids:
        sp = specs[i]
        c = oracle_lib.surface(sp.D, sp.L, sp.size, sp.part)
        c.load_dense(sp.phi())
        fn = workloads.speed_of(sp.speed)
        for step in range(1, steps + 1):
            c.update(fn)
            if step in checkpoints:
                out[(i, step)] = layer_digests(c.layer_positions, sp.size, sp.L)
        out[(i, "phi")] = phi_digest(c.snapshot())
        c.close()
    return out


def served_surface_job(args):
    """Oracle of one served surface running its requests alone, in order: layer digests and
    snapshot digest at the end.  requests: [(lo, hi, speed index, steps)]."""
    import oracle_lib
    from felt_amd import workloads
    spec, reqs = args
    c = oracle_lib.surface(spec.D, spec.L, spec.size, spec.part)
    c.load_dense(spec.phi())
    for lo, hi, k, steps in reqs:
        fn = workloads.speed_of(k)
        for _ in range(steps):
            c.update_box(lo, hi, fn)
    res = (layer_digests(c.layer_positions, spec.size, spec.L), phi_digest(c.snapshot()))
    c.close()
    return res


def state_digest(s):
    """Layers (as layer_digests), active partitions and the isogrid / delta / status-change
    partition lists of a Surface mirror (GPU or oracle)."""
    from felt_amd._binding import LIST_DELTA, LIST_ISOGRID, LIST_STATUS_CHANGE
    parts = [hashlib.sha1(np.packbits(s.partition_active()).tobytes()).hexdigest()]
    for kind in (LIST_ISOGRID, LIST_DELTA, LIST_STATUS_CHANGE):
        for k in range(2 * s.L + 1):
            parts.append(hashlib.sha1(np.sort(s.partition_list(kind, k)).tobytes()).hexdigest())
    return layer_digests(s.layer_positions, s.size, s.L), tuple(parts)


# The single large surface (C2) at a quarter of its lengths: 128^3 in 8^3 partitions, a sphere of
# radius 37.5 perturbed by 8 spherical harmonics of amplitude up to 1.5, value noise on a 16-voxel
# lattice; the speed and the step count are the config's.
C2Q = dict(size=(128, 128, 128), part=(8, 8, 8), radius=37.5, seed=0xFE1702, amp=1.5, cell=16,
           steps=200, march_every=10, mesh_steps=(10, 100, 200))


def c2q_field():
    from felt_amd import workloads
    return workloads.harmonic_sphere_phi(C2Q["size"], 3, C2Q["radius"], C2Q["seed"], amp=C2Q["amp"])


def c2q_speed():
    from felt_amd._binding import speed_noise_curv
    return speed_noise_curv(-0.3, 0.1, 0.5, C2Q["seed"], C2Q["cell"])


def c2q_job(fp64):
    """The oracle (fp32, or the FP64 oracle) through the reduced C2 run: state digests every
    step and meshes at mesh_steps (fp32), snapshots every march_every steps."""
    import oracle_lib
    c = oracle_lib.surface(3, 3, C2Q["size"], C2Q["part"], fp64=fp64)
    c.load_dense(c2q_field())
    pc = None
    if not fp64:
        pc = oracle_lib.polys(c)
        pc.invalidate()
        pc.march()
    sp = c2q_speed()
    out = {"state": [], "phi": {}, "mesh": {}}
    for step in range(1, C2Q["steps"] + 1):
        c.update(sp)
        if pc is not None:
            out["state"].append(state_digest(c))
            pc.notify()
            if step % C2Q["march_every"] == 0:
                pc.march()
            if step in C2Q["mesh_steps"]:
                out["mesh"][step] = parity.mesh_sets(pc, c)
        if step % C2Q["march_every"] == 0:
            out["phi"][step] = c.snapshot()
    return out


# The large topology-changing run (C5a) at an eighth of its lengths: 128^3 in 4^3 partitions,
# 64 spheres of radius U[20, 60] / 8, 16 boxes of 8^3 per step.  The speed is -0.5 until the
# inside region (phi < 0) is one face-connected component -- every initial sphere merged -- and
# 0.2 * curvature from the next step on; 300 steps.
C5AQ = dict(N=128, part=4, L=3, seed=0xFE1705, steps=300, boxes=16, box=8)


def c5aq_setup():
    """Centres, radii and the boxes of every step, from the config's SplitMix64 stream."""
    from felt_amd import workloads
    N = C5AQ["N"]
    rng = workloads.SplitMix64(C5AQ["seed"])
    cs = [tuple(rng.uniform(-N / 2, N / 2) for _ in range(3)) for _ in range(64)]
    rs = [rng.uniform(20, 60) / 8 for _ in range(64)]
    boxes = []
    for _ in range(C5AQ["steps"]):
        step_boxes = []
        for _ in range(C5AQ["boxes"]):
            lo = [rng.randint(-N // 2, N // 2 - C5AQ["box"]) for _ in range(3)]
            step_boxes.append((lo, [x + C5AQ["box"] - 1 for x in lo]))
        boxes.append(step_boxes)
    return cs, rs, boxes


def one_component(snapshot) -> bool:
    from scipy import ndimage
    return ndimage.label(snapshot < 0)[1] == 1


def c5aq_run(surface, polys, on_step):
    """Drives the reduced C5a run on a Surface mirror and its Polys (GPU or oracle); on_step(step,
    switched) after every step.  Returns the step after which the speed switched."""
    from felt_amd._binding import speed_constant, speed_curvature
    _, _, boxes = c5aq_setup()
    switch = None
    for step in range(C5AQ["steps"]):
        sp = speed_constant(-0.5) if switch is None else speed_curvature(0.2)
        surface.update(sp)
        polys.notify()
        for lo, hi in boxes[step]:
            surface.update_box(lo, hi, sp)
            polys.notify()
        polys.march()
        if switch is None and one_component(surface.snapshot()):
            switch = step
        on_step(step, switch)
    return switch


def c5aq_job(_):
    import oracle_lib
    from felt_amd import workloads
    cs, rs, _ = c5aq_setup()
    size = (C5AQ["N"],) * 3
    c = oracle_lib.surface(3, C5AQ["L"], size, (C5AQ["part"],) * 3)
    c.load_dense(workloads.spheres_phi(size, C5AQ["L"], cs, rs))
    pc = oracle_lib.polys(c)
    pc.invalidate()
    pc.march()
    out = {"state": [], "phi": {}, "mesh": {}, "band": []}

    def on_step(step, switch):
        out["state"].append(state_digest(c))
        out["band"].append(int(c.stats().band_voxels))
        if step % 50 == 49:
            out["phi"][step] = c.snapshot()
        if step == switch or step in (100, 299):
            out["mesh"][step] = parity.mesh_sets(pc, c)

    out["switch"] = c5aq_run(c, pc, on_step)
    return out
