#!/usr/bin/env python3
"""Benchmark of mesh_to_flexible_dual_grid on one MI355X.

One step = one call of the public drop-in API
``o_voxel.convert.mesh_to_flexible_dual_grid`` on a synthetic mesh already
resident in HBM (float32 vertices, int64 faces, as a caller passes them),
including the wrapper's dtype cast and aabb shift.  Prints ONE JSON line.

  python bench.py [--steps K] [--warmup W]
                  [--workload scene|scene_large|ico_cap|scene_rankdef|batch|to_mesh|serialize|vxz] [--grid G]

Default workload: the 10.3M-face seeded scene (the ~280k-face scene generator,
two 1-to-4 subdivisions with vertex jitter, trimmed to 10.3M faces) at grid
512, weights face 1.0 / boundary 0.2 / regularization 1e-2; headline in the
reference float path, the compensated precision timed beside it.

Reported next to the throughput:
  roofline      the conversion's algorithmic bytes (12V + 12F + 27N) over the
                dominant kernel's average launch time (HIP events on the
                conversion's stream around every launch) vs 8 TB/s, beside the
                kernel's own byte model and the whole-step figure
  cpu_baseline  the CPU oracle (float path, single thread) on the same
                workload (a seeded 300k-face sample of meshes above 1M faces)
  parity        occupancy / flag mismatches and dual error against the oracle
                float path (on the same sample)
  precision     the other precision's time, and the compensated path's dual
                error against the FP64 oracle (on the same sample)
"""

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "o-voxel-gpu_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "M faces/s + ms/mesh at grid 512; batched/FP32-comp vs plain at dual RMS≤2e-4"
HBM_PEAK_GBS = 8000.0
WEIGHTS = (1.0, 0.2, 1e-2)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name, grid):
    import meshgen

    if name == "ico_cap":
        v, f = meshgen.ico_cap()
        g = grid or 64
        seed = None
    elif name == "scene":
        v, f = meshgen.scene(seed=2)
        g = grid or 512
        seed = 2
    elif name == "scene_large":
        v, f = meshgen.scene_large(seed=3)
        g = grid or 512
        seed = 3
    elif name == "scene_rankdef":
        v, f = meshgen.scene_rankdef(seed=4)
        g = grid or 1024
        seed = 4
    else:
        raise SystemExit(f"unknown workload {name}")
    return v, f, g, seed


def algorithmic_bytes(slot, st, F, V, n_words, precision="reference"):
    """Compulsory HBM bytes of one launch of a kernel slot: what the slot's
    kernels must read and write at least once, per the comments below (the
    roofline's `achieved` figure divides this by the slot's launch time)."""
    N, R, E, ev = st["voxels"], st["records"], st["boundary_edges"], st["events"]
    FR, BR = st["face_records"], st["boundary_records"]
    R0 = max(0, R - FR - BR)  # crossing records: one per (crossing, brick of its voxels)
    comp = precision == "compensated"
    # the fold reads every record once, each face plane and boundary segment
    # once, the (first record, count) pair of each chunk and the occupancy word
    # + first voxel id of each brick row (64 rows of 12 B per brick), and writes
    # the flags (3 B) and, on the reference path, the dual vertex (12 B; the
    # voxel is solved where it is folded) and the queued box-search items (72 B),
    # on the precision variant the compensated quadric (80 B); the fold tiers
    # split it in proportion to the voxel entries each folds
    fold_out = (80 * N + 3 * N) if comp else (15 * N + 72 * st["box_voxels"])
    fold_all = (16 * R + 16 * F + 88 * E + 8 * st["chunks"] + 768 * st["active_bricks"] + fold_out)
    # voxel entries: 4 per crossing, 1 per face or boundary record; the large
    # tiers' count is in entries
    large_share = min(1.0, st["large_brick_records"] / max(1, 4 * ev + FR + BR))
    table = {
        # faces (12 B) + vertices (12 B) read once, 16 B plane + 36 B gathered triangle per face written
        "prep": 12 * F + 12 * V + 16 * F + 36 * F,
        # faces + vertices read, occupancy bitset written once, 16 B per crossing recorded
        "mark": 12 * F + 12 * V + 8 * n_words + 16 * ev,
        # bitset read, word offsets written
        "word_scan": 8 * n_words + 4 * n_words,
        # half-edges binned by smaller vertex: faces read by the count and place passes, 3F 8 B half-edges
        # written and read back, 8 B per boundary edge written
        "edges": 24 * F + 48 * F + 8 * E,
        # bitset + offsets read, coords written
        "coords": 12 * n_words + 12 * N,
        # crossings read (16 B), one 16 B record per (crossing, brick) written
        "gen_events": 16 * ev + 16 * R0,
        # faces + vertices + bitset read, one 12 B hit per face / voxel overlap written
        "face": 12 * F + 12 * V + 8 * n_words + 12 * FR,
        # hits read, one 16 B record per hit written
        "gen_fhits": 12 * FR + 16 * FR,
        # boundary segments (88 B) read, one 16 B record per voxel on the segment written
        "bdry": 88 * E + 16 * BR,
        "fold": (1.0 - large_share) * fold_all,
        "fold_large": large_share * fold_all,
        # precision variant: quadric (80 B) + coords (12 B) read, dual (12 B) written, 72 B per queued voxel
        "solve": 104 * N + 72 * st["box_voxels"],
        # queued voxels (72 B) read, dual (12 B) written
        "solve_box": 84 * st["box_voxels"],
    }
    return table.get(slot)


def meshgen_grid(grid):
    import meshgen

    return meshgen.grid_args(grid)


def cpu_baseline(v, f, grid, workload, max_faces=None):
    """Oracle float path (single thread) on the workload or a bounded face sample."""
    import meshgen
    from oracle import oracle as O

    O.build()
    vs, gr, a0 = meshgen.grid_args(grid)
    vl = (v - a0).astype(np.float32)
    fi = f.astype(np.int32)
    sample = f"full {workload} mesh ({len(f)} faces) at grid {grid}"
    if max_faces is not None and len(fi) > max_faces:
        rng = np.random.default_rng(0)
        idx = np.sort(rng.choice(len(fi), size=max_faces, replace=False))
        fi = fi[idx]
        sample = f"{max_faces} of {len(f)} faces (seeded sample) of {workload} at grid {grid}"
    # repeated until at least 10 s of CPU work; dt is the mean time of one run
    runs, total = 0, 0.0
    while runs == 0 or total < 10.0:
        t0 = time.perf_counter()
        ref = O.flexible_dual_grid(vl, fi, vs, gr, *WEIGHTS)
        total += time.perf_counter() - t0
        runs += 1
    dt = total / runs
    sample += f", mean of {runs} runs ({total:.1f} s)"
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return ref, dt, len(fi), sample, cpu, vs, gr, fi


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="scene_large")
    ap.add_argument("--grid", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-other-precision", action="store_true",
                    help="skip timing the precision not selected by --precision")
    ap.add_argument("--precision", default="reference", choices=["reference", "compensated"],
                    help="arithmetic of the headline steps (compensated: voxel-frame compensated FP32, FP64 "
                         "fallback); the other one is timed too and reported in the `precision` block with its "
                         "dual error against the FP64 oracle")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed steps, copy a 256 MiB buffer 5 times (a kernel of known traffic "
                         "for scaling hardware byte counters)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build measures one GPU only (the conversion of one mesh runs on one device); "
              f"--gpus {args.gpus} / WORLD_SIZE {world} is not supported", file=sys.stderr)
        sys.exit(2)

    import torch

    import o_voxel
    from o_voxel import _C

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)

    if args.workload == "batch":
        return run_batch(args, torch, o_voxel, _C, dev)
    if args.workload == "to_mesh":
        return run_to_mesh(args, torch, o_voxel, _C, dev)
    if args.workload in ("serialize", "vxz"):
        return run_voxel_io(args, torch, o_voxel, _C, dev)
    return run_conversion(args, torch, o_voxel, _C, dev)


def _device_warmup(torch, dev, seconds=0.3):
    """Torch's own first launches, before the cold call: the public API
    shifts the vertices by the aabb origin with a torch subtraction, and the
    first launch of a torch kernel in a process loads its code object (tens
    of ms on a freshly started machine, where torch's library is not yet in
    the page cache).  About `seconds` of copies and that subtraction on
    tensors unrelated to the conversion, so the cold call measures this
    library's first-call costs (its code loading, workspace growth)."""
    a = torch.empty(64 << 20, dtype=torch.uint8, device=dev)
    b = torch.empty_like(a)
    v = torch.zeros((1 << 16, 3), dtype=torch.float32, device=dev)
    o = torch.zeros((2, 3), dtype=torch.float32).pin_memory().to(dev, non_blocking=True)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(16):
            b.copy_(a)
            w = v - o[0].reshape(1, 3)
        torch.cuda.synchronize()
    del a, b, v, o, w


def _calibrate_copy(torch, dev):
    """A 256 MiB device copy, 5 times: a kernel of known traffic for scaling hardware byte counters."""
    src = torch.ones(64 * 2 ** 20, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def _dual_error(d, ref, vs):
    """(rms, max) of the per-voxel Euclidean dual difference, in voxels."""
    if len(d) == 0:
        return 0.0, 0.0
    e = (np.asarray(d, np.float64) - np.asarray(ref, np.float64)) / np.asarray(vs, np.float64)[None, :]
    n = np.sqrt((e ** 2).sum(1))
    return float(np.sqrt((n ** 2).mean())), float(n.max())


def run_conversion(args, torch, o_voxel, _C, dev):
    """One step = one public-API conversion of the workload's mesh (inputs in
    HBM).  The headline precision is --precision (default: the reference
    float path, the drop-in); the other precision is timed on the same mesh
    and reported in the line's `precision` block, with dual errors against the
    FP64 oracle on a seeded face subset (the whole mesh when it has at most
    300k faces)."""
    v, f, grid, seed = make_workload(args.workload, args.grid)
    F, V = int(f.shape[0]), int(v.shape[0])
    log(f"[bench] workload={args.workload} F={F} V={V} grid={grid}")
    vt = torch.from_numpy(v).to(dev)
    ft = torch.from_numpy(f).to(dev)
    aabb = [[-0.5, -0.5, -0.5], [0.5, 0.5, 0.5]]
    other = "compensated" if args.precision == "reference" else "reference"

    def step(precision, fts=ft):
        return o_voxel.convert.mesh_to_flexible_dual_grid(vt, fts, grid_size=grid, aabb=aabb, face_weight=WEIGHTS[0],
                                                          boundary_weight=WEIGHTS[1],
                                                          regularization_weight=WEIGHTS[2], precision=precision)

    def timed(precision):
        """K timed steps back to back: ms per step over the whole region."""
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for _ in range(args.steps):
            out = step(precision)
        torch.cuda.synchronize()
        return (time.perf_counter() - t_start) * 1e3 / args.steps, out

    def latencies(precision, timing):
        """K more steps, each waited for: per-step latency (ms) and, with
        timing, per-kernel device times (HIP events around every launch on the
        conversion's stream)."""
        lat, per_kernel = [], {}
        if timing:
            _C.set_timing(dev, True)
        for _ in range(args.steps):
            ts = time.perf_counter()
            step(precision)
            torch.cuda.synchronize()
            lat.append((time.perf_counter() - ts) * 1e3)
            if timing:
                for k, ms in _C.kernel_times(dev).items():
                    per_kernel.setdefault(k, []).append(ms)
        if timing:
            _C.set_timing(dev, False)
        return lat, per_kernel

    # cold call (first conversion in the process, on a device brought out of idle)
    _device_warmup(torch, dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = step(args.precision)
    torch.cuda.synchronize()
    cold_ms = (time.perf_counter() - t0) * 1e3
    for _ in range(args.warmup):
        out = step(args.precision)
    torch.cuda.reset_peak_memory_stats(dev)
    # timed region: K steps of the headline precision
    ms_per_step, out = timed(args.precision)
    st = _C.last_stats(dev)
    peak_mb = torch.cuda.max_memory_allocated(dev) / 2 ** 20 + st["workspace_bytes"] / 2 ** 20
    step_ms, per_kernel = latencies(args.precision, True)
    if args.calibrate_copy:
        _calibrate_copy(torch, dev)
    # the other precision on the same mesh
    o_ms = o_lat = None
    if not args.no_other_precision:
        for _ in range(max(1, args.warmup)):
            step(other)
        o_ms, _ = timed(other)
        o_st = _C.last_stats(dev)
        o_lat, _ = latencies(other, False)
    coords, dual, inter = (t.cpu().numpy() for t in out)
    value = F / (ms_per_step * 1e-3) / 1e6

    G = grid
    n_words = G * G * ((G + 63) // 64)
    kern_avg = {k: float(np.mean(x)) for k, x in per_kernel.items()}
    dom = max(kern_avg, key=kern_avg.get)
    # canonical algorithmic bytes of one conversion: 12 B per vertex and per
    # face read, 27 B per voxel written (coords 12, dual 12, flags 3)
    b_alg = 12 * V + 12 * F + 27 * st["voxels"]
    model = algorithmic_bytes(dom, st, F, V, n_words, args.precision)
    achieved = b_alg / (kern_avg[dom] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{dom}_{args.workload}_traffic.json")
    if os.path.exists(tpath) and args.precision == "reference":
        with open(tpath) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    roofline = {
        "kernel": dom,
        "bound": "hbm",
        "achieved": round(achieved, 3),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 6),
        "traffic": traffic,
        "alg_bytes": int(b_alg),
        "alg_bytes_basis": "12V + 12F + 27N of the conversion, all charged to the dominant kernel",
        "avg_ms": round(kern_avg[dom], 4),
        "kernel_model": {"alg_bytes": int(model) if model else None,
                         "achieved": round(model / (kern_avg[dom] * 1e-3) / 1e9, 3) if model else None,
                         "frac": round(model / (kern_avg[dom] * 1e-3) / 1e9 / HBM_PEAK_GBS, 6) if model else None,
                         "basis": "the kernel's own compulsory reads and writes (records, planes, quadrics)"},
        "path": {"achieved": round(b_alg / (ms_per_step * 1e-3) / 1e9, 3),
                 "frac": round(b_alg / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 6),
                 "basis": "12V + 12F + 27N over the whole step"},
    }

    # CPU baseline, parity and the precision block on one face sample
    cpu = parity = None
    max_faces = 300_000 if F > 1_000_000 else None
    precision = {"mode": other if o_ms is not None else None}
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        ref, dt, nf, sample, cpu_model, vs, gr, fi = cpu_baseline(v, f, grid, args.workload, max_faces)
        cpu = {"value": round(nf / dt / 1e6, 4), "unit": "M faces/s", "cores": 1, "kind": "port",
               "sample": sample, "seconds": round(dt, 3), "cpu": cpu_model, "nproc": os.cpu_count()}
        if nf == F:
            s_out = {args.precision: (coords, dual, inter)}
        else:  # the same face sample converted on the GPU
            fs = torch.from_numpy(fi.astype(np.int64)).to(dev)
            s_out = {args.precision: tuple(t.cpu().numpy() for t in step(args.precision, fs))}
        if o_ms is not None:
            fs = ft if nf == F else torch.from_numpy(fi.astype(np.int64)).to(dev)
            s_out[other] = tuple(t.cpu().numpy() for t in step(other, fs))
        perm = O.canonical_order(ref.coords, gr)
        c_r, d_r, i_r = s_out["reference"] if "reference" in s_out else (None, None, None)
        on = "full mesh" if nf == F else f"{nf}-face seeded sample"
        if c_r is not None:
            occ_ok = c_r.shape == ref.coords.shape and bool(np.array_equal(c_r, ref.coords[perm]))
            parity = {"on": on, "occupancy_match": occ_ok, "vs": "oracle float path"}
            if occ_ok:
                rms, mx = _dual_error(d_r, ref.dual[perm], vs)
                parity.update({
                    "intersected_mismatch": int((i_r != ref.flags[perm]).sum()),
                    "dual_rms_voxel": rms, "dual_max_voxel": mx,
                    "dual_bit_identical": float(np.mean(np.all(d_r == ref.dual[perm], axis=1))) if len(d_r) else 1.0,
                })
        if "compensated" in s_out:
            from oracle import oracle as O

            r64 = O.flexible_dual_grid((v - meshgen_grid(grid)[2]).astype(np.float32), fi, vs, gr, *WEIGHTS,
                                       fp64=True)
            p64 = O.canonical_order(r64.coords, gr)
            c_c, d_c, i_c = s_out["compensated"]
            occ_ok = c_c.shape == r64.coords.shape and bool(np.array_equal(c_c, r64.coords[p64]))
            rms, mx = _dual_error(d_c, r64.dual64[p64], vs) if occ_ok else (None, None)
            precision.update({"vs": "FP64 oracle", "on": on, "occupancy_match": occ_ok,
                              "intersected_mismatch": int((i_c != r64.flags[p64]).sum()) if occ_ok else None,
                              "dual_rms_voxel": rms, "dual_max_voxel": mx})
            if c_r is not None and occ_ok:
                rms_p, mx_p = _dual_error(d_r, r64.dual64[p64], vs)
                precision.update({"reference_path_dual_rms_voxel": rms_p, "reference_path_dual_max_voxel": mx_p})
    if o_ms is not None:
        precision.update({"ms_per_step": round(o_ms, 4), "M_faces_per_s": round(F / (o_ms * 1e-3) / 1e6, 4),
                          "p50_ms": round(float(np.percentile(o_lat, 50)), 4),
                          "fp64_fraction": o_st["fp64_voxels"] / max(1, o_st["voxels"])})

    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "M faces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}-{F}f@{grid}", "faces": F, "vertices": V, "grid": grid,
                   "voxels": st["voxels"], "records": st["records"], "seed": seed, "precision": args.precision,
                   "weights": {"face": WEIGHTS[0], "boundary": WEIGHTS[1], "regularization": WEIGHTS[2]}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "precision": precision,
        "cold_ms": round(cold_ms, 3),
        "p50_ms": round(float(np.percentile(step_ms, 50)), 4),
        "p99_ms": round(float(np.percentile(step_ms, 99)), 4),
        "kernels_ms": {k: round(x, 4) for k, x in sorted(kern_avg.items(), key=lambda kv: -kv[1])},
        "peak_mem_mb": round(peak_mb, 1),
        "counts": {k: v for k, v in st.items() if k != "workspace_bytes"},
    }
    print(json.dumps(line), flush=True)


def run_batch(args, torch, o_voxel, _C, dev):
    """256 seeded meshes, faces log-uniform in [500, 500k], grid per mesh from
    {64, 128, 256, 512}: one batched conversion per step vs the same meshes
    converted by 256 plain calls."""
    import meshgen

    meshes = meshgen.batch(n_meshes=256, seed=5)
    F = sum(len(f) for _, f, _ in meshes)
    log(f"[bench] batch: {len(meshes)} meshes, {F} faces")
    aabb = [[-0.5, -0.5, -0.5], [0.5, 0.5, 0.5]]
    ms_in = [(torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev)) for v, f, _ in meshes]
    grids = [g for _, _, g in meshes]
    kw = dict(face_weight=WEIGHTS[0], boundary_weight=WEIGHTS[1], regularization_weight=WEIGHTS[2])

    def batched(precision):
        return o_voxel.convert.mesh_to_flexible_dual_grid_batch(ms_in, grid_sizes=grids, aabb=aabb,
                                                                precision=precision, **kw)

    def sequential(precision):
        out, lat = [], []
        for (vt, ft), g in zip(ms_in, grids):
            t0 = time.perf_counter()
            out.append(o_voxel.convert.mesh_to_flexible_dual_grid(vt, ft, grid_size=g, aabb=aabb,
                                                                  precision=precision, **kw))
            torch.cuda.synchronize()
            lat.append((time.perf_counter() - t0) * 1e3)
        return out, lat

    def measure(precision):
        """K batched steps timed back to back, then the same meshes by one
        plain call each (median over K/4 passes, per-call latency recorded)."""
        for _ in range(args.warmup):
            batched(precision)
            sequential(precision)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out_b = batched(precision)
        torch.cuda.synchronize()
        b_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        s_ms, lat_all = [], []
        for _ in range(max(1, args.steps // 4)):
            t0 = time.perf_counter()
            out_s, lat = sequential(precision)
            s_ms.append((time.perf_counter() - t0) * 1e3)
            lat_all += lat
        s_ms = float(np.median(s_ms))
        same = all(all(torch.equal(a, b) for a, b in zip(ob, os_)) for ob, os_ in zip(out_b, out_s))
        return {
            "batched_ms": round(b_ms, 3),
            "batched_meshes_per_s": round(len(meshes) / (b_ms * 1e-3), 1),
            "batched_M_faces_per_s": round(F / (b_ms * 1e-3) / 1e6, 3),
            "sequential_ms": round(s_ms, 3),
            "sequential_meshes_per_s": round(len(meshes) / (s_ms * 1e-3), 1),
            "speedup": round(s_ms / b_ms, 3),
            "sequential_p50_ms_per_mesh": round(float(np.percentile(lat_all, 50)), 4),
            "sequential_p99_ms_per_mesh": round(float(np.percentile(lat_all, 99)), 4),
            "batched_ms_per_mesh": round(b_ms / len(meshes), 4),
            "outputs_identical_to_sequential": bool(same),
        }

    _device_warmup(torch, dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batched("reference")
    torch.cuda.synchronize()
    cold_ms = (time.perf_counter() - t0) * 1e3
    res = measure("reference")
    batch_ms = res["batched_ms"]
    res_c = None if args.no_other_precision else measure("compensated")
    # graph mode: the same batched steps with the device work after the voxel
    # set replayed from HIP graphs (captured at the first step of a plan)
    eager = [[t.clone() for t in o] for o in batched("reference")]
    _C.set_graph(dev, True)
    for _ in range(max(2, args.warmup)):
        batched("reference")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out_g = batched("reference")
    torch.cuda.synchronize()
    g_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    gstats = _C.graph_stats(dev)
    _C.set_graph(dev, False)
    same_g = all(all(torch.equal(a, b) for a, b in zip(og, oe)) for og, oe in zip(out_g, eager))
    del eager, out_g
    graph = {"batched_ms": round(g_ms, 3), "eager_batched_ms": round(batch_ms, 3),
             "gain_ms": round(batch_ms - g_ms, 3), "outputs_identical_to_eager": bool(same_g), **gstats}
    # per-kernel device times of one more batched step
    _C.set_timing(dev, True)
    batched("reference")
    kt = _C.kernel_times(dev)
    _C.set_timing(dev, False)
    st = _C.last_stats(dev)
    line = {
        "metric": METRIC,
        "value": round(F / (batch_ms * 1e-3) / 1e6, 4),
        "unit": "M faces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batch_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batch-{len(meshes)}meshes-{F}f", "meshes": len(meshes), "faces": F, "seed": 5,
                   "grids": {str(g): grids.count(g) for g in sorted(set(grids))},
                   "weights": {"face": WEIGHTS[0], "boundary": WEIGHTS[1], "regularization": WEIGHTS[2]}},
        "batch": res,
        "batch_compensated": res_c,
        "batch_graph": graph,
        "cold_ms": round(cold_ms, 3),
        "kernels_ms": {k: round(x, 4) for k, x in sorted(kt.items(), key=lambda kv: -kv[1]) if x >= 0},
        "counts": {k: v for k, v in st.items() if k != "workspace_bytes"},
    }
    print(json.dumps(line), flush=True)


def run_to_mesh(args, torch, o_voxel, _C, dev):
    """flexible_dual_grid_to_mesh on the flexible dual grid of the ~280k-face
    scene (grid 512 unless --grid), dual vertices in voxel units as the
    reference's example stores them.  Steps time the public API (fused HIP
    path); beside it, the reference's own recipe for the same function (its
    hashmap calls and torch glue, flexible_dual_grid.py:419-451) composed from
    this build's generic hashmap on the same GPU, and the CPU oracle on a
    z-slab sample of the voxels."""
    import meshgen
    from oracle import dual_mesh as DM

    grid = args.grid or 512
    v, f = meshgen.scene(seed=2)
    aabb = [[-0.5, -0.5, -0.5], [0.5, 0.5, 0.5]]
    coords, dual, inter = o_voxel.convert.mesh_to_flexible_dual_grid(
        torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev), grid_size=grid, aabb=aabb,
        face_weight=WEIGHTS[0], boundary_weight=WEIGHTS[1], regularization_weight=WEIGHTS[2])
    dual = dual * grid - coords
    N = int(coords.shape[0])
    flags = int(inter.sum())
    log(f"[bench] to_mesh: N={N} flags={flags} grid={grid}")

    def fused():
        return o_voxel.convert.flexible_dual_grid_to_mesh(coords, dual, inter, None, aabb=aabb, grid_size=grid)

    off = torch.from_numpy(DM.EDGE_NEIGHBORS.astype(np.int32)).to(dev).unsqueeze(0)
    vs = torch.full((3,), 1.0 / grid, dtype=torch.float32, device=dev)
    a0 = torch.full((1, 3), -0.5, dtype=torch.float32, device=dev)

    def recipe():
        tk = torch.full((2 * N,), -1, dtype=torch.int32, device=dev).view(torch.uint32)
        tv = torch.empty((2 * N,), dtype=torch.uint32, device=dev)
        _C.hashmap_insert_3d_idx_as_val_cuda(tk, tv, torch.cat([torch.zeros_like(coords[:, :1]), coords], 1),
                                             grid, grid, grid)
        cand = (coords.reshape(N, 1, 1, 3) + off)[inter]
        M = cand.shape[0]
        keys = torch.cat([torch.zeros((M * 4, 1), dtype=torch.int32, device=dev), cand.reshape(-1, 3)], 1)
        idx = _C.hashmap_lookup_3d_cuda(tk, tv, keys, grid, grid, grid).reshape(M, 4).int()
        q = idx[(idx != -1).all(1)]
        verts = (coords.float() + dual) * vs + a0

        def align(t):
            p0, p1, p2, p3 = (verts[t[:, k]] for k in range(4))
            n0 = torch.cross(p1 - p0, p2 - p0, dim=1)
            n1 = torch.cross(p2 - p1, p3 - p1, dim=1)
            return (n0 * n1).sum(1, keepdim=True).abs()

        t1, t2 = q[:, [0, 1, 2, 0, 2, 3]], q[:, [0, 1, 3, 3, 1, 2]]
        return verts, torch.where(align(t1) > align(t2), t1, t2).reshape(-1, 3)

    _device_warmup(torch, dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fused()
    torch.cuda.synchronize()
    cold_ms = (time.perf_counter() - t0) * 1e3
    for _ in range(args.warmup):
        fused()
        recipe()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        vf, ff = fused()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        vr, fr = recipe()
    torch.cuda.synchronize()
    recipe_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    same_v = bool(torch.equal(vf, vr))
    face_mismatch = int((ff != fr).any(1).sum()) if ff.shape == fr.shape else -1
    # the quads split differently: area of their first triangle (q0, q1, q2)
    # relative to the quad's extent, |(v1 - v0) x (v2 - v0)| / diam^2 in float64
    # (where it is within rounding of zero, torch's FMA-contracted cross of a
    # vector with itself no longer gives exactly zero and the recipe's choice
    # can flip)
    diff_area = None
    if face_mismatch > 0:
        qd = (ff != fr).reshape(-1, 6).any(1).nonzero().flatten()
        _, _, quads = _C.flexible_dual_grid_to_mesh_cuda(coords, dual, inter, None, [grid] * 3, [1.0 / grid] * 3,
                                                         [-0.5] * 3, faces=False, quads=True)
        t1 = quads[qd].long()
        p = vf.double()
        e1, e2 = p[t1[:, 1]] - p[t1[:, 0]], p[t1[:, 2]] - p[t1[:, 0]]
        corners = p[t1]  # [m, 4, 3]
        diam = torch.cdist(corners, corners).amax((1, 2))
        diff_area = float((torch.linalg.norm(torch.cross(e1, e2, dim=1), dim=1) / diam.clamp_min(1e-300) ** 2).max())
    L = int(ff.shape[0]) // 2
    # per-kernel device times (HIP events on the call's stream), averaged over steps
    _C.set_timing(dev, True)
    acc = {}
    for _ in range(args.steps):
        fused()
        for k, x in _C.kernel_times(dev).items():
            if k.startswith("dm_"):
                acc[k] = acc.get(k, 0.0) + x
    _C.set_timing(dev, False)
    kt = {k: x / args.steps for k, x in acc.items()}
    if args.calibrate_copy:
        src = torch.ones(64 * 2 ** 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    cap = 64
    while cap < 2 * N:
        cap *= 2
    alg = {"dm_insert": 12 * N + 8 * cap + 8 * N,
           "dm_count": 44 * N + 32 * flags,
           "dm_scan": 8 * N,
           "dm_emit": 17 * N + 104 * L}
    dom = max(kt, key=kt.get)
    achieved = alg[dom] / (kt[dom] * 1e-3) / 1e9
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "algorithmic_bytes_per_launch": alg[dom], "avg_launch_ms": round(kt[dom], 4)}
    tpath = os.path.join(ROOT, "profiles", f"{dom}_to_mesh_traffic.json")
    if os.path.exists(tpath):
        roof["traffic"] = json.load(open(tpath))["hbm_bytes_per_launch"]
    cpu = None
    if not args.no_cpu_baseline:
        n_s = min(N, 1_000_000)  # voxels of the lowest z slabs (canonical order is z-major)
        c_np, d_np, i_np = coords[:n_s].cpu().numpy(), dual[:n_s].cpu().numpy(), inter[:n_s].cpu().numpy()
        t0 = time.perf_counter()
        DM.dual_grid_to_mesh(c_np, d_np, i_np, None, (grid,) * 3, [1.0 / grid] * 3, [-0.5] * 3)
        dt = time.perf_counter() - t0
        cpu = {"value": round(n_s / dt / 1e6, 4), "unit": "M voxels/s", "cores": 1, "kind": "port",
               "sample": f"numpy/Python oracle on the first {n_s} voxels (lowest z slabs), {dt:.2f} s",
               "cpu_model": platform.processor() or platform.machine(), "nproc": os.cpu_count()}
    line = {
        "metric": "flexible_dual_grid_to_mesh M voxels/s",
        "value": round(N / (ms * 1e-3) / 1e6, 4),
        "unit": "M voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"to_mesh-scene280k@{grid}", "voxels": N, "intersected_flags": flags, "quads": L,
                   "grid": grid, "seed": 2},
        "reference_recipe_ms": round(recipe_ms, 4),
        "speedup_vs_reference_recipe": round(recipe_ms / ms, 3),
        "parity_vs_recipe": {"vertices_identical": same_v, "faces_differing": face_mismatch,
                             "max_relative_area_of_differing_first_triangles": diff_area},
        "cold_ms": round(cold_ms, 3),
        "kernels_ms": {k: round(x, 4) for k, x in sorted(kt.items(), key=lambda kv: -kv[1])},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def _scene_voxels(torch, o_voxel, dev, grid):
    """The conversion of the ~280k-face scene: coords, duals, flags."""
    import meshgen

    v, f = meshgen.scene(seed=2)
    return o_voxel.convert.mesh_to_flexible_dual_grid(
        torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev), grid_size=grid, aabb=[[-0.5] * 3, [0.5] * 3],
        face_weight=WEIGHTS[0], boundary_weight=WEIGHTS[1], regularization_weight=WEIGHTS[2])


def run_voxel_io(args, torch, o_voxel, _C, dev):
    """serialize: Z-order and Hilbert encode + decode of the scene
    conversion's voxels (grid 512).  vxz: the reference example's O-voxel
    file of that conversion (Morton-sorted voxels, duals quantised to uint8,
    flags packed to 3 bits, parent filter) written and read back through
    o_voxel.io, uncompressed and deflate; CPU oracle beside each."""
    import io as _io

    grid = args.grid or 512
    coords, dual, inter = _scene_voxels(torch, o_voxel, dev, grid)
    N = int(coords.shape[0])

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, out

    line = {"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "data": "synthetic"}
    if args.workload == "serialize":
        from oracle import serialize as S

        res = {}
        for mode in ("z_order", "hilbert"):
            enc_ms, codes = timed(lambda: o_voxel.serialize.encode_seq(coords, mode=mode), args.steps)
            dec_ms, back = timed(lambda: o_voxel.serialize.decode_seq(codes, mode=mode), args.steps)
            c_np = coords.cpu().numpy()
            t0 = time.perf_counter()
            ref = (S.z_order_encode if mode == "z_order" else S.hilbert_encode)(c_np[:, 0], c_np[:, 1], c_np[:, 2])
            cpu_s = time.perf_counter() - t0
            res[mode] = {"encode_ms": round(enc_ms, 4), "decode_ms": round(dec_ms, 4),
                         "encode_M_voxels_per_s": round(N / enc_ms / 1e3, 1),
                         "codes_identical_to_oracle": bool(np.array_equal(codes.cpu().numpy(), ref)),
                         "decode_round_trip": bool(torch.equal(back, coords)),
                         "cpu_oracle_encode_s": round(cpu_s, 3)}
        hb = res["hilbert"]
        alg = 16 * N  # 12 B of coordinates read + 4 B of code written per voxel
        achieved = alg / (hb["encode_ms"] * 1e-3) / 1e9
        line.update({
            "metric": "voxel codes M voxels/s (Hilbert encode)", "value": hb["encode_M_voxels_per_s"],
            "unit": "M voxels/s", "ms_per_step": hb["encode_ms"], "dtype": "int32",
            "config": {"workload": f"serialize-scene280k@{grid}", "voxels": N, "grid": grid, "seed": 2},
            "modes": res,
            "roofline": {"kernel": "k_seq_encode<hilbert>", "bound": "hbm", "achieved": round(achieved, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": None, "algorithmic_bytes_per_launch": alg,
                         "note": "whole call (one kernel + wrapper), not the kernel alone"},
            "cpu_baseline": {"value": round(N / res["hilbert"]["cpu_oracle_encode_s"] / 1e6, 3),
                             "unit": "M voxels/s", "cores": 1, "kind": "port",
                             "sample": f"numpy oracle, all {N} voxels (Hilbert encode)"},
        })
        print(json.dumps(line), flush=True)
        return
    # vxz
    from oracle import vxz as X

    order = torch.argsort(o_voxel.serialize.encode_seq(coords))
    coords, dual, inter = coords[order], dual[order], inter[order]
    attr = {"dual_vertices": (torch.clamp(dual * grid - coords, 0, 1) * 255).type(torch.uint8),
            "intersected": (inter[:, 0:1] + 2 * inter[:, 1:2] + 4 * inter[:, 2:3]).type(torch.uint8)}
    res = {}
    for comp in ("none", "deflate"):
        def write():
            b = _io.BytesIO()
            o_voxel.io.write_vxz(b, coords, attr, filter="parent", compression=comp)
            return b.getvalue()

        steps = max(1, args.steps // 4) if comp == "deflate" else args.steps
        w_ms, data = timed(write, steps)
        r_ms, (c2, a2) = timed(lambda: o_voxel.io.read_vxz(_io.BytesIO(data)), steps)
        same = bool(torch.equal(c2, coords.cpu())) and all(torch.equal(a2[k], attr[k].cpu()) for k in attr)
        res[comp] = {"write_ms": round(w_ms, 3), "read_ms": round(r_ms, 3), "bytes": len(data),
                     "write_M_voxels_per_s": round(N / w_ms / 1e3, 2), "read_M_voxels_per_s": round(N / r_ms / 1e3, 2),
                     "round_trip_exact": same}
    # CPU oracle writer on one chunk-sized sample (voxels of the lowest Morton codes)
    n_s = min(N, 60000)
    c_np = coords[:n_s].cpu().numpy()
    a_np = {k: v[:n_s].cpu().numpy() for k, v in attr.items()}
    t0 = time.perf_counter()
    ref = X.write_vxz_bytes(c_np, a_np, filter="parent", compression="none")
    cpu_s = time.perf_counter() - t0
    b = _io.BytesIO()
    o_voxel.io.write_vxz(b, coords[:n_s], {k: v[:n_s] for k, v in attr.items()}, filter="parent",
                         compression="none")
    un = res["none"]
    line.update({
        "metric": "O-voxel .vxz write M voxels/s (parent filter, uncompressed)", "value": un["write_M_voxels_per_s"],
        "unit": "M voxels/s", "ms_per_step": un["write_ms"], "dtype": "u8",
        "config": {"workload": f"vxz-scene280k@{grid}", "voxels": N, "grid": grid, "chunk_size": 256, "seed": 2},
        "compression": res,
        "sample_bytes_identical_to_oracle": b.getvalue() == ref,
        "roofline": None,
        "cpu_baseline": {"value": round(n_s / cpu_s / 1e6, 4), "unit": "M voxels/s", "cores": 1, "kind": "port",
                         "sample": f"Python oracle writer on the first {n_s} Morton-ordered voxels, uncompressed, "
                                   f"{cpu_s:.2f} s"},
    })
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
