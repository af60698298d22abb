"""Benchmark of the projective SPH smoothing path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1|c3]

A step is one pass of the hot path over the workload's synthetic snapshot
(already resident in HBM): per view, the per-particle quantity (X-ray
emissivity or mass), the camera projection with depth mask, the reference
particle-grid build and the projected-kernel gather, plus the retrieval of
the image.  The metric is particle-kernel evaluations per second, where an
evaluation is a (view, pixel, particle) triple with w != 0 and the pixel
inside the particle's support, counted exactly (independent of culling).

Prints ONE JSON line.  See the "measurement" section of DESIGN.md for the
roofline and CPU-baseline definitions.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-kernel evals/s, SPH projection: batched vs plain on MI355X @1e-5 rel err"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
FLOP_PER_EVAL = 12


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(name):
    from illcrawl_amd import synth
    if name == "c1":
        snap = synth.cached("c1")
        views = [dict(synth.centred_view(snap.box, width=2000.0, nx=512), quantity=0)]
        desc = dict(workload="plain projected gas density (column density)", n_cells=snap.n,
                    box_kpc_h=snap.box, resolution=[512, 512], width_kpc_h=2000.0,
                    kernel="quartic_ref", mode="reference", views=1)
    elif name == "c2":
        snap = synth.cached("c2")
        views = [dict(synth.centred_view(snap.box, width=5000.0, nx=2048), quantity=1)]
        desc = dict(workload="X-ray bremsstrahlung projection", n_cells=snap.n,
                    box_kpc_h=snap.box, halos=8, resolution=[2048, 2048], width_kpc_h=5000.0,
                    kernel="quartic_ref", mode="reference", views=1, accumulator="fp64")
    else:
        raise SystemExit(f"unknown workload {name}")
    return snap, views, desc


class Renderer:
    """Plain path: one view per launch sequence through the C ABI."""

    def __init__(self, snap, views, mode="reference", tile=None):
        import torch
        from illcrawl_amd import backend as B, synth
        self.B = B
        self.views = views
        dev = torch.device("cuda", 0)
        self.pos4 = torch.empty((snap.n, 4), dtype=torch.float32, device=dev)
        self.pos4[:, :3] = torch.from_numpy(snap.pos).to(dev)
        self.mass = torch.from_numpy(snap.mass).to(dev)
        self.h = torch.from_numpy(snap.hsml).to(dev)
        self.has_gas = snap.temp is not None
        if self.has_gas:
            self.rho = torch.from_numpy(snap.rho).to(dev)
            self.temp = torch.from_numpy(snap.temp).to(dev)
            self.xe = torch.from_numpy(snap.xe).to(dev)
            self.gtab, self.gx0, self.gdx = synth.gaunt_table_1d()
        self.backend = B.ProjectiveSmoothingGrid(mode=mode, tile=tile)
        self.backend.init_backend(snap.n)
        self.pts = [B.camera_pixel_points(v) for v in views]
        self.gather_ms = []
        self.images = [None] * len(views)

    def step(self, record=False):
        import torch
        B = self.B
        for k, v in enumerate(self.views):
            if v.get("quantity", 0) == 1:
                w = B.quantity_xray_bolometric(self.rho, self.temp, self.xe, self.gtab,
                                               self.gx0, self.gdx)
            else:
                w = B.quantity_scaled(self.mass, 1.0)
            self.pos4[:, 3] = w
            proj = B.project_particles(v, self.pos4)
            self.backend.setup_projected_particles(proj, self.h)
            self.backend.setup_evaluation_points(self.pts[k], layout_nx=v["nx"])
            if record:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
            self.backend.run()
            if record:
                e1.record()
                self.gather_ms.append((e0, e1))
            self.images[k] = self.backend.retrieve_results(1.0)

    def count_evals(self):
        B = self.B
        total, per = 0, []
        for k, v in enumerate(self.views):
            w = (B.quantity_xray_bolometric(self.rho, self.temp, self.xe, self.gtab, self.gx0,
                                            self.gdx) if v.get("quantity", 0) == 1
                 else B.quantity_scaled(self.mass, 1.0))
            self.pos4[:, 3] = w
            proj = B.project_particles(v, self.pos4)
            self.backend.setup_projected_particles(proj, self.h)
            self.backend.setup_evaluation_points(self.pts[k], layout_nx=v["nx"])
            t, pp = self.backend.count_evals(per_point=True)
            total += t
            per.append(pp)
        return total, per


def cpu_baseline(snap, view, per_point_counts, seconds=15.0, threads=None):
    """The oracle (the reference algorithm restated in C, OpenMP over
    pixels) on a bounded random sample of the view's pixels; evals from the
    exact per-pixel pair counts of the same pixels."""
    import numpy as np
    import oracle
    from illcrawl_amd import synth
    threads = threads or int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count())
    oracle.set_num_threads(threads)
    cam = oracle.Camera(view["position"], view["look_at"], view["roll"], view["width"],
                        view["nx"], view["ny"])
    if view.get("quantity", 0) == 1:
        tab, x0, dx = synth.gaunt_table_1d()
        w = oracle.xray_bolometric(snap.rho, snap.temp, snap.xe, tab, x0, dx)
    else:
        w = snap.mass
    pp = cam.project_particles(snap.particles4(w), view["depth"])
    pts = cam.pixel_points()
    t0 = time.perf_counter()
    g = oracle.Grid(pp, snap.hsml, target=1000)
    t_grid = time.perf_counter() - t0
    rng = np.random.default_rng(123)
    order = rng.permutation(pts.shape[0])
    counts = per_point_counts.cpu().numpy()
    done, evals, elapsed, batch = 0, 0, 0.0, 16
    while elapsed < seconds and done < order.size:
        idx = order[done:done + batch]
        t0 = time.perf_counter()
        g.gather_ref(pts[idx], acc_mode=0)
        elapsed += time.perf_counter() - t0
        evals += int(counts[idx].sum())
        done += idx.size
        batch = min(batch * 2, 4096)
    return dict(value=evals / elapsed, unit="evals/s", cores=threads, kind="port",
                sample=f"{done} random pixels of view 0 ({elapsed:.1f} s of the reference "
                       f"gather restated in C, FP32 accumulator); grid build "
                       f"{t_grid:.2f} s single-threaded not included")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=os.environ.get("ILC_BENCH_WORKLOAD", "c2"))
    ap.add_argument("--mode", default="reference")
    ap.add_argument("--tile", default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for "
              f"--gpus {args.gpus}); views would shard across GPUs with no exchange",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")

    t0 = time.time()
    snap, views, desc = workload(args.workload)
    log(f"[bench] generated {snap.n} cells in {time.time() - t0:.1f}s")
    tile = tuple(int(x) for x in args.tile.split("x")) if args.tile else None
    r = Renderer(snap, views, mode=args.mode, tile=tile)
    total_evals, per_point = r.count_evals()
    log(f"[bench] exact evals per step: {total_evals:.6g}")
    for _ in range(args.warmup):
        r.step()
    torch.cuda.synchronize()
    log("[bench] warmup done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.step(record=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    gather = [a.elapsed_time(b) for a, b in r.gather_ms]
    gather_avg_ms = float(np.mean(gather))
    log(f"[bench] {args.steps} steps in {dt:.3f}s, gather avg {gather_avg_ms:.2f} ms")

    ms_per_step = dt * 1e3 / args.steps
    value = total_evals / (ms_per_step * 1e-3)
    evals_per_launch = total_evals / len(views)
    achieved_tflops = evals_per_launch * FLOP_PER_EVAL / (gather_avg_ms * 1e-3) / 1e12
    npix = sum(v["nx"] * v["ny"] for v in views)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 accumulation)",
        "data": "synthetic (seeded generator: uniform + NFW halos, 32-NN density)",
        "config": dict(desc, path="plain", parallelism="single GPU",
                       evals_per_step=total_evals, pixels_per_step=npix),
        "pixels_per_s": npix / (ms_per_step * 1e-3),
        "gather_ms": gather_avg_ms,
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_FP32_TFLOPS,
                     "traffic": None, "kernel": "k_gather (reference mode)",
                     "flop_per_eval": FLOP_PER_EVAL},
    }
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(snap, views[0], per_point[0], args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
