"""Benchmark of the projective SPH smoothing path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2|c1|c4|c5|v1]

Workloads (synthetic, seeded; see illcrawl_amd/synth.py):
  c3 (default)  the batched multi-view workload: the 16M-cell X-ray snapshot
                (75000 kpc/h box, 8 NFW halos) projected into 64 views of
                random orientation (16 each at 256^2, 512^2, 1024^2, 2048^2,
                widths 1000-5000 kpc/h, quantity alternating column density /
                X-ray), one persistent gather launch per batch;
  c2            one 2048^2 X-ray view of the same snapshot (plain path);
  c1            the 2^20-cell column-density view, 512^2 (plain path);
  c5            32 X-ray slabs (500 kpc/h, centre mask) of a 64M-cell snapshot at
                4096^2 each, batched;
  c4            forward and adjoint of a 1024^2 column-density image of 4M
                cells (exact mode, --kernel quartic|cubic): adjoint cost
                relative to forward and gradient error against the FP64 oracle
                on a sample.
  v1            volumetric smoothing (the 3-D grid reconstruction): gas
                density on 8 slices of 2048^2 points, one through each halo
                of the 16M-cell snapshot, 3000 kpc/h wide; a step is the 3-D
                grid build over all 16M cells, one gather launch over the
                8 slices and the retrieval.

A step is one pass of the hot path over the workload, with the snapshot
already resident in HBM: per view the quantity functor, the camera projection
with depth mask, the reference particle-grid build, and the projected-kernel
gather, then the image retrieval.  The metric counts particle-kernel
evaluations: (view, pixel, particle) triples with w != 0 and the pixel inside
the particle's support, counted exactly on the GPU (independent of culling).

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-kernel evals/s, SPH projection: batched vs plain on MI355X @1e-5 rel err"
PEAK_FP32_TFLOPS = 157.3  # MI355X vector FP32 (the spec figure counts packed FP32)
FLOP_PER_EVAL = 12
# 3-D pair: dx, dy, dz; r^2 (mul + 2 FMA); s = 1 - r^2/h^2 (FMA); s^2; w s^2 + a (FMA)
FLOP_PER_EVAL_3D = 13
# HBM bytes per k_gather_batch launch, from rocprofv3 FETCH_SIZE / WRITE_SIZE
# runs of this workload scaled by a 1 GiB copy (profiles/k_gather_batch_xray_16M_64views.md);
# keyed by (workload, mode, tile)
MEASURED_TRAFFIC = {("c3", "reference", (16, 16, 256)): dict(read=22.63e9, write=3.250e9),
                    # profiles/k_gather_batch_xray_64M_32slabs.md
                    ("c5", "reference", (16, 16, 256)): dict(read=21.15e9, write=19.59e9),
                    # profiles/k_gather_batch_colden_1M_512_and_xray_16M_2048.md
                    ("c1", "reference", (8, 8, 64)): dict(read=28.80e6, write=10.97e6),
                    ("c2", "reference", (16, 16, 256)): dict(read=252.97e6, write=161.38e6),
                    # k_smooth3d, profiles/k_smooth3d_slices_16M.md
                    ("v1", "reference", (16, 16)): dict(read=1.244e9, write=0.461e9)}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# every gather tile shape (w, h, threads per workgroup)
PLAIN_TILES = [(16, 16, 256), (32, 16, 256), (32, 32, 256), (64, 16, 256), (8, 8, 64),
               (16, 16, 64)]


def parse_tile(text):
    """'WxH' (one tile per 256-thread workgroup) or 'WxH/G' (G threads per
    workgroup: 256, or 64 for one-wave 16x16 tiles) -> (W, H, G)."""
    shape, _, wg = text.partition("/")
    w, h = (int(x) for x in shape.split("x"))
    return w, h, int(wg) if wg else 256


KINDS = {"quartic_ref": 0, "quartic": 1, "cubic": 2}


def sampled_parity(snap, view, img_f64, n=1500, seed=7, kernel="quartic_ref", mode="reference"):
    """Relative L-inf / sum error of the GPU image against the oracle on a
    random pixel sample (plus the maximum): the reference-semantics gather
    (reference mode), or the FP64 sum over every particle (exact mode)."""
    import numpy as np
    import oracle
    from illcrawl_amd import synth
    cam = oracle.Camera(view["position"], view["look_at"], view["roll"], view["width"],
                        view["nx"], view["ny"])
    if view.get("quantity", 0) == 1:
        tab, x0, dx = synth.gaunt_table_1d()
        w = oracle.xray_bolometric(snap.rho, snap.temp, snap.xe, tab, x0, dx)
    else:
        w = snap.mass
    pp = cam.project_particles(snap.particles4(w), view["depth"])
    pts = cam.pixel_points()
    g = oracle.Grid(pp, snap.hsml, target=1000)
    img = img_f64.reshape(-1)
    rng = np.random.default_rng(seed)
    idx = np.unique(np.concatenate([rng.choice(img.size, min(n, img.size), replace=False),
                                    [int(np.argmax(img))]]))
    if mode == "reference":
        ref = g.gather_ref(pts[idx], acc_mode=1)
        ref32 = g.gather_ref(pts[idx], acc_mode=0)
    else:
        ref = oracle.gather_exact(pp, snap.hsml, pts[idx], kind=KINDS[kernel])
        ref32 = None
    m = ref > 1e-6 * ref.max()
    linf = float(np.max(np.abs(img[idx][m] - ref[m]) / ref[m])) if m.any() else 0.0
    l32 = float(np.max(np.abs(ref32[m] - ref[m]) / ref[m])) if m.any() and ref32 is not None \
        else None
    return dict(pixels=int(idx.size), linf_vs_oracle=linf,
                oracle_fp32_accumulator_vs_fp64=l32)


def cpu_baseline(snap, view, per_point_counts, seconds=15.0, kernel="quartic_ref"):
    """The oracle (the reference algorithm restated in C, OpenMP over
    pixels) on a bounded random sample of one view's pixels; evals from the
    exact per-pixel pair counts of the same pixels (GPU count kernel, which
    the tests pin to the oracle's count).  Kernels other than the
    reference's quartic have no reference algorithm: the oracle's FP64 sum
    over every particle stands in."""
    import numpy as np
    import oracle
    from illcrawl_amd import synth
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count())
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
    ref_kernel = kernel == "quartic_ref"
    if not ref_kernel:
        batch = 1
    while elapsed < seconds and done < order.size:
        idx = order[done:done + batch]
        t0 = time.perf_counter()
        if ref_kernel:
            g.gather_ref(pts[idx], acc_mode=0)
        else:
            oracle.gather_exact(pp, snap.hsml, pts[idx], kind=KINDS[kernel])
        elapsed += time.perf_counter() - t0
        evals += int(counts[idx].sum())
        done += idx.size
        batch = min(batch * 2, 4096)
    what = ("the reference gather restated in C, FP32 accumulator" if ref_kernel else
            f"the oracle's FP64 {kernel} sum over every particle (no reference algorithm)")
    return dict(value=evals / elapsed, unit="evals/s", cores=threads, kind="port",
                sample=f"{done} random pixels of one {view['nx']}x{view['ny']} view "
                       f"({elapsed:.1f} s of {what}, {threads} OpenMP threads); its grid "
                       f"build ({t_grid:.2f} s, one thread) not included")


def clip_parity(snap, view, img_f64, n=240, seed=7):
    """Line-of-sight clip: the oracle's FP64 quadrature of the clipped 3-D
    quartic over every particle (orc_gather_clip) at a random pixel sample
    plus the maximum; its timing is the CPU baseline (no reference
    algorithm exists for the clip)."""
    import numpy as np
    import oracle
    from illcrawl_amd import synth
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count())
    oracle.set_num_threads(threads)
    cam = oracle.Camera(view["position"], view["look_at"], view["roll"], view["width"],
                        view["nx"], view["ny"])
    if view.get("quantity", 0) == 1:
        tab, x0, dx = synth.gaunt_table_1d()
        w = oracle.xray_bolometric(snap.rho, snap.temp, snap.xe, tab, x0, dx)
    else:
        w = snap.mass
    pz = cam.project_particles_depth(snap.particles4(w))
    pts = cam.pixel_points()
    img = img_f64.reshape(-1)
    rng = np.random.default_rng(seed)
    idx = np.unique(np.concatenate([rng.choice(img.size, n, replace=False),
                                    [int(np.argmax(img))]]))
    t0 = time.perf_counter()
    ref = oracle.gather_clip(pz, snap.hsml, pts[idx], view["depth"], 0)
    elapsed = time.perf_counter() - t0
    m = ref > 1e-6 * ref.max()
    linf = float(np.max(np.abs(img[idx][m] - ref[m]) / ref[m]))
    s = float(abs(img[idx].sum() - ref.sum()) / ref.sum())
    # pairs inside the support at the sampled pixels with w != 0 after the
    # depth cut (the eval unit; the clip keeps the particles whose support
    # overlaps the depth range)
    p4 = pz.copy()
    over = (pz[:, 2] + snap.hsml >= 0) & (pz[:, 2] - snap.hsml <= np.float32(view["depth"]))
    p4[~over, 3] = 0.0
    p4[:, 2] = 0.0
    _, cnt = oracle.gather_exact(p4, snap.hsml, pts[idx], with_counts=True)
    evals = int(cnt.sum())
    cpu = dict(value=evals / elapsed, unit="evals/s", cores=threads, kind="port",
               sample=f"{idx.size} random pixels of one {view['nx']}x{view['ny']} slab "
                      f"({elapsed:.1f} s of the oracle's FP64 clipped quadrature over every "
                      f"particle, {threads} OpenMP threads)")
    return dict(pixels=int(idx.size), linf_vs_oracle=linf, sum_vs_oracle=s), cpu


def single_view_counts(snap, view, kernel="quartic_ref"):
    """Per-pixel exact eval counts of one view (single-view backend)."""
    import torch
    from illcrawl_amd import backend as B, synth
    if view.get("quantity", 0) == 1:
        tab, x0, dx = synth.gaunt_table_1d()
        w = B.quantity_xray_bolometric(snap.rho, snap.temp, snap.xe, tab, x0, dx)
    else:
        w = B.quantity_scaled(snap.mass, 1.0)
    p4 = torch.empty((snap.n, 4), dtype=torch.float32, device="cuda")
    p4[:, :3] = torch.from_numpy(snap.pos).cuda()
    p4[:, 3] = w
    proj = B.project_particles(view, p4)
    del p4
    b = B.ProjectiveSmoothingGrid(mode="exact", kernel=kernel)
    b.init_backend(snap.n)
    b.setup_projected_particles(proj, snap.hsml)
    b.setup_evaluation_points(B.camera_pixel_points(view), layout_nx=view["nx"])
    total, per = b.count_evals(per_point=True)
    return total, per


def workload(name):
    from illcrawl_amd import synth
    if name == "c1":
        snap = synth.cached("c1")
        views = [dict(synth.centred_view(snap.box, width=2000.0, nx=512), quantity=0)]
        desc = dict(workload="plain projected gas density (column density), 1 view",
                    n_cells=snap.n, box_kpc_h=snap.box, resolution=[512, 512],
                    width_kpc_h=2000.0)
    elif name == "c2":
        snap = synth.cached("c2")
        views = [dict(synth.centred_view(snap.box, width=5000.0, nx=2048), quantity=1)]
        desc = dict(workload="X-ray bremsstrahlung projection, 1 view", n_cells=snap.n,
                    box_kpc_h=snap.box, halos=8, resolution=[2048, 2048], width_kpc_h=5000.0)
    elif name == "c3":
        snap = synth.cached("c2")
        views = synth.random_views(snap.box, n_views=64, seed=2)
        desc = dict(workload="batched multi-view: 64 random views of the 16M-cell X-ray "
                             "snapshot, one persistent launch", n_cells=snap.n,
                    box_kpc_h=snap.box, halos=8, views=64,
                    resolutions={"256^2": 16, "512^2": 16, "1024^2": 16, "2048^2": 16},
                    width_kpc_h=[1000.0, 5000.0], quantity="alternating column density / X-ray")
    elif name == "c5":
        snap = synth.cached("c5")
        views = synth.slab_views(snap.box, n_slabs=32, thickness=500.0, nx=4096)
        desc = dict(workload="slab tomography: 32 slabs of 500 kpc/h along z through the "
                             "64M-cell X-ray snapshot, 4096^2 each, one persistent launch",
                    n_cells=snap.n, box_kpc_h=snap.box, halos=32, views=32,
                    resolution=[4096, 4096], width_kpc_h=snap.box, slab_kpc_h=500.0,
                    quantity="X-ray", depth_cut="centre mask")
    else:
        raise SystemExit(f"unknown workload {name}")
    return snap, views, desc


def adjoint_main(args):
    """c4: forward and adjoint of L = 1/2 |I(m) - I(m')|^2, m' = m * lognormal(0.05),
    for the 4M-cell column-density snapshot at 1024^2 (exact mode, quartic).
    Forward = projection + grid build + gather + retrieval; adjoint =
    ilc_psg_backward (gradient clear, adjoint gather, un-permute)."""
    import numpy as np
    import torch
    from illcrawl_amd import backend as B, synth
    t0 = time.time()
    snap = synth.cached("c4")
    view = synth.centred_view(snap.box, width=2000.0, nx=1024)
    log(f"[bench] c4: {snap.n} cells ({time.time() - t0:.1f}s)")
    dev = torch.device("cuda", 0)
    p4 = torch.empty((snap.n, 4), dtype=torch.float32, device=dev)
    p4[:, :3] = torch.from_numpy(snap.pos).to(dev)
    p4[:, 3] = torch.from_numpy(snap.mass).to(dev)
    hs = torch.from_numpy(snap.hsml).to(dev)
    pts = B.camera_pixel_points(view)
    rng = np.random.default_rng(7)
    m_t = torch.from_numpy((snap.mass * rng.lognormal(0.0, 0.05, snap.n)).astype(np.float32))
    b = B.ProjectiveSmoothingGrid(mode="exact", kernel=args.kernel)
    b.init_backend(snap.n)

    def forward(w=None):
        q = p4 if w is None else torch.cat([p4[:, :3], w.to(dev)[:, None]], 1)
        proj = B.project_particles(view, q)
        b.setup_projected_particles(proj, hs)
        b.setup_evaluation_points(pts, layout_nx=view["nx"])
        b.run()
        return b.retrieve_results(1.0), proj

    target, _ = forward(m_t)
    img, proj = forward()
    G = (img - target).contiguous()
    total = b.count_evals()
    for _ in range(args.warmup):
        forward()
        b.backward(G)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        forward()
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        grads = b.backward(G)
    torch.cuda.synchronize()
    bwd_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    log(f"[bench] c4 forward {fwd_ms:.2f} ms, adjoint {bwd_ms:.2f} ms, evals {total:.6g}")
    line = {
        "metric": f"adjoint vs forward cost, projected SPH image (exact {args.kernel})",
        "value": total / (fwd_ms * 1e-3), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": fwd_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 per-pixel accumulation)",
        "data": "synthetic (seeded: uniform + NFW halo, 32-NN density, lognormal masses)",
        "config": dict(workload="adjoint of a 1024^2 column-density image of 4M cells",
                       n_cells=snap.n, box_kpc_h=snap.box, resolution=[1024, 1024],
                       width_kpc_h=2000.0, mode="exact", kernel=args.kernel,
                       evals_per_step=total, parallelism="single GPU"),
        "adjoint": dict(ms=bwd_ms, forward_ms=fwd_ms, ratio=bwd_ms / fwd_ms,
                        evals_per_s=total / (bwd_ms * 1e-3)),
    }
    if not args.no_parity:
        from oracle import backward_exact
        log("[bench] c4 gradient check vs FP64 oracle on a particle sample ...")
        sel = np.sort(rng.choice(snap.n, size=2000, replace=False))
        pp = proj.cpu().numpy()[sel]
        ref = backward_exact(pp, snap.hsml[sel], pts.cpu().numpy(),
                             G.cpu().numpy().astype(np.float64),
                             {"quartic": 1, "quartic_ref": 0, "cubic": 2}[args.kernel])
        errs = {}
        for name, g, r in zip(("m", "h", "x", "y"), grads, ref):
            got = g.cpu().numpy()[sel].astype(np.float64)
            errs[name] = float(np.max(np.abs(got - r)) / max(np.max(np.abs(r)), 1e-300))
        line["adjoint"]["grad_rel_to_max_err"] = errs
        line["adjoint"]["grad_sample"] = "2000 random particles, all 1024^2 pixels, FP64"
    print(json.dumps(line), flush=True)


def batch_adjoint_main(args):
    """c5 --adjoint: forward and adjoint of L = 1/2 sum over the 32 slabs of
    |I_v(w) - I_v(w')|^2, w' = w * lognormal(0.05) (densities scaled by the
    square root: the emission goes as rho^2), exact mode; one batched render
    and one batched backward (gradients w.r.t. w, h and positions)."""
    import numpy as np
    import torch
    from illcrawl_amd import backend as B, synth
    t0 = time.time()
    snap, views, desc = workload(args.workload)
    log(f"[bench] {args.workload} adjoint: {snap.n} cells, {len(views)} views "
        f"({time.time() - t0:.1f}s)")
    tab, x0, dx = synth.gaunt_table_1d()
    rng = np.random.default_rng(7)
    rho_t = (snap.rho * np.sqrt(rng.lognormal(0.0, 0.05, snap.n))).astype(np.float32)
    r = B.BatchRenderer(mode="exact")
    r.set_gaunt_table(tab, x0, dx)
    if args.tile:
        r.set_tile(*parse_tile(args.tile))
    r.set_particles(snap.pos, snap.mass, snap.hsml, rho_t, snap.temp, snap.xe)
    r.set_views(views)
    r.render()
    targets = [r.image(k).clone() for k in range(len(views))]
    r.set_particles(snap.pos, snap.mass, snap.hsml, snap.rho, snap.temp, snap.xe)
    r.render()
    grads = [(r.image(k) - targets[k]).contiguous() for k in range(len(views))]
    total = int(sum(r.count_evals()))
    for _ in range(args.warmup):
        r.render()
        r.backward(grads)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.render()
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gw, gh, gp = r.backward(grads)
    torch.cuda.synchronize()
    bwd_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    log(f"[bench] {args.workload} forward {fwd_ms:.2f} ms, adjoint {bwd_ms:.2f} ms")
    line = {
        "metric": "adjoint vs forward cost, batched slab projection (exact quartic_ref)",
        "value": total / (fwd_ms * 1e-3), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": fwd_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 per-pixel accumulation)",
        "data": "synthetic (seeded: uniform + NFW halos, 32-NN density, lognormal masses)",
        "config": dict(desc, mode="exact", evals_per_step=total, parallelism="single GPU"),
        "adjoint": dict(ms=bwd_ms, forward_ms=fwd_ms, ratio=bwd_ms / fwd_ms,
                        nonzero_dw=int((gw != 0).sum().item()),
                        finite=bool(torch.isfinite(gw).all() and torch.isfinite(gh).all()
                                    and torch.isfinite(gp).all())),
    }
    print(json.dumps(line), flush=True)


def volumetric_slices(snap, nx=2048, width=3000.0):
    """One slice through each halo, normals cycling through z, y, x and two
    oblique directions; all slices as one point array of nx-wide rows."""
    import numpy as np
    from illcrawl_amd import backend as B
    normals = [(0, 0, 1), (0, 1, 0), (1, 0, 0), (1, 1, 1), (1, -2, 0.5)]
    pts = [B.slice_points(np.asarray(c, np.float64), normals[k % len(normals)], width, nx)
           for k, (c, _) in enumerate(snap.halos)]
    return np.concatenate(pts, 0)


def volumetric_main(args):
    import numpy as np
    import torch
    import oracle
    from illcrawl_amd import backend as B, synth
    t0 = time.time()
    snap = synth.cached("c2")
    nx, width = 2048, 3000.0
    pts = volumetric_slices(snap, nx, width)
    log(f"[bench] v1: {snap.n} cells, {pts.shape[0] // (nx * nx)} slices of {nx}^2 "
        f"({time.time() - t0:.1f}s)")
    dev = torch.device("cuda", 0)
    p4 = torch.from_numpy(snap.particles4()).to(dev)
    h = torch.from_numpy(snap.hsml).to(dev)
    pts_d = torch.from_numpy(pts).to(dev)
    out = torch.empty(pts.shape[0], dtype=torch.float32, device=dev)
    g = B.GridSmoothing(mode=args.mode, kernel="quartic_ref" if args.mode == "reference"
                        else "quartic")
    g.init_backend(snap.n)
    g.setup_particles(p4, [h])
    g.setup_evaluation_points(pts_d, layout_nx=nx)
    total, per = g.count_evals(per_point=True)
    log(f"[bench] exact evals per step: {total:.6g}")

    def step(ev=None):
        g.setup_particles(p4, [h])
        g.setup_evaluation_points(pts_d, layout_nx=nx)
        if ev:
            ev[0].record()
        g.run()
        if ev:
            ev[1].record()
        return g.retrieve_results()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    gather_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    log(f"[bench] v1: {ms_per_step:.2f} ms/step, gather {gather_ms:.2f} ms")
    achieved = total * FLOP_PER_EVAL_3D / (gather_ms * 1e-3) / 1e12
    traffic = MEASURED_TRAFFIC.get(("v1", args.mode, (16, 16)))
    gi = g.grid_info()
    line = {
        "metric": "particle-kernel evals/s, volumetric SPH smoothing (3-D grid) on MI355X",
        "value": total / (ms_per_step * 1e-3), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 per-point accumulation)",
        "data": "synthetic (seeded: uniform + NFW halos, 32-NN density, lognormal masses)",
        "config": dict(workload="volumetric smoothing: gas density on 8 slices through the "
                                "halos of the 16M-cell snapshot", n_cells=snap.n,
                       box_kpc_h=snap.box, slices=len(snap.halos), resolution=[nx, nx],
                       width_kpc_h=width, mode=args.mode, particles_per_cell=64,
                       evals_per_step=total,
                       points_per_step=int(pts.shape[0]), grid_cells=gi["num_cells"],
                       parallelism="single GPU"),
        "points_per_s": pts.shape[0] / (ms_per_step * 1e-3),
        "gather_ms": gather_ms,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic["read"] + traffic["write"] if traffic else None,
                     "traffic_split": traffic,
                     "kernel": "k_smooth3d", "flop_per_eval": FLOP_PER_EVAL_3D,
                     "evals_per_launch": total},
    }
    rng = np.random.default_rng(5)
    ocache = {}

    def ogrid():
        if "g" not in ocache:
            t = time.perf_counter()
            ocache["g"] = oracle.Grid(snap.particles4(), snap.hsml, target=64)
            ocache["t"] = time.perf_counter() - t
        return ocache["g"]

    if not args.no_parity:
        log("[bench] sampled parity vs oracle ...")
        img = g.retrieve_results_f64().cpu().numpy()
        idx = rng.choice(pts.shape[0], 3000, replace=False)
        ref = (ogrid().smooth3d_ref(pts[idx], form=1) if args.mode == "reference"
               else oracle.smooth3d_exact(snap.particles4(), snap.hsml, pts[idx]))
        m = ref > 1e-6 * ref.max()
        line["parity"] = dict(points=int(idx.size),
                              linf=float(np.max(np.abs(img[idx][m] - ref[m]) / ref[m])),
                              sum=float(abs(img[idx].sum() - ref.sum()) / ref.sum()))
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count())
        oracle.set_num_threads(threads)
        og = ogrid()
        counts = per.cpu().numpy()
        order = rng.permutation(pts.shape[0])
        done, evals, elapsed, batch = 0, 0, 0.0, 64
        while elapsed < args.cpu_seconds and done < order.size:
            sel = order[done:done + batch]
            t = time.perf_counter()
            og.smooth3d_ref(pts[sel], acc_mode=0)
            elapsed += time.perf_counter() - t
            evals += int(counts[sel].sum())
            done += sel.size
            batch = min(batch * 2, 16384)
        line["cpu_baseline"] = dict(
            value=evals / elapsed, unit="evals/s", cores=threads, kind="port",
            sample=f"{done} random points of the 8 slices ({elapsed:.1f} s of the reference "
                   f"3-D grid gather restated in C, FP32 accumulator, {threads} OpenMP "
                   f"threads); its grid build ({ocache['t']:.2f} s, one thread) not included")
    print(json.dumps(line), flush=True)


def breakdown_main(args, snap, views, kernel, tile):
    """Where the batched-vs-plain gap comes from: every resolution group of
    the workload rendered alone, batched (one persistent launch) and plain
    (one view per launch sequence), with the gather share of each step and
    the group's exact evals; then the whole batch."""
    import numpy as np
    import torch
    from illcrawl_amd import backend as B, synth
    tab, x0, dx = synth.gaunt_table_1d()
    groups = {}
    for k, v in enumerate(views):
        groups.setdefault((v["nx"], v["ny"]), []).append(k)
    groups["all"] = list(range(len(views)))

    def timed(r, plain, steps):
        r.render(plain=plain)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gs = []
        for _ in range(steps):
            r.render(plain=plain)
            gs.append(r.last_gather_ms())
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, float(np.mean(gs))

    out = {}
    for key, idx in groups.items():
        r = B.BatchRenderer(mode=args.mode, kernel=kernel, tile=tile)
        r.set_particles(snap.pos, snap.mass, snap.hsml, snap.rho, snap.temp, snap.xe)
        r.set_gaunt_table(tab, x0, dx)
        r.set_views([views[k] for k in idx])
        r.render()
        evals = int(sum(r.count_evals()))
        b_ms, b_g = timed(r, False, args.steps)
        p_ms, p_g = timed(r, True, max(1, args.plain_steps))
        name = key if key == "all" else f"{key[0]}x{key[1]}"
        out[name] = dict(views=len(idx), evals=evals, batched_ms=b_ms, batched_gather_ms=b_g,
                         plain_ms=p_ms, plain_gather_ms=p_g, ratio=p_ms / b_ms,
                         batched_gather_evals_per_s=evals / (b_g * 1e-3),
                         plain_gather_evals_per_s=evals / (p_g * 1e-3))
        log(f"[bench] {name}: batched {b_ms:.2f} ms (gather {b_g:.2f}), plain {p_ms:.2f} ms "
            f"(gathers {p_g:.2f}), evals {evals:.4g}")
        del r
        torch.cuda.empty_cache()
    print(json.dumps({"workload": args.workload, "tile": list(tile), "breakdown": out}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=os.environ.get("ILC_BENCH_WORKLOAD", "c3"))
    ap.add_argument("--mode", default="reference", choices=["reference", "exact"])
    ap.add_argument("--adjoint", action="store_true",
                    help="batched forward + adjoint timing of the workload (exact mode)")
    ap.add_argument("--kernel", default=None, choices=["quartic", "quartic_ref", "cubic"],
                    help="projected kernel: default quartic_ref (the reference's), quartic "
                         "for c4; a kernel other than quartic_ref renders in exact mode")
    ap.add_argument("--tile", default=None,
                    help="WxH[/G] gather tile (G threads per workgroup, 256 or 64); "
                         "default: autotuned")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--clip", action="store_true",
                    help="line-of-sight kernel clip of each view's depth range instead of the "
                         "reference's centre mask (exact mode; no plain leg)")
    ap.add_argument("--breakdown", action="store_true",
                    help="time every resolution group alone, batched and plain (diagnostic)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one device copy of MIB MiB (a known-traffic kernel for "
                         "scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              f"(views would shard across GPUs with no exchange between them)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from illcrawl_amd import backend as B, synth
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    if args.workload == "c4":
        args.kernel = args.kernel or "quartic"
        return adjoint_main(args)
    kernel = args.kernel or "quartic_ref"
    if kernel != "quartic_ref" and args.mode == "reference":
        log(f"[bench] the {kernel} kernel has no reference search: exact mode")
        args.mode = "exact"
    if args.clip:
        if args.mode == "reference":
            log("[bench] the line-of-sight clip has no reference search: exact mode")
        args.mode = "exact"
        args.plain_steps = 0
    if args.workload == "v1":
        return volumetric_main(args)
    if args.adjoint:
        return batch_adjoint_main(args)

    t0 = time.time()
    snap, views, desc = workload(args.workload)
    log(f"[bench] {args.workload}: {snap.n} cells, {len(views)} views ({time.time() - t0:.1f}s)")
    r = B.BatchRenderer(mode=args.mode, kernel=kernel)
    r.set_particles(snap.pos, snap.mass, snap.hsml, snap.rho, snap.temp, snap.xe)
    tab, x0, dx = synth.gaunt_table_1d()
    r.set_gaunt_table(tab, x0, dx)
    if args.clip:
        r.set_depth_clip(True)
        desc = dict(desc, depth_cut="line-of-sight kernel clip")
    r.set_views(views)
    if args.tile:
        tile = parse_tile(args.tile)
        r.set_tile(*tile)
        tuned = None
    else:
        _, best_ms = r.autotune()
        tile = r.tile()
        tuned = dict(tile=list(tile), best_ms=best_ms)
        log(f"[bench] autotuned tile {tile} ({best_ms:.1f} ms)")
    if args.breakdown:
        del r
        return breakdown_main(args, snap, views, kernel, tile)
    r.render()
    counts = r.count_evals()
    total_evals = int(sum(counts))
    log(f"[bench] exact evals per step: {total_evals:.6g}")
    for _ in range(args.warmup):
        r.render()
    torch.cuda.synchronize()

    # timed (the headline): hipGraph replays of the step's device work, one
    # persistent gather launch per step
    r.render(graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.render(graph=True)
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    log(f"[bench] batched, graph replay: {ms_per_step:.2f} ms/step")

    # the same steps launched eagerly, with HIP events around the gather
    # launch on its stream
    gather = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.render()
        gather.append(r.last_gather_ms())
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    gather_ms = float(np.mean(gather))
    log(f"[bench] batched, eager: {eager_ms:.2f} ms/step, gather {gather_ms:.2f} ms")

    # plain: one view per launch sequence, at the tile shape that is fastest
    # for the plain path itself (timed over every shape, one render each),
    # not at the batch's
    plain_ms, plain_gather, plain_tile = None, None, None
    if args.plain_steps > 0:
        batch_tile = r.tile()
        best = None
        for cand in PLAIN_TILES:
            r.set_tile(*cand)
            r.render(plain=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r.render(plain=True)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            if best is None or ms < best[0]:
                best = (ms, cand)
        plain_tile = best[1]
        r.set_tile(*plain_tile)
        r.render(plain=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pg = []
        for _ in range(args.plain_steps):
            r.render(plain=True)
            pg.append(r.last_gather_ms())
        torch.cuda.synchronize()
        plain_ms = (time.perf_counter() - t0) * 1e3 / args.plain_steps
        plain_gather = float(np.mean(pg))
        log(f"[bench] plain: {plain_ms:.2f} ms/step, gathers {plain_gather:.2f} ms "
            f"(its fastest tile {plain_tile})")
        r.set_tile(*batch_tile)
        r.render()

    value = total_evals / (ms_per_step * 1e-3)
    traffic = MEASURED_TRAFFIC.get((args.workload, args.mode, tuple(tile)))
    achieved = total_evals * FLOP_PER_EVAL / (gather_ms * 1e-3) / 1e12
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
        "dtype": "f32 (f64 per-pixel accumulation)",
        "data": "synthetic (seeded: uniform + NFW halos, 32-NN density, lognormal masses)",
        "config": dict(desc, mode=args.mode,
                       kernel="quartic_ref (the reference's projected quartic)"
                       if kernel == "quartic_ref" else kernel, tile=list(tile), autotune=tuned,
                       evals_per_step=total_evals, pixels_per_step=npix,
                       parallelism="single GPU"),
        "pixels_per_s": npix / (ms_per_step * 1e-3),
        "batched_vs_plain": (plain_ms / ms_per_step) if plain_ms else None,
        "plain": dict(ms_per_step=plain_ms, evals_per_s=(total_evals / (plain_ms * 1e-3))
                      if plain_ms else None, gather_ms_sum=plain_gather,
                      tile=list(plain_tile) if plain_tile else None),
        "timing": "hipGraph replay of each step; eager launches below",
        "eager": dict(ms_per_step=eager_ms, evals_per_s=total_evals / (eager_ms * 1e-3)),
        "gather_ms": gather_ms,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic["read"] + traffic["write"] if traffic else None,
                     "traffic_split": traffic,
                     "kernel": "k_gather_batch (persistent gather, all views)",
                     "flop_per_eval": FLOP_PER_EVAL, "evals_per_launch": total_evals},
    }
    if args.copy_calibration > 0:
        nb = args.copy_calibration << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        line["copy_calibration_bytes"] = nb
        del src, dst
    big = max(range(len(views)), key=lambda k: views[k]["nx"] * views[k]["ny"])
    if args.clip:
        big = len(views) // 2
        if not args.no_parity:
            log("[bench] sampled parity vs the oracle's clipped quadrature ...")
            line["parity"], cpu = clip_parity(snap, views[big], r.image_f64(big).cpu().numpy())
            line["parity"]["view"] = big
            if not args.no_cpu_baseline:
                line["cpu_baseline"] = cpu
        print(json.dumps(line), flush=True)
        return
    if not args.no_parity:
        log("[bench] sampled parity vs oracle ...")
        img = r.image_f64(big).cpu().numpy()
        line["parity"] = dict(view=big, **sampled_parity(snap, views[big], img, kernel=kernel,
                                                         mode=args.mode))
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        _, per = single_view_counts(snap, views[big], kernel)
        line["cpu_baseline"] = cpu_baseline(snap, views[big], per, args.cpu_seconds, kernel)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
