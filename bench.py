"""Benchmark: eigenpairs/sec (k=50) of the FEM hot path on one MI355X.

Workload (one "step"): icosphere level 8 scaled to a seeded ellipsoid (axes
U[0.5, 2]^3, radial noise 0.01; 655,362 vertices, 1,310,720 triangles).
The mesh arrays are resident in HBM when the timed region starts; a step
copies them into a fresh device mesh (so the symbolic phase runs every
step), assembles A and B in FP64 (lapy Solver(mesh, lump=False)) and
computes the 50 smallest eigenpairs of A x = lambda B x (Solver.eigs(k=50));
the eigenpairs stay in HBM.  value = 50 * steps / timed seconds.

Also reported: the eigenvalue error against FP64 oracle eigenvalues of the
same mesh (tests/golden/ellipsoid_l8_k50.json, written by the oracle in the
build container), the roofline of the dominant kernel measured with HIP
events, and the CPU baseline (the oracle pipeline on a bounded sample).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_eigs_seconds(v, t, k, threads):
    """One run of the oracle pipeline (numpy assembly, scipy splu + eigsh
    shift-invert, as lapy/solver.py:106-194 and :667-716) with the BLAS/OpenMP
    pools limited to `threads` (None: all).  Returns (assembly s, factor+eigsh s)."""
    import threadpoolctl

    from oracle import fem

    with threadpoolctl.threadpool_limits(threads):
        t0 = time.perf_counter()
        a, b = fem.fem_tria(v, t)
        t1 = time.perf_counter()
        fem.eigs(a, b, k=k)
        t2 = time.perf_counter()
    return t1 - t0, t2 - t1


def _oracle_mesh_seconds(v, t, k):
    """Oracle pipeline on one triangle mesh with one BLAS thread (a pool
    worker of the batch baseline; module level so that spawn can pickle it)."""
    a, e = _oracle_eigs_seconds(v, t, k, 1)
    return a + e


def _host_info():
    return f"CPU: {_cpu_model()}, os.cpu_count()={os.cpu_count()}, sched_getaffinity={len(os.sched_getaffinity(0))}"


def cpu_baseline_small(v, t, k, runs=5):
    """The oracle on the small sphere itself: median of `runs` runs, one thread."""
    secs = sorted(sum(_oracle_eigs_seconds(v, t, k, 1)) for _ in range(runs))
    med = secs[len(secs) // 2]
    return {"value": round(k / med, 4), "unit": "eigenpairs/s", "cores": 1, "kind": "port",
            "sample": f"oracle pipeline (numpy assembly + splu + eigsh shift-invert) on the same L5 sphere "
                      f"({len(v)} vertices), k={k}, median of {runs} runs ({med:.3f} s), BLAS threads=1. "
                      + _host_info()}


def cpu_baseline_tet(k, cells=24):
    """The oracle on a bounded tetrahedral cube of `cells`^3 cells (the same
    construction as the 64^3 workload): assembly, splu + eigsh, and the heat
    solve timed separately, one thread.  SuperLU's fill on a 3-D Laplacian
    grows much faster than the vertex count, so the full 64^3 cube is not a
    bounded sample."""
    import threadpoolctl

    from lapy_amd import meshgen
    from oracle import fem

    v, t = meshgen.volume_cube(cells=cells)
    with threadpoolctl.threadpool_limits(1):
        t0 = time.perf_counter()
        a, b = fem.fem_tetra(v, t)
        t1 = time.perf_counter()
        fem.eigs(a, b, k=k)
        t2 = time.perf_counter()
        fem.diffusion(v, t, 0, m=1e-3 / fem.avg_edge_length(v, t) ** 2)
        t3 = time.perf_counter()
    return {"value": round(k / (t2 - t0), 4), "unit": "eigenpairs/s", "cores": 1, "kind": "port",
            "sample": f"oracle pipeline on the {cells}^3-cell tetrahedral cube ({len(v)} vertices, {len(t)} tets), "
                      f"k={k}, one run, BLAS threads=1: assembly {t1 - t0:.2f} s, splu + eigsh {t2 - t1:.2f} s, "
                      f"heat solve (splu of M + tA) {t3 - t2:.2f} s. " + _host_info(),
            "heat_seconds": round(t3 - t2, 4)}


def _stratified(meshes, stride):
    """Every `stride`-th mesh of each level (vertex count) of the batch, so
    that the sample's level mix is the batch's."""
    by_level = {}
    for v, t in meshes:
        by_level.setdefault(len(v), []).append((v, t))
    return [m for nv in sorted(by_level) for m in by_level[nv][::stride]]


def cpu_baseline_batch(meshes, k, seq_stride=16, pool_stride=4, max_procs=16):
    """The oracle on level-stratified samples of the batch: the reference's
    sequential loop (one thread) over every `seq_stride`-th mesh of each
    level, and a spawn pool of one-thread workers over the host's cores (at
    most `max_procs`) on every `pool_stride`-th mesh of each level, enough
    meshes per worker to balance the sizes."""
    import multiprocessing as mp

    seq_s = _stratified(meshes, seq_stride)
    t0 = time.perf_counter()
    for v, t in seq_s:
        _oracle_mesh_seconds(v, t, k)
    seq = time.perf_counter() - t0
    pool_s = _stratified(meshes, pool_stride)
    procs = max(1, min(max_procs, len(os.sched_getaffinity(0))))
    with mp.get_context("spawn").Pool(procs) as pool:
        pool.starmap(_oracle_mesh_seconds, [(seq_s[0][0], seq_s[0][1], 2)] * procs)  # workers import scipy
        t0 = time.perf_counter()
        pool.starmap(_oracle_mesh_seconds, [(v, t, k) for v, t in pool_s[::-1]], chunksize=1)  # largest first
        par = time.perf_counter() - t0
    avg = sum(len(v) for v, _ in meshes) / len(meshes)
    return {"value": round(k * len(seq_s) / seq, 4), "unit": "eigenpairs/s", "cores": 1, "kind": "port",
            "sample": f"oracle pipeline (as lapy shapedna.compute_shapedna per mesh), k={k}, on level-stratified "
                      f"samples of the batch ({avg:.0f} vertices per mesh on average): sequential loop over "
                      f"{len(seq_s)} meshes ({sum(len(v) for v, _ in seq_s) / len(seq_s):.0f} vertices on average) "
                      f"{seq:.2f} s; pool of {procs} one-thread workers over {len(pool_s)} meshes "
                      f"({sum(len(v) for v, _ in pool_s) / len(pool_s):.0f} on average) {par:.2f} s. " + _host_info(),
            "pool": {"value": round(k * len(pool_s) / par, 4), "processes": procs}}


def cpu_baseline(level: int, k: int, thread_level: int):
    """The oracle timed on this host's cores, on the metric's own mesh
    (icosphere level `level`, one run with BLAS threads = 1: SuperLU, ARPACK
    and numpy's element arithmetic are single-threaded, which is how LaPy runs).
    The all-cores leg (BLAS and OpenMP pools unrestricted) is timed beside a
    second one-thread run on the smaller level `thread_level`, to show what
    extra threads give the reference's path."""
    from lapy_amd import meshgen

    v, t = meshgen.large_ellipsoid(level=level)
    a1, e1 = _oracle_eigs_seconds(v, t, k, 1)
    vs, ts = meshgen.large_ellipsoid(level=thread_level)
    sa1, se1 = _oracle_eigs_seconds(vs, ts, k, 1)
    san, sen = _oracle_eigs_seconds(vs, ts, k, None)
    ncpu = os.cpu_count()
    aff = len(os.sched_getaffinity(0))
    return {
        "value": round(k / (a1 + e1), 4),
        "unit": "eigenpairs/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle pipeline (numpy assembly + scipy splu + eigsh shift-invert, as lapy Solver.eigs) on the "
                   f"metric's mesh, icosphere level {level} ellipsoid ({len(v)} vertices), k={k}, one run, BLAS "
                   f"threads=1: assembly {a1:.2f} s, factor+eigsh {e1:.2f} s. All-cores check on level "
                   f"{thread_level} ({len(vs)} vertices): {k / (sa1 + se1):.3f} eigenpairs/s with 1 thread, "
                   f"{k / (san + sen):.3f} with all {aff} threads. CPU: {_cpu_model()}, os.cpu_count()={ncpu}, "
                   f"sched_getaffinity={aff}"),
        "all_cores": {"value": round(k / (san + sen), 4), "threads": aff, "level": thread_level,
                      "one_thread_value": round(k / (sa1 + se1), 4)},
    }


def profiled_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 summary (FETCH_SIZE +
    WRITE_SIZE, copy-calibrated), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def golden(name):
    path = os.path.join(ROOT, "tests", "golden", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def rel_err(evals, gold_evals):
    gold = np.asarray(gold_evals)
    scale = np.abs(gold).max()
    den = np.where(np.abs(gold) < 1e-6 * scale, scale, np.abs(gold))
    return float(np.max(np.abs(np.asarray(evals) - gold) / den))


def fem_oracle_diffusion(v, t):
    from oracle import fem

    return fem.diffusion(v, t, 0, m=1.0)


def run_workload(args, ctx, _lib, meshgen):
    """The other workloads of the benchmark plan (one JSON line each; the
    default L8 line is the contract line):
      small: icosphere L5 sphere (10,242 vertices), radial noise 0.02, k = 10;
      tet:   64^3-cell tetrahedral cube (274,625 vertices), k = 30, then heat
             diffusion from vertex 0 with t = 1e-3 (m = 1e-3 / avg_edge^2),
             PCG to ||r|| / ||b|| <= 1e-10;
      pcie:  the L8 workload from mesh arrays in host memory to eigenvalues
             and eigenvectors in host memory (upload, assembly, eigensolve,
             download);
      batch: 256 seeded ellipsoids of icosphere levels 3-6, k = 50 each,
             assembled per mesh and solved by the batched eigensolver with
             the autotuned worker count;
      geodesic: heat-method distance on the L8 ellipsoid from vertex 0: heat
             diffusion (m = 1), then compute_geodesic_f (unit gradient,
             divergence, Poisson solve by PCG), host arrays in and out."""
    from lapy_amd import TetMesh, heat

    out = {"unit": "eigenpairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic"}
    if args.workload in ("small", "tet"):
        if args.workload == "small":
            v, t = meshgen.small_sphere(level=5)
            kind, k, gold = _lib.LH_TRIA, 10, golden("sphere_l5_k10.json")
            name = "icosphere L5 sphere (radial noise 0.02), FP64 assembly + k=10 eigenpairs"
        else:
            v, t = meshgen.volume_cube(cells=64)
            kind, k, gold = _lib.LH_TET, 30, golden("volume_c64_k30.json")
            name = "64^3-cell tetrahedral cube, FP64 assembly + k=30 eigenpairs"
        base = _lib.DeviceMesh(ctx, v, t)

        def step():
            m = base.clone()
            A, B = m.assemble(kind, False, np.float64)
            ev, eh, st = _lib.eigs(ctx, A, B, k, -0.01, keep_device=True)
            return ev, eh, st, (m, A, B)

        for _ in range(args.warmup):
            ev, eh, st, keep = step()
            ctx.sync()
            del eh, keep
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ev, eh, st, keep = step()
            ctx.sync()
            del eh, keep
        total = time.perf_counter() - t0
        out.update({"metric": f"eigenpairs/sec (k={k})", "value": round(k * args.steps / total, 4),
                    "ms_per_step": round(1000 * total / args.steps, 2),
                    "config": {"workload": name, "vertices": len(v), "elements": len(t), "k": k,
                               "lobpcg_iterations": st.iterations, "max_rel_resid": st.max_resid,
                               "eig_rel_err_vs_oracle": rel_err(ev, gold["evals"]) if gold else None}})
        if args.workload == "tet":
            mesh = TetMesh(v, t)
            ael = float(mesh.avg_edge_length())
            mh = 1e-3 / (ael * ael)
            heat.diffusion(mesh, 0, m=mh, rtol=1e-10)  # warm the device mesh
            t1 = time.perf_counter()
            u = heat.diffusion(mesh, 0, m=mh, rtol=1e-10)
            th = time.perf_counter() - t1
            herr = None
            if gold:
                idx = np.asarray(gold["probe_index"])
                ref = np.asarray(gold["heat_probe_values"])
                herr = float(np.max(np.abs(u[idx] - ref)) / np.max(np.abs(ref)))
            out["config"].update({"heat_seconds": round(th, 4), "heat_rtol": 1e-10, "heat_probe_rel_err": herr})
        if not args.no_cpu_baseline:
            out["cpu_baseline"] = cpu_baseline_small(v, t, k) if args.workload == "small" else cpu_baseline_tet(k)
    elif args.workload == "geodesic":
        from lapy_amd import TriaMesh, diffgeo

        v, t = meshgen.large_ellipsoid(level=args.level)
        mesh = TriaMesh(v, t)
        u = heat.diffusion(mesh, 0, m=1.0)
        g = diffgeo.compute_geodesic_f(mesh, u)  # warm: device mesh, topology
        for _ in range(args.warmup):
            g = diffgeo.compute_geodesic_f(mesh, u)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g = diffgeo.compute_geodesic_f(mesh, u)
        total = time.perf_counter() - t0
        out.update({"metric": "geodesic distance fields/sec (heat method)", "unit": "fields/s",
                    "value": round(args.steps / total, 4), "ms_per_step": round(1000 * total / args.steps, 2),
                    "config": {"workload": f"icosphere L{args.level} ellipsoid, geodesic distance from vertex 0 "
                                           "(host arrays in and out: PCIe inclusive)",
                               "vertices": len(v), "elements": len(t), "source_distance": float(g[0]),
                               "max_distance": float(g.max())}})
        if not args.no_cpu_baseline:
            from oracle import diffgeo as odg

            vs, ts = meshgen.large_ellipsoid(level=6)
            us = fem_oracle_diffusion(vs, ts)
            c0 = time.perf_counter()
            odg.geodesic_f(vs, ts, us)
            ct = time.perf_counter() - c0
            out["cpu_baseline"] = {"value": round(1.0 / ct, 4), "unit": "fields/s", "cores": 1, "kind": "port",
                                   "sample": f"oracle geodesic_f (splu) on the L6 ellipsoid, {len(vs)} vertices"}
    elif args.workload == "batch":
        nb = args.batch_count
        meshes = meshgen.shape_batch(count=nb)
        k = 50
        gold = golden("batch_256_k50.json") if nb == 256 else None
        base = [_lib.DeviceMesh(ctx, v, t) for v, t in meshes]
        # preconditioner: Chebyshev-Jacobi (default), plain Jacobi, or the
        # learned network, trained first (untimed) on six other meshes of
        # the same family (levels 3 and 4, seeds outside the batch)
        popts = {"cheb": {}, "jacobi": {"poly_degree": 1},
                 "learned": {"precond": _lib.LH_PRECOND_LEARNED},
                 "auto": {"precond": _lib.LH_PRECOND_AUTO}}[args.precond]
        net = None
        if args.precond in ("learned", "auto"):
            from lapy_amd import learned

            t0 = time.perf_counter()
            net = learned.LearnedPreconditioner(ctx=ctx)
            train = meshgen.shape_batch(count=6, seed=50_000, levels=(3, 4))
            hist = net.train(train, steps=300, ncol=8, lr=3e-3, seed=1)
            net.use()
            log(f"[bench] learned preconditioner trained: loss {hist[:10].mean():.3f} -> {hist[-30:].mean():.3f} "
                f"({time.perf_counter() - t0:.1f} s)")

        def assemble_all():
            ms = [b.clone() for b in base]
            ab = [m.assemble(_lib.LH_TRIA, False, np.float64) for m in ms]
            return ms, [a for a, _ in ab], [b for _, b in ab]

        # autotune the worker count on a sample of every eighth mesh (unless
        # a count is given)
        if args.batch_workers > 0:
            best, secs = args.batch_workers, None
        else:
            ms, As, Bs = assemble_all()
            sample = list(range(0, nb, 8))
            t0 = time.perf_counter()
            best, secs = _lib.eigs_autotune(ctx, [As[i] for i in sample], [Bs[i] for i in sample],
                                            [k] * len(sample), **popts)
            log(f"[bench] autotune: best workers {best}, sample seconds {secs} ({time.perf_counter() - t0:.1f} s)")
            del ms, As, Bs
        for w in range(args.warmup):
            t1 = time.perf_counter()
            ms, As, Bs = assemble_all()
            res = _lib.eigs_batched(ctx, As, Bs, [k] * nb, workers=best, keep_device=True, **popts)
            ctx.sync()
            del res, ms, As, Bs
            log(f"[bench] batch warmup {w}: {time.perf_counter() - t1:.2f} s")
        t0 = time.perf_counter()
        for st_i in range(args.steps):
            t1 = time.perf_counter()
            ms, As, Bs = assemble_all()
            res = _lib.eigs_batched(ctx, As, Bs, [k] * nb, workers=best, keep_device=True, **popts)
            ctx.sync()
            log(f"[bench] batch step {st_i}: {time.perf_counter() - t1:.2f} s")
        total = time.perf_counter() - t0
        err = None
        if gold:
            err = max(rel_err(r[0], g) for r, g in zip(res, gold["evals"]))
        its = [r[2].iterations for r in res]
        out.update({"metric": f"eigenpairs/sec (k=50), {nb}-mesh ShapeDNA batch",
                    "value": round(nb * k * args.steps / total, 4), "ms_per_step": round(1000 * total / args.steps, 2),
                    "config": {"workload": f"{nb} ellipsoids, icosphere levels 3-6, FP64 assembly + k=50 eigenpairs each",
                               "vertices_total": int(sum(len(v) for v, _ in meshes)), "k": k, "workers": best,
                               "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
                               "precond": args.precond,
                               "fallbacks": int(sum(r[2].fallback_used for r in res)),
                               "autotune_seconds": secs, "lobpcg_iterations_min_max": [min(its), max(its)],
                               "eig_rel_err_vs_oracle": err}})
        if not args.no_cpu_baseline:
            out["cpu_baseline"] = cpu_baseline_batch(meshes, k)
    elif args.workload == "learned":
        run_learned(args, ctx, _lib, meshgen, out)
        return
    else:  # pcie
        v, t = meshgen.large_ellipsoid(level=8)
        k = 50
        gold = golden("ellipsoid_l8_k50.json")

        def step():
            m = _lib.DeviceMesh(ctx, v, t)  # host -> device
            A, B = m.assemble(_lib.LH_TRIA, False, np.float64)
            return _lib.eigs(ctx, A, B, k, -0.01)  # eigenvalues and eigenvectors downloaded

        for _ in range(args.warmup):
            step()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ev, vec, st = step()
        total = time.perf_counter() - t0
        out.update({"metric": "eigenpairs/sec (k=50), host arrays to host eigenpairs",
                    "value": round(k * args.steps / total, 4), "ms_per_step": round(1000 * total / args.steps, 2),
                    "config": {"workload": "icosphere L8 ellipsoid, host mesh -> host eigenpairs (PCIe inclusive)",
                               "vertices": len(v), "k": k, "lobpcg_iterations": st.iterations,
                               "eigenvector_bytes_downloaded": int(vec.nbytes),
                               "eig_rel_err_vs_oracle": rel_err(ev, gold["evals"]) if gold else None}})
    print(json.dumps(out), flush=True)


def run_learned(args, ctx, _lib, meshgen, out):
    """The learned preconditioner at its specified scale: trained on the
    solver's own residuals of 512 seeded meshes of icosphere levels 4-7, then
    applied inside the batched LOBPCG (k = 50) and the heat PCG on 64
    held-out level-7 meshes (163,842 vertices each), against plain Jacobi
    (the fallback and the plain form) and Chebyshev-Jacobi (the default).
    Every solve ends at the same FP64 residual test, so the eigenvalues of
    all forms agree; the line reports throughput, iterations and fallbacks of
    each form and the roofline of the network's layer kernel."""
    from lapy_amd import TriaMesh, learned

    k = 50
    t0 = time.perf_counter()
    net = learned.LearnedPreconditioner(ctx=ctx)
    if args.kernels_only:  # only the layer kernel's timing (for the profiler)
        v, t = meshgen.held_out_mesh(0)
        A0, B0 = _lib.DeviceMesh(ctx, v, t).assemble(_lib.LH_TRIA, False, np.float64)
        us = net.net.time_layer(A0, B0, 0.01, 1.0, 64, iters=30)
        us_u = net.net.time_layer(A0, B0, 0.01, 1.0, 64, iters=30, reorder=False)
        print(json.dumps({"kernel": "k_mp_fwd", "avg_us_locality_order": us, "avg_us_mesh_order": us_u,
                          "vertices": A0.n, "columns": 64}), flush=True)
        return
    hist = net.train_collection(meshgen.training_mesh, args.train_count, epochs=1, steps_per_mesh=2, ncol=8,
                                lr=3e-3, seed=1, log=log)
    train_s = time.perf_counter() - t0
    log(f"[bench] trained on {args.train_count} meshes in {train_s:.1f} s: loss {np.mean(hist[:32]):.3f} -> "
        f"{np.mean(hist[-64:]):.3f}")
    held = [meshgen.held_out_mesh(i) for i in range(args.eval_count)]
    dms = [_lib.DeviceMesh(ctx, v, t) for v, t in held]
    mats = [m.assemble(_lib.LH_TRIA, False, np.float64) for m in dms]
    # roofline of the layer kernel: one forward layer on an n x 64 block
    A0, B0 = mats[0]
    n0 = A0.n
    nnz0 = A0.nnz
    us_layer = net.net.time_layer(A0, B0, 0.01, 1.0, 64, iters=30)  # the solver's locality order
    us_layer_mesh_order = net.net.time_layer(A0, B0, 0.01, 1.0, 64, iters=30, reorder=False)
    pairs = n0 * 64
    # algorithmic bytes: H_l read once and H_{l+1} written (64 B each per
    # pair, bfloat16), the graph (rowptr, columns, FP32 weights) and the
    # vertex features; flops: the two 32 x 32 products per pair
    layer_bytes = 128 * pairs + 4 * (n0 + 1) + 8 * nnz0 + 16 * n0
    layer_flops = 2 * 2 * 32 * 32 * pairs
    ach = layer_bytes / (us_layer * 1e-6) / 1e9
    log(f"[bench] layer kernel: {us_layer:.1f} us, {ach:.0f} GB/s, {layer_flops / (us_layer * 1e-6) / 1e12:.1f} TFLOP/s")
    net.use()
    # (name, options, meshes): the form without the fallback runs on the
    # first chunk only, to show how the network itself fares
    chunk = 16
    nall = len(mats)
    forms = [("chebyshev_fp64", {}, nall), ("jacobi_fp64", {"poly_degree": 1}, nall),
             ("learned_fp64", {"precond": _lib.LH_PRECOND_LEARNED}, nall),
             ("learned_mixed", {"precond": _lib.LH_PRECOND_LEARNED, "precision": _lib.LH_PREC_MIXED}, nall),
             ("jacobi_mixed", {"poly_degree": 1, "precision": _lib.LH_PREC_MIXED}, nall),
             ("learned_fp64_no_fallback", {"precond": _lib.LH_PRECOND_LEARNED, "no_fallback": True,
                                           "max_iter": 2500}, min(chunk, nall))]
    results = {}
    ref = None
    first_chunk = {}
    for name, opts, cnt in forms:
        t1 = time.perf_counter()
        evs, its, fb, nconv = [], [], 0, 0
        for c0 in range(0, cnt, chunk):
            As = [a for a, _ in mats[c0:c0 + chunk]]
            Bs = [b for _, b in mats[c0:c0 + chunk]]
            try:
                res = _lib.eigs_batched(ctx, As, Bs, [k] * len(As), workers=16, keep_device=True, **opts)
            except _lib.NoConvergence as e:  # some problem stopped at max_iter: its last Ritz values
                res = e.results
            ctx.sync()
            for r in res:
                evs.append(r[0])
                its.append(r[2].iterations)
                fb += int(r[2].fallback_used)
                nconv += int(r[2].converged < k)
            del res
            if c0 == 0:
                first_chunk[name] = {"seconds": round(time.perf_counter() - t1, 3),
                                     "lobpcg_iterations_median": int(np.median(its))}
            log(f"[bench] {name}: {c0 + len(As)}/{cnt} meshes, {time.perf_counter() - t1:.1f} s")
        dt = time.perf_counter() - t1
        if ref is None:
            ref = evs
        err = max(rel_err(e, r) for e, r in zip(evs, ref))
        results[name] = {"meshes": cnt, "eigenpairs_per_s": round(k * cnt / dt, 3), "seconds": round(dt, 3),
                         "lobpcg_iterations_min_median_max": [int(min(its)), int(np.median(its)), int(max(its))],
                         "fallbacks_to_jacobi": fb, "not_converged": nconv,
                         "eig_rel_diff_vs_chebyshev_fp64": err}
    # heat PCG on each held-out mesh (lumped mass, t = avg edge length^2)
    heat = {"jacobi": [0, 0.0], "learned": [0, 0.0]}
    hfb = 0
    for (v, t), dm in zip(held, dms):
        Al, Bl = dm.assemble(_lib.LH_TRIA, True, np.float64)
        tt = TriaMesh(v, t).avg_edge_length() ** 2
        b = np.zeros(Al.n)
        b[0] = 1.0
        for name, pc in (("jacobi", _lib.LH_PRECOND_JACOBI), ("learned", _lib.LH_PRECOND_LEARNED)):
            t1 = time.perf_counter()
            _, st = _lib.pcg(ctx, Al, Bl, 1.0, tt, b, rtol=1e-10, precond=pc)
            heat[name][0] += st.iterations
            heat[name][1] += time.perf_counter() - t1
            if name == "learned":
                hfb += int(st.fallback_used)
    net.use(False)
    m = len(mats)
    heat_out = {name: {"mean_iterations": round(v[0] / m, 2), "mean_ms": round(1000 * v[1] / m, 3)}
                for name, v in heat.items()}
    heat_out["learned"]["fallbacks_to_jacobi"] = hfb
    plain = results["jacobi_fp64"]["eigenpairs_per_s"]
    out.update({
        "metric": "eigenpairs/sec (k=50), 64 held-out icosphere L7 meshes, learned preconditioner in batched LOBPCG",
        "value": results["learned_fp64"]["eigenpairs_per_s"], "ms_per_step": round(1000 * results["learned_fp64"][
            "seconds"], 2), "steps": 1, "warmup": 0, "dtype": "f64 (network bf16, fp32 accumulation)",
        "config": {"workload": f"train on {args.train_count} meshes (icosphere levels 4-7, solver residuals), "
                               f"evaluate on {m} held-out level-7 meshes ({n0} vertices each), k={k}",
                   "train_seconds": round(train_s, 2), "train_steps": int(len(hist)),
                   "train_loss_first_last": [float(np.mean(hist[:32])), float(np.mean(hist[-64:]))],
                   "forms": results,
                   "speedup_vs_plain_fp64_jacobi": {name: round(r["eigenpairs_per_s"] / plain, 3)
                                                    for name, r in results.items()},
                   "first_16_meshes": first_chunk,
                   "heat_pcg_rtol_1e-10": heat_out},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_mp_fwd (one forward layer, 16-pair tiles on v_mfma_f32_16x16x32_bf16), m=64, "
                               "in the eigensolver's locality order",
                     "avg_us": round(us_layer, 2), "avg_us_in_mesh_order": round(us_layer_mesh_order, 2),
                     "algorithmic_bytes": int(layer_bytes),
                     "tflops_bf16": round(layer_flops / (us_layer * 1e-6) / 1e12, 2)},
    })
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--k", type=int, default=50)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--tol", type=float, default=0.0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--poly-degree", type=int, default=0)
    ap.add_argument("--poly-ratio", type=float, default=0.0)
    ap.add_argument("--poly-storage", default="auto", choices=["auto", "f32", "f16"])
    ap.add_argument("--cpu-level", type=int, default=8,
                    help="icosphere level of the CPU baseline's main run (the metric's mesh: 8)")
    ap.add_argument("--cpu-thread-level", type=int, default=7,
                    help="level of the CPU baseline's one-thread vs all-cores comparison")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the PCIe-inclusive and mixed-precision figures of the L8 line")
    ap.add_argument("--verbose", type=int, default=0)
    ap.add_argument("--train-count", type=int, default=512, help="learned workload: training meshes")
    ap.add_argument("--eval-count", type=int, default=64, help="learned workload: held-out level-7 meshes")
    ap.add_argument("--workload", default="l8",
                    choices=["l8", "small", "tet", "pcie", "batch", "geodesic", "learned"],
                    help="l8: the benchmark line (default); small / tet / pcie / batch: the other workloads")
    ap.add_argument("--batch-count", type=int, default=256, help="batch workload: number of meshes")
    ap.add_argument("--batch-workers", type=int, default=0,
                    help="batch workload: solver workers (0: autotuned on every eighth mesh)")
    ap.add_argument("--precond", default="cheb", choices=["cheb", "jacobi", "learned", "auto"],
                    help="batch workload: Chebyshev-Jacobi (default), plain Jacobi, the learned network, or auto "
                         "(the network on the meshes where the polynomial would be plain Jacobi)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="skip the eigensolves; only the roofline timings of the two measured kernels")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run one device-to-device copy of 3 x 22,369,056 bytes (for PMC scale factors)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "lapy_amd runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    # the batched eigensolver runs one stream per worker: for the workloads
    # that use it, let HIP map up to 16 streams to hardware queues (its
    # default is 4, which serialises the streams sharing a queue); read by the
    # HIP runtime at its initialisation, so it is set before the library loads
    if args.workload in ("batch", "learned"):
        os.environ["GPU_MAX_HW_QUEUES"] = "16"
    from lapy_amd import _lib, meshgen

    ctx = _lib.default_context()
    if args.copy_calibration:
        # a device mesh clone is three plain device-to-device copies of the
        # vertex planes: 3 x 2,796,132 doubles = 3 x 22,369,056 bytes
        nvc = 2796132
        cm = _lib.DeviceMesh(ctx, np.random.default_rng(0).random((nvc, 3)), np.zeros((1, 3), np.int64))
        cm.clone()
        ctx.sync()
        print(json.dumps({"copy_calibration_bytes_per_copy": nvc * 8, "copies": 3}), flush=True)
        return
    if args.workload != "l8":
        run_workload(args, ctx, _lib, meshgen)
        return
    storage = {"auto": _lib.LH_STORE_AUTO, "f32": _lib.LH_STORE_F32, "f16": _lib.LH_STORE_F16}[args.poly_storage]
    v, t = meshgen.large_ellipsoid(level=args.level)
    nv, nf = len(v), len(t)
    log(f"[bench] mesh level {args.level}: V={nv} F={nf}")
    base = _lib.DeviceMesh(ctx, v, t)  # resident in HBM before timing
    prec = {"fp64": _lib.LH_PREC_FP64, "mixed": _lib.LH_PREC_MIXED}[args.precision]

    def step():
        m = base.clone()
        A, B = m.assemble(_lib.LH_TRIA, False, np.float64)
        evals, eh, st = _lib.eigs(ctx, A, B, args.k, -0.01, precision=prec, tol=args.tol, block=args.block,
                                  verbose=args.verbose, keep_device=True, poly_degree=args.poly_degree,
                                  poly_ratio=args.poly_ratio, poly_storage=storage)
        return evals, eh, st, (m, A, B)

    if args.kernels_only:
        args.warmup, args.steps = 0, 0
    for w in range(args.warmup):
        t0 = time.perf_counter()
        evals, eh, st, keep = step()
        ctx.sync()
        log(f"[bench] warmup {w}: {time.perf_counter() - t0:.3f} s, iterations {st.iterations}, "
            f"max resid {st.max_resid:.2e}")
        del eh, keep
    ctx.sync()
    times = []
    t_start = time.perf_counter()
    results = []
    for s in range(args.steps):
        t0 = time.perf_counter()
        evals, eh, st, keep = step()
        ctx.sync()
        dt = time.perf_counter() - t0
        times.append(dt)
        results.append((evals, st))
        log(f"[bench] step {s}: {dt:.3f} s, iterations {st.iterations}, rr {st.rr_calls}, spmm {st.spmm_calls}, "
            f"t_rr {st.t_rr_us / 1e6:.3f} s")
        del eh, keep
    ctx.sync()
    total = time.perf_counter() - t_start
    value = args.k * args.steps / total if args.steps else None
    evals, st = results[-1] if results else (None, None)

    # accuracy against the FP64 oracle eigenvalues of the same mesh
    gold_path = os.path.join(ROOT, "tests", "golden", f"ellipsoid_l{args.level}_k{args.k}.json")
    eig_err = None
    if evals is not None and os.path.exists(gold_path):
        with open(gold_path) as f:
            gold = np.array(json.load(f)["evals"])
        scale = np.abs(gold).max()
        den = np.where(np.abs(gold) < 1e-6 * scale, scale, np.abs(gold))
        eig_err = float(np.max(np.abs(evals - gold) / den))

    # extra figures beside the contract value (not part of it): the same
    # workload PCIe-inclusive (host mesh arrays in, host eigenpairs out), and
    # the mixed-precision solve on the same mesh
    extras = {}
    if args.steps and not args.no_extras:
        def pcie_step():
            mh = _lib.DeviceMesh(ctx, v, t)  # host -> device
            Ah, Bh = mh.assemble(_lib.LH_TRIA, False, np.float64)
            return _lib.eigs(ctx, Ah, Bh, args.k, -0.01, precision=prec, tol=args.tol, block=args.block,
                             poly_degree=args.poly_degree, poly_ratio=args.poly_ratio, poly_storage=storage)

        pcie_step()
        n_extra = 3
        t0 = time.perf_counter()
        for _ in range(n_extra):
            ev_p, vec_p, st_p = pcie_step()
        dt = time.perf_counter() - t0
        extras["pcie_inclusive"] = {
            "value": round(args.k * n_extra / dt, 4), "unit": "eigenpairs/s", "steps": n_extra,
            "ms_per_step": round(1000 * dt / n_extra, 2),
            "what": "host mesh arrays -> device mesh, symbolic + numeric assembly, eigensolve, eigenvalues and "
                    "eigenvectors downloaded to host memory",
            "eigenvector_bytes_downloaded": int(vec_p.nbytes),
            "eig_rel_err_vs_oracle": rel_err(ev_p, gold) if os.path.exists(gold_path) else None}
        log(f"[bench] pcie-inclusive: {1000 * dt / n_extra:.1f} ms per step")
        if args.precision == "fp64":
            def mixed_step():
                mm = base.clone()
                Am, Bm = mm.assemble(_lib.LH_TRIA, False, np.float64)
                ev_m, eh_m, st_m = _lib.eigs(ctx, Am, Bm, args.k, -0.01, precision=_lib.LH_PREC_MIXED,
                                             keep_device=True)
                ctx.sync()
                return ev_m, st_m

            mixed_step()
            t0 = time.perf_counter()
            for _ in range(n_extra):
                ev_m, st_m = mixed_step()
            dt = time.perf_counter() - t0
            mixed_err = None
            if os.path.exists(gold_path):
                mixed_err = rel_err(ev_m, gold)
            extras["mixed_precision"] = {
                "value": round(args.k * n_extra / dt, 4), "unit": "eigenpairs/s", "steps": n_extra,
                "ms_per_step": round(1000 * dt / n_extra, 2), "lobpcg_iterations": st_m.iterations,
                "lobpcg_iterations_fp32": st_m.iterations_fp32, "eig_rel_err_vs_oracle": mixed_err,
                "what": "same workload, precision='mixed' (FP32 phase, then FP64 to the same tolerance)"}
            log(f"[bench] mixed precision: {1000 * dt / n_extra:.1f} ms per step, "
                f"{st_m.iterations_fp32}/{st_m.iterations} iterations in FP32")

    # rooflines, kernel durations from HIP events on the library's stream
    m = base.clone()
    A, B = m.assemble(_lib.LH_TRIA, False, np.float64)
    nnz = A.nnz
    # dominant kernel: one Chebyshev step of the preconditioner (FP32 arithmetic, block of m = 64 columns)
    mcols = args.block if args.block else 64
    half = storage == _lib.LH_STORE_F16 or (storage == _lib.LH_STORE_AUTO and mcols % 8 == 0)
    us_c = _lib.time_cheb_step(ctx, A, B, mcols, iters=50, reorder=True,
                               storage=_lib.LH_STORE_F16 if half else _lib.LH_STORE_F32)
    # x_k (gathered, counted once), b, x_{k-1} read + x_{k+1} write: 4 values per entry of the block
    # (2 bytes each in FP16 storage, 4 in FP32); K values (FP32) + cols (int32): 8 nnz; row structure
    # (row pointers or ELL padding) + D^-1: 8 n
    es = 2 if half else 4
    cheb_bytes = 4 * es * nv * mcols + 8 * nnz + 8 * nv
    ach_c = cheb_bytes / (us_c * 1e-6) / 1e9
    kname = "k_cheb_step_e (ELL, FP16 storage)" if half else "k_cheb_step_v (FP32 storage)"
    # traffic: measured by rocprofv3 on exactly this configuration (L8, m = 64, FP16 storage)
    profiled = args.level == 8 and mcols == 64 and half
    roofline = {"bound": "hbm", "achieved": round(ach_c, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach_c / HBM_PEAK_GBS, 4),
                "traffic": profiled_traffic("cheb_step_e_l8_m64_traffic.txt") if profiled else None,
                "kernel": f"{kname}, Chebyshev-Jacobi preconditioner step, m={mcols}",
                "avg_us": round(us_c, 2), "algorithmic_bytes": int(cheb_bytes)}
    # the metric's second figure: numeric assembly
    # (rows in the pattern's locality order for meshes of >= 65,536 vertices, else in mesh order)
    us = m.time_assembly(_lib.LH_TRIA, False, A, B, iters=20)
    ctx.set_assembly_order(_lib.LH_ASM_ORDER_MESH)
    us_mesh = m.time_assembly(_lib.LH_TRIA, False, A, B, iters=20)
    ctx.set_assembly_order(_lib.LH_ASM_ORDER_AUTO)
    alg_bytes = 24 * nv + 12 * nf + 16 * nnz  # v (FP64 x3), t (int32 x3), A and B values (FP64)
    achieved = alg_bytes / (us * 1e-6) / 1e9
    sched = nv >= 65536
    roofline_asm = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": profiled_traffic("tria_rows_sched_l8_traffic.txt") if args.level == 8 else None,
                    "kernel": ("k_tria_rows_sched (numeric assembly, rows in locality order, FP64)" if sched
                               else "k_tria_rows (numeric assembly, FP64)"),
                    "avg_us": round(us, 2), "avg_us_mesh_order": round(us_mesh, 2),
                    "algorithmic_bytes": int(alg_bytes)}

    cpu = None
    if not args.no_cpu_baseline:
        log(f"[bench] cpu baseline: oracle on level {args.cpu_level}")
        cpu = cpu_baseline(args.cpu_level, args.k, args.cpu_thread_level)

    out = {
        "metric": "eigenpairs/sec (k=50) at lambda rel err <=1e-8 vs FP64 oracle",
        "value": round(value, 4) if value is not None else None,
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * total / args.steps, 2) if args.steps else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-spmm/f64-rr",
        "data": "synthetic",
        "config": {"workload": f"icosphere L{args.level} ellipsoid, FP64 assembly + k={args.k} eigenpairs",
                   "vertices": nv, "triangles": nf, "k": args.k, "precision": args.precision,
                   "lobpcg_iterations": st.iterations if st else None,
                   "lobpcg_iterations_fp32": st.iterations_fp32 if st else None,
                   "chebyshev_block_columns": st.poly_columns if st else None, "eig_rel_err_vs_oracle": eig_err,
                   "max_rel_resid": st.max_resid if st else None},
        "roofline": roofline,
        "roofline_assembly": roofline_asm,
        "cpu_baseline": cpu,
    }
    out.update(extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
