"""Benchmark of the matrix-free Newton-Krylov hot path on one MI355X.

Headline (the JSON ``value``): algorithmic HBM bandwidth of the FP64 tangent product A(v) on the 400x400
unit-square Q4 mesh (160,000 elements, 321,602 dofs), state = 5 % uniaxial stretch + N(0,(1e-3 h)^2)
noise, probe v ~ N(0,1), left edge clamped.  One step = one A(v) over the whole mesh, inputs resident in
HBM.  Algorithmic bytes per A(v):

    B_alg = N_qp (10 s_C + 2 nen s_G) + 4 n_el nen + 2 n_dof s_v

(10 unique moduli and nen x 2 shape gradients per Gauss point, int32 connectivity, v read once, y written
once).  Extra fields report the FP32-moduli product, the 200x200 Q9 product, and the CPU baseline: the
oracle's op-for-op restatement of the reference's ``qp`` product on the host cores.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 20]
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "differentiable-fea_amd")]

import torch  # noqa: E402

METRIC = "A(v) HBM GB/s at 160k Q4; solves/s + p50/p99 latency, mixed vs FP64"
HBM_PEAK_GBS = 8000.0
PMC_SUMMARY = os.path.join(ROOT, "profiles", "apply_q4_400_fp64_pmc.md")


def measured_traffic():
    """HBM bytes per FP64 A(v) at 400x400 Q4 from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (scaled by the factors stated there); None when the summary is absent."""
    try:
        with open(PMC_SUMMARY) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_step:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def b_alg(n_el, nen, nq, n_dof, s):
    return nq * n_el * (10 * s + 2 * nen * s) + 4 * n_el * nen + 2 * n_dof * s


def problem(et, n):
    from difffea_amd.problem import build_problem
    p = build_problem(dict(element_type=et, nx=n, ny=n, E=1.0, nu=0.3, load_mode="displacement",
                           load_value=(0.1, None), fixed_side="left", load_side="right"))
    g = torch.Generator().manual_seed(0)
    X = p["nodes"]
    u = torch.zeros_like(X)
    u[:, 0] = 0.05 * X[:, 0]
    u = u.reshape(-1) + 1e-3 * (1.0 / n) * torch.randn(u.numel(), dtype=torch.float64, generator=g)
    v = torch.randn(u.numel(), dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    theta = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    return p, u, v, theta


def time_apply(p, u, v, theta, precision, steps, warmup):
    from difffea_amd.operators import make_tangent_operator
    dev = torch.device("cuda")
    A = make_tangent_operator("hip", u.to(dev), theta.to(dev), p, precision=precision)
    vd = v.to(dev, torch.float64 if precision == "fp64" else torch.float32)
    y = torch.empty_like(vd)
    sysm = A.system
    for _ in range(warmup):
        sysm.apply(vd, y)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        sysm.apply(vd, y)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    m = sysm.mesh
    s = 8 if precision == "fp64" else 4
    B = b_alg(m.n_el, m.nen, m.nq, m.n_dof, s)
    return dict(ms=ms, gbs=B / (ms * 1e-3) / 1e9, bytes=B, products_per_s=1e3 / ms, y=y)


def time_apply_cold(p, u, v, theta, precision, steps=100, flush_mb=512):
    """A(v) with the caches cold: before every product a 512 MB buffer (twice the 256 MB Infinity Cache) is
    rewritten, then the product alone is timed with HIP events on the kernels' stream; median over steps."""
    from difffea_amd.operators import make_tangent_operator
    dev = torch.device("cuda")
    A = make_tangent_operator("hip", u.to(dev), theta.to(dev), p, precision=precision)
    vd = v.to(dev, torch.float64 if precision == "fp64" else torch.float32)
    y = torch.empty_like(vd)
    sysm = A.system
    junk = torch.empty(flush_mb * (1 << 20) // 4, dtype=torch.float32, device=dev)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for i in range(steps + 3):
        junk.fill_(float(i))
        if i >= 3:
            ev[i - 3][0].record()
        sysm.apply(vd, y)
        if i >= 3:
            ev[i - 3][1].record()
    torch.cuda.synchronize()
    ms = statistics.median(a.elapsed_time(b) for a, b in ev)
    m = sysm.mesh
    B = b_alg(m.n_el, m.nen, m.nq, m.n_dof, 8 if precision == "fp64" else 4)
    return dict(us_per_product=round(ms * 1e3, 3), gbs=round(B / (ms * 1e-3) / 1e9, 2), flush_mb=flush_mb,
                steps=steps)


def time_pcg_iteration(p, u, v, theta, precision, iters=400, chunk=50):
    """One fused PCG iteration (one kernel: scalar step, update, next product) on the same mesh and state,
    replayed from HIP graphs of `chunk` kernels; rtol = 0 keeps the solve running.  Algorithmic bytes per
    iteration: the A(v) bytes plus 11 dof vectors (x, r, p, s = A p, w = A M r read and written, diag read;
    v read / y written are already in the A(v) bytes, so 9 more)."""
    from difffea_amd import _lib
    from difffea_amd.operators import make_system
    dev = torch.device("cuda")
    s = make_system(u.to(dev), theta.to(dev), p, precision)
    s.moduli(u.to(dev))
    b = (v * p["m"]).to(dev)
    s.pcg_start(b, rtol=0.0, max_iter=iters + 2 * chunk + 1)
    s.pcg_iterate(chunk)                       # warm-up (graph capture + first replay)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    n = (iters // chunk) * chunk
    e0.record()
    for _ in range(n // chunk):
        s.pcg_iterate(chunk)
    e1.record()
    torch.cuda.synchronize()
    st, _, k = s.pcg_status()
    if st != _lib.PCG_RUNNING or k != n + chunk - 1:        # the first kernel only forms A M b
        raise RuntimeError(f"PCG timing run stopped early (status {st}, k {k})")
    us = e0.elapsed_time(e1) * 1e3 / n
    m = s.mesh
    sz = 8 if precision == "fp64" else 4
    B = b_alg(m.n_el, m.nen, m.nq, m.n_dof, sz) + 9 * m.n_dof * sz
    return dict(us_per_iteration=round(us, 3), gbs=round(B / (us * 1e-6) / 1e9, 2), bytes_per_iteration=B,
                iterations=n)


def time_postprocess(p, u, theta, steps=200):
    """dfea_fields (element fields + nodal averages) and dfea_reactions on the same mesh and state.
    Algorithmic bytes: per Gauss point the shape gradients and w det J (2 nen 8 + 8); per element theta,
    int32 connectivity credit and, for the fields, 7 outputs; per node u read and, for the fields, 6 outputs."""
    from difffea_amd.operators import make_system
    dev = torch.device("cuda")
    ud = u.to(dev)
    s = make_system(ud, theta.to(dev), p)
    m = s.mesh
    bufs = s.fields_buffers()
    r2 = torch.empty(2, dtype=torch.float64, device=dev)
    out = {}
    for name, fn, extra_el, extra_node in (("fields", lambda: s.fields(ud, bufs), 7 * 8, 6 * 8),
                                            ("reactions", lambda: s.reactions(ud, r2), 0, 0)):
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / steps
        B = m.nq * m.n_el * (2 * m.nen * 8 + 8) + m.n_el * (8 + 4 * m.nen + extra_el) + m.n_nodes * (16 + extra_node)
        out[name] = dict(us_per_call=round(us, 3), gbs=round(B / (us * 1e-6) / 1e9, 2), bytes_per_call=B)
    return out


def element_products(steps=200):
    """FP64 A(v) of the other element types and of unstructured meshes at about the headline's size:
    structured tri3 283x283 (160,178 triangles), tri6 141x141 (39,762), quad8 200x200 (40,000), and Delaunay
    plates with a hole r = 0.2 (irregular node valence): quad4 h = 0.0025 (209,811 quadrilaterals) and tri3
    h = 0.0025 (279,801 triangles).  State: 5 % stretch + N(0,(1e-3 h)^2) noise; B_alg as the headline."""
    from difffea_amd.problem import build_problem
    out = {}
    cases = (("tri3_283", dict(element_type="tri3", nx=283, ny=283), 1 / 283),
             ("tri6_141", dict(element_type="tri6", nx=141, ny=141), 1 / 141),
             ("quad8_200", dict(element_type="quad8", nx=200, ny=200), 1 / 200),
             ("quad4_unstructured_h0.0025", dict(element_type="quad4", mesher="unstructured", geometry="plate_hole",
                                                 h=0.0025), 0.0025),
             ("tri3_unstructured_h0.0025", dict(element_type="tri3", mesher="unstructured", geometry="plate_hole",
                                                h=0.0025), 0.0025))
    for name, spec, h in cases:
        p = build_problem(dict(spec, E=1.0, nu=0.3, load_mode="displacement", load_value=(0.1, None),
                               fixed_side="left", load_side="right"))
        g = torch.Generator().manual_seed(0)
        X = p["nodes"]
        u = torch.zeros_like(X)
        u[:, 0] = 0.05 * X[:, 0]
        u = u.reshape(-1) + 1e-3 * h * torch.randn(u.numel(), dtype=torch.float64, generator=g)
        v = torch.randn(u.numel(), dtype=torch.float64, generator=torch.Generator().manual_seed(1))
        th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
        r = time_apply(p, u, v, th, "fp64", steps, 10)
        out[name] = dict(n_elements=int(p["conn"].shape[0]), n_dofs=int(u.numel()), us_per_product=round(r["ms"] * 1e3, 3),
                         gbs=round(r["gbs"], 2), bytes_per_step=r["bytes"])
        print(f"# {name} fp64 A(v): {out[name]}", file=sys.stderr, flush=True)
    return out


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform's processor string."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(p, u, v, theta, seconds=10.0):
    """The oracle's 'qp' restatement (gather, einsum, batched 4x4 matmul, einsum, index_add_), FP64."""
    from oracle import fem
    nthr = os.cpu_count() or 1
    torch.set_num_threads(nthr)
    p = fem.precompute_geometry(p)
    C, _ = fem.qp_setup(u, theta, p)
    A = fem.make_A_operator_qp(u, theta, p, C=C)
    for _ in range(2):
        A(v)
    times = []
    t_end = time.perf_counter() + seconds
    while len(times) < 7 or time.perf_counter() < t_end:
        t0 = time.perf_counter()
        A(v)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    n_el, nen = p["conn"].shape
    B = b_alg(n_el, nen, p["dN_dX"].shape[1], u.numel(), 8)
    return dict(value=B / med / 1e9, unit="GB/s", cores=torch.get_num_threads(), kind="port",
                cpu_model=cpu_model(), os_cpu_count=os.cpu_count(),
                sample=f"400x400 Q4 FP64 A(v), oracle qp restatement, median of {len(times)} products "
                       f"({med * 1e3:.2f} ms each) after 2 warm-ups",
                ms_per_product=med * 1e3)


def solve_config1(precision, cg_rtol=None, cg_max_iter=500, preconditioner="jacobi"):
    """Unit-square 100x100 Q4, E=1 nu=0.3, left edge clamped, right edge u_x = 0.1 in 3 load steps; Newton
    tolerance 1e-10 relative to ||K(0) u_D|| (the linear out-of-balance force of the prescribed motion)."""
    from difffea_amd.problem import build_problem
    from difffea_amd.newton_krylov import newton_krylov_solve, HipNewton
    p = build_problem(dict(element_type="quad4", nx=100, ny=100, E=1.0, nu=0.3, load_mode="displacement",
                           load_value=(0.1, None)))
    from difffea_amd.serve import reference_residual_norm
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64, device="cuda")
    probe = HipNewton(p, th)
    tol = 1e-10 * reference_residual_norm(probe.sys, p, probe.device)
    kw = dict(verbose=False, precision=precision, cg_rtol=cg_rtol, cg_max_iter=cg_max_iter,
              preconditioner=preconditioner)
    newton_krylov_solve(p, th, 3, tol, **kw)
    u, info = newton_krylov_solve(p, th, 3, tol, **kw)
    return u, dict(seconds=round(info["seconds"], 4), solves_per_s=round(1.0 / info["seconds"], 3),
                   newton=info["total_newton"], cg=info["total_cg"], converged=info["converged"], tol_abs=tol)


def solve_config3(precision, preconditioner="jacobi"):
    """O-grid plate with a hole r=0.2, 64 x 256 Q9, E_e lognormal (mean 1, sigma 0.3, seed 2), nu=0.3,
    left edge clamped, dead traction on the right edge (total 0.1) in 5 load steps; Newton tolerance
    1e-10 relative to ||m f||."""
    import math
    from difffea_amd.problem import build_problem
    from difffea_amd.newton_krylov import newton_krylov_solve
    p = build_problem(dict(element_type="quad9", geometry="plate_hole_ogrid",
                           geom_params=dict(width=1.0, radius=0.2, n_rad=64, n_circ=256), E=1.0, nu=0.3,
                           load_mode="force", load_value=(0.1, 0.0)))
    g = torch.Generator().manual_seed(2)
    n_el = p["conn"].shape[0]
    th = torch.exp(torch.randn(n_el, dtype=torch.float64, generator=g) * 0.3 - 0.045).cuda()
    tol = 1e-10 * float(torch.norm(p["f"] * p["m"]))
    u, info = newton_krylov_solve(p, th, 5, tol, verbose=False, precision=precision, cg_max_iter=20000,
                                  preconditioner=preconditioner)
    return u, dict(seconds=round(info["seconds"], 4), newton=info["total_newton"], cg=info["total_cg"],
                   converged=info["converged"], tol_abs=tol)


def sequential_fp64(n_requests, seed=5, side=(32, 512), rate=None):
    """The plain FP64 GPU solver: the same requests solved one after another with newton_krylov_solve.
    With ``rate``, also the latency from arrival of a first-in-first-out queue in front of it (Poisson arrivals
    of that rate, the stream's seed): request i starts at max(arrival_i, end of request i-1) and takes the
    service time measured for it here (the solver serves one request at a time, so the queue is exact)."""
    import numpy as np
    from difffea_amd.serve import make_stream, reference_residual_norm, request_problem
    from difffea_amd.newton_krylov import HipNewton
    reqs = make_stream(n_requests, rate=1e9, seed=seed, side=side)
    lat = []
    t0 = time.perf_counter()
    for r in reqs:
        t1 = time.perf_counter()
        p, th = request_problem(r)
        solver = HipNewton(p, th.cuda())
        ref = reference_residual_norm(solver.sys, p, solver.device)
        u, info = solver.solve(r.n_load_steps, 1e-8 * ref, cg_max_iter=20000)
        if r.adjoint:
            from difffea_amd.adjoint import solve_adjoint, compute_implicit_vjp
            lam, _ = solve_adjoint(solver, u, p["f"].cuda(), rtol=1e-8, max_iter=20000)
            compute_implicit_vjp(solver, u, lam)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t1)
        if len(lat) % 32 == 0:
            print(f"# plain FP64 solver: {len(lat)} of {len(reqs)} requests, {time.perf_counter() - t0:.1f} s",
                  file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    out = dict(precision="fp64", mode="sequential newton_krylov_solve", n_requests=len(reqs), wall_s=wall,
               solves_per_s=len(reqs) / wall, p50_latency_s=float(np.percentile(lat, 50)),
               p99_latency_s=float(np.percentile(lat, 99)), latency_from="start of the request's solve")
    if rate is not None:
        from difffea_amd.serve import fifo_latencies
        qlat = fifo_latencies([r.arrival for r in make_stream(n_requests, rate=rate, seed=seed, side=side)], lat)
        out.update(arrival_rate_per_s=rate, p50_latency_from_arrival_s=float(np.percentile(qlat, 50)),
                   p99_latency_from_arrival_s=float(np.percentile(qlat, 99)))
    return out


def serve_stream(n_requests, precision, rate=None, seed=5, side=(32, 512), preconditioner="jacobi"):
    from difffea_amd.serve import Server, make_stream
    reqs = make_stream(n_requests, rate=rate or 1e9, seed=seed, side=side)
    srv = Server(precision, max_slots=64, preconditioner=preconditioner)

    def progress(done, total):
        if done % 32 == 0:
            print(f"# served {precision} {preconditioner}: {done} of {total} requests", file=sys.stderr,
                  flush=True)

    recs, summ = srv.run(reqs, realtime=rate is not None, progress=progress)
    summ["n_newton"] = int(sum(r["total_newton"] for r in recs))
    summ["n_cg"] = int(sum(r["total_cg"] for r in recs))
    return summ


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true", help="skip the complete-solve workloads")
    ap.add_argument("--serve-requests", type=int, default=64,
                    help="requests of the served stream (the full workload is 2000)")
    ap.add_argument("--serve-side", type=int, nargs=2, default=(32, 256),
                    help="side range of the served meshes (the full workload is 32 512)")
    ap.add_argument("--config1-only", action="store_true",
                    help="only the 100x100 Q4 solves (FP64 and mixed), one JSON line; for profiling the solve")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(2)

    if args.config1_only:
        print(json.dumps({prec: solve_config1(prec)[1] for prec in ("fp64", "mixed")}))
        return

    p, u, v, th = problem("quad4", 400)
    r64 = time_apply(p, u, v, th, "fp64", args.steps, args.warmup)
    print(f"# Q4 400x400 fp64 A(v): {r64['ms'] * 1e3:.2f} us, {r64['gbs']:.0f} GB/s", file=sys.stderr, flush=True)
    r32 = time_apply(p, u, v, th, "mixed", args.steps, args.warmup)
    print(f"# Q4 400x400 fp32 A(v): {r32['ms'] * 1e3:.2f} us, {r32['gbs']:.0f} GB/s", file=sys.stderr, flush=True)
    cold64 = time_apply_cold(p, u, v, th, "fp64")
    cold32 = time_apply_cold(p, u, v, th, "mixed")
    print(f"# Q4 400x400 A(v), caches flushed before each product: fp64 {cold64}, fp32 {cold32}", file=sys.stderr,
          flush=True)
    pcg64 = time_pcg_iteration(p, u, v, th, "fp64")
    pcg32 = time_pcg_iteration(p, u, v, th, "mixed")
    print(f"# Q4 400x400 PCG iteration: fp64 {pcg64['us_per_iteration']:.2f} us, "
          f"fp32 {pcg32['us_per_iteration']:.2f} us", file=sys.stderr, flush=True)
    post = time_postprocess(p, u, th)
    print(f"# Q4 400x400 post-processing: {post}", file=sys.stderr, flush=True)
    p9, u9, v9, th9 = problem("quad9", 200)
    q9 = time_apply(p9, u9, v9, th9, "fp64", args.steps, args.warmup)
    print(f"# Q9 200x200 fp64 A(v): {q9['ms'] * 1e3:.2f} us, {q9['gbs']:.0f} GB/s", file=sys.stderr, flush=True)

    elems = element_products()

    solves = None
    if not args.no_solve:
        solves = {}
        u64, solves["c1_fp64_ew"] = solve_config1("fp64")
        u32, solves["c1_mixed_ew"] = solve_config1("mixed")
        solves["c1_mixed_vs_fp64_rel"] = float((u32 - u64).abs().max() / u64.abs().max())
        _, solves["c1_fp64_cg1e-12"] = solve_config1("fp64", cg_rtol=1e-12, cg_max_iter=5000)
        ub, solves["c1_fp64_block_jacobi"] = solve_config1("fp64", preconditioner="block_jacobi")
        _, solves["c1_mixed_block_jacobi"] = solve_config1("mixed", preconditioner="block_jacobi")
        solves["c1_block_vs_jacobi_rel"] = float((ub - u64).abs().max() / u64.abs().max())
        print(f"# config-1 solves: {solves}", file=sys.stderr, flush=True)
        v64, solves["c3_fp64"] = solve_config3("fp64")
        v32, solves["c3_mixed"] = solve_config3("mixed")
        solves["c3_mixed_vs_fp64_rel"] = float((v32 - v64).abs().max() / v64.abs().max())
        vb, solves["c3_fp64_block_jacobi"] = solve_config3("fp64", "block_jacobi")
        _, solves["c3_mixed_block_jacobi"] = solve_config3("mixed", "block_jacobi")
        solves["c3_block_vs_jacobi_rel"] = float((vb - v64).abs().max() / v64.abs().max())
        print(f"# config-3 solves: {solves['c3_fp64']} {solves['c3_mixed']} {solves['c3_fp64_block_jacobi']} "
              f"{solves['c3_mixed_block_jacobi']}", file=sys.stderr, flush=True)
        if args.serve_requests > 0:
            sat64 = serve_stream(args.serve_requests, "fp64", side=tuple(args.serve_side))
            print(f"# served fp64 (saturated): {sat64}", file=sys.stderr, flush=True)
            rate = 0.8 * sat64["solves_per_s"]
            seq = sequential_fp64(args.serve_requests, side=tuple(args.serve_side))
            print(f"# plain FP64 GPU solver (sequential): {seq}", file=sys.stderr, flush=True)
            sat32 = serve_stream(args.serve_requests, "mixed", side=tuple(args.serve_side))
            print(f"# served mixed (saturated): {sat32}", file=sys.stderr, flush=True)
            sat32b = serve_stream(args.serve_requests, "mixed", side=tuple(args.serve_side),
                                  preconditioner="block_jacobi")
            print(f"# served mixed block-Jacobi (saturated): {sat32b}", file=sys.stderr, flush=True)
            srv = {"saturated_fp64": sat64, "saturated_mixed": sat32, "saturated_mixed_block_jacobi": sat32b,
                   "plain_fp64_sequential": seq,
                   "saturated_mixed_vs_plain_fp64": round(sat32["solves_per_s"] / seq["solves_per_s"], 3),
                   "saturated_mixed_block_jacobi_vs_plain_fp64": round(sat32b["solves_per_s"] / seq["solves_per_s"],
                                                                       3),
                   "fp64": serve_stream(args.serve_requests, "fp64", rate, side=tuple(args.serve_side)),
                   "mixed": serve_stream(args.serve_requests, "mixed", rate, side=tuple(args.serve_side)),
                   "arrival_rate_per_s": rate}
            print(f"# served: {srv}", file=sys.stderr, flush=True)
            solves["served"] = srv

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(p, u, v, th, args.cpu_seconds)
        print(f"# CPU baseline: {cpu['ms_per_product']:.2f} ms on {cpu['cores']} threads", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC, "value": round(r64["gbs"], 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r64["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "tangent product A(v), 400x400 unit-square Q4 (160,000 el, 321,602 dofs, 2x2 Gauss), "
                               "neo-Hookean mu=0.3846 lambda=0.5769, 5% stretch + noise, left edge clamped, "
                               "FP64 moduli", "n_elements": 160000, "n_dofs": 321602,
                   "bytes_per_step": r64["bytes"]},
        "products_per_s": round(r64["products_per_s"], 1),
        "roofline": {"bound": "hbm", "achieved": round(r64["gbs"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r64["gbs"] / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                     "step": "one A(v) = k_apply + k_apply_fin (shared nodes); traffic from "
                             "profiles/apply_q4_400_fp64_pmc.md"},
        "fp32_moduli": {"us_per_product": round(r32["ms"] * 1e3, 3), "gbs": round(r32["gbs"], 2),
                        "bytes_per_step": r32["bytes"]},
        "cold_cache_fp64": cold64, "cold_cache_fp32": cold32,
        "q9_200x200_fp64": {"us_per_product": round(q9["ms"] * 1e3, 3), "gbs": round(q9["gbs"], 2),
                            "bytes_per_step": q9["bytes"]},
        "pcg_iteration_fp64": pcg64, "pcg_iteration_mixed": pcg32, "postprocess_q4_400": post,
        "element_products_fp64": elems,
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
        "solves": solves,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
