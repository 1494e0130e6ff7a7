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
    s.pcg_iterate(chunk, True)                 # warm-up (graph capture + first replay)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    n = (iters // chunk) * chunk
    e0.record()
    for _ in range(n // chunk):
        s.pcg_iterate(chunk, True)
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


def time_events(fn, steps=200, warmup=5):
    """Mean time of fn() in us (HIP events on torch's current stream, the kernels' stream)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e3 / steps


def newton_step_kernels(p, u, theta):
    """The per-Newton-step kernels on the 400x400 Q4 mesh, FP64 state: the moduli + Jacobi diagonal kernel
    (FP64 and FP32 moduli), the residual and the adjoint-gradient kernel, with their algorithmic bytes
    (per Gauss point: shape gradients 2 nen 8 B + w det J 8 B; per element: int32 connectivity 4 nen B and
    theta 8 B where read; per dof: the FP64 vectors read and written; the moduli kernel writes 10 moduli per
    Gauss point and the diagonal per dof in its precision; the gradient kernel writes 8 B per element)."""
    from difffea_amd.operators import make_system
    dev = torch.device("cuda")
    ud = u.to(dev)
    lam = torch.randn(u.numel(), dtype=torch.float64, generator=torch.Generator().manual_seed(6)).to(dev)
    out = {}
    for prec, sC in (("fp64", 8), ("mixed", 4)):
        s = make_system(ud, theta.to(dev), p, prec)
        m = s.mesh
        nqp, nd = m.n_el * m.nq, m.n_dof
        t = time_events(lambda: s.moduli(ud))
        B = nqp * (2 * m.nen * 8 + 8 + 10 * sC) + m.n_el * (4 * m.nen + 8) + nd * (8 + sC)
        out["moduli_" + prec] = dict(us_per_call=round(t, 3), bytes_per_call=B, gbs=round(B / t / 1e3, 1))
    s = make_system(ud, theta.to(dev), p, "fp64")
    m = s.mesh
    nqp, nd = m.n_el * m.nq, m.n_dof
    R = torch.empty_like(ud)
    t = time_events(lambda: s.residual(ud, out=R, lf=1.0))
    B = nqp * (2 * m.nen * 8 + 8) + m.n_el * (4 * m.nen + 8) + nd * 3 * 8
    out["residual"] = dict(us_per_call=round(t, 3), bytes_per_call=B, gbs=round(B / t / 1e3, 1))
    t = time_events(lambda: s.adjoint_grad(ud, lam))
    B = nqp * (2 * m.nen * 8 + 8) + m.n_el * (4 * m.nen + 8) + nd * 2 * 8
    out["adjoint_grad"] = dict(us_per_call=round(t, 3), bytes_per_call=B, gbs=round(B / t / 1e3, 1))
    return out


def solve_config4(precision):
    """Adjoint sensitivity at 256x256 Q4 (65,536 elements): rho_e ~ U(0.1, 1) (seed 3) scaling mu and lambda,
    left edge clamped, dead vertical load on the right edge with total 0.05 mean(rho) E H (E = 1, H = 1),
    compliance f.u; forward Newton solve to 1e-10 ||f||, adjoint PCG to 1e-10 relative (mixed: refined against
    an FP64 tangent to the same tolerance), then the gradient kernel.  Times in seconds, HIP-synchronised."""
    from difffea_amd.adjoint import compute_implicit_vjp, solve_adjoint
    from difffea_amd.newton_krylov import HipNewton
    from difffea_amd.problem import build_problem
    n = 256
    rho = 0.1 + 0.9 * torch.rand(n * n, dtype=torch.float64, generator=torch.Generator().manual_seed(3))
    F = 0.05 * float(rho.mean())
    p = build_problem(dict(element_type="quad4", nx=n, ny=n, E=1.0, nu=0.3, load_mode="force",
                           load_value=(0.0, -F)))
    dev = torch.device("cuda")
    f = p["f"].to(dev)
    tol = 1e-10 * float(torch.norm(p["f"]))
    res = {}
    for rep in range(2):                          # the first run warms up (graphs, allocator)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solver = HipNewton(p, rho.to(dev), precision)
        u, info = solver.solve(1, tol, cg_max_iter=20000)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        lam, its = solve_adjoint(solver, u, f, rtol=1e-10, max_iter=20000)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        g = compute_implicit_vjp(solver, u, lam)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        res = dict(forward_s=round(t1 - t0, 4), adjoint_solve_s=round(t2 - t1, 4), gradient_s=round(t3 - t2, 5),
                   total_s=round(t3 - t0, 4), gradients_per_s=round(1.0 / (t3 - t0), 3), converged=info["converged"],
                   newton=info["total_newton"], cg=info["total_cg"], adjoint_cg=int(its))
    res["grad_kernel_us"] = round(time_events(lambda: solver.sys.adjoint_grad(u, lam)), 3)
    return g, res


def served_comparison(n_requests, side=(32, 512), seed=5, preconditioner="jacobi",
                      legs=("plain", "fp64", "mixed", "poisson"), part=None):
    """The served request stream (config-5 distribution: sides log-uniform in ``side`` for Q4 and Q9 alike,
    70 % Q4, mu ~ U(0.2, 0.6), lambda/mu ~ U(1, 3), stretch ~ U(1 %, 15 %) in 3 load steps, 20 % with the
    compliance gradient; seed 5), every request to 1e-8 relative residual, in four runs on the same requests
    (``legs`` picks some of them; the Poisson run needs the plain one; ``part`` = (start, stop) serves that slice
    of the stream, arrivals shifted to start at 0, so a stream too long for one run is measured in slices):

    * plain: the FP64 GPU solver (``HipNewton``, the single-solve API) one request after another;
    * fp64 / mixed: the continuous-batching server, all requests queued at t = 0 (capacity);
    * poisson: the mixed server with Poisson arrivals at 0.8x the plain solver's capacity (latency from
      arrival), beside the plain solver behind a first-in-first-out queue with the same arrivals (latency from
      its measured service times).

    The host-only set-up of every request (problem, mesh plan) is done once beforehand and given to all runs,
    and its time is reported apart, so the ratios compare GPU work: ``mixed_vs_fp64_served`` is the precision
    effect, ``fp64_served_vs_plain`` the batching effect, ``mixed_served_vs_plain`` both.  Each mixed result is
    compared with the FP64 server's (max-norm relative error of u, and of the gradient where asked)."""
    import numpy as np
    from difffea_amd.adjoint import compute_implicit_vjp, solve_adjoint
    from difffea_amd.device import HipMesh
    from difffea_amd.newton_krylov import HipNewton
    from difffea_amd.serve import Server, fifo_latencies, make_stream, reference_residual_norm
    unknown = set(legs) - {"plain", "fp64", "mixed", "poisson"}
    if unknown:
        raise ValueError(f"unknown runs of the served comparison: {sorted(unknown)}")
    if "poisson" in legs and "plain" not in legs:
        raise ValueError("the Poisson run's arrival rate comes from the plain run")
    import dataclasses
    dev = torch.device("cuda")

    def stream(rate):
        rs = make_stream(n_requests, rate=rate, seed=seed, side=side)
        if part is None:
            return rs
        rs = rs[part[0]:part[1]]
        return [dataclasses.replace(r, arrival=r.arrival - rs[0].arrival) for r in rs]

    reqs = stream(1e9)
    t0 = time.perf_counter()
    prepared = {r.rid: Server.prepare(r) for r in reqs}
    setup_s = time.perf_counter() - t0
    print(f"# served stream: {len(reqs)} requests, host set-up {setup_s:.1f} s", file=sys.stderr, flush=True)

    def summary(lat, wall, n, **kw):
        return dict(n_requests=n, wall_s=round(wall, 3), solves_per_s=round(n / wall, 3),
                    p50_latency_s=round(float(np.percentile(lat, 50)), 4),
                    p99_latency_s=round(float(np.percentile(lat, 99)), 4), **kw)

    def progress(label):
        # every 16th result, and every one of the last 16 (the largest requests finish last)
        return lambda d, t: print(f"# {label}: {d} of {t}", file=sys.stderr, flush=True) \
            if d % 16 == 0 or d > t - 16 else None

    out = {
        "workload": f"{len(reqs)} requests" + ("" if part is None else f" ({part[0]}..{part[1] - 1} of {n_requests})") +
                    f", sides log-uniform in [{side[0]}, {side[1]}] (Q4 and Q9), 70 % Q4, "
                    f"20 % with the compliance gradient, seed {seed}, 1e-8 relative residual",
        "n_q9": int(sum(r.element_type == "quad9" for r in reqs)), "max_dofs": int(max(
            prepared[r.rid][0]["nodes"].shape[0] for r in reqs) * 2),
        "host_setup_s": round(setup_s, 2), "host_setup_s_per_request": round(setup_s / len(reqs), 4),
    }
    plain = service = plain_u = None
    if "plain" in legs:
        # plain FP64 solver, one request after another (mesh upload + geometry and the solve timed)
        lat, plain_u, plain_cg, plain_not_run = [], {}, 0, 0
        report = progress("plain FP64 solver")
        t0 = time.perf_counter()
        for r in reqs:
            t1 = time.perf_counter()
            p, th, plan = prepared[r.rid]
            solver = HipNewton(p, th.to(dev), mesh=HipMesh(p, dev, plan=plan), preconditioner=preconditioner)
            u, info = solver.solve(r.n_load_steps, 1e-8 * reference_residual_norm(solver.sys, p, dev),
                                   cg_max_iter=20000)
            plain_cg += info["total_cg"]
            plain_not_run += info["pcg_not_run"]
            if r.adjoint:
                lam, _ = solve_adjoint(solver, u, p["f"].to(dev), rtol=1e-8, max_iter=20000)
                compute_implicit_vjp(solver, u, lam)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t1)
            plain_u[r.rid] = u.cpu()
            del solver
            report(len(lat), len(reqs))
        plain = summary(lat, time.perf_counter() - t0, len(reqs), latency_from="start of the request's solve",
                        mode="HipNewton, one request after another", n_cg=plain_cg, n_pcg_not_run=plain_not_run)
        service = list(lat)
        out["plain_fp64"] = plain

    def serve(precision, realtime, rate=None):
        rs = stream(rate or 1e9)
        srv = Server(precision, max_slots=64, preconditioner=preconditioner)
        label = f"served {precision}" + (" (Poisson arrivals)" if realtime else "")
        recs, summ = srv.run(rs, realtime=realtime, keep_solution=True, prepared=prepared, progress=progress(label))
        summ["n_newton"] = int(sum(x.get("total_newton", 0) for x in recs))
        summ["n_cg"] = int(sum(x.get("total_cg", 0) for x in recs))
        summ["n_pcg_not_run"] = int(sum(x.get("pcg_not_run", 0) for x in recs))
        return {x["rid"]: x for x in recs}, summ

    keys = ("wall_s", "solves_per_s", "p50_latency_s", "p99_latency_s", "latency_from", "converged", "errors",
            "n_newton", "n_cg", "n_pcg_not_run", "peak_admitted")
    r64 = r32 = sat64 = sat32 = None
    if "fp64" in legs:
        r64, sat64 = serve("fp64", False)
        out["served_fp64"] = {k: sat64[k] for k in keys if k in sat64}
    if "mixed" in legs:
        r32, sat32 = serve("mixed", False)
        out["served_mixed"] = {k: sat32[k] for k in keys if k in sat32}
    if sat64 and sat32:
        out["mixed_vs_fp64_served"] = round(sat32["solves_per_s"] / sat64["solves_per_s"], 3)
    if plain and sat64:
        out["fp64_served_vs_plain"] = round(sat64["solves_per_s"] / plain["solves_per_s"], 3)
    if plain and sat32:
        out["mixed_served_vs_plain"] = round(sat32["solves_per_s"] / plain["solves_per_s"], 3)
        out["mixed_vs_plain_incl_host_setup"] = round(sat32["solves_per_s"] * (plain["wall_s"] + setup_s)
                                                      / len(reqs), 3)
    ref_u = r64 if r64 else ({i: {"u": u} for i, u in plain_u.items()} if plain_u else None)
    if r32 and ref_u:
        out["mixed_vs_fp64_max_rel_error_u"] = max(float((r32[i]["u"] - ref_u[i]["u"]).abs().max()
                                                         / ref_u[i]["u"].abs().max()) for i in ref_u)
    if r32 and r64:
        err_g = [float((r32[i]["grad"] - r64[i]["grad"]).abs().max() / r64[i]["grad"].abs().max())
                 for i in r64 if "grad" in r64[i]]
        out["mixed_vs_fp64_max_rel_error_grad"] = max(err_g) if err_g else None
    if plain_u and r64:
        out["plain_vs_served_fp64_max_rel_error_u"] = max(float((plain_u[i] - r64[i]["u"]).abs().max()
                                                                / r64[i]["u"].abs().max()) for i in r64)
    if "poisson" in legs:
        rate = 0.8 * plain["solves_per_s"]
        _, poi32 = serve("mixed", True, rate)
        arrivals = [r.arrival for r in stream(rate)]
        q = fifo_latencies(arrivals, service)
        out["poisson_rate_per_s"] = round(rate, 3)
        out["served_mixed_poisson"] = {k: poi32[k] for k in keys if k in poi32}
        out["plain_fp64_fifo_poisson"] = dict(p50_latency_s=round(float(np.percentile(q, 50)), 4),
                                              p99_latency_s=round(float(np.percentile(q, 99)), 4),
                                              latency_from="arrival (service times of the plain run)")
    return out


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
    ap.add_argument("--serve-side", type=int, nargs=2, default=(32, 512),
                    help="side range of the served meshes (the configured workload: 32 512)")
    ap.add_argument("--served-only", action="store_true", help="only the served-stream comparison, one JSON line")
    ap.add_argument("--served-legs", default="plain,fp64,mixed,poisson",
                    help="runs of the served comparison (comma-separated subset of plain, fp64, mixed, poisson)")
    ap.add_argument("--serve-slice", default=None,
                    help="start:stop -- serve only that slice of the --serve-requests stream (long streams in parts)")
    ap.add_argument("--config4-only", action="store_true",
                    help="only the 256x256 Q4 adjoint sensitivity (FP64 and mixed), one JSON line; for profiling it")
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
    if args.config4_only:
        print(json.dumps({prec: solve_config4(prec)[1] for prec in ("fp64", "mixed")}))
        return
    if args.served_only:
        part = None if args.serve_slice is None else tuple(int(x) for x in args.serve_slice.split(":"))
        print(json.dumps(served_comparison(args.serve_requests, tuple(args.serve_side),
                                           legs=tuple(args.served_legs.split(",")), part=part)))
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
    steps = newton_step_kernels(p, u, th)
    print(f"# Q4 400x400 Newton-step kernels: {steps}", file=sys.stderr, flush=True)
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
        g64, solves["c4_fp64"] = solve_config4("fp64")
        g32, solves["c4_mixed"] = solve_config4("mixed")
        solves["c4_mixed_vs_fp64_grad_rel"] = float((g32 - g64).abs().max() / g64.abs().max())
        print(f"# config-4 adjoint: {solves['c4_fp64']} {solves['c4_mixed']} "
              f"{solves['c4_mixed_vs_fp64_grad_rel']:.2e}", file=sys.stderr, flush=True)
        if args.serve_requests > 0:
            srv = served_comparison(args.serve_requests, tuple(args.serve_side))
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
        "newton_step_kernels_q4_400": steps,
        "element_products_fp64": elems,
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample",
                                                                      "cpu_model", "os_cpu_count")},
        "solves": solves,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
