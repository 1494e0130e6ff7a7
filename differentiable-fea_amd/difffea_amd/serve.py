"""Served solve stream with continuous batching on one GPU.

Requests (meshes of uneven size, any element type, own material and load, optionally a compliance gradient)
arrive over time.  The server admits them as slots and HBM budget free up, keeps every solve's state on
the device, and advances all admitted solves together in *rounds*: in one round every solve performs its
next device operation (lift, residual + norm, moduli + diagonal, PCG start, a chunk of PCG iterations,
line-search trial, accept, adjoint gradient), and operations of the same kind, element type and precision
from all solves go through ONE launch of the batched kernels (descriptor table + work items).  Requests that
arrive while others run are admitted between rounds, so their work overlaps on the device.

Each solve follows the reference's state machine (`src/difffea/newton_krylov.py:126-264`): load continuation
with cut-backs, Eisenstat-Walker forcing, Armijo line search; the adjoint follows `src/difffea/adjoint.py`.
Per-solve reductions use only the solve's own blocks, and no kernel's arithmetic depends on the batch, so a
solve's result is bitwise the same alone and in any mix.

Latency = completion time - arrival time; the report holds solves/s, p50 / p99 latency and per-request
Newton / CG counts and final residuals.  A request that cannot be set up or admitted (invalid parameters, a
footprint beyond the admission budget, device memory exhausted) gets a record with ``error`` and takes no slot
or memory; the other requests run as if it had not come.
"""
import ctypes
import heapq
import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .device import HipMesh, HipSystem, _ptr, _stream
from .elements import FAMILY, NEN, ORDER, default_nq, element_type_of
from .materials import material_spec
from .plan import MeshPlan, chunking_for
from .problem import build_problem
from .adjoint import MAX_REFINE
from .newton_krylov import MIXED_RTOL_FLOOR, nan_pcg_result

HBM_SHARE = 0.8          # share of the free device memory the server admits requests into by default


# ------------------------------------------------------------------------------------------------------------
# workload
# ------------------------------------------------------------------------------------------------------------
@dataclass
class Request:
    rid: int
    arrival: float                 # seconds from the start of the stream
    element_type: str
    n: int                         # cells per side of the unit square
    mu: float
    lam: float
    stretch: float                 # prescribed x-displacement of the right edge (displacement requests)
    n_load_steps: int = 3
    adjoint: bool = False          # compliance gradient w.r.t. per-element rho (force-loaded request)
    seed: int = 0
    # a caller's own problem instead of the unit square above: the reference's params dictionary and theta
    # (element_type, n, mu, lam and stretch are then unused)
    params: dict = None
    theta: torch.Tensor = None
    tol: float = None              # absolute tolerance on ||R|| (None: the server's relative tolerance)
    max_newton: int = 25
    max_cutbacks: int = 8


def make_stream(n_requests=2000, rate=50.0, seed=5, side=(32, 512), q4_fraction=0.7, adjoint_fraction=0.2):
    """The served-stream workload: Poisson arrivals, side log-uniform, 70 % Q4 / 30 % Q9,
    mu ~ U(0.2, 0.6), lambda/mu ~ U(1, 3), stretch ~ U(1 %, 15 %), 3 load steps, 20 % with adjoint."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    out = []
    for i in range(n_requests):
        et = "quad4" if rng.random() < q4_fraction else "quad9"
        n = int(round(math.exp(rng.uniform(math.log(side[0]), math.log(side[1])))))   # cells per side, Q4 and Q9
        mu = rng.uniform(0.2, 0.6)
        lam = mu * rng.uniform(1.0, 3.0)
        out.append(Request(i, float(t[i]), et, n, mu, lam, rng.uniform(0.01, 0.15),
                           adjoint=bool(rng.random() < adjoint_fraction), seed=int(rng.integers(1 << 30))))
    return out


def fifo_latencies(arrivals, service_times):
    """Latency from arrival of each request in a first-in-first-out queue in front of a server that handles
    one request at a time: request i starts at max(arrival_i, end of request i-1) and ends service_times[i]
    later.  ``arrivals`` must be sorted."""
    end, out = 0.0, []
    for a, d in zip(arrivals, service_times):
        end = max(a, end) + d
        out.append(end - a)
    return out


def request_problem(req):
    """params + theta of a request (its own problem, or the unit square, left edge clamped, right edge
    loaded)."""
    if req.params is not None:
        theta = req.theta if req.theta is not None else torch.ones(req.params["conn"].shape[0], dtype=torch.float64)
        return req.params, theta.detach().to("cpu", torch.float64)
    if req.adjoint:
        # dead traction on the right edge: total force = stretch * E (uniaxial estimate), so strains match
        E = req.mu * (3 * req.lam + 2 * req.mu) / (req.lam + req.mu)
        spec = dict(element_type=req.element_type, nx=req.n, ny=req.n, material_params=dict(mu=req.mu, lmbda=req.lam),
                    load_mode="force", load_value=(req.stretch * E * 0.5, 0.0))
    else:
        spec = dict(element_type=req.element_type, nx=req.n, ny=req.n, material_params=dict(mu=req.mu, lmbda=req.lam),
                    load_mode="displacement", load_value=(req.stretch, None))
    p = build_problem(spec)
    n_el = p["conn"].shape[0]
    if req.adjoint:
        g = torch.Generator().manual_seed(req.seed)
        theta = 0.1 + 0.9 * torch.rand(n_el, dtype=torch.float64, generator=g)
    else:
        theta = torch.ones(n_el, dtype=torch.float64)
    return p, theta


LIFT_PROBE = 1e-6


def reference_residual_norm(sys_, params, device):
    """The norm that relative residual tolerances refer to.

    Force-driven problems: ||m f||.  Displacement-driven problems: ||K(0) u_D|| restricted to the free dofs,
    the linear out-of-balance force of the bare prescribed displacement, evaluated as ||R(d lift)|| / d at
    d = 1e-6.  (The residual of the full lift itself can be NaN: moving only the boundary nodes by the whole
    prescribed displacement inverts the boundary elements of fine meshes.)
    """
    uD = params["u_D"].to(device, torch.float64)
    if not bool(torch.any(uD != 0)):
        return float(torch.norm((params["f"] * params["m"]).to(device)))
    R = sys_.residual(sys_.lift(torch.zeros_like(uD), uD, LIFT_PROBE), lf=LIFT_PROBE)
    return float(torch.norm(R)) / LIFT_PROBE


def footprint_bytes(element_type, n, precision, preconditioner="jacobi", n_el=None, n_dof=None):
    """HBM bytes a request holds while admitted (geometry, moduli, vectors, preconditioner, plan), an upper
    estimate; n_el / n_dof: the sizes of a caller's own mesh (else those of the n x n unit square)."""
    nen, nq = NEN[element_type], default_nq(element_type)
    if n_el is None:
        n_el = (2 if FAMILY[element_type] == "tri" else 1) * n * n
        side = n + 1 if ORDER[element_type] == 1 else 2 * n + 1
        n_dof = 2 * (side * side - (n * n if element_type == "quad8" else 0))
    s = 8 if precision == "fp64" else 4
    per_qp = 2 * nen * 12 + 8 + 10 * s
    pre = n_dof * s // 2 if preconditioner == "block_jacobi" else 0     # the off-diagonal inverse entries
    return n_el * nq * per_qp + n_dof * (12 * s + 5 * 8) + pre + n_el * nen * 8 + 64 * n_el + (1 << 20)


def request_footprint(req, precision, preconditioner="jacobi"):
    """footprint_bytes of a request (plus, for a mixed-precision adjoint, its FP64 refinement tangent)."""
    if req.params is not None:
        conn = req.params["conn"]
        et = req.params.get("element_type") or element_type_of(conn.shape[1])
        n_el, n_dof = conn.shape[0], 2 * req.params["nodes"].shape[0]
        b = footprint_bytes(et, 0, precision, preconditioner, n_el=n_el, n_dof=n_dof)
    else:
        et = req.element_type
        b = footprint_bytes(et, req.n, precision, preconditioner)
        n_el = (2 if FAMILY[et] == "tri" else 1) * req.n * req.n
        side = req.n + 1 if ORDER[et] == 1 else 2 * req.n + 1
        n_dof = 2 * side * side
    if req.adjoint and precision == "mixed":
        b += n_el * default_nq(et) * 10 * 8 + n_dof * 16 * 8
    return b


# ------------------------------------------------------------------------------------------------------------
# one solve as a generator of device operations
# ------------------------------------------------------------------------------------------------------------
@dataclass
class Job:
    req: Request
    slot: int
    mesh: HipMesh
    sys: HipSystem
    params: dict
    theta: torch.Tensor
    tol: float = 0.0
    gen: object = None
    op: tuple = None
    result: dict = field(default_factory=dict)
    items: dict = field(default_factory=dict)


def solve_ops(job, tol, max_newton=25, max_cutbacks=8, cg_max_iter=500, adjoint_rtol=1e-8):
    """The reference's Newton-Krylov state machine, as a generator of device operations.

    yields ("lift", lf) -> None, ("resnorm", lf) -> ||R||, ("moduli",) -> None,
    ("pcg", rtol, max_iter, neg_res) -> iterations, ("trial", alpha) -> None, ("accept",) -> None,
    ("restore",) -> None, ("nan_pcg", max_iter) -> iterations (a NaN residual: the PCG's known result,
    ``nan_pcg_result``), ("adjoint",) -> iterations (PCG on the adjoint right-hand side or, in mixed precision,
    on the refinement residual), ("adjoint_residual",) -> ||m f - A lam|| / ||m f|| (FP64 tangent),
    ("adjoint_grad",) -> None.  Returns the reference's ``info`` entries (`newton_krylov.py:226-264`: converged,
    load_factor, history of per-iteration records, total_newton, total_cg, n_cutbacks) plus the last
    residual norm.
    """
    info = {"total_newton": 0, "total_cg": 0, "n_cutbacks": 0, "converged": False, "residual": float("nan"),
            "load_factor": 0.0, "history": [], "pcg_not_run": 0}
    mixed = job.sys.precision == "mixed"
    d_nom = 1.0 / job.req.n_load_steps
    d, lf, in_row, step = d_nom, 0.0, 0, 0
    while lf < 1.0 - 1e-12:
        lf_try = min(1.0, lf + d)
        step += 1
        # ---- newton_at_load ----
        yield ("lift", lf_try)                    # vec[u] = where(fixed, lf u_D, u_acc)
        yield ("accept",)                         # the lifted state is the Newton start (u_acc = vec[u])
        rn = yield ("resnorm", lf_try)
        r0 = rn
        ok = False
        for it in range(max_newton + 1):
            if rn < tol:
                ok = True
                break
            if it == max_newton:
                break
            eta = min(0.5, math.sqrt(rn / (r0 + 1e-14)))
            if mixed:
                eta = max(eta, MIXED_RTOL_FLOOR)
            if math.isnan(rn):
                k = yield ("nan_pcg", cg_max_iter)
                info["pcg_not_run"] += 1
            else:
                yield ("moduli",)
                k = yield ("pcg", eta, cg_max_iter, True)
            info["total_newton"] += 1
            info["total_cg"] += k
            alpha, accepted = 1.0, False
            while alpha >= 2.0 ** -5:
                yield ("trial", alpha)            # vec[u] = u_acc + alpha x
                rc = yield ("resnorm", lf_try)
                if rc < (1.0 - 1e-4 * alpha) * rn:
                    yield ("accept",)
                    rn, accepted = rc, True
                    break
                alpha *= 0.5
            info["history"].append(dict(step=step, load_factor=lf_try, iteration=it, residual=float(rn),
                                        cg_iters=int(k), alpha=alpha))
            if not accepted:
                break
        if ok:
            lf = lf_try
            info["load_factor"] = lf
            in_row = 0
            d = min(d_nom, 1.5 * d)
            # keep the converged state as the base of the next load step (u_acc already holds it)
        else:
            in_row += 1
            info["n_cutbacks"] += 1
            if in_row > max_cutbacks:
                info["residual"] = rn
                info["load_factor"] = lf
                yield ("restore",)                # the result is the last converged state
                return info
            d *= 0.5
            yield ("restore",)                    # u_acc = last converged state
    info["converged"] = True
    info["residual"] = rn
    info["load_factor"] = 1.0
    if job.req.adjoint:
        yield ("moduli",)                         # tangent at u* (vec[u] holds the accepted state)
        k = yield ("adjoint",)                    # A lam = m f (compliance: dL/du = f)
        info["adjoint_cg"] = k
        if mixed:
            # iterative refinement to the FP64 tolerance: residual m f - A lam with an FP64 tangent, FP32
            # corrections A d = r
            for _ in range(MAX_REFINE):
                rel = yield ("adjoint_residual",)
                if not rel > adjoint_rtol:
                    break
                info["adjoint_cg"] += yield ("adjoint",)
        yield ("adjoint_grad",)                   # the hand-written gradient kernel
    return info


# ------------------------------------------------------------------------------------------------------------
# the server
# ------------------------------------------------------------------------------------------------------------
class _Done:
    """A finished future holding a value (host set-up the caller did beforehand)."""

    def __init__(self, value):
        self.value = value

    def done(self):
        return True

    def result(self):
        return self.value


def _kinds(jobs):
    """Element kinds (nen plus plan flags) present among the jobs, in a fixed order."""
    return sorted({j.mesh.kind for j in jobs})


def _mat_kind(jobs, nen):
    """The kind argument of the material kernels for a group: with the library flag when any job needs it."""
    return nen | (_lib.MATLIB if any(j.sys.mat.library for j in jobs) else 0)


class Server:
    """Continuous-batching solver for a stream of requests on one GPU (see module docstring)."""

    def __init__(self, precision="fp64", max_slots=64, hbm_budget=None, pcg_chunk=8, rel_tol=1e-8,
                 device=None, adjoint_rtol=1e-8, cg_max_iter=20000, preconditioner="jacobi"):
        self.lib = _lib.load()
        self.device = torch.device(device or "cuda")
        if preconditioner not in _lib.PRECOND:
            raise ValueError(f"Unknown preconditioner '{preconditioner}'. Choose one of {tuple(_lib.PRECOND)}.")
        self.preconditioner = preconditioner
        self.precision = precision
        self.prec_id = _lib.FP64 if precision == "fp64" else _lib.FP32
        self.max_slots = max_slots
        if hbm_budget is None:
            # admission budget: a share of the device memory free when the server starts (the rest stays for
            # the caching allocator's slack and other users of the device)
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(HBM_SHARE * free)
        self.hbm_budget = hbm_budget
        self.pcg_chunk = pcg_chunk
        self.rel_tol = rel_tol
        self.adjoint_rtol = adjoint_rtol
        self.cg_max_iter = cg_max_iter
        dsz = self.lib.dfea_desc_size()
        self.dsz = dsz
        self.table = torch.zeros(max_slots * dsz, dtype=torch.uint8, device=self.device)
        self.sc = torch.zeros(max_slots, _lib.NSC, dtype=torch.float64, device=self.device)
        self.sarg = torch.zeros(max_slots, dtype=torch.float64, device=self.device)
        self.free = list(range(max_slots))[::-1]
        self.jobs = {}
        self.used_bytes = 0
        self._item_cache = {}
        self._streams = {}
        self.pcg_chunk_max = 64
        self.prep_ahead = 16         # requests whose host set-up runs ahead on worker threads
        self.prep_workers = 4

    def _side_stream(self, idx):
        """Extra stream number idx (default priority) for launches that may overlap the main stream's."""
        s = self._streams.get(idx)
        if s is None:
            s = self._streams[idx] = torch.cuda.Stream(device=self.device)
        return s

    # -- host <-> device without stalls ----------------------------------------------------------------------
    def _upload(self, a):
        """Host array -> device tensor through pinned memory, ordered on the stream, without a host wait."""
        return torch.from_numpy(np.ascontiguousarray(a)).pin_memory().to(self.device, non_blocking=True)

    def _set_sc(self, slots, field, values):
        """sc[slot, field] = value for every pair, as one device scatter (no host wait)."""
        idx = np.asarray(slots, np.int64) * _lib.NSC + _lib.SC[field]
        self.sc.view(-1).index_copy_(0, self._upload(idx), self._upload(np.asarray(values, np.float64)))

    # -- admission -----------------------------------------------------------------------------------------
    @staticmethod
    def prepare(req):
        """The host-only part of admitting a request (problem set-up and mesh plan); it needs no GPU and
        runs on a worker thread while the GPU works."""
        params, theta = request_problem(req)
        cij = params.get("cell_ij")
        conn = params["conn"].numpy()
        plan = MeshPlan(params["nodes"].numpy(), conn, params["fixed"].numpy(),
                        chunk_elems=chunking_for(conn.shape[1], conn.shape[0])[0],
                        cell_ij=None if cij is None else cij.numpy())
        return params, theta, plan

    def _admit(self, req, prepared=None):
        slot = self.free.pop()
        try:
            return self._admit_to(slot, req, prepared)
        except BaseException:
            self.free.append(slot)             # nothing of this request was left on the device
            raise

    def _admit_to(self, slot, req, prepared):
        params, theta, plan = prepared if prepared is not None else self.prepare(req)
        mesh = HipMesh(params, self.device, plan=plan)
        sys_ = HipSystem(mesh, theta, material_spec(params), params["f"], self.precision, uD=params["u_D"],
                         sc=self.sc[slot], preconditioner=self.preconditioner)
        sys_._ensure_moduli_buffers()
        self.table[slot * self.dsz:(slot + 1) * self.dsz].copy_(
            torch.frombuffer(bytearray(bytes(sys_.desc)), dtype=torch.uint8))
        job = Job(req, slot, mesh, sys_, params, theta)
        job.bytes = request_footprint(req, self.precision, self.preconditioner)
        pl = mesh.plan
        job.items = {"chunk": self._items(slot, pl.n_chunks), "fin": self._items(slot, mesh.n_finblk),
                     "vec": self._items(slot, mesh.n_vblk), "solve": self._items(slot, 1)}
        job.tol = req.tol if req.tol is not None else self.rel_tol * reference_residual_norm(sys_, params, self.device)
        if req.adjoint:              # right-hand side of the compliance adjoint, m f
            job.bvec64 = (params["m"] * params["f"]).to(self.device, torch.float64).contiguous()
            job.pcg_rhs = job.bvec64.to(sys_.vdt)
            job.lam = None
            job.sys64 = None
        self._item_cache.clear()
        job.saved = torch.zeros_like(sys_.vec["uacc"])
        job.gen = solve_ops(job, job.tol, max_newton=req.max_newton, max_cutbacks=req.max_cutbacks,
                            cg_max_iter=self.cg_max_iter, adjoint_rtol=self.adjoint_rtol)
        job.op = next(job.gen)
        job.t_admit = time.perf_counter()
        self.used_bytes += job.bytes           # counted once nothing of the admission can fail any more
        self.jobs[slot] = job
        return job

    def _retire(self, job, info):
        self._item_cache.clear()
        self.free.append(job.slot)
        self.used_bytes -= job.bytes
        del self.jobs[job.slot]
        job.sys.release_graphs()

    @staticmethod
    def _items(slot, n):
        it = np.zeros((n, 2), np.int32)
        it[:, 0] = slot
        it[:, 1] = np.arange(n)
        return it

    def _dev_items(self, jobs, kind):
        """Concatenated work items of the jobs on the device (cached while the set of jobs is unchanged)."""
        key = (kind, tuple(j.slot for j in jobs))
        hit = self._item_cache.get(key)
        if hit is not None:
            return hit
        arrs = [j.items[kind] for j in jobs if len(j.items[kind])]
        if not arrs:
            hit = (None, 0)
        else:
            a = np.concatenate(arrs)
            hit = (self._upload(a), len(a))
        self._item_cache[key] = hit
        return hit

    # -- one round -------------------------------------------------------------------------------------------
    def _round(self, host_work=None):
        """One round: enqueue every job's next operation (no host waits), run ``host_work`` while the GPU
        executes, then one sync reads every job's scalar block."""
        lib, T, st = self.lib, _ptr(self.table), _stream()
        by = {}
        for j in self.jobs.values():
            by.setdefault(j.op[0], []).append(j)
        sarg_host = None
        # lift / trial: per-solve scalar (lf or alpha) in sarg
        for kind, mode in (("lift", 0), ("trial", 1)):
            js = by.get(kind, [])
            if not js:
                continue
            if sarg_host is None:
                sarg_host = np.zeros(self.max_slots)
            for j in js:
                sarg_host[j.slot] = j.op[1]
        if sarg_host is not None:
            self.sarg.copy_(torch.from_numpy(sarg_host).pin_memory(), non_blocking=True)
        for kind, mode in (("lift", 0), ("trial", 1), ("accept", 2)):
            js = by.get(kind, [])
            for prec in (_lib.FP64, _lib.FP32):
                sub = [j for j in js if j.sys.prec_id == prec]
                it, n = self._dev_items(sub, "vec")
                if n:
                    _lib.check(lib.dfea_axpy(T, _ptr(it), n, mode, None, None, None, _ptr(self.sarg), None, prec, st),
                               "dfea_axpy")
        # restore: u_acc = saved converged state (rare, per solve)
        for j in by.get("restore", []):
            j.sys.vec["uacc"].copy_(j.saved)
        for j in by.get("nan_pcg", []):
            j.nan_iters = nan_pcg_result(j.sys.vec["x"], j.op[1])[1]
        # residual + norm (lf per solve in sc[LF])
        js = by.get("resnorm", [])
        if js:
            self._set_sc([j.slot for j in js], "lf", [float(j.op[1]) for j in js])
            for nen in _kinds(js):
                sub = [j for j in js if j.mesh.kind == nen]
                it, n = self._dev_items(sub, "chunk")
                if not n:
                    continue
                fi, nf = self._dev_items(sub, "fin")
                _lib.check(lib.dfea_residual(T, _ptr(it), n, _ptr(fi), nf, None, None, -1.0, _mat_kind(sub, nen), st),
                           "dfea_residual")
            it, n = self._dev_items(js, "vec")
            si, ns = self._dev_items(js, "solve")
            _lib.check(lib.dfea_dot(T, _ptr(it), n, _ptr(si), ns, None, None, st), "dfea_dot")
        # moduli + diagonal at vec[u]
        js = by.get("moduli", [])
        for nen in _kinds(js):
            sub = [j for j in js if j.mesh.kind == nen]
            it, n = self._dev_items(sub, "chunk")
            if n:
                fi, nf = self._dev_items(sub, "fin")
                _lib.check(lib.dfea_moduli(T, _ptr(it), n, _ptr(fi), nf, None, _mat_kind(sub, nen), self.prec_id, st),
                           "dfea_moduli")
        # PCG start (new solves) then a chunk of iterations for every running solve
        js = by.get("pcg", []) + by.get("adjoint", [])
        new = [j for j in js if not getattr(j, "pcg_running", False)]
        for j in new:
            j.pcg_rounds = 0
        if new:
            adj_rtol = max(self.adjoint_rtol, MIXED_RTOL_FLOOR) if self.precision == "mixed" else self.adjoint_rtol
            slots = [j.slot for j in new]
            self._set_sc(slots, "rtol", [float(j.op[1]) if j.op[0] == "pcg" else adj_rtol for j in new])
            self._set_sc(slots, "maxit", [float(j.op[2]) if j.op[0] == "pcg" else float(self.cg_max_iter)
                                          for j in new])
            for neg in (True, False):
                sub = [j for j in new if (j.op[0] == "pcg") == neg]
                if not sub:
                    continue
                if not neg:                      # adjoint right-hand side (m f, or a refinement residual)
                    for j in sub:
                        j.sys.vec["b"].copy_(j.pcg_rhs)
                it, n = self._dev_items(sub, "vec")
                si, ns = self._dev_items(sub, "solve")
                _lib.check(lib.dfea_pcg_init(T, _ptr(it), n, _ptr(si), ns, int(neg), self.prec_id, st), "dfea_pcg_init")
            for j in new:
                j.pcg_running = True
        if js:
            # one stream per element kind: the kinds' PCG launches are independent and overlap on the GPU.
            # Every group's work items are uploaded (on the main stream) before the fork event, so the side
            # streams' kernels see them.
            groups = []
            for nen in _kinds(js):
                sub = [j for j in js if j.mesh.kind == nen]
                # chunk length: grows while a solve keeps iterating (finished solves skip on the device)
                chunk = min(self.pcg_chunk << min(max(j.pcg_rounds for j in sub), 3), self.pcg_chunk_max)
                for j in sub:
                    j.pcg_rounds += 1
                it, n = self._dev_items(sub, "chunk")
                groups.append((nen, chunk, it, n))
            main = torch.cuda.current_stream(self.device)
            fork = None
            if len(groups) > 1:
                fork = torch.cuda.Event()
                fork.record(main)
            joins = []
            for idx, (nen, chunk, it, n) in enumerate(groups):
                stream = main if idx == 0 else self._side_stream(idx)
                if idx > 0:
                    stream.wait_event(fork)
                _lib.check(lib.dfea_pcg_iterate(T, _ptr(it), n, chunk, nen, self.prec_id,
                                                ctypes.c_void_p(stream.cuda_stream)), "dfea_pcg_iterate")
                if idx > 0:
                    done = torch.cuda.Event()
                    done.record(stream)
                    joins.append(done)
            for done in joins:
                main.wait_event(done)
        # adjoint refinement residuals (mixed precision) and gradients: per job, on the main stream
        for j in by.get("adjoint_residual", []):
            x = j.sys.vec["x"].to(torch.float64)
            j.lam = x.clone() if j.lam is None else j.lam + x
            if j.sys64 is None:                  # the FP64 tangent at u*, for the refinement residuals
                j.sys64 = HipSystem(j.mesh, j.theta, j.sys.mat, j.sys.f, "fp64", preconditioner=self.preconditioner)
                j.sys64.moduli(j.sys.vec["uacc"])
            r = j.bvec64 - j.sys64.apply(j.lam)
            j.pcg_rhs = r.to(j.sys.vdt)
            j.adj_rel = float(torch.norm(r)) / max(float(torch.norm(j.bvec64)), 1e-300)
        for j in by.get("adjoint_grad", []):
            lam = j.lam if j.lam is not None else j.sys.vec["x"].to(torch.float64)
            j.result["grad"] = j.sys.adjoint_grad(j.sys.vec["uacc"], lam)
            j.sys64 = None
        if host_work is not None:
            host_work()
        # one host sync: status, iterations, norms
        sc = self.sc.cpu().numpy()
        finished = []
        for j in list(self.jobs.values()):
            kind = j.op[0]
            if kind in ("pcg", "adjoint"):
                status = int(sc[j.slot, _lib.SC["status"]])
                if status == _lib.PCG_RUNNING:
                    continue
                j.pcg_running = False
                val = int(sc[j.slot, _lib.SC["iters"]])
            elif kind == "nan_pcg":
                val = j.nan_iters
            elif kind == "adjoint_residual":
                val = j.adj_rel
            elif kind == "resnorm":
                val = math.sqrt(sc[j.slot, _lib.SC["norm"]])
            else:
                val = None
            try:
                j.op = j.gen.send(val)
                if j.op[0] == "lift":
                    # a new load level starts from the last converged state: remember it for cut-backs
                    j.saved.copy_(j.sys.vec["uacc"])
            except StopIteration as stop:
                finished.append((j, stop.value))
        return finished

    # -- driver ------------------------------------------------------------------------------------------------
    def run(self, requests, realtime=True, keep_solution=False, progress=None, prepared=None):
        """Serve ``requests`` (sorted by arrival).  Returns the per-request records and a summary.
        ``progress``: optional callable(n_done, n_total), called as requests complete.  ``prepared``: optional
        dict rid -> the ``prepare(req)`` result, host set-up done beforehand (else worker threads do it)."""
        pending = sorted(requests, key=lambda r: r.arrival)
        records = []
        futures = {}                 # request index -> host set-up running on the worker thread
        pool = ThreadPoolExecutor(max_workers=self.prep_workers)
        t0 = time.perf_counter()
        i = 0
        peak_jobs = peak_bytes = 0

        def prepare_arrived():
            # start the host-only admission work of the next arrived requests on the worker thread
            now, k = time.perf_counter() - t0, i
            while k < len(pending) and k < i + self.prep_ahead and (not realtime or pending[k].arrival <= now):
                if k not in futures:
                    pre = prepared.get(pending[k].rid) if prepared else None
                    futures[k] = pool.submit(self.prepare, pending[k]) if pre is None else _Done(pre)
                k += 1

        try:
            while i < len(pending) or self.jobs:
                now = time.perf_counter() - t0
                prepare_arrived()
                while i < len(pending) and (not realtime or pending[i].arrival <= now) and self.free:
                    r = pending[i]
                    try:
                        fb = request_footprint(r, self.precision, self.preconditioner)
                    except Exception as exc:      # a request the set-up cannot size: its own error record
                        fb, err = 0, exc
                    else:
                        err = None if fb <= self.hbm_budget else MemoryError(
                            f"the request needs {fb} bytes of device memory, more than the admission budget "
                            f"{self.hbm_budget}")
                    if err is None and self.jobs and self.used_bytes + fb > self.hbm_budget:
                        break
                    fut = futures.get(i)
                    if err is None and self.jobs and (fut is None or not fut.done()):
                        break                      # admitted once its host set-up is done; the GPU keeps working
                    if err is None:
                        try:
                            job = self._admit(r, futures.pop(i).result() if fut is not None else None)
                        except Exception as exc:  # invalid request, device memory exhausted, ...
                            err = exc
                    if err is not None:
                        # the request fails alone: an error record, no slot or memory kept, the others go on
                        futures.pop(i, None)
                        t_now = time.perf_counter()
                        records.append(dict(rid=r.rid, element_type=r.element_type, n=r.n, adjoint=r.adjoint,
                                            converged=False, error=f"{type(err).__name__}: {err}",
                                            latency=t_now - (t0 + r.arrival if realtime else t_now), queue=0.0))
                        i += 1
                        if progress is not None:
                            progress(len(records), len(pending))
                        prepare_arrived()
                        continue
                    job.t_arrival = t0 + r.arrival if realtime else job.t_admit
                    peak_jobs = max(peak_jobs, len(self.jobs))
                    peak_bytes = max(peak_bytes, self.used_bytes)
                    i += 1
                    prepare_arrived()
                if not self.jobs:
                    if i < len(pending):
                        time.sleep(max(0.0, min(0.01, pending[i].arrival - (time.perf_counter() - t0))))
                    continue
                for job, info in self._round(prepare_arrived):
                    t_done = time.perf_counter()
                    rec = dict(rid=job.req.rid, element_type=job.req.element_type, n=job.req.n, adjoint=job.req.adjoint,
                               n_dof=job.mesh.n_dof, tol_abs=job.tol, latency=t_done - job.t_arrival,
                               queue=job.t_admit - job.t_arrival, **info)
                    if keep_solution:
                        rec["u"] = job.sys.vec["uacc"].detach().cpu().clone()
                        if "grad" in job.result:
                            rec["grad"] = job.result["grad"].detach().cpu().clone()
                    records.append(rec)
                    self._retire(job, info)
                    if progress is not None:
                        progress(len(records), len(pending))
        finally:
            pool.shutdown(wait=True)
        wall = time.perf_counter() - t0
        solved = [r for r in records if "error" not in r]
        lat = np.array([r["latency"] for r in solved]) if solved else np.zeros(1)
        summary = dict(precision=self.precision, preconditioner=self.preconditioner, n_requests=len(records),
                       wall_s=wall,
                       latency_from="arrival" if realtime else "admission",
                       solves_per_s=len(solved) / wall, p50_latency_s=float(np.percentile(lat, 50)),
                       p99_latency_s=float(np.percentile(lat, 99)),
                       converged=int(sum(r["converged"] for r in records)), errors=len(records) - len(solved),
                       hbm_budget=self.hbm_budget, peak_admitted=peak_jobs, peak_admitted_bytes=peak_bytes)
        return records, summary
