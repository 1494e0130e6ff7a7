"""Device-resident state of the hot path and thin wrappers around the C ABI.

``HipMesh``    plan tables + geometry cache (dN_dX planes in FP64 and FP32, w_detJ) of one mesh + mask.
``HipSystem``  one solve on a mesh: theta, precision, moduli, Jacobi diagonal, PCG / Newton vectors and
               its scratch (shared-node partials, reduction partials, scalar block); owns a descriptor.
Several systems are advanced by one launch through a descriptor table of their descriptors and concatenated
work-item tables (``serve.Server`` builds those).

All memory is allocated through torch (the caching allocator) and lives on the current CUDA(HIP) device;
kernels are enqueued on torch's current stream.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .materials import MaterialSpec
from .elements import NEN, default_nq, element_type_of, reference_gradient_table
from .plan import MeshPlan, chunking_for

F64, F32 = torch.float64, torch.float32

# the scalar-block entries the host reads after a round of PCG kernels (status, k, maxit, iters) are contiguous,
# so they are one copy
_SC_STATUS = _lib.SC["status"]
assert (_lib.SC["k"], _lib.SC["maxit"], _lib.SC["iters"]) == (_SC_STATUS + 1, _SC_STATUS + 2, _SC_STATUS + 3)


# torch's current stream of the current device as a raw handle: the launch path asks for it on every call, and
# the stream object of torch.cuda.current_stream() costs about 3 us of host time against 0.1 us for the handle
_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_CUR_DEVICE = getattr(torch._C, "_cuda_getDevice", None)


def _stream():
    if _RAW_STREAM is not None and _CUR_DEVICE is not None:
        return ctypes.c_void_p(_RAW_STREAM(_CUR_DEVICE()))
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _dev_int32(a, device):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).to(device)


class HipMesh:
    """Mesh plan + geometry on the device.  ``params``: the reference's params contract."""

    def __init__(self, params, device=None, plan_mode="auto", plan=None, chunking="auto"):
        """plan_mode: "auto" / "partials" / "halo" (quad4); chunking: "auto" / "small" / "large" -- the chunk
        size of partials plans (plan.chunking_for); plan: a MeshPlan of these tables built beforehand."""
        self.lib = _lib.load()
        device = torch.device(device or "cuda")
        if device.type != "cuda":
            raise ValueError("HipMesh lives on a GPU device")
        self.device = device
        nodes = params["nodes"].detach().to("cpu", F64)
        conn = params["conn"].detach().to("cpu", torch.int64)
        fixed = params["fixed"].detach().to("cpu", torch.bool)
        self.element_type = params.get("element_type") or element_type_of(conn.shape[1])
        self.nen = conn.shape[1]
        GP = params["GP"].detach().to("cpu", F64).numpy()
        GW = params["GW"].detach().to("cpu", F64).numpy()
        self.nq = GP.shape[0]
        if self.nen != NEN[self.element_type] or self.nq != default_nq(self.element_type):
            raise ValueError(f"the hip kernels take {self.element_type} elements with {NEN[self.element_type]} "
                             f"nodes and a {default_nq(self.element_type)}-point rule, got nen={self.nen}, "
                             f"n_q={self.nq}")
        cij = params.get("cell_ij")
        # plan: a MeshPlan of exactly these tables, built beforehand (e.g. while the GPU was busy)
        ce, _ = chunking_for(self.nen, conn.shape[0], chunking) if plan_mode != "halo" else (None, 0)
        self.plan = p = plan if plan is not None else MeshPlan(
            nodes.numpy(), conn.numpy(), fixed.numpy(), mode=plan_mode, chunk_elems=ce,
            cell_ij=None if cij is None else cij.detach().cpu().numpy())
        if p.n_el != conn.shape[0] or p.n_nodes != nodes.shape[0]:
            raise ValueError("the plan was built for another mesh")
        # element kind of every launch: nen, flagged for halo plans, small-chunk quad4 plans and tail-balanced plans
        small = p.mode == "partials" and self.nen == 4 and p.chunk_elems == self.lib.dfea_chunk_elems(4 | _lib.SMALL)
        tail = p.mode == "partials" and getattr(p, "tail", False)
        # quadratic meshes of at most PCG_WIDE_CHUNKS chunks: PCG workgroups with more Gauss-point groups (a property
        # of the mesh, not of the launch, so a solve gives the same bits alone and in a batch)
        wide = p.mode == "partials" and self.nen > 4 and p.n_chunks <= _lib.PCG_WIDE_CHUNKS
        self.kind = self.nen | (_lib.HALO if p.mode == "halo" else 0) | (_lib.SMALL if small else 0) | \
            (_lib.TAIL if tail else 0) | (_lib.WIDE if wide else 0)
        inv_off, inv_slot = p.inverse_map() if p.mode == "partials" else (np.zeros(1, np.int32), np.zeros(1, np.uint16))
        if p.validate(self.lib.dfea_chunk_nodes(), (inv_off, inv_slot) if p.mode == "partials" else None) > \
                self.lib.dfea_chunk_threads(self.kind):
            raise ValueError("chunk larger than the kernel's workgroup")
        E = self.lib.dfea_chunk_elems(self.kind)
        if p.mode == "partials":
            ne = p.chunks[:, 1].astype(np.int64)
            if not np.array_equal(p.chunks[:, 0], np.cumsum(ne) - ne) or ne.max() > E or ne.min() < 1:
                raise ValueError(f"partials plans must cut the element order into consecutive chunks of at most "
                                 f"{E} elements")
            if not tail and not np.array_equal(p.chunks[:, 0], np.arange(p.n_chunks) * E):
                raise ValueError(f"partials plans other than tail-balanced ones must cut the element order into "
                                 f"chunks of {E} elements")
            if tail and self.nen > 4:
                raise ValueError("tail-balanced plans are for linear elements")
        self.n_el, self.n_nodes = p.n_el, p.n_nodes
        self.n_dof = 2 * p.n_nodes
        dv = device
        # local-node tables at the fixed stride the kernels address them by (partials plans)
        dt = p.device_tables(inv_off if p.mode == "partials" else None)
        self.lstride = dt["lstride"]
        self.n_lnodes = len(dt["lnode"])
        self.t_chunks = _dev_int32(dt["chunks"], dv)
        self.t_lnode = _dev_int32(dt["lnode"], dv)
        self.t_lelem = torch.from_numpy(p.lelem.view(np.int16)).to(dv)
        self.t_sh_node = _dev_int32(p.sh_node, dv)
        self.t_sh_off = _dev_int32(p.sh_off, dv)
        self.t_pslot = _dev_int32(dt["pslot"] if len(dt["pslot"]) else np.zeros(1, np.int32), dv)
        self.t_hel = _dev_int32(p.hel if len(p.hel) else np.zeros(1, np.int32), dv)
        self.t_prange = _dev_int32(dt["prange"] if len(dt["pslot"]) else np.zeros((1, 2), np.int32), dv)
        self.t_inv_off = _dev_int32(dt["inv_off"] if p.mode == "partials" else inv_off, dv)
        self.t_inv_slot = torch.from_numpy(inv_slot.view(np.int16)).to(dv)
        self.t_hlelem = torch.from_numpy(np.ascontiguousarray(p.hlelem).view(np.int16).reshape(-1) if p.hlelem.size
                                         else np.zeros(1, np.int16)).to(dv)
        self.t_conn = _dev_int32(p.conn_int, dv)
        self.t_perm = _dev_int32(p.perm, dv)
        self.perm = torch.from_numpy(p.perm.astype(np.int64)).to(dv)
        self.t_nodes = nodes.contiguous().to(dv)
        self.t_dNdxi = torch.from_numpy(reference_gradient_table(self.element_type, GP)).contiguous().to(dv)
        self.t_gw = torch.from_numpy(GW).to(dv)
        self.t_nfix = torch.from_numpy(p.nfix).to(dv)
        # per-element planes: n_el columns + one copy column per halo entry
        self.n_cols = self.n_el + len(p.hel)
        self.dN64 = torch.empty(self.nq * self.nen * 2, self.n_cols, dtype=F64, device=dv)
        self.dN32 = torch.empty(self.nq * self.nen * 2, self.n_cols, dtype=F32, device=dv)
        self.w = torch.empty(self.nq, self.n_cols, dtype=F64, device=dv)
        # original element id of every plane column (core columns, then halo copies)
        self.col_elem = torch.from_numpy(np.r_[p.perm, p.perm[p.hel]].astype(np.int64)).to(dv)
        self.n_vblk = self.lib.dfea_vblk(self.n_dof)
        self.n_finblk = self.lib.dfea_finblk(p.n_shared)
        self.red_size = self.lib.dfea_red_size(p.n_chunks, self.n_finblk, self.n_vblk)
        # single-solve item tables (solve index 0)
        # one solve: NULL item tables (solve 0, blocks 0 .. n-1)
        self.items = self.fin_items = self.vitems = self.solve_items = None
        # geometry, once
        d = self.fill_desc(_lib.Desc())
        tab = self.upload(d)
        _lib.check(self.lib.dfea_geometry(_ptr(tab), _ptr(self.items), p.n_chunks, self.kind, _stream()),
                   "dfea_geometry")
        self._geom_table = tab

    def fill_desc(self, d):
        p = self.plan
        d.chunks, d.lnode, d.lelem = self.t_chunks.data_ptr(), self.t_lnode.data_ptr(), self.t_lelem.data_ptr()
        d.sh_node, d.sh_off, d.pslot = self.t_sh_node.data_ptr(), self.t_sh_off.data_ptr(), self.t_pslot.data_ptr()
        d.hel, d.hlelem, d.n_halo = self.t_hel.data_ptr(), self.t_hlelem.data_ptr(), len(p.hel)
        d.prange = self.t_prange.data_ptr()
        d.inv_off, d.inv_slot = self.t_inv_off.data_ptr(), self.t_inv_slot.data_ptr()
        d.n_cols = self.n_cols
        d.conn, d.perm, d.nodes = self.t_conn.data_ptr(), self.t_perm.data_ptr(), self.t_nodes.data_ptr()
        d.dNdxi, d.gw, d.nfix = self.t_dNdxi.data_ptr(), self.t_gw.data_ptr(), self.t_nfix.data_ptr()
        d.dN64, d.dN32, d.w = self.dN64.data_ptr(), self.dN32.data_ptr(), self.w.data_ptr()
        d.n_chunks, d.n_lnodes, d.n_shared, d.n_part = p.n_chunks, self.n_lnodes, p.n_shared, p.n_part
        d.lstride = self.lstride
        d.n_el, d.n_nodes, d.nen, d.nq = p.n_el, p.n_nodes, self.nen, self.nq
        d.n_finblk, d.n_vblk = self.n_finblk, self.n_vblk
        return d

    def upload(self, d):
        raw = torch.frombuffer(bytearray(bytes(d)), dtype=torch.uint8)
        return raw.to(self.device)

    # -- element-order helpers ---------------------------------------------------------------------------
    def to_internal(self, per_element):
        """Per-element values (original order) -> plane columns (internal order + halo copies)."""
        return per_element.to(self.device, F64)[self.col_elem].contiguous()


class HipSystem:
    """One solve (mesh + theta + material + precision) with all its device buffers."""

    def __init__(self, mesh, theta, mat, f, precision="fp64", uD=None, sc=None, preconditioner="jacobi"):
        """mat: a ``materials.MaterialSpec`` or, for neo-Hookean with both parameters theta-scaled, the pair
        (mu0, lmbda0)."""
        self.mesh = m = mesh
        self.lib = mesh.lib
        dv = mesh.device
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
        if preconditioner not in _lib.PRECOND:
            raise ValueError(f"Unknown preconditioner '{preconditioner}'. Choose one of {tuple(_lib.PRECOND)}.")
        self.precision = precision
        self.preconditioner = preconditioner
        self.prec_id = _lib.FP64 if precision == "fp64" else _lib.FP32
        self.vdt = F64 if precision == "fp64" else F32
        self.theta = mesh.to_internal(theta)
        if not isinstance(mat, MaterialSpec):
            mp = torch.as_tensor(mat, dtype=F64).cpu().reshape(-1)
            if mp.numel() != 2:
                raise ValueError("a bare parameter vector means neo-Hookean (mu0, lmbda0); pass a MaterialSpec")
            mat = MaterialSpec(0, (float(mp[0]), float(mp[1])) + (0.0,) * (_lib.MAT_NP - 2), 3)
        self.mat = mat
        self.kind = m.kind | (_lib.MATLIB if mat.library else 0)   # element kind of the material kernels
        self.f = f.to(dv, F64).contiguous()
        nq, n_el, nd = m.nq, m.n_el, m.n_dof
        self.C = None                    # moduli and preconditioner: allocated by the first moduli() call
        self.diag = None                 # Jacobi: diag(A); block Jacobi: inverse blocks {m00, m11} per node, m01
        self.part = torch.empty(max(1, 2 * m.plan.n_part) * 2, dtype=F64, device=dv)   # two parities of pairs
        self.red = torch.zeros(max(1, m.red_size), dtype=F64, device=dv)
        # the scalar block may be a row of a pool shared by a batch of solves
        self.sc = torch.zeros(_lib.NSC, dtype=F64, device=dv) if sc is None else sc
        self.vec = {k: torch.zeros(nd, dtype=self.vdt, device=dv) for k in _lib.PCG_VECS}
        self.vec["u"] = torch.zeros(nd, dtype=F64, device=dv)
        self.vec["R"] = torch.zeros(nd, dtype=F64, device=dv)
        self.vec["uacc"] = torch.zeros(nd, dtype=F64, device=dv)
        self.vec["uD"] = torch.zeros(nd, dtype=F64, device=dv) if uD is None else uD.to(dv, F64).contiguous()
        self._graphs = {}
        # pinned host staging of scalars read back: [0:4] PCG status, [8] dot products; the event marks the
        # copy of a dot product
        self._hst = torch.empty(9, dtype=F64, pin_memory=True)
        self._dot_ev = torch.cuda.Event()
        self._st_ev = torch.cuda.Event()
        self.desc = self._make_desc()
        self.table = m.upload(self.desc)

    def _make_desc(self):
        d = self.mesh.fill_desc(_lib.Desc())
        d.theta = self.theta.data_ptr()
        if self.C is not None:
            if self.prec_id == _lib.FP64:
                d.C64, d.diag64 = self.C.data_ptr(), self.diag.data_ptr()
            else:
                d.C32, d.diag32 = self.C.data_ptr(), self.diag.data_ptr()
        d.f, d.part, d.red, d.sc = self.f.data_ptr(), self.part.data_ptr(), self.red.data_ptr(), self.sc.data_ptr()
        for k, i in _lib.VEC.items():
            d.vec[i] = self.vec[k].data_ptr()
        for k, v in enumerate(self.mat.p):
            d.mat_p[k] = v
        d.material, d.mat_scal = self.mat.mid, self.mat.scal
        d.precision = self.prec_id
        d.precond = _lib.PRECOND[self.preconditioner]
        return d

    # -- kernels on this single system -------------------------------------------------------------------
    def _ensure_moduli_buffers(self):
        if self.C is None:
            m = self.mesh
            self.C = torch.empty(10 * m.nq, m.n_cols, dtype=self.vdt, device=m.device)
            n = m.n_dof + (m.n_nodes if self.preconditioner == "block_jacobi" else 0)
            self.diag = torch.empty(n, dtype=self.vdt, device=m.device)
            self.release_graphs()             # graphs hold the previous table's address
            self.desc = self._make_desc()
            self.table = m.upload(self.desc)

    def moduli(self, u):
        self._ensure_moduli_buffers()
        m = self.mesh
        if u.dtype != F64 or not u.is_contiguous():
            u = u.to(F64).contiguous()
        _lib.check(self.lib.dfea_moduli(_ptr(self.table), _ptr(m.items), m.plan.n_chunks, _ptr(m.fin_items),
                                        m.n_finblk, _ptr(u), self.kind, self.prec_id, _stream()), "dfea_moduli")

    def precondition(self, r):
        """u = M r with the preconditioner of the last moduli() call (device tensors, the system's precision)."""
        r = r.to(self.vdt)
        if self.preconditioner == "jacobi":
            return r / self.diag
        nd = self.mesh.n_dof
        rn = r.reshape(-1, 2)
        d = self.diag[:nd].reshape(-1, 2)
        o = self.diag[nd:]
        return torch.stack((d[:, 0] * rn[:, 0] + o * rn[:, 1], o * rn[:, 0] + d[:, 1] * rn[:, 1]), 1).reshape(-1)

    def residual(self, u, out=None, lf=1.0):
        m = self.mesh
        if u.dtype != F64 or not u.is_contiguous():
            u = u.to(F64).contiguous()
        out = torch.empty_like(u) if out is None else out
        _lib.check(self.lib.dfea_residual(_ptr(self.table), _ptr(m.items), m.plan.n_chunks, _ptr(m.fin_items),
                                          m.n_finblk, _ptr(u), _ptr(out), float(lf), self.kind, _stream()),
                   "dfea_residual")
        return out

    def apply(self, v, out=None):
        m = self.mesh
        if v.dtype != self.vdt or not v.is_contiguous():
            v = v.to(self.vdt).contiguous()
        out = torch.empty_like(v) if out is None else out
        # one solve: the host copy of the descriptor lets the shared-node finish take its pointers as arguments
        _lib.check(self.lib.dfea_apply_host(ctypes.byref(self.desc), _ptr(self.table), m.plan.n_chunks, m.n_finblk,
                                            _ptr(v), _ptr(out), self.kind, self.prec_id, _stream()), "dfea_apply_host")
        return out

    def dot(self, x, y=None):
        """Deterministic FP64 dot product on the device; returns a 0-d device tensor (no host sync)."""
        m = self.mesh
        x = x.to(F64).contiguous()
        y = x if y is None else y.to(F64).contiguous()
        _lib.check(self.lib.dfea_dot(_ptr(self.table), _ptr(m.vitems), m.n_vblk, _ptr(m.solve_items), 1,
                                     _ptr(x), _ptr(y), _stream()), "dfea_dot")
        return self.sc[_lib.SC["norm"]].clone()

    def dot_host(self, x, y=None):
        """The same dot product read to the host (one copy through pinned memory, one synchronisation)."""
        return self.dot_start(x, y)()

    def dot_start(self, x, y=None):
        """Enqueue the dot product and its copy to pinned host memory; returns a function that waits for that
        copy alone (work enqueued after this call keeps running) and gives the value."""
        m = self.mesh
        if x.dtype != F64 or not x.is_contiguous():
            x = x.to(F64).contiguous()
        if y is None:
            y = x
        elif y.dtype != F64 or not y.is_contiguous():
            y = y.to(F64).contiguous()
        _lib.check(self.lib.dfea_dot(_ptr(self.table), _ptr(m.vitems), m.n_vblk, _ptr(m.solve_items), 1,
                                     _ptr(x), _ptr(y), _stream()), "dfea_dot")
        n = _lib.SC["norm"]
        h = self._hst[8:9]
        h.copy_(self.sc[n:n + 1], non_blocking=True)
        ev = self._dot_ev
        ev.record()

        def wait():
            ev.synchronize()
            return float(h[0])
        return wait

    def norm(self, x):
        return torch.sqrt(self.dot(x))

    def lift(self, u0, uD, lf):
        """where(fixed, lf * u_D, u0) on the device."""
        m = self.mesh
        out = torch.empty(m.n_dof, dtype=F64, device=m.device)
        a = torch.full((1,), float(lf), dtype=F64, device=m.device)
        _lib.check(self.lib.dfea_axpy(_ptr(self.table), _ptr(m.vitems), m.n_vblk, 0, _ptr(u0.to(F64).contiguous()),
                                      None, _ptr(uD.to(F64).contiguous()), _ptr(a), _ptr(out), self.prec_id,
                                      _stream()), "dfea_axpy")
        return out

    def reactions(self, u, out=None):
        """(Rx, Ry): the internal force summed over the fixed dofs (device tensor of 2)."""
        m = self.mesh
        out = torch.empty(2, dtype=F64, device=m.device) if out is None else out
        _lib.check(self.lib.dfea_reactions(_ptr(self.table), _ptr(m.items), m.plan.n_chunks, _ptr(m.solve_items), 1,
                                           _ptr(u.to(F64).contiguous()), _ptr(out), self.kind, _stream()),
                   "dfea_reactions")
        return out

    def fields_buffers(self):
        """(eout, nout, work) for fields(); nout starts at zero (nodes that no element touches stay zero)."""
        m = self.mesh
        return (torch.empty(7, m.n_el, dtype=F64, device=m.device),
                torch.zeros(6, m.n_nodes, dtype=F64, device=m.device),
                torch.empty(max(1, 8 * m.plan.n_part), dtype=F64, device=m.device))

    def fields(self, u, bufs=None):
        """Element fields [7, n_el] (area, J, sxx, syy, sxy, mises, psi; original element order) and nodal
        averages [6, N] (J, sxx, syy, sxy, mises, psi) of the state u; bufs from fields_buffers() may be
        reused across calls."""
        m = self.mesh
        eout, nout, work = self.fields_buffers() if bufs is None else bufs
        _lib.check(self.lib.dfea_fields(_ptr(self.table), _ptr(m.items), m.plan.n_chunks, _ptr(m.fin_items),
                                        m.n_finblk, _ptr(u.to(F64).contiguous()), _ptr(eout), _ptr(nout),
                                        _ptr(work), self.kind, _stream()), "dfea_fields")
        return eout, nout

    def adjoint_grad(self, u, lam):
        m = self.mesh
        g = torch.empty(m.n_el, dtype=F64, device=m.device)
        _lib.check(self.lib.dfea_adjoint_grad(_ptr(self.table), _ptr(m.items), m.plan.n_chunks,
                                              _ptr(u.to(F64).contiguous()), _ptr(lam.to(F64).contiguous()), _ptr(g),
                                              self.kind, _stream()), "dfea_adjoint_grad")
        return g

    # -- fused PCG ---------------------------------------------------------------------------------------------
    def pcg_start(self, b=None, rtol=1e-5, max_iter=500, neg_res=False):
        """Set up PCG for A x = b (b None & neg_res: b = -vec['R']); needs moduli() first."""
        if self.C is None:
            raise RuntimeError("pcg_start needs the moduli of the linearisation point (call moduli(u) first)")
        m = self.mesh
        if b is not None:
            self.vec["b"].copy_(b)
        # one launch pair; the tolerance and the cap travel as kernel arguments
        _lib.check(self.lib.dfea_pcg_start(_ptr(self.table), m.n_vblk, float(rtol), int(max_iter), int(neg_res),
                                           self.prec_id, _stream()), "dfea_pcg_start")

    # solve rounds as direct launches (HIP graphs on request): a graph launch costs 35-60 us of host time before its
    # first kernel starts, while direct launches cost ~3 us each and keep ahead of kernels of >= ~7 us; measured on
    # MI355X, the 100x100 Q4 solve (2,953 PCG kernels in 26 Newton steps) 27.7 ms with graphs, 26.7 ms direct
    use_graphs = False

    def _graph(self, n_iter):
        """Executable HIP graph of n_iter fused PCG iterations (captured once per chunk length)."""
        g = self._graphs.get(n_iter)
        if g is None:
            m = self.mesh
            h = ctypes.c_void_p()
            _lib.check(self.lib.dfea_pcg_graph_create(
                _ptr(self.table), _ptr(m.items), m.plan.n_chunks, int(n_iter), self.kind, self.prec_id,
                ctypes.byref(h)), "dfea_pcg_graph_create")
            self._graphs[n_iter] = g = h
        return g

    def pcg_iterate(self, n_iter, graph=None):
        m = self.mesh
        if n_iter <= 0:
            return
        if self.use_graphs if graph is None else graph:
            _lib.check(self.lib.dfea_graph_launch(self._graph(int(n_iter)), _stream()), "dfea_graph_launch")
            return
        _lib.check(self.lib.dfea_pcg_iterate(_ptr(self.table), _ptr(m.items), m.plan.n_chunks, int(n_iter), self.kind,
                                             self.prec_id, _stream()), "dfea_pcg_iterate")

    def release_graphs(self):
        for g in self._graphs.values():
            self.lib.dfea_graph_destroy(g)
        self._graphs = {}

    def __del__(self):
        try:
            self.release_graphs()
        except Exception:
            pass

    def pcg_status(self):
        """(status, iterations, k) of the solve (one copy through pinned memory, one synchronisation)."""
        return self.pcg_status_start()()

    def pcg_status_start(self):
        """Enqueue the copy of the status block to pinned host memory; returns a function that waits for that
        copy alone (work enqueued after this call keeps running) and gives (status, iterations, k)."""
        h = self._hst[0:4]
        h.copy_(self.sc[_SC_STATUS:_SC_STATUS + 4], non_blocking=True)    # status, k, maxit, iters
        ev = self._st_ev
        ev.record()

        def wait():
            ev.synchronize()
            st, k, _, its = h.tolist()
            return int(st), int(its), int(k)
        return wait

    ROUND_GRAPHS = (128, 64, 32, 16)
    LEAD_DIRECT = 2          # graph rounds: kernels launched directly at the start of a solve (even: pairs)

    def pcg_round(self, n_iter, graph=None):
        """n_iter PCG kernels (rounded up to even): replays of the graphs of ROUND_GRAPHS kernels (each
        captured once per system), then direct launches of the remainder."""
        if not (self.use_graphs if graph is None else graph):
            self.pcg_iterate(n_iter, False)
            return
        n = n_iter + (n_iter & 1)
        for g in self.ROUND_GRAPHS:
            while n >= g:
                self.pcg_iterate(g, True)
                n -= g
        if n:
            self.pcg_iterate(n, False)

    def pcg_run(self, max_iter, chunk=16, graph=None, first=None, after_round=None):
        """Iterate a started PCG until it stops; returns (iterations, status).

        Kernels run in rounds and the status block is read after each round; kernels of a solve that has
        stopped skip their work on the device.  The first round has `first` kernels (default `chunk`: an
        estimate of the iterations saves rounds) and is launched without reading the status before it; later
        rounds have `chunk` kernels, doubling up to 128.  The first kernel forms A M b and the test on the last
        update is made by the kernel after it, so k iterations take k + 2 kernels.  ``after_round``: a function
        called after each round's status copy is enqueued and before the host waits for it -- work it enqueues
        (that must not change the PCG's inputs) runs on the GPU meanwhile, on the round's answer so far."""
        k = 0
        n = chunk if first is None else max(2, int(first))
        n = min(n, max_iter + 2)
        if self.LEAD_DIRECT and n > self.LEAD_DIRECT and (self.use_graphs if graph is None else graph):
            # the first kernels of the solve as direct launches: the GPU starts on them while the host submits the
            # graphs of the rest of the round (a graph launch takes tens of microseconds of host time)
            self.pcg_iterate(self.LEAD_DIRECT, False)
            n -= self.LEAD_DIRECT
        while True:
            self.pcg_round(min(n, max_iter - k + 2), graph)
            status = self.pcg_status_start()
            if after_round is not None:
                after_round()
            st, its, k = status()
            if st != _lib.PCG_RUNNING:
                return its, st
            n = chunk
            chunk = min(2 * chunk, 128)

    def pcg_solve(self, b=None, rtol=1e-5, max_iter=500, neg_res=False, chunk=16, graph=None, first=None):
        """Run the fused PCG to completion; returns (x view, iterations, status)."""
        self.pcg_start(b, rtol, max_iter, neg_res)
        its, st = self.pcg_run(max_iter, chunk, graph, first)
        return self.vec["x"], its, st
