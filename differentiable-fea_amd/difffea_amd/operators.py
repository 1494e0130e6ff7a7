"""The ``hip`` tangent / preconditioner / residual variants behind the reference's operator API.

Mirrors `src/difffea/operators.py`:

* ``make_tangent_operator(kind, u, theta, params) -> A``      (`operators.py:415-435`)
  ``A(v) = m * J(u)(m v) + (1 - m) v``; kinds ``TANGENT_OPERATORS = ("auto", "hip")``; unknown kind ->
  ``ValueError``.  ``"auto"`` is ``"hip"`` (this package only ships the hip variant).
* ``make_AT_operator(u, theta, params)``                       (`operators.py:441-471`)
  the hyperelastic tangent is symmetric, so the transpose product is the same kernel.
* ``make_preconditioner(kind, u, theta, params)``              (`operators.py:585, 702-710`)
  ``"jacobi"``: ``r -> r / diag`` with the diagonal from the fused moduli kernel (clamp 1e-6, pivot 1);
  ``"block_jacobi"``: ``r -> B_n^-1 r_n`` per node with the inverse nodal 2x2 blocks from the same kernel
  (`operators.py:588-638`).  ``"ilu"`` needs the assembled sparse matrix on the host and is not offered.
* ``compute_jacobi_diagonal(u, theta, params)``                (`operators.py:477-556`)
* ``global_residual(u, theta, params, load_factor)``           (`operators.py:159-220`)
* ``precompute_geometry(params)`` / ``get_geometry(params)``    (`operators.py:85-128`)
  ``dN_dX`` (n_el, n_q, nen, 2) and ``w_detJ`` (n_el, n_q) from the device geometry kernel, in the
  original element order, on the device of ``params["nodes"]``.

Every closure keeps its own device buffers (moduli, diagonal), so closures built at different states stay
valid side by side, as the reference's functional closures do.
"""
import weakref

import torch

from . import _lib
from .device import HipMesh, HipSystem
from .materials import material_spec

TANGENT_OPERATORS = ("auto", "hip")
PRECONDITIONERS = ("jacobi", "block_jacobi")
PRECISIONS = ("fp64", "mixed")

_MESHES = {}


def _check_material(params):
    """The material of ``params`` as the kernels take it (id, nominal parameters, theta-scaled entries)."""
    return material_spec(params)


def precompute_geometry(params, device=None):
    """``params`` plus the geometry cache ``dN_dX`` (n_el, n_q, nen, 2) = dN_a/dX_j and ``w_detJ`` (n_el, n_q)
    = w_q det J, computed by the geometry kernel (`src/difffea/operators.py:85-120`) and returned in the
    original element order on the device of ``params["nodes"]``."""
    mesh = get_mesh(params, device)
    n_el, nq, nen = mesh.n_el, mesh.nq, mesh.nen
    dN = mesh.dN64[:, :n_el].reshape(nq, nen, 2, n_el).permute(3, 0, 1, 2)     # internal element order
    w = mesh.w[:, :n_el].t()
    dN_dX = torch.empty_like(dN)
    w_detJ = torch.empty_like(w)
    dN_dX[mesh.perm] = dN
    w_detJ[mesh.perm] = w
    dev = params["nodes"].device
    return {**params, "dN_dX": dN_dX.contiguous().to(dev), "w_detJ": w_detJ.contiguous().to(dev)}


def get_geometry(params, device=None):
    """(dN_dX, w_detJ) of ``params``: the cached entries when present, else the geometry kernel's
    (`src/difffea/operators.py:123-128`)."""
    if "dN_dX" in params and "w_detJ" in params:
        return params["dN_dX"], params["w_detJ"]
    g = precompute_geometry(params, device)
    return g["dN_dX"], g["w_detJ"]


def _mesh_tensors(params):
    return params["nodes"], params["conn"], params["fixed"]


def get_mesh(params, device=None):
    """Device plan + geometry of ``params``: built once and reused while the mesh tensors (nodes, conn, fixed)
    are alive and unmodified (an in-place edit bumps a tensor's version and the mesh is rebuilt, as the
    reference recomputes from the tensors on every call); an entry is dropped as soon as one of its tensors
    is freed, so the device memory of meshes the caller no longer holds is released."""
    device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    if device.type != "cuda":
        raise ValueError("the hip variant runs on a GPU: move the state to a cuda (ROCm) device")
    ts = _mesh_tensors(params)
    key = (tuple(id(t) for t in ts), str(device))
    versions = tuple(t._version for t in ts)
    ent = _MESHES.get(key)
    if ent is not None:
        refs, vers, mesh = ent
        if vers == versions and all(r() is t for r, t in zip(refs, ts)):
            return mesh
        del _MESHES[key]
    with torch.cuda.device(device):
        mesh = HipMesh(params, device)
    # the callback runs when a tensor is freed (before its id can be reused), so a key never outlives them
    refs = [weakref.ref(t, lambda _r, k=key: _MESHES.pop(k, None)) for t in ts]
    _MESHES[key] = (refs, versions, mesh)
    return mesh


def cached_meshes():
    """Number of device meshes held by the cache (for tests and diagnostics)."""
    return len(_MESHES)


def make_system(u, theta, params, precision="fp64", preconditioner="jacobi"):
    if precision not in PRECISIONS:
        raise ValueError(f"Unknown precision '{precision}'. Choose one of {PRECISIONS}.")
    mesh = get_mesh(params, u.device)
    return HipSystem(mesh, theta.detach(), _check_material(params), params["f"], precision,
                     preconditioner=preconditioner)


class HipTangent:
    """Callable closure ``v -> A(v)`` holding the moduli and preconditioner at its linearisation point."""

    def __init__(self, u, theta, params, precision="fp64", preconditioner="jacobi"):
        self.system = make_system(u, theta, params, precision, preconditioner)
        self.system.moduli(u.detach())
        self._u = u

    def __call__(self, v):
        y = self.system.apply(v)
        return y if y.dtype == v.dtype else y.to(v.dtype)

    @property
    def diag(self):
        return self.system.diag


_LAST = {}


def make_A_operator_hip(u, theta, params, precision="fp64"):
    A = HipTangent(u, theta, params, precision)
    _LAST["op"] = (weakref.ref(u), u._version, id(theta), id(params), A)
    return A


def make_tangent_operator(kind, u, theta, params, precision="fp64"):
    """Build the forward tangent operator by name (``"auto"`` or ``"hip"``)."""
    if kind in ("auto", "hip"):
        return make_A_operator_hip(u, theta, params, precision)
    raise ValueError(f"Unknown tangent operator '{kind}'. Choose one of {TANGENT_OPERATORS}.")


def make_AT_operator(u, theta, params, precision="fp64"):
    """Transpose product; the hyperelastic tangent with the symmetric pivot is symmetric, so A^T = A."""
    return make_A_operator_hip(u, theta, params, precision)


def _diag_at(u, theta, params, precision="fp64"):
    ent = _LAST.get("op")
    if ent is not None:
        ref, ver, tid, pid, A = ent
        if ref() is u and u._version == ver and tid == id(theta) and pid == id(params) and \
                A.system.precision == precision:
            return A.diag
    return HipTangent(u, theta, params, precision).diag


def compute_jacobi_diagonal(u, theta, params):
    """diag(A): free dofs clamped to >= 1e-6, constrained dofs 1 (FP64)."""
    return _diag_at(u, theta, params).clone()


def make_jacobi_preconditioner(u, theta, params, precision="fp64"):
    diag = _diag_at(u, theta, params, precision)
    return lambda r: r / (diag if diag.dtype == r.dtype else diag.to(r.dtype))


def make_block_jacobi_preconditioner(u, theta, params, precision="fp64"):
    """r -> B_n^-1 r_n per node, the inverse blocks from the moduli kernel (`operators.py:618-638`)."""
    system = HipTangent(u, theta, params, precision, "block_jacobi").system
    return lambda r: system.precondition(r).to(r.dtype)


def make_preconditioner(kind, u, theta, params, precision="fp64"):
    """Build a preconditioner by name: ``"jacobi"`` or ``"block_jacobi"``."""
    if kind == "jacobi":
        return make_jacobi_preconditioner(u, theta, params, precision)
    if kind == "block_jacobi":
        return make_block_jacobi_preconditioner(u, theta, params, precision)
    raise ValueError(f"Unknown preconditioner '{kind}'. Choose one of {PRECONDITIONERS}.")


def global_residual(u, theta, params, load_factor=1.0):
    """R = m * (R_int(u) - lf f) with the analytic neo-Hookean stress (FP64)."""
    sys_ = make_system(u, theta, params, "fp64")
    return sys_.residual(u.detach(), lf=float(load_factor))


__all__ = ["TANGENT_OPERATORS", "PRECONDITIONERS", "make_tangent_operator", "make_AT_operator",
           "make_preconditioner", "make_jacobi_preconditioner", "make_block_jacobi_preconditioner",
           "compute_jacobi_diagonal", "global_residual",
           "get_mesh", "make_system", "HipTangent", "_lib"]
