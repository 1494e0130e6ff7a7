"""The ``hip`` post-processing of a solution, mirroring `src/difffea/postprocess.py`.

* ``compute_fields(u, theta, params) -> dict``   (`postprocess.py:51-118`)
  per element: ``J``, ``sxx``, ``syy``, ``sxy`` (Cauchy stress), ``mises`` (with the plane-strain
  ``sigma_zz``), ``energy_density`` -- quadrature-weighted means over the element -- and their nodal averages
  (suffix ``_n``); nodal ``ux``, ``uy``, ``u_mag`` (and ``_n`` aliases); scalars ``strain_energy`` and
  ``min_J``.  One kernel per call (``dfea_fields``: element states, LDS accumulation of the nodal sums,
  shared nodes summed in slot order); the scalars are device reductions of its output.
* ``support_reactions(u, theta, params) -> (2,)``  (`postprocess.py:121-137`)
  the internal force (all-ones mask, no external load) summed over the constrained dofs
  (``dfea_reactions``: per-chunk sums in a fixed order, no atomics).

Every element type (tri3, quad4, tri6, quad8, quad9) and every library material (neo-Hookean, Mooney-Rivlin,
Saint Venant-Kirchhoff, Demiray, HGO fibres), as the rest of the ``hip`` variant.
"""
import torch

from .operators import make_system

_ELEM = ("area", "J", "sxx", "syy", "sxy", "mises", "energy_density")
_NODE = ("J", "sxx", "syy", "sxy", "mises", "energy_density")


def compute_fields(u, theta, params) -> dict:
    """All post-processing fields of a solution u (device tensors; same keys as the reference)."""
    u = u.detach()
    s = make_system(u, theta, params)
    eout, nout = s.fields(u)
    e = dict(zip(_ELEM, eout.unbind(0)))
    un = u.reshape(-1, 2).to(torch.float64)
    out = {k: e[k] for k in _ELEM[1:]}
    out.update(ux=un[:, 0], uy=un[:, 1], u_mag=torch.linalg.norm(un, dim=1),
               strain_energy=torch.sum(e["energy_density"] * e["area"]), min_J=e["J"].min())
    for k in ("ux", "uy", "u_mag"):
        out[k + "_n"] = out[k]
    for k, v in zip(_NODE, nout.unbind(0)):
        out[k + "_n"] = v
    return out


def support_reactions(u, theta, params) -> torch.Tensor:
    """Total reaction (Rx, Ry) at the constrained dofs; sum(reactions) + sum(applied loads) = 0 at
    equilibrium."""
    u = u.detach()
    return make_system(u, theta, params).reactions(u)


__all__ = ["compute_fields", "support_reactions"]
