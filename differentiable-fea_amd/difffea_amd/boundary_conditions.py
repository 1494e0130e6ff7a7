"""Dirichlet selection and consistent edge loads (host-side set-up, CPU float64).

Restates the conventions of `src/difffea/boundary_conditions.py:141-277`:
bounding-box side selection with tolerance ``tol * max(extent)``, node-major dof numbering ``2a+i``,
boundary edges = edges used by exactly one element (keyed by their end nodes), and work-equivalent nodal
forces of a uniform traction integrated with a 3-point Gauss rule per edge.
"""
import numpy as np
import torch

from .elements import EDGES, edge_shape, element_order, gauss_1d


def select_boundary_nodes(nodes, side, tol=1e-6):
    """(N,) bool mask of the nodes on one side ('left','right','bottom','top') of the bounding box."""
    x, y = nodes[:, 0], nodes[:, 1]
    eps = tol * torch.max(torch.max(x) - torch.min(x), torch.max(y) - torch.min(y))
    if side == "left":
        return x <= torch.min(x) + eps
    if side == "right":
        return x >= torch.max(x) - eps
    if side == "bottom":
        return y <= torch.min(y) + eps
    if side == "top":
        return y >= torch.max(y) - eps
    raise ValueError("side must be 'left', 'right', 'bottom' or 'top'")


def dofs_from_node_mask(node_mask, components=(0, 1)):
    """Node mask (N,) -> flat dof mask (2N,) for the given components; dof(a, i) = 2a + i."""
    per_node = torch.zeros((node_mask.shape[0], 2), dtype=torch.bool)
    for c in components:
        per_node[:, c] = node_mask
    return per_node.reshape(-1)


def boundary_edges(conn, edge_local_nodes):
    """Edges that belong to exactly one element: (n_boundary_edges, k) global ids [start, end(, mid)]."""
    c = np.asarray(conn)
    all_edges = np.stack([c[:, list(e)] for e in edge_local_nodes], axis=1).reshape(-1, len(edge_local_nodes[0]))
    a, b = all_edges[:, 0], all_edges[:, 1]
    big = int(c.max()) + 1
    key = np.minimum(a, b).astype(np.int64) * big + np.maximum(a, b)
    _, inverse, counts = np.unique(key, return_inverse=True, return_counts=True)
    return torch.from_numpy(all_edges[counts[inverse] == 1].astype(np.int64))


def edge_traction_load(nodes, edges, order, node_mask, total_force, n_gauss=3):
    """Consistent nodal forces (N, 2) of a uniform traction giving ``total_force`` on the masked segment.

    f_a = sum_g w_g N_a(s_g) |dX/ds|_g t, with t = total_force / (segment length).
    """
    X = np.asarray(nodes, dtype=np.float64)
    f = np.zeros_like(X)
    E = np.asarray(edges)
    on = np.asarray(node_mask)[E].all(axis=1)
    E = E[on]
    if E.shape[0] == 0:
        return torch.from_numpy(f)
    s, w = gauss_1d(n_gauss)
    N, dN = edge_shape(order, s)                              # (g, k)
    Xe = X[E]                                                 # (E, k, 2)
    tangent = np.einsum("gk,ekd->egd", dN, Xe)
    speed = np.linalg.norm(tangent, axis=-1)                  # (E, g)
    length = np.einsum("eg,g->", speed, w)
    t = np.asarray(total_force, dtype=np.float64) / length
    f_edge = np.einsum("g,ga,eg,i->eai", w, N, speed, t)      # (E, k, 2)
    np.add.at(f, E.reshape(-1), f_edge.reshape(-1, 2))
    return torch.from_numpy(f)


def element_edges(element_type):
    return EDGES[element_type]


def traction_order(element_type):
    return element_order(element_type)
