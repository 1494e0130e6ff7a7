"""The set-up on the path's input side (meshes, boundary selection, dof numbering, consistent edge loads),
checked on the CPU against what the reference's own tests demand of it (`tests/test_mesh_boundary.py:40-80`
of the reference: boundary edges lie on the boundary, the consistent load sums to the applied force and
vanishes off the loaded side, a quadratic edge splits a uniform traction 1/6 : 4/6 : 1/6) and against closed
forms (node counts, positive element Jacobians, node-major dof numbering, the O-grid's hole and outer
square)."""
import numpy as np
import pytest
import torch

from difffea_amd.boundary_conditions import (boundary_edges, dofs_from_node_mask, edge_traction_load,
                                             element_edges, select_boundary_nodes, traction_order)
from difffea_amd.mesh import create_plate_hole_ogrid_q9, create_rectangle_mesh
from difffea_amd.problem import build_problem

ETS = ["quad4", "quad8", "quad9", "tri3", "tri6"]
# nodes per element and element count per cell of the structured meshes
NEN = {"quad4": 4, "quad8": 8, "quad9": 9, "tri3": 3, "tri6": 6}
PER_CELL = {"quad4": 1, "quad8": 1, "quad9": 1, "tri3": 2, "tri6": 2}


def _n_nodes(et, nx, ny):
    if et in ("quad4", "tri3"):
        return (nx + 1) * (ny + 1)
    return (2 * nx + 1) * (2 * ny + 1) - (nx * ny if et == "quad8" else 0)


@pytest.mark.parametrize("et", ETS)
def test_mesh_sizes_and_counter_clockwise_elements(et):
    nodes, conn = create_rectangle_mesh(et, 2.0, 1.0, 3, 2)
    nc = 3 if et.startswith("tri") else 4
    assert nodes.shape == (_n_nodes(et, 3, 2), 2) and conn.shape == (6 * PER_CELL[et], NEN[et])
    assert torch.equal(torch.unique(conn), torch.arange(nodes.shape[0]))    # every node used, ids consecutive
    X = nodes[conn[:, :nc]].numpy()                      # corners, counter-clockwise
    area = 0.5 * np.sum(X[:, :, 0] * np.roll(X[:, :, 1], -1, 1) - np.roll(X[:, :, 0], -1, 1) * X[:, :, 1], axis=1)
    assert np.allclose(area, 2.0 / (6 * PER_CELL[et]))   # every element positive, total = the rectangle
    if NEN[et] > nc:                                     # mid-side nodes at the edge midpoints, in edge order
        mids = nodes[conn[:, nc:2 * nc]].numpy()
        assert np.allclose(mids, 0.5 * (X + np.roll(X, -1, 1)))
    if et == "quad9":
        assert np.allclose(nodes[conn[:, 8]].numpy(), X.mean(axis=1))
    if et.startswith("tri"):                             # lower-left triangles first: (n0, n1, n3) of each cell
        qn, q = create_rectangle_mesh("quad4", 2.0, 1.0, 3, 2)
        assert np.allclose(X[:6], qn[q[:, [0, 1, 3]]].numpy()) and np.allclose(X[6:], qn[q[:, [1, 2, 3]]].numpy())


@pytest.mark.parametrize("et", ETS)
def test_boundary_edges_lie_on_the_boundary(et):
    nodes, conn = create_rectangle_mesh(et, 1.0, 1.0, 3, 2)
    edges = boundary_edges(conn, element_edges(et))
    assert edges.shape[0] == 2 * (3 + 2)                 # one boundary edge per cell side on the perimeter
    assert edges.shape[1] == (2 if et in ("quad4", "tri3") else 3)
    xy = nodes[edges].numpy()
    on = (xy[..., 0] < 1e-12) | (xy[..., 0] > 1 - 1e-12) | (xy[..., 1] < 1e-12) | (xy[..., 1] > 1 - 1e-12)
    assert on.all()


@pytest.mark.parametrize("et", ETS)
def test_consistent_load_total_force_and_support(et):
    nodes, conn = create_rectangle_mesh(et, 2.0, 1.0, 3, 2)
    edges = boundary_edges(conn, element_edges(et))
    mask = select_boundary_nodes(nodes, "right")
    f = edge_traction_load(nodes, edges, traction_order(et), mask, (5.0, -3.0))
    assert torch.allclose(f.sum(0), torch.tensor([5.0, -3.0], dtype=torch.float64), atol=1e-12)
    assert float(f[~mask].abs().max()) == 0.0


def test_edge_load_split_linear_and_quadratic():
    for et, expect in (("quad4", [3.0, 3.0]), ("quad9", [1.0, 1.0, 4.0]), ("quad8", [1.0, 1.0, 4.0]),
                       ("tri3", [3.0, 3.0]), ("tri6", [1.0, 1.0, 4.0])):
        nodes, conn = create_rectangle_mesh(et, 1.0, 1.0, 1, 1)
        edges = boundary_edges(conn, element_edges(et))
        mask = select_boundary_nodes(nodes, "right")
        f = edge_traction_load(nodes, edges, traction_order(et), mask, (6.0, 0.0))
        right = torch.where(mask)[0]
        assert sorted(f[right, 0].tolist()) == pytest.approx(expect, abs=1e-12)


def test_node_major_dof_numbering_and_side_selection():
    nodes, _ = create_rectangle_mesh("quad4", 1.0, 1.0, 2, 2)
    left = select_boundary_nodes(nodes, "left")
    assert torch.equal(left, nodes[:, 0] == 0.0) and int(left.sum()) == 3
    d = dofs_from_node_mask(left, components=(1,))
    assert torch.equal(torch.where(d)[0], 2 * torch.where(left)[0] + 1)
    with pytest.raises(ValueError):
        select_boundary_nodes(nodes, "front")


def test_problem_masks_and_prescribed_displacement():
    p = build_problem(dict(element_type="quad9", nx=4, ny=3, E=1.0, nu=0.3, load_mode="displacement",
                           load_value=(0.1, None), fixed_side="left", load_side="right"))
    X = p["nodes"]
    left = X[:, 0] < 1e-12
    right = X[:, 0] > 1 - 1e-12
    fixed = p["fixed"].reshape(-1, 2)
    assert fixed[left].all() and fixed[right, 0].all() and not fixed[right, 1].any()
    assert torch.equal(p["m"], (~p["fixed"]).to(torch.float64))
    uD = p["u_D"].reshape(-1, 2)
    assert torch.allclose(uD[right, 0], torch.full((int(right.sum()),), 0.1, dtype=torch.float64))
    assert float(uD[~right].abs().max()) == 0.0 and float(p["f"].abs().max()) == 0.0


def test_ogrid_hole_and_outer_square():
    p = create_plate_hole_ogrid_q9(width=1.0, radius=0.2, n_rad=4, n_circ=16)
    nodes, conn = p[0], p[1]
    r = torch.linalg.norm(nodes, dim=1)
    assert float(r.min()) == pytest.approx(0.2, abs=1e-12)
    assert float(nodes.abs().max()) == pytest.approx(0.5, abs=1e-12)
    assert conn.shape == (4 * 16, 9)
    X = nodes[conn[:, :4]].numpy()
    area = 0.5 * np.sum(X[:, :, 0] * np.roll(X[:, :, 1], -1, 1) - np.roll(X[:, :, 0], -1, 1) * X[:, :, 1], axis=1)
    assert (area > 0).all()


def test_fifo_queue_latencies():
    """The queue in front of a one-request-at-a-time solver: idle server, back-to-back arrivals, a burst."""
    from difffea_amd.serve import fifo_latencies
    assert fifo_latencies([0.0, 10.0], [1.0, 2.0]) == [1.0, 2.0]              # no waiting
    assert fifo_latencies([0.0, 0.5, 0.5], [1.0, 1.0, 1.0]) == [1.0, 1.5, 2.5]  # each waits for the one before
    assert fifo_latencies([], []) == []


def test_element_library_entries():
    """The library rows of the reference (`elements.py:38-57`): node counts, families, orders, corner counts,
    edge tables, default rules; reference node coordinates are where the shape functions are Kronecker."""
    from difffea_amd import elements
    assert elements.element_names() == ["quad4", "quad8", "quad9", "tri3", "tri6"]
    for name in elements.element_names():
        e = elements.get_element(name)
        X = elements.reference_node_coordinates(name)
        assert X.shape == (e["n_nodes"], 2)
        assert np.allclose(elements.shape_values(name, X[:, 0], X[:, 1]), np.eye(e["n_nodes"]), atol=1e-15)
        assert len(e["edges"]) == e["n_corners"] and all(len(ed) == (2 if e["order"] == 1 else 3) for ed in e["edges"])
    assert elements.get_element("tri6")["gauss"] == ("triangle", 4)
    with pytest.raises(ValueError):
        elements.get_element("hex8")
