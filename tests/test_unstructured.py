"""Unstructured meshes on the CPU: the mesher's exchange format (`src/difffea/gmsh_mesh.py:135-188` of the
reference: an ``.npz`` with ``xy`` and ``conn``), unused-node filtering and the counter-clockwise repair
(`gmsh_mesh.py:41-103`), checked against what fixes them independently of their code: a mirrored mesh is
clockwise everywhere and must come back with det J > 0 at every Gauss point and every mid-side node at the
midpoint of its edge; renumbered and padded files must give the same elements; the Gmsh-free mesher gives
valid, irregular meshes; the chunk planner assembles them like the oracle."""
import numpy as np
import pytest
import torch

from oracle import fem
from difffea_amd import elements
from difffea_amd.mesh import (N_CORNERS, check_and_fix_ccw, create_rectangle_mesh, create_unstructured_plate_hole,
                              filter_unused_nodes, load_mesh_npz, save_mesh_npz, signed_corner_area)
from difffea_amd.problem import build_problem

ETS = list(elements.ELEMENT_TYPES)


def _geometry_ok(nodes, conn, et):
    """det J > 0 at every Gauss point (oracle geometry) and mid-side / centre nodes where the element
    library puts them (straight edges)."""
    GP, GW = fem.element_rule(et)
    p = fem.precompute_geometry(dict(nodes=nodes, conn=conn, GP=GP, GW=GW, element_type=et))
    assert float(p["w_detJ"].min()) > 0
    nc = N_CORNERS[et]
    X = nodes[conn].numpy()
    for k, e in enumerate(elements.EDGES[et]):
        if len(e) == 3:
            assert np.allclose(X[:, e[2]], 0.5 * (X[:, e[0]] + X[:, e[1]]), atol=1e-14)
    if et == "quad9":
        assert np.allclose(X[:, 8], X[:, :nc].mean(axis=1), atol=1e-14)
    return p


@pytest.mark.parametrize("et", ETS)
def test_ccw_repair_of_a_mirrored_mesh(et):
    """Mirroring x -> -x turns every element clockwise; the repair must restore det J > 0 everywhere and keep
    mid-side nodes on their edges; elements already counter-clockwise are left alone."""
    nodes, conn = create_rectangle_mesh(et, 2.0, 1.0, 3, 2)
    mirrored = nodes * torch.tensor([-1.0, 1.0], dtype=torch.float64)
    assert (signed_corner_area(mirrored, conn, N_CORNERS[et]) < 0).all()
    fixed = check_and_fix_ccw(mirrored, conn, et)
    assert (signed_corner_area(mirrored, fixed, N_CORNERS[et]) > 0).all()
    assert torch.equal(fixed[:, 0], conn[:, 0])                             # node 0 stays first
    assert torch.equal(torch.sort(fixed, 1).values, torch.sort(conn, 1).values)  # same nodes per element
    _geometry_ok(mirrored, fixed, et)
    assert torch.equal(check_and_fix_ccw(nodes, conn, et), conn)


def test_filter_unused_nodes_keeps_elements_and_order():
    nodes, conn = create_rectangle_mesh("quad9", 1.0, 1.0, 2, 2)
    extra = torch.tensor([[5.0, 5.0], [6.0, 6.0]], dtype=torch.float64)
    padded = torch.cat([extra[:1], nodes, extra[1:]])
    n2, c2 = filter_unused_nodes(padded, conn + 1)
    assert torch.equal(n2, nodes) and torch.equal(c2, conn)


@pytest.mark.parametrize("et", ETS)
def test_npz_roundtrip_renumbered_padded_and_flipped(et, tmp_path):
    nodes, conn = create_unstructured_plate_hole(et, h=0.25, seed=1)
    path = tmp_path / "m.npz"
    save_mesh_npz(path, nodes, conn)
    n1, c1 = load_mesh_npz(path, et)
    assert torch.equal(n1, nodes) and torch.equal(c1, conn)
    # the same mesh written as a mesher might: shuffled node ids, unused points, some elements clockwise
    rng = np.random.default_rng(2)
    N = nodes.shape[0]
    perm = rng.permutation(N + 5)
    xy = np.zeros((N + 5, 2))
    xy[perm[:N]] = nodes.numpy()
    xy[perm[N:]] = rng.uniform(5, 6, (5, 2))                                # points no element uses
    cw = conn.numpy().copy()
    flip = rng.random(cw.shape[0]) < 0.5
    rev = {"tri3": [0, 2, 1], "tri6": [0, 2, 1, 5, 4, 3], "quad4": [0, 3, 2, 1],
           "quad8": [0, 3, 2, 1, 7, 6, 5, 4], "quad9": [0, 3, 2, 1, 7, 6, 5, 4, 8]}[et]
    cw[flip] = cw[flip][:, rev]
    np.savez(tmp_path / "raw.npz", xy=xy, conn=perm[cw])
    n2, c2 = load_mesh_npz(tmp_path / "raw.npz", et)
    assert n2.shape == nodes.shape
    # element by element the same coordinates in the same local order
    assert torch.allclose(n2[c2], nodes[conn], rtol=0, atol=0)
    _geometry_ok(n2, c2, et)


def test_npz_loader_rejects_bad_files(tmp_path):
    nodes, conn = create_rectangle_mesh("quad4", 1.0, 1.0, 2, 2)
    np.savez(tmp_path / "a.npz", xy=nodes.numpy(), conn=conn.numpy())
    with pytest.raises(ValueError):
        load_mesh_npz(tmp_path / "a.npz", "quad9")                         # wrong nodes per element
    np.savez(tmp_path / "b.npz", xy=nodes.numpy(), conn=conn.numpy() + 100)
    with pytest.raises(ValueError):
        load_mesh_npz(tmp_path / "b.npz", "quad4")                         # node ids out of range
    np.savez(tmp_path / "c.npz", points=nodes.numpy(), conn=conn.numpy())
    with pytest.raises(ValueError):
        load_mesh_npz(tmp_path / "c.npz", "quad4")                         # arrays missing
    np.savez(tmp_path / "d.npz", xy=np.array([{"a": 1}], dtype=object), conn=conn.numpy())
    with pytest.raises(ValueError):
        load_mesh_npz(tmp_path / "d.npz", "quad4")                         # nothing is unpickled


@pytest.mark.parametrize("et", ETS)
def test_unstructured_mesher_valid_and_irregular(et):
    nodes, conn = create_unstructured_plate_hole(et, h=0.08, seed=3)
    p = _geometry_ok(nodes, conn, et)
    r = torch.linalg.norm(nodes - 0.5, dim=1)
    n_hole = int((torch.abs(r - 0.2) < 1e-12).sum())                       # hole vertices on the circle,
    assert n_hole >= 8 and float(r.min()) > 0.2 * np.cos(np.pi / n_hole) - 1e-12   # mid-sides on its chords
    # the elements tile the unit plate minus the regular n_hole-gon inscribed in the hole
    assert float(p["w_detJ"].sum()) == pytest.approx(1 - 0.5 * n_hole * 0.04 * np.sin(2 * np.pi / n_hole), abs=1e-12)
    assert torch.equal(torch.unique(conn), torch.arange(nodes.shape[0]))
    val = np.bincount(conn[:, :N_CORNERS[et]].numpy().ravel())
    assert val.max() >= (6 if N_CORNERS[et] == 4 else 7)                    # irregular valence
    p2 = build_problem(dict(element_type=et, mesher="unstructured", geometry="plate_hole",
                            geom_params=dict(width=1.0, height=1.0, radius=0.2), h=0.08, mesh_seed=3,
                            load_mode="force", load_value=(0.1, 0.0)))
    assert torch.equal(p2["conn"], conn) and float(p2["f"].reshape(-1, 2)[:, 0].sum()) == pytest.approx(0.1)


def test_build_problem_from_a_mesh_file(tmp_path):
    nodes, conn = create_unstructured_plate_hole("tri6", h=0.2, seed=4)
    save_mesh_npz(tmp_path / "t6.npz", nodes, conn)
    p = build_problem(dict(element_type="tri6", mesher="npz", mesh_file=str(tmp_path / "t6.npz"),
                           load_mode="displacement", load_value=(0.05, None)))
    assert torch.equal(p["nodes"], nodes) and torch.equal(p["conn"], conn)
    assert bool(p["fixed"].reshape(-1, 2)[nodes[:, 0] < 1e-9].all())
    with pytest.raises(ValueError):
        build_problem(dict(element_type="tri6", mesher="gmsh"))
