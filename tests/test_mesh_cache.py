"""The device-mesh cache of the hip operators (CPU: the device mesh is replaced by a stand-in object, so only
the cache's bookkeeping runs): one entry per live, unmodified set of mesh tensors; entries die with their
tensors; an in-place edit of nodes / fixed rebuilds the mesh."""
import contextlib
import gc

import torch

from helpers import unit_square


class _FakeMesh:
    built = 0

    def __init__(self, params, device):
        _FakeMesh.built += 1


def _patch(monkeypatch):
    from difffea_amd import operators
    monkeypatch.setattr(operators, "HipMesh", _FakeMesh)
    monkeypatch.setattr(torch.cuda, "device", lambda d: contextlib.nullcontext())
    operators._MESHES.clear()
    return operators


def test_cache_reuses_and_releases(monkeypatch):
    ops = _patch(monkeypatch)
    p = unit_square("quad4", 4)
    m1 = ops.get_mesh(p, "cuda:0")
    assert ops.get_mesh(p, "cuda:0") is m1 and ops.get_mesh(dict(p), "cuda:0") is m1
    assert ops.cached_meshes() == 1
    for i in range(50):                  # fresh problems, dropped after use: the cache does not grow
        q = unit_square("quad4", 3 + i % 4)
        ops.get_mesh(q, "cuda:0")
        del q
        gc.collect()
        assert ops.cached_meshes() == 1
    del p, m1
    gc.collect()
    assert ops.cached_meshes() == 0


def test_in_place_edit_rebuilds(monkeypatch):
    ops = _patch(monkeypatch)
    p = unit_square("quad4", 4)
    m1 = ops.get_mesh(p, "cuda:0")
    p["nodes"].mul_(2.0)                 # a shape update in place
    m2 = ops.get_mesh(p, "cuda:0")
    assert m2 is not m1 and ops.cached_meshes() == 1
    p["fixed"][0] = ~p["fixed"][0]       # new supports in place
    assert ops.get_mesh(p, "cuda:0") is not m2
    assert ops.get_mesh(p, "cuda:0") is ops.get_mesh(p, "cuda:0")
