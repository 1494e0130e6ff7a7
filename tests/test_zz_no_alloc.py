"""dfea_apply_noalloc (the product that allocates nothing: shared-node partials in a static device scratch)
against dfea_apply on the same descriptor: the same kernels and the same order of additions, so the outputs
are the same bits.  Two calls in a row, the second on a smaller mesh (the scratch keeps the first call's
partials; every partial is written before it is read), in FP64 and FP32, quad4 and quad9; sizes beyond the
scratch are refused before anything is launched."""
import ctypes

import pytest
import torch

from helpers import state, unit_square

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _system(et, n, precision):
    from difffea_amd.operators import make_system
    p = unit_square(et, n)
    u, v = state(p, seed=n, scale=(0.1 if et == "quad4" else 0.02) / n)     # no inverted Gauss point
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    s = make_system(u.to(DEV), th.to(DEV), p, precision)
    s.moduli(u.to(DEV))
    return s, v.to(DEV).to(s.vdt).contiguous()


def _noalloc(s, v):
    from difffea_amd import _lib
    m = s.mesh
    y = torch.full_like(v, float("nan"))
    rc = m.lib.dfea_apply_noalloc(ctypes.c_void_p(s.table.data_ptr()), m.plan.n_chunks, m.n_finblk, m.plan.n_part,
                                  ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(y.data_ptr()), m.kind, s.prec_id,
                                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    _lib.check(rc, "dfea_apply_noalloc")
    return y


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("et,sizes", [("quad4", (150, 37)), ("quad9", (70, 21))])
def test_noalloc_product_equals_dfea_apply(et, sizes, precision):
    for n in sizes:                          # the second call on a smaller mesh
        s, v = _system(et, n, precision)
        assert s.mesh.plan.n_part > 0
        ref = s.apply(v)
        y = _noalloc(s, v)
        torch.cuda.synchronize()
        assert bool(torch.isfinite(ref).all())
        assert torch.equal(y, ref)


def test_noalloc_refuses_more_partials_than_the_scratch_holds():
    from difffea_amd import _lib
    s, v = _system("quad4", 16, "fp64")
    m = s.mesh
    y = torch.zeros_like(v)
    too_many = _lib.NOALLOC_SCRATCH_BYTES // 16 + 1
    rc = m.lib.dfea_apply_noalloc(ctypes.c_void_p(s.table.data_ptr()), m.plan.n_chunks, m.n_finblk, too_many,
                                  ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(y.data_ptr()), m.kind, s.prec_id,
                                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == -1 and b"scratch" in m.lib.dfea_last_error()
    torch.cuda.synchronize()
    assert float(y.abs().max()) == 0.0
