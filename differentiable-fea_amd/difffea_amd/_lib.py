"""ctypes binding of libdfea_hip.so (the C ABI declared in include/dfea.h).

The library is loaded after ``import torch`` so that it binds to the HIP runtime torch already loaded
(both carry the soname ``libamdhip64.so.7``): one runtime, shared streams and device pointers.
Loading fails loudly when the library has not been built -- there is no fallback path.
"""
import ctypes
import os

import torch  # noqa: F401  (must be imported first: it loads the HIP runtime the library binds to)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libdfea_hip.so"
LIB_PATH = os.path.join(_HERE, LIB_NAME)

FP64, FP32 = 0, 1
PRECOND = {"jacobi": 0, "block_jacobi": 1}     # dfea_desc.precond
MAXCOL = 16
FIXED_X, FIXED_Y, NODE_MASK = 0x40000000, 0x80000000, 0x3FFFFFFF
HALO = 0x100          # element-kind flag of halo plans
SMALL = 0x400         # element-kind flag of small-chunk quad4 plans (64-element chunks)
TAIL = 0x800          # element-kind flag of tail-balanced linear partials plans (chunks of two sizes)
MATLIB = 0x200        # element-kind flag: material kernels with the whole material library
WIDE = 0x1000         # element-kind flag: PCG workgroups of small quadratic-element meshes (more groups)
PCG_WIDE_CHUNKS = 256  # quadratic meshes of at most this many chunks get DFEA_WIDE
MAT_NP = 8            # dfea_desc.mat_p slots
NOALLOC_SCRATCH_BYTES = 8 << 20   # DFEA_NOALLOC_SCRATCH_BYTES: static device scratch of dfea_apply_noalloc
VEC = dict(x=0, r=1, w=2, p=3, s=4, b=5, u=6, R=7, uacc=8, uD=9, x2=10, r2=11, w2=12, p2=13, s2=14)
NVEC = 15
PCG_VECS = ("x", "r", "w", "p", "s", "b", "x2", "r2", "w2", "p2", "s2")   # in the solve's precision
SC = dict(rz=0, pAp=1, alpha=2, beta=3, rr=4, bb=5, rtol=6, status=7, k=8, maxit=9, iters=10, norm=11, lf=12,
          step=13)
NSC = 24
PCG_RUNNING, PCG_CONVERGED, PCG_CURVATURE, PCG_MAXITER = 0, 1, 2, 3

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


class Desc(ctypes.Structure):
    """Mirror of ``dfea_desc`` (field order = declaration order in include/dfea.h)."""
    _fields_ = [
        ("chunks", _P), ("lnode", _P), ("lelem", _P), ("sh_node", _P), ("sh_off", _P), ("pslot", _P), ("hel", _P), ("hlelem", _P), ("prange", _P),
        ("inv_off", _P), ("inv_slot", _P),
        ("conn", _P), ("perm", _P), ("nodes", _P), ("dNdxi", _P), ("gw", _P), ("nfix", _P),
        ("dN64", _P), ("dN32", _P), ("w", _P), ("theta", _P), ("C64", _P), ("C32", _P),
        ("diag64", _P), ("diag32", _P), ("f", _P), ("part", _P), ("red", _P), ("sc", _P),
        ("vec", _P * NVEC), ("mat_p", _D * MAT_NP),
        ("n_chunks", ctypes.c_int32), ("n_lnodes", ctypes.c_int32), ("n_shared", ctypes.c_int32),
        ("n_part", ctypes.c_int32), ("n_el", ctypes.c_int32), ("n_nodes", ctypes.c_int32),
        ("nen", ctypes.c_int32), ("nq", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("n_finblk", ctypes.c_int32), ("n_vblk", ctypes.c_int32), ("n_halo", ctypes.c_int32),
        ("n_cols", ctypes.c_int32), ("precond", ctypes.c_int32), ("material", ctypes.c_int32),
        ("mat_scal", ctypes.c_int32), ("lstride", ctypes.c_int32),
    ]


CHUNK_INTS = 64     # sizeof(dfea_chunk) / 4
COL_OFF, HCOL_OFF = 12, 12 + MAXCOL + 1   # int offsets of col[] and hcol[] in a chunk record

_SIGNATURES = {
    "dfea_version": (_I, []),
    "dfea_last_error": (ctypes.c_char_p, []),
    "dfea_desc_size": (_I, []),
    "dfea_desc_offset": (_I, [_I]),
    "dfea_red_size": (_I, [_I, _I, _I]),
    "dfea_vblk": (_I, [_I]),
    "dfea_finblk": (_I, [_I]),
    "dfea_chunk_elems": (_I, [_I]),
    "dfea_chunk_nodes": (_I, []),
    "dfea_chunk_threads": (_I, [_I]),
    "dfea_geometry": (_I, [_P, _P, _I, _I, _P]),
    "dfea_moduli": (_I, [_P, _P, _I, _P, _I, _P, _I, _I, _P]),
    "dfea_reactions": (_I, [_P, _P, _I, _P, _I, _P, _P, _I, _P]),
    "dfea_fields": (_I, [_P, _P, _I, _P, _I, _P, _P, _P, _P, _I, _P]),
    "dfea_residual": (_I, [_P, _P, _I, _P, _I, _P, _P, _D, _I, _P]),
    "dfea_apply": (_I, [_P, _P, _I, _P, _I, _P, _P, _I, _I, _P]),
    "dfea_apply_noalloc": (_I, [_P, _I, _I, _I, _P, _P, _I, _I, _P]),
    "dfea_dot": (_I, [_P, _P, _I, _P, _I, _P, _P, _P]),
    "dfea_axpy": (_I, [_P, _P, _I, _I, _P, _P, _P, _P, _P, _I, _P]),
    "dfea_pcg_init": (_I, [_P, _P, _I, _P, _I, _I, _I, _P]),
    "dfea_pcg_start": (_I, [_P, _I, _D, _I, _I, _I, _P]),
    "dfea_apply_host": (_I, [_P, _P, _I, _I, _P, _P, _I, _I, _P]),
    "dfea_pcg_iterate": (_I, [_P, _P, _I, _I, _I, _I, _P]),
    "dfea_adjoint_grad": (_I, [_P, _P, _I, _P, _P, _P, _I, _P]),
    "dfea_pcg_graph_create": (_I, [_P, _P, _I, _I, _I, _I, ctypes.POINTER(_P)]),
    "dfea_graph_launch": (_I, [_P, _P]),
    "dfea_graph_destroy": (_I, [_P]),
}

_lib = None


class DfeaError(RuntimeError):
    pass


def load(path=None):
    """Load (once) and return the ctypes handle; raises if the library is missing."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("DFEA_LIB", LIB_PATH)
    if not os.path.exists(path):
        raise DfeaError(f"{LIB_NAME} is not built (looked for {path}); run __graft_entry__.build() first")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.dfea_desc_size() != ctypes.sizeof(Desc):
        raise DfeaError(f"dfea_desc layout mismatch: C {lib.dfea_desc_size()} B, ctypes {ctypes.sizeof(Desc)} B")
    for i, (fname, _) in enumerate(Desc._fields_):
        if lib.dfea_desc_offset(i) != getattr(Desc, fname).offset:
            raise DfeaError(f"dfea_desc field {fname}: C offset {lib.dfea_desc_offset(i)}, "
                            f"ctypes {getattr(Desc, fname).offset}")
    _lib = lib
    return lib


def check(rc, what):
    """Map a C return code to a Python exception (ValueError for invalid arguments)."""
    if rc == 0:
        return
    msg = _lib.dfea_last_error().decode() if _lib is not None else ""
    if rc == -1:
        raise ValueError(f"{what}: {msg}")
    raise DfeaError(f"{what} failed ({rc}): {msg}")


def exported_symbols():
    return sorted(_SIGNATURES)
