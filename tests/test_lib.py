"""The C-ABI library loads and exports every entry point include/dfea.h declares (no GPU calls)."""
import ctypes
import os
import re

from difffea_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "dfea.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(dfea_\w+)\s*\(", src, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.exported_symbols())


def test_descriptor_layout_matches_ctypes_mirror():
    lib = _lib.load()
    assert lib.dfea_desc_size() == ctypes.sizeof(_lib.Desc)
    for i, (name, _) in enumerate(_lib.Desc._fields_):
        assert lib.dfea_desc_offset(i) == getattr(_lib.Desc, name).offset, name
    assert lib.dfea_desc_offset(len(_lib.Desc._fields_)) == -1


def test_size_queries():
    lib = _lib.load()
    assert lib.dfea_vblk(1024) == 1 and lib.dfea_vblk(1025) == 2
    assert [lib.dfea_chunk_elems(n) for n in (3, 4, 6, 8, 9)] == [256, 256, 64, 64, 64]
    assert [lib.dfea_chunk_threads(n) for n in (3, 4, 6, 8, 9)] == [256, 256, 64, 64, 64]
    assert lib.dfea_chunk_elems(5) == -1 and lib.dfea_chunk_threads(7) == -1
    assert lib.dfea_red_size(3, 2, 1) == 6 * 3 + 6 + 2 + 3


def test_missing_library_fails_loudly():
    """Without the HIP library the path raises (no CPU fallback): a fresh process pointed at a missing file."""
    import subprocess
    import sys
    code = ("import sys; sys.path[:0] = [sys.argv[1]]\n"
            "from difffea_amd import _lib\n"
            "try:\n    _lib.load()\nexcept _lib.DfeaError as e:\n    print('raised:', e)\n")
    env = dict(os.environ, DFEA_LIB=os.path.join(ROOT, "no_such_dir", "libdfea_hip.so"))
    r = subprocess.run([sys.executable, "-c", code, os.path.join(ROOT, "differentiable-fea_amd")], env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "raised:" in r.stdout and "is not built" in r.stdout


def test_noalloc_product_refuses_bad_sizes_before_any_launch():
    """The no-allocation product checks its sizes on the host and returns DFEA_EINVAL before touching the
    device (the table pointer is never dereferenced on the host)."""
    lib = _lib.load()
    fake = ctypes.c_void_p(256)
    too_many = _lib.NOALLOC_SCRATCH_BYTES // 16 + 1
    assert lib.dfea_apply_noalloc(fake, 1, 1, too_many, fake, fake, 4, _lib.FP64, None) == -1
    assert b"scratch" in lib.dfea_last_error()
    assert lib.dfea_apply_noalloc(fake, 1, 1, 2 * too_many, fake, fake, 4, _lib.FP32, None) == -1
    assert lib.dfea_apply_noalloc(fake, 1, 1, 10, None, fake, 4, _lib.FP64, None) == -1
    assert lib.dfea_apply_noalloc(fake, 1, 1, 10, fake, fake, 4, 7, None) == -1
