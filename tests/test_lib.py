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
    assert lib.dfea_red_size(3, 2, 1) == 6 * 3 + 2 + 3
