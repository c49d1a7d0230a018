"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

import libsupermesh_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for line in text.splitlines():
            if line.lstrip().startswith(("static", "#", "//")):
                continue  # inline helpers and macros are not exported
            m = re.match(r"^\s*[\w\s\*]+?\b(libsupermesh_\w+)\s*\(", line)
            if m:
                names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_reference_symbols():
    names = declared_functions()
    # every function of the reference's include/libsupermesh-c.h:18-31
    for ref in ["libsupermesh_sort_intersection_finder_set_input",
                "libsupermesh_sort_intersection_finder_query_output",
                "libsupermesh_sort_intersection_finder_get_output",
                "libsupermesh_tree_intersection_finder_set_input",
                "libsupermesh_tree_intersection_finder_query_output",
                "libsupermesh_tree_intersection_finder_get_output",
                "libsupermesh_intersect_tris_real", "libsupermesh_intersect_tets_real",
                "libsupermesh_triangle_area", "libsupermesh_tetrahedron_volume",
                "libsupermesh_interval_size", "libsupermesh_intersect_intervals",
                "libsupermesh_intersect_tet_hex", "libsupermesh_intersect_tri_quad"]:
        assert ref in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(libsupermesh_amd.get_library())
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(libsupermesh_amd, "_PKG", str(tmp_path))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        libsupermesh_amd.get_library()
