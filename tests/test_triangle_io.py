"""Triangle / TetGen .node / .ele reader (Read_Triangle.F90:92-225) on files
written by the test itself: attributes, boundary markers, 1-based ids, and
the reference's abort cases as ValueError."""
import numpy as np
import pytest

from libsupermesh_amd import meshgen, triangle_io


def _write(path, text):
    path.write_text(text)
    return str(path)


def test_node_with_attributes_and_markers(tmp_path):
    f = _write(tmp_path / "m.node",
               "4 2 2 1\n"
               "1 0.0 0.0 1.5 -2.0 1\n"
               "2 1.0 0.0 2.5 -3.0 0\n"
               "3 0.0 1.0 3.5 -4.0 1\n"
               "4 0.25 0.75 4.5 -5.0 0\n")
    pos, attrs, bm = triangle_io.read_node(f, 2, extras=True)
    assert pos.dtype == np.float64 and pos.shape == (4, 2)
    assert np.array_equal(pos, [[0, 0], [1, 0], [0, 1], [0.25, 0.75]])
    assert np.array_equal(attrs, [[1.5, -2.0], [2.5, -3.0], [3.5, -4.0], [4.5, -5.0]])
    assert np.array_equal(bm, [[1], [0], [1], [0]])
    # positions alone, as the finders take them
    assert np.array_equal(triangle_io.read_node(f, 2), pos)


def test_ele_one_based_with_attributes(tmp_path):
    f = _write(tmp_path / "m.ele",
               "2 3 1\n"
               "1 1 2 3 7.0\n"
               "2 2 4 3 8.0\n")
    enl, attrs = triangle_io.read_ele(f, nnodes=4, extras=True)
    assert enl.dtype == np.int32
    assert np.array_equal(enl, [[0, 1, 2], [1, 3, 2]])
    assert np.array_equal(attrs, [[7.0], [8.0]])


def test_tetgen_mesh_3d(tmp_path):
    _write(tmp_path / "t.node", "4 3 0 0\n1 0 0 0\n2 1 0 0\n3 0 1 0\n4 0 0 1\n")
    _write(tmp_path / "t.ele", "1 4 0\n1 1 2 3 4\n")
    pos, enl = triangle_io.read_mesh(str(tmp_path / "t"), 3)
    assert pos.shape == (4, 3) and np.array_equal(enl, [[0, 1, 2, 3]])


@pytest.mark.parametrize("text,msg", [
    ("1 3 0 0\n1 0 0\n", "Invalid dimension"),
    ("-1 2 0 0\n", "Invalid number of vertices"),
    ("1 2 -1 0\n1 0 0\n", "Invalid number of attributes"),
    ("1 2 0 2\n1 0 0 1 1\n", "Invalid number of boundary markers"),
    ("2 2 0 0\n1 0 0\n3 1 1\n", "Invalid vertex number"),
])
def test_node_errors(tmp_path, text, msg):
    with pytest.raises(ValueError, match=msg):
        triangle_io.read_node(_write(tmp_path / "bad.node", text), 2)


@pytest.mark.parametrize("text,msg", [
    ("-1 3 0\n", "Invalid number of elements"),
    ("1 3 -1\n1 1 2 3\n", "Invalid number of attributes"),
    ("2 3 0\n1 1 2 3\n1 1 2 3\n", "Invalid element number"),
    ("1 3 0\n1 0 2 3\n", "Invalid node"),
    ("1 3 0\n1 1 2 5\n", "Invalid node"),
])
def test_ele_errors(tmp_path, text, msg):
    with pytest.raises(ValueError, match=msg):
        triangle_io.read_ele(_write(tmp_path / "bad.ele", text), nnodes=4)


@pytest.mark.parametrize("dim", [2, 3])
def test_round_trip_is_exact(tmp_path, dim):
    pos, enl = meshgen.split_grid(dim, 5, 7)
    base = str(tmp_path / "grid")
    triangle_io.write_mesh(base, pos, enl)
    p2, e2 = triangle_io.read_mesh(base, dim)
    assert np.array_equal(p2, pos) and np.array_equal(e2, enl)
