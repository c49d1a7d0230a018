"""The Fortran interface (libsupermesh_amd/fortran/libsupermesh.F90) driven like
the reference's serial tests, per pair through the C ABI on the GPU."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "libsupermesh_amd", "lib", "test_intersector_gpu")


def write_mesh(path, pos, enl):
    with open(path, "wb") as f:
        f.write(np.array([pos.shape[0], pos.shape[1], enl.shape[0], enl.shape[1]], dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(pos, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(enl + 1, dtype=np.int32).tobytes())


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran interface not built (no amdflang)")
@pytest.mark.parametrize("a,b,npairs,npieces,total", [
    ("triangle_0_01", "square_0_01", 114323, 116295, 0.5),
    ("cube.1", "cube.2", None, None, 1.0),
])
def test_fortran_driver(tmp_path, fixture_meshes, a, b, npairs, npieces, total):
    _run_driver(tmp_path, fixture_meshes, a, b, npairs, npieces, total, "pairs", closed=False)


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran interface not built (no amdflang)")
@pytest.mark.parametrize("a,b,total", [("triangle_0_01", "square_0_01", 0.5), ("cube.1", "cube.2", 1.0)])
def test_fortran_callback_driver(tmp_path, fixture_meshes, a, b, total):
    # supermesh(): every closed-overlap candidate reaches the callback
    _run_driver(tmp_path, fixture_meshes, a, b, None, None, total, "callback", closed=True)


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran interface not built (no amdflang)")
@pytest.mark.parametrize("a,b,total", [("square.1", "square.4", 1.0), ("cube.1", "cube.3", 0.5),
                                       ("line.1", "line.2", 1.0)])
def test_fortran_mixed_and_1d(tmp_path, fixture_meshes, a, b, total):
    # src/tests/test_element_intersector.F90 / test_interval_intersector.F90 totals
    pa, ea = fixture_meshes[a]
    pb, eb = fixture_meshes[b]
    fa, fb = tmp_path / "a.bin", tmp_path / "b.bin"
    write_mesh(fa, pa, ea)
    write_mesh(fb, pb, eb)
    r = subprocess.run([EXE, str(fa), str(fb), "mixed"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    dim, n, pieces, s = r.stdout.split()
    if int(dim) == 1:
        ip, ix = O.sort_finder(pa, ea, pb, eb)
    else:
        ip, ix = O.brute_force(pa, ea, pb, eb, closed=True)
    assert int(n) == len(ix)
    assert abs(float(s) - total) < 100.0 * max(np.finfo(float).eps, O.spacing(total))


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran interface not built (no amdflang)")
@pytest.mark.parametrize("a,b,total", [("square.1", "square.4", 1.0), ("cube.1", "cube.3", 0.5),
                                       ("square.3", "square.4", 1.0)])
def test_fortran_callback_driver_mixed(tmp_path, fixture_meshes, a, b, total):
    # parallel_supermesh takes any element type (dim x loc_a, dim x loc_b) and
    # clips with intersect_elements: quads / hexahedra reach the callback whole
    pa, ea = fixture_meshes[a]
    pb, eb = fixture_meshes[b]
    fa, fb = tmp_path / "a.bin", tmp_path / "b.bin"
    write_mesh(fa, pa, ea)
    write_mesh(fb, pb, eb)
    r = subprocess.run([EXE, str(fa), str(fb), "callback"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    dim, n, pieces, s = r.stdout.split()
    ip, ix = O.brute_force(pa, ea, pb, eb, closed=True)
    assert int(n) == len(ix)
    assert abs(float(s) - total) < 100.0 * max(np.finfo(float).eps, O.spacing(total))


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran interface not built (no amdflang)")
def test_fortran_callback_driver_reports_errors(tmp_path, fixture_meshes):
    # an out-of-range node id: supermesh() returns with the message readable
    # through libsupermesh_gpu_error() after its buffers are freed
    pa, ea = fixture_meshes["cube.1"]
    pb, eb = fixture_meshes["cube.2"]
    fa, fb = tmp_path / "a.bin", tmp_path / "b.bin"
    write_mesh(fa, pa, ea)
    write_mesh(fb, pb, eb)
    r = subprocess.run([EXE, str(fa), str(fb), "badnode"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split()[1] == "0"


def _run_driver(tmp_path, fixture_meshes, a, b, npairs, npieces, total, mode, closed):
    pa, ea = fixture_meshes[a]
    pb, eb = fixture_meshes[b]
    fa, fb = tmp_path / "a.bin", tmp_path / "b.bin"
    write_mesh(fa, pa, ea)
    write_mesh(fb, pb, eb)
    r = subprocess.run([EXE, str(fa), str(fb), mode], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    dim, n, pieces, s = r.stdout.split()
    ip, ix = O.brute_force(pa, ea, pb, eb, closed=closed)
    rows, cols = O.csr_to_pairs(ip, ix)
    cnt, meas = O.clip_pairs(pa, ea, pb, eb, rows, cols)
    assert int(n) == len(ix)
    assert int(pieces) == int(cnt.sum())
    if npairs is not None:
        assert int(n) == npairs and int(pieces) == npieces
    tol = 100.0 * max(np.finfo(float).eps, O.spacing(total))
    assert abs(float(s) - total) < tol


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran interface not built (no amdflang)")
@pytest.mark.parametrize("mode,a,b", [("front", "triangle_0_01", "square_0_01"), ("front", "cube.1", "cube.3"),
                                      ("finder", "square.1", "square.2"), ("finder", "line.1", "line.2")])
def test_fortran_front_and_default_finder(tmp_path, fixture_meshes, mode, a, b):
    # advancing_front_intersection_finder / intersection_finder through the
    # Fortran module: the CSR (1-based) checksum equals the oracle's
    pa, ea = fixture_meshes[a]
    pb, eb = fixture_meshes[b]
    fa, fb = tmp_path / "a.bin", tmp_path / "b.bin"
    write_mesh(fa, pa, ea)
    write_mesh(fb, pb, eb)
    r = subprocess.run([EXE, str(fa), str(fb), mode], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    dim, n, chk, _ = r.stdout.split()
    if mode == "front":
        ip, ix = O.advancing_front(pa, ea, pb, eb)
    elif pa.shape[1] == 1:
        ip, ix = O.sort_finder(pa, ea, pb, eb)
    else:
        ip, ix = O.brute_force(pa, ea, pb, eb, closed=True)
    ix = np.asarray(ix, dtype=np.int64) + 1
    ref = int(np.sum(ix * np.arange(1, len(ix) + 1)) + np.sum((np.asarray(ip) + 1) * np.arange(8, len(ip) + 8)))
    assert int(n) == len(ix) and int(chk) == ref


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran interface not built (no amdflang)")
@pytest.mark.parametrize("a,b", [("square.1", "square.2"), ("cube.1", "cube.2"), ("square.1", "square.4")])
def test_fortran_tree_types_and_stateful_finders(tmp_path, fixture_meshes, a, b):
    # tree_type / rtree_type allocate-query-deallocate and the stateful
    # *_intersection_finder_set_input / find / query_output / get_output /
    # reset (Intersection_Finder.F90:481-760), element by element against the
    # CSR finders inside the driver; the counts against the oracle here
    pa, ea = fixture_meshes[a]
    pb, eb = fixture_meshes[b]
    fa, fb = tmp_path / "a.bin", tmp_path / "b.bin"
    write_mesh(fa, pa, ea)
    write_mesh(fb, pb, eb)
    r = subprocess.run([EXE, str(fa), str(fb), "tree_api"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    dim, n_strict, n_closed, _ = r.stdout.split()
    assert int(n_strict) == len(O.brute_force(pa, ea, pb, eb)[1])
    assert int(n_closed) == len(O.brute_force(pa, ea, pb, eb, closed=True)[1])


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran interface not built (no amdflang)")
@pytest.mark.parametrize("a,b,total", [("square.1", "square.2", 1.0), ("square.4", "square.1", 1.0),
                                       ("square.3", "square.4", 1.0), ("cube.1", "cube.2", 1.0),
                                       ("cube.3", "cube.2", 0.5), ("cube.3", "cube.4", 0.5),
                                       ("line.1", "line.2", 1.0)])
def test_fortran_intersect_elements(tmp_path, fixture_meshes, a, b, total):
    # src/tests/test_element_intersector.F90: intersect_elements over every
    # pair, simplex_volume summed; the piece count against the oracle
    pa, ea = fixture_meshes[a]
    pb, eb = fixture_meshes[b]
    fa, fb = tmp_path / "a.bin", tmp_path / "b.bin"
    write_mesh(fa, pa, ea)
    write_mesh(fb, pb, eb)
    r = subprocess.run([EXE, str(fa), str(fb), "elements"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    dim, n, pieces, s = r.stdout.split()
    if int(dim) == 1:
        ip, ix = O.sort_finder(pa, ea, pb, eb)
    else:
        ip, ix = O.brute_force(pa, ea, pb, eb, closed=True)
    assert int(n) == len(ix)
    assert abs(float(s) - total) < 1e-12
