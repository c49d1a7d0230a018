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
    pa, ea = fixture_meshes[a]
    pb, eb = fixture_meshes[b]
    fa, fb = tmp_path / "a.bin", tmp_path / "b.bin"
    write_mesh(fa, pa, ea)
    write_mesh(fb, pb, eb)
    r = subprocess.run([EXE, str(fa), str(fb)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    dim, n, pieces, s = r.stdout.split()
    ip, ix = O.brute_force(pa, ea, pb, eb)
    rows, cols = O.csr_to_pairs(ip, ix)
    cnt, meas = O.clip_pairs(pa, ea, pb, eb, rows, cols)
    assert int(n) == len(ix)
    assert int(pieces) == int(cnt.sum())
    if npairs is not None:
        assert int(n) == npairs and int(pieces) == npieces
    tol = 100.0 * max(np.finfo(float).eps, O.spacing(total))
    assert abs(float(s) - total) < tol
