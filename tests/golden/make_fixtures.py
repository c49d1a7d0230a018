"""Convert the reference's test meshes into tests/golden/meshes.npz.

Run in a container where the reference is mounted (it is not needed at test
time).  Reads the Triangle/TetGen files of src/tests/data in the reference
(the small meshes directly, the 0_01 / 0_05 meshes from the two tarballs the
reference's CMakeLists.txt:316-319 extracts) and stores, per mesh, ``<name>/pos``
(float64[nnodes, dim]) and ``<name>/enl`` (0-based int32[nelements, loc]).
Only data enters the repository.
"""
import os
import sys
import tarfile
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from libsupermesh_amd.triangle_io import read_mesh  # noqa: E402

REF_DATA = "/root/reference/src/tests/data"

SMALL = [("line.1", 1), ("line.2", 1), ("triangle.1", 2), ("triangle.2", 2), ("square.1", 2), ("square.2", 2),
         ("square.3", 2), ("square.4", 2), ("cube.1", 3), ("cube.2", 3),
         ("cube.3", 3), ("cube.4", 3), ("tet", 3)]
LARGE = [("2D_0_01.tar.bz2", [("triangle_0_01", 2), ("square_0_01", 2)]),
         ("3D_0_05.tar.bz2", [("pyramid_0_05", 3), ("cube_0_05", 3)])]


def main():
    out = {}
    for name, dim in SMALL:
        pos, enl = read_mesh(os.path.join(REF_DATA, name), dim)
        out[name + "/pos"], out[name + "/enl"] = pos, enl
    with tempfile.TemporaryDirectory() as tmp:
        for tarname, meshes in LARGE:
            with tarfile.open(os.path.join(REF_DATA, tarname), "r:bz2") as tf:
                members = [m for m in tf.getmembers()
                           if any(m.name in (n + ".node", n + ".ele") for n, _ in meshes)]
                tf.extractall(tmp, members=members)
            for name, dim in meshes:
                pos, enl = read_mesh(os.path.join(tmp, name), dim)
                out[name + "/pos"], out[name + "/enl"] = pos, enl
    np.savez_compressed(os.path.join(HERE, "meshes.npz"), **out)
    for k in sorted(out):
        print(k, out[k].shape, out[k].dtype)


if __name__ == "__main__":
    main()
