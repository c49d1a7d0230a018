"""Triangle / TetGen ``.node`` and ``.ele`` mesh files.

Follows the reference reader ``src/Read_Triangle.F90:92-225``:

* ``.node``: header ``nnodes dim nattrs nbm`` then ``id x y [z] [attrs] [bm]``
  per vertex, ids 1..nnodes in order;
* ``.ele``: header ``nelements loc nattrs`` then ``id n1 .. nloc [attrs]``,
  1-based vertex ids.

Positions come back as ``float64[nnodes, dim]`` (the C ABI layout) and the
element-node list as 0-based ``int32[nelements, loc]``.  Malformed files raise
``ValueError`` where the reference aborts (``libsupermesh_abort``).
"""
from __future__ import annotations

import numpy as np


def _tokens(path):
    with open(path, "r") as f:
        text = f.read()
    lines = [ln.split("#", 1)[0].split() for ln in text.splitlines()]
    return [ln for ln in lines if ln]


def read_node(path, dim):
    """Vertex coordinates, ``float64[nnodes, dim]`` (Read_Triangle.F90:92-159)."""
    lines = _tokens(path)
    nnodes, ldim, nattrs, nbm = (int(x) for x in lines[0][:4])
    if nnodes < 0:
        raise ValueError("Invalid number of vertices")
    if ldim != dim:
        raise ValueError("Invalid dimension")
    if nattrs < 0:
        raise ValueError("Invalid number of attributes")
    if nbm < 0 or nbm > 1:
        raise ValueError("Invalid number of boundary markers")
    pos = np.empty((nnodes, dim), dtype=np.float64)
    for i in range(nnodes):
        row = lines[1 + i]
        if int(row[0]) != i + 1:
            raise ValueError("Invalid vertex number")
        pos[i] = [float(x) for x in row[1:1 + dim]]
    return pos


def read_ele(path, nnodes=None):
    """Element-node list, 0-based ``int32[nelements, loc]`` (Read_Triangle.F90:168-225)."""
    lines = _tokens(path)
    nele, loc, nattrs = (int(x) for x in lines[0][:3])
    if nele < 0:
        raise ValueError("Invalid number of elements")
    if nattrs < 0:
        raise ValueError("Invalid number of attributes")
    enl = np.empty((nele, loc), dtype=np.int64)
    for i in range(nele):
        row = lines[1 + i]
        if int(row[0]) != i + 1:
            raise ValueError("Invalid element number")
        enl[i] = [int(x) for x in row[1:1 + loc]]
    if enl.size and enl.min() < 1:
        raise ValueError("Invalid node")
    if nnodes is not None and enl.size and enl.max() > nnodes:
        raise ValueError("Invalid node")
    return (enl - 1).astype(np.int32)


def read_mesh(basename, dim):
    """``(positions, enlist)`` from ``basename.node`` and ``basename.ele``."""
    pos = read_node(basename + ".node", dim)
    enl = read_ele(basename + ".ele", nnodes=pos.shape[0])
    return pos, enl


def write_mesh(basename, positions, enlist):
    """Write ``basename.node`` / ``basename.ele`` with 1-based ids and 17
    significant digits, so a round trip is exact."""
    positions = np.asarray(positions, dtype=np.float64)
    enlist = np.asarray(enlist)
    nn, dim = positions.shape
    ne, loc = enlist.shape
    with open(basename + ".node", "w") as f:
        f.write(f"{nn} {dim} 0 0\n")
        for i in range(nn):
            f.write(f"{i + 1} " + " ".join(f"{x:.16e}" for x in positions[i]) + "\n")
    with open(basename + ".ele", "w") as f:
        f.write(f"{ne} {loc} 0\n")
        for i in range(ne):
            f.write(f"{i + 1} " + " ".join(str(int(v) + 1) for v in enlist[i]) + "\n")
