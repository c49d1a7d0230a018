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


def read_node(path, dim, extras=False):
    """Vertex coordinates, ``float64[nnodes, dim]`` (Read_Triangle.F90:92-159).

    With ``extras`` also the optional outputs of the reference reader:
    ``(positions, attributes float64[nnodes, nattrs], boundary_markers
    int64[nnodes, nbm])``."""
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
    if len(lines) < 1 + nnodes:
        raise ValueError("Too few vertex lines")
    pos = np.empty((nnodes, dim), dtype=np.float64)
    attrs = np.empty((nnodes, nattrs), dtype=np.float64)
    bms = np.empty((nnodes, nbm), dtype=np.int64)
    for i in range(nnodes):
        row = lines[1 + i]
        if len(row) < 1 + dim + nattrs + nbm:
            raise ValueError("Too few values on a vertex line")
        if int(row[0]) != i + 1:
            raise ValueError("Invalid vertex number")
        pos[i] = [float(x) for x in row[1:1 + dim]]
        attrs[i] = [float(x) for x in row[1 + dim:1 + dim + nattrs]]
        bms[i] = [int(x) for x in row[1 + dim + nattrs:1 + dim + nattrs + nbm]]
    if extras:
        return pos, attrs, bms
    return pos


def read_ele(path, nnodes=None, extras=False):
    """Element-node list, 0-based ``int32[nelements, loc]`` (Read_Triangle.F90:168-225).

    With ``extras`` also the element attributes, ``float64[nelements, nattrs]``."""
    lines = _tokens(path)
    nele, loc, nattrs = (int(x) for x in lines[0][:3])
    if nele < 0:
        raise ValueError("Invalid number of elements")
    if nattrs < 0:
        raise ValueError("Invalid number of attributes")
    if len(lines) < 1 + nele:
        raise ValueError("Too few element lines")
    enl = np.empty((nele, loc), dtype=np.int64)
    attrs = np.empty((nele, nattrs), dtype=np.float64)
    for i in range(nele):
        row = lines[1 + i]
        if len(row) < 1 + loc + nattrs:
            raise ValueError("Too few values on an element line")
        if int(row[0]) != i + 1:
            raise ValueError("Invalid element number")
        enl[i] = [int(x) for x in row[1:1 + loc]]
        attrs[i] = [float(x) for x in row[1 + loc:1 + loc + nattrs]]
        if enl[i].min() < 1 or (nnodes is not None and enl[i].max() > nnodes):
            raise ValueError("Invalid node")
    out = (enl - 1).astype(np.int32)
    if extras:
        return out, attrs
    return out


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
