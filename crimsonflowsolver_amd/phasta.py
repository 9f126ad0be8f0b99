"""PHASTA binary files (geombc.dat.<n>, restart.<step>.<n>): reader and
restart writer, written from the file layout (host side, numpy).

Layout (what cvSolverIO.cxx:127-370 reads and :388-520 writes): text header
lines ``# ...``; every field is a header line ``<keyword> : < nbytes >
i1 i2 ...`` followed, when nbytes > 0, by nbytes of binary data whose last
byte is a newline.  Integers are int32, reals float64, in the byte order
announced by the ``byteorder magic number`` block (the int 362436 in the
writer's order).  2-d arrays are column-major: ``co-ordinates : < . > numnp
nsd`` holds x(numnp,nsd); ``connectivity interior ... : < . > npro nshl ...``
holds ien(npro,nshl) with 1-based node numbers (readnblk.f90:98-408 reads
the same fields).
"""
from __future__ import annotations

import re

import numpy as np

MAGIC = 362436
_HDR = re.compile(r"(.*?):\s*<\s*(\d+)\s*>(.*)")


def read_blocks(path):
    """{keyword: (header ints, raw bytes without the trailing newline)} in
    file order, and the byte order ('<' or '>')."""
    data = open(path, "rb").read()
    pos, blocks = 0, {}
    while pos < len(data):
        nl = data.find(b"\n", pos)
        if nl < 0:
            break
        line = data[pos:nl].decode("latin1")
        pos = nl + 1
        if line.startswith("#") or not line.strip():
            continue
        m = _HDR.match(line)
        if not m:
            raise ValueError(f"{path}: malformed header line {line[:60]!r}")
        key, nb = m.group(1).strip(), int(m.group(2))
        ints = [int(v) for v in m.group(3).split()]
        raw = data[pos:pos + nb]
        if nb and raw[-1:] != b"\n":
            raise ValueError(f"{path}: block {key!r} does not end with a newline")
        blocks[key] = (ints, raw[:-1] if nb else b"")
        pos += nb
    order = "<"
    if "byteorder magic number" in blocks:
        raw = blocks["byteorder magic number"][1][:4]
        if np.frombuffer(raw, "<i4")[0] == MAGIC:
            order = "<"
        elif np.frombuffer(raw, ">i4")[0] == MAGIC:
            order = ">"
        else:
            raise ValueError(f"{path}: bad byte-order magic number")
    return blocks, order


def _arr(blocks, order, key, dtype, shape):
    ints, raw = blocks[key]
    a = np.frombuffer(raw, order + dtype)
    n = int(np.prod(shape))
    if a.size < n:
        raise ValueError(f"block {key!r}: {a.size} values, expected {n}")
    return np.asfortranarray(a[:n].reshape(shape, order="F"))


def read_geombc(path):
    """Mesh and boundary data of one part: x(numnp,3), ien(nel,4) 1-based
    (interior blocks concatenated), ienb(nelb,4), iBCB(nelb,2), BCB(nelb,
    nshlb, ndBCB), iBC(nshg) and BC(nshg, ndofBC) expanded through the bc
    mapping array (zero rows where no code is set)."""
    b, o = read_blocks(path)
    numnp = b["number of nodes"][0][0]
    x = _arr(b, o, "co-ordinates", "f8", (numnp, 3))
    ien = [_arr(b, o, k, "i4", (v[0][0], v[0][1])) for k, v in b.items() if k.startswith("connectivity interior")]
    ienb = [_arr(b, o, k, "i4", (v[0][0], v[0][1])) for k, v in b.items() if k.startswith("connectivity boundary")]
    out = {"x": x, "numnp": numnp,
           "ien": np.asfortranarray(np.concatenate(ien)) if ien else np.zeros((0, 4), np.int32, order="F"),
           "ienb": np.asfortranarray(np.concatenate(ienb)) if ienb else np.zeros((0, 4), np.int32, order="F")}
    codes = [(k, v) for k, v in b.items() if k.startswith("nbc codes")]
    vals = [(k, v) for k, v in b.items() if k.startswith("nbc values")]
    if codes:
        k, v = codes[0]
        out["iBCB"] = _arr(b, o, k, "i4", (v[0][0], 2))
    if vals:
        k, v = vals[0]
        npro = v[0][0]
        nbytes = len(v[1])
        ndbcb = nbytes // 8 // npro // 3  # nshlb = 3 for a tet face
        out["BCB"] = _arr(b, o, k, "f8", (npro, 3, ndbcb))
    if "bc mapping array" in b:
        nshg = b["bc mapping array"][0][0]
        mp = _arr(b, o, "bc mapping array", "i4", (nshg,))
        nb = b["bc codes array"][0][0]
        codes = _arr(b, o, "bc codes array", "i4", (nb,))
        nvals = b["boundary condition array"][0][0]
        ndof = nvals // max(nb, 1)
        bcv = _arr(b, o, "boundary condition array", "f8", (nb, ndof)) if nb else np.zeros((0, ndof))
        iBC = np.zeros(nshg, np.int32)
        BC = np.zeros((nshg, ndof), order="F")
        has = mp > 0
        iBC[has] = codes[mp[has] - 1]
        BC[has] = bcv[mp[has] - 1]
        out["iBC"], out["BC"] = iBC, BC
    return out


def read_restart(path):
    """solution y(nshg,ndof) and 'time derivative of solution' ac(nshg,ndof),
    columns (p,u1,u2,u3,...) as PHASTA stores them, and the step number."""
    b, o = read_blocks(path)
    out = {}
    for key, name in (("solution", "y"), ("time derivative of solution", "ac")):
        if key in b:
            n, ndof = b[key][0][0], b[key][0][1]
            out[name] = _arr(b, o, key, "f8", (n, ndof))
            if len(b[key][0]) > 2:
                out["step"] = b[key][0][2]
    return out


def _block(key, ints, payload=b""):
    head = f"{key} : < {len(payload) + 1 if payload else 0} > " + " ".join(str(i) for i in ints) + " \n"
    return head.encode("latin1") + (payload + b"\n" if payload else b"")


def write_restart(path, y, ac, step):
    """restart.<step>.<part> with 'solution' and 'time derivative of
    solution' (nshg,ndof) in the reader's layout, little endian."""
    y = np.asfortranarray(y, dtype="<f8")
    ac = np.asfortranarray(ac, dtype="<f8")
    n, ndof = y.shape
    parts = [b"# PHASTA Input File Version 2.0\n", f"# Byte Order Magic Number : {MAGIC} \n".encode(),
             _block("byteorder magic number", [1], np.array([MAGIC], "<i4").tobytes()),
             _block("number of modes", [n]), _block("number of variables", [ndof]),
             _block("solution", [n, ndof, step], y.tobytes(order="F")),
             _block("time derivative of solution", [n, ndof, step], ac.tobytes(order="F"))]
    with open(path, "wb") as f:
        f.write(b"".join(parts))


def to_solver_order(y_phasta):
    """PHASTA's (p,u1,u2,u3,...) columns -> the solver arrays' (u1,u2,u3,p),
    as readnblk.f90:541-542 permutes them."""
    y = np.asfortranarray(y_phasta)
    return np.asfortranarray(np.concatenate([y[:, 1:4], y[:, 0:1]], axis=1))
