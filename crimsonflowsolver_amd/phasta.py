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
    """[(keyword, header ints, raw bytes without the trailing newline)] in
    file order (a keyword repeats once per element block), and the byte
    order ('<' or '>')."""
    data = open(path, "rb").read()
    pos, blocks = 0, []
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
        blocks.append((key, ints, raw[:-1] if nb else b""))
        pos += nb
    order = "<"
    magic = _find(blocks, "byteorder magic number")
    if magic is not None:
        raw = magic[2][:4]
        if np.frombuffer(raw, "<i4")[0] == MAGIC:
            order = "<"
        elif np.frombuffer(raw, ">i4")[0] == MAGIC:
            order = ">"
        else:
            raise ValueError(f"{path}: bad byte-order magic number")
    return blocks, order


def _find(blocks, key):
    """The first block named `key`, or None."""
    for b in blocks:
        if b[0] == key:
            return b
    return None


def _arr(block, order, dtype, shape):
    key, ints, raw = block
    a = np.frombuffer(raw, order + dtype)
    n = int(np.prod(shape))
    if a.size < n:
        raise ValueError(f"block {key!r}: {a.size} values, expected {n}")
    return np.asfortranarray(a[:n].reshape(shape, order="F"))


def read_geombc(path):
    """Mesh and boundary data of one part: x(numnp,3), ien(nel,4) 1-based,
    ienb(nelb,4), iBCB(nelb,2), BCB(nelb, nshlb, ndBCB), each with its element
    blocks concatenated in file order (readnblk.f90:98-408 reads one block
    per topology; the boundary blocks' connectivity, codes and values come
    in the same order and must agree in length), and iBC(nshg) and BC(nshg,
    ndofBC) expanded through the bc mapping array (zero rows where no code
    is set)."""
    b, o = read_blocks(path)
    numnp = _find(b, "number of nodes")[1][0]
    x = _arr(_find(b, "co-ordinates"), o, "f8", (numnp, 3))

    def blocks_of(prefix):
        return [blk for blk in b if blk[0].startswith(prefix)]

    ien = [_arr(blk, o, "i4", (blk[1][0], blk[1][1])) for blk in blocks_of("connectivity interior")]
    bcon = blocks_of("connectivity boundary")
    ienb = [_arr(blk, o, "i4", (blk[1][0], blk[1][1])) for blk in bcon]
    out = {"x": x, "numnp": numnp,
           "ien": np.asfortranarray(np.concatenate(ien)) if ien else np.zeros((0, 4), np.int32, order="F"),
           "ienb": np.asfortranarray(np.concatenate(ienb)) if ienb else np.zeros((0, 4), np.int32, order="F")}
    codes = blocks_of("nbc codes")
    vals = blocks_of("nbc values")
    if codes or vals:
        if len(codes) != len(bcon) or len(vals) != len(bcon):
            raise ValueError(f"{path}: {len(bcon)} boundary blocks but {len(codes)} code and {len(vals)} value "
                             f"blocks")
        ib, bv, ndbcb = [], [], None
        for con, cb, vb in zip(bcon, codes, vals):
            npro = con[1][0]
            if cb[1][0] != npro or vb[1][0] != npro:
                raise ValueError(f"{path}: boundary block of {npro} elements with {cb[1][0]} codes and "
                                 f"{vb[1][0]} value rows")
            ib.append(_arr(cb, o, "i4", (npro, 2)))
            nd = len(vb[2]) // 8 // max(npro, 1) // 3  # nshlb = 3 for a tet face
            if ndbcb is not None and nd != ndbcb:
                raise ValueError(f"{path}: boundary value blocks of {ndbcb} and {nd} values per face node")
            ndbcb = nd
            bv.append(_arr(vb, o, "f8", (npro, 3, nd)))
        out["iBCB"] = np.asfortranarray(np.concatenate(ib))
        out["BCB"] = np.asfortranarray(np.concatenate(bv))
    tags = _find(b, "node tags")
    if tags is not None:  # readnblk.f90:216-228 (bit 0: ring node of a deformable wall)
        out["nodetag"] = _arr(tags, o, "i4", (tags[1][0],)).ravel()
    mapping = _find(b, "bc mapping array")
    if mapping is not None:
        nshg = mapping[1][0]
        mp = _arr(mapping, o, "i4", (nshg,))
        cblk = _find(b, "bc codes array")
        nb = cblk[1][0]
        bcodes = _arr(cblk, o, "i4", (nb,))
        vblk = _find(b, "boundary condition array")
        ndof = vblk[1][0] // max(nb, 1)
        bcv = _arr(vblk, o, "f8", (nb, ndof)) if nb else np.zeros((0, ndof))
        iBC = np.zeros(nshg, np.int32)
        BC = np.zeros((nshg, ndof), order="F")
        has = mp > 0
        iBC[has] = bcodes[mp[has] - 1]
        BC[has] = bcv[mp[has] - 1]
        out["iBC"], out["BC"] = iBC, BC
    return out


def read_restart(path):
    """solution y(nshg,ndof) and 'time derivative of solution' ac(nshg,ndof),
    columns (p,u1,u2,u3,...) as PHASTA stores them, the step number, and
    when present the deformable wall's 'displacement' and 'displacement_ref'
    (nshg,3) (readnblk.f90:589-640)."""
    b, o = read_blocks(path)
    out = {}
    for key, name in (("solution", "y"), ("time derivative of solution", "ac"), ("displacement", "u"),
                      ("displacement_ref", "uref")):
        blk = _find(b, key)
        if blk is not None:
            n, ndof = blk[1][0], blk[1][1]
            out[name] = _arr(blk, o, "f8", (n, ndof))
            if len(blk[1]) > 2:
                out["step"] = blk[1][2]
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
