"""Writes tests/golden/basic_case.npz from the reference's mainTests/basic
input files (testbin/mainTests/basic/geombc.dat.1 and restart.0.1, the
1072-node / 4157-tet cylinder its own end-to-end tests run), read with the
package's PHASTA reader.  Data only: the mesh, its boundary codes and the
restart state.  Run in a container that holds the reference:

    python tests/golden/make_basic_case.py /root/reference
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from crimsonflowsolver_amd import phasta  # noqa: E402


def main(ref):
    d = os.path.join(ref, "testbin", "mainTests", "basic")
    g = phasta.read_geombc(os.path.join(d, "geombc.dat.1"))
    r = phasta.read_restart(os.path.join(d, "restart.0.1"))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basic_case.npz")
    np.savez_compressed(out, x=g["x"], ien=g["ien"], ienb=g["ienb"], iBCB=g["iBCB"], BCB=g["BCB"],
                        iBC=g["iBC"], BC=g["BC"], y=phasta.to_solver_order(r["y"]),
                        ac=phasta.to_solver_order(r["ac"]), step=np.int32(r.get("step", 0)))
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
