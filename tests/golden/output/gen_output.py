"""Golden vectors of the field output (plot_f, visu) from the flang-built
reference harness, in this container only.

The input is a seeded random state of the 8x6x5 configuration (velocities,
pressure, level set and the eddy viscosity of that velocity field); the
harness runs plot_f and visu on it in a scratch directory, and the record
file arquivos/campos_00011 (bytes, record markers included) and the index
arquivos/visu.xdmf (text) are stored with the input state in
tests/golden/output/g8_frame.npz.

    python tests/golden/output/gen_output.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402


def main():
    p = O.Params()
    O.ref_build(p)
    s = O.random_state(p, seed=31, perturb=0.4)
    s.run("visco")
    s.run("contorno", 1)
    wd = tempfile.mkdtemp()
    os.makedirs(os.path.join(wd, "arquivos"))
    O.ref_run(s.copy(), "plot", workdir=wd)
    O.ref_run(s.copy(), "visu", workdir=wd)
    rec = open(os.path.join(wd, "arquivos", "campos_00011"), "rb").read()
    xdmf = open(os.path.join(wd, "arquivos", "visu.xdmf")).read()
    data = {"__scal_in": np.array([s.scal[k] for k in O.SCALARS]),
            "__params": np.array([p.nx, p.ny, p.nz, p.ccx0, p.ccxf, p.ccy0, p.ccyf, p.ccz0, p.cczf]),
            "__h": np.array([p.dx, p.dy, p.dz]),
            "record_file": np.frombuffer(rec, dtype=np.uint8),
            "xdmf": np.array(xdmf)}
    for name, _ in O.ARRAYS:
        data["in_" + name] = s.arr(name)
    np.savez_compressed(os.path.join(HERE, "g8_frame.npz"), **data)
    print("record", len(rec), "bytes; xdmf", len(xdmf), "chars; dt", s.scal["dt"])


if __name__ == "__main__":
    main()
