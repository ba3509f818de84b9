"""Writes tests/golden/su2_*.npz: small SU2 boxes (the text itself, as bytes)
with the arrays the reference computes for them -- Read_SU2_mesh
(cgeometry_supply.h:201-433), CFVM::create_FVM_from_geometry (cfvm.cpp:100-170)
and CFVM::define_matrix_structure (cfvm.cpp:74-98) -- from oracle/_ref/
libocsl_ref.so, which oracle/Makefile builds from the reference's sources.

    python tests/golden/make_su2_golden.py
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from opencslsolver_amd import mesh  # noqa: E402

CASES = {"hex3_s5": ("hex", 3, 5), "prism2_s7": ("prism", 2, 7), "tet2_s9": ("tet", 2, 9)}
H = 0.25


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    for name, (kind, n, seed) in CASES.items():
        text = mesh.box_su2(n, kind, h=H, seed=seed)
        with tempfile.NamedTemporaryFile("w", suffix=".su2", delete=False) as f:
            f.write(text)
            path = f.name
        try:
            arrays = oracle.ref_su2_arrays(path)
        finally:
            os.unlink(path)
        np.savez_compressed(os.path.join(here, f"su2_{name}.npz"),
                            su2_text=np.frombuffer(text.encode(), np.uint8), **arrays)
        print(name, {k: v.shape for k, v in arrays.items() if k in ("cell_volume", "face_left", "bc_left", "col")})


if __name__ == "__main__":
    main()
