#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference's own CPU
path, compiled from its sources by oracle/Makefile.ref (oracle/_ref).

Each fixture holds outputs of the reference on inputs that are regenerated
from a seed (oracle.uniform_inputs), plus SHA-256 digests of those inputs and
of the operators (oracle.operators, FP64, cast to the precision), so a test
can rebuild the inputs and know they are the ones the reference saw.
Outputs: q*phi (the reference's output convention), the far-field q*phi,
proxysval (M, levels 2..maxlev) and fieldval after post-M2L (F), plus the
leaf particle lists and the interaction-list lengths.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

CASES = [
    # name, N, lmod, precisions
    ("u4096_d0", 4096, 0, ("s", "d")),
    ("u4096_d1", 4096, 1, ("s", "d")),
    ("u32768_d0", 32768, 0, ("s",)),
]


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    if not O.ref_available("s"):
        O.build_ref()
    U, K, _ = O.operators(4)
    for name, N, lmod, precs in CASES:
        pos, q = O.uniform_inputs(N, seed=0)
        for prec in precs:
            dt = O.dtype(prec)
            Uc, Kc = U.astype(dt), K.astype(dt)
            out = O.ref_run(pos.astype(dt), q.astype(dt), 4, Uc, Kc, prec=prec, lmod=lmod, dump=True)
            ml = out["info"]["maxlev"]
            s = O.levsta(ml)[2]
            path = os.path.join(HERE, f"{name}_{prec}.npz")
            np.savez_compressed(
                path, N=N, lmod=lmod, n=4, maxlev=ml, seed=0,
                input_digest=digest(pos.astype(dt), q.astype(dt)),
                operator_digest=digest(Uc, Kc),
                phi=out["phi"], phi_far=out["phi_far"],
                M=out["M"][s * 32:], F=out["F"][s * 64:],
                list=out["list"], sta=out["sta"], end=out["end"],
                iinter=out["iinter"], ineigh=out["ineigh"])
            print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
