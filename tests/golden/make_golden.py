"""Write the golden vectors under tests/golden/ from the reference engine.

Runs the reference's own MC3D/MC2D engine (compiled from /root/reference by
oracle/Makefile into oracle/_ref/libref{2,3}d.so) on every case in
cases.py at OpenMP thread counts 1 and 4, and stores the inputs with the
normalised outputs (ER, EI, EBR, EBI, DEBR, DEBI, HN, loss) as
<case>.npz.  Error cases store the status the reference reports
(1 + mcerror).  Usage:

    make -C oracle && python tests/golden/make_golden.py [case ...]

With case names, only those cases are written.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import cases  # noqa: E402
import pyoracle  # noqa: E402


def main() -> None:
    if not pyoracle.have_reference():
        raise SystemExit("oracle/_ref is not built: run `make -C oracle` with the reference present")
    only = set(sys.argv[1:])
    for name, build in cases.CASES.items():
        if only and name not in only:
            continue
        p = build()
        d = cases.to_arrays(p)
        for t in cases.THREADS:
            res = pyoracle.run_reference(p, t)
            for k in ("ER", "EI", "EBR", "EBI", "DEBR", "DEBI", "HN"):
                d[f"out_T{t}_{k}"] = getattr(res, k)
            d[f"out_T{t}_loss"] = np.asarray(res.loss)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print(f"{name}: Ne={p.Ne} Nb={p.Nb} N={p.Nphoton} loss={res.loss}")
    for name, build in cases.ERROR_CASES.items():
        if only and name not in only:
            continue
        p = build()
        d = cases.to_arrays(p)
        try:
            pyoracle.run_reference(p, 1)
            status = 0
        except pyoracle.OracleError as e:
            status = e.status
        d["out_status"] = np.asarray(status)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print(f"{name}: status {status}")


if __name__ == "__main__":
    main()
