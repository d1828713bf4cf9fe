"""Writes tests/golden/spinodal224_2000.npz: the oracle's run of the 2-D
spinodal dewetting film at its full size (224^2 periodic cells, side 0.64,
capillary + disjoining, adaptive dt from 1e-5, 2000 steps;
yuzen_amd.cases.spinodal_2d).

The oracle's linear solves run the device's structured multigrid
(OracleStep lin_mode 1), so that the device's Newton counts can be asserted
equal at every step: two multigrids leave different errors at the
Eisenstat-Walker tolerance, and a pass's correction is that error.

Stored per step: the return code, the Newton passes, the linear
iterations, the clock and the step in force after it, and the accepted
attempt's pass log (the correction the exit test took and the linear
tolerance of each pass); the interior film after steps 100, 500, 1000 and
2000. The GPU test advances the device step over the same 2000 steps and
checks it against these.

    python tests/golden/make_spinodal224.py [--steps 2000]

The oracle runs its row loops over OpenMP threads (OMP_NUM_THREADS); the
result does not depend on their number. The run takes about an hour.
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "spinodal224_2000.npz")
KEEP = (100, 500, 1000, 2000)
MAXP = 32


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--out", default=OUT)
    ap.add_argument("--lin-mode", type=int, default=1, help="0 the reference's algebraic multigrid, 1 the device's")
    args = ap.parse_args()
    from yuzen_amd import cases
    from oracle import binding as ob
    c, h, s, cc, dt, _ = cases.spinodal_2d(n=224)
    o = ob.OracleStep(c, h, s, cc, dt, lin_mode=args.lin_mode)
    n = args.steps
    rc = np.zeros(n, np.int32)
    nk = np.zeros(n, np.int32)
    lin = np.zeros(n, np.int64)
    clock = np.zeros((n, 3))
    pcorr = np.zeros((n, MAXP))
    peta = np.zeros((n, MAXP))
    films = {}
    t0 = time.time()
    for k in range(n):
        r = o.advance()
        rc[k], nk[k], lin[k] = r[0], r[1], r[2]
        clock[k] = o.clock()
        a, e, _ = o.pass_log()
        pcorr[k, :len(a)] = a
        peta[k, :len(e)] = e
        if r[0] != 0:
            raise SystemExit(f"step {k} failed (rc {r[0]})")
        if (k + 1) in KEEP:
            hh = o.h().reshape(cases.padded_shape(c))
            films[f"h{k + 1}"] = hh[2:-2, 2:-2].copy()
        if k % 20 == 0 or k == n - 1:
            print(f"step {k}: nk {nk[k]} lin {lin[k]} t {clock[k, 0]:.4e} dt {clock[k, 1]:.3e} "
                  f"({time.time() - t0:.0f} s)", flush=True)
    np.savez_compressed(args.out, rc=rc, nk=nk, lin=lin, clock=clock, pass_corr=pcorr, pass_eta=peta,
                        lin_mode=np.int32(args.lin_mode), **films)
    print(f"wrote {args.out}")


if __name__ == "__main__":
    main()
