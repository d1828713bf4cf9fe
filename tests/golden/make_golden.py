"""Writes the golden fixtures of tests/golden from the reference's own code.

Runs oracle/_ref/ref_golden (built by oracle/build_ref.sh from the reference
sources: its term kernels, Mobility, the Koren and van Leer limiters and
NonDimensionalNumbers, driven the way the reference's test_jacobian.cpp
drives them) into a scratch directory and packs what it wrote:

  ref_terms.npz   per case: CRS of A and J (rowptr/col/val), the fields
                  h, s, c on every row, A(h)h, the FaceFlux of every face
                  of every real cell and the SideRate
  ref_scalars.npz mobility table, limiter table, Dp table, the unlimited
                  face schemes' FaceValue and FaceWeights, the limited
                  ones' with the row's cell at L and at R, the evaporation
                  rate of C_Evaporation, the nozzle supply of Source_WM2D/1D
  manifest.json   the dimensionless parameters of every case
  ref_state_wm1d.state  a restart state file (field*.state) written by the
                  reference's C_RestartState for a 1-D film of 8 cells
  ref_neutral.npz, neutral/*.dat, neutral/*.vts  ASCII field files
                  written by the reference's OutputNeutral and VTK files
                  written by its OutputVTK, their inputs (the terms' own
                  velocity fields and the clock among them), and what its
                  ASCII restart read back from the ASCII files
  eq/eq_n*.dat.gz the reference's analytic equilibrium droplets
                  (cases/eq/eq_n1000, 2000, 4000.dat), copied as data

    python tests/golden/make_golden.py
"""
import gzip
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "ref_golden")
REF = os.environ.get("YUZEN_REFERENCE", "/root/reference")


def read_case(path):
    with open(path, "rb") as f:
        n, nnz, nnzj, nxp = np.fromfile(f, dtype=np.int32, count=4)
        rp = np.fromfile(f, dtype=np.int32, count=n + 1)
        col = np.fromfile(f, dtype=np.int32, count=nnz)
        av = np.fromfile(f, dtype=np.float64, count=nnz)
        rpj = np.fromfile(f, dtype=np.int32, count=n + 1)
        colj = np.fromfile(f, dtype=np.int32, count=nnzj)
        jv = np.fromfile(f, dtype=np.float64, count=nnzj)
        h, s, c, ax = (np.fromfile(f, dtype=np.float64, count=n) for _ in range(4))
        ndir = 2 if nxp == n else 4
        ff = np.fromfile(f, dtype=np.float64, count=n * ndir).reshape(n, ndir)
        side = np.fromfile(f, dtype=np.float64, count=n)
        vx = np.fromfile(f, dtype=np.float64, count=n)
        vy = np.fromfile(f, dtype=np.float64, count=n)
        en = np.fromfile(f, dtype=np.float64, count=n + 1)
    return dict(rp=rp, col=col, av=av, rpj=rpj, colj=colj, jv=jv, h=h, s=s, c=c, ax=ax, ff=ff, side=side,
                vx=vx, vy=vy, energy=en[:1], phi=en[1:])


def read_table(path):
    with open(path, "rb") as f:
        rows, cols = np.fromfile(f, dtype=np.int32, count=2)
        return np.fromfile(f, dtype=np.float64, count=rows * cols).reshape(rows, cols)


def main():
    if not os.path.exists(EXE):
        subprocess.check_call(["bash", os.path.join(ROOT, "oracle", "build_ref.sh")])
    with tempfile.TemporaryDirectory() as tmp:
        subprocess.check_call([EXE, tmp], stdout=subprocess.DEVNULL)
        man = json.load(open(os.path.join(tmp, "manifest.json")))
        arrays = {}
        for case in man["cases"]:
            d = read_case(os.path.join(tmp, case["name"] + ".bin"))
            for k, v in d.items():
                arrays[f"{case['name']}__{k}"] = v
        np.savez_compressed(os.path.join(HERE, "ref_terms.npz"), **arrays)
        np.savez_compressed(os.path.join(HERE, "ref_scalars.npz"),
                            mobility=read_table(os.path.join(tmp, "mobility.bin")),
                            limiter=read_table(os.path.join(tmp, "limiter.bin")),
                            limiter_vl=read_table(os.path.join(tmp, "limiter_vl.bin")),
                            dp=read_table(os.path.join(tmp, "dp.bin")),
                            schemes=read_table(os.path.join(tmp, "schemes.bin")),
                            limited_weights=read_table(os.path.join(tmp, "limited_weights.bin")),
                            evap=read_table(os.path.join(tmp, "evap.bin")),
                            source=read_table(os.path.join(tmp, "source.bin")))
        with open(os.path.join(HERE, "manifest.json"), "w") as f:
            json.dump(man, f, indent=1)
        # the ASCII field files OutputNeutral wrote, their inputs and what
        # the reference's ASCII restart read back from them
        neutral = {}
        os.makedirs(os.path.join(HERE, "neutral"), exist_ok=True)
        for d in ("1d", "2d"):
            with open(os.path.join(tmp, f"neutral_{d}_in.bin"), "rb") as f:
                n, nxp = np.fromfile(f, dtype=np.int32, count=2)
                for k in ("h", "s", "c", "vx", "vy"):
                    neutral[f"{d}__{k}"] = np.fromfile(f, dtype=np.float64, count=n)
                # the terms' own velocity fields (capillary, Marangoni,
                # vertical gravity, in-plane gravity, centrifugal) and the
                # clock the .vts carries (the time, the time scale)
                for t in range(5):
                    neutral[f"{d}__tx{t}"] = np.fromfile(f, dtype=np.float64, count=n)
                    neutral[f"{d}__ty{t}"] = np.fromfile(f, dtype=np.float64, count=n)
                neutral[f"{d}__clock"] = np.fromfile(f, dtype=np.float64, count=2)
                # FieldInfo: volume, solute, mean c, max/min h+s, max/min
                # h, max/min c
                neutral[f"{d}__info"] = np.fromfile(f, dtype=np.float64, count=9)
                assert neutral[f"{d}__info"].size == 9 and not f.read()
                neutral[f"{d}__nxp"] = np.array(nxp)
            with open(os.path.join(tmp, f"neutral_{d}_read.bin"), "rb") as f:
                n, ok = np.fromfile(f, dtype=np.int32, count=2)
                assert ok == 1
                for k in ("h", "s", "c"):
                    neutral[f"{d}__read_{k}"] = np.fromfile(f, dtype=np.float64, count=n)
            for kind in ("nd", "dim"):
                for ext in ("dat", "vts"):
                    src = os.path.join(tmp, f"neutral_{d}_{kind}00000.{ext}")
                    with open(src, "rb") as f, open(os.path.join(HERE, "neutral", f"{d}_{kind}.{ext}"), "wb") as g:
                        g.write(f.read())
        np.savez_compressed(os.path.join(HERE, "ref_neutral.npz"), **neutral)
        # the reference's analytic equilibrium droplets (cases/eq, ASCII
        # initial conditions at dx = 0.6, 0.3, 0.15 um), as data, gzipped
        os.makedirs(os.path.join(HERE, "eq"), exist_ok=True)
        for n in (1000, 2000, 4000):
            with open(os.path.join(REF, "cases", "eq", f"eq_n{n}.dat"), "rb") as f:
                raw = f.read()
            with gzip.GzipFile(os.path.join(HERE, "eq", f"eq_n{n}.dat.gz"), "wb", mtime=0) as g:
                g.write(raw)
        with open(os.path.join(tmp, "ref_state_wm1d.state"), "rb") as f:
            state = f.read()
        with open(os.path.join(HERE, "ref_state_wm1d.state"), "wb") as f:
            f.write(state)
    print("wrote", os.path.join(HERE, "ref_terms.npz"))


if __name__ == "__main__":
    sys.exit(main())
