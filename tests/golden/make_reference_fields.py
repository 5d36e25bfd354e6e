"""Write per-sweep field fixtures from the reference binary.

Runs the reference POREMAPS, compiled from its own sources by
``make -C oracle ref`` into ``oracle/_ref/POREMAPS`` (MPICH from the image),
on small seeded geometries and stores what it writes: the interior fields
after N sweeps, the neighbourhood codes, the per-voxel rank of the domain
decomposition and the log rows.  Only these outputs are committed
(``reference_fields.npz``, ``reference_fields.json``); the binary and its run
directories stay outside the repository's history.

    python tests/golden/make_reference_fields.py

Cases:
  box     13 x 9 x 11 (x, y, z) random medium, 35 % solid; solvers 1, 2, 3 x
          boundary_method 0..4 x N in {1, 7, 1101} sweeps, one rank.
  decomp  10 x 8 x 18 random medium; `mpiexec -n P*Q*R` with
          dom_decomposition P Q R for z-slabs (1 1 2, 1 1 3, 1 1 4 uneven,
          1 1 9 two-plane slabs), x / y splits (2 1 1, 1 2 1, 3 1 1 uneven)
          and blocks (2 2 1, 2 1 2, 2 2 2, 3 2 2), boundary_method 0, 1, 3,
          solver 2 after 7 sweeps; solvers 1 and 3 on 2 2 2; 1101 sweeps on
          1 1 2 (bc 0) and 2 2 2 (bc 1).
  channel the shipped 24 x 14 x 50 channel (porosity -1.0, eps 1e-6) under
          1, 2, 3 and 8 ranks: the log row at t = 3000, and the last row of
          the solve to eps (known answers for the decompositions' lagged
          halos and for solver 1 against solver 2).
  stop    the box medium solved to eps 1e-6 with solvers 1, 2, 3 (bc 0 and
          2): sweep count and last log row.

voxel_size is 1.0 in the field cases, so the velocity files hold the
velocities unscaled (src/output.cc:86 multiplies by the voxel size).
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "POREMAPS")
MPIEXEC = os.environ.get("MPIEXEC", "/opt/conda/bin/mpiexec")
CHANNEL = os.path.join(HERE, "rect_channel_24_14_50_vs_5e-05.raw")

BOX = (13, 9, 11)      # nx, ny, nz
DEC = (10, 8, 18)


def medium(shape_xyz, seed, solid=0.35):
    nx, ny, nz = shape_xyz
    rng = np.random.default_rng(seed)
    return (rng.random((nz, ny, nx)) < solid).astype(np.uint8)


def inp_text(geom, size, dims, bc, dx, max_iter, it_eval, it_write, solver, eps, porosity, dom, wout):
    return (f"dom_decomposition {dims[0]} {dims[1]} {dims[2]}\n"
            f"boundary_method {bc}\n"
            f"geometry_file_name {geom}\n"
            f"size_x_y_z  {size[0]} {size[1]} {size[2]}\n"
            f"voxel_size  {dx!r}\n"
            f"max_iter    {max_iter}\n"
            f"it_eval    {it_eval}\n"
            f"it_write   {it_write}\n"
            f"log_file_name run.log\n"
            f"solving_algorithm {solver}\n"
            f"eps {eps!r}\n"
            f"porosity {porosity!r}\n"
            f"dom_interest {' '.join(str(d) for d in dom)}\n"
            f"write_output {' '.join(str(w) for w in wout)}")


def run(raw, size, *, dims=(0, 0, 0), bc=0, dx=1.0, max_iter=7, it_eval=100, it_write=100, solver=2,
        eps=1e-6, porosity=-1.0, dom=(0, 0, 0, 0, 0, 0), wout=(1, 1, 1, 1), raw_path=None):
    """Run the reference in a scratch directory; returns its outputs."""
    nproc = max(1, dims[0] * dims[1] * dims[2])
    with tempfile.TemporaryDirectory(prefix="pmref_") as d:
        geom = "geom.raw"
        if raw_path is not None:
            shutil.copy(raw_path, os.path.join(d, geom))
        else:
            raw.tofile(os.path.join(d, geom))
        with open(os.path.join(d, "run.inp"), "w") as f:
            f.write(inp_text(geom, size, dims, bc, dx, max_iter, it_eval, it_write, solver, eps, porosity, dom,
                             wout))
        cmd = [MPIEXEC, "-n", str(nproc), REF_BIN, "run.inp"]
        r = subprocess.run(cmd, cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed ({cmd}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        nx, ny, nz = size
        out = {}
        for key, pre, dt in (("u", "velx_", np.float64), ("v", "vely_", np.float64), ("w", "velz_", np.float64),
                             ("p", "press_", np.float64), ("nhood", "voxel_neighborhood_", np.int32),
                             ("decomp", "domain_decomp_", np.int32)):
            fn = os.path.join(d, pre + geom)
            if os.path.exists(fn):
                out[key] = np.fromfile(fn, dtype=dt).reshape(nz, ny, nx)
        logf = os.path.join(d, "run.log")
        out["log"] = open(logf).read().splitlines() if os.path.exists(logf) else []
        last = [ln for ln in r.stdout.splitlines() if ln.startswith("Iteration:")]
        out["last_progress"] = last[-1] if last else ""
    return out


def main() -> int:
    if not os.path.exists(REF_BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    arrays = {}
    meta = {"box": {"size_xyz": BOX, "seed": 11, "solid": 0.35, "runs": []},
            "decomp": {"size_xyz": DEC, "seed": 12, "solid": 0.35, "runs": []},
            "channel": [], "stop": []}

    box = medium(BOX, 11)
    arrays["box_raw"] = box
    for bc in range(5):
        for solver in (1, 2, 3):
            for n in (1, 7, 1101):
                o = run(box, BOX, bc=bc, solver=solver, max_iter=n, wout=(1, 1, 1, 0))
                tag = f"box_s{solver}_bc{bc}_n{n}"
                for fld in "puvw":
                    arrays[f"{tag}_{fld}"] = o[fld]
                if solver == 2 and n == 1:
                    arrays[f"box_bc{bc}_nhood"] = o["nhood"]
                meta["box"]["runs"].append({"tag": tag, "solver": solver, "bc": bc, "sweeps": n, "log": o["log"]})
                print(tag, flush=True)

    dec = medium(DEC, 12)
    arrays["decomp_raw"] = dec
    layouts = [(1, 1, 2), (1, 1, 3), (1, 1, 4), (1, 1, 9), (2, 1, 1), (1, 2, 1), (3, 1, 1), (2, 2, 1), (2, 1, 2),
               (2, 2, 2), (3, 2, 2)]
    cases = [(dims, bc, 2, 7) for dims in layouts for bc in (0, 1, 3)]
    cases += [((2, 2, 2), 0, 1, 7), ((2, 2, 2), 0, 3, 7), ((1, 1, 2), 0, 2, 1101), ((2, 2, 2), 1, 2, 1101)]
    cases += [((1, 1, 1), bc, 2, 7) for bc in (0, 1, 3)]
    for dims, bc, solver, n in cases:
        o = run(dec, DEC, dims=dims, bc=bc, solver=solver, max_iter=n)
        tag = f"dec_{dims[0]}{dims[1]}{dims[2]}_s{solver}_bc{bc}_n{n}"
        for fld in "puvw":
            arrays[f"{tag}_{fld}"] = o[fld]
        arrays[f"{tag}_decomp"] = o["decomp"]
        meta["decomp"]["runs"].append({"tag": tag, "dims": dims, "solver": solver, "bc": bc, "sweeps": n,
                                       "log": o["log"]})
        print(tag, flush=True)

    for dims, solver, n in (((1, 1, 1), 2, 3001), ((1, 1, 1), 1, 3001), ((1, 1, 2), 2, 3001), ((2, 1, 1), 2, 3001),
                            ((1, 1, 3), 2, 3001), ((1, 2, 1), 2, 3001), ((2, 2, 2), 2, 3001),
                            ((1, 1, 1), 2, 200001), ((1, 1, 1), 1, 200001), ((1, 1, 2), 2, 200001),
                            ((2, 1, 1), 2, 200001)):
        o = run(None, (24, 14, 50), dims=dims, bc=0, dx=5e-05, max_iter=n, solver=solver, eps=1e-06, porosity=-1.0,
                wout=(0, 0, 0, 0), raw_path=CHANNEL)
        rows = [ln for ln in o["log"] if not ln.startswith("#")]
        row = [ln for ln in rows if ln.startswith("3000,")] if n == 3001 else rows[-1:]
        meta["channel"].append({"dims": dims, "solver": solver, "max_iter": n, "row": row[0] if row else None})
        print("channel", dims, solver, n, row, flush=True)

    for bc in (0, 2):
        for solver in (1, 2, 3):
            o = run(box, BOX, bc=bc, solver=solver, max_iter=200001, eps=1e-6, dx=1e-5, wout=(0, 0, 0, 0))
            rows = [ln for ln in o["log"] if not ln.startswith("#")]
            meta["stop"].append({"bc": bc, "solver": solver, "last_row": rows[-1] if rows else None,
                                 "last_progress": o["last_progress"]})
            print("stop", bc, solver, rows[-1] if rows else None, flush=True)

    np.savez_compressed(os.path.join(HERE, "reference_fields.npz"), **arrays)
    with open(os.path.join(HERE, "reference_fields.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("wrote", os.path.join(HERE, "reference_fields.npz"))
    return 0


if __name__ == "__main__":
    sys.exit(main())
