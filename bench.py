"""Benchmark of the POREMAPS Stokes sweep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2]
                    [--mode jacobi|sor3|reference] [--precision f32dc|f64] [--omega w] [--omega-p wp]
                    [--tte] [--no-cpu]
    python bench.py --config 3 [--count 64] [--compare]      (batch)
    python bench.py --config 4 [--count 24] [--rate 4] [--compare]  (serving)

A step is one it_eval interval of the solver (100 sweeps of the bulk sweep,
with the defect-correction fold and residual of that interval) over the whole
synthetic field, fields resident in HBM.  value = GLUP/s = nx*ny*nz*sweeps/s
(dense count); fluid_glups counts the fluid voxels only.  The default solver
is the FP32+FP64-refined 3-colour SOR (relaxation 1.8 / 6.0): the red-black
SOR of the target in a race-free colouring, and the fastest refined solver to
eps on the default field (time_to_eps reports the refined Jacobi beside it).

Workloads (synthetic, generated from a seed):
  config 0: the reference's own CPU case, a tube 54x54x50 (r 20, bc 0, dx 2e-5);
            with --mode reference it times the exact-order FP64 sweep
  config 1: overlapping-sphere pack 256^3, phi 0.35, r 6-12, bc 2, dx 1e-5
  config 2: thresholded Gaussian random field 512^3, corr. length 8, phi 0.20,
            bc 1, dx 5e-6 (default: the largest single-GPU configuration)
  config 3: 64 media solved to eps in one batch -> solves/h (whole job timed)
  config 4: Poisson stream of tensor requests -> solves/h, p50/p99 latency

roofline: the dominant kernel group ("sweep") timed with HIP events on the
library's own stream over the timed region; achieved = algorithmic bytes per
launch, 8*S+1 B per voxel of the interior (read and write p, u, v, w in the
swept precision S = 4 or 8 B, read a 1-byte code), over the average launch
duration; peak = 8 TB/s HBM.  achieved_with_residual_GBps adds the 16 B of
FP32 residual the defect-correction sweep also reads.  traffic = the
measured HBM bytes per launch (rocprofv3 read requests by size + WRITE_SIZE,
calibrated on a device copy of known size) from the committed counter summary
named in traffic_source, when the run uses its workload and kernel form.
time_to_eps (default on; --no-tte skips it): eps 1e-6 from fields resident in
HBM to the stop rule, for the plain FP64 Jacobi solver, the FP32+FP64-refined
Jacobi and 3-colour SOR, and the FP64 solves at the same relaxations (k_zz of
each refined solve relative to its FP64 twin).
batch / serving (default run): the batch of 64 media (config 3) and the
serving stream of 24 tensor requests (config 4) with their defaults,
summarised in the same line (solves/h; p50 / p99 latency); --no-batch skips
them.
cpu_baseline: the reference POREMAPS itself (oracle/_ref/POREMAPS, built from
the reference's sources by `make -C oracle ref`) under mpiexec on this
host's cores (at most 16), on the same field: the rate of the difference of a
short and a longer run (start-up and geometry read cancel); the CPU oracle's
threaded block sweep when the binary or mpiexec is missing.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GLUP/s + time-to-eps (FP32-refined vs FP64 plain); solves/hr, p99 latency"
PEAK_HBM_GBS = 8000.0


def workload(cfg: int):
    from poremaps_amd import geometry as geo

    if cfg == 0:
        spec = dict(name="tube_54", n=54, bc=0, dx=2e-5,
                    desc="tube 54x54x50, radius 20, bc 0 (periodic), the reference's own CPU case")
        return geo.config0(), spec
    cache = f"/tmp/poremaps_bench_cfg{cfg}.npy"
    if cfg == 1:
        spec = dict(name="sphere_pack_256", n=256, bc=2, dx=1e-5,
                    desc="overlapping-sphere pack 256^3, phi 0.35, r U[6,12], seed 1, bc 2 (no-slip x/y)")
    elif cfg == 2:
        spec = dict(name="grf_512", n=512, bc=1, dx=5e-6,
                    desc="Gaussian random field 512^3, corr. length 8, phi 0.20, seed 1, bc 1 (slip x/y)")
    else:
        raise SystemExit(f"unknown config {cfg}")
    if os.path.exists(cache):
        g = np.load(cache)
    else:
        t0 = time.time()
        if cfg == 1:
            g = geo.sphere_pack(256, 0.35, 6, 12, seed=1)
        else:
            g = geo.gaussian_field(512, 8.0, 0.20, seed=1)
        np.save(cache, g)
        print(f"[bench] generated {spec['name']} in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    return g, spec


def cpu_baseline(g, bc, budget_s=20.0):
    """Reference-order FP64 sweeps (solving_algorithm 2) of the CPU oracle on
    this host: the interior cut into z-slabs, one thread each, the way
    `mpirun -np P` with dom_decomposition 1 1 P runs the reference, on all
    host cores (at most 16, the GPU box's share); the rate is the difference
    of a 1-sweep and an n-sweep run, so that copying the state into the slabs
    does not count.  The one-thread rate of the serial sweep is kept beside."""
    import oracle as O

    G = O.geom_halo(g, bc)
    code = O.eval_geometry(G)
    st = [a.copy() for a in O.init_fields(G, bc)]
    O.pins(G, st[0])
    n = g.size
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    P = max(1, min(cores, 16))

    def run(k):
        t0 = time.perf_counter()
        O.sweeps_ref_slabs(2, G, code, *st, bc, k, P)
        return time.perf_counter() - t0

    t1 = run(1)
    t3 = run(3)
    per = max((t3 - t1) / 2, 1e-3)
    k = int(max(4, min(400, (budget_s * 0.6) / per)))
    tk = run(k)
    rate = n * (k - 1) / (tk - t1) / 1e9
    # one thread, serial sweep
    done, t0 = 0, time.perf_counter()
    while True:
        O.sweep_ref(2, G, code, *st)
        for f in st:
            O.exchange(f, bc)
        O.pins(G, st[0])
        done += 1
        el = time.perf_counter() - t0
        if el > budget_s * 0.25 or done >= 50:
            break
    return {"value": round(rate, 4), "unit": "GLUP/s", "cores": P, "kind": "port",
            "sample": f"{k} minus 1 reference-order (solving_algorithm 2) FP64 sweeps of the same "
                      f"{g.shape[2]}x{g.shape[1]}x{g.shape[0]} field, CPU oracle in {P} z-slabs on {P} threads "
                      f"(as mpirun -np {P}, dom_decomposition 1 1 {P}), {tk - t1:.1f} s",
            "value_1core": round(n * done / el / 1e9, 5),
            "sample_1core": f"{done} serial sweeps on 1 thread, {el:.1f} s"}


def cpu_baseline_reference(g, spec, budget_s=25.0):
    """The reference binary on this host's cores: `mpiexec -n P POREMAPS`
    with dom_decomposition 0 0 0 (MPI_Dims_create picks the blocks), solver 2,
    no output files; two runs of k1 and k2 sweeps, rate = cells * (k2 - k1) /
    (t2 - t1).  None when the binary or mpiexec is missing."""
    import subprocess
    import tempfile

    exe = os.path.join(ROOT, "oracle", "_ref", "POREMAPS")
    mpiexec = os.environ.get("MPIEXEC", "/opt/conda/bin/mpiexec")
    if not (os.path.exists(exe) and os.path.exists(mpiexec)):
        return None
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    P = max(1, min(cores, 16))
    nz, ny, nx = g.shape
    with tempfile.TemporaryDirectory(prefix="pm_cpu_") as d:
        np.ascontiguousarray(g, dtype=np.uint8).tofile(os.path.join(d, "geom.raw"))

        def run(k):
            with open(os.path.join(d, "run.inp"), "w") as f:
                f.write(f"dom_decomposition 0 0 0\nboundary_method {spec['bc']}\ngeometry_file_name geom.raw\n"
                        f"size_x_y_z  {nx} {ny} {nz}\nvoxel_size  {spec['dx']!r}\nmax_iter    {k}\n"
                        f"it_eval    100\nit_write   100\nlog_file_name run.log\nsolving_algorithm 2\n"
                        f"eps 1e-06\nporosity -1.0\ndom_interest 0 0 0 0 0 0\nwrite_output 0 0 0 0")
            t0 = time.perf_counter()
            r = subprocess.run([mpiexec, "-n", str(P), exe, "run.inp"], cwd=d, capture_output=True, text=True,
                               timeout=600)
            el = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError(f"reference run failed: {r.stderr[-500:]}")
            print(f"[bench] cpu baseline: reference binary, {P} ranks, {k} sweeps in {el:.1f}s", file=sys.stderr,
                  flush=True)
            return el

        t1 = run(2)
        t2 = run(12)
        per = max((t2 - t1) / 10, 1e-4)
        k = int(max(12, min(2000, 2 + budget_s / per)))
        tk = run(k) if k > 12 else t2
    rate = g.size * (k - 2) / (tk - t1) / 1e9
    return {"value": round(rate, 4), "unit": "GLUP/s", "cores": P, "kind": "reference",
            "sample": f"reference POREMAPS (oracle/_ref, g++ -O3, MPICH) under mpiexec -n {P}, dom_decomposition "
                      f"0 0 0, solving_algorithm 2, FP64, the same {nx}x{ny}x{nz} field: {k} minus 2 sweeps, "
                      f"{tk - t1:.1f} s"}


def time_kernels(s, n):
    import ctypes as C
    import poremaps_amd as pm

    L = pm.lib()

    class KT(C.Structure):
        _fields_ = [("name", C.c_char * 48), ("launches", C.c_int), ("avg_ms", C.c_double),
                    ("bytes_per_launch", C.c_double)]

    out = (KT * 16)()
    cnt = C.c_int()
    L.pm_time_kernels.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    rc = L.pm_time_kernels(s._h, n, out, 16, C.byref(cnt))
    if rc != 0:
        raise RuntimeError(L.pm_last_error().decode())
    return {out[i].name.decode(): dict(launches=out[i].launches, avg_ms=out[i].avg_ms,
                                       bytes=out[i].bytes_per_launch) for i in range(cnt.value)}


def measured_traffic(spec, form, precision):
    """HBM bytes per sweep of the sweep kernels (one launch of the Jacobi
    record pass, three colour launches of the SOR), from the committed
    rocprofv3 counter summary (read requests by size + WRITE_SIZE, calibrated
    on a device copy of known size), when the run uses the workload and kernel
    form it was measured on; else None."""
    known = {"interleaved-lagged-records-skip2": ("jacobi_f32dc_grf512_hbm_traffic.md",
                                                  r"traffic_bytes_per_launch k_jac4_rec_skip2: (\d+)"),
             "interleaved-lagged-records": ("jacobi_f32dc_grf512_hbm_traffic.md",
                                            r"traffic_bytes_per_launch k_jac4_rec: (\d+)"),
             "interleaved-colour-split-skip2": ("sor3_f32dc_grf512_hbm_traffic.md",
                                                r"traffic_bytes_per_sweep k_colour4_split_skip2 x3: (\d+)"),
             "interleaved-colour-split": ("sor3_f32dc_grf512_hbm_traffic.md",
                                          r"traffic_bytes_per_sweep k_colour4_split x3: (\d+)"),
             "interleaved-colour-compact-packed": ("sor3_packed_f32dc_grf512_hbm_traffic.md",
                                                   r"traffic_bytes_per_sweep k_colour4_packed x3: (\d+)"),
             "interleaved-colour-compact": ("sor3_compact_f32dc_grf512_hbm_traffic.md",
                                            r"traffic_bytes_per_sweep k_colour4_compact x3: (\d+)"),
             "interleaved-compact-packed-skipall-occ6": ("jacobi_packed_f32dc_grf512_hbm_traffic.md",
                                                         r"traffic_bytes_per_launch k_jac4c_lagp: (\d+)"),
             "interleaved-compact-skipall": ("jacobi_compact_f32dc_grf512_hbm_traffic.md",
                                             r"traffic_bytes_per_launch k_jac4c_lag: (\d+)")}
    if spec["name"] != "grf_512" or form not in known or precision != "f32dc":
        return None
    path = os.path.join(ROOT, "profiles", known[form][0])
    if not os.path.exists(path):
        return None
    m = re.search(known[form][1], open(path).read())
    return (int(m.group(1)), "profiles/" + known[form][0]) if m else None


def default_tte(g, kw):
    """Time to eps 1e-6 on the field: the plain FP64 Jacobi solver, the
    FP32+FP64-refined Jacobi (2.5, 7) and 3-colour SOR (1.8, 6), and the FP64
    solves at the same relaxations; speedups against the plain solver and each
    refined k_zz relative to its FP64 twin."""
    import poremaps_amd as pm

    runs = [("fp64_plain_jacobi", dict(mode="jacobi", precision="f64", omega=1.0, omega_p=0.0)),
            ("f32dc_jacobi", dict(mode="jacobi", precision="f32dc", omega=2.5, omega_p=7.0)),
            ("f32dc_sor3", dict(mode="sor3", precision="f32dc", omega=1.8, omega_p=6.0)),
            ("fp64_jacobi_same_relaxation", dict(mode="jacobi", precision="f64", omega=2.5, omega_p=7.0)),
            ("fp64_sor3_same_relaxation", dict(mode="sor3", precision="f64", omega=1.8, omega_p=6.0))]
    out = {}
    for label, mk in runs:
        s = pm.Solver(g, **dict(kw, **mk, eps=1e-6))
        last = [time.time()]

        def progress(it, conv, secs, label=label):
            if time.time() - last[0] > 20:
                last[0] = time.time()
                print(f"[bench] tte {label}: sweep {it}, conv {conv:.3e}", file=sys.stderr, flush=True)

        res, _ = s.solve(progress=progress)
        s.close()
        out[label] = {"seconds": round(res.seconds, 3), "sweeps": res.sweeps, "wk33": res.last.wk33,
                      "glups": round(res.glups, 2), "omega": mk["omega"], "omega_p": mk["omega_p"]}
        print(f"[bench] tte {label}: {out[label]}", file=sys.stderr, flush=True)
    plain = out["fp64_plain_jacobi"]["seconds"]
    for lab, twin in (("f32dc_jacobi", "fp64_jacobi_same_relaxation"), ("f32dc_sor3", "fp64_sor3_same_relaxation")):
        v = out[lab]
        v["speedup_vs_fp64_plain"] = round(plain / v["seconds"], 2)
        v["wk33_rel_vs_fp64_same_relaxation"] = (v["wk33"] - out[twin]["wk33"]) / out[twin]["wk33"]
        v["wk33_rel_vs_fp64_plain"] = (v["wk33"] - out["fp64_plain_jacobi"]["wk33"]) / out["fp64_plain_jacobi"]["wk33"]
    return out


def _solve_alone(raw, bc, dx, **kw):
    import poremaps_amd as pm

    t0 = time.perf_counter()
    last = [t0]

    def progress(it, conv, secs):
        # a line every 30 s keeps long baseline solves visibly alive
        if time.perf_counter() - last[0] > 30:
            last[0] = time.perf_counter()
            print(f"[bench] one-at-a-time {raw.shape}: sweep {it}, conv {conv:.3e}", file=sys.stderr, flush=True)

    with pm.Solver(raw, boundary_method=bc, voxel_size=dx, eps=1e-6, **kw) as s:
        res, _ = s.solve(progress=progress)
    return time.perf_counter() - t0, res


def bench_batch(a):
    """Config 3: 64 cubic media (sides log-uniform 64-192, half sphere packs,
    half Gaussian fields, porosity 0.15-0.45), bc 0, all solved to eps 1e-6
    in one batch.  A step is one it_eval interval of every running sample;
    the timed region is the whole job: admission (upload, geometry
    classification, first sweep) and every interval until the last sample
    converges."""
    import poremaps_amd as pm
    from poremaps_amd import geometry as geo

    t0 = time.time()
    media = geo.random_media(a.count, 64, 192, seed=3)
    print(f"[bench] generated {len(media)} media in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    dense = sum(m["raw"].size for m in media)
    with pm.Batch(precision=a.precision, omega=a.omega, omega_p=a.omega_p, it_eval=a.it_eval, eps=1e-6,
                  max_iter=a.max_iter) as b:
        sync = time.perf_counter()
        slots = [b.add(m["raw"], boundary_method=0, voxel_size=1e-5) for m in media]
        t_load = time.perf_counter() - sync
        steps = 0
        while True:
            n = b.step()
            steps += 1
            if steps % 50 == 0:
                print(f"[bench] interval {steps}: {n} running", file=sys.stderr, flush=True)
            if n == 0:
                break
        el = time.perf_counter() - sync
        st = [b.status(s) for s in slots]
    sweeps = [r.sweeps for _, r in st]
    lups = sum(m["raw"].size * r.sweeps for m, (_, r) in zip(media, st))
    line = {
        "metric": METRIC, "value": round(len(media) / el * 3600.0, 1), "unit": "solves/h", "n_gpus": 1,
        "steps": steps, "warmup": 0, "ms_per_step": round(el / steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if a.precision == "f32dc" else "f64",
        "data": "synthetic",
        "config": {"workload": f"batch of {len(media)} cubic media, sides log-uniform 64-192, half sphere "
                               f"packs (r 3-10) half Gaussian fields (corr. 4-12), phi U[0.15,0.45], bc 0, "
                               f"eps 1e-6, seed 3", "precision": a.precision, "omega": a.omega,
                   "omega_p": a.omega_p,
                   "it_eval": a.it_eval, "dense_voxels": dense},
        "seconds": round(el, 3), "load_seconds": round(t_load, 3),
        "aggregate_glups": round(lups / el / 1e9, 2),
        "sweeps": {"min": min(sweeps), "median": int(np.median(sweeps)), "max": max(sweeps)},
        "states": sorted({s for s, _ in st}),
    }
    if a.compare:
        for label, kw in (("one_at_a_time_same_mode", dict(mode="jacobi", precision=a.precision, omega=a.omega,
                                                           omega_p=a.omega_p)),
                          ("one_at_a_time_fp64_plain", dict(mode="jacobi", precision="f64", omega=1.0))):
            tot = 0.0
            for i, m in enumerate(media):
                dt, _ = _solve_alone(m["raw"], 0, 1e-5, it_eval=a.it_eval, **kw)
                tot += dt
                print(f"[bench] {label} {i}: {dt:.2f}s", file=sys.stderr, flush=True)
            line[label] = {"seconds": round(tot, 3), "solves_per_hour": round(len(media) / tot * 3600.0, 1),
                           "batch_speedup": round(tot / el, 2)}
    if a.accuracy:
        # each sample against its FP64 solve with the same relaxation pair
        # (omega 0: the library picks the same pair from the porosity), alone
        errs = []
        for i, (m, (_, r)) in enumerate(zip(media, st)):
            _, ref = _solve_alone(m["raw"], 0, 1e-5, it_eval=a.it_eval, mode="jacobi", precision="f64",
                                  omega=a.omega, omega_p=a.omega_p)
            errs.append(abs(r.last.wk33 / ref.last.wk33 - 1.0))
            print(f"[bench] fp64 same relaxation {i}: k rel err {errs[-1]:.2e}, sweeps {r.sweeps} / {ref.sweeps}",
                  file=sys.stderr, flush=True)
        line["k_rel_err_vs_fp64_same_relaxation"] = {"max": max(errs), "median": float(np.median(errs)),
                                                     "samples": len(errs)}
    return line


def bench_serve(a):
    """Config 4: Poisson arrivals (4 requests/s) of permeability-tensor
    requests, sides log-uniform 48-320, sphere packs / Gaussian fields as in
    the batch workload, bc uniform in {0, 1, 2}, eps 1e-6, three rotated
    solves each, continuous batching on one GPU.  Baseline: the same requests
    solved one at a time with the plain FP64 Jacobi solver."""
    import poremaps_amd as pm
    from poremaps_amd import geometry as geo

    rng = np.random.default_rng(4)
    t0 = time.time()
    media = geo.random_media(a.count, 48, 320, seed=4, axes=(0, 1, 2))
    print(f"[bench] generated {len(media)} media in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    arrivals = np.cumsum(rng.exponential(1.0 / a.rate, len(media)))
    arrivals -= arrivals[0]
    bcs = rng.integers(0, 3, len(media))
    reqs = [(m["raw"], int(bc), 1e-5, 1e-6, float(t)) for m, bc, t in zip(media, bcs, arrivals)]
    def progress(done, running, t):
        print(f"[bench] t={t:.1f}s completed {done}/{len(reqs)}, {running} solves running", file=sys.stderr,
              flush=True)

    small = pm.SMALL_CELLS if a.small_cells is None else a.small_cells
    resp, st = pm.serve(reqs, realtime=True, progress=progress, max_solves=a.max_solves, small_cells=small,
                        precision=a.precision, omega=a.omega, omega_p=a.omega_p, it_eval=a.it_eval,
                        max_iter=a.max_iter)
    print(f"[bench] served: {json.dumps(st.as_dict())}", file=sys.stderr, flush=True)
    line = {
        "metric": METRIC, "value": round(st.solves_per_hour, 1), "unit": "solves/h", "n_gpus": 1,
        "steps": st.intervals, "warmup": 0, "ms_per_step": round(st.wall_s / max(st.intervals, 1) * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if a.precision == "f32dc" else "f64", "data": "synthetic",
        "config": {"workload": f"{len(media)} tensor requests, Poisson {a.rate}/s, sides log-uniform 48-320, "
                               f"sphere packs / Gaussian fields, phi U[0.15,0.45], bc U{{0,1,2}}, eps 1e-6, "
                               f"seed 4", "precision": a.precision, "omega": a.omega, "omega_p": a.omega_p,
                   "it_eval": a.it_eval, "small_cells": small},
        "latency_p50_s": round(st.latency_p50_s, 3), "latency_p99_s": round(st.latency_p99_s, 3),
        "latency_max_s": round(st.latency_max_s, 3), "wall_s": round(st.wall_s, 3),
        "completed": st.completed, "max_running_solves": st.max_running, "solve_cap": a.max_solves,
        "offered_solves_per_hour": round(3 * a.rate * 3600.0, 1),
        "peak_device_GB": round(st.peak_bytes / 1e9, 2),
    }
    if a.compare:
        n_base = min(a.baseline_requests or len(reqs), len(reqs))
        tot = 0.0
        for i in range(n_base):
            raw, bc = reqs[i][0], reqs[i][1]
            for axes in ((0, 1, 2), (2, 0, 1), (1, 2, 0)):
                dt, _ = _solve_alone(np.ascontiguousarray(raw.transpose(axes)), bc, 1e-5, mode="jacobi",
                                     precision="f64", omega=1.0, it_eval=a.it_eval)
                tot += dt
            print(f"[bench] plain one-at-a-time request {i}: {tot:.1f}s so far", file=sys.stderr, flush=True)
        sph = 3 * n_base / tot * 3600.0
        line["one_at_a_time_fp64_plain"] = {"requests": n_base, "seconds": round(tot, 3),
                                            "solves_per_hour": round(sph, 1)}
        line["speedup_vs_plain"] = round(st.solves_per_hour / sph, 2)
    if a.accuracy:
        # responses against the FP64 tensor with the same relaxation pairs
        n_acc = min(a.baseline_requests or 8, len(reqs))
        diag, off = [], []
        for i in range(n_acc):
            raw, bc = reqs[i][0], reqs[i][1]
            k64, _ = pm.tensor(raw, boundary_method=bc, voxel_size=1e-5, precision="f64", omega=a.omega,
                               omega_p=a.omega_p, it_eval=a.it_eval, max_iter=a.max_iter)
            k = resp[i].tensor
            d = np.diag(k64)
            diag.append(float(np.max(np.abs(np.diag(k) / d - 1.0))))
            off.append(float(np.max(np.abs((k - k64) - np.diag(np.diag(k - k64)))) / np.max(np.abs(d))))
            print(f"[bench] fp64 tensor {i}: diagonal rel err {diag[-1]:.2e}, off-diagonal {off[-1]:.2e}",
                  file=sys.stderr, flush=True)
        line["k_err_vs_fp64_same_relaxation"] = {"requests": n_acc, "diagonal_rel_max": max(diag),
                                                 "off_diagonal_rel_to_diag_max": max(off)}
    return line


def batch_and_serving(a):
    """The batch (64 media) and serving (24 requests at 4/s) workloads with
    their defaults, summarised for the default bench line, which then carries
    the whole metric: GLUP/s and time to eps of one field, solves/h of the
    batch, solves/h and p50 / p99 latency of the serving stream."""
    import argparse as _ap

    out = {}
    b = _ap.Namespace(**vars(a))
    b.omega, b.omega_p, b.mode, b.compare, b.accuracy = 0.0, 0.0, "jacobi", False, False
    b.count = 64
    ln = bench_batch(b)
    out["batch"] = {"solves_per_hour": ln["value"], "seconds": ln["seconds"], "load_seconds": ln["load_seconds"],
                    "aggregate_glups": ln["aggregate_glups"], "sweeps": ln["sweeps"],
                    "workload": ln["config"]["workload"], "relaxation": "pair by porosity"}
    b.count = 24
    ln = bench_serve(b)
    out["serving"] = {"solves_per_hour": ln["value"], "latency_p50_s": ln["latency_p50_s"],
                      "latency_p99_s": ln["latency_p99_s"], "latency_max_s": ln["latency_max_s"],
                      "wall_s": ln["wall_s"], "completed": ln["completed"],
                      "offered_solves_per_hour": ln["offered_solves_per_hour"],
                      "peak_device_GB": ln["peak_device_GB"], "workload": ln["config"]["workload"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--mode", default="sor3",
                    help="the FP32+DC 3-colour SOR by default (the fastest refined solver to eps; jacobi, "
                         "reference)")
    ap.add_argument("--precision", default="f32dc")
    ap.add_argument("--omega", type=float, default=None,
                    help="velocity relaxation of the fast modes (default for one field: 1.8 SOR, 2.5 Jacobi; "
                         "batch and serving: 0 = the pair chosen per geometry from its porosity)")
    ap.add_argument("--omega-p", type=float, default=None,
                    help="pressure relaxation of the fast modes (default for one field: 6.0 SOR, 7.0 Jacobi; "
                         "0 with omega 0)")
    ap.add_argument("--it-eval", type=int, default=100)
    ap.add_argument("--tte", action="store_true", help="time-to-eps of the selected mode beside the plain solver")
    ap.add_argument("--no-tte", action="store_true", help="skip the default time-to-eps set")
    ap.add_argument("--accuracy", action="store_true",
                    help="with --tte: FP64 at the same relaxation and an eps 1e-9 reference solve; configs 3 "
                         "and 4: every batch sample / the first --baseline-requests (8) responses against the "
                         "FP64 solves with the same relaxation")
    ap.add_argument("--reference-eps", type=float, default=1e-9,
                    help="--accuracy: eps of the tightly converged reference solve (0: none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batch and serving summaries of the default run")
    ap.add_argument("--count", type=int, default=None, help="media in the batch / serving workloads")
    ap.add_argument("--rate", type=float, default=4.0, help="serving arrival rate, requests/s")
    ap.add_argument("--compare", action="store_true", help="also time the one-at-a-time baselines")
    ap.add_argument("--baseline-requests", type=int, default=0)
    ap.add_argument("--max-solves", type=int, default=0, help="serving: most solves on the GPU at once")
    ap.add_argument("--small-cells", type=int, default=None,
                    help="serving: requests of at most this many voxels run in their own batch (0: one batch)")
    ap.add_argument("--max-iter", type=int, default=400_000, help="sweep cap per solve (batch, serving)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build solves one geometry per GPU and has no "
              f"multi-GPU path (asked for --gpus {a.gpus})", file=sys.stderr)
        sys.exit(2)
    if a.omega is None:
        a.omega = 0.0 if a.config in (3, 4) else 1.8 if a.mode == "sor3" else 2.5
    if a.omega_p is None:
        a.omega_p = 0.0 if a.omega == 0.0 else 6.0 if a.mode == "sor3" else 7.0
    if a.config == 3:
        a.count = a.count or 64
        print(json.dumps(bench_batch(a)), flush=True)
        return None
    if a.config == 4:
        a.count = a.count or 24
        print(json.dumps(bench_serve(a)), flush=True)
        return None

    import poremaps_amd as pm

    g, spec = workload(a.config)
    if a.mode == "reference":
        # the exact-order FP64 sweep takes no precision or relaxation
        a.precision, a.omega, a.omega_p = "f64", 1.0, 0.0
    n = g.size
    kw = dict(boundary_method=spec["bc"], voxel_size=spec["dx"], mode=a.mode, precision=a.precision,
              omega=a.omega, omega_p=a.omega_p, it_eval=a.it_eval, eps=1e-6)
    print(f"[bench] {spec['desc']}: loading", file=sys.stderr, flush=True)
    s = pm.Solver(g, **kw)
    for _ in range(a.warmup):
        s.step(a.it_eval)
        s.permeability()
    t0 = time.perf_counter()
    kt = time_kernels(s, a.steps * a.it_eval)
    s.permeability()
    el = time.perf_counter() - t0
    sweeps = a.steps * a.it_eval
    glups = n * sweeps / el / 1e9
    k = s.permeability()
    choice, tuned = s.kernel_choice()
    s.close()
    group = "sweep" if "sweep" in kt else "reference_sweep"
    sw = kt[group]
    # algorithmic bytes per sweep: read + write p, u, v, w (S bytes each) and
    # a 1-byte code per interior voxel (the FP32 correction state has S = 4)
    S = 4 if a.precision == "f32dc" else 8
    algo = (8 * S + 1) * n
    achieved = algo / (sw["avg_ms"] * 1e-3) / 1e9
    print(f"[bench] {sweeps} sweeps in {el:.3f}s: {glups:.2f} GLUP/s; kernels {json.dumps(kt)}",
          file=sys.stderr, flush=True)
    traffic = measured_traffic(spec, choice, a.precision)
    line = {
        "metric": METRIC,
        "value": round(glups, 3),
        "unit": "GLUP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(el / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if a.precision == "f32dc" else "f64",
        "data": "synthetic",
        "fluid_glups": round(glups * float((g == 0).mean()), 3),
        "config": {"workload": spec["desc"], "nx": g.shape[2], "ny": g.shape[1], "nz": g.shape[0],
                   "mode": a.mode, "precision": a.precision, "omega": a.omega, "omega_p": a.omega_p,
                   "sweeps_per_step": a.it_eval, "porosity": float((g == 0).mean())},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic[0] if traffic else None,
                     "traffic_source": traffic[1] if traffic else None,
                     # the measured HBM bytes per launch over this run's
                     # launch time: the fraction of the peak the kernel really
                     # moves (the dense algorithmic count above charges every
                     # cell of the domain, while the passes touch the fluid
                     # cells and their neighbours only, so it can exceed 1)
                     "traffic_frac": (round(traffic[0] / (sw["avg_ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)
                                      if traffic else None),
                     "kernel": group, "avg_ms": round(sw["avg_ms"], 4),
                     "bytes_per_launch": algo, "bytes_per_voxel": 8 * S + 1,
                     "achieved_with_residual_GBps": round(sw["bytes"] / (sw["avg_ms"] * 1e-3) / 1e9, 1)},
        "kernels": {k2: {"avg_ms": round(v["avg_ms"], 4), "launches": v["launches"],
                         "GBps": round(v["bytes"] / (v["avg_ms"] * 1e-3) / 1e9, 1)} for k2, v in kt.items()},
        "sweep_kernel": {"form": choice, "autotuner_ms_per_sweep": {k: v for k, v in tuned.items()
                                                                    if k != "packed-far"},
                         **({"packed_far_entries": int(tuned["packed-far"])} if "packed-far" in tuned else {})},
    }
    if not (a.tte or a.accuracy or a.no_tte) and a.mode != "reference":
        line["time_to_eps"] = default_tte(g, kw)
    if not (a.tte or a.accuracy or a.no_tte or a.no_batch) and a.mode != "reference" and a.config == 2:
        line.update(batch_and_serving(a))
    if a.tte or a.accuracy:
        tte = {}
        runs = [("fp64_plain_jacobi", dict(mode="jacobi", precision="f64", omega=1.0, omega_p=0.0)),
                (f"{a.precision}_{a.mode}_w{a.omega}_wp{a.omega_p}",
                 dict(mode=a.mode, precision=a.precision, omega=a.omega, omega_p=a.omega_p))]
        if a.accuracy:
            # the FP64 solver with the same relaxation (the oracle's semantics,
            # bit-exact on small cases) and a tightly converged reference
            runs += [(f"fp64_{a.mode}_w{a.omega}_wp{a.omega_p}",
                      dict(mode=a.mode, precision="f64", omega=a.omega, omega_p=a.omega_p)),
                     ]
            if a.reference_eps > 0:
                runs.append((f"reference_eps{a.reference_eps:g}",
                             dict(mode=a.mode, precision=a.precision, omega=a.omega, omega_p=a.omega_p,
                                  eps=a.reference_eps)))
        for label, mk in runs:
            s = pm.Solver(g, **dict(kw, **mk))
            last = [time.time()]

            def progress(it, conv, secs, label=label):
                if time.time() - last[0] > 20:
                    last[0] = time.time()
                    print(f"[bench] tte {label}: sweep {it}, conv {conv:.3e}, {secs:.0f}s", file=sys.stderr,
                          flush=True)

            res, _ = s.solve(progress=progress)
            s.close()
            tte[label] = {"seconds": round(res.seconds, 3), "sweeps": res.sweeps,
                          "wk33": res.last.wk33, "glups": round(res.glups, 2)}
            print(f"[bench] tte {label}: {tte[label]}", file=sys.stderr, flush=True)
        if a.accuracy:
            # the refined solve against the FP64 solve with the same relaxation
            same = tte[f"fp64_{a.mode}_w{a.omega}_wp{a.omega_p}"]["wk33"]
            v = tte[f"{a.precision}_{a.mode}_w{a.omega}_wp{a.omega_p}"]
            v["wk33_rel_vs_fp64_same_relaxation"] = (v["wk33"] - same) / same
            if a.reference_eps > 0:
                key = f"reference_eps{a.reference_eps:g}"
                ref = tte[key]["wk33"]
                for v in tte.values():
                    v[f"wk33_rel_err_vs_eps{a.reference_eps:g}"] = (v["wk33"] - ref) / ref
        line["time_to_eps"] = tte
    if not a.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cb = None
        try:
            cb = cpu_baseline_reference(g, spec)
        except Exception as e:  # the binary failed: fall back to the oracle port
            print(f"[bench] reference binary unavailable ({e}); CPU oracle instead", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cb if cb is not None else cpu_baseline(g, spec["bc"])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
