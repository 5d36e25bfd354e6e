"""The benchmark configurations' media on the GPU, checked against the oracle.

* Media shaped like the sphere-pack workload (bc 2) and the Gaussian-field
  workload (bc 1), small enough for the CPU oracle: every fast mode bit for bit
  per sweep, and whole solves (sweep count, fields, k) against the oracle's
  own solve to the stop rule (orc_solve_fast).
* Batched samples, tensor columns and served responses against the oracle's
  solves of the same (rotated) geometries -- not against other GPU solves.
* Reduced forms of the batch workload (16 media of mixed kind and size in one
  batch) and of the serving workload (a Poisson stream of tensor requests,
  bc 0 / 1 / 2).
* Full size: the 256^3 sphere pack and the 512^3 Gaussian field, FP32 +
  FP64-refined k_zz against the FP64 solve at the same relaxation (bar 1e-5).
* The reference-order Gauss-Seidel (solving_algorithm 2) to eps 1e-6 against
  the fast modes at eps 1e-6 (bar 1e-4 on k_zz), on small media of both kinds
  and on the 256^3 sphere pack itself.

Reductions are compensated parallel sums on the GPU and serial sums in the
oracle, so k is compared to 1e-12 relative; fields and sweep counts exactly.
"""
import functools

import numpy as np
import pytest

import oracle as O
import poremaps_amd as pm
from poremaps_amd import geometry as geo
from test_gpu_parity import nonc, oracle_dc, oracle_fast

pytestmark = pytest.mark.gpu
INTERIOR = (slice(2, -2),) * 3
ROTATIONS = ((0, 1, 2), (2, 0, 1), (1, 2, 0))


def say(*a):
    print("[case]", *a, flush=True)


def c1_like(n=24, seed=5):
    return geo.sphere_pack(n, 0.35, 2, 4, seed=seed)


def c2_like(n=24, seed=5):
    return geo.gaussian_field(n, 3.0, 0.25, seed=seed)


MEDIA = {"sphere_pack_bc2": (c1_like, 2), "gaussian_field_bc1": (c2_like, 1)}


def pair_by_porosity(raw):
    # the library's choice at omega = 0 (pm_relaxation): (0.6, 30) below
    # porosity 0.20, else (2.5, 7)
    return (0.6, 30.0) if float((raw == 0).mean()) < 0.20 else (2.5, 7.0)


def oracle_k(raw, bc, dx, mode="jacobi", precision="f32dc", pair=None, eps=1e-6):
    om, omp = pair or pair_by_porosity(raw)
    t, _, st, G = O.solve_fast(raw, bc, mode, precision, om, omp, eps)
    k = O.permeability(G, st[1], st[2], st[3], st[0], (0,) * 6, dx, O.porosity(G))
    return t, k, st


def oracle_tensor(raw, bc, dx):
    """The three rotated solves of the tensor procedure (README.md:217-229),
    scattered to k[velocity axis, driving axis]."""
    k = np.zeros((3, 3))
    sweeps = []
    for rot in ROTATIONS:
        t, kk, _ = oracle_k(np.ascontiguousarray(raw.transpose(rot)), bc, dx)
        sweeps.append(t)
        wk13, wk23, wk33 = kk[3], kk[4], kk[5]
        if rot == (0, 1, 2):
            k[0, 2], k[1, 2], k[2, 2] = wk13, wk23, wk33
        elif rot == (2, 0, 1):
            k[1, 0], k[2, 0], k[0, 0] = wk13, wk23, wk33
        else:
            k[2, 1], k[0, 1], k[1, 1] = wk13, wk23, wk33
    return k, sweeps


def close(a, b, rel=1e-12):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.all(np.abs(a - b) <= rel * np.max(np.abs(b)))


# ---------------------------------------------------------------------------
# per sweep and whole solves on the workloads' kinds of media
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("medium", sorted(MEDIA))
@pytest.mark.parametrize("mode,precision", [("jacobi", "f64"), ("sor3", "f64"), ("jacobi", "f32dc"),
                                            ("sor3", "f32dc")])
def test_workload_media_sweeps_bit_exact(medium, mode, precision):
    make, bc = MEDIA[medium]
    raw = make()
    say("workload medium per sweep", medium, mode, precision)
    if precision == "f64":
        ref = oracle_fast(raw, bc, mode, 2.0, 30, 7.0)
        kw = {}
    else:
        ref = oracle_dc(raw, bc, mode, 2.0, 31, 10, 7.0)
        kw = {"it_eval": 10}
    with pm.Solver(raw, boundary_method=bc, mode=mode, precision=precision, omega=2.0, omega_p=7.0, **kw) as s:
        s.step(30 if precision == "f64" else 31)
        m = nonc(ref[0])
        for name, a, b in zip("puvw", s.state_halo(), ref):
            assert np.array_equal(a[m], b[m]), f"{name} differs"


@pytest.mark.parametrize("medium", sorted(MEDIA))
@pytest.mark.parametrize("mode,precision", [("jacobi", "f32dc"), ("sor3", "f32dc"), ("jacobi", "f64")])
def test_workload_media_solve_matches_oracle_solve(medium, mode, precision):
    make, bc = MEDIA[medium]
    raw = make(20, 7)
    pair = (1.8, 6.0) if mode == "sor3" else (2.5, 7.0)
    say("workload medium solve", medium, mode, precision)
    t, k, st = oracle_k(raw, bc, 1e-5, mode, precision, pair)
    with pm.Solver(raw, boundary_method=bc, voxel_size=1e-5, mode=mode, precision=precision, omega=pair[0],
                   omega_p=pair[1], eps=1e-6) as s:
        res, _ = s.solve()
        velz, press = s.field("velz"), s.field("press")
    assert res.sweeps == t
    assert np.array_equal(velz, st[3][INTERIOR]) and np.array_equal(press, st[0][INTERIOR])
    assert close([res.last.wk13, res.last.wk23, res.last.wk33], k[3:6])


@pytest.mark.parametrize("medium", sorted(MEDIA))
def test_workload_media_sor_compact_solve_matches_oracle_solve(medium, monkeypatch):
    """FP32+DC 3-colour SOR on the compact state (fluid cells and periodic
    halo copies only): the whole solve equals the oracle's, sweep count,
    fields and k."""
    make, bc = MEDIA[medium]
    raw = make(20, 7)
    pair = (1.8, 6.0)
    say("workload medium solve, compact SOR", medium)
    t, k, st = oracle_k(raw, bc, 1e-5, "sor3", "f32dc", pair)
    monkeypatch.setenv("POREMAPS_SOR", "compact")
    with pm.Solver(raw, boundary_method=bc, voxel_size=1e-5, mode="sor3", precision="f32dc", omega=pair[0],
                   omega_p=pair[1], eps=1e-6) as s:
        assert s.kernel_choice()[0] == "interleaved-colour-compact-packed"
        res, _ = s.solve()
        velz, press = s.field("velz"), s.field("press")
    assert res.sweeps == t
    assert np.array_equal(velz, st[3][INTERIOR]) and np.array_equal(press, st[0][INTERIOR])
    assert close([res.last.wk13, res.last.wk23, res.last.wk33], k[3:6])


# ---------------------------------------------------------------------------
# batch, tensor and serving against the oracle
# ---------------------------------------------------------------------------
def batch_media():
    return [(c1_like(18, 1), 0), (c2_like(16, 2), 1), (geo.sphere_pack((14, 18, 16), 0.3, 2, 3, seed=4), 2),
            (geo.gaussian_field((20, 12, 14), 2.5, 0.18, seed=6), 0)]


def test_batch_samples_match_oracle_solves():
    media = batch_media()
    say("batch vs oracle", [m.shape for m, _ in media])
    with pm.Batch(precision="f32dc", omega=pm.AUTO) as b:
        slots = [b.add(raw, boundary_method=bc, voxel_size=1e-5) for raw, bc in media]
        while b.step() > 0:
            pass
        for slot, (raw, bc) in zip(slots, media):
            state, res = b.status(slot)
            t, k, st = oracle_k(raw, bc, 1e-5)
            assert state == "converged"
            assert res.sweeps == t
            assert np.array_equal(b.field(slot, "velz"), st[3][INTERIOR])
            assert np.array_equal(b.field(slot, "press"), st[0][INTERIOR])
            assert close(res.last.wk33, k[5])


def test_tensor_matches_oracle_rotated_solves():
    raw = geo.sphere_pack((18, 14, 22), 0.55, 2, 4, seed=3, axes=(0, 1, 2))
    say("tensor vs oracle", raw.shape)
    k, res = pm.tensor(raw, boundary_method=0, voxel_size=1e-5)
    want, sweeps = oracle_tensor(raw, 0, 1e-5)
    assert [r.sweeps for r in res] == sweeps
    assert close(k, want)
    # a mis-scattered column would swap a diagonal with an off-diagonal entry
    assert np.all(np.diag(k) > 10 * np.abs(k[~np.eye(3, dtype=bool)]).max())


def poisson_stream(count, rate, seed, sides=(12, 20)):
    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(sides[0]), np.log(sides[1])))))
        phi = rng.uniform(0.3, 0.45)
        if i % 2 == 0:
            raw = geo.sphere_pack(n, phi, 1.5, 3, seed=100 + i, axes=(0, 1, 2))
        else:
            raw = geo.gaussian_field(n, 3.0, phi, seed=100 + i, axes=(0, 1, 2))
        reqs.append((raw, int(rng.integers(0, 3)), 1e-5, 1e-6, t))
        t += rng.exponential(1.0 / rate)
    return reqs


def test_serving_stream_matches_oracle_tensors():
    """A reduced serving workload: six tensor requests arriving as a Poisson
    stream, boundary_method uniform in {0, 1, 2}, sphere packs and Gaussian
    fields, served with continuous batching; every response equals the
    oracle's tensor of its geometry."""
    reqs = poisson_stream(6, 20.0, 11)
    say("serving vs oracle", [r[0].shape for r in reqs], [r[1] for r in reqs])
    resp, st = pm.serve(reqs, realtime=True)
    assert st.completed == len(reqs)
    for r, q in zip(resp, reqs):
        want, sweeps = oracle_tensor(q[0], q[1], q[2])
        assert list(r.sweeps) == sweeps
        assert close(r.tensor, want)


def test_batch_workload_reduced():
    """The batch workload reduced to 16 cubes (sides log-uniform 12-24, half
    sphere packs, half Gaussian fields, porosity 0.15-0.45, bc 0) in one
    batch: every sample converges, bit-equal to the oracle's solve."""
    media = geo.random_media(16, 12, 24, seed=8)
    say("batch workload reduced", [m["raw"].shape for m in media])
    with pm.Batch(precision="f32dc", omega=pm.AUTO) as b:
        slots = [b.add(m["raw"], boundary_method=0, voxel_size=1e-5) for m in media]
        while b.step() > 0:
            pass
        for slot, m in zip(slots, media):
            state, res = b.status(slot)
            t, k, st = oracle_k(m["raw"], 0, 1e-5)
            assert state == "converged", (m["raw"].shape, state)
            assert res.sweeps == t
            assert np.array_equal(b.field(slot, "velz"), st[3][INTERIOR])


# ---------------------------------------------------------------------------
# full size: FP32 + FP64-refined k against FP64 at the same relaxation
# ---------------------------------------------------------------------------
def _refined_vs_fp64(raw, bc, dx, mode, pair):
    out = {}
    for prec in ("f32dc", "f64"):
        with pm.Solver(raw, boundary_method=bc, voxel_size=dx, mode=mode, precision=prec, omega=pair[0],
                       omega_p=pair[1], eps=1e-6) as s:
            res, _ = s.solve()
            out[prec] = res
        say(" ", prec, res.sweeps, res.last.wk33, f"{res.seconds:.2f}s")
    return out


@pytest.mark.parametrize("mode,pair", [("jacobi", (2.5, 7.0)), ("sor3", (1.8, 6.0))])
def test_sphere_pack_256_refined_k_matches_fp64(mode, pair):
    raw = geo.sphere_pack(256, 0.35, 6, 12, seed=1)
    say("sphere pack 256^3 bc 2", mode, pair)
    r = _refined_vs_fp64(raw, 2, 1e-5, mode, pair)
    assert r["f32dc"].sweeps == r["f64"].sweeps
    assert abs(r["f32dc"].last.wk33 / r["f64"].last.wk33 - 1) < 1e-5


@functools.lru_cache(maxsize=1)
def _grf512():
    return geo.gaussian_field(512, 8.0, 0.20, seed=1)


@pytest.mark.parametrize("mode", ["sor3", "jacobi"])
def test_gaussian_field_512_packed_links_equal_link_arrays(mode, monkeypatch):
    """The Gaussian-field workload at full size (512^3, bc 1): 300 sweeps (three
    intervals, folds and residuals included) with the passes reading packed
    link records (the default; for the SOR also with windows of one value, so
    that most entries read the link arrays instead, and with those windows
    without the keep setting, where the records are given up for the link
    arrays) leave the same velocity and pressure fields, bit for bit, as with
    the passes reading the link arrays."""
    raw = _grf512()
    pair = (1.8, 6.0) if mode == "sor3" else (2.5, 7.0)
    if mode == "jacobi":
        monkeypatch.setenv("POREMAPS_JACOBI", "compact")
    full = {"POREMAPS_SOR_LINKS" if mode == "sor3" else "POREMAPS_JAC_LINKS": "full"}
    forms = [("packed", {}), ("full", full)]
    if mode == "sor3":
        forms.append(("span0", {"POREMAPS_PACK_SPAN": "0", "POREMAPS_PACK_KEEP": "1"}))
        # the same windows without the keep setting: 71 % of the entries fall
        # back, so the records are given up for the link arrays
        forms.append(("span0-fallback", {"POREMAPS_PACK_SPAN": "0"}))
    out = {}
    for name, envs in forms:
        with monkeypatch.context() as m:
            for k, v in envs.items():
                m.setenv(k, v)
            with pm.Solver(raw, boundary_method=1, voxel_size=5e-6, mode=mode, precision="f32dc", omega=pair[0],
                           omega_p=pair[1], it_eval=100) as s:
                kname, tried = s.kernel_choice()
                s.step(300)
                out[name] = (kname, tried.get("packed-far"), s.field("velz"), s.field("press"))
        say("gaussian field 512^3", mode, name, out[name][0], "packed-far", out[name][1])
    assert "-packed" in out["packed"][0] and "-packed" not in out["full"][0]
    for name in out:
        assert np.array_equal(out[name][2], out["packed"][2]), name
        assert np.array_equal(out[name][3], out["packed"][3]), name
    if mode == "sor3":
        assert out["span0"][1] > 1_000_000
        assert "-packed" not in out["span0-fallback"][0] and out["span0-fallback"][1] > 1_000_000


def test_gaussian_field_512_refined_k_matches_fp64():
    raw = _grf512()
    say("gaussian field 512^3 bc 1")
    r = _refined_vs_fp64(raw, 1, 5e-6, "jacobi", (2.5, 7.0))
    assert r["f32dc"].sweeps == r["f64"].sweeps
    assert abs(r["f32dc"].last.wk33 / r["f64"].last.wk33 - 1) < 1e-5


# ---------------------------------------------------------------------------
# the reference's Gauss-Seidel against the fast modes at eps 1e-6
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("medium", sorted(MEDIA))
def test_fast_modes_within_1e4_of_reference_gs(medium):
    make, bc = MEDIA[medium]
    raw = make(40, 9)
    say("reference GS vs fast modes", medium)
    with pm.Solver(raw, boundary_method=bc, voxel_size=1e-5, solving_algorithm=2, eps=1e-6) as s:
        gs, _ = s.solve()
    say("  reference GS", gs.sweeps, gs.last.wk33)
    for mode, pair in (("jacobi", pair_by_porosity(raw)), ("sor3", (1.8, 6.0))):
        with pm.Solver(raw, boundary_method=bc, voxel_size=1e-5, mode=mode, precision="f32dc", omega=pair[0],
                       omega_p=pair[1], eps=1e-6) as s:
            res, _ = s.solve()
        say(" ", mode, pair, res.sweeps, res.last.wk33)
        assert abs(res.last.wk33 / gs.last.wk33 - 1) < 1e-4, (mode, res.last.wk33, gs.last.wk33)


def test_sphere_pack_256_fast_modes_within_1e4_of_reference_gs():
    """The sphere-pack workload at full size (256^3, phi 0.35, bc 2): the
    reference's lexicographic Gauss-Seidel (solving_algorithm 2, exact order,
    about a minute) to eps 1e-6, and the FP32 + FP64-refined Jacobi and
    3-colour SOR within 1e-4 of its k_zz."""
    raw = geo.sphere_pack(256, 0.35, 6, 12, seed=1)
    say("sphere pack 256^3: reference GS to eps")
    with pm.Solver(raw, boundary_method=2, voxel_size=1e-5, solving_algorithm=2, eps=1e-6) as s:
        gs, _ = s.solve()
    say("  reference GS", gs.sweeps, gs.last.wk33, f"{gs.seconds:.1f}s")
    for mode, pair in (("jacobi", (2.5, 7.0)), ("sor3", (1.8, 6.0))):
        with pm.Solver(raw, boundary_method=2, voxel_size=1e-5, mode=mode, precision="f32dc", omega=pair[0],
                       omega_p=pair[1], eps=1e-6) as s:
            res, _ = s.solve()
        say(" ", mode, pair, res.sweeps, res.last.wk33)
        assert abs(res.last.wk33 / gs.last.wk33 - 1) < 1e-4, (mode, res.last.wk33, gs.last.wk33)
