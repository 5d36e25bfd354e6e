"""Batch engine on the GPU: every sample of a batch (any mix of sizes and
boundary methods, samples joining and leaving between intervals) ends with the
sweep count, k and fields of the same solve run alone -- bit for bit."""
import os

import numpy as np
import pytest

import poremaps_amd as pm
from poremaps_amd import geometry as geo
from conftest import random_geometry

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def cases(channel_raw):
    return [
        (random_geometry((20, 18, 22), 0.3, 3), 0, 1e-5),
        (channel_raw, 0, 5e-5),
        (geo.tube(30, 30, 24, 15, 15, 9), 2, 2e-5),
        (geo.sphere_pack((22, 16, 18), 0.45, 2, 4, seed=5, require_percolation=False), 1, 1e-5),
    ]


def alone(raw, bc, dx, precision, omega, omega_p=pm.OMEGA_P):
    with pm.Solver(raw, boundary_method=bc, voxel_size=dx, mode="jacobi", precision=precision, omega=omega,
                   omega_p=omega_p, eps=1e-6) as s:
        res, _ = s.solve()
        return res, s.field("velz"), s.field("press")


def run_batch(b, slots):
    for _ in range(100000):
        if b.step() == 0:
            break
    return {s: b.status(s) for s in slots}


@pytest.mark.parametrize("precision,graph,skip2,queue", [("f32dc", True, "1", "1"), ("f32dc", False, "1", "1"),
                                                        ("f32dc", True, "0", "1"), ("f32dc", True, "1", "0"),
                                                        ("f32dc", False, "1", "0"), ("f64", True, "1", "1")])
def test_batch_equals_single_solves(channel_raw, precision, graph, skip2, queue):
    """skip2 "0": the lagged pass with every stencil load (POREMAPS_BATCH_SKIP2=0);
    queue "1": the interval's passes as persistent kernels over the tile queues
    (POREMAPS_BATCH_QUEUE=1), "0": one workgroup per tile (the default)."""
    say("batch vs single", precision, graph, skip2, queue)
    cs = cases(channel_raw)
    saved = {k: os.environ.get(k) for k in ("POREMAPS_BATCH_SKIP2", "POREMAPS_BATCH_QUEUE")}
    os.environ["POREMAPS_BATCH_SKIP2"] = skip2
    os.environ["POREMAPS_BATCH_QUEUE"] = queue
    try:
        b = pm.Batch(precision=precision, omega=2.0, omega_p=pm.OMEGA_P, use_graph=graph)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    with b:
        slots = [b.add(raw, boundary_method=bc, voxel_size=dx) for raw, bc, dx in cs]
        st = run_batch(b, slots)
        for slot, (raw, bc, dx) in zip(slots, cs):
            state, res = st[slot]
            ref, velz, press = alone(raw, bc, dx, precision, 2.0)
            assert state == "converged"
            assert res.sweeps == ref.sweeps
            assert res.last.wk33 == ref.last.wk33
            assert np.array_equal(b.field(slot, "velz"), velz)
            assert np.array_equal(b.field(slot, "press"), press)


@pytest.mark.parametrize("links", ["packed-span0", "full"])
def test_batch_compact_link_forms_equal_single_solves(channel_raw, links, monkeypatch):
    """Every sample on the compact state (POREMAPS_JACOBI=compact): its lagged
    pass from packed records whose windows hold one value each (span 0, so
    that most entries read the link arrays, in the batch's packed kernel) and
    from the link arrays themselves (the batch's general kernel); each sample
    ends as the same solve alone, bit for bit."""
    say("batch compact links", links)
    monkeypatch.setenv("POREMAPS_JACOBI", "compact")
    if links == "full":
        monkeypatch.setenv("POREMAPS_JAC_LINKS", "full")
    else:
        monkeypatch.setenv("POREMAPS_PACK_SPAN", "0")
        monkeypatch.setenv("POREMAPS_PACK_KEEP", "1")
    cs = cases(channel_raw)
    with pm.Batch(precision="f32dc", omega=2.0, omega_p=pm.OMEGA_P) as b:
        slots = [b.add(raw, boundary_method=bc, voxel_size=dx) for raw, bc, dx in cs]
        st = run_batch(b, slots)
        for slot, (raw, bc, dx) in zip(slots, cs):
            state, res = st[slot]
            ref, velz, press = alone(raw, bc, dx, "f32dc", 2.0)
            assert state == "converged"
            assert res.sweeps == ref.sweeps
            assert res.last.wk33 == ref.last.wk33
            assert np.array_equal(b.field(slot, "velz"), velz)
            assert np.array_equal(b.field(slot, "press"), press)


def test_batch_join_and_evict(channel_raw):
    """Continuous batching: a sample admitted while others run, and samples
    evicted as they finish, still reproduce their solo results."""
    say("batch join/evict")
    cs = cases(channel_raw)
    with pm.Batch(precision="f32dc", omega=2.0, omega_p=pm.OMEGA_P) as b:
        s0 = b.add(cs[0][0], boundary_method=cs[0][1], voxel_size=cs[0][2])
        s1 = b.add(cs[1][0], boundary_method=cs[1][1], voxel_size=cs[1][2])
        for _ in range(7):
            b.step()
        s2 = b.add(cs[2][0], boundary_method=cs[2][1], voxel_size=cs[2][2])
        done = {}
        for _ in range(100000):
            n = b.step()
            for s in (s0, s1, s2):
                if s not in done and b.status(s)[0] != "running":
                    done[s] = (b.status(s)[1], b.field(s, "velz"))
                    b.remove(s)
            if n == 0 and len(done) == 3:
                break
        for s, (raw, bc, dx) in zip((s0, s1, s2), cs[:3]):
            res, velz = done[s]
            ref, rvelz, _ = alone(raw, bc, dx, "f32dc", 2.0)
            assert res.sweeps == ref.sweeps and res.last.wk33 == ref.last.wk33
            assert np.array_equal(velz, rvelz)
        assert b.memory()[0] == 0


def test_batch_admission_by_memory(channel_raw):
    say("batch admission")
    with pm.Batch(precision="f32dc", budget_bytes=1) as b:
        with pytest.raises(pm.PoremapsError) as e:
            b.add(channel_raw)
        assert e.value.code == -4
    one = pm.Batch(precision="f32dc").sample_bytes(channel_raw.shape)
    with pm.Batch(precision="f32dc", budget_bytes=int(one * 1.5)) as b:
        b.add(channel_raw)
        with pytest.raises(pm.PoremapsError):
            b.add(channel_raw)
        used, budget = b.memory()
        assert used == one and budget == int(one * 1.5)


def test_relaxation_by_porosity_and_batch_equals_solo():
    """omega = 0: the pair is chosen at load from the porosity, (0.6, 30)
    below 0.20 and (2.5, 7) otherwise, the same way alone and in a batch (the
    batch default), so each batched sample still equals its solo solve."""
    tight = geo.tube(20, 18, 24, 9.5, 8.5, 3.0)  # porosity 0.08
    open_ = geo.sphere_pack((18, 16, 20), 0.5, 2, 4, seed=5)
    say("relaxation by porosity")
    with pm.Batch(precision="f32dc") as b:
        slots = [b.add(m, boundary_method=0, voxel_size=1e-5) for m in (tight, open_)]
        while b.step() > 0:
            pass
        st = [b.status(s)[1] for s in slots]
    for m, r, pair in zip((tight, open_), st, ((0.6, 30.0), (2.5, 7.0))):
        with pm.Solver(m, boundary_method=0, voxel_size=1e-5, mode="jacobi", precision="f32dc", omega=pm.AUTO,
                       eps=1e-6) as s:
            assert s.relaxation() == pair
            res, _ = s.solve()
        assert res.sweeps == r.sweeps
        assert (res.last.wk33, res.last.wmean_velz) == (r.last.wk33, r.last.wmean_velz)
        with pm.Solver(m, boundary_method=0, voxel_size=1e-5, mode="jacobi", precision="f32dc", omega=pair[0],
                       omega_p=pair[1], eps=1e-6) as s:
            res2, _ = s.solve()
        assert (res2.sweeps, res2.last.wk33) == (res.sweeps, res.last.wk33)


def test_two_batches_overlapped_on_their_streams(channel_raw):
    """Batch A's interval launched, batch B stepped on its own stream while it
    runs, then A finished (the serving loop's two lanes): every sample of both
    ends as its solve run alone, bit for bit.  While A's interval is in flight
    A takes no other call."""
    c = cases(channel_raw)
    say("overlapped batches")
    with pm.Batch(precision="f32dc", omega=pm.AUTO) as a, pm.Batch(precision="f32dc", omega=pm.AUTO) as b:
        sa = [a.add(raw, boundary_method=bc, voxel_size=dx) for raw, bc, dx in c[:2]]
        sb = [b.add(raw, boundary_method=bc, voxel_size=dx) for raw, bc, dx in c[2:]]
        with pytest.raises(pm.PoremapsError) as e:
            a.finish()
        assert e.value.code == -3
        na = nb = 1
        for _ in range(100000):
            if na:
                a.launch()
                with pytest.raises(pm.PoremapsError) as e:
                    a.add(c[0][0], boundary_method=0, voxel_size=1e-5)
                assert e.value.code == -3
            k = 0
            while nb and k < 8 and not (na and a.ready()):
                nb = b.step()
                k += 1
            if na:
                na = a.finish()
            if not na and not nb:
                break
        done = [(a, s, case) for s, case in zip(sa, c[:2])] + [(b, s, case) for s, case in zip(sb, c[2:])]
        for batch, slot, (raw, bc, dx) in done:
            state, res = batch.status(slot)
            ref, velz, _ = alone(raw, bc, dx, "f32dc", pm.AUTO)
            assert state == "converged" and res.sweeps == ref.sweeps and res.last.wk33 == ref.last.wk33
            assert np.array_equal(batch.field(slot, "velz"), velz)
