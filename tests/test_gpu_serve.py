"""Permeability tensor and the serving loop on the GPU.

* pm_tensor's columns are the solves of the numpy-rotated geometry run alone,
  bit for bit, scattered to the physical axes (README.md:217-229 procedure).
* A geometry invariant under the cyclic axis permutation gives three equal
  diagonal entries and cyclically equal off-diagonals, bitwise.
* pm_serve: every response equals the tensor of its geometry solved alone,
  whatever the arrival pattern, memory budget and co-running requests.
"""
import numpy as np
import pytest

import poremaps_amd as pm
from poremaps_amd import geometry as geo

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def solo(raw, bc, dx):
    # the serving default: the relaxation pair chosen from the porosity
    with pm.Solver(raw, boundary_method=bc, voxel_size=dx, mode="jacobi", precision="f32dc", omega=pm.AUTO,
                   eps=1e-6) as s:
        res, _ = s.solve()
        return res


def aniso():
    return geo.sphere_pack((18, 14, 22), 0.55, 2, 4, seed=3, axes=(0, 1, 2))


def test_tensor_columns_are_rotated_solves():
    raw = aniso()
    say("tensor columns", raw.shape)
    k, res = pm.tensor(raw, boundary_method=0, voxel_size=1e-5, precision="f32dc")
    # flow along z: the geometry itself
    r = solo(raw, 0, 1e-5)
    assert res[0].sweeps == r.sweeps
    assert (k[0, 2], k[1, 2], k[2, 2]) == (r.last.wk13, r.last.wk23, r.last.wk33)
    # flow along x: array axes (z, y, x) -> (x, z, y)
    r = solo(np.ascontiguousarray(raw.transpose(2, 0, 1)), 0, 1e-5)
    assert res[1].sweeps == r.sweeps
    assert (k[1, 0], k[2, 0], k[0, 0]) == (r.last.wk13, r.last.wk23, r.last.wk33)
    # flow along y: array axes (z, y, x) -> (y, x, z)
    r = solo(np.ascontiguousarray(raw.transpose(1, 2, 0)), 0, 1e-5)
    assert res[2].sweeps == r.sweeps
    assert (k[2, 1], k[0, 1], k[1, 1]) == (r.last.wk13, r.last.wk23, r.last.wk33)
    assert np.all(np.diag(k) > 0)


def cross(n, h):
    """Three orthogonal square ducts of half-width h through the centre of an
    n-cube: invariant under every axis permutation and mirror."""
    c = (n - 1) / 2.0
    z, y, x = np.ogrid[:n, :n, :n]
    ax, ay, az = abs(x - c) < h, abs(y - c) < h, abs(z - c) < h
    return np.where((ax & ay) | (ay & az) | (ax & az), 0, 1).astype(np.uint8)


def test_tensor_of_cubic_cross_is_isotropic():
    raw = cross(24, 4)
    say("tensor cubic cross")
    k, res = pm.tensor(raw, boundary_method=0, voxel_size=1e-5)
    assert k[0, 0] == k[1, 1] == k[2, 2] > 0
    assert res[0].sweeps == res[1].sweeps == res[2].sweeps
    # the three rotated arrays are identical, so the cyclic images agree bitwise
    assert k[1, 0] == k[2, 1] == k[0, 2] and k[2, 0] == k[0, 1] == k[1, 2]
    # the staggered scheme is not exactly mirror-symmetric: the CPU oracle's
    # reference-order solve of a 16-cube cross gives wk13 = wk23 = -9.70e-15
    # against wk33 = 2.14e-11 (ratio 4.5e-4); a mis-scattered column would
    # put a diagonal-sized value here
    off = k[~np.eye(3, dtype=bool)]
    assert np.all(np.abs(off) < 1e-3 * k[2, 2])


def test_serve_matches_requests_solved_alone():
    media = [aniso(), cross(16, 3), geo.sphere_pack((16, 16, 16), 0.5, 2, 3, seed=9, axes=(0, 1, 2)),
             geo.gaussian_field(16, 4.0, 0.4, seed=4, axes=(0, 1, 2))]
    bcs = [0, 1, 2, 0]
    reqs = [(m, bc, 1e-5, 1e-6, 0.03 * i) for i, (m, bc) in enumerate(zip(media, bcs))]
    with pm.Batch() as b:
        needs = [sum(b.sample_bytes(np.transpose(m, t).shape) for t in ((0, 1, 2), (2, 0, 1), (1, 2, 0)))
                 for m in media]
    # room for one request at a time: later arrivals queue for memory
    assert 2 * min(needs) > max(needs)
    # small_cells 5000: the 16^3 requests (4,096 voxels) in the second lane,
    # the others in the first, both lanes busy at once
    for budget, realtime, small in ((0, False, pm.SMALL_CELLS), (max(needs), True, pm.SMALL_CELLS),
                                    (0, False, 5000), (0, False, 0)):
        say("serve", "budget", budget, "realtime", realtime, "small_cells", small)
        resp, st = pm.serve(reqs, realtime=realtime, budget_bytes=budget, small_cells=small)
        assert st.completed == len(reqs)
        assert st.latency_p50_s <= st.latency_p99_s <= st.latency_max_s
        assert st.solves_per_hour > 0
        if budget:
            assert st.peak_bytes <= budget and st.max_running <= 3
        else:
            assert st.max_running > 3
        for (m, bc, dx, _, t), r in zip(reqs, resp):
            k, res = pm.tensor(m, boundary_method=bc, voxel_size=dx)
            assert np.array_equal(r.tensor, k)
            assert list(r.sweeps) == [x.sweeps for x in res]
            assert list(r.state) == [1, 1, 1]
            assert r.arrival_s == t and r.done_s > r.admit_s
            if realtime:
                assert r.admit_s >= t


def test_serve_per_request_eps():
    m = aniso()
    say("serve eps")
    resp, _ = pm.serve([(m, 0, 1e-5, 1e-6, 0.0), (m, 0, 1e-5, 1e-9, 0.0)], realtime=False, max_iter=200_000)
    say("sweeps", list(resp[0].sweeps), list(resp[1].sweeps))
    assert list(resp[1].state) == [1, 1, 1]
    assert all(a <= b for a, b in zip(resp[0].sweeps, resp[1].sweeps))
    assert sum(resp[0].sweeps) < sum(resp[1].sweeps)
    for r, eps in zip(resp, (1e-6, 1e-9)):
        k, _ = pm.tensor(m, boundary_method=0, voxel_size=1e-5, eps=eps)
        assert np.array_equal(r.tensor, k)


def test_serve_dom_interest_tensor():
    """A request with a dom_interest box: each rotated solve reports the
    sub-domain (k13, k23, k33) of the box in its own axes, and the response's
    tensor_dom equals those of the rotated arrays solved alone with the
    rotated boxes (the reference's sub-domain permeability, pressure gradient
    along each solve's flow axis)."""
    raw = aniso()  # (nz, ny, nx) = (18, 14, 22)
    box = (3, 17, 2, 11, 4, 13)  # x0 x1 y0 y1 z0 z1
    say("serve dom_interest")
    resp, _ = pm.serve([(raw, 0, 1e-5, 1e-6, 0.0, box)], realtime=False)
    kd = resp[0].tensor_dom
    assert np.array_equal(resp[0].tensor, pm.tensor(raw, boundary_method=0, voxel_size=1e-5)[0])
    # rotation r: array axes (z, y, x) -> transpose perm; box of the rotated
    # array's (x, y, z) = box of physical axes new_axis[r]
    perms = ((0, 1, 2), (2, 0, 1), (1, 2, 0))
    new_axis = ((0, 1, 2), (1, 2, 0), (2, 0, 1))
    for perm, na in zip(perms, new_axis):
        a = np.ascontiguousarray(raw.transpose(perm))
        rb = sum(((box[2 * ax], box[2 * ax + 1]) for ax in na), ())
        with pm.Solver(a, boundary_method=0, voxel_size=1e-5, mode="jacobi", precision="f32dc", omega=pm.AUTO,
                       eps=1e-6, dom_interest=rb) as s:
            _, rows = s.solve()
        last = rows[-1]
        ax, ay, az = na
        assert (kd[ax, az], kd[ay, az], kd[az, az]) == (last["k13"], last["k23"], last["k33"])
        assert last["k33"] != 0.0



def test_serve_rejects_only_the_requests_it_cannot_admit():
    """A request larger than the whole memory budget, and one whose geometry
    holds a byte other than 0 / 1, are answered PM_SAMPLE_REJECTED with zero k;
    every other request in the stream is still served, equal to its tensor
    solved alone, and the rejected ones stay out of the latency figures."""
    small = [geo.sphere_pack((16, 16, 16), 0.5, 2, 3, seed=s, axes=(0, 1, 2)) for s in (21, 22)]
    big = geo.sphere_pack((40, 40, 40), 0.5, 2, 3, seed=23, axes=(0, 1, 2))
    bad = small[0].copy()
    bad[3, 4, 5] = 2
    with pm.Batch() as b:
        needs = [sum(b.sample_bytes(np.transpose(m, t).shape) for t in ((0, 1, 2), (2, 0, 1), (1, 2, 0)))
                 for m in (small[0], big)]
    assert needs[1] > needs[0]
    reqs = [(small[0], 0, 1e-5, 1e-6, 0.0), (big, 0, 1e-5, 1e-6, 0.01), (bad, 0, 1e-5, 1e-6, 0.02),
            (small[1], 0, 1e-5, 1e-6, 0.03)]
    say("serve rejects", needs)
    resp, st = pm.serve(reqs, realtime=False, budget_bytes=needs[0])
    assert st.rejected == 2 and st.completed == 2
    for i in (1, 2):
        assert list(resp[i].state) == [pm.SAMPLE_REJECTED] * 3 and not resp[i].tensor.any()
    for i in (0, 3):
        k, res = pm.tensor(reqs[i][0], boundary_method=0, voxel_size=1e-5)
        assert list(resp[i].state) == [1, 1, 1] and np.array_equal(resp[i].tensor, k)
