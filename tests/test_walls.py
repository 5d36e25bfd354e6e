"""Opt-in no-slip domain walls (pm_params.walls = 1) in the oracle.

The reference has no physical walls at the domain faces: the halo of a
non-periodic face is fluid (src/geometry.cc:98-106), so flow there sees a
frozen open boundary.  walls = 1 makes the halo layers of the non-periodic x
and y faces solid instead.  The stencil reaches two cells, so a medium with
two-voxel solid walls of its own, swept periodically, is the same problem:
the shipped channel (rect_channel_24_14_50, a 20 x 10 fluid duct inside
two-voxel walls, boundary_method 0) against a fully fluid 20 x 10 x 50 box
with boundary_method 2 (z periodic only) and walls = 1.  Same codes, the same
fields sweep by sweep in the reference order and in the Jacobi modes, the
reference binary's sweep count and k (scaled by the porosity, the only
difference between the two runs' k), and the duct series through it.
"""
import numpy as np
import pytest

import oracle as O
from poremaps_amd import geometry as geo

FLUID = (slice(2, -2), slice(4, -4), slice(4, -4))   # the channel's fluid region, haloed
INNER = (slice(2, -2),) * 3
PHI_CHANNEL = 200 / 336


def box():
    return np.zeros((50, 10, 20), np.uint8)


def walled(raw, pad=2):
    """raw inside explicit solid x / y walls of ``pad`` voxels."""
    out = np.ones((raw.shape[0], raw.shape[1] + 2 * pad, raw.shape[2] + 2 * pad), np.uint8)
    out[:, pad:-pad, pad:-pad] = raw
    return out


def test_channel_is_the_walled_box(channel_raw):
    assert np.array_equal(channel_raw, walled(box()))


@pytest.mark.parametrize("bc", range(5))
def test_wall_halos(bc):
    raw = geo.sphere_pack((12, 9, 11), 0.35, 2, 3, seed=bc)
    G0, G1 = O.geom_halo(raw, bc, 0), O.geom_halo(raw, bc, 1)
    assert np.array_equal(G0[INNER], G1[INNER])
    px = py = bc == 0
    if px and py:
        assert np.array_equal(G0, G1)         # walls only on non-periodic faces
        return
    # the x / y halo columns are solid over the full z extent, halo planes included
    assert (G1[:, :2, :] == 1).all() and (G1[:, -2:, :] == 1).all()
    assert (G1[:, :, :2] == 1).all() and (G1[:, :, -2:] == 1).all()
    # the z halo planes inside the walls are what the reference makes them
    assert np.array_equal(G1[:2, 2:-2, 2:-2], G0[:2, 2:-2, 2:-2])
    assert np.array_equal(G1[-2:, 2:-2, 2:-2], G0[-2:, 2:-2, 2:-2])


def test_walled_box_codes_and_reference_sweeps_equal_channel(channel_raw):
    Gc, Gb = O.geom_halo(channel_raw, 0), O.geom_halo(box(), 2, walls=1)
    cc, cb = O.eval_geometry(Gc), O.eval_geometry(Gb)
    assert np.array_equal(cc[FLUID], cb[INNER])
    sc = [a.copy() for a in O.init_fields(Gc, 0)]
    sb = [a.copy() for a in O.init_fields(Gb, 2)]
    O.pins(Gc, sc[0])
    O.pins(Gb, sb[0])
    for _ in range(40):
        O.sweep_ref(2, Gc, cc, *sc)
        O.sweep_ref(2, Gb, cb, *sb)
        for a, b in zip(sc, sb):
            O.exchange(a, 0)
            O.exchange(b, 2)
        O.pins(Gc, sc[0])
        O.pins(Gb, sb[0])
    for name, a, b in zip("puvw", sc, sb):
        assert np.array_equal(a[FLUID], b[INNER]), name


def test_walled_box_reference_solve_matches_binary_channel(known):
    """Gauss-Seidel to eps 1e-6: the reference binary's channel run (9301
    sweeps, wk33 8.368958e-09) from the fully fluid box.  The fluid sums are
    the channel's with its solid zeros left out, so conv is the same and k
    differs by the porosity alone."""
    kn = known["channel_gs"]
    G = O.geom_halo(box(), 2, walls=1)
    t, rows = O.calc_flow(G, O.eval_geometry(G), 2, O.init_fields(G, 2), solver=2, eps=1e-6, dx=5e-5)
    assert t == kn["sweeps"]
    assert "%e" % (rows[-1][10] * PHI_CHANNEL) == kn["wk33"]
    k_an = geo.duct_k(20, 10, 5e-5, 50, 1.0)
    assert abs(rows[-1][10] / k_an - 1) < 0.01


@pytest.mark.parametrize("mode,precision", [("jacobi", "f64"), ("jacobi", "f32dc"), ("sor3", "f64"),
                                            ("sor3", "f32dc")])
def test_walled_box_fast_solves_equal_channel(channel_raw, mode, precision):
    """Whole fast-mode solves.  Jacobi: bit for bit.  The SOR's colour of a
    cell is (i + j + k) mod 3 in the array's own coordinates, which the
    channel's walls shift by two voxels in x and y, so its update order -- and
    the last bits -- differ: sweep count and k agree to 1e-12, velz to 1e-7
    of its maximum (the two iterations meet at the eps = 1e-6 stop)."""
    pair = (1.8, 6.0) if mode == "sor3" else (2.5, 7.0)
    tc, kc, sc, _ = O.solve_fast(channel_raw, 0, mode, precision, *pair)
    tb, kb, sb, _ = O.solve_fast(box(), 2, mode, precision, *pair, walls=1)
    assert tb == tc
    if mode == "jacobi":
        assert kb == kc
        for name, a, b in zip("puvw", sc, sb):
            assert np.array_equal(a[FLUID], b[INNER]), name
    else:
        assert kb == pytest.approx(kc, rel=1e-12)
        assert np.allclose(sb[3][INNER], sc[3][FLUID], rtol=0, atol=1e-7 * np.abs(sc[3]).max())


@pytest.mark.parametrize("bc", [2, 3])
def test_walls_equal_explicit_solid_walls_on_a_pack(bc):
    """A sphere pack with walls = 1 against the same pack inside explicit
    two-voxel solid x / y walls: 300 FP32 + defect-correction Jacobi sweeps
    bit for bit.  bc 2 (z periodic): the walled pack is swept as an all
    periodic domain (bc 0).  bc 3 (no periodic axis): under bc 3 as well; its
    fluid x / y halos lie beyond the two-voxel walls, out of the stencil's
    reach."""
    raw = geo.sphere_pack((16, 12, 14), 0.3, 2, 3, seed=11)
    tb, kb, sb, _ = O.solve_fast(raw, bc, "jacobi", "f32dc", 2.5, 7.0, max_iter=300, walls=1)
    te, ke, se, _ = O.solve_fast(walled(raw), {2: 0, 3: 3}[bc], "jacobi", "f32dc", 2.5, 7.0, max_iter=300)
    assert tb == te and kb == ke
    for name, a, b in zip("puvw", se, sb):
        assert np.array_equal(a[FLUID], b[INNER]), name


# ---------------------------------------------------------------------------
# slip walls (walls = 2): mirror halos of the tangential velocities, the
# normal velocity held at 0 on the wall face
# ---------------------------------------------------------------------------
def plates(axis, n=6, h=10, nz=20):
    """A plane channel between two-voxel solid plates: plates at the y ends
    (axis "y", open along x, n voxels wide) or at the x ends (axis "x")."""
    if axis == "y":
        raw = np.zeros((nz, h + 4, n), np.uint8)
        raw[:, :2, :] = 1
        raw[:, -2:, :] = 1
    else:
        raw = np.zeros((nz, n, h + 4), np.uint8)
        raw[:, :, :2] = 1
        raw[:, :, -2:] = 1
    return raw


@pytest.mark.parametrize("axis", ["y", "x"])
@pytest.mark.parametrize("mode,precision", [("jacobi", "f64"), ("jacobi", "f32dc"), ("sor3", "f64"),
                                            ("sor3", "f32dc")])
def test_slip_plates_are_plane_poiseuille(axis, mode, precision):
    """Between plates 10 voxels apart whose open faces are slip walls
    (walls = 2, bc 2) the flow does not depend on the open direction: the
    same solve as with that axis periodic (bc 0) to rounding, and the
    discrete plane Poiseuille profile, whose cell-centre mean is
    (h^2 + 1/2) / 12 times the driving gradient (nz - 1) / nz (the one-sided
    wall stencils are exact on the parabola).  Plates at the y ends test the x
    faces (u held at 0 on the wall face, v and w mirrored), plates at the x
    ends the y faces.  No-slip faces and the reference's open faces both slow
    the flow."""
    raw = plates(axis)
    nz, h = raw.shape[0], 10
    pair = (1.8, 6.0) if mode == "sor3" else (2.5, 7.0)
    ts, ks, ss, _ = O.solve_fast(raw, 2, mode, precision, *pair, eps=1e-10, walls=2)
    assert ks == pytest.approx((h * h + 0.5) / 12 * (nz - 1) / nz, rel=1e-8)
    if mode == "jacobi":
        # the SOR's colour passes keep a row uniform only with periodic copies,
        # so only Jacobi runs the same iterates as the periodic axis
        tp, kp, sp, _ = O.solve_fast(raw, 0, mode, precision, *pair, eps=1e-10)
        assert ts == tp and ks == pytest.approx(kp, rel=1e-12)
        for a, b in zip(sp, ss):
            assert np.abs(a[INNER] - b[INNER]).max() <= 1e-11 * np.abs(sp[3][INNER]).max()
    if mode == "jacobi" and precision == "f64":
        _, kn, _, _ = O.solve_fast(raw, 2, mode, precision, *pair, eps=1e-10, walls=1)
        _, ko, _, _ = O.solve_fast(raw, 2, mode, precision, *pair, eps=1e-10, walls=0)
        assert kn < ko < ks


@pytest.mark.parametrize("bc", range(5))
def test_slip_halo_geometry_is_the_mirror_image(bc):
    """walls = 2: the halo layers of the non-periodic x / y faces hold the
    mirror image of the first two interior layers (a slip wall is a symmetry
    plane of the geometry too); periodic axes keep their periodic copies."""
    raw = geo.sphere_pack((12, 9, 11), 0.35, 2, 3, seed=bc)
    G0, G2 = O.geom_halo(raw, bc, 0), O.geom_halo(raw, bc, 2)
    assert np.array_equal(G0[INNER], G2[INNER])
    if bc == 0:
        assert np.array_equal(G0, G2)
        return
    zi = slice(2, -2)
    assert np.array_equal(G2[zi, :, 1], G2[zi, :, 2]) and np.array_equal(G2[zi, :, 0], G2[zi, :, 3])
    assert np.array_equal(G2[zi, :, -2], G2[zi, :, -3]) and np.array_equal(G2[zi, :, -1], G2[zi, :, -4])
    assert np.array_equal(G2[zi, 1, :], G2[zi, 2, :]) and np.array_equal(G2[zi, 0, :], G2[zi, 3, :])
    assert np.array_equal(G2[zi, -2, :], G2[zi, -3, :]) and np.array_equal(G2[zi, -1, :], G2[zi, -4, :])
