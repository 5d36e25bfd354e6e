"""Pins the CPU oracle before it is trusted as the checker of the HIP path.

* The reference's own unit-test known answers (tests/test_solver.py,
  tests/test_geometry.py, tests/test_evaluation.py) for get_2nd_derivation,
  the neighbourhood codes, porosity, convergence and permeability.
* The reference binary's measured outputs on its own benchmark inputs
  (tests/golden/reference_runs.json): sweep counts and the log's k values.
* Closed forms: Hagen-Poiseuille and the duct series, the neighbourhood-case
  characterisation, linear initial pressure, pin values.
* The fast modes' fixed point equals the reference's, and the defect-
  correction identity (FP32 correction + FP64 residual == FP64 sweep).
"""
import numpy as np
import pytest

import oracle as O
from poremaps_amd import geometry as geo
from conftest import random_geometry

KINDS = ["ref", "f64", "f32"]


# ---------------------------------------------------------------------------
# get_2nd_derivation: reference tests/test_solver.py:67-165 known answers
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", KINDS)
def test_d2_known_answers(kind):
    d = lambda *a: O.d2(*a, kind=kind)
    tol = 1e-6 if kind == "f32" else 1e-14
    assert d(1, 0.0, 0.0, 1.0, 0.0, 0.0) == pytest.approx(-8.0, abs=tol)
    assert d(1, 5.0, 5.0, 0.0, 5.0, 5.0) == 0.0
    assert d(1, 99.0, 99.0, 2.0, 99.0, 99.0) == pytest.approx(-16.0, abs=tol)
    assert d(2, 0.0, 0.0, 1.0, 3.0, 0.0) == pytest.approx(8.0 / 3.0 * 3.0 - 16.0 / 3.0, abs=tol)
    assert d(2, 0.0, 0.0, 1.0, 2.0, 0.0) == 0.0  # r = 2c cancels exactly
    assert d(3, 0.0, 3.0, 1.0, 0.0, 0.0) == pytest.approx(8.0 / 3.0 * 3.0 - 16.0 / 3.0, abs=tol)
    assert d(2, 0.0, 0.0, 2.0, 4.0, 0.0) == d(3, 0.0, 4.0, 2.0, 0.0, 0.0)
    assert d(4, 0.0, 0.0, 1.0, 2.0, 3.0) == pytest.approx(2.0 * 2.0 - 5.0 - 0.6, abs=tol)
    assert d(5, 3.0, 2.0, 1.0, 0.0, 0.0) == pytest.approx(2.0 * 2.0 - 0.6 - 5.0, abs=tol)
    assert d(4, 0.0, 0.0, 2.0, 4.0, 6.0) == pytest.approx(d(5, 6.0, 4.0, 2.0, 0.0, 0.0), abs=tol)
    assert d(6, 0.0, 1.0, 0.0, 1.0, 0.0) == 2.0       # x^2
    assert d(6, 0.0, -1.0, 0.0, 1.0, 0.0) == 0.0      # x
    assert d(6, 5.0, 5.0, 5.0, 5.0, 5.0) == 0.0       # const
    assert d(6, 99.0, 1.0, 2.0, 3.0, 99.0) == d(6, -1.0, 1.0, 2.0, 3.0, -1.0)


def test_d2_one_sided_stencils_are_exact_for_their_polynomials():
    # Wall-aware second derivatives with the no-slip wall half a cell beyond
    # the last fluid cell: each one-sided stencil is exact for the quadratic
    # that vanishes on its wall(s) (cell centres at 0.5, 1.5, 2.5); a dropped
    # term, a wrong sign or swapped ll/rr operands breaks these.
    # case 4: wall left of c (at c - 1/2), fluid c, r, rr; case 5 mirrored
    g = lambda x: x * (4.0 - x)
    got_g = O.d2(4, 0.0, 0.0, g(0.5), g(1.5), g(2.5))
    assert got_g == pytest.approx(-2.0, rel=1e-12)
    got_g5 = O.d2(5, g(2.5), g(1.5), g(0.5), 0.0, 0.0)
    assert got_g5 == pytest.approx(-2.0, rel=1e-12)
    # case 2: wall left, one fluid neighbour r, solid beyond -> exact for the
    # parabola vanishing on both walls (channel of width 2 cells)
    h = lambda x: x * (2.0 - x)
    assert O.d2(2, 0.0, 0.0, h(0.5), h(1.5), 0.0) == pytest.approx(-2.0, rel=1e-12)
    assert O.d2(3, 0.0, h(0.5), h(1.5), 0.0, 0.0) == pytest.approx(-2.0, rel=1e-12)
    # case 1: single cell between two walls, parabola x(1-x)
    k = lambda x: x * (1.0 - x)
    assert O.d2(1, 0.0, 0.0, k(0.5), 0.0, 0.0) == pytest.approx(-2.0, rel=1e-12)


# ---------------------------------------------------------------------------
# eval_geometry: reference tests/test_geometry.py:127-191 + characterisation
# ---------------------------------------------------------------------------
def _codes(raw, bc):
    G = O.geom_halo(raw, bc)
    return G, O.eval_geometry(G)


def test_isolated_fluid_voxel_is_111():
    raw = np.ones((7, 7, 7), np.uint8)
    raw[3, 3, 3] = 0
    for bc in range(5):
        _, c = _codes(raw, bc)
        assert c[5, 5, 5] == 111


@pytest.mark.parametrize("bc", range(5))
def test_all_fluid_interior_is_666(bc):
    # the binary's wall halos are fluid (src/geometry.cc:98-106) and periodic
    # halos are copies, so an all-fluid box is 666 everywhere
    _, c = _codes(np.zeros((8, 8, 8), np.uint8), bc)
    assert (c[2:-2, 2:-2, 2:-2] == 666).all()


def test_channel_wall_neighbour_and_solids():
    raw = np.zeros((10, 6, 6), np.uint8)
    raw[:, 0, :] = 1; raw[:, -1, :] = 1; raw[:, :, 0] = 1; raw[:, :, -1] = 1
    G, c = _codes(raw, 0)
    assert c[2 + 5, 2 + 3, 2 + 1] // 100 == 4      # first fluid x beside a wall
    assert (c[G == 1] == -1).all()


def test_adjacent_pair_is_case_2_and_3():
    raw = np.ones((5, 5, 5), np.uint8)
    raw[2, 2, 1] = 0
    raw[2, 2, 2] = 0
    _, c = _codes(raw, 0)
    assert c[4, 4, 3] // 100 == 2 and c[4, 4, 4] // 100 == 3


@pytest.mark.parametrize("bc", range(5))
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_codes_match_their_definition(bc, seed):
    """Every axis digit is the case its four neighbours define (src/geometry.cc
    :246-289): {1,2,4} <=> -1 solid, {1,3,5} <=> +1 solid, 2 => +2 solid,
    4 => +2 fluid, 3 => -2 solid, 5 => -2 fluid."""
    raw = random_geometry((6, 5, 7), 0.6, seed)
    G, c = _codes(raw, bc)
    nz, ny, nx = raw.shape
    for i in range(2, nz + 2):
        for j in range(2, ny + 2):
            for k in range(2, nx + 2):
                if G[i, j, k]:
                    assert c[i, j, k] == -1
                    continue
                digits = (c[i, j, k] // 100, c[i, j, k] // 10 % 10, c[i, j, k] % 10)
                for cs, (di, dj, dk) in zip(digits, [(0, 0, 1), (0, 1, 0), (1, 0, 0)]):
                    s = lambda m: G[i + m * di, j + m * dj, k + m * dk] != 0
                    assert (cs in (1, 2, 4)) == s(-1)
                    assert (cs in (1, 3, 5)) == s(1)
                    if cs == 2: assert s(2)
                    if cs == 4: assert not s(2)
                    if cs == 3: assert s(-2)
                    if cs == 5: assert not s(-2)


def test_geometry_halo_periodic_copy_and_fluid_walls():
    raw = random_geometry((6, 5, 7), 0.5, 9)
    G0 = O.geom_halo(raw, 0)
    assert (G0[2:-2, 2:-2, 0] == raw[:, :, -2]).all() and (G0[2:-2, 2:-2, -1] == raw[:, :, 1]).all()
    assert (G0[0, 2:-2, 2:-2] == raw[-2]).all()
    G3 = O.geom_halo(raw, 3)
    assert (G3[:, :, :2] == 0).all() and (G3[:2] == 0).all()
    G1 = O.geom_halo(raw, 1)
    assert (G1[:, :, :2] == 0).all() and (G1[0, 2:-2, 2:-2] == raw[-2]).all()


def test_pack_code_round_trip():
    raw = random_geometry((6, 5, 7), 0.5, 4)
    _, c = _codes(raw, 0)
    pc = O.pack_code(c)
    fl = c > 0
    dec = (pc[fl].astype(int) // 36 + 1) * 100 + ((pc[fl].astype(int) // 6) % 6 + 1) * 10 + pc[fl] % 6 + 1
    assert (dec == c[fl]).all()
    assert (pc[~fl] == 255).all()


# ---------------------------------------------------------------------------
# porosity, initial state, pins
# ---------------------------------------------------------------------------
def test_porosity_closed_forms(channel_raw):
    assert O.porosity(O.geom_halo(np.zeros((4, 4, 4), np.uint8), 3)) == 1.0
    assert O.porosity(O.geom_halo(np.ones((6, 6, 6), np.uint8), 0)) == 0.0
    half = np.zeros((6, 6, 6), np.uint8); half[:, :, :3] = 1
    assert O.porosity(O.geom_halo(half, 0)) == 0.5
    one = np.zeros((6, 6, 6), np.uint8); one[3, 3, 3] = 1
    assert O.porosity(O.geom_halo(one, 0)) == 215 / 216
    assert O.porosity(O.geom_halo(channel_raw, 0)) == 200 / 336


@pytest.mark.parametrize("bc", [0, 2, 4])
def test_initial_pressure_is_linear_and_pins(bc):
    raw = random_geometry((9, 5, 6), 0.3, 5)
    G = O.geom_halo(raw, bc)
    p, u, v, w = O.init_fields(G, bc)
    nz = raw.shape[0]
    for i in range(2, nz + 2):
        sl = p[i, 2:-2, 2:-2]
        assert (sl[raw[i - 2] == 0] == nz - i + 4).all()
        assert (sl[raw[i - 2] == 1] == 0.0).all()
    assert (u == 0).all() and (v == 0).all() and (w == 0).all()
    if bc != 0:
        assert (p[2:-2, 2:-2, :2] == -1.0).all()
    O.pins(G, p)
    inl = p[1, 2:-2, 2:-2][G[1, 2:-2, 2:-2] == 0]
    out = p[-3, 2:-2, 2:-2][G[-3, 2:-2, 2:-2] == 0]
    assert (inl == nz + 2).all() and (out == 3.0).all()


# ---------------------------------------------------------------------------
# evaluation: reference tests/test_evaluation.py:45-203 known answers
# ---------------------------------------------------------------------------
def _fluid_state(n=6, val=(0.0, 0.0, 0.0)):
    G = O.geom_halo(np.zeros((n, n, n), np.uint8), 0)
    shp = G.shape
    return G, [np.full(shp, val[0]), np.full(shp, val[1]), np.full(shp, val[2]), np.zeros(shp)]


def test_permeability_known_answers():
    G, (u, v, w, p) = _fluid_state(val=(0.0, 0.0, 1.0))
    k = O.permeability(G, u, v, w, p, (0,) * 6, 1e-6, 1.0)
    assert k[5] == pytest.approx((1.0 / 0.01) * 1e-12) and k[3] == 0.0 and k[4] == 0.0
    k05 = O.permeability(G, u, v, w, p, (0,) * 6, 1e-6, 0.5)
    assert k05[5] == pytest.approx(0.5 * k[5])
    k2 = O.permeability(G, u, v, w, p, (0,) * 6, 2e-6, 1.0)
    assert k2[5] == pytest.approx(4.0 * k[5])
    G, (u, v, w, p) = _fluid_state(n=4)
    w[2:4] = 1.0; w[4:6] = 2.0
    k = O.permeability(G, u, v, w, p, (0,) * 6, 1.0, 1.0)
    assert k[6] == 2.0 and k[7] == 1.5          # wmax, wmean
    G, (u, v, w, p) = _fluid_state(val=(1.0, 0.0, 0.0))
    k = O.permeability(G, u, v, w, p, (0,) * 6, 1e-6, 1.0)
    assert k[3] == pytest.approx(100.0 * 1e-12)


def test_convergence_known_answers():
    G, (u, v, w, p) = _fluid_state(val=(0.0, 0.0, 1.0))
    conv, k = O.convergence(G, w, 0.0)
    assert conv == 1.0 and k == pytest.approx(1.0 / 0.01)
    conv, k2 = O.convergence(G, 2.0 * w, k)
    assert conv == pytest.approx(abs(200.0 - 100.0) / 200.0) and k2 == pytest.approx(200.0)
    conv, _ = O.convergence(G, w, 1.0 / 0.01)
    assert conv == 0.0


def test_dom_interest_pressure_gradient_sign():
    # dom_interest k uses the signed plane-mean pressure difference, so a
    # linear pressure falling along +z gives a negative gradient
    G, (u, v, w, p) = _fluid_state(n=6, val=(0.0, 0.0, 1.0))
    for i in range(p.shape[0]):
        p[i] = 10.0 - i
    k = O.permeability(G, u, v, w, p, (0, 5, 0, 5, 1, 4), 1.0, 1.0)
    assert k[2] == pytest.approx(-100.0) and k[5] == pytest.approx(100.0)


# ---------------------------------------------------------------------------
# whole solver: the reference binary's measured runs
# ---------------------------------------------------------------------------
def _run_channel(channel_raw, **kw):
    G = O.geom_halo(channel_raw, 0)
    code = O.eval_geometry(G)
    st = O.init_fields(G, 0)
    return O.calc_flow(G, code, 0, st, dx=5e-5, **kw)


@pytest.mark.parametrize("case", ["channel_gs", "channel_jacobi_ref"])
def test_channel_matches_reference_binary(channel_raw, known, case):
    kn = known[case]
    t, rows = _run_channel(channel_raw, solver=kn["solver"], eps=1e-6)
    assert t == kn["sweeps"]
    assert "%e" % rows[-1][10] == kn["wk33"]
    assert rows[-1][0] == kn["sweeps"] - 1
    if kn["solver"] == known["channel_gs_t3000"]["solver"]:
        r3 = rows[rows[:, 0] == 3000][0]
        assert "%e" % r3[10] == known["channel_gs_t3000"]["wk33"]


def test_channel_dom_interest_matches_reference_binary(channel_raw, known):
    kn = known["channel_dom_interest_t1500"]
    _, rows = _run_channel(channel_raw, solver=2, eps=1e-6, dom=tuple(kn["dom_interest"]), max_iter=1501)
    r = rows[rows[:, 0] == kn["iteration"]][0]
    assert "%e" % r[7] == kn["k33"] and "%e" % r[10] == kn["wk33"]


def test_channel_reference_sor_matches_reference_binary(channel_raw, known):
    kn = known["channel_sor_ref"]
    t, rows = _run_channel(channel_raw, solver=3, eps=1e-6)
    assert t == kn["sweeps"] and ("%e" % rows[-1][10]).startswith(kn["wk33_prefix"])


def test_channel_against_duct_series(channel_raw, known):
    _, rows = _run_channel(channel_raw, solver=2, eps=1e-9)
    k_an = geo.duct_k(20, 10, 5e-5, 50, 200 / 336)
    assert "%e" % k_an == known["channel_duct_series"]["k33"]
    assert abs(rows[-1][10] / k_an - 1) < 0.01


@pytest.mark.slow
def test_tube_matches_reference_binary_and_hagen_poiseuille(known):
    raw = geo.config0()
    G = O.geom_halo(raw, 0)
    code = O.eval_geometry(G)
    for case in ("tube_sor_ref", "tube_gs"):
        kn = known[case]
        t, rows = O.calc_flow(G, code, 0, O.init_fields(G, 0), solver=kn["solver"], eps=1e-6, dx=2e-5)
        assert t == kn["sweeps"] and "%e" % rows[-1][10] == kn["wk33"]
    k_hp = geo.hagen_poiseuille_k(raw, 20, 2e-5, 50)
    assert "%e" % k_hp == known["tube_hagen_poiseuille"]["k33"]
    assert abs(rows[-1][10] / k_hp - 1) < 0.03


def test_tube_short_run_against_hagen_poiseuille_scale():
    """Cheap stand-in for the slow tube run: the first log row's wk33 is
    positive and below the analytic value (the flow is still developing)."""
    raw = geo.config0()
    G = O.geom_halo(raw, 0)
    t, rows = O.calc_flow(G, O.eval_geometry(G), 0, O.init_fields(G, 0), solver=2, eps=1e-6,
                          dx=2e-5, max_iter=1201)
    k_hp = geo.hagen_poiseuille_k(raw, 20, 2e-5, 50)
    assert 0 < rows[-1][10] < k_hp


# ---------------------------------------------------------------------------
# fast modes: same fixed point as the reference; defect-correction identity
# ---------------------------------------------------------------------------
def fast_solve(G, pc, bc, mode, omega, eps, maxit=400000):
    A = [a.copy() for a in O.init_fields(G, bc)]
    O.pins(G, A[0])
    B = [a.copy() for a in A]
    qa = O.outlet_plane(pc, bc, A)
    qb = qa.copy()
    pins = O.solution_pins(G)
    kprev, t, conv = 0.0, 0, 1 + eps
    while t < maxit and conv >= eps:
        if mode == "jacobi":
            O.sweep_jacobi(pc, bc, omega, A, B, qa, qb, 3.0)
            O.ghost(G, bc, B, qb, *pins)
            A, B, qa, qb = B, A, qb, qa
        else:
            O.sweep_colour(pc, omega, A, qa, 3.0)
            O.ghost(G, bc, A, qa, *pins)
        if t % 100 == 0 and t > 1000:
            conv, kprev = O.convergence(G, A[3], kprev)
        t += 1
    return t, kprev, A


@pytest.mark.parametrize("mode,omega", [("jacobi", 1.0), ("jacobi", 2.5), ("sor3", 1.8)])
def test_fast_modes_share_the_reference_fixed_point(channel_raw, mode, omega):
    G = O.geom_halo(channel_raw, 0)
    code = O.eval_geometry(G)
    _, rows = O.calc_flow(G, code, 0, O.init_fields(G, 0), solver=2, eps=1e-11, dx=5e-5)
    k_ref = rows[-1][10] / ((200 / 336) * 5e-5 ** 2)   # wk33 -> mean(w)/Re
    _, k, _ = fast_solve(G, O.pack_code(code), 0, mode, omega, 1e-11)
    assert k == pytest.approx(k_ref, rel=2e-9)


@pytest.mark.parametrize("mode", ["jacobi", "sor3"])
@pytest.mark.parametrize("bc", [0, 2])
def test_defect_correction_reproduces_fp64_sweeps(mode, bc):
    """m FP32 correction sweeps on (X, R) plus a fold equal m FP64 sweeps of
    the same ordering up to FP32 rounding of the correction (early sweeps,
    where corrections are as large as the state: a few 1e-6 relative; a
    missing residual, pin or outlet term is an O(1) difference)."""
    raw = random_geometry((10, 8, 9), 0.35, 11)
    G = O.geom_halo(raw, bc)
    pc = O.pack_code(O.eval_geometry(G))
    pins = O.solution_pins(G)
    om = 1.5
    # FP64 reference trajectory
    A = [a.copy() for a in O.init_fields(G, bc)]
    O.pins(G, A[0])
    B = [a.copy() for a in A]
    qa = O.outlet_plane(pc, bc, A); qb = qa.copy()
    X = [a.copy() for a in A]
    d = [np.zeros(G.shape, np.float32) for _ in range(4)]
    d2 = [np.zeros(G.shape, np.float32) for _ in range(4)]
    qd = np.zeros(G.shape[1:], np.float32); qd2 = qd.copy()
    R = O.residual(pc, X, O.outlet_plane(pc, bc, X))
    for t in range(60):
        if mode == "jacobi":
            O.sweep_jacobi(pc, bc, om, A, B, qa, qb, 3.0); O.ghost(G, bc, B, qb, *pins)
            A, B, qa, qb = B, A, qb, qa
            O.sweep_jacobi(pc, bc, om, d, d2, qd, qd2, np.float32(0), R=R)
            O.ghost(G, bc, d2, qd2, 0.0, 0.0, 0.0)
            d, d2, qd, qd2 = d2, d, qd2, qd
        else:
            O.sweep_colour(pc, om, A, qa, 3.0); O.ghost(G, bc, A, qa, *pins)
            O.sweep_colour(pc, om, d, qd, np.float32(0), R=R)
            O.ghost(G, bc, d, qd, 0.0, 0.0, 0.0)
        if t % 20 == 0:
            O.fold(G, bc, X, d)
            R = O.residual(pc, X, O.outlet_plane(pc, bc, X))
            qd[:] = 0; qd2[:] = 0
    O.fold(G, bc, X, d)
    fl = (G == 0)
    fl[:2] = fl[-2:] = False; fl[:, :2] = fl[:, -2:] = False; fl[:, :, :2] = fl[:, :, -2:] = False
    for a, x in zip(A[1:], X[1:]):
        scale = np.abs(a[fl]).max()
        assert np.abs(a[fl] - x[fl]).max() <= 1e-5 * scale
    assert np.abs(A[0][fl] - X[0][fl]).max() <= 1e-5 * np.abs(A[0][fl]).max()


# ---------------------------------------------------------------------------
# threaded block sweeps (the CPU baseline on all host cores): one block is
# the serial sweep; decompositions against the reference binary are pinned
# in tests/test_reference_fields.py
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("bc", [0, 2, 4])
def test_one_block_is_the_serial_sweep(bc):
    raw = random_geometry((17, 10, 12), 0.4, 11 + bc)
    G = O.geom_halo(raw, bc)
    code = O.eval_geometry(G)
    st = [f.copy() for f in O.init_fields(G, bc)]
    O.pins(G, st[0])
    got = [f.copy() for f in st]
    O.sweeps_ref_slabs(2, G, code, *got, bc, 6, 1)
    ser = [f.copy() for f in st]
    for _ in range(6):
        O.sweep_ref(2, G, code, *ser)
        for f in ser:
            O.exchange(f, bc)
        O.pins(G, ser[0])
    for a, b in zip(got, ser):
        assert np.array_equal(a, b)
