"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bit-exact (np.array_equal on every compared cell) for the integer geometry
codes, the initial state, the reference-order sweeps and the FP64/FP32 fast
sweeps; the reductions are compensated parallel sums, compared to 1e-12
relative and on the log's printed digits.  Full-size runs are checked through
sweep counts and k against the reference binary's known answers.
"""
import contextlib
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
import poremaps_amd as pm
from poremaps_amd import geometry as geo
from conftest import GOLDEN, random_geometry

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def nonc(a):
    """Interior and face-halo cells (corner halo cells are never read)."""
    m = np.zeros(a.shape, bool)
    h = [np.zeros(n, bool) for n in a.shape]
    for ax, n in enumerate(a.shape):
        h[ax][:2] = True
        h[ax][-2:] = True
    Z, Y, X = np.meshgrid(h[0], h[1], h[2], indexing="ij")
    m = (Z.astype(int) + Y.astype(int) + X.astype(int)) <= 1
    return m


def oracle_state(raw, bc):
    G = O.geom_halo(raw, bc)
    code = O.eval_geometry(G)
    st = [a.copy() for a in O.init_fields(G, bc)]
    O.pins(G, st[0])
    return G, code, st


# ---------------------------------------------------------------------------
# geometry, porosity, initial state
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("bc", range(5))
def test_geometry_codes_porosity_and_initial_state(bc):
    raw = random_geometry((9, 11, 13), 0.45, 100 + bc)
    say("geometry", bc, raw.shape)
    G, code, st = oracle_state(raw, bc)
    with pm.Solver(raw, boundary_method=bc) as s:
        assert np.array_equal(s.field("nhood"), code[2:-2, 2:-2, 2:-2])
        assert s.porosity == O.porosity(G)
        for a, b in zip(s.state_halo(), st):
            assert np.array_equal(a, b)


def test_geometry_rejects_non_binary_bytes():
    raw = np.zeros((4, 4, 4), np.uint8)
    raw[1, 2, 3] = 2
    with pytest.raises(pm.PoremapsError) as e:
        pm.Solver(raw)
    assert e.value.code == -5


# ---------------------------------------------------------------------------
# reference order, per iteration
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("solver", [1, 2, 3])
@pytest.mark.parametrize("bc", range(5))
def test_reference_sweeps_bit_exact(solver, bc):
    raw = random_geometry((13, 9, 11), 0.4, 7 * bc + solver)
    say("reference sweeps", solver, bc)
    G, code, st = oracle_state(raw, bc)
    with pm.Solver(raw, boundary_method=bc, solving_algorithm=solver) as s:
        done = 0
        for n in (1, 6, 33):
            for _ in range(n):
                O.sweep_ref(solver, G, code, *st)
                for f in st:
                    O.exchange(f, bc)
                O.pins(G, st[0])
            s.step(n)
            done += n
            for name, a, b in zip("puvw", s.state_halo(), st):
                assert np.array_equal(a, b), f"{name} differs after {done} sweeps"


@contextlib.contextmanager
def env(name, value):
    """An environment knob of the library, for solvers created inside."""
    old = os.environ.get(name)
    os.environ[name] = value
    try:
        yield
    finally:
        if old is None:
            os.environ.pop(name, None)
        else:
            os.environ[name] = old


def ref_form(form):
    """The exact-order sweep's kernel form for solvers created inside."""
    return env("POREMAPS_REF", form)


@pytest.mark.parametrize("form", ["onewg", "planes"])
@pytest.mark.parametrize("solver,bc", [(2, 0), (3, 2), (1, 4)])
def test_reference_sweep_forms_bit_exact(form, solver, bc):
    # the default LDS-tiled wavefront is covered above; the single-workgroup
    # and per-hyperplane forms must give the same bits
    raw = random_geometry((13, 9, 11), 0.4, 7 * bc + solver)
    say("reference form", form, solver, bc)
    G, code, st = oracle_state(raw, bc)
    with ref_form(form), pm.Solver(raw, boundary_method=bc, solving_algorithm=solver) as s:
        for _ in range(7):
            O.sweep_ref(solver, G, code, *st)
            for f in st:
                O.exchange(f, bc)
            O.pins(G, st[0])
        s.step(7)
        for name, a, b in zip("puvw", s.state_halo(), st):
            assert np.array_equal(a, b), f"{name} differs ({form})"


@pytest.mark.parametrize("form", ["tiles", "planes"])
@pytest.mark.parametrize("solver", [2, 3])
def test_reference_sweeps_large_domain_bit_exact(form, solver):
    # many tiles per diagonal, partial tiles on every axis (60 x 70 x 98)
    raw = random_geometry((60, 70, 98), 0.35, 5)
    say("reference large domain", form, solver, raw.shape)
    G, code, st = oracle_state(raw, 0)
    with ref_form(form), pm.Solver(raw, boundary_method=0, solving_algorithm=solver) as s:
        for _ in range(2):
            O.sweep_ref(solver, G, code, *st)
            for f in st:
                O.exchange(f, 0)
            O.pins(G, st[0])
        s.step(2)
        for a, b in zip(s.state_halo(), st):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("P", [2, 3, 4])
@pytest.mark.parametrize("solver,bc", [(2, 0), (1, 2), (3, 4), (2, 1)])
def test_reference_z_slabs_match_slab_oracle(P, solver, bc):
    """dom_decomposition 1 1 P on one GPU: P z-slabs, each swept against
    frozen copies of its neighbours' edge planes, then the exchange and the
    pins -- bit for bit the oracle's restatement of `mpirun -np P`
    (orc_sweeps_ref_slabs, pinned to the reference binary's decomposed runs
    in tests/test_reference_fields.py), with slabs of unequal thickness
    (nz 13)."""
    raw = random_geometry((13, 9, 11), 0.35, 11 * P + bc)
    say("reference z-slabs", P, solver, bc)
    G, code, st = oracle_state(raw, bc)
    O.sweeps_ref_slabs(solver, G, code, *st, bc, 5, P)
    with pm.Solver(raw, boundary_method=bc, solving_algorithm=solver, z_slabs=P) as s:
        s.step(5)
        for name, a, b in zip("puvw", s.state_halo(), st):
            assert np.array_equal(a, b), f"{name} differs"


def test_reference_z_slabs_several_tiles_per_slab():
    """Slabs of 14 / 13 / 13 planes (two tile layers each) on a 40 x 20 x 30
    field; the decomposition changes the iterates (frozen slab edges), and the
    GPU follows the oracle's slabs, not the one-domain sweep."""
    raw = random_geometry((40, 20, 30), 0.35, 77)
    say("reference z-slabs large")
    G, code, st = oracle_state(raw, 0)
    one = [a.copy() for a in st]
    O.sweeps_ref_slabs(2, G, code, *st, 0, 3, 3)
    O.sweeps_ref_slabs(2, G, code, *one, 0, 3, 1)
    assert not np.array_equal(st[1], one[1])
    with pm.Solver(raw, boundary_method=0, solving_algorithm=2, z_slabs=3) as s:
        s.step(3)
        for name, a, b in zip("puvw", s.state_halo(), st):
            assert np.array_equal(a, b), f"{name} differs"


def test_reference_solve_channel_matches_binary_and_oracle(channel_raw, known):
    kn = known["channel_gs"]
    say("reference solve channel")
    G = O.geom_halo(channel_raw, 0)
    code = O.eval_geometry(G)
    t, rows = O.calc_flow(G, code, 0, O.init_fields(G, 0), solver=2, eps=1e-6, dx=5e-5)
    with pm.Solver(channel_raw, boundary_method=0, solving_algorithm=2, eps=1e-6, voxel_size=5e-5) as s:
        res, grows = s.solve()
    assert res.sweeps == kn["sweeps"] == t
    assert len(grows) == len(rows)
    # the oracle sums serially like the reference (error up to n*eps relative,
    # n = 10^4 fluid voxels); the GPU sums are compensated.  conv is a
    # relative difference of two such sums, so it is compared absolutely.
    for g, r in zip(grows, rows):
        assert g["iteration"] == r[0]
        assert g["conv"] == pytest.approx(r[1], rel=0, abs=1e-11)
        for key, col in (("wmax_velz", 3), ("wmean_velz", 4), ("wk13", 8), ("wk23", 9), ("wk33", 10)):
            assert g[key] == pytest.approx(r[col], rel=1e-11, abs=1e-300)
        assert "%e" % g["wk33"] == "%e" % r[10]
    assert "%e" % grows[-1]["wk33"] == kn["wk33"]


def test_reference_solve_channel_dom_interest(channel_raw, known):
    kn = known["channel_dom_interest_t1500"]
    say("reference dom_interest")
    with pm.Solver(channel_raw, boundary_method=0, solving_algorithm=2, eps=1e-6, voxel_size=5e-5,
                   dom_interest=tuple(kn["dom_interest"]), max_iter=1501) as s:
        _, rows = s.solve()
    r = [x for x in rows if x["iteration"] == kn["iteration"]][0]
    assert "%e" % r["k33"] == kn["k33"] and "%e" % r["wk33"] == kn["wk33"]


def test_reference_solve_tube_matches_binary(known):
    """Config 0 (54x54x50 tube) to eps, exact order: the reference binary's
    sweep count and k, and Hagen-Poiseuille within discretisation error."""
    raw = geo.config0()
    for case in ("tube_sor_ref", "tube_gs"):
        kn = known[case]
        say("reference tube", case)
        with pm.Solver(raw, boundary_method=0, solving_algorithm=kn["solver"], eps=1e-6,
                       voxel_size=2e-5) as s:
            res, rows = s.solve()
        assert res.sweeps == kn["sweeps"]
        assert "%e" % rows[-1]["wk33"] == kn["wk33"]
    assert "%e" % rows[-1]["wmean_velz"] == known["tube_gs"]["wmean_velz"]
    k_hp = geo.hagen_poiseuille_k(raw, 20, 2e-5, 50)
    assert abs(rows[-1]["wk33"] / k_hp - 1) < 0.03


# ---------------------------------------------------------------------------
# fast modes, per iteration
# ---------------------------------------------------------------------------
def oracle_fast(raw, bc, mode, omega, n, omega_p=None):
    G, code, A = oracle_state(raw, bc)
    pc = O.pack_code(code)
    B = [a.copy() for a in A]
    qa = O.outlet_plane(pc, bc, A)
    qb = qa.copy()
    pins = O.solution_pins(G)
    for _ in range(n):
        if mode == "jacobi":
            O.sweep_jacobi(pc, bc, omega, A, B, qa, qb, 3.0, omega_p=omega_p)
            O.ghost(G, bc, B, qb, *pins)
            A, B, qa, qb = B, A, qb, qa
        else:
            O.sweep_colour(pc, omega, A, qa, 3.0, omega_p=omega_p)
            O.ghost(G, bc, A, qa, *pins)
    return A


def oracle_dc(raw, bc, mode, omega, n, it_eval, omega_p=None):
    G, code, X = oracle_state(raw, bc)
    pc = O.pack_code(code)
    d = [np.zeros(G.shape, np.float32) for _ in range(4)]
    d2 = [np.zeros(G.shape, np.float32) for _ in range(4)]
    qd = np.zeros(G.shape[1:], np.float32)
    qd2 = qd.copy()
    R = O.residual(pc, X, O.outlet_plane(pc, bc, X))
    z = np.float32(0)
    for t in range(n):
        if mode == "jacobi":
            O.sweep_jacobi(pc, bc, omega, d, d2, qd, qd2, z, R=R, omega_p=omega_p)
            O.ghost(G, bc, d2, qd2, z, z, z)
            d, d2, qd, qd2 = d2, d, qd2, qd
        else:
            O.sweep_colour(pc, omega, d, qd, z, R=R, omega_p=omega_p)
            O.ghost(G, bc, d, qd, z, z, z)
        if t % it_eval == 0:
            O.fold(G, bc, X, d)
            R = O.residual(pc, X, O.outlet_plane(pc, bc, X))
            qd[:] = 0
            qd2[:] = 0
    return X


@pytest.mark.parametrize("mode", ["jacobi", "sor3"])
@pytest.mark.parametrize("bc", [0, 2, 4])
@pytest.mark.parametrize("omega,omega_p", [(1.7, 0.0), (2.0, 7.0)])
def test_fast_fp64_sweeps_bit_exact(mode, bc, omega, omega_p):
    raw = random_geometry((12, 10, 14), 0.35, 31 + bc)
    say("fast f64", mode, bc, omega, omega_p)
    ref = oracle_fast(raw, bc, mode, omega, 25, omega_p or None)
    with pm.Solver(raw, boundary_method=bc, mode=mode, omega=omega, omega_p=omega_p) as s:
        s.step(25)
        m = nonc(ref[0])
        for name, a, b in zip("puvw", s.state_halo(), ref):
            assert np.array_equal(a[m], b[m]), f"{name} differs"


@pytest.mark.parametrize("mode", ["jacobi", "sor3"])
@pytest.mark.parametrize("bc", [0, 1, 3])
@pytest.mark.parametrize("omega,omega_p", [(1.6, 0.0), (2.5, 6.0)])
def test_fast_f32dc_sweeps_bit_exact(mode, bc, omega, omega_p):
    raw = random_geometry((12, 10, 14), 0.35, 41 + bc)
    say("fast f32dc", mode, bc, omega, omega_p)
    ref = oracle_dc(raw, bc, mode, omega, 31, 10, omega_p or None)
    with pm.Solver(raw, boundary_method=bc, mode=mode, omega=omega, omega_p=omega_p, precision="f32dc",
                   it_eval=10) as s:
        s.step(31)
        m = nonc(ref[0])
        for name, a, b in zip("puvw", s.state_halo(), ref):
            assert np.array_equal(a[m], b[m]), f"{name} differs"


@pytest.mark.parametrize("layout", ["split", "split-all-loads", "natural", "compact", "compact-full", "compact-span0",
                                    "compact-span3", "compact-span0-fallback"])
@pytest.mark.parametrize("shape,bc", [((6, 5, 28), 0), ((5, 7, 27), 1), ((4, 6, 2), 2), ((7, 4, 60), 0),
                                      ((5, 6, 13), 3), ((9, 12, 130), 1), ((10, 9, 97), 0), ((6, 7, 20), 4)])
def test_sor_f32dc_state_layouts_bit_exact(layout, shape, bc):
    """FP32+DC 3-colour SOR with the correction state in colour-split rows
    (the default) and in natural rows: both equal the oracle bit for bit.
    nx 27, 28 and 60 are widths whose split rows need a wider row pitch than
    the natural ones; nx 2 gives blocks of two cells; nx 97 and 130 rows of
    several lines per block.  Split rows with the +-2 loads made only where
    the case reads them (default) and with every load; the compact state
    (fluid cells and periodic halo copies only, linked entries) with its
    links packed (default), in full, and packed with windows of 1 and 4
    (span 0, 3) so that most entries fall back to the full links, the records
    kept whatever the count (POREMAPS_PACK_KEEP=1) or, by default, given up
    for the link arrays when more than a quarter of the entries fall back."""
    raw = random_geometry(shape, 0.3, sum(shape) + bc)
    say("sor f32dc layout", layout, shape, bc)
    ref = oracle_dc(raw, bc, "sor3", 1.8, 16, 5, 6.0)
    split = layout.startswith("split")
    span = layout[len("compact-span"):].split("-")[0] if layout.startswith("compact-span") else "4095"
    keep = "0" if layout.endswith("fallback") else "1"
    with env("POREMAPS_SOR", "split" if split else "natural" if layout == "natural" else "compact"), \
            env("POREMAPS_SOR_SKIP2", "0" if layout == "split-all-loads" else "1"), \
            env("POREMAPS_SOR_LINKS", "full" if layout == "compact-full" else "packed"), \
            env("POREMAPS_PACK_SPAN", span), env("POREMAPS_PACK_KEEP", keep), \
            pm.Solver(raw, boundary_method=bc, mode="sor3", omega=1.8, omega_p=6.0, precision="f32dc",
                      it_eval=5) as s:
        want = {"split": "interleaved-colour-split-skip2", "split-all-loads": "interleaved-colour-split",
                "natural": "interleaved-colour", "compact-full": "interleaved-colour-compact"}.get(
                    layout, "interleaved-colour-compact-packed")
        name, tried = s.kernel_choice()
        if layout == "compact-span0-fallback":
            gave_up = tried["packed-far"] * 4 > int((raw == 0).sum())
            want = "interleaved-colour-compact" if gave_up else want
        assert name == want
        if layout == "compact-span0" and raw.size >= 500:
            # (the smallest boxes have at most three distinct offsets a slot)
            assert tried["packed-far"] > 0
        s.step(16)
        m = nonc(ref[0])
        for name, a, b in zip("puvw", s.state_halo(), ref):
            assert np.array_equal(a[m], b[m]), f"{name} differs"


@pytest.mark.parametrize("stage", ["fused", "seq"])
@pytest.mark.parametrize("chunk", [1, 3, 16])
@pytest.mark.parametrize("shape,bc", [((13, 9, 11), 0), ((40, 6, 7), 1), ((27, 5, 9), 2), ((7, 8, 6), 3),
                                      ((3, 6, 7), 4), ((64, 5, 4), 0)])
def test_sor_f32dc_z_chunks_bit_exact(stage, chunk, shape, bc):
    """FP32+DC 3-colour SOR in z-chunks (colour 1 trailing colour 0 by 3
    planes, colour 2 by 6), the three colours of a stage in one launch or one
    launch each: the oracle's three whole passes bit for bit, for chunks of
    one plane, of a few planes and of more planes than the domain holds."""
    raw = random_geometry(shape, 0.3, sum(shape) + bc + chunk)
    say("sor f32dc z-chunks", stage, chunk, shape, bc)
    ref = oracle_dc(raw, bc, "sor3", 1.8, 16, 5, 6.0)
    with env("POREMAPS_SOR_CHUNK", str(chunk)), env("POREMAPS_SOR_STAGE", stage), \
            pm.Solver(raw, boundary_method=bc, mode="sor3", omega=1.8, omega_p=6.0, precision="f32dc",
                      it_eval=5) as s:
        assert s.kernel_choice()[0] == f"interleaved-colour-split-z{chunk}" + ("-seq" if stage == "seq" else "") + \
            "-skip2"
        s.step(16)
        m = nonc(ref[0])
        for name, a, b in zip("puvw", s.state_halo(), ref):
            assert np.array_equal(a[m], b[m]), f"{name} differs"


@pytest.mark.parametrize("halos", ["fused", "separate", "march4", "march8", "occ6", "occ8", "skip2", "skipall",
                                   "compact", "compact-skipall", "compact-aos", "compact-full", "compact-span0"])
@pytest.mark.parametrize("shape,bc", [((2, 5, 6), 0), ((3, 3, 7), 1), ((6, 2, 3), 0), ((5, 6, 7), 2),
                                      ((4, 5, 6), 4), ((7, 6, 9), 3), ((9, 8, 40), 0), ((11, 13, 70), 0),
                                      ((6, 9, 130), 1)])
def test_jacobi_f32dc_record_pass_halo_fill_bit_exact(halos, shape, bc):
    """The FP32+DC record pass writing its cells' periodic halo copies and
    halo pins itself (default) and followed by the separate ghost fill: both
    equal the oracle bit for bit.  Sides of 2 and 3 give cells with two
    copies on one axis (planes 2 and nz both, or 3 and nz+1).  The compact
    state (no halo copies: links across a periodic face go to the image) with
    every load and with each load only where used, the lagged pass reading
    packed records of its twelve links (default; span0: windows of one value,
    so that most entries fall back to the link arrays), the link arrays
    (full) or one record per entry (aos)."""
    raw = random_geometry(shape, 0.3, 3 * sum(shape) + bc)
    say("record pass halos", halos, shape, bc)
    ref = oracle_dc(raw, bc, "jacobi", 2.5, 16, 5, 6.0)
    march = halos.startswith("march")
    compact = halos.startswith("compact")
    occ = halos[3:] if halos.startswith("occ") else "0"
    form = "march" if march else "compact" if compact else "records"
    skip = {"skip2": "1", "skipall": "2", "compact-skipall": "2", "compact-aos": "2", "compact-full": "2",
            "compact-span0": "2"}.get(halos, "0")
    links = {"compact-aos": "aos", "compact-full": "full"}.get(halos, "packed")
    with env("POREMAPS_HALOS", "separate" if halos == "separate" else "fused"), env("POREMAPS_JACOBI", form), \
            env("POREMAPS_JAC_LINKS", links), env("POREMAPS_PACK_SPAN", "0" if halos == "compact-span0" else "4095"), \
            env("POREMAPS_PACK_KEEP", "1"), \
            env("POREMAPS_MARCH_TY", halos[5:] if march else "4"), env("POREMAPS_MARCH_ZSEG", "3"), \
            env("POREMAPS_REC_OCC", occ), env("POREMAPS_REC_SKIP2", skip), \
            pm.Solver(raw, boundary_method=bc, mode="jacobi", omega=2.5, omega_p=6.0, precision="f32dc",
                      it_eval=5) as s:
        s.step(16)
        name, tried = s.kernel_choice()
        if compact:
            want = "interleaved-compact" + ("-packed" if links == "packed" else "") + ("-skipall" if skip == "2" else "")
            assert name == want, name
            if halos == "compact-span0" and raw.size >= 500:
                assert tried["packed-far"] > 0
        else:
            want = f"interleaved-zmarch-ty{halos[5:]}-zs3" if march else "interleaved-lagged-records"
            assert name == want + (f"-occ{occ}" if occ != "0" else "") + \
                {"skip2": "-skip2", "skipall": "-skipall"}.get(halos, ""), name
        m = nonc(ref[0])
        for name, a, b in zip("puvw", s.state_halo(), ref):
            assert np.array_equal(a[m], b[m]), f"{name} differs"


# ---------------------------------------------------------------------------
# edge shapes: the smallest domains, thin ones, all fluid, all solid
# ---------------------------------------------------------------------------
EDGE_SHAPES = [
    ("minimal all-fluid 2x2x2", lambda: np.zeros((2, 2, 2), np.uint8)),
    ("thin random 2x3x9", lambda: random_geometry((2, 3, 9), 0.3, 5)),
    ("all-fluid 5x4x6", lambda: np.zeros((5, 4, 6), np.uint8)),
    ("thin random 3x7x2", lambda: random_geometry((3, 7, 2), 0.3, 6)),
]


@pytest.mark.parametrize("name,make", EDGE_SHAPES, ids=[e[0] for e in EDGE_SHAPES])
@pytest.mark.parametrize("bc", [0, 2, 4])
def test_edge_shapes_reference_bit_exact(name, make, bc):
    raw = make()
    say("edge reference", name, bc)
    G, code, st = oracle_state(raw, bc)
    with pm.Solver(raw, boundary_method=bc, solving_algorithm=2) as s:
        for _ in range(7):
            O.sweep_ref(2, G, code, *st)
            for f in st:
                O.exchange(f, bc)
            O.pins(G, st[0])
        s.step(7)
        for n_, a, b in zip("puvw", s.state_halo(), st):
            assert np.array_equal(a, b), f"{n_} differs ({name}, bc {bc})"


@pytest.mark.parametrize("name,make", EDGE_SHAPES, ids=[e[0] for e in EDGE_SHAPES])
@pytest.mark.parametrize("mode,prec", [("jacobi", "f64"), ("jacobi", "f32dc"), ("sor3", "f32dc")])
@pytest.mark.parametrize("bc", [0, 1, 3])
def test_edge_shapes_fast_modes_bit_exact(name, make, mode, prec, bc):
    raw = make()
    say("edge fast", name, mode, prec, bc)
    # FP32+DC: 11 sweeps with a fold every 5, so that the last sweep (index
    # 10) ends with a fold and both sides hold the FP64 state
    n = 13 if prec == "f64" else 11
    if prec == "f64":
        ref = oracle_fast(raw, bc, mode, 2.0, n, 5.0)
        kw = {}
    else:
        ref = oracle_dc(raw, bc, mode, 2.0, n, 5, 5.0)
        kw = dict(precision="f32dc", it_eval=5)
    with pm.Solver(raw, boundary_method=bc, mode=mode, omega=2.0, omega_p=5.0, **kw) as s:
        s.step(n)
        m = nonc(ref[0])
        for n_, a, b in zip("puvw", s.state_halo(), ref):
            assert np.array_equal(a[m], b[m]), f"{n_} differs ({name}, {mode} {prec}, bc {bc})"


@pytest.mark.parametrize("mode,prec", [("reference", "f64"), ("jacobi", "f64"), ("jacobi", "f32dc"),
                                       ("sor3", "f32dc")])
def test_all_solid_medium_stops_at_first_convergence_test(mode, prec):
    """No fluid: the convergence value is 0/0 at the first test (sweep 1100);
    the reference's loop condition is false for it and the run ends there.
    The solver ends at the same point with PM_E_DIVERGED and no restart."""
    raw = np.ones((6, 5, 7), np.uint8)
    say("all solid", mode, prec)
    kw = {} if mode == "reference" else dict(mode=mode, precision=prec, omega=2.5, omega_p=7.0)
    with pm.Solver(raw, boundary_method=0, eps=1e-6, max_iter=50_000, **kw) as s:
        with pytest.raises(pm.PoremapsError) as e:
            s.solve()
        assert e.value.code == -7 and "iteration 1100" in str(e.value)
        assert s.porosity == 0.0


# ---------------------------------------------------------------------------
# fast modes, to convergence: the reference's fixed point
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode,omega,prec,omega_p", [("jacobi", 1.0, "f64", 0.0), ("jacobi", 2.5, "f32dc", 0.0),
                                                     ("jacobi", 2.5, "f32dc", 6.0), ("jacobi", 2.0, "f64", 8.0),
                                                     ("sor3", 1.8, "f64", 0.0), ("sor3", 1.8, "f32dc", 0.0)])
def test_fast_modes_reach_reference_fixed_point(channel_raw, mode, omega, prec, omega_p):
    say("fast fixed point", mode, omega, prec, omega_p)
    G = O.geom_halo(channel_raw, 0)
    code = O.eval_geometry(G)
    _, tight = O.calc_flow(G, code, 0, O.init_fields(G, 0), solver=2, eps=1e-11, dx=5e-5)
    _, at_eps = O.calc_flow(G, code, 0, O.init_fields(G, 0), solver=2, eps=1e-6, dx=5e-5)
    with pm.Solver(channel_raw, boundary_method=0, mode=mode, omega=omega, precision=prec, omega_p=omega_p,
                   eps=1e-11, voxel_size=5e-5) as s:
        res, _ = s.solve()
    tol = 1e-9 if prec == "f64" else 1e-6
    assert res.last.wk33 == pytest.approx(tight[-1][10], rel=tol)
    with pm.Solver(channel_raw, boundary_method=0, mode=mode, omega=omega, precision=prec, omega_p=omega_p,
                   eps=1e-6, voxel_size=5e-5) as s:
        res, _ = s.solve()
    assert res.last.wk33 == pytest.approx(at_eps[-1][10], rel=1e-4)


# ---------------------------------------------------------------------------
# the POREMAPS command line
# ---------------------------------------------------------------------------
def test_cli_drop_in_channel(tmp_path, channel_raw):
    say("cli channel")
    (tmp_path / "rect_channel_24_14_50_vs_5e-05.raw").write_bytes(channel_raw.tobytes())
    inp = open(os.path.join(GOLDEN, "input_rect_channel_24_14_50_vs_5e-05.inp")).read()
    inp = inp.replace("eps 1e-07", "eps 1e-06").replace("porosity 1.0", "porosity -1.0")
    (tmp_path / "input.inp").write_text(inp)
    r = subprocess.run([pm.CLI_PATH, "input.inp"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Iteration: 9300\t Convergence:" in r.stdout
    G = O.geom_halo(channel_raw, 0)
    code = O.eval_geometry(G)
    st = O.init_fields(G, 0)
    t, rows = O.calc_flow(G, code, 0, st, solver=2, eps=1e-6, dx=5e-5)
    geom = "rect_channel_24_14_50_vs_5e-05.raw"
    log = (tmp_path / "permeability_rect_channel_24_14_50_vs_5e-05.log").read_text().splitlines()
    assert log[0].startswith("# iteration, conv, TPS")
    assert len(log) == 1 + len(rows)
    for line, row in zip(log[1:], rows):
        f = line.split(", ")
        exp = ["%i" % row[0]] + ["%e" % x for x in row[1:]]
        assert [f[0], f[1]] + f[3:] == [exp[0], exp[1]] + exp[3:]
    nh = np.fromfile(tmp_path / ("voxel_neighborhood_" + geom), np.int32).reshape(50, 14, 24)
    assert np.array_equal(nh, code[2:-2, 2:-2, 2:-2])
    p = np.fromfile(tmp_path / ("press_" + geom), np.float64).reshape(50, 14, 24)
    assert np.array_equal(p, st[0][2:-2, 2:-2, 2:-2])
    for name, f in zip(("velx_", "vely_", "velz_"), st[1:]):
        v = np.fromfile(tmp_path / (name + geom), np.float64).reshape(50, 14, 24)
        assert np.array_equal(v, f[2:-2, 2:-2, 2:-2] * 5e-5)
    dd = np.fromfile(tmp_path / ("domain_decomp_" + geom), np.int32)
    assert dd.size == 24 * 14 * 50 and (dd == 0).all()


def test_cli_fast_mode_default_relaxation(tmp_path, channel_raw):
    """--mode=jacobi --precision=f32dc without --omega: the pair chosen from
    the porosity, the same solve as the library's with omega=AUTO (sweep
    count, last log row and velocity field)."""
    say("cli fast mode")
    (tmp_path / "rect_channel_24_14_50_vs_5e-05.raw").write_bytes(channel_raw.tobytes())
    inp = open(os.path.join(GOLDEN, "input_rect_channel_24_14_50_vs_5e-05.inp")).read()
    inp = inp.replace("eps 1e-07", "eps 1e-06").replace("porosity 1.0", "porosity -1.0")
    (tmp_path / "input.inp").write_text(inp)
    r = subprocess.run([pm.CLI_PATH, "input.inp", "--mode=jacobi", "--precision=f32dc"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    with pm.Solver(channel_raw, boundary_method=0, voxel_size=5e-5, mode="jacobi", precision="f32dc",
                   omega=pm.AUTO, eps=1e-6) as s:
        phi = s.porosity
        assert s.relaxation() == ((0.6, 30.0) if phi < 0.2 else (2.5, 7.0))
        res, rows = s.solve()
        velz = s.field("velz")
    log = (tmp_path / "permeability_rect_channel_24_14_50_vs_5e-05.log").read_text().splitlines()
    last = log[-1].split(", ")
    assert int(last[0]) == rows[-1]["iteration"]
    assert last[-1] == "%e" % rows[-1]["wk33"]
    v = np.fromfile(tmp_path / "velz_rect_channel_24_14_50_vs_5e-05.raw", np.float64).reshape(50, 14, 24)
    assert np.array_equal(v, velz * 5e-5)  # the files hold velocities times the voxel size


def test_cli_dom_decomposition_z_slabs(tmp_path, channel_raw):
    """dom_decomposition 1 1 3 in the .inp: the CLI sweeps as `mpirun -np 3`
    does (three z-slabs), the same solve as the library's z_slabs=3, and
    writes each voxel's rank into domain_decomp_."""
    say("cli z-slabs")
    (tmp_path / "rect_channel_24_14_50_vs_5e-05.raw").write_bytes(channel_raw.tobytes())
    inp = open(os.path.join(GOLDEN, "input_rect_channel_24_14_50_vs_5e-05.inp")).read()
    inp = inp.replace("eps 1e-07", "eps 1e-06").replace("porosity 1.0", "porosity -1.0")
    inp = inp.replace("dom_decomposition 0 0 0", "dom_decomposition 1 1 3")
    (tmp_path / "input.inp").write_text(inp)
    r = subprocess.run([pm.CLI_PATH, "input.inp"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "its blocks reproduced" in r.stdout
    # each voxel's rank (src/output.cc:156-190): slabs of 17, 17, 16 planes
    dd = np.fromfile(tmp_path / "domain_decomp_rect_channel_24_14_50_vs_5e-05.raw", np.int32).reshape(50, 14, 24)
    assert np.array_equal(dd, pm.rank_map((50, 14, 24), (1, 1, 3)))
    assert (dd[:17] == 0).all() and (dd[17:34] == 1).all() and (dd[34:] == 2).all()
    with pm.Solver(channel_raw, boundary_method=0, solving_algorithm=2, eps=1e-6, voxel_size=5e-5,
                   z_slabs=3) as s:
        res, rows = s.solve()
        velz = s.field("velz")
    log = (tmp_path / "permeability_rect_channel_24_14_50_vs_5e-05.log").read_text().splitlines()
    last = log[-1].split(", ")
    assert int(last[0]) == rows[-1]["iteration"]
    assert last[-1] == "%e" % rows[-1]["wk33"]
    v = np.fromfile(tmp_path / "velz_rect_channel_24_14_50_vs_5e-05.raw", np.float64).reshape(50, 14, 24)
    assert np.array_equal(v, velz * 5e-5)

