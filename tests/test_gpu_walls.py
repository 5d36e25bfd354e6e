"""Opt-in no-slip domain walls (pm_params.walls = 1) on the GPU.

The HIP path with walls against the oracle with walls (codes, initial state,
sweeps and whole solves bit for bit), and against the reference binary's
channel run through the fully fluid 20 x 10 x 50 box: the same 9301
Gauss-Seidel sweeps and wk33 = 8.368958e-09 once scaled by the channel's
porosity (the box's fluid sums are the channel's with its solid zeros left
out).  The command line's --walls=noslip gives the library's solve.
"""
import contextlib
import subprocess

import numpy as np
import pytest

import oracle as O
import poremaps_amd as pm
from poremaps_amd import geometry as geo
from test_gpu_parity import env, nonc, say
from test_walls import plates

pytestmark = pytest.mark.gpu
INNER = (slice(2, -2),) * 3
PHI_CHANNEL = 200 / 336


def box():
    return np.zeros((50, 10, 20), np.uint8)


@pytest.mark.parametrize("bc", range(5))
def test_walls_geometry_codes_and_initial_state(bc):
    raw = geo.sphere_pack((12, 9, 11), 0.35, 2, 3, seed=bc)
    say("walls geometry", bc)
    G = O.geom_halo(raw, bc, walls=1)
    code = O.eval_geometry(G)
    st = [a.copy() for a in O.init_fields(G, bc)]
    O.pins(G, st[0])
    with pm.Solver(raw, boundary_method=bc, walls=1) as s:
        assert np.array_equal(s.field("nhood"), code[INNER])
        assert s.porosity == O.porosity(G)
        for a, b in zip(s.state_halo(), st):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("bc", [2, 3])
@pytest.mark.parametrize("mode,precision", [("jacobi", "f64"), ("sor3", "f64"), ("jacobi", "f32dc"),
                                            ("sor3", "f32dc")])
def test_walls_fast_solves_match_oracle(bc, mode, precision):
    raw = geo.sphere_pack((18, 14, 16), 0.3, 2, 3, seed=3 + bc)
    pair = (1.8, 6.0) if mode == "sor3" else (2.5, 7.0)
    say("walls fast solve", bc, mode, precision)
    t, _, st, G = O.solve_fast(raw, bc, mode, precision, *pair, walls=1)
    k = O.permeability(G, st[1], st[2], st[3], st[0], (0,) * 6, 1e-5, O.porosity(G))
    with pm.Solver(raw, boundary_method=bc, voxel_size=1e-5, mode=mode, precision=precision, omega=pair[0],
                   omega_p=pair[1], eps=1e-6, walls=1) as s:
        res, _ = s.solve()
        halo = s.state_halo()
    assert res.sweeps == t
    m = nonc(st[0])
    for name, a, b in zip("puvw", halo, st):
        assert np.array_equal(a[m], b[m]), name
    assert res.last.wk33 == pytest.approx(k[5], rel=1e-12)


@pytest.mark.parametrize("solver", [1, 2, 3])
def test_walls_reference_sweeps_bit_exact(solver):
    raw = geo.sphere_pack((13, 9, 11), 0.35, 2, 3, seed=solver)
    say("walls reference sweeps", solver)
    G = O.geom_halo(raw, 3, walls=1)
    code = O.eval_geometry(G)
    st = [a.copy() for a in O.init_fields(G, 3)]
    O.pins(G, st[0])
    for _ in range(20):
        O.sweep_ref(solver, G, code, *st)
        for f in st:
            O.exchange(f, 3)
        O.pins(G, st[0])
    with pm.Solver(raw, boundary_method=3, solving_algorithm=solver, walls=1) as s:
        s.step(20)
        m = nonc(st[0])
        for name, a, b in zip("puvw", s.state_halo(), st):
            assert np.array_equal(a[m], b[m]), name


def test_walled_box_reference_solve_is_the_binary_channel_run(known):
    kn = known["channel_gs"]
    say("walls box reference solve")
    with pm.Solver(box(), boundary_method=2, solving_algorithm=2, eps=1e-6, voxel_size=5e-5, walls=1) as s:
        res, rows = s.solve()
    assert res.sweeps == kn["sweeps"]
    assert "%e" % (rows[-1]["wk33"] * PHI_CHANNEL) == kn["wk33"]
    assert abs(rows[-1]["wk33"] / geo.duct_k(20, 10, 5e-5, 50, 1.0) - 1) < 0.01


def test_cli_walls_noslip(tmp_path):
    say("cli walls")
    (tmp_path / "box.raw").write_bytes(box().tobytes())
    (tmp_path / "input.inp").write_text(
        "dom_decomposition 0 0 0\nboundary_method 2\ngeometry_file_name box.raw\nsize_x_y_z  20 10 50\n"
        "voxel_size  5e-05\nmax_iter    100000001\nit_eval    100\nit_write   100\n"
        "log_file_name permeability_box.log\nsolving_algorithm 2\neps 1e-06\nporosity -1.0\n"
        "dom_interest 0 0 0 0 0 0\nwrite_output 1 0 0 0\n")
    r = subprocess.run([pm.CLI_PATH, "input.inp", "--mode=jacobi", "--precision=f32dc", "--walls=noslip"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    with pm.Solver(box(), boundary_method=2, voxel_size=5e-5, mode="jacobi", precision="f32dc",
                   omega=pm.AUTO, eps=1e-6, walls=1) as s:
        res, rows = s.solve()
        velz = s.field("velz")
    last = (tmp_path / "permeability_box.log").read_text().splitlines()[-1].split(", ")
    assert int(last[0]) == rows[-1]["iteration"] and last[-1] == "%e" % rows[-1]["wk33"]
    v = np.fromfile(tmp_path / "velz_box.raw", np.float64).reshape(50, 10, 20)
    assert np.array_equal(v, velz * 5e-5)
    r = subprocess.run([pm.CLI_PATH, "input.inp", "--walls=closed"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "walls" in r.stderr


# ---------------------------------------------------------------------------
# slip walls (walls = 2): the fast modes against the oracle's slip walls
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("bc", range(5))
def test_slip_geometry_codes_and_initial_state(bc):
    """The halo of a slip face holds the mirror image of the geometry."""
    raw = geo.sphere_pack((12, 9, 11), 0.35, 2, 3, seed=bc)
    say("slip geometry", bc)
    G = O.geom_halo(raw, bc, walls=2)
    code = O.eval_geometry(G)
    st = [a.copy() for a in O.init_fields(G, bc)]
    O.pins(G, st[0])
    with pm.Solver(raw, boundary_method=bc, mode="jacobi", walls=2) as s:
        assert np.array_equal(s.field("nhood"), code[INNER])
        assert s.porosity == O.porosity(G)
        for a, b in zip(s.state_halo(), st):
            assert np.array_equal(a, b)


SLIP_FORMS = {
    # (mode, precision): {form: (environment knobs, kernel choice)}
    ("jacobi", "f32dc"): {
        "records-skip2": ({"POREMAPS_JACOBI": "records", "POREMAPS_REC_SKIP2": "1"},
                          "interleaved-lagged-records-skip2"),
        "records": ({"POREMAPS_JACOBI": "records", "POREMAPS_REC_SKIP2": "0"}, "interleaved-lagged-records"),
        "records-separate-halos": ({"POREMAPS_JACOBI": "records", "POREMAPS_REC_SKIP2": "0",
                                    "POREMAPS_HALOS": "separate"}, "interleaved-lagged-records"),
        "pair": ({"POREMAPS_JACOBI": "pair"}, "interleaved-pair"),
        "list": ({"POREMAPS_JACOBI": "list", "POREMAPS_SWEEP_LOADS": "branch"}, "interleaved-lagged-branch"),
        # forms that do not hold slip walls fall back to the record pass
        "march->records": ({"POREMAPS_JACOBI": "march", "POREMAPS_REC_SKIP2": "2"},
                           "interleaved-lagged-records-skip2"),
        "tuned": ({}, None),
    },
    ("jacobi", "f64"): {
        "records": ({"POREMAPS_JACOBI": "records"}, "list-lagged-records"),
        "pair": ({"POREMAPS_JACOBI": "pair"}, "list-pair-select"),
        "list": ({"POREMAPS_JACOBI": "list", "POREMAPS_SWEEP_LOADS": "branch"}, "list-lagged-branch"),
        "stream->records": ({"POREMAPS_JACOBI": "stream"}, "list-lagged-records"),
        "tuned": ({}, None),
    },
    ("sor3", "f32dc"): {
        "split-skip2": ({"POREMAPS_SOR": "split", "POREMAPS_SOR_SKIP2": "1"}, "interleaved-colour-split-skip2"),
        "split": ({"POREMAPS_SOR": "split", "POREMAPS_SOR_SKIP2": "0"}, "interleaved-colour-split"),
        "split-z3": ({"POREMAPS_SOR": "split", "POREMAPS_SOR_CHUNK": "3"}, "interleaved-colour-split-z3-skip2"),
        "natural": ({"POREMAPS_SOR": "natural"}, "interleaved-colour"),
        # the compact state does not hold slip walls: colour-split rows
        "compact->split": ({"POREMAPS_SOR": "compact"}, "interleaved-colour-split-skip2"),
    },
    ("sor3", "f64"): {"default": ({}, None)},
}


@contextlib.contextmanager
def knobs(d):
    with contextlib.ExitStack() as es:
        for k, v in d.items():
            es.enter_context(env(k, v))
        yield


@pytest.mark.parametrize("bc", [2, 3])
@pytest.mark.parametrize("mode,precision,form", [(m, p, f) for (m, p), fs in SLIP_FORMS.items() for f in fs])
def test_slip_fast_solves_match_oracle(bc, mode, precision, form):
    """Every kernel form of the fast modes with slip walls: the whole solve
    (sweep count, k, state) equals the oracle's slip-wall solve bit for bit."""
    raw = geo.sphere_pack((18, 14, 16), 0.3, 2, 3, seed=7 + bc)
    pair = (1.8, 6.0) if mode == "sor3" else (2.5, 7.0)
    say("slip fast solve", bc, mode, precision, form)
    t, _, st, G = O.solve_fast(raw, bc, mode, precision, *pair, walls=2)
    k = O.permeability(G, st[1], st[2], st[3], st[0], (0,) * 6, 1e-5, O.porosity(G))
    kn, want = SLIP_FORMS[(mode, precision)][form]
    with knobs(kn), pm.Solver(raw, boundary_method=bc, voxel_size=1e-5, mode=mode, precision=precision,
                              omega=pair[0], omega_p=pair[1], eps=1e-6, walls=2) as s:
        res, _ = s.solve()
        if want is not None:
            assert s.kernel_choice()[0] == want
        halo = s.state_halo()
    assert res.sweeps == t
    m = nonc(st[0])
    for name, a, b in zip("puvw", halo, st):
        assert np.array_equal(a[m], b[m]), name
    assert res.last.wk33 == pytest.approx(k[5], rel=1e-12)


@pytest.mark.parametrize("axis", ["y", "x"])
@pytest.mark.parametrize("mode,precision", [("jacobi", "f32dc"), ("sor3", "f64")])
def test_slip_plates_are_plane_poiseuille(axis, mode, precision):
    """Plates 10 voxels apart, the open faces slip walls: the discrete plane
    Poiseuille permeability (h^2 + 1/2) / 12 * (nz - 1) / nz in voxel units."""
    raw = plates(axis)
    nz, h = raw.shape[0], 10
    pair = (1.8, 6.0) if mode == "sor3" else (2.5, 7.0)
    say("slip plates", axis, mode, precision)
    with pm.Solver(raw, boundary_method=2, voxel_size=1.0, mode=mode, precision=precision, omega=pair[0],
                   omega_p=pair[1], eps=1e-10, walls=2) as s:
        s.solve()
        velz = s.field("velz")
    # the convergence test's permeability: the mean fluid w over Re = 0.01
    k = velz[raw == 0].mean() / 0.01
    assert k == pytest.approx((h * h + 0.5) / 12 * (nz - 1) / nz, rel=1e-8)


def test_slip_reference_mode_and_unknown_walls_rejected():
    raw = plates("y")
    say("slip rejected")
    for kw in ({"walls": 2}, {"walls": 3, "mode": "jacobi"}, {"walls": -1, "mode": "jacobi"}):
        with pytest.raises(pm.PoremapsError) as e:
            pm.Solver(raw, boundary_method=2, **kw)
        assert e.value.code == -1, kw


def test_cli_walls_slip(tmp_path):
    say("cli walls slip")
    raw = plates("y")
    nz, ny, nx = raw.shape
    (tmp_path / "pl.raw").write_bytes(raw.tobytes())
    (tmp_path / "input.inp").write_text(
        f"dom_decomposition 0 0 0\nboundary_method 2\ngeometry_file_name pl.raw\nsize_x_y_z  {nx} {ny} {nz}\n"
        "voxel_size  1e-05\nmax_iter    100000001\nit_eval    100\nit_write   100\n"
        "log_file_name permeability_pl.log\nsolving_algorithm 2\neps 1e-06\nporosity -1.0\n"
        "dom_interest 0 0 0 0 0 0\nwrite_output 1 0 0 0\n")
    r = subprocess.run([pm.CLI_PATH, "input.inp", "--mode=jacobi", "--precision=f32dc", "--walls=slip"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    with pm.Solver(raw, boundary_method=2, voxel_size=1e-5, mode="jacobi", precision="f32dc",
                   omega=pm.AUTO, eps=1e-6, walls=2) as s:
        res, rows = s.solve()
        velz = s.field("velz")
    last = (tmp_path / "permeability_pl.log").read_text().splitlines()[-1].split(", ")
    assert int(last[0]) == rows[-1]["iteration"] and last[-1] == "%e" % rows[-1]["wk33"]
    v = np.fromfile(tmp_path / "velz_pl.raw", np.float64).reshape(nz, ny, nx)
    assert np.array_equal(v, velz * 1e-5)
