"""Opt-in no-slip domain walls (pm_params.walls = 1) on the GPU.

The HIP path with walls against the oracle with walls (codes, initial state,
sweeps and whole solves bit for bit), and against the reference binary's
channel run through the fully fluid 20 x 10 x 50 box: the same 9301
Gauss-Seidel sweeps and wk33 = 8.368958e-09 once scaled by the channel's
porosity (the box's fluid sums are the channel's with its solid zeros left
out).  The command line's --walls=noslip gives the library's solve.
"""
import subprocess

import numpy as np
import pytest

import oracle as O
import poremaps_amd as pm
from poremaps_amd import geometry as geo
from test_gpu_parity import nonc, say

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
