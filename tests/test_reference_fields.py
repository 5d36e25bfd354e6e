"""Pins the CPU oracle against what the reference binary itself computed.

The fixtures (tests/golden/reference_fields.npz / .json) were written by
tests/golden/make_reference_fields.py, which runs the reference POREMAPS
compiled from its own sources (oracle/Makefile, target ref) under mpiexec:

* per-sweep fields after 1, 7 and 1101 sweeps for solvers 1, 2, 3 and
  boundary_method 0..4 on one rank -- the oracle's reference sweep
  (src/solver.cc:112-223 with the exchange and pins of src/calc_flow.cc:
  109-117) must reproduce them bit for bit.  Solvers 1 and 2 differ already
  after one sweep (the divergence taken before or after the velocity update,
  src/solver.cc:187-189 vs :216-218), so a solver-1 mistake fails here;
* fields after 7 and 1101 sweeps under `mpiexec -n P*Q*R` with
  dom_decomposition P Q R (z-slabs down to two planes, x and y splits,
  blocks) -- the oracle's block sweep must reproduce them bit for bit;
* the stop rule: sweep count and log row at eps 1e-6 (src/calc_flow.cc:
  101-128), on the box medium and on the shipped channel, one rank and
  decomposed; the log rows of the channel at t = 3000 under five
  decompositions (SURVEY finding F7: 8.192931e-09 one rank, 8.192801e-09 with
  1 1 2, 8.192311e-09 with 2 1 1).
"""
import numpy as np
import pytest

import oracle as O
from conftest import ref_run

INTERIOR = (slice(2, -2),) * 3


def _box_tags():
    return [f"box_s{s}_bc{bc}_n{n}" for bc in range(5) for s in (1, 2, 3) for n in (1, 7, 1101)]


def _dec_tags():
    layouts = ["112", "113", "114", "119", "211", "121", "311", "221", "212", "222", "322"]
    tags = [f"dec_{d}_s2_bc{bc}_n7" for d in layouts for bc in (0, 1, 3)]
    tags += ["dec_222_s1_bc0_n7", "dec_222_s3_bc0_n7", "dec_112_s2_bc0_n1101", "dec_222_s2_bc1_n1101"]
    tags += [f"dec_111_s2_bc{bc}_n7" for bc in (0, 1, 3)]
    return tags


def fmt_row(r):
    """A log row as the reference prints it (src/output.cc:216-219), TPS left
    out (it is a timing)."""
    it, conv, _tps, *vals = r
    return ", ".join([f"{int(it)}", f"{conv:e}"] + [f"{x:e}" for x in vals])


def drop_tps(line):
    parts = [s.strip() for s in line.split(",")]
    return ", ".join(parts[:2] + parts[3:])


def _initial(raw, bc):
    G = O.geom_halo(raw, bc)
    code = O.eval_geometry(G)
    st = [a.copy() for a in O.init_fields(G, bc)]
    O.pins(G, st[0])
    return G, code, st


@pytest.mark.parametrize("tag", _box_tags())
def test_reference_sweep_matches_binary(ref_fields, ref_meta, tag):
    r = ref_run(ref_meta, "box", tag)
    G, code, st = _initial(ref_fields["box_raw"], r["bc"])
    for _ in range(r["sweeps"]):
        O.sweep_ref(r["solver"], G, code, *st)
        for f in st:
            O.exchange(f, r["bc"])
        O.pins(G, st[0])
    for name, a in zip("puvw", st):
        assert np.array_equal(a[INTERIOR], ref_fields[f"{tag}_{name}"]), f"{tag}: {name}"


@pytest.mark.parametrize("bc", range(5))
def test_neighbourhood_codes_match_binary(ref_fields, bc):
    G = O.geom_halo(ref_fields["box_raw"], bc)
    assert np.array_equal(O.eval_geometry(G)[INTERIOR], ref_fields[f"box_bc{bc}_nhood"])


def test_solver_1_and_2_differ_after_one_sweep(ref_fields):
    # the fixture tells them apart, so the sweep test above pins solver 1
    for bc in range(5):
        a, b = ref_fields[f"box_s1_bc{bc}_n1_p"], ref_fields[f"box_s2_bc{bc}_n1_p"]
        assert not np.array_equal(a, b)


@pytest.mark.parametrize("solver", [1, 2, 3])
@pytest.mark.parametrize("bc", [0, 3])
def test_log_row_after_1101_sweeps_matches_binary(ref_fields, ref_meta, solver, bc):
    # first evaluation at t = 1100 (conv = 1 from k_prev = 0), voxel size 1
    r = ref_run(ref_meta, "box", f"box_s{solver}_bc{bc}_n1101")
    G, code, st = _initial(ref_fields["box_raw"], bc)
    n, rows = O.calc_flow(G, code, bc, st, solver=solver, max_iter=1101, dx=1.0)
    assert n == 1101
    want = [drop_tps(x) for x in r["log"] if not x.startswith("#")]
    assert [fmt_row(x) for x in rows] == want


@pytest.mark.parametrize("tag", _dec_tags())
def test_block_sweeps_match_binary(ref_fields, ref_meta, tag):
    r = ref_run(ref_meta, "decomp", tag)
    G, code, st = _initial(ref_fields["decomp_raw"], r["bc"])
    O.sweeps_ref_blocks(r["solver"], G, code, *st, r["bc"], r["sweeps"], r["dims"])
    for name, a in zip("puvw", st):
        assert np.array_equal(a[INTERIOR], ref_fields[f"{tag}_{name}"]), f"{tag}: {name}"


@pytest.mark.parametrize("tag", ["dec_119_s2_bc0_n7", "dec_222_s2_bc1_n7", "dec_311_s2_bc3_n7",
                                 "dec_121_s2_bc0_n7"])
def test_block_sweeps_compose_across_calls(ref_fields, ref_meta, tag):
    # the outlet plane before its pins is part of the state: halo copies of
    # it in other blocks (x / y splits of the top layer) hold the swept value
    r = ref_run(ref_meta, "decomp", tag)
    G, code, st = _initial(ref_fields["decomp_raw"], r["bc"])
    outlet = np.ascontiguousarray(st[0][-3]).copy()
    for n in (2, 1, 4):
        O.sweeps_ref_blocks(r["solver"], G, code, *st, r["bc"], n, r["dims"], outlet=outlet)
    for name, a in zip("puvw", st):
        assert np.array_equal(a[INTERIOR], ref_fields[f"{tag}_{name}"]), f"{tag}: {name}"


@pytest.mark.parametrize("tag", ["dec_211_s2_bc0_n7", "dec_121_s2_bc3_n7", "dec_222_s2_bc1_n7"])
def test_block_sweeps_without_the_outlet_plane_would_differ(ref_fields, ref_meta, tag):
    # a plausible mistake -- resuming with the pinned outlet plane in the
    # blocks' halo copies -- is caught: across an x or y split of the top
    # layer a block reads its neighbour's outlet-plane pressure (the -x / -y
    # neighbour), which the reference holds as swept, before the pins.  (A
    # lower z-slab never reads its +z halo pressure.)
    r = ref_run(ref_meta, "decomp", tag)
    G, code, st = _initial(ref_fields["decomp_raw"], r["bc"])
    for n in (2, 5):
        O.sweeps_ref_blocks(2, G, code, *st, r["bc"], n, r["dims"])
    assert not np.array_equal(st[0][INTERIOR], ref_fields[tag + "_p"])


def test_rank_map_matches_binary(ref_fields, ref_meta):
    # get_dom_limits (src/geometry.cc:129-171): the first size % dims blocks
    # one cell longer; MPICH's Cartesian ranks with z fastest
    import poremaps_amd as pm
    for r in ref_meta["decomp"]["runs"]:
        nx, ny, nz = ref_meta["decomp"]["size_xyz"]
        got = pm.rank_map((nz, ny, nx), r["dims"])
        assert np.array_equal(got, ref_fields[r["tag"] + "_decomp"]), r["tag"]


@pytest.mark.parametrize("bc", [0, 2])
@pytest.mark.parametrize("solver", [1, 2, 3])
def test_stop_rule_matches_binary(ref_fields, ref_meta, bc, solver):
    e = [x for x in ref_meta["stop"] if x["bc"] == bc and x["solver"] == solver][0]
    G, code, st = _initial(ref_fields["box_raw"], bc)
    n, rows = O.calc_flow(G, code, bc, st, solver=solver, eps=1e-6, max_iter=200001, dx=1e-5)
    assert fmt_row(rows[-1]) == drop_tps(e["last_row"])
    assert n == int(e["last_row"].split(",")[0]) + 1


def _channel_rows_blocks(channel_raw, dims, upto):
    """Log rows of the channel under dom_decomposition dims at t = upto - 1
    (and the rate test against t = upto - 101)."""
    G, code, st = _initial(channel_raw, 0)
    outlet = np.ascontiguousarray(st[0][-3]).copy()
    O.sweeps_ref_blocks(2, G, code, *st, 0, upto - 100, dims, outlet=outlet)
    _, kprev = O.convergence(G, st[3], 0.0)
    O.sweeps_ref_blocks(2, G, code, *st, 0, 100, dims, outlet=outlet)
    conv, _ = O.convergence(G, st[3], kprev)
    phi = O.porosity(G)
    k = O.permeability(G, st[1], st[2], st[3], st[0], (0,) * 6, 5e-05, phi)
    return fmt_row([upto - 1, conv, 0.0, k[6], k[7]] + list(k[:6]))


@pytest.mark.parametrize("dims,solver,max_iter", [((1, 1, 1), 2, 3001), ((1, 1, 1), 1, 3001),
                                                  ((1, 1, 1), 2, 200001), ((1, 1, 1), 1, 200001)])
def test_channel_one_rank_matches_binary(channel_raw, ref_meta, dims, solver, max_iter):
    e = [x for x in ref_meta["channel"] if tuple(x["dims"]) == dims and x["solver"] == solver
         and x["max_iter"] == max_iter][0]
    G, code, st = _initial(channel_raw, 0)
    n, rows = O.calc_flow(G, code, 0, st, solver=solver, eps=1e-6, max_iter=max_iter, dx=5e-05)
    row = [fmt_row(x) for x in rows if int(x[0]) == int(e["row"].split(",")[0])]
    assert row == [drop_tps(e["row"])]


@pytest.mark.parametrize("dims", [(1, 1, 2), (2, 1, 1), (1, 1, 3), (1, 2, 1), (2, 2, 2)])
def test_channel_decompositions_at_t3000_match_binary(channel_raw, ref_meta, dims):
    e = [x for x in ref_meta["channel"] if tuple(x["dims"]) == dims and x["max_iter"] == 3001][0]
    assert _channel_rows_blocks(channel_raw, dims, 3001) == drop_tps(e["row"])


def test_channel_decomposition_known_answers(ref_meta):
    # SURVEY finding F7, measured independently during the survey
    rows = {tuple(x["dims"]): x["row"] for x in ref_meta["channel"] if x["max_iter"] == 3001 and x["solver"] == 2}
    assert rows[(1, 1, 1)].endswith("8.192931e-09")
    assert rows[(1, 1, 2)].endswith("8.192801e-09")
    assert rows[(2, 1, 1)].endswith("8.192311e-09")
