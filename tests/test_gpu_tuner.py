"""Jacobi kernel forms and the autotuner on the GPU.

The velocity + pressure pair, the lagged single pass over the fluid list
(all loads up front, case-predicated, or from per-voxel records) and the
lagged dense 2.5-D pass (two
z-segment lengths) -- in FP32+DC the pair and lagged passes over the
interleaved state, and the lagged pass on the compact state -- must produce the same
iterates bit for bit, and a solve whose first sweeps the autotuner spreads
over all of them must end exactly where a fixed-form solve ends."""
import os

import numpy as np
import pytest

import poremaps_amd as pm
from poremaps_amd import geometry as geo
from conftest import random_geometry

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def solve(raw, bc, precision, form=None, zs=None, loads=None, march=None):
    old = {k: os.environ.get(k) for k in ("POREMAPS_JACOBI", "POREMAPS_STREAM_ZS", "POREMAPS_SWEEP_LOADS",
                                          "POREMAPS_MARCH_TY", "POREMAPS_MARCH_ZSEG", "POREMAPS_REC_SKIP2")}
    try:
        for k in ("POREMAPS_MARCH_TY", "POREMAPS_MARCH_ZSEG"):
            os.environ.pop(k, None)
        if march:
            os.environ["POREMAPS_MARCH_TY"], os.environ["POREMAPS_MARCH_ZSEG"] = str(march[0]), str(march[1])
        if loads:
            os.environ["POREMAPS_SWEEP_LOADS"] = loads
        else:
            os.environ.pop("POREMAPS_SWEEP_LOADS", None)
        os.environ.pop("POREMAPS_REC_SKIP2", None)
        if form and (form.startswith("records-") or form.startswith("compact-")):
            # the record pass skipping its unused +-2 loads (skip2) or every
            # unused load (skipall); the compact pass skipping every unused
            # load
            os.environ["POREMAPS_REC_SKIP2"] = "2" if form.endswith("skipall") else "1"
            form = form.split("-")[0]
        if form:
            os.environ["POREMAPS_JACOBI"] = form
        else:
            os.environ.pop("POREMAPS_JACOBI", None)
        if zs:
            os.environ["POREMAPS_STREAM_ZS"] = str(zs)
        else:
            os.environ.pop("POREMAPS_STREAM_ZS", None)
        with pm.Solver(raw, boundary_method=bc, mode="jacobi", precision=precision, omega=pm.OMEGA,
                       omega_p=pm.OMEGA_P, eps=1e-6, voxel_size=1e-5, max_iter=200_000) as s:
            res, _ = s.solve()
            fields = [s.field(n) for n in ("velx", "vely", "velz", "press")]
            return res, fields, s.kernel_choice()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("precision,bc", [("f32dc", 0), ("f32dc", 1), ("f64", 2), ("f64", 4)])
def test_kernel_forms_and_tuned_solve_are_bit_identical(precision, bc):
    # 70 x 20 x 36: two x tiles and three y tiles of the dense kernel, a
    # partial tile on each axis and (zs 16) three z segments
    raw = random_geometry((36, 20, 70), 0.4, 71 + bc)
    say("kernel forms", precision, bc)
    ref, ref_fields, _ = solve(raw, bc, precision, "pair")
    forms = [("list", None, None, None), ("list", None, "branch", None), ("records", None, None, None),
             ("records-skip2", None, None, None), ("stream", 16, None, None), ("stream", 64, None, None)]
    if precision == "f32dc":
        # the z-march of the interleaved lagged pass: tiles of 64 x 4 / 64 x 8,
        # z segments of 5 planes (several per solve), 128 and the whole depth
        forms += [("march", None, None, (4, 5)), ("march", None, None, (8, 128)), ("march", None, None, (4, 0)),
                  ("records-skipall", None, None, None), ("compact", None, None, None),
                  ("compact-skipall", None, None, None)]
    for form, zs, loads, march in forms + [(None, None, None, None)]:
        res, fields, (name, tried) = solve(raw, bc, precision, form, zs, loads, march)
        say(form, zs, loads, march, name, res.sweeps)
        assert res.sweeps == ref.sweeps and res.last.wk33 == ref.last.wk33, (form, zs)
        for a, b, n in zip(fields, ref_fields, ("velx", "vely", "velz", "press")):
            assert np.array_equal(a, b), (form, zs, n)
        if form is None:
            # FP32+DC: the layout (compact or dense) timed first, then the
            # forms of that layout: the compact pass with every load and with
            # each load only where used (also at 6 waves per SIMD), or
            # the dense interleaved forms; FP64: pair, lagged list, records
            # and the dense 2.5-D forms
            forms_tried = {k: v for k, v in tried.items() if not k.startswith("layout-") and k != "packed-far"}
            assert name in forms_tried
            if precision == "f32dc":
                assert set(tried) >= {"layout-compact", "layout-dense"}
                compact = tried["layout-compact"] <= tried["layout-dense"]
                assert name.startswith("interleaved-compact") == compact
                assert len(forms_tried) == (3 if compact else 9)
            else:
                assert len(forms_tried) == 7


def test_divergent_relaxation_restarts_slower():
    """A relaxation pair beyond the stability bound (omega * omega_p = 22.5)
    diverges; the solve starts over from the initial state with omega_p cut
    by a quarter and ends exactly where a solve started with that pair ends
    -- alone and in a batch."""
    # open pores: the unstable mode is a high-frequency mode of bulk fluid
    raw = geo.sphere_pack((20, 16, 18), 0.55, 2, 4, seed=3)
    say("restart after divergence")
    kw = dict(boundary_method=0, voxel_size=1e-5, mode="jacobi", precision="f32dc", eps=1e-6, max_iter=100_000)
    with pm.Solver(raw, omega=2.5, omega_p=9.0, **kw) as s:
        res, _ = s.solve()
    with pm.Solver(raw, omega=2.5, omega_p=9.0 * 0.75, **kw) as s:
        ref, _ = s.solve()
    assert res.restarts == 1 and ref.restarts == 0
    assert res.sweeps == ref.sweeps and res.last.wk33 == ref.last.wk33
    with pm.Batch(omega=2.5, omega_p=9.0) as b:
        slot = b.add(raw, boundary_method=0, voxel_size=1e-5)
        while b.step():
            pass
        state, bres = b.status(slot)
    assert state == "converged" and bres.restarts == 1
    assert bres.sweeps == ref.sweeps and bres.last.wk33 == ref.last.wk33
