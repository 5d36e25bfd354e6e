"""C-ABI library checks that need no GPU: every symbol the headers declare is
exported, and the host-side file formats (ordered .inp, raw dumps, log rows)
behave like the reference's (src/read_problem_params.cc, src/output.cc)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import poremaps_amd as pm
from conftest import GOLDEN, ROOT


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(pm_[a-z0-9_]+)\s*\(", txt)) - {"pm_log_fn", "pm_progress_fn"}


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", pm.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (pm_[a-z0-9_]+)", out))
    for h in ("poremaps_amd.h", "poremaps_amd_io.h", "poremaps_amd_batch.h", "poremaps_amd_serve.h"):
        if not os.path.exists(os.path.join(ROOT, "include", h)):
            continue
        missing = _declared(h) - exported
        assert not missing, f"{h}: not exported: {sorted(missing)}"


def test_library_loads_and_reports_errors_without_gpu_work():
    L = pm.lib()
    p = pm.make_params(8, 8, 8)
    assert pm.device_bytes(p) > 8 * 8 * 8 * 38
    bad = pm.make_params(8, 8, 8, boundary_method=7)
    assert pm.device_bytes(bad) == -1
    assert b"boundary_method" in L.pm_last_error()
    assert pm.device_bytes(pm.make_params(8, 8, 8, walls=1)) == pm.device_bytes(p)
    assert pm.device_bytes(pm.make_params(8, 8, 8, walls=2)) == -1
    assert b"walls" in L.pm_last_error()


def test_read_inp_reference_files():
    p, x = pm.read_inp(os.path.join(GOLDEN, "input_rect_channel_24_14_50_vs_5e-05.inp"))
    assert (p.nx, p.ny, p.nz) == (24, 14, 50)
    assert p.boundary_method == 0 and p.voxel_size == 5e-05
    assert p.max_iter == 100000001 and p.it_eval == 100 and p.it_write == 100
    assert p.solving_algorithm == 2 and p.eps == 1e-07 and p.porosity == 1.0
    assert list(p.dom_interest) == [0] * 6 and list(x.write_output) == [1, 1, 1, 1]
    assert x.geometry_file == b"rect_channel_24_14_50_vs_5e-05.raw"
    assert x.log_file == b"permeability_rect_channel_24_14_50_vs_5e-05.log"
    assert p.mode == 0 and p.precision == 0          # reference behaviour by default
    p, x = pm.read_inp(os.path.join(GOLDEN, "input_template.inp"))
    assert list(x.dom_decomposition) == [1, 2, 4] and p.eps == 1e-05


def test_read_inp_rejects_out_of_order_keys(tmp_path):
    src = open(os.path.join(GOLDEN, "input_template.inp")).read()
    bad = src.replace("voxel_size", "voxel_sizes")
    f = tmp_path / "bad.inp"
    f.write_text(bad)
    with pytest.raises(pm.PoremapsError) as e:
        pm.read_inp(str(f))
    assert "voxel_size" in str(e.value)
    with pytest.raises(pm.PoremapsError):
        pm.read_inp(str(tmp_path / "missing.inp"))


def test_raw_writer_scales_and_does_not_truncate(tmp_path):
    L = pm.lib()
    f = tmp_path / "velz_x.raw"
    f.write_bytes(b"\xff" * 64)                       # longer pre-existing file
    v = np.arange(4, dtype=np.float64)
    assert L.pm_write_raw_f64(str(f).encode(), v.ctypes.data, 4, 2e-5, 1) == 0
    data = f.read_bytes()
    assert len(data) == 64                            # MPI_MODE_RDWR|CREATE keeps the tail
    assert np.array_equal(np.frombuffer(data[:32], np.float64), v * 2e-5)
    g = tmp_path / "nh.raw"
    c = np.array([111, -1, 666], np.int32)
    assert L.pm_write_raw_i32(str(g).encode(), c.ctypes.data, 3) == 0
    assert np.array_equal(np.fromfile(g, np.int32), c)


def test_log_row_format(tmp_path):
    L = pm.lib()
    f = tmp_path / "run.log"
    r = pm.LogRow(1100, 1.0, 123.5, 1e-5, 2e-6, 0.0, 0.0, 0.0, 1e-20, -2e-20, 8.368958e-09)
    assert L.pm_write_log_row(str(f).encode(), C.byref(r), 1) == 0
    assert L.pm_write_log_row(str(f).encode(), C.byref(r), 0) == 0
    lines = f.read_text().splitlines()
    assert lines[0].startswith("# iteration, conv, TPS [1/s], wmax_velz [m/s]")
    assert lines[1] == ("1100, 1.000000e+00, 1.235000e+02, 1.000000e-05, 2.000000e-06, "
                        "0.000000e+00, 0.000000e+00, 0.000000e+00, 1.000000e-20, -2.000000e-20, "
                        "8.368958e-09")
    assert lines[2] == lines[1] and len(lines) == 3


def test_cli_reports_missing_input_without_gpu(tmp_path):
    r = subprocess.run([pm.CLI_PATH, str(tmp_path / "nope.inp")], capture_output=True, text=True,
                       cwd=tmp_path, timeout=60)
    assert r.returncode == 2 and "doesnt exist" in r.stderr
