"""The headless driver (sigmasph_amd/run_simulation.py) and the constraint recorder.

CPU: the recorder's files and columns on a stand-in solver (constraint_recorder.py:
30-122 fixes them: the wide table with c{k}_{name}_* columns plus u_rel_mag and
F_axial_avg for two grips, one focused table per grip); the energy and momentum tables
on a given budget; the driver's refusals.
GPU: a solid deck with grips run through main() for two render cycles of two steps;
its final budget, momenta and the last CSV row against the oracle stepped the same way.
"""
import csv
import json
import os

import numpy as np
import pytest

from sigmasph_amd import run_simulation as R
from sigmasph_amd.constraint_recorder import ConstraintRecorder
from sigmasph_amd.scenes import solid_block


class _Scalar:
    def __init__(self, v):
        self.v = v

    def __getitem__(self, k):
        return self.v


class _FakeSolver:
    num_constraints = 2
    constraints_list = [{"name": "left"}, {"name": "constraint_right"}]

    def __init__(self):
        self.load_s = _Scalar(0.25)
        self.k = 0

    def get_constraint_data(self):
        self.k += 1
        return [{"displacement": np.array([-0.1 * self.k, 0.0, 0.0]), "force": np.array([-3.0, 0.5, 0.0])},
                {"displacement": np.array([0.1 * self.k, 0.0, 0.0]), "force": np.array([4.0, -0.5, 0.0])}]


def test_constraint_recorder_files(tmp_path):
    rec = ConstraintRecorder("deck", _FakeSolver(), interval=1, output_dir=str(tmp_path))
    rec.record(1, 1e-5)
    rec.record(2, 2e-5)
    rec.close()
    rows = list(csv.reader(open(tmp_path / "deck_constraints.csv")))
    assert rows[0][:4] == ["step", "time", "load_s", "c0_left_ux"]
    assert rows[0][-2:] == ["u_rel_mag", "F_axial_avg"]
    assert len(rows[0]) == 3 + 2 * 8 + 2 and len(rows) == 3
    r = rows[2]
    assert r[0] == "2" and r[1] == "2.00000000e-05" and r[2] == "2.500000e-01"
    assert float(r[-2]) == pytest.approx(0.4) and float(r[-1]) == pytest.approx(3.5)
    assert float(r[3 + 3]) == pytest.approx(0.2)          # |u| of grip 0
    left = list(csv.reader(open(tmp_path / "deck_constraint_left.csv")))
    right = list(csv.reader(open(tmp_path / "deck_constraint_right.csv")))   # prefix not doubled
    assert left[0] == ["step", "time", "load_s", "u_x", "u_y", "u_z", "u_mag", "F_x", "F_y", "F_z", "F_mag"]
    assert float(right[1][7]) == 4.0 and float(right[1][10]) == pytest.approx(np.hypot(4.0, 0.5))


class _PS:
    is_solid = True
    axisymmetric = True
    walls = False


def test_energy_and_momentum_tables():
    b0 = {"ke": 1.0, "pe": 0.0, "ie": 0.0, "hg": 0.0, "leftover": 0.0, "wall": 0.0, "visc": 0.0,
          "pw": 0.0, "wd": 0.0, "se": 0.0, "total": 1.0}
    b1 = dict(b0, ke=0.5, ie=0.49, hg=0.005, leftover=0.001, visc=-0.1, pw=0.2)
    txt = R.format_energy_summary(_PS(), b0, b1, step=10, t=1e-3)
    assert "Internal Energy (IE)" in txt and "Hourglass Work (HE)" in txt
    tot1 = 0.5 + 0.49 + 0.005 - 0.001
    assert ("%13.4e" % tot1).strip() in txt

    class S:
        g = [0.0, 0.0, 0.0]
        num_constraints = 2
    m = R.format_momentum_summary(_PS(), S(), np.zeros(3), 1.0, np.array([1e-3, 0.2, 0.0]), 1.0)
    assert "p_x  (axial, CONSERVED)" in m and "2 displacement grip(s)" in m
    assert "0.10000% of the gross momentum" in m


def test_driver_refuses_the_window_and_unbuilt_outputs(tmp_path):
    with pytest.raises(SystemExit, match="--no-gui"):
        R.main(["--scene_file", "x.json"])
    sc = solid_block(n=(6, 4, 1))
    sc["Output"] = {"ply": True}
    p = tmp_path / "d.json"
    p.write_text(json.dumps(sc))
    with pytest.raises(SystemExit, match="ply"):
        R.main(["--scene_file", str(p), "--no-gui"])


@pytest.mark.gpu
def test_driver_run_matches_the_oracle(gpu, tmp_path):
    from oracle import oracle as O
    from sigmasph_amd import _native as N
    from sigmasph_amd.SOLID import SolidSettings
    from helpers import solid_layout, solid_oracle_state
    sc = solid_block(n=(24, 8, 1), two_d="axisymmetric", grips=0.05, ramp_time=1e-4,
                     div_velocity=(300.0, 0.0, 0.0))
    sc["Time"].update({"stepsPerRenderUpdate": 2, "maxSteps": 2})
    sc["Output"] = {"constraints": True, "txt": True, "netcdf": True, "interval": 1}
    path = tmp_path / "grips.json"
    path.write_text(json.dumps(sc))
    res = R.main(["--scene_file", str(path), "--no-gui"], output_dir=str(tmp_path), quiet=True)
    assert res["steps"] == 4 and res["cycles"] == 2 and res["stop"] == "hit maxSteps"
    lay, st, P, S = solid_layout(sc)
    s = solid_oracle_state(lay, st)
    for _ in range(4):
        O.solid_step(lay.dims, P, S, s)
    b = res["budget1"]
    m = s.m
    ke = 0.5 * np.sum(m * np.sum(s.v * s.v, axis=1))
    assert b["ke"] == pytest.approx(ke, rel=1e-9)
    assert b["ie"] == pytest.approx(O.solid_reduce(s, N.SPH_R_INTERNAL), rel=1e-8)
    assert b["comp"] == pytest.approx(O.solid_reduce(s, N.SPH_R_COMPRESSION), rel=1e-8, abs=1e-300)
    assert b["e_neg"] == pytest.approx(O.solid_reduce(s, N.SPH_R_NEG_INTERNAL), rel=1e-8, abs=1e-300)
    for key, slot in (("hg", N.SPH_S_HG_WORK), ("leftover", N.SPH_S_LEFTOVER), ("visc", N.SPH_S_VISC_WORK)):
        scale = max(abs(s.scal[N.SPH_S_HG_WORK]), abs(s.scal[N.SPH_S_VISC_WORK]), abs(s.scal[N.SPH_S_LEFTOVER]))
        assert abs(b[key] - s.scal[slot]) <= 1e-8 * scale
    p = np.sum(m[:, None] * s.v, axis=0)
    np.testing.assert_allclose(res["p1"], p, rtol=1e-9, atol=1e-12 * np.sum(m * np.abs(s.v).sum(1)))
    rows = list(csv.reader(open(tmp_path / "grips_constraints.csv")))
    assert len(rows) == 5 and rows[-1][0] == "4"
    fx = float(rows[-1][3 + 4])                       # c0_left_Fx
    assert fx == pytest.approx(s.cons[0, 0], rel=1e-7, abs=1e-9 * np.abs(s.cons[:2, :3]).max())
    assert os.path.exists(tmp_path / "dump_0000.xyz") and os.path.exists(tmp_path / "dump_0001.xyz")
    head = open(tmp_path / "dump_0001.xyz").read().splitlines()[:2]
    assert int(head[0]) == lay.dims.n and "von_mises:R:1" in head[1]
    # the trajectory: one frame per dump, the same dump the XYZ file was written from
    from scipy.io import netcdf_file
    xyz = np.loadtxt(tmp_path / "dump_0001.xyz", skiprows=2)
    with netcdf_file(str(tmp_path / "grips.nc"), "r", mmap=False) as ds:
        assert ds.variables["coordinates"].shape == (2, lay.dims.n, 3)
        assert ds.variables["time"][0] == 0.0 and ds.variables["time"][1] > 0.0
        np.testing.assert_allclose(ds.variables["coordinates"][1], xyz[:, :3], rtol=2e-3, atol=1e-9)
        np.testing.assert_allclose(ds.variables["von_mises"][1], xyz[:, 11], rtol=2e-3, atol=1.0)
        assert "sigma_xz" not in ds.variables


@pytest.mark.gpu
def test_driver_fluid_dambreak_with_duration(gpu, tmp_path):
    """A fluid deck stopped by Time.duration: the last step is clipped to the time left,
    the budget's terms are the oracle's after the same steps."""
    from oracle import oracle as O
    from sigmasph_amd.scenes import dam_break_3d
    from helpers import layout, oracle_state
    sc = dam_break_3d(dx=0.04, column=(0.4, 0.4, 0.2), tank=(1.6, 0.6, 0.2))
    lay, fs, P = layout(sc)
    dur = 2.5 * fs.dt_initial
    sc["Time"].update({"stepsPerRenderUpdate": 5, "duration": dur})
    path = tmp_path / "dam.json"
    path.write_text(json.dumps(sc))
    res = R.main(["--scene_file", str(path), "--no-gui"], output_dir=str(tmp_path), quiet=True)
    assert res["stop"] == "reached duration"
    assert res["sim_time"] == pytest.approx(dur, rel=1e-12)
    assert res["steps"] >= 3 and res["cycles"] == 1
    s = oracle_state(lay, fs)
    ke0 = 0.5 * np.sum(s.m * np.sum(s.v * s.v, axis=1))
    assert res["budget0"]["ke"] == pytest.approx(ke0, rel=1e-12, abs=1e-300)
    assert np.isfinite(res["budget1"]["total"]) and res["budget1"]["ke"] > 0.0
