"""AMBER-NetCDF trajectory output and ParticlesFile input (sigmasph_amd/netcdf_writer.py).

The writer encodes the classic 64-bit-offset container itself; every check reads its
files back through scipy.io.netcdf_file, an independent decoder of the same format, so a
header, offset or padding mistake fails to open or reads wrong values.  The variable
set follows netcdf_writer.py:46-97 and 134-188 of the reference: the AMBER dimensions and
labels, f4 coordinates/velocities, f8 cell description, i4 id / object_id / eroded, the
solid fields (the out-of-plane shears in 3-D only) and the optional fields the first
dump carries.  The input side is checked on files scipy writes, and through the host
layout against the same particles laid from blocks.
"""
import json
import struct

import numpy as np
import pytest
from scipy.io import netcdf_file

from sigmasph_amd.netcdf_writer import NetCDFTrajectoryWriter, read_particles_file


def _dump(n, rng, solid=True, extra=()):
    d = {"position": rng.normal(size=(n, 3)), "velocity": rng.normal(size=(n, 3)),
         "object_id": rng.integers(0, 3, n).astype(np.int32), "density": rng.uniform(7000, 8000, n),
         "pressure": rng.normal(size=n), "lam": rng.uniform(0, 1, n), "divergence": rng.normal(size=n)}
    if solid:
        for k in ("xx", "yy", "zz", "xy", "xz", "yz"):
            d["sigma_" + k] = rng.normal(size=n) * 1e8
        d["von_mises"] = rng.uniform(0, 1e9, n)
        d["eps_plastic"] = rng.uniform(0, 0.1, n)
        d["eroded"] = np.zeros(n, dtype=np.int32)
    for k in extra:
        d[k] = rng.normal(size=n)
    return d


def _read(path):
    with netcdf_file(str(path), "r", mmap=False) as ds:
        out = {"version": ds.version_byte, "dims": dict(ds.dimensions), "attrs": dict(ds._attributes),
               "vars": {k: (v.dimensions, v.typecode(), np.array(v[:]), dict(v._attributes))
                        for k, v in ds.variables.items()}}
    return out


def test_trajectory_frames_read_back_by_an_independent_decoder(tmp_path):
    rng = np.random.default_rng(3)
    path = tmp_path / "deck.nc"
    w = NetCDFTrajectoryWriter(str(path), [0.2, 0.1, 0.05], is_solid=True, two_d=False)
    dumps = [_dump(37, rng, extra=("e_int", "temperature")) for _ in range(3)]
    for k, d in enumerate(dumps):
        w.write(d, 1e-5 * k)
        f = _read(path)                       # readable after every frame, before close
        assert f["version"] == 2
        assert f["vars"]["time"][2].shape == (k + 1,)
    w.close()
    f = _read(path)
    assert f["attrs"] == {"Conventions": b"AMBER", "ConventionVersion": b"1.0", "program": b"sigmaSPH",
                          "programVersion": b"0.1"}
    assert f["dims"] == {"frame": None, "atom": 37, "spatial": 3, "cell_spatial": 3, "cell_angular": 3,
                         "label": 5}
    V = f["vars"]
    assert b"".join(V["spatial"][2]) == b"xyz" and b"".join(V["cell_spatial"][2]) == b"abc"
    assert [b"".join(r) for r in V["cell_angular"][2]] == [b"alpha", b"beta ", b"gamma"]
    assert V["coordinates"][0] == ("frame", "atom", "spatial") and V["coordinates"][1] == "f"
    assert V["coordinates"][3]["units"] == b"angstrom"
    assert V["velocities"][3]["units"] == b"angstrom/picosecond"
    assert V["cell_lengths"][1] == "d" and V["id"][1] == "i" and V["object_id"][1] == "i"
    names = [k for k in V if V[k][0][:2] == ("frame", "atom") and V[k][0] == ("frame", "atom")]
    assert names == ["id", "object_id", "density", "pressure", "lam", "sigma_xx", "sigma_yy", "sigma_zz",
                     "sigma_xy", "von_mises", "eps_plastic", "eroded", "sigma_xz", "sigma_yz", "e_int",
                     "temperature"]
    np.testing.assert_array_equal(V["time"][2], np.float32([0.0, 1e-5, 2e-5]))
    for k, d in enumerate(dumps):
        np.testing.assert_array_equal(V["coordinates"][2][k], d["position"].astype(np.float32))
        np.testing.assert_array_equal(V["velocities"][2][k], d["velocity"].astype(np.float32))
        np.testing.assert_array_equal(V["object_id"][2][k], d["object_id"])
        np.testing.assert_array_equal(V["id"][2][k], np.arange(37))
        np.testing.assert_array_equal(V["sigma_yz"][2][k], d["sigma_yz"].astype(np.float32))
        np.testing.assert_array_equal(V["temperature"][2][k], d["temperature"].astype(np.float32))
        np.testing.assert_array_equal(V["cell_lengths"][2][k], [0.2, 0.1, 0.05])
        np.testing.assert_array_equal(V["cell_angles"][2][k], [90.0, 90.0, 90.0])


def test_single_layer_and_fluid_field_sets(tmp_path):
    rng = np.random.default_rng(4)
    w = NetCDFTrajectoryWriter(str(tmp_path / "a.nc"), [1, 1, 1], is_solid=True, two_d=True)
    w.write(_dump(5, rng), 0.0)                 # odd atom count: i4/f4 slabs stay aligned
    w.close()
    V = _read(tmp_path / "a.nc")["vars"]
    assert "sigma_zz" in V and "sigma_xz" not in V and "sigma_yz" not in V and "e_int" not in V
    w = NetCDFTrajectoryWriter(str(tmp_path / "b.nc"), [1, 1, 1], is_solid=False)
    w.write(_dump(6, rng, solid=False), 0.5)
    w.write(_dump(6, rng, solid=False), 1.0)
    w.close()
    V = _read(tmp_path / "b.nc")["vars"]
    assert "von_mises" not in V and V["lam"][2].shape == (2, 6)


def test_particle_count_is_fixed_for_the_file(tmp_path):
    rng = np.random.default_rng(5)
    w = NetCDFTrajectoryWriter(str(tmp_path / "c.nc"), [1, 1, 1], is_solid=False)
    w.write(_dump(8, rng, solid=False), 0.0)
    with pytest.raises(RuntimeError, match="particle count changed"):
        w.write(_dump(9, rng, solid=False), 1.0)
    w.close()


def _particles_file(path, x, v, rho, obj, version=2, **extra):
    with netcdf_file(str(path), "w", version=version) as ds:
        ds.createDimension("n", len(x))
        ds.createDimension("xyz", 3)
        for name, arr, dims, t in (("position", x, ("n", "xyz"), "d"), ("velocity", v, ("n", "xyz"), "d"),
                                   ("density", rho, ("n",), "d"), ("object_id", obj, ("n",), "i")):
            ds.createVariable(name, t, dims)[:] = arr
        for name, arr in extra.items():
            ds.createVariable(name, "d", ("n",))[:] = arr


def test_read_particles_file_and_its_refusals(tmp_path):
    rng = np.random.default_rng(6)
    x, v = rng.normal(size=(11, 3)), rng.normal(size=(11, 3))
    rho, obj = rng.uniform(1, 2, 11), rng.integers(1, 4, 11)
    for version in (1, 2):
        p = tmp_path / ("p%d.nc" % version)
        _particles_file(p, x, v, rho, obj, version=version)
        d = read_particles_file(str(p), two_d=False)
        np.testing.assert_array_equal(d["position"], x)
        np.testing.assert_array_equal(d["object_id"], obj)
        assert "size" not in d
    p = tmp_path / "sized.nc"
    _particles_file(p, x, v, rho, obj, volume=np.full(11, 1e-9))
    assert read_particles_file(str(p), two_d=False)["size"][0] == 1e-9
    with pytest.raises(SystemExit, match="needs `area`"):
        read_particles_file(str(p), two_d=True)
    h5 = tmp_path / "h5.nc"
    h5.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(SystemExit, match="NetCDF-4"):
        read_particles_file(str(h5), two_d=False)
    with pytest.raises(SystemExit, match="does not exist"):
        read_particles_file(str(tmp_path / "none.nc"), two_d=False)
    bad = tmp_path / "bad.nc"
    with netcdf_file(str(bad), "w") as ds:
        ds.createDimension("n", 2)
        ds.createVariable("position", "d", ("n",))[:] = [0, 0]
    with pytest.raises(SystemExit, match="no `velocity`"):
        read_particles_file(str(bad), two_d=False)


def test_layout_from_a_particles_file_matches_the_blocks(tmp_path):
    """A ParticlesFile holding two blocks' particles (file order shuffled, ids 7 and 3)
    lays them grouped by ascending object id in file order, rounded through float32,
    with each object's SolidBlock material; grips select on the loaded positions."""
    from sigmasph_amd.config import SimConfig
    from sigmasph_amd.particle_system import SceneLayout
    from sigmasph_amd.scenes import solid_block
    sc = solid_block(n=(10, 6, 1), grips=0.05)
    blk = dict(sc["SolidBlocks"][0])
    b2 = dict(blk, objectId=7, start=[blk["start"][0], blk["end"][1] + 0.01, blk["start"][2]],
              end=[blk["end"][0], 2 * blk["end"][1] - blk["start"][1] + 0.01, blk["end"][2]])
    sc["SolidBlocks"] = [dict(blk, objectId=3), b2]
    ref = SceneLayout(SimConfig(sc))
    x, obj = ref.x_init, ref.obj_init
    v = np.random.default_rng(7).normal(size=x.shape)
    rho = np.full(len(x), 7850.0) + np.arange(len(x)) * 1e-3
    order = np.random.default_rng(8).permutation(len(x))
    pf = tmp_path / "parts.nc"
    _particles_file(pf, x[order], v[order], rho[order], obj[order])
    sc2 = json.loads(json.dumps(sc))
    sc2["ParticlesFile"] = str(pf)
    lay = SceneLayout(SimConfig(sc2))
    want = []
    for oid in (3, 7):
        want += [k for k in order if obj[k] == oid]
    want = np.array(want)
    np.testing.assert_array_equal(lay.obj_init, obj[want])
    np.testing.assert_array_equal(lay.x_init, x[want].astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(lay.v_init, v[want].astype(np.float32).astype(np.float64))
    rho32 = rho[want].astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(lay.m_init / lay.V_init, rho32, rtol=1e-15)
    np.testing.assert_array_equal(lay.eos_init, ref.eos_init[want])
    assert lay.object_collection[3]["particleNum"] == int((obj == 3).sum())
    assert lay.bc_flag_init.any()
    sized = tmp_path / "sized.nc"
    _particles_file(sized, x, v, rho, obj, area=np.full(len(x), 1e-9))   # a single-layer deck
    sc2["ParticlesFile"] = str(sized)
    with pytest.raises(SystemExit, match="variable h"):
        SceneLayout(SimConfig(sc2))


def test_header_layout_is_the_classic_grammar(tmp_path):
    """The first bytes of a trajectory: magic "CDF" 2, the record count, then the
    dimension list tag 10 with six entries, the first the unlimited `frame` (length 0)."""
    rng = np.random.default_rng(9)
    w = NetCDFTrajectoryWriter(str(tmp_path / "h.nc"), [1, 1, 1], is_solid=False)
    w.write(_dump(4, rng, solid=False), 0.0)
    w.write(_dump(4, rng, solid=False), 0.0)
    w.close()
    b = (tmp_path / "h.nc").read_bytes()
    assert b[:4] == b"CDF\x02"
    assert struct.unpack(">iii", b[4:16]) == (2, 10, 6)
    assert struct.unpack(">i", b[16:20])[0] == 5 and b[20:25] == b"frame"
    assert struct.unpack(">i", b[28:32])[0] == 0
