"""Binary trajectory output in the AMBER NetCDF convention, and the ParticlesFile reader.

`NetCDFTrajectoryWriter` keeps the reference's writer (netcdf_writer.py:1-225): one
growing `<scene>.nc` for the run, created on the first frame, the AMBER dimensions and
label variables, `coordinates`/`velocities` as (frame, atom, spatial) f4, the cell
description, `id`, and one (frame, atom) variable per per-particle field from the same
four field lists (always, solid, solid 3-D, optional-if-dumped), i4 for `object_id` and
`eroded`.  The container is the classic 64-bit-offset format (CDF-2), the one the
reference asks netCDF4 for.  netCDF4 is not installed here, so this module encodes the
container itself, from the published classic-format grammar: a big-endian header
(magic "CDF\\x02", record count, dimension, global-attribute and variable lists with
8-byte data offsets), the fixed-size variables, then one record per frame holding every
record variable's slab in header order, each padded to 4 bytes.  A frame is appended by
writing its record and then the new record count, and flushed, so a run cut off at any
point leaves a file that any NetCDF-3 reader opens with every frame written so far.

`read_particles_file` reads a ParticlesFile (particle_system.py:712-854): `position`
(n, 3), `velocity` (n, 3), `density` (n,), `object_id` (n,), and optionally `area`
(single layer) or `volume` (3-D), the per-particle sizes that switch on variable h.
It reads classic NetCDF (CDF-1, CDF-2) through scipy; a NetCDF-4 (HDF5) file is refused
with the conversion that makes it readable.
"""
import os
import struct

import numpy as np

# (dump key, variable name): the per-particle fields (netcdf_writer.py:46-97)
FIELDS_ALWAYS = [("object_id", "object_id"), ("density", "density"), ("pressure", "pressure"),
                 ("lam", "lam")]
FIELDS_SOLID = [("sigma_xx", "sigma_xx"), ("sigma_yy", "sigma_yy"), ("sigma_zz", "sigma_zz"),
                ("sigma_xy", "sigma_xy"), ("von_mises", "von_mises"), ("eps_plastic", "eps_plastic"),
                ("eroded", "eroded")]
FIELDS_SOLID_3D = [("sigma_xz", "sigma_xz"), ("sigma_yz", "sigma_yz")]
FIELDS_OPTIONAL = [("e_int", "e_int"), ("temperature", "temperature"), ("burn_f", "burn_f"),
                   ("damage", "damage"), ("porosity", "porosity"), ("ale_e_work", "ale_e_work"),
                   ("damage_t", "damage_t"), ("flaw_eps_min", "flaw_eps_min"), ("jc_omega", "jc_omega"),
                   ("h", "h")]
INTEGER_FIELDS = {"object_id", "eroded"}

# classic-format tags and types
_NC_DIMENSION, _NC_VARIABLE, _NC_ATTRIBUTE = 10, 11, 12
_TYPES = {"S1": (2, 1, ">S1"), "i4": (4, 4, ">i4"), "f4": (5, 4, ">f4"), "f8": (6, 8, ">f8")}


def _pad4(n):
    return (n + 3) & ~3


def _name(s):
    b = s.encode("utf-8")
    return struct.pack(">i", len(b)) + b + b"\0" * (_pad4(len(b)) - len(b))


def _attrs(items):
    if not items:
        return struct.pack(">ii", 0, 0)
    out = [struct.pack(">ii", _NC_ATTRIBUTE, len(items))]
    for k, v in items:
        b = str(v).encode("utf-8")
        out.append(_name(k) + struct.pack(">ii", 2, len(b)) + b + b"\0" * (_pad4(len(b)) - len(b)))
    return b"".join(out)


class _Var:
    def __init__(self, name, dtype, dims, attrs=()):
        self.name, self.dtype, self.dims, self.attrs = name, dtype, tuple(dims), list(attrs)
        self.begin = 0


class CDF2Writer:
    """A classic 64-bit-offset NetCDF file with one unlimited dimension, written once
    (`create`) and then appended to one record at a time (`append`)."""

    def __init__(self, path):
        self.path = path
        self.dims = []            # (name, length or None for the record dimension)
        self.gattrs = []
        self.vars = []
        self.numrecs = 0
        self._f = None

    def dim(self, name, length):
        self.dims.append((name, length))

    def var(self, name, dtype, dims, attrs=()):
        v = _Var(name, dtype, dims, attrs)
        self.vars.append(v)
        return v

    def _shape(self, v):
        return [dict(self.dims)[d] for d in v.dims]

    def _is_rec(self, v):
        return bool(v.dims) and dict(self.dims)[v.dims[0]] is None

    def _vsize(self, v):
        shape = self._shape(v)
        n = 1
        for s in (shape[1:] if self._is_rec(v) else shape):
            n *= s
        return _pad4(n * _TYPES[v.dtype][1])

    def _header(self):
        ids = {name: k for k, (name, _) in enumerate(self.dims)}
        out = [b"CDF\x02", struct.pack(">i", self.numrecs)]
        out.append(struct.pack(">ii", _NC_DIMENSION, len(self.dims)))
        for name, length in self.dims:
            out.append(_name(name) + struct.pack(">i", 0 if length is None else length))
        out.append(_attrs(self.gattrs))
        out.append(struct.pack(">ii", _NC_VARIABLE, len(self.vars)))
        for v in self.vars:
            out.append(_name(v.name) + struct.pack(">i", len(v.dims)))
            out.append(b"".join(struct.pack(">i", ids[d]) for d in v.dims))
            out.append(_attrs(v.attrs))
            out.append(struct.pack(">iiq", _TYPES[v.dtype][0], self._vsize(v), v.begin))
        return b"".join(out)

    def create(self, fixed_data):
        """Lay the file out and write the header and the fixed-size variables
        (`fixed_data`: name -> array)."""
        off = len(self._header())          # begins are fixed-width: the size is final
        for v in self.vars:
            if not self._is_rec(v):
                v.begin = off
                off += self._vsize(v)
        self.rec_start = off
        self.recsize = 0
        for v in self.vars:
            if self._is_rec(v):
                v.begin = off + self.recsize
                self.recsize += self._vsize(v)
        self._f = open(self.path, "wb")
        self._f.write(self._header())
        for v in self.vars:
            if not self._is_rec(v):
                a = np.ascontiguousarray(fixed_data[v.name], dtype=_TYPES[v.dtype][2]).tobytes()
                self._f.write(a + b"\0" * (self._vsize(v) - len(a)))
        self._f.flush()

    def append(self, record):
        """Write record `numrecs` (`record`: name -> array for every record variable),
        then the new record count."""
        parts = []
        for v in self.vars:
            if self._is_rec(v):
                a = np.ascontiguousarray(record[v.name], dtype=_TYPES[v.dtype][2]).tobytes()
                if len(a) != self._vsize(v) and _pad4(len(a)) != self._vsize(v):
                    raise ValueError("record variable %s: %d bytes for a slab of %d"
                                     % (v.name, len(a), self._vsize(v)))
                parts.append(a + b"\0" * (self._vsize(v) - len(a)))
        self._f.seek(self.rec_start + self.numrecs * self.recsize)
        self._f.write(b"".join(parts))
        self.numrecs += 1
        self._f.seek(4)
        self._f.write(struct.pack(">i", self.numrecs))
        self._f.flush()

    def close(self):
        if self._f is not None:
            self._f.close()
            self._f = None


class NetCDFTrajectoryWriter:
    """One growing AMBER-NetCDF file for the whole run (NetCDFTrajectoryWriter,
    netcdf_writer.py:100-225): created on the first write(), from that dump's particle
    count and keys; each write() appends a frame and flushes it."""

    def __init__(self, path, domain_size, is_solid, two_d=False, extra_fields=()):
        self.path = path
        self.domain_size = np.asarray(domain_size, dtype=float)
        self.fields = list(FIELDS_ALWAYS)
        if is_solid:
            self.fields += list(FIELDS_SOLID)
            if not two_d:
                self.fields += list(FIELDS_SOLID_3D)
        self.fields += list(extra_fields)
        self._w = None
        self._n = None
        self.frames_written = 0

    def _create(self, n, keys=()):
        self.fields += [(k, name) for k, name in FIELDS_OPTIONAL if k in keys]
        w = CDF2Writer(self.path)
        w.gattrs = [("Conventions", "AMBER"), ("ConventionVersion", "1.0"), ("program", "sigmaSPH"),
                    ("programVersion", "0.1")]
        w.dim("frame", None)
        w.dim("atom", n)
        w.dim("spatial", 3)
        w.dim("cell_spatial", 3)
        w.dim("cell_angular", 3)
        w.dim("label", 5)
        w.var("spatial", "S1", ("spatial",))
        w.var("cell_spatial", "S1", ("cell_spatial",))
        w.var("cell_angular", "S1", ("cell_angular", "label"))
        w.var("time", "f4", ("frame",), [("units", "picosecond")])
        w.var("coordinates", "f4", ("frame", "atom", "spatial"), [("units", "angstrom")])
        w.var("velocities", "f4", ("frame", "atom", "spatial"), [("units", "angstrom/picosecond")])
        w.var("cell_lengths", "f8", ("frame", "cell_spatial"), [("units", "angstrom")])
        w.var("cell_angles", "f8", ("frame", "cell_angular"), [("units", "degree")])
        w.var("id", "i4", ("frame", "atom"))
        for _key, name in self.fields:
            w.var(name, "i4" if name in INTEGER_FIELDS else "f4", ("frame", "atom"))
        w.create({"spatial": np.array(list("xyz"), "S1"), "cell_spatial": np.array(list("abc"), "S1"),
                  "cell_angular": np.array([list("alpha"), list("beta "), list("gamma")], "S1")})
        self._w = w
        self._n = n

    def write(self, obj_data, sim_time):
        """Append one frame; `obj_data` is what ParticleSystem.dump() returned."""
        pos = np.asarray(obj_data["position"])
        if self._w is None:
            self._create(pos.shape[0], obj_data.keys())
        elif pos.shape[0] != self._n:
            raise RuntimeError("particle count changed from %d to %d; the NetCDF atom dimension is "
                               "fixed when the file is created" % (self._n, pos.shape[0]))
        rec = {"time": np.float32(sim_time), "coordinates": pos.astype(np.float32),
               "velocities": np.asarray(obj_data["velocity"]).astype(np.float32),
               "cell_lengths": self.domain_size, "cell_angles": np.array([90.0, 90.0, 90.0]),
               "id": np.arange(self._n, dtype=np.int32)}
        for key, name in self.fields:
            rec[name] = np.asarray(obj_data[key]).astype(np.int32 if name in INTEGER_FIELDS else np.float32)
        self._w.append(rec)
        self.frames_written += 1

    def close(self):
        if self._w is not None:
            self._w.close()
            self._w = None


def read_particles_file(path, two_d):
    """The arrays of a ParticlesFile: {"position", "velocity", "density", "object_id"}
    and "size" (the `area` or `volume` column, when the file carries one), as float64 /
    int32 numpy arrays in file order (_load_particles_from_file, _read_particle_sizes,
    particle_system.py:666-854).  Raises SystemExit with the reason when the file is
    missing, not classic NetCDF, lacks a column, or carries the wrong size column."""
    from scipy.io import netcdf_file

    if not os.path.exists(path):
        raise SystemExit("ParticlesFile '%s' does not exist" % path)
    with open(path, "rb") as f:
        magic = f.read(4)
    if magic[:3] != b"CDF":
        kind = "a NetCDF-4 (HDF5) file" if magic == b"\x89HDF" else "not a NetCDF file"
        raise SystemExit("ParticlesFile '%s' is %s; this build reads classic NetCDF (CDF-1/CDF-2): "
                         "convert it with `nccopy -k '64-bit offset' in.nc out.nc`" % (path, kind))
    out = {}
    with netcdf_file(path, "r", mmap=False) as ds:
        for name in ("position", "velocity", "density", "object_id"):
            if name not in ds.variables:
                raise SystemExit("ParticlesFile '%s' has no `%s` variable" % (path, name))
        out["position"] = np.array(ds.variables["position"][:], dtype=np.float64).reshape(-1, 3)
        n = len(out["position"])
        out["velocity"] = np.array(ds.variables["velocity"][:], dtype=np.float64).reshape(-1, 3)
        out["density"] = np.array(ds.variables["density"][:], dtype=np.float64).reshape(-1)
        out["object_id"] = np.array(ds.variables["object_id"][:]).astype(np.int32).reshape(-1)
        for k in ("velocity", "density", "object_id"):
            if len(out[k]) != n:
                raise SystemExit("ParticlesFile '%s': `%s` holds %d rows for %d positions"
                                 % (path, k, len(out[k]), n))
        name, other = ("area", "volume") if two_d else ("volume", "area")
        if name in ds.variables:
            size = np.array(ds.variables[name][:], dtype=np.float64).reshape(-1)
            if len(size) != n or not np.all(np.isfinite(size)) or size.min() <= 0.0:
                raise SystemExit("ParticlesFile '%s': every `%s` must be positive and finite" % (path, name))
            out["size"] = size
        elif other in ds.variables:
            raise SystemExit("ParticlesFile '%s' carries `%s`, but a %s deck needs `%s` to give each "
                             "particle its size" % (path, other, "single-layer" if two_d else "3D", name))
    return out
