"""Scene builders for the workloads this build is measured and tested on.

Each returns a scene dict in the reference's card format (readable by `SimConfig`).

- `dam_break_3d`: a 0.4 x 0.4 x 0.2 m water column in a closed 1.6 x 0.6 x 0.2 m tank on a
  dx = 0.01 m lattice (32 000 particles), +-0.01 dx seeded jitter, Tait gamma = 7,
  c0 = 10 sqrt(2 g H), support 2.6 dx (Wendland C2 with smoothing length 1.3 dx),
  delta = 0.1, shifting on, CFL 0.2.  The tank is the reference's reflecting box
  (`Domain.boundary: "reflect"`, sph_base.py:267-296); the reference has no wall
  particles.
- `tgv_block`: an n^3 lattice at dx in a periodic cube of side n dx, +-0.05 dx jitter,
  Taylor-Green vortex u = U sin(kx)cos(ky)cos(kz), v = -U cos(kx)sin(ky)cos(kz), w = 0
  with k = 2 pi / L, Tait c0 = 20 m/s, support 2.6 dx, delta = 0.1, shifting on, CFL 0.2.
  n = 160 at dx = 1 mm is the 4 096 000-particle precision-scaling block.
"""
import math


def dam_break_3d(dx=0.01, jitter=0.01, seed=0, column=(0.4, 0.4, 0.2), tank=(1.6, 0.6, 0.2),
                 delta=0.1, pst=True, viscosity=0.0, cfl=0.2):
    g = 9.81
    H = column[1]
    c0 = 10.0 * math.sqrt(2.0 * g * H)
    factor = 5.2                     # support radius = 2.6 dx = 5.2 particle radii
    h = 0.5 * dx * factor
    pad = h                          # the reflecting walls sit one support radius inside
    size = [tank[k] + 2.0 * pad for k in range(3)]
    lo = [pad + 0.5 * dx] * 3
    hi = [lo[k] + column[k] - 0.5 * dx for k in range(3)]
    return {
        "Domain": {"start": [0.0, 0.0, 0.0], "end": size, "particleRadius": 0.5 * dx,
                   "supportRadiusFactor": factor, "boundary": "reflect", "restitution": 0.5},
        "Time": {"CFL": cfl},
        "Solver": {"type": "deltaPlusSPH", "viscosity": viscosity},
        "Load": {"gravitation": [0.0, -g, 0.0]},
        "Diffusion": {"delta": delta},
        "Shifting": {"enabled": bool(pst), "mode": "sun", "maMode": "local"},
        "Materials": [{"name": "water", "density0": 1000.0,
                       "eos": {"type": "tait", "c0": c0, "exponent": 7.0}}],
        "FluidBlocks": [{"objectId": 0, "start": lo, "end": [hi[k] + 0.25 * dx for k in range(3)],
                         "density": 1000.0, "jitter": jitter, "seed": seed}],
    }


def tgv_block(n=160, dx=1.0e-3, jitter=0.05, seed=0, U=1.0, c0=20.0, delta=0.1, pst=True,
              viscosity=0.0, cfl=0.2):
    L = n * dx
    factor = 5.2
    return {
        "Domain": {"start": [0.0, 0.0, 0.0], "end": [L, L, L], "particleRadius": 0.5 * dx,
                   "supportRadiusFactor": factor, "periodic": "xyz"},
        "Time": {"CFL": cfl},
        "Solver": {"type": "deltaPlusSPH", "viscosity": viscosity},
        "Load": {"gravitation": [0.0, 0.0, 0.0]},
        "Diffusion": {"delta": delta},
        "Shifting": {"enabled": bool(pst), "mode": "sun", "maMode": "local"},
        "Materials": [{"name": "water", "density0": 1000.0,
                       "eos": {"type": "tait", "c0": c0, "exponent": 7.0}}],
        "FluidBlocks": [{"objectId": 0, "start": [0.5 * dx] * 3, "end": [L - 0.25 * dx] * 3,
                         "density": 1000.0, "jitter": jitter, "seed": seed,
                         "taylorGreen": {"U": U, "L": L, "origin": [0.0, 0.0, 0.0]}}],
    }


def free_block(n=(10, 10, 10), dx=0.01, jitter=0.05, seed=0, velocity=(0.0, 0.0, 0.0),
               delta=0.1, pst=True, viscosity=0.0, c0=20.0, gravity=(0.0, 0.0, 0.0),
               symmetry="", two_d=None, extra=None):
    """A small open-boundary block for tests: free surfaces on every face, optional
    symmetry planes at y = 0 / z = 0 (the block then starts at dx/2 from them) or a single
    plane-strain / axisymmetric layer (`two_d` = "planeStrain" / "axisymmetric")."""
    lo = [0.5 * dx + (0.3 if k == 0 else 0.0) for k in range(3)]
    hi = [lo[k] + n[k] * dx - 0.5 * dx for k in range(3)]
    if two_d:
        lo[2], hi[2] = 0.0, 0.25 * dx
    dom = {"start": [0.0, 0.0, 0.0], "end": [1.0, 1.0, 1.0], "particleRadius": 0.5 * dx,
           "supportRadiusFactor": 5.2 if not two_d else 6.0}
    if symmetry:
        dom["symmetryPlanes"] = symmetry
    if two_d:
        dom[two_d] = True
    sc = {
        "Domain": dom,
        "Time": {"CFL": 0.2},
        "Solver": {"type": "deltaPlusSPH", "viscosity": viscosity},
        "Load": {"gravitation": list(gravity)},
        "Diffusion": {"delta": delta},
        "Shifting": {"enabled": bool(pst), "mode": "sun", "maMode": "local"},
        "Materials": [{"name": "water", "density0": 1000.0,
                       "eos": {"type": "tait", "c0": c0, "exponent": 7.0}}],
        "FluidBlocks": [{"objectId": 0, "start": lo, "end": [hi[k] + 0.25 * dx for k in range(3)],
                         "density": 1000.0, "jitter": jitter, "seed": seed,
                         "velocity": list(velocity)}],
    }
    if extra:
        for k, v in extra.items():
            sc.setdefault(k, {}).update(v) if isinstance(v, dict) else sc.__setitem__(k, v)
    return sc
