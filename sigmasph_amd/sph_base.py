"""SPHBase: what every solver shares (sph_base.py:10-655 of the reference).

Holds gravity, viscosity and reference density, the `dt[None]` scalar, and the f64
energy and momentum diagnostics.  Every diagnostic is a deterministic f64 reduction
on the GPU (`sph_reduce`), read back once: a device-to-host sync, as the reference's
`-> ti.f64` kernels are.
"""
import ctypes

import numpy as np
import torch

from . import _native as N
from .fields import Scalar


class SPHBase:
    def __init__(self, particle_system):
        self.ps = particle_system
        c = self.ps.cfg.get_cfg
        self.g = np.array(c("gravitation"), dtype=np.float64)
        self.viscosity = float(c("viscosity"))
        self.density_0 = float(c("density0"))
        self.dt = Scalar(self.ps.scal, N.SPH_S_DT)
        self.dt[None] = 0.0
        self.wall_work = Scalar(self.ps.scal, N.SPH_S_WALL_WORK)
        self._out = torch.zeros(4, dtype=torch.float64, device=self.ps.device)
        self._g_dev = torch.tensor(self.g, dtype=torch.float64)

    # ----------------------------------------------------------------------------------
    def _reduce(self, which, params=None):
        ps = self.ps
        g = (ctypes.c_double * 3)(*self.g)
        rc = ps._lib.sph_reduce(ps.ctx, ctypes.byref(ps.state), which, ctypes.cast(g, ctypes.c_void_p),
                                ctypes.byref(params) if params is not None else None,
                                self._out.data_ptr(), ps.stream())
        N.check(rc, ps.ctx, "sph_reduce")
        return self._out[:3].cpu().numpy()

    def initialize(self):
        self.ps.initialize_particle_system()

    def substep(self):
        pass

    def compute_total_mass(self):
        return float(self._reduce(N.SPH_R_MASS)[0])

    def compute_kinetic_energy(self):
        """Sum 1/2 m |v|^2 in f64 (sph_base.py:445-451)."""
        return float(self._reduce(N.SPH_R_KINETIC)[0])

    def compute_kinetic_energy_off_grid(self):
        return float(self._reduce(N.SPH_R_KINETIC_OFFGRID)[0])

    def compute_potential_energy(self):
        """-Sum m g.x, datum at the origin (sph_base.py:473-501)."""
        return float(self._reduce(N.SPH_R_POTENTIAL)[0])

    def compute_wall_energy(self):
        return float(self.wall_work[None])

    def compute_viscous_energy(self):
        return 0.0

    def compute_internal_energy(self):
        return 0.0

    def compute_elastic_energy(self):
        return 0.0

    def compute_hourglass_energy(self):
        return 0.0

    def compute_momentum(self):
        """Sum m v as a float64 3-vector (sph_base.py:551-590)."""
        return np.array(self._reduce(N.SPH_R_MOMENTUM), dtype=np.float64)

    def compute_gross_momentum(self):
        """Sum m |v| (sph_base.py:592-620)."""
        return float(self._reduce(N.SPH_R_GROSS_MOMENTUM)[0])

    def step(self):
        raise NotImplementedError
