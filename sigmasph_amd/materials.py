"""Materials of the solid step: one resolved row per declared material.

Follows the reference's resolution (materials.py:148-449, material_models/*.py): an
entry declares `density0`, an `eos` (linear or Tait here) and, for a solid, a `strength`
card with `youngsModulus`, `poissonRatio` and a `plasticity` model (linear J2 hardening
or Johnson-Cook).  Each particle carries the integer row of its material (`eos_id`);
every per-material constant is read through it, as in the reference's `mat_tab`,
`eos_tab` and `plastic_tab` (SOLID.py:426-451).  The rows are float64 here.

`library` names a calibrated preset (material_models/material_library.py); the
entry's own keys are deep-merged over it.  The preset constants are data and are kept
as the reference states them.
"""
import copy
import ctypes
import math

from ._native import SphMaterial, EOS_LINEAR, EOS_TAIT, PLASTIC_NONE, PLASTIC_LINEAR, PLASTIC_JC


class MaterialError(ValueError):
    """A declared material whose constants cannot be resolved as written."""


def _jc_preset(rho0, cv, E, nu, A, B, n, C, Tm, m):
    return {"density0": rho0, "specificHeat": cv, "eos": {"type": "linear"},
            "strength": {"type": "hypoElastic", "youngsModulus": E, "poissonRatio": nu,
                         "plasticity": {"model": "johnson_cook", "A": A, "B": B, "n": n,
                                        "C": C, "eps0_dot": 0.001, "T0": 293.0, "Tm": Tm,
                                        "m": m, "Cp": cv, "chi": 0.9}}}


# material_models/material_library.py:28-180 (mm / ms / GPa / kg units)
LIBRARY = {
    "aluminium": _jc_preset(2.70e-06, 896.0, 69.0, 0.33, 0.324, 0.114, 0.42, 0.002, 925.0, 1.34),
    "construction_steel": _jc_preset(7.85e-06, 486.0, 210.0, 0.30, 0.350, 0.275, 0.36, 0.022,
                                     1780.0, 1.00),
    "stainless_steel": _jc_preset(7.90e-06, 500.0, 193.0, 0.30, 0.310, 1.000, 0.65, 0.020,
                                  1720.0, 1.00),
    "copper": _jc_preset(8.96e-06, 383.0, 124.0, 0.34, 0.090, 0.292, 0.31, 0.025, 1356.0, 1.09),
}

ENTRY_KEYS = {"name", "library", "density0", "specificHeat", "delta", "eos", "strength"}


def _merge(base, over):
    out = copy.deepcopy(base)
    for k, v in over.items():
        if isinstance(v, dict) and isinstance(out.get(k), dict):
            out[k] = _merge(out[k], v)
        else:
            out[k] = copy.deepcopy(v)
    return out


def expand_preset(entry):
    """expand_material_preset (material_library.py:548-580)."""
    lib = entry.get("library")
    if lib is not None:
        if lib not in LIBRARY:
            raise MaterialError("unknown library material %r (have %s)"
                                % (lib, ", ".join(sorted(LIBRARY))))
        merged = _merge(LIBRARY[lib], {k: v for k, v in entry.items() if k != "library"})
        return merged
    name = entry.get("name")
    if name in LIBRARY and ("density0" not in entry or "eos" not in entry
                            or "strength" not in entry):
        return _merge(LIBRARY[name], entry)
    return dict(entry)


class Material:
    """One resolved material (materials.py:148-449)."""

    def __init__(self, name, index, entry, default_delta=0.0):
        entry = expand_preset(entry)
        where = "Materials[%r]" % name
        for k in entry:
            if k not in ENTRY_KEYS:
                raise MaterialError("%s: unknown key %r" % (where, k))
        self.name, self.index = name, index
        try:
            self.rho0 = float(entry["density0"])
        except (KeyError, TypeError, ValueError):
            raise MaterialError("%s: density0 is required and must be a number" % where)
        if not self.rho0 > 0.0:
            raise MaterialError("%s: density0 must be positive" % where)
        d = entry.get("delta")
        self.delta = float(default_delta if d is None else d)
        self.cv = float(entry.get("specificHeat") or 0.0)
        eos = entry.get("eos") or {}
        kind = str(eos.get("type", "")).lower()
        strength = entry.get("strength")
        self.has_strength = strength is not None
        self.plastic_kind = PLASTIC_NONE
        self.sigma_y0 = self.hardening = 0.0
        self.jc = dict(A=0.0, B=0.0, n=0.0, C=0.0, eps0_dot=1e-3, T0=293.0, Tm=0.0, m=1.0)
        K_from_E = None
        self.G = 0.0
        if strength is not None:
            E, nu = strength.get("youngsModulus"), strength.get("poissonRatio")
            if E is None or nu is None:
                raise MaterialError("%s.strength: youngsModulus and poissonRatio are required"
                                    % where)
            E, nu = float(E), float(nu)
            K_from_E, self.G = E / (3.0 * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))
            pl = strength.get("plasticity") or {}
            model = str(pl.get("model") or "linear").lower()
            if model in ("johnson_cook", "hollomon", "power_law"):
                A = pl.get("A", pl.get("yieldStress"))
                if A is None:
                    raise MaterialError("%s: Johnson-Cook needs A" % where)
                self.jc = dict(A=float(A), B=float(pl.get("B", pl.get("hardeningModulus", 0.0))),
                               n=float(pl.get("n", 0.0)), C=float(pl.get("C", 0.0)),
                               eps0_dot=float(pl.get("eps0_dot", 1e-3)),
                               T0=float(pl.get("T0", 293.0)), Tm=float(pl.get("Tm", 0.0)),
                               m=float(pl.get("m", 1.0)))
                if self.jc["eps0_dot"] <= 0.0:
                    raise MaterialError("%s: eps0_dot must be positive" % where)
                self.plastic_kind = PLASTIC_JC
                self.sigma_y0, self.hardening = self.jc["A"], self.jc["B"]
                if self.cv == 0.0 and pl.get("Cp") is not None:
                    self.cv = float(pl["Cp"])
            elif model == "linear":
                sy = pl.get("yieldStress")
                if sy is None and pl.get("yieldStrain") is not None:
                    sy = float(pl["yieldStrain"]) * E
                self.sigma_y0 = 0.0 if sy is None else float(sy)
                H = pl.get("hardeningModulus")
                if H is None:
                    r = pl.get("tangentModulusRatio")
                    if r is None:
                        H = 0.0
                    else:
                        r = float(r)
                        if not 0.0 <= r < 1.0:
                            raise MaterialError("%s: tangentModulusRatio must be in [0, 1)" % where)
                        H = E * r / (1.0 - r)
                self.hardening = float(H)
                if self.sigma_y0 > 0.0:
                    self.plastic_kind = PLASTIC_LINEAR
            else:
                raise MaterialError("%s: plasticity model %r is not part of this step" % (where, model))
        if kind == "tait":
            c0, g = eos.get("c0"), eos.get("exponent")
            if c0 is None or g is None:
                raise MaterialError("%s.eos: a Tait equation of state needs c0 and exponent" % where)
            c0, g = float(c0), float(g)
            self.eos_kind, self.eos_gamma = EOS_TAIT, g
            self.eos_S = c0 * c0 * self.rho0 / g
            self.K, self.c_eos = c0 * c0 * self.rho0, c0
        elif kind == "linear":
            if strength is not None:
                if eos.get("bulkModulus") is not None:
                    raise MaterialError("%s.eos: bulkModulus is derived from the strength card"
                                        % where)
                K = K_from_E
            else:
                if eos.get("bulkModulus") is None:
                    raise MaterialError("%s.eos: a linear EOS without strength needs bulkModulus"
                                        % where)
                K = float(eos["bulkModulus"])
            self.eos_kind, self.eos_gamma, self.eos_S = EOS_LINEAR, 1.0, K
            self.K, self.c_eos = K, math.sqrt(K / self.rho0)
        else:
            raise MaterialError("%s.eos: type %r is not part of this step (linear, tait)"
                                % (where, kind))
        self.c_p = math.sqrt((self.K + 4.0 * self.G / 3.0) / self.rho0)
        # reference_temperature / reference_energy (materials.py:429-445)
        self.T0 = self.jc["T0"] if self.plastic_kind == PLASTIC_JC else 293.0
        self.e0 = 0.0

    def row(self):
        m = SphMaterial()
        m.rho0, m.K, m.G, m.c_p, m.c0 = self.rho0, self.K, self.G, self.c_p, self.c_eos
        m.sigma_y0, m.hardening, m.cv, m.T0, m.e0 = (self.sigma_y0, self.hardening, self.cv,
                                                     self.T0, self.e0)
        m.eos_S, m.eos_gamma, m.delta = self.eos_S, self.eos_gamma, self.delta
        j = self.jc
        (m.jc_A, m.jc_B, m.jc_n, m.jc_C, m.jc_eps0_dot, m.jc_T0, m.jc_Tm,
         m.jc_m) = j["A"], j["B"], j["n"], j["C"], j["eps0_dot"], j["T0"], j["Tm"], j["m"]
        m.eos_kind, m.plastic_kind = self.eos_kind, self.plastic_kind
        return m

    def describe(self):
        return ("%s: rho0 %.4g, K %.4g, G %.4g, c_p %.4g, delta %.3g, plastic %d"
                % (self.name, self.rho0, self.K, self.G, self.c_p, self.delta, self.plastic_kind))


def resolve(entries, default_delta=0.0):
    """name -> Material, rows in declaration order."""
    out = {}
    for k, e in enumerate(entries or []):
        if not isinstance(e, dict):
            raise MaterialError("Materials[%d]: expected an object" % k)
        name = e.get("name")
        if not isinstance(name, str) or not name:
            raise MaterialError("Materials[%d]: name is required" % k)
        if name in out:
            raise MaterialError("Materials[%d]: duplicate name %r" % (k, name))
        out[name] = Material(name, len(out), e, default_delta)
    return out


def table(materials):
    """ctypes array of rows, in index order."""
    rows = sorted(materials.values(), key=lambda m: m.index)
    arr = (SphMaterial * len(rows))()
    for k, m in enumerate(rows):
        arr[k] = m.row()
    return arr
