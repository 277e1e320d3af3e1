"""The oracle's shock terms against closed forms and invariants: the polynomial
Mie-Grueneisen EOS, the VNR bulk viscosity q and Monaghan's pair viscosity Pi.

- Mie-Grueneisen from shock constants (c0, s, Gamma0): the cold curve is the Hugoniot
  p_H = rho0 c0^2 eta / (1 - s eta)^2 (eta = 1 - rho0/rho) to third order in
  mu = rho/rho0 - 1 (the error falls as mu^4); in tension with linear expansion it is
  C1 mu exactly; (dp/de)_rho = Gamma0 rho; the sound speed is the isentropic
  derivative (checked against a finite difference along de = p/rho^2 drho) and is c0^2
  at the reference state; the cold energy vanishes at rho0 and its derivative is
  p_cold / rho^2 (finite differences on both branches);
- q and Pi conserve momentum and their work conjugates close the energy to round-off;
  Pi vanishes on a separating field; the dt limits take their closed forms on an affine
  converging field (corrected grad v is exact there): VNR
  CFL h / (Q + sqrt(Q^2 + c_p^2)), Q = C_Q l |div v|; Monaghan at beta = 0
  CFL h / ((1 + 1.2 alpha) c_p).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from sigmasph_amd import _native as N
from sigmasph_amd.materials import Material, LIBRARY, normalize_name
from sigmasph_amd.scenes import solid_block
from helpers import solid_layout, solid_oracle_state

AL2017 = {"name": "projectile", "library": "al2017_t4"}


def mg_row(entry=AL2017):
    m = Material(entry["name"], 0, entry)
    assert m.eos_kind == N.EOS_MG
    return m, m.row()


def test_library_aliases_and_mg_constants():
    assert normalize_name(" AL2017-T4 ") == "al2017_t4" and normalize_name("bg20") == "gelatin_20"
    m, row = mg_row()
    rho0, c0, s = 2.79e-06, 5328.0, 1.34
    assert m.rho0 == rho0 and m.K == pytest.approx(rho0 * c0 * c0, rel=1e-15)   # K0 = C1
    assert m.c_eos == c0
    assert row.mg_C2 == pytest.approx((2 * s - 1) * row.mg_C1, rel=1e-15)
    G = 72.4 / (2 * 1.33)
    assert m.c_p == pytest.approx(math.sqrt((rho0 * c0 * c0 + 4 * G / 3) / rho0), rel=1e-14)
    g = Material("g", 0, {"name": "gelatin"})       # implicit by name, through the alias
    assert g.eos_kind == N.EOS_MG and g.rho0 == LIBRARY["gelatin_10"]["density0"]


def hugoniot(m, mu):
    eta = mu / (1.0 + mu)
    return m.rho0 * m.c_eos ** 2 * eta / (1.0 - 1.34 * eta) ** 2


def test_mg_cold_curve_is_the_hugoniot_to_third_order():
    m, row = mg_row()
    errs = []
    for mu in (2e-2, 1e-2, 5e-3):
        p, _ = O.mg_pressure(row, m.rho0 * (1 + mu), 0.0)
        errs.append(abs(p - hugoniot(m, mu)) / hugoniot(m, mu))
    assert errs[0] < 1e-4
    for a, b in zip(errs, errs[1:]):
        assert 6.0 < a / b < 10.0        # relative error ~ mu^3, i.e. absolute ~ mu^4
    assert O.mg_pressure(row, m.rho0, 0.0)[0] == 0.0
    for mu in (-1e-3, -0.05):           # tension: C1 mu exactly (linear expansion)
        assert O.mg_pressure(row, m.rho0 * (1 + mu), 0.0)[0] == pytest.approx(row.mg_C1 * mu, rel=1e-13)
    _, cubic = mg_row({"name": "p", "library": "al2017_t4",
                       "eos": {"linearExpansion": False}})
    mu = -0.05
    want = cubic.mg_C1 * mu + cubic.mg_C2 * mu ** 2 + cubic.mg_C3 * mu ** 3
    assert O.mg_pressure(cubic, m.rho0 * (1 + mu), 0.0)[0] == pytest.approx(want, rel=1e-13)


def test_mg_thermal_term_and_sound_speed():
    m, row = mg_row()
    for mu in (0.0, 0.1, -0.02):
        rho = m.rho0 * (1 + mu)
        p0, _ = O.mg_pressure(row, rho, 0.0)
        p1, _ = O.mg_pressure(row, rho, 1e5)
        assert (p1 - p0) / 1e5 == pytest.approx(2.0 * rho, rel=1e-12)   # Gamma0 rho
    _, c2 = O.mg_pressure(row, m.rho0, 0.0)
    assert c2 == pytest.approx(5328.0 ** 2, rel=1e-14)

    def p_isentrope(rho1, e1, rho, n=64):
        # integrate de/drho = p / rho^2 from (rho1, e1) by RK4
        e, r, hstep = e1, rho1, (rho - rho1) / n
        f = lambda r_, e_: O.mg_pressure(row, r_, e_)[0] / (r_ * r_)
        for _ in range(n):
            k1 = f(r, e); k2 = f(r + hstep / 2, e + hstep * k1 / 2)
            k3 = f(r + hstep / 2, e + hstep * k2 / 2); k4 = f(r + hstep, e + hstep * k3)
            e += hstep * (k1 + 2 * k2 + 2 * k3 + k4) / 6
            r += hstep
        return O.mg_pressure(row, r, e)[0]

    for mu, e1 in ((0.15, 2e5), (0.05, 0.0), (-0.01, 5e4)):
        rho1 = m.rho0 * (1 + mu)
        d = 1e-5 * rho1
        fd = (p_isentrope(rho1, e1, rho1 + d) - p_isentrope(rho1, e1, rho1 - d)) / (2 * d)
        _, c2 = O.mg_pressure(row, rho1, e1)
        assert c2 == pytest.approx(fd, rel=1e-6)


@pytest.mark.parametrize("linear", [True, False])
def test_mg_cold_energy_integrates_the_cold_pressure(linear):
    _, row = mg_row({"name": "p", "library": "al2017_t4", "eos": {"linearExpansion": linear}})
    rho0 = row.rho0
    assert O.cold_energy(row, rho0) == 0.0
    for mu in (0.2, 0.01, -0.03):
        rho = rho0 * (1 + mu)
        d = 1e-6 * rho
        fd = (O.cold_energy(row, rho + d) - O.cold_energy(row, rho - d)) / (2 * d)
        assert fd == pytest.approx(O.mg_pressure(row, rho, 0.0)[0] / rho ** 2, rel=1e-6)
    # clamp without negative pressure: no cold energy in tension
    assert O.cold_energy(row, 0.99 * rho0, allow_negative_pressure=False) == 0.0


def _shock_state(two_d="planeStrain", q=0.0, alpha=0.0, beta=0.0, jitter=0.1, **kw):
    solver = {}
    if q:
        solver["bulkViscosity"] = {"quadratic": q, "lengthScale": kw.pop("length", "dx0")}
    if alpha or beta:
        solver["artificialViscosity"] = {"alpha": alpha, "beta": beta}
    sc = solid_block(n=(20, 20, 1) if two_d else (10, 10, 10), jitter=jitter, two_d=two_d, pst=False,
                     viscosity=0.0, hourglass=0.0, sts=False, delta=0.0, extra={"Solver": solver}, **kw)
    lay, st, P, S = solid_layout(sc)
    s = solid_oracle_state(lay, st)
    O.sort_only(lay.dims, s)
    return lay, st, P, S, s


@pytest.mark.parametrize("two_d", ["planeStrain", None])
@pytest.mark.parametrize("which", ["q", "pi"])
def test_shock_viscosities_conserve_momentum_and_energy(two_d, which):
    kw = {"q": 2.0} if which == "q" else {"alpha": 1.0, "beta": 2.0}
    lay, st, P, S, s = _shock_state(two_d, **kw)
    rng = np.random.default_rng(5)
    s.v[:] = rng.normal(size=(s.n, 3)) * 100.0
    g = rng.normal(size=(s.n, 9)) * 300.0
    if two_d:
        s.v[:, 2] = 0.0
        G = g.reshape(-1, 3, 3); G[:, 2, :] = 0.0; G[:, :, 2] = 0.0
    s.grad_v[:] = g
    s.V *= 1.0 + 0.01 * rng.normal(size=s.n)
    O.solid_phase("stress_accelerations", lay.dims, P, S, s)
    if which == "q":
        tr = g.reshape(-1, 3, 3).trace(axis1=1, axis2=2)
        assert np.count_nonzero(s.q) == np.count_nonzero(tr < 0) > 0
        assert np.all(s.q >= 0.0)
    else:
        assert np.count_nonzero(s.av_mu) > 0
    f = s.m[:, None] * s.acc
    scale = np.sum(np.abs(f), axis=0).max()
    assert np.all(np.abs(f.sum(axis=0)) < 1e-12 * scale)
    power = s.m * np.einsum("ij,ij->i", s.v, s.acc)
    assert abs(power.sum() + np.sum(s.m * s.de_int)) < 1e-12 * np.sum(np.abs(power))


def test_monaghan_leaves_a_separating_field_alone():
    outs = []
    for alpha in (0.0, 1.0):
        lay, st, P, S, s = _shock_state(alpha=alpha, beta=2.0 * alpha, jitter=0.0)
        c = s.x.mean(axis=0)
        s.v[:] = 50.0 * (s.x - c)
        s.v[:, 2] = 0.0
        O.solid_phase("stress_accelerations", lay.dims, P, S, s)
        outs.append((s.acc.copy(), s.de_int.copy()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("mode", ["q_dx0", "q_h", "monaghan"])
def test_dt_limits_on_a_converging_field(mode):
    if mode == "monaghan":
        lay, st, P, S, s = _shock_state(alpha=1.5, jitter=0.0, kernel_correction=True)
    else:
        lay, st, P, S, s = _shock_state(q=2.0, length=mode[2:], jitter=0.0, kernel_correction=True)
    k = 300.0
    c = s.x.mean(axis=0)
    s.v[:] = -k * (s.x - c)
    s.v[:, 2] = 0.0
    O.solid_step(lay.dims, P, S, s)
    dt = s.scal[N.SPH_S_DT]
    h, cfl, cp = lay.support_radius, P.cfl, st.c_p
    acoustic = cfl * h / math.sqrt(s.scal[N.SPH_S_MAX_C2])
    force = cfl * math.sqrt(h / s.scal[N.SPH_S_MAX_ACC])
    if mode == "monaghan":
        lim = cfl * h / ((1.0 + 1.2 * 1.5) * cp)
    else:
        ell = st.dx0 if mode == "q_dx0" else lay.support_radius
        Q = 2.0 * ell * 2.0 * k                      # C_Q l |tr grad v|, tr = -2k
        lim = cfl * h / (Q + math.sqrt(Q * Q + cp * cp))
    assert lim < min(acoustic, force)
    assert dt == pytest.approx(lim, rel=1e-12)


@pytest.mark.parametrize("mat", ["linear", "mg_gamma0", "mg"])
def test_compression_energy_reduction(mat):
    """Sum m e_cold(rho) over a uniformly compressed block: the Tait integral at gamma = 1
    (S/rho0 (ln x + 1/x - 1)) for the linear row, the Mie-Grueneisen cold integral at
    Gamma0 = 0, nothing for a row whose pressure reads e (Gamma0 > 0); the clamped form
    is zero in tension; the negative internal energy is Sum m min(e, 0)."""
    material = None
    if mat != "linear":
        g = 0.0 if mat == "mg_gamma0" else 2.0
        material = {"name": "p", "density0": 7.9e-6,
                    "eos": {"type": "mie_gruneisen", "c0": 4500.0, "s": 1.4, "gamma0": g},
                    "strength": {"type": "hypoElastic", "youngsModulus": 200.0, "poissonRatio": 0.3}}
    lay, st, P, S = solid_layout(solid_block(n=(6, 5, 1), material=material))
    s = solid_oracle_state(lay, st)
    row = st.mat_table[0]
    for x in (1.02, 0.97):
        s.V[:] = s.m / (row.rho0 * x)
        got = O.solid_reduce(s, N.SPH_R_COMPRESSION)
        clamped = O.solid_reduce(s, N.SPH_R_COMPRESSION_CLAMPED)
        if mat == "linear":
            want = s.m.sum() * row.eos_S / row.rho0 * (math.log(x) + 1.0 / x - 1.0)
        elif mat == "mg_gamma0":
            want = s.m.sum() * O.cold_energy(row, row.rho0 * x)
        else:
            want = 0.0
        assert got == pytest.approx(want, rel=1e-12, abs=1e-300)
        assert clamped == (pytest.approx(want, rel=1e-12, abs=1e-300) if x > 1.0 else 0.0)
    rng = np.random.default_rng(3)
    s.e_int[:] = rng.normal(size=s.n)
    assert O.solid_reduce(s, N.SPH_R_NEG_INTERNAL) == pytest.approx(
        np.sum(s.m * np.minimum(s.e_int, 0.0)), rel=1e-12)
