"""Pins of the oracle's remap geometry (src/advect.c:1226-1408) and strength
models (src/strength.c:379-854) by what the mathematics fixes:

* get_flux_volcyl: the swept quadrilateral's signed area times the mean
  radius of its corners (the Pappus volume /(2 pi) for a parallelogram,
  checked against direct quadrature), exactly antisymmetric in its sign;
* cut_length: inverse of the area fraction cut from an a x b rectangle by
  the line x + y = d (computed here by exact piecewise integration);
* geometric_flux_ratio: the fraction of an edge where a linear level set
  is positive;
* vofMSTACS: continuity at the switching points and boundedness
  (nD <= f <= min(nD/cd, 1) for Courant numbers below 1/3);
* strength: limits, slopes, continuity and special cases of each model;
* the epsilon-alpha porosity rate (WunnemannPorosity): its switches, and its
  compaction laws integrated along a compressive path against their closed
  forms (exponential above alpha_x, parabolic below).
"""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
import pysalec as ps

D2 = C.c_double * 2


def volcyl(a, b, da, db):
    return po.olib().oracle_get_flux_volcyl(D2(*a), D2(*b), D2(*da), D2(*db))


def shoelace(P):
    x, y = P[:, 0], P[:, 1]
    return 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)


def test_flux_volcyl_area_times_mean_radius():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        a = rng.uniform([0, -50], [100, 50])
        b = a + rng.uniform(-10, 10, 2)
        da, db = rng.uniform(-3, 3, 2), rng.uniform(-3, 3, 2)
        quad = np.array([a, b, b - db, a - da])
        S = shoelace(quad)
        xbar = np.mean(quad[:, 0])
        v = volcyl(a, b, da, db)
        # swept areas here are O(10); allow rounding of O(1e-12) of that scale
        assert abs(v - S * abs(xbar)) <= 1e-12 * 100.0 * abs(xbar), (v, S, xbar)


def test_flux_volcyl_pappus_quadrature():
    """Parallelogram swept by a rigid face shift: V/(2 pi) = integral of r dA."""
    rng = np.random.default_rng(2)
    g, w = np.polynomial.legendre.leggauss(8)
    for _ in range(200):
        a = rng.uniform([5, -50], [100, 50])
        b = a + rng.uniform(-10, 10, 2)
        d = rng.uniform(-3, 3, 2)
        # bilinear map of the unit square onto (a, b, b-d, a-d)
        P = np.array([a, b, b - d, a - d])
        tot = 0.0
        for u, wu in zip(g, w):
            for s, ws in zip(g, w):
                uu, ss = 0.5 * (u + 1), 0.5 * (s + 1)
                x = (1 - uu) * (1 - ss) * P[0] + uu * (1 - ss) * P[1] + uu * ss * P[2] + (1 - uu) * ss * P[3]
                dxu = (1 - ss) * (P[1] - P[0]) + ss * (P[2] - P[3])
                dxs = (1 - uu) * (P[3] - P[0]) + uu * (P[2] - P[1])
                jac = dxu[0] * dxs[1] - dxu[1] * dxs[0]
                tot += 0.25 * wu * ws * x[0] * jac
        v = volcyl(a, b, d, d)
        assert abs(abs(v) - abs(tot)) <= 1e-10 * abs(tot) + 1e-12
        assert np.sign(v) == np.sign(tot) or tot == 0.0


def test_flux_volcyl_antisymmetric_sign():
    rng = np.random.default_rng(3)
    for _ in range(5000):
        a = rng.uniform([0, -50], [100, 50])
        b = a + rng.uniform(-10, 10, 2)
        da, db = rng.uniform(-1, 1, 2) * 10.0 ** rng.uniform(-12, 1), rng.uniform(-1, 1, 2) * 10.0 ** rng.uniform(-12, 1)
        v1, v2 = volcyl(a, b, da, db), volcyl(b, a, db, da)
        assert np.sign(v1) == -np.sign(v2)
        assert abs(v1 + v2) <= 4e-16 * abs(v1)


def rect_fraction(d, a, b):
    """Area of {x + y <= d} inside [0, a] x [0, b], divided by a b."""
    # integrand clamp(d - x, 0, b) is piecewise linear in x with kinks at d - b and d
    xs = np.unique(np.clip([0.0, d - b, d, a], 0.0, a))
    area = 0.0
    for x0, x1 in zip(xs[:-1], xs[1:]):
        f0, f1 = np.clip(d - x0, 0, b), np.clip(d - x1, 0, b)
        area += 0.5 * (f0 + f1) * (x1 - x0)
    return area / (a * b)


def test_cut_length_inverts_area_fraction():
    L = po.olib()
    rng = np.random.default_rng(4)
    for _ in range(3000):
        a, b = 10.0 ** rng.uniform(-1, 2, 2)
        if max(a, b) * 1e-5 > min(a, b):
            continue
        vf = rng.uniform(0, 1)
        d = L.oracle_cut_length(vf, a, b)
        assert abs(rect_fraction(d, max(a, b), min(a, b)) - vf) <= 1e-12
    assert L.oracle_cut_length(0.0, 3.0, 2.0) == 0.0
    assert L.oracle_cut_length(1.0, 3.0, 2.0) == 5.0
    assert L.oracle_cut_length(1.5, 3.0, 2.0) == 5.0
    # below 0 the general branch takes sqrt of a negative number, as the
    # reference does (src/advect.c:1380-1381); callers only pass mixed cells
    assert np.isnan(L.oracle_cut_length(-0.5, 3.0, 2.0))
    # degenerate (one side below 1e-5 of the other): the strip limit d = vf a + b
    for vf in (0.0, 0.3, 1.0):
        d = L.oracle_cut_length(vf, 1.0, 1e-7)
        assert abs(rect_fraction(d, 1.0, 1e-7) - vf) <= 1e-6


def test_geometric_flux_ratio_level_set_fraction():
    L = po.olib()
    rng = np.random.default_rng(5)
    s = (np.arange(200000) + 0.5) / 200000
    for _ in range(300):
        p1, p2 = rng.uniform(-1, 1, 2)
        got = L.oracle_geometric_flux_ratio(p1, p2)
        frac = np.mean(p1 + (p2 - p1) * s > 0)
        assert abs(got - frac) <= 2e-5
    assert L.oracle_geometric_flux_ratio(0.0, 0.0) == 1.0
    assert L.oracle_geometric_flux_ratio(-1.0, -2.0) == 0.0


def test_mstacs_continuity_and_boundedness():
    L = po.olib()
    f = L.oracle_vof_mstacs
    for cd in (0.05, 0.2, 1.0 / 3.0, 0.6, 0.9):
        for gf in (0.0, 0.3, 1.0):
            for x in (0.2, 0.5):
                lo, hi = f(x * (1 - 1e-12), gf, cd), f(x * (1 + 1e-12), gf, cd)
                assert abs(lo - hi) <= 1e-9
            for nv in np.linspace(1e-6, 1 - 1e-6, 101):
                v = f(nv, gf, cd)
                # bounded: never below the donor value, never above 1, and for
                # cd < 1/3 never above the HYPER-C limit min(nD/cd, 1); above
                # cd = 1/3 the compressive part is min(3 nD, 1)
                upper = min(nv / cd, 1.0) if cd < 1.0 / 3.0 else min(3.0 * nv, 1.0)
                assert nv - 1e-15 <= v <= upper + 1e-15, (nv, gf, cd, v)
    # the compressive limit (gf = 1) is HYPER-C, the smooth limit (gf = 0) STOIC
    assert f(0.1, 1.0, 0.2) == pytest.approx(0.5)
    assert f(0.1, 0.0, 0.2) == pytest.approx(0.3)
    assert f(0.9, 0.0, 0.2) == 1.0


# ----------------------------------------------------------- strength ---
def material(**kw):
    m = ps.Material()
    m.Yint0, m.Yfricint, m.Ylimint = 1.0e7, 1.2, 2.5e9
    m.Ydam0, m.Yfricdam, m.Ylimdam = 1.0e4, 0.6, 2.5e9
    m.IvanA, m.IvanB, m.IvanC = 1.0e-4, 1.0e-11, 3.0e8
    m.Pbd, m.Pbp = 2.7e9, 4.1e9
    m.Asimon, m.Csimon, m.Tmelt0, m.Tfrac, m.Tdelta = 1.0 / 6.0e9, 1.0 / 3.0, 1373.0, 1.2, 0.0
    m.Yten0 = 3.0e8
    m.JCA, m.JCB, m.JCN, m.JCC, m.JCM, m.JCTREF, m.JCMaxPlastic = 4.9e7, 1.57e8, 0.167, 0.016, 1.7, 800.0, 10.0
    for k, v in kw.items():
        setattr(m, k, v)
    return m


def Y(fn, m, p=0.0, D=0.0, tps=0.0, rate=0.0, T=293.0):
    pa = (C.c_double * 5)(p, D, tps, rate, T)
    return po.olib().oracle_strength_fn(fn, pa, C.byref(m))


def test_strength_lundborg_and_drucker():
    m = material(Tfrac=1.0e6)  # softening factor 1
    F = ps.SALEC_FN_ROCKSTRENGTHT
    assert Y(F, m, p=0.0) == pytest.approx(m.Yint0, rel=1e-15)
    # slope mu at p = 0 and the Ylim asymptote
    h = 1.0
    assert (Y(F, m, p=h) - Y(F, m, p=0.0)) / h == pytest.approx(m.Yfricint, rel=1e-6)
    assert Y(F, m, p=1e15) == pytest.approx(m.Ylimint, rel=1e-5)
    # fully damaged: min(Lundborg, Drucker-Prager capped at Ylimdam)
    for p in (0.0, 1e6, 1e8, 1e9, 1e10):
        yi = Y(F, m, p=p)
        yd = min(m.Ydam0 + m.Yfricdam * p, m.Ylimdam)
        assert Y(F, m, p=p, D=1.0) == pytest.approx(min(yi, yd), rel=1e-14)
        assert Y(F, m, p=p, D=0.5) == pytest.approx(0.5 * (yi + min(yi, yd)), rel=1e-14)


def test_strength_simple_rock_equals_rock_without_softening():
    m = material(Tfrac=1.0e6)
    for p in (0.0, 3e7, 2e9):
        for D in (0.0, 0.4, 1.0):
            assert Y(ps.SALEC_FN_SIMPLEROCK, m, p=p, D=D) == pytest.approx(
                Y(ps.SALEC_FN_ROCKSTRENGTHT, m, p=p, D=D), rel=1e-15)


def test_strength_thermal_softening_limits():
    m = material()
    F = ps.SALEC_FN_ROCKSTRENGTHT
    p = 1e9
    Tm = m.Tmelt0 * (m.Asimon * p + 1.0) ** m.Csimon   # Simon melting curve
    assert Y(F, m, p=p, T=Tm * 1.001) == 0.0           # above the solidus: no strength
    assert Y(F, m, p=p, T=1.0) == pytest.approx(Y(ps.SALEC_FN_SIMPLEROCK, m, p=p), rel=1e-12)
    mid = Y(F, m, p=p, T=0.5 * Tm)
    assert mid == pytest.approx(np.tanh((Tm / (0.5 * Tm) - 1.0) * m.Tfrac) * Y(ps.SALEC_FN_SIMPLEROCK, m, p=p),
                                rel=1e-12)


def test_ivanov_and_collins_damage_rates():
    m = material()
    for p in (-1e9, 0.0, 1e8, 3e8, 1e9, 1e10):
        assert Y(ps.SALEC_FN_IVANOV, m, p=p) == pytest.approx(1.0 / max(m.IvanA, m.IvanB * (p - m.IvanC)),
                                                              rel=1e-15)
    # Collins: continuous at the brittle-ductile and brittle-plastic pressures
    for pb in (m.Pbd, m.Pbp):
        lo = Y(ps.SALEC_FN_COLLINS, m, p=pb * (1 - 1e-12))
        hi = Y(ps.SALEC_FN_COLLINS, m, p=pb * (1 + 1e-12))
        assert lo == pytest.approx(hi, rel=1e-9)
    assert Y(ps.SALEC_FN_COLLINS, m, p=0.0) == pytest.approx(100.0)


def test_tensile_vonmises_johnsoncook_special_cases():
    m = material(Tfrac=1.0e6)
    assert Y(ps.SALEC_FN_SIMPLETENSILE, m, D=1.0) == 0.0
    assert Y(ps.SALEC_FN_SIMPLETENSILE, m, D=0.0) == pytest.approx(m.Yten0)
    assert Y(ps.SALEC_FN_VONMISES1, m, p=1e9, D=0.0) == m.Yint0
    assert Y(ps.SALEC_FN_VONMISES1, m, p=1e9, D=1.0) == min(m.Yint0, m.Ydam0)
    # Johnson-Cook at zero plastic strain and quasi-static rate: A (times the model's constant factor)
    y0 = Y(ps.SALEC_FN_JOHNSONCOOK1, m)
    y1 = Y(ps.SALEC_FN_JOHNSONCOOK1, m, tps=1.0)
    assert y1 / y0 == pytest.approx((m.JCA + m.JCB) / m.JCA, rel=1e-14)
    y2 = Y(ps.SALEC_FN_JOHNSONCOOK1, m, rate=np.e)
    assert y2 / y0 == pytest.approx(1.0 + m.JCC, rel=1e-14)
    assert Y(ps.SALEC_FN_NONE, m, p=1e9) == 0.0


def wunnemann_mat(alpha0=1.5, alphax=1.1, kappa=0.98, chi=0.5, epse0=-1e-5, avel=1e-4):
    """material with the epsilon-alpha parameters and the derived ones of
    src/sale2d_back.c:371-385"""
    m = ps.Material()
    m.alpha0, m.alphax, m.kappa, m.chi, m.epse0, m.avel = alpha0, alphax, kappa, chi, epse0, avel
    m.alphamax = alpha0
    m.epsex = epse0 + np.log(alphax / alpha0) / kappa
    m.epsec = m.epsex + 2.0 * (1.0 - alphax) / (kappa * alphax)
    m.epsec = min(m.epsec, m.epsex)
    m.alphae = alpha0 * (1.0 + (1.0 - chi * chi) * epse0)
    m.alphamax2 = alpha0
    m.pfunc = ps.SALEC_FN_WUNNEMANN
    return m


def wrate(m, pty, cvs, pre, dcvs, u=10.0, w=0.0):
    p = (C.c_double * 6)(pty, cvs, pre, dcvs, u, w)
    return po.olib().oracle_porosity_wunnemann(p, C.byref(m))


def test_porosity_rate_switches():
    m = wunnemann_mat()
    # no change of fully compacted, over-distended, tensile or resting cells
    assert wrate(m, 1.0, -0.1, 1e8, -1.0) == 0.0
    assert wrate(m, 1.6, -0.1, 1e8, -1.0) == 0.0
    assert wrate(m, 1.3, -0.1, 0.0, -1.0) == 0.0
    assert wrate(m, 1.3, -0.1, 1e8, -1.0, u=1e-5, w=0.0) == 0.0
    # compaction removes distension under compression, and the rate is linear in the strain rate
    r1, r2 = wrate(m, 1.3, -1.0, 1e8, -1.0), wrate(m, 1.3, -1.0, 1e8, -2.0)
    assert r1 < 0.0 and r2 == 2.0 * r1
    # above alpha_x the slope is kappa * alpha (exponential compaction), below it
    # kappa * alpha_x * sqrt((alpha - 1)/(alpha_x - 1)); the two meet at alpha_x
    assert np.isclose(r1, -m.kappa * 1.3, rtol=1e-15)
    lo = wrate(m, m.alphax - 1e-12, -1.0, 1e8, -1.0)
    hi = wrate(m, m.alphax + 1e-12, -1.0, 1e8, -1.0)
    assert np.isclose(lo, hi, rtol=1e-5)
    # the elastic regime (strain above the compaction threshold eps_c(alpha),
    # here -0.146 at alpha 1.3) scales with 1 - c_ratio^2 and vanishes for a
    # matrix sound speed ratio chi = 1
    e = wrate(m, 1.3, -0.1, 1e8, -1.0)
    assert wrate(m, 1.3, -0.15, 1e8, -1.0) == -m.kappa * 1.3
    ratio = 1.0 + (m.chi - 1.0) * (1.3 - 1.0) / (m.alphamax - 1.0)
    assert np.isclose(e, -1.3 * (1.0 - ratio * ratio), rtol=1e-15)
    assert wrate(wunnemann_mat(chi=1.0), 1.3, 1.0, 1e8, -1.0) == 0.0


def test_porosity_compaction_laws_integrate_to_closed_forms():
    """Along a compressive path (strain rate -1, strain below the threshold),
    d(alpha)/d(eps) = kappa alpha gives alpha = alpha_a exp(kappa (eps - eps_a))
    above alpha_x, and kappa alpha_x sqrt((alpha-1)/(alpha_x-1)) gives
    sqrt(alpha - 1) = sqrt(alpha_b - 1) + kappa alpha_x (eps - eps_b) / (2 sqrt(alpha_x - 1))
    below it.  Integrated with classical RK4 in the strain."""
    m = wunnemann_mat()
    f = lambda a: wrate(m, a, -1.0, 1e8, -1.0) / -1.0      # d(alpha)/d(eps)
    def rk4(a, h, n):
        for _ in range(n):
            k1 = f(a); k2 = f(a + 0.5 * h * k1); k3 = f(a + 0.5 * h * k2); k4 = f(a + h * k3)
            a += h / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        return a
    h, n = -1e-4, 1000            # eps decreases by 0.1
    a = rk4(1.45, h, n)
    assert np.isclose(a, 1.45 * np.exp(m.kappa * h * n), rtol=1e-10)
    b0 = 1.08
    b = rk4(b0, h, 200)
    want = (np.sqrt(b0 - 1.0) + m.kappa * m.alphax * (h * 200) / (2.0 * np.sqrt(m.alphax - 1.0))) ** 2 + 1.0
    assert np.isclose(b, want, rtol=1e-8)
