"""Host set-up of spherical targets, the convective and lunar temperature
profiles and target additions (src/target_profile.c), pinned by closed forms.

The reference's target set-up needs its MPI field layer and does not build
here, so these checks pin the host's restatement against what the
mathematics fixes:

* a spherical target of uniform density (an EOS stiff enough that the
  lithostatic pressure changes the density by < 1e-8) has central gravity
  g(r) = -(4/3) pi G rho r inside and -(4/3) pi G rho R (R/r)^2 outside, and
  pressure p(r) = p_s + (2/3) pi G rho^2 (R^2 - r^2); the body force of every
  vertex points at the planet's centre;
* the conductive part of a convective profile follows
  T = T_s + s (z + z^2 / 2R) and its convective part the adiabat
  dT/dz = T gamma_frac |g|; the lunar profile's conductive part is the same
  curve;
* target additions (cube and sphere objects) put their material, state and
  velocity into exactly the cells whose centres they hold, with the EOS state
  of the given (T, p) divided by the distension, and leave the other cells
  as they were.
"""
import math
import os

import numpy as np
import pytest

import pysalec as ps

G = 6.674e-11


def stiff_sphere(tmp_path, R=2005.0, nx=250, ny=500, rho=3000.0, **over):
    so = {"target.type": "sphere", "target.r_planet": R, "target.depth": "[%g]" % (2 * R), "mesh.Oy": 0.9,
          "projectile.center": "[0.0, 100.0]", "projectile.radiu": 30.0}
    so.update(over)
    inp, eos = ps.write_config(str(tmp_path / "sph"), "c0", nx=nx, ny=ny, sections_over=so)
    ps.write_eos(eos, "granite", ps.tillotson(rho, A=1.0e12, B=1.0e12))
    return ps.HostRun(inp, eos)


def profile_grid(R, dy):
    """Profile spacing of build_target_profile (src/target_profile.c:1167-1169)."""
    dz = min(dy * 0.5, R / 3001)
    n = math.ceil(R / dz)
    return dz, n


def test_sphere_central_gravity_and_pressure(tmp_path):
    R, rho = 2005.0, 3000.0          # the centre (0, -R) is no vertex of the 10 m mesh
    h = stiff_sphere(tmp_path, R=R, rho=rho)
    dz, n = profile_grid(R, 10.0)
    vp, bf = h.field("V_POS"), h.field("V_BF")
    d = np.sqrt((vp[0]) ** 2 + (vp[1] + R) ** 2)
    li = np.clip(np.floor(d / dz), 0, n - 1)
    r = R * li / (n - 1)
    gin = -4.0 / 3.0 * math.pi * G * rho * r
    gout = -4.0 / 3.0 * math.pi * G * rho * R * (R / d) ** 2
    expect = np.where(d >= R, gout, gin)
    mag = -np.sqrt(bf[0] ** 2 + bf[1] ** 2)
    deep = (d >= R) | (r >= 0.1 * R)       # the midpoint rule's 1/(4k^2) term stays below 3e-6 there
    assert deep.sum() > 1000 and (d >= R).sum() > 1000
    np.testing.assert_allclose(mag[deep], expect[deep], rtol=1e-5)
    # direction: anti-parallel to the radius vector
    cross = bf[0] * (vp[1] + R) - bf[1] * vp[0]
    dot = bf[0] * vp[0] + bf[1] * (vp[1] + R)
    assert np.all(np.abs(cross) <= 1e-12 * np.abs(mag) * d)
    assert np.all(dot < 0)
    # ANC is off on a sphere
    assert not np.any(h.field("V_LAPLACE"))
    # cells: target inside r_planet (outside the projectile), vacuum outside
    ep, vof = h.field("E_POS"), h.field("M_VOF")
    cd = np.sqrt(ep[0] ** 2 + (ep[1] + R) ** 2)
    pd = np.sqrt(ep[0] ** 2 + (ep[1] - 100.0) ** 2)
    tgt = (cd <= R) & (pd > 30.0)
    assert np.all(vof[1][tgt] == 1.0) and np.all(vof[0][tgt] == 0.0)
    assert np.all(vof[0][(cd > R) & (pd > 30.0)] == 1.0)
    # lithostatic pressure of a uniform sphere
    lc = np.clip(np.floor(cd / dz), 0, n - 1)
    rc = R * lc / (n - 1)
    pexp = 1.0e-5 + 2.0 / 3.0 * math.pi * G * rho ** 2 * (R ** 2 - rc ** 2)
    sel = tgt & (rc <= 0.9 * R)
    assert sel.sum() > 1000
    np.testing.assert_allclose(h.field("E_PRE")[0][sel], pexp[sel], rtol=1e-5)
    np.testing.assert_allclose(h.field("M_DEN")[1][sel], rho, rtol=1e-8)
    assert np.all(h.field("E_TEM")[0][sel] == 293.0)


def test_sphere_layers(tmp_path):
    """Layer thicknesses count down from the surface along the radius."""
    R, top = 2005.0, 200.0
    inp, eos = ps.write_config(str(tmp_path / "lay"), "c1", nx=250, ny=500, sections_over={
        "target.type": "sphere", "target.r_planet": R, "target.depth": "[%g, %g]" % (top, 2 * R), "mesh.Oy": 0.9,
        "projectile.center": "[0.0, 100.0]", "projectile.radiu": 30.0})
    h = ps.HostRun(inp, eos)
    dz, n = profile_grid(R, 10.0)
    ep, vof = h.field("E_POS"), h.field("M_VOF")
    cd = np.sqrt(ep[0] ** 2 + (ep[1] + R) ** 2)
    outer = (cd <= R) & (cd >= R - top + 4 * dz) & (ep[1] < 0)
    inner = cd <= R - top - 4 * dz
    assert outer.sum() > 100 and inner.sum() > 1000
    assert np.all(vof[1][outer] == 1.0) and np.all(vof[2][inner] == 1.0)


DLITH = 1000.0      # the 400-row mesh reaches 2800 m below the surface


def convective_run(tmp_path, kind, ny=400, **over):
    """A plane granite target with a convective or lunar profile; no melt
    cap (SimonT0 1e6 K)."""
    so = {"target.temperature": "[%s, 293.0]" % kind, "target.surf_dtdz": 0.01, "target.d_lith": DLITH,
          "target.r_planet": 1.0e5, "condition.gravity": "[0.0, -100.0]", "material.SimonT0": "[, 1.0e6]"}
    so.update(over)
    return ps.HostRun(*ps.write_config(str(tmp_path / kind), "c0", nx=16, ny=ny, sections_over=so))


def plane_profile_depths(h, ny, dy=10.0):
    """Depth below the surface of the profile node each target cell takes
    (init_material_plane_distribution, src/target_profile.c:640-652)."""
    max_depth = ny * dy * 1.5
    dz = min(dy * 0.5, max_depth / 3001)
    n = math.ceil(max_depth / dz)
    ey = h.field("E_POS")[1][:, 4]
    li = n - 1 - (np.floor(np.abs(ey) / max_depth * (n - 1)) + 1)
    li = np.clip(li, 0, n - 1)
    rad = max_depth * li / (n - 1)
    kd = max_depth * (n - 1) / (n - 1) - rad
    return ey, kd


def test_convective_profile(tmp_path):
    ny, R, s, Ts = 400, 1.0e5, 0.01, 293.0
    h = convective_run(tmp_path, "convective", ny=ny)
    ey, kd = plane_profile_depths(h, ny)
    T = h.field("E_TEM")[0][:, 4]
    tgt = ey < -1.0
    lith = tgt & (kd < 0.85 * DLITH)
    assert lith.sum() > 50
    np.testing.assert_allclose(T[lith], Ts + s * (kd[lith] + kd[lith] ** 2 / (2 * R)), rtol=1e-5)
    # below the lithosphere: the adiabat, d ln T / dz = gamfrac |g|
    mat = h.cfg.mat[1]
    cc = mat.gamfrac * 100.0
    conv = tgt & (kd > DLITH + 50.0)
    assert conv.sum() > 100
    z, t = kd[conv], T[conv]
    slope = np.diff(np.log(t)) / np.diff(z)
    np.testing.assert_allclose(slope[np.diff(z) > 0], cc, rtol=1e-3)
    assert cc * (z.max() - z.min()) > 0.04          # the adiabat changes T measurably


def test_lunar_profile_conductive_part(tmp_path):
    # a steep gradient, so that 0.9 t_lith lies well above the surface temperature
    ny, R, s, Ts = 400, 1.0e5, 1.0, 293.0
    h = convective_run(tmp_path, "lunar", ny=ny, **{"target.surf_dtdz": s})
    ey, kd = plane_profile_depths(h, ny)
    T = h.field("E_TEM")[0][:, 4]
    # t_lith from d_lith (ConsistentLith); conduction only while T < 0.9 t_lith
    t_lith = Ts + s * DLITH * (1.0 - 0.5 * DLITH / R)
    sel = (ey < -1.0) & (T < 0.85 * t_lith)
    assert sel.sum() > 50
    np.testing.assert_allclose(T[sel], Ts + s * (kd[sel] + kd[sel] ** 2 / (2 * R)), rtol=1e-5)
    assert np.all(np.isfinite(T))


def test_profile_options_rejected(tmp_path):
    for over, what in (({"target.type": "torus"}, "target.type"),
                       ({"addition.nobject": 1, "addition.material": "[granite]", "addition.type": "[cone]",
                         "addition.parameter": "[p]", "addition.p": "[1, 2, 3]"}, "addition"),
                       ({"addition.nobject": 1, "addition.material": "[granite]", "addition.type": "[sphere]",
                         "addition.parameter": "[p]"}, "parameter")):
        with pytest.raises(RuntimeError, match=what):
            ps.HostRun(*ps.write_config(str(tmp_path / "bad"), "c0", nx=16, ny=24, sections_over=over))


def test_target_additions(tmp_path):
    base_over = {"projectile.center": "[0.0, 300.0]"}
    add = dict(base_over)
    add.update({"addition.nobject": 2, "addition.material": "[granite, vacuum_]", "addition.type": "[sphere, cube]",
                "addition.parameter": "[ball, box]",
                "addition.ball": "[300.0, 200.0, 80.0, 5.0, -7.0, 400.0, 2.0e6, 0.5, 1.25]",
                "addition.box": "[400.0, -100.0, 100.0, 0.0, 0.0, 50.0, 0.0, 0.0, -1, -1, -1, -1]"})
    h0 = ps.HostRun(*ps.write_config(str(tmp_path / "a0"), "c0", nx=64, ny=96, sections_over=base_over))
    h1 = ps.HostRun(*ps.write_config(str(tmp_path / "a1"), "c0", nx=64, ny=96, sections_over=add))
    ep = h1.field("E_POS")
    ball = np.hypot(300.0 - ep[0], 200.0 - ep[1]) <= 80.0
    xl, yl = (ep[0] - 400.0) / 100.0, (ep[1] + 100.0) / 50.0
    box = (xl >= 0) & (xl <= 1) & (yl >= 0) & (yl <= 1)
    assert ball.sum() > 100 and box.sum() > 20 and not np.any(ball & box)
    vof, den, eng, pty = h1.field("M_VOF"), h1.field("M_DEN"), h1.field("M_ENG"), h1.field("M_PTY")
    d, e = ps.eos_tp(h1.cfg.mat[1], 400.0, 2.0e6)
    assert np.all(vof[1][ball] == 1.0) and np.all(vof[0][ball] == 0.0)
    assert np.all(den[1][ball] == d / 1.25) and np.all(eng[1][ball] == e / 1.25)
    assert np.all(pty[1][ball] == 1.25)
    assert np.all(h1.field("E_TEM")[0][ball] == 400.0) and np.all(h1.field("E_PRE")[0][ball] == 2.0e6)
    assert np.all(h1.field("E_DAM")[0][ball] == 0.5)
    vel = h1.field("E_VEL")
    assert np.all(vel[0][ball] == 5.0) and np.all(vel[1][ball] == -7.0)
    assert np.all(h1.field("E_WPT")[:, ball] == -0.25)
    assert np.all(vof[0][box] == 1.0) and np.all(vof[1][box] == 0.0) and np.all(den[:, box] == 0.0)
    assert np.all(h1.field("E_WPT")[:, box] == 1.0)
    rest = ~(ball | box)
    for name in ("M_VOF", "M_DEN", "M_ENG", "M_PTY", "E_PRE", "E_TEM", "E_DAM", "E_VEL", "E_WPT"):
        assert np.array_equal(h0.field(name)[..., rest], h1.field(name)[..., rest]), name
    # the box keeps the cells' temperature and pressure (negative entries)
    assert np.array_equal(h0.field("E_PRE")[0][box], h1.field("E_PRE")[0][box])


def test_complex_crater_config_sets_up(tmp_path):
    """The 3000 x 4000 complex-crater deck, scaled down: four materials, a
    spherical target with central gravity and a conductive profile."""
    h = ps.HostRun(*ps.write_config(str(tmp_path / "c4"), "c4", nx=60, ny=80))
    assert h.nmat == 4
    vof = h.field("M_VOF")
    for m in range(4):
        assert vof[m].max() == 1.0, m
    bf = h.field("V_BF")
    g = np.hypot(bf[0], bf[1])
    assert 1.4 < g.min() and g.max() < 1.7                # lunar surface gravity
    assert np.any(bf[0] != 0.0)                           # not uniform: it points at the centre


def test_vertex_on_the_planet_centre(tmp_path):
    """vec_norm turns the 0/0 direction of a vertex on the centre into
    (1, 1) (src/geometry/linear.c:116-131); g there is 0, so its body force
    is zero and finite."""
    R = 2000.0                                  # the centre (0, -2000) is a vertex of the 10 m mesh
    h = stiff_sphere(tmp_path, R=R)
    vp, bf = h.field("V_POS"), h.field("V_BF")
    at = (vp[0] == 0.0) & (vp[1] == -R)
    assert at.sum() == 1
    assert np.all(bf[:, at] == 0.0) and np.all(np.isfinite(bf))
