"""Closed-form pins of the oracle's step functions, one function at a time
(oracle_phase), on the real mesh geometry of the host set-up.

Each test builds a state whose answer the mathematics fixes and checks the
function against it; the cases are chosen so that a dropped term, a wrong
sign, a swapped component or a wrong index fails:

* velocity gradient (update_e_gradvel, src/sale2d.c:1144-1235): exact for
  linear velocity fields on uniform and stretched meshes; the hoop term is
  v_r / r at the cell centre; artificial viscosity zero in expansion,
  quadratic (linear) in the compression rate for the quadratic (linear)
  coefficient alone;
* CFL time step (calculate_dt, src/sale2d.c:1959-2016): 0.2 h / (|u| + c)
  on a uniform mesh, with the 1.1 growth cap;
* cell velocity (update_e_vel, src/sale2d.c:838-927): the vertex value of a
  uniform field, the centre value of a linear one;
* vertex volume fractions (update_v_vof): the average of the four cells, a
  partition of unity;
* vertex momentum (update_v_vel_inc, src/sale2d.c:930-1048): a uniform
  increment adds to a uniform velocity;
* nodal acceleration (calculate_acc_default, src/velocity.c:9-155): zero
  net force of a uniform pressure, -dp/dz / rho of a linear one, and the
  axisymmetric divergence ((S_rr - S_tt) / r, S_rz / r) / rho of a uniform
  deviatoric stress, exactly on a uniform mesh and to 1 % on a stretched one;
* displacement (calculate_v_dis): constant-acceleration kinematics summed
  over steps;
* internal energy (update_energy, src/sale2d.c:564-712): pressure work
  -(p + q) div v and deviatoric power S_rz (dv_r/dz + dv_z/dr);
* block vibration (block_vibration, src/strength.c:625-695): decay close to
  exp(-2 dt / T_d), excitation to (C_vib |v|)^2 capped at VibMax^2;
* tracers (fl_update_node, src/tracer.c:52-175): constant-acceleration
  kinematics in a cell open to vacuum, the mesh displacement of a linear
  field inside the material;
* crater profile (trace_profile, src/ejecta.c:8-63): a flat surface gives
  the centre of the first vacuum row in every column;
* ejecta flow (src/ejecta.c:65-134): linear drag and Coulomb friction
  increments;
* damping layers (src/velocity.c, src/advect.c:304-593): the split
  divergence and stress power of the pml2/pml3 cell auxiliaries, the split
  stress force of the vertex auxiliary, the pml3 flux auxiliaries and their
  coupling into the remap, both sponges;
* remap: flat and inclined interfaces, a row of mixed cells, a uniform
  state.
"""
import ctypes as C
import math
import os
import tempfile

import numpy as np
import pytest

import pyoracle as po
import pysalec as ps

IN = (slice(2, -2), slice(2, -2))          # physical cells
VIN = (slice(3, -3), slice(3, -3))         # vertices away from the padding
STRETCH = {"mesh.ext": "on", "mesh.ex": "[1.04, 6, 6]", "mesh.ey": "[1.04, 6, 6]"}


def fresh(which="c0", nx=24, ny=36, **over):
    h = ps.HostRun(*ps.write_config(os.path.join(tempfile.mkdtemp(prefix="pin_"), which), which, nx=nx, ny=ny,
                                    sections_over=over))
    return h, po.Oracle(h.cfg, h.state())


@pytest.mark.parametrize("mesh", ["uniform", "stretched"])
def test_gradient_of_linear_velocity(mesh):
    h, o = fresh(**(STRETCH if mesh == "stretched" else {}))
    S = o.state
    r, z = S["V_POS"][0], S["V_POS"][1]
    a, b, c0, d, e, c1 = 0.3, -0.7, 2.0, 1.1, 0.5, -3.0
    S["V_VEL"][0] = a * r + b * z + c0
    S["V_VEL"][1] = d * r + e * z + c1
    o.phase("GRADVEL")
    g = S["E_GRAD_VEL"]
    for i, x in enumerate((a, b, d, e)):          # dvr/dr, dvr/dz, dvz/dr, dvz/dz
        np.testing.assert_allclose(g[i][IN], x, rtol=0, atol=1e-13)
    rc, zc = S["E_POS"][0][IN], S["E_POS"][1][IN]
    hoop = (a * rc + b * zc + c0) / rc            # v_r / r at the centre
    np.testing.assert_allclose(g[4][IN], hoop, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(S["E_DIV_VEL"][0][IN], a + e + hoop, rtol=1e-12, atol=1e-13)


def test_artificial_viscosity_scaling():
    def q_of(rate, avis1, avis2):
        h, o = fresh(**{"numerical.AVISL": str(avis1), "numerical.AVISQ": str(avis2)})
        S = o.state
        S["M_VOF"][0] = 0.0
        S["M_VOF"][1] = 1.0
        S["E_PRE"][0] = 1.0e8
        S["E_CSD"][0] = 3000.0
        S["E_DEN"][0] = 2500.0
        r, z = S["V_POS"][0], S["V_POS"][1]
        S["V_VEL"][0] = 0.0 * r
        S["V_VEL"][1] = rate * z                  # uniaxial strain rate, div = rate
        o.phase("GRADVEL")
        return S["E_Q"][0][IN].copy()
    assert np.all(q_of(+5.0, 0.5, 1.2) == 0.0)    # expansion: no viscosity
    q1, q2 = q_of(-5.0, 0.0, 1.2), q_of(-10.0, 0.0, 1.2)
    assert np.all(q1 > 0)
    np.testing.assert_allclose(q2, 4.0 * q1, rtol=1e-12)      # quadratic term ~ div^2
    l1, l2 = q_of(-5.0, 0.5, 0.0), q_of(-10.0, 0.5, 0.0)
    np.testing.assert_allclose(l2, 2.0 * l1, rtol=1e-12)      # linear term ~ |div|


def test_cfl_time_step():
    h, o = fresh(**{"cycle.maxdt": "10.0"})
    S = o.state
    c, u = 2500.0, 400.0
    S["E_CSD"][0] = c
    S["E_VEL"][0] = u
    S["E_VEL"][1] = 0.0
    o.phase("DT", dt=1.0)
    assert o.dt == pytest.approx(0.2 * 10.0 / (u + c), rel=1e-15)    # uniform 10 m cells
    o.phase("DT", dt=1.0e-4)
    assert o.dt == pytest.approx(1.1e-4, rel=1e-15)                   # growth cap


def test_cell_velocity_from_vertices():
    h, o = fresh(**STRETCH)
    S = o.state
    S["E_DEN"][0] = 1000.0
    S["E_WPT"][:] = -0.25
    S["V_MASS"][0] = 1.0 + np.random.default_rng(1).random(S["V_MASS"][0].shape)
    S["V_MASS"][1] = 0.0
    S["V_VEL"][0] = 7.5
    S["V_VEL"][1] = -2.25
    o.phase("E_VEL")
    np.testing.assert_allclose(S["E_VEL"][0][IN], 7.5, rtol=1e-14)
    np.testing.assert_allclose(S["E_VEL"][1][IN], -2.25, rtol=1e-14)
    # a linear field with equal masses: the average of the corners
    r, z = S["V_POS"][0], S["V_POS"][1]
    S["V_MASS"][0] = 1.0
    S["V_VEL"][0] = 0.2 * r - 0.1 * z
    o.phase("E_VEL")
    vv = S["V_VEL"][0]
    corner = 0.25 * (vv[:-1, :-1] + vv[:-1, 1:] + vv[1:, 1:] + vv[1:, :-1])
    np.testing.assert_allclose(S["E_VEL"][0][IN], corner[IN], rtol=1e-13)


def test_vertex_volume_fractions_partition_of_unity():
    h, o = fresh("c1")
    S = o.state
    rng = np.random.default_rng(2)
    w = rng.random(S["M_VOF"].shape)
    S["M_VOF"][:] = w / w.sum(axis=0)
    o.phase("V_VOF")
    vof = S["V_VOF"]
    np.testing.assert_allclose(vof.sum(axis=0)[1:-1, 1:-1], 1.0, rtol=0, atol=4e-16)
    m = S["M_VOF"]
    avg = 0.25 * (m[:, :-1, :-1] + m[:, :-1, 1:] + m[:, 1:, 1:] + m[:, 1:, :-1])
    np.testing.assert_allclose(vof[:, 1:-1, 1:-1], avg, rtol=1e-15, atol=1e-16)


def test_vertex_momentum_increment():
    h, o = fresh(**{"numerical.MaxVelocity": "1.0e6", "numerical.MinVelocity": "1.0e-8"})
    S = o.state
    S["E_DEN"][0] = 2000.0
    S["E_WPT"][:] = -0.25
    S["E_DVEL"][0] = 3.0
    S["E_DVEL"][1] = -1.5
    vol = S["V_VOL"][1:5].sum(axis=0)
    S["V_MASS"][0] = 2000.0 * vol                  # the mass the vertex will get
    S["V_VEL"][0] = 10.0
    S["V_VEL"][1] = 20.0
    o.phase("V_VEL_INC")
    np.testing.assert_allclose(S["V_VEL"][0][VIN], 13.0, rtol=1e-14)
    np.testing.assert_allclose(S["V_VEL"][1][VIN], 18.5, rtol=1e-14)


@pytest.mark.parametrize("mesh", ["uniform", "stretched"])
def test_nodal_acceleration_closed_forms(mesh):
    over = {"numerical.damping": "off", "numerical.ANC": "0.0"}
    if mesh == "stretched":
        over.update(STRETCH)
    h, o = fresh(**over)
    S = o.state
    rho = 2000.0
    S["E_DEN"][0] = rho
    S["E_Q"][:] = 0.0
    S["E_STE"][:] = 0.0
    S["V_MASS"][:] = 0.0
    bf = S["V_BF"]
    acc = S["V_ACC"]
    # a uniform pressure exerts no net force
    S["E_PRE"][0] = 1.0e9
    o.phase("ACC", dt=1e-3)
    np.testing.assert_allclose(acc[:, VIN[0], VIN[1]], bf[:, VIN[0], VIN[1]], rtol=0, atol=1e-11)
    # a linear pressure p = p0 + k z: a_z = g_z - k / rho
    k = 1.0e5
    S["E_PRE"][0] = 1.0e9 + k * S["E_POS"][1]
    o.phase("ACC", dt=1e-3)
    np.testing.assert_allclose(acc[1][VIN], bf[1][VIN] - k / rho, rtol=0, atol=5e-11)
    assert np.all(acc[0][VIN] == bf[0][VIN])
    # a uniform deviatoric stress: ((S_rr - S_tt) / r, S_rz / r) / rho
    S["E_PRE"][0] = 0.0
    S["E_STE"][0], S["E_STE"][1], S["E_STE"][2], S["E_STE"][3] = 3.0e6, 1.0e6, -1.0e6, -2.0e6
    o.phase("ACC", dt=1e-3)
    r = S["V_POS"][0][VIN]
    tol = 1e-13 if mesh == "uniform" else 2e-2
    np.testing.assert_allclose(acc[0][VIN] - bf[0][VIN], (3.0e6 + 2.0e6) / (rho * r), rtol=tol)
    np.testing.assert_allclose(acc[1][VIN] - bf[1][VIN], 1.0e6 / (rho * r), rtol=tol)


def test_displacement_kinematics():
    h, o = fresh()
    S = o.state
    a, v0, dt, n = np.array([1.5, -9.0]), np.array([4.0, 2.0]), 1.0e-3, 40
    S["V_VEL"][0], S["V_VEL"][1] = v0
    total = np.zeros_like(S["V_DIS"])
    for _ in range(n):
        S["V_ACC"][0], S["V_ACC"][1] = a
        o.phase("V_DIS", dt=dt)
        total += S["V_DIS"]
    T = n * dt
    for c in range(2):
        np.testing.assert_allclose(total[c][1:-1, 1:-1], v0[c] * T + 0.5 * a[c] * T * T, rtol=1e-12)
        np.testing.assert_allclose(S["V_VEL"][c][1:-1, 1:-1], v0[c] + a[c] * T, rtol=1e-13)


def test_energy_work_terms():
    h, o = fresh("c1")
    S = o.state
    S["M_VOF"][:] = 0.0
    S["M_VOF"][1] = 1.0
    S["M_DEN"][1] = 1800.0
    S["E_DEN"][0] = 1800.0
    S["M_ENG"][1] = 5.0e8
    dt = 1e-4
    # pure shear: S_rz (dv_r/dz + dv_z/dr)
    tau, g1, g2 = 2.0e6, 0.3, -0.1
    S["E_STE"][:] = 0.0
    S["E_STE"][1] = tau
    S["E_GRAD_VEL"][:] = 0.0
    S["E_GRAD_VEL"][1], S["E_GRAD_VEL"][2] = g1, g2
    S["E_DIV_VEL"][0] = 0.0
    S["E_PRE"][0] = 0.0
    S["E_Q"][0] = 0.0
    o.phase("ENERGY", dt=dt)
    np.testing.assert_allclose(S["M_ENG"][1][IN], 5.0e8 + tau * (g1 + g2) * dt, rtol=1e-15)
    # pure compression: -(p + q) div v
    S["M_ENG"][1] = 5.0e8
    S["E_STE"][:] = 0.0
    S["E_GRAD_VEL"][:] = 0.0
    S["E_DIV_VEL"][0] = -2.0
    S["E_PRE"][0] = 3.0e7
    S["E_Q"][0] = 1.0e6
    o.phase("ENERGY", dt=dt)
    np.testing.assert_allclose(S["M_ENG"][1][IN], 5.0e8 + 3.1e7 * 2.0 * dt, rtol=1e-15)


def test_block_vibration_decay_and_excitation():
    over = {"material.Toff": "[, 16.0]", "material.Cvib": "[, 0.1]", "material.VibMax": "[, 200.0]"}
    h, o = fresh(**over)
    S = o.state
    S["M_VOF"][0] = 0.0
    S["M_VOF"][1] = 1.0
    S["E_DAM"][0] = 1.0
    S["E_Q"][0] = 0.0
    S["E_VIB"][0] = 100.0
    Td = h.cfg.mat[1].Tdamp
    for x in (1e-3, 1e-2):
        S["E_VIB"][0] = 100.0
        o.phase("VIBRATION", dt=x * Td, t=1.0)
        f = S["E_VIB"][0][IN] / 100.0
        assert np.all(np.abs(f - math.exp(-2 * x)) <= 4.0 / 3.0 * x ** 3 + 1e-15)
    # excitation while t <= Toff, q > 0 and C_vib |v| >= 1
    S["E_VIB"][0] = 0.0
    S["E_Q"][0] = 1.0
    S["E_VEL"][0], S["E_VEL"][1] = 30.0, 40.0          # |v| = 50, C_vib |v| = 5
    o.phase("VIBRATION", dt=1e-3, t=1.0)
    np.testing.assert_allclose(S["E_VIB"][0][IN], 25.0, rtol=1e-15)
    S["E_VIB"][0] = 0.0
    S["E_VEL"][0], S["E_VEL"][1] = 3.0e4, 4.0e4        # capped at VibMax^2
    o.phase("VIBRATION", dt=1e-3, t=1.0)
    np.testing.assert_allclose(S["E_VIB"][0][IN], 200.0 ** 2, rtol=1e-15)


def test_tracer_kinematics():
    h, o0 = fresh("c1")
    tr = h.tracers(o0.state["V_VEL"])
    o = po.Oracle(h.cfg, o0.state, tracers=tr)
    S = o.state
    n = o.tr.n
    assert n > 10
    dt = 1e-3
    p0 = np.array(o.tr.pos, copy=True).reshape(2, n)
    # inside the material: the mesh displacement, a linear field, interpolated exactly
    S["M_VOF"][:] = 0.0
    S["M_VOF"][1] = 1.0
    r, z = S["V_POS"][0], S["V_POS"][1]
    S["V_DIS"][0] = 1e-3 * r + 2e-3 * z + 0.01
    S["V_DIS"][1] = -3e-3 * r + 0.02
    S["V_VEL"][0] = 5.0
    S["V_VEL"][1] = -1.0
    o.phase("TRACERS", dt=dt)
    p1 = np.array(o.tr.pos, copy=True).reshape(2, n)
    alive = np.array(o.tr.alive, copy=True) == 1
    np.testing.assert_allclose(p1[0][alive] - p0[0][alive], 1e-3 * p0[0][alive] + 2e-3 * p0[1][alive] + 0.01,
                               rtol=1e-10)
    np.testing.assert_allclose(p1[1][alive] - p0[1][alive], -3e-3 * p0[0][alive] + 0.02, rtol=1e-10)
    v1 = np.array(o.tr.vel, copy=True).reshape(2, n)
    assert np.allclose(v1[0][alive], 5.0) and np.allclose(v1[1][alive], -1.0)
    # open to vacuum: ballistic, x + u dt + a dt^2 / 2 and u + a dt
    S["M_VOF"][0] = 0.5
    S["M_VOF"][1] = 0.5
    S["V_ACC"][0] = 2.0
    S["V_ACC"][1] = -9.0
    o.phase("TRACERS", dt=dt)
    p2 = np.array(o.tr.pos, copy=True).reshape(2, n)
    v2 = np.array(o.tr.vel, copy=True).reshape(2, n)
    alive = np.array(o.tr.alive, copy=True) == 1
    np.testing.assert_allclose(p2[0][alive] - p1[0][alive], 5.0 * dt + 0.5 * 2.0 * dt * dt, rtol=0, atol=1e-12)
    np.testing.assert_allclose(p2[1][alive] - p1[1][alive], -1.0 * dt - 0.5 * 9.0 * dt * dt, rtol=0, atol=1e-12)
    np.testing.assert_allclose(v2[0][alive], 5.0 + 2.0 * dt, rtol=1e-14)
    np.testing.assert_allclose(v2[1][alive], -1.0 - 9.0 * dt, rtol=1e-14)


def test_crater_profile_of_a_flat_surface():
    h, o = fresh("c1")
    S = o.state
    o.phase("PROFILE")
    yp = S["Y_PROFILE"][0][2:-2]
    ez = S["E_POS"][1][:, 2]
    first_vac = ez[ez > 0].min()                   # centre of the first vacuum row: 5 m
    assert first_vac == 5.0
    # the projectile (vacuum around it) sits above the surface near the axis
    np.testing.assert_array_equal(yp[h.nx // 2:], first_vac)


def test_ejecta_drag_and_friction():
    over = {"ejecta.resistance": "0.5", "ejecta.friction": "0.2"}
    h, o = fresh(**over)
    S = o.state
    S["Y_PROFILE"][:] = 1.0e9                      # every cell lies below the profile
    S["E_EJT"][0] = 1.0
    S["E_DVEL"][:] = 0.0
    S["E_VEL"][0] = 3.0
    S["E_VEL"][1] = -4.0
    S["E_STE"][:] = 1.0
    dt = 1e-3
    o.phase("EJECTA", dt=dt)
    gy = 1.62                                      # condition.gravity, away from the bottom rows (no body force there)
    dv0 = -0.5 * 3.0 * dt - 0.2 * gy * dt          # drag against u, friction against sign(u)
    dv1 = -0.5 * -4.0 * dt
    up = (slice(5, -2), slice(2, -2))
    np.testing.assert_allclose(S["E_DVEL"][0][up], dv0, rtol=1e-14)
    np.testing.assert_allclose(S["E_DVEL"][1][IN], dv1, rtol=1e-14)
    S["E_DVEL"][:] = 0.0
    S["E_VEL"][0] = -3.0                           # friction flips with the direction of motion
    o.phase("EJECTA", dt=dt)
    np.testing.assert_allclose(S["E_DVEL"][0][up], 0.5 * 3.0 * dt + 0.2 * gy * dt, rtol=1e-14)
    assert np.all(S["E_STE"][:, IN[0], IN[1]] == 0.0)


# ---------------------------------------------------------------- stress --
VONMISES = {"material.yshear": "[, VonMises1]", "material.damage": "[, None]", "material.ytens": "[, None]"}


def solid(Y=1.0e14, **over):
    """c0 granite, solid (density rho0), constant yield strength Y."""
    o2 = dict(VONMISES)
    o2["material.yint0"] = "[, %r]" % Y
    o2["material.ydam0"] = "[, %r]" % Y
    o2.update(over)
    h, o = fresh(**o2)
    S = o.state
    S["M_VOF"][0] = 0.0
    S["M_VOF"][1] = 1.0
    S["E_DEN"][0] = h.cfg.mat[1].den0
    S["E_DAM"][0] = 0.0
    S["E_TPS"][0] = 0.0
    S["E_STE2S"][0] = 0.0
    S["E_STA2S"][0] = 0.0
    S["E_GRAD_VEL"][:] = 0.0
    return h, o, S


def j2(s):
    return ((s[0] - s[2]) ** 2 + (s[2] - s[3]) ** 2 + (s[3] - s[0]) ** 2) / 6.0 + s[1] ** 2


def test_stress_rotation_keeps_the_invariant():
    """Pure spin with no elastic increment (mu = 0): the Jaumann rotation
    turns the in-plane deviator and leaves J2 unchanged."""
    h, o, S = solid()
    S["E_MU"][0] = 0.0
    st = np.array([3.0e6, 1.2e6, -1.0e6, -2.0e6])
    for i in range(4):
        S["E_STE"][i] = st[i]
    S["E_GRAD_VEL"][1], S["E_GRAD_VEL"][2] = -40.0, 40.0          # dv_r/dz = -dv_z/dr: rotation only
    o.phase("STRESS", dt=1e-3)
    new = S["E_STE"][:, IN[0], IN[1]]
    np.testing.assert_allclose(j2(new), j2(st), rtol=1e-12)
    assert np.all(new[3] == st[3])                                  # the hoop component does not turn
    np.testing.assert_allclose(new[0] + new[2], st[0] + st[2], rtol=1e-14)
    assert not np.allclose(new[1], st[1])                           # it did turn
    np.testing.assert_allclose(S["E_STE2S"][0][IN], np.sqrt(j2(st)), rtol=1e-12)


def test_stress_hooke_deviator():
    """Elastic increment from zero stress, no spin: 2 mu dt times the
    deviatoric strain rate (uniaxial in z, then simple shear)."""
    h, o, S = solid()
    mu, dt, ez = 3.0e10, 1e-4, 2.0
    S["E_MU"][0] = mu
    S["E_STE"][:] = 0.0
    S["E_GRAD_VEL"][3] = ez
    o.phase("STRESS", dt=dt)
    st = S["E_STE"][:, IN[0], IN[1]]
    np.testing.assert_allclose(st[2], 4.0 / 3.0 * mu * ez * dt, rtol=1e-14)
    np.testing.assert_allclose(st[0], -2.0 / 3.0 * mu * ez * dt, rtol=1e-14)
    np.testing.assert_allclose(st[3], -2.0 / 3.0 * mu * ez * dt, rtol=1e-14)
    assert np.all(st[1] == 0.0)
    S["E_STE"][:] = 0.0
    S["E_GRAD_VEL"][:] = 0.0
    S["E_GRAD_VEL"][1] = S["E_GRAD_VEL"][2] = 0.5                   # engineering shear rate 1
    o.phase("STRESS", dt=dt)
    st = S["E_STE"][:, IN[0], IN[1]]
    np.testing.assert_allclose(st[1], mu * 1.0 * dt, rtol=1e-14)
    assert np.all(st[0] == 0.0) and np.all(st[2] == 0.0)


def test_stress_radial_return_to_the_yield_surface():
    Y = 5.0e6
    h, o, S = solid(Y=Y)
    S["E_MU"][0] = 3.0e10
    st = np.array([3.0e7, 1.0e7, -1.0e7, -2.0e7])                   # far outside the surface
    for i in range(4):
        S["E_STE"][i] = st[i]
    o.phase("STRESS", dt=1e-6)
    new = S["E_STE"][:, IN[0], IN[1]]
    np.testing.assert_allclose(np.sqrt(j2(new)), Y, rtol=1e-12)
    np.testing.assert_allclose(S["E_STE2S"][0][IN], Y, rtol=1e-14)
    np.testing.assert_allclose(S["E_STRENGTH"][0][IN], Y, rtol=1e-14)
    ratio = new / st[:, None, None]
    np.testing.assert_allclose(ratio, np.broadcast_to(ratio[0], ratio.shape), rtol=1e-9)   # scaled, not turned


# --------------------------------------------------------------- mixture --
def test_mixture_closure():
    h, o = fresh("c1")
    S = o.state
    m1, m2 = h.cfg.mat[1], h.cfg.mat[2]
    d1, e1, d2, e2 = 1900.0, 1900.0 * 2.0e5, 2900.0, 2900.0 * 5.0e5
    ref1 = np.zeros(3)
    ref2 = np.zeros(3)
    L = po.olib()
    for mat, d, e, ref in ((m1, d1, e1, ref1), (m2, d2, e2, ref2)):
        t, p, c = C.c_double(), C.c_double(), C.c_double()
        L.oracle_calculate_state_correct(C.byref(mat), d, e, -1.0e30, C.byref(t), C.byref(p), C.byref(c), 1.0)
        ref[:] = (t.value, p.value, c.value)
    S["E_DAM"][0] = 0.0
    S["M_PTY"][:] = 1.0
    S["M_DEN"][1], S["M_ENG"][1], S["M_DEN"][2], S["M_ENG"][2] = d1, e1, d2, e2
    # one material: its own state
    S["M_VOF"][:] = 0.0
    S["M_VOF"][1] = 1.0
    o.update_state()
    assert np.all(S["E_TEM"][0][IN] == ref1[0]) and np.all(S["E_DEN"][0][IN] == d1)
    # two materials: volume averages of p, T, rho; harmonic average of c
    S["M_VOF"][1], S["M_VOF"][2] = 0.25, 0.75
    o.update_state()
    np.testing.assert_allclose(S["E_TEM"][0][IN], 0.25 * ref1[0] + 0.75 * ref2[0], rtol=1e-14)
    np.testing.assert_allclose(S["E_DEN"][0][IN], 0.25 * d1 + 0.75 * d2, rtol=1e-15)
    assert ref1[1] > 0 and ref2[1] > 0
    np.testing.assert_allclose(S["E_PRE"][0][IN], 0.25 * ref1[1] + 0.75 * ref2[1], rtol=1e-14)
    np.testing.assert_allclose(S["E_CSD"][0][IN], 1.0 / (0.25 / ref1[2] + 0.75 / ref2[2]), rtol=1e-14)
    # open to vacuum (no partial pressure): no pressure, unit sound speed
    S["M_VOF"][0], S["M_VOF"][1], S["M_VOF"][2] = 0.1, 0.2, 0.7
    o.update_state()
    assert np.all(S["E_PRE"][0][IN] == 0.0) and np.all(S["E_CSD"][0][IN] == 1.0) and np.all(S["E_MU"][0][IN] == 0.0)


# --------------------------------------------------------------- damping --
def test_pml_velocity_relaxation():
    """pml2: in a damping cell the velocity increment relaxes the velocity at
    the rate sigma_x + sigma_y (du/dt = -sigma u)."""
    h, o = fresh(**{"numerical.damping": "pml2"})
    S = o.state
    S["M_VOF"][0] = 0.0
    S["M_VOF"][1] = 1.0
    S["E_DVEL"][:] = 0.0
    S["E_VEL"][0], S["E_VEL"][1] = 2.0, -3.0
    h.cfg.damp_time = 1.0e9
    o.phase("DAMP", dt=1e-3, t=0.0)
    sig = S["E_SIGMA"][0] + S["E_SIGMA"][1]
    lay = (sig > 0)[IN]
    assert lay.sum() > 20
    np.testing.assert_allclose(S["E_DVEL"][0][IN][lay], -sig[IN][lay] * 2.0 * 1e-3, rtol=1e-14)
    np.testing.assert_allclose(S["E_DVEL"][1][IN][lay], sig[IN][lay] * 3.0 * 1e-3, rtol=1e-14)
    assert np.all(S["E_DVEL"][:, IN[0], IN[1]][:, ~lay] == 0.0)


# ---------------------------------------------------------------- ghosts --
def test_axis_reflection_and_outflow_ghosts():
    """On the axis (mirror boundary) the ghost columns reflect the first two
    physical columns, with the radial velocity increment reversed; on an
    outflow side they repeat them as the same reflection without a sign."""
    h, o = fresh("c1", **{"boundary.right": "outflow"})
    S = o.state
    rng = np.random.default_rng(4)
    f = rng.standard_normal(S["E_DVEL"].shape)
    S["E_DVEL"][:] = f
    o.ghost_fill("E_DVEL", 2)                       # velocity increment
    rows = slice(2, -2)
    np.testing.assert_array_equal(S["E_DVEL"][0][rows, 1], -S["E_DVEL"][0][rows, 2])
    np.testing.assert_array_equal(S["E_DVEL"][0][rows, 0], -S["E_DVEL"][0][rows, 3])
    np.testing.assert_array_equal(S["E_DVEL"][1][rows, 1], S["E_DVEL"][1][rows, 2])
    np.testing.assert_array_equal(S["E_DVEL"][1][rows, 0], S["E_DVEL"][1][rows, 3])
    g = rng.standard_normal(S["E_PRE"].shape)
    S["E_PRE"][:] = g
    o.ghost_fill("E_PRE", 0)                        # copied on every side
    np.testing.assert_array_equal(S["E_PRE"][0][rows, -2], S["E_PRE"][0][rows, -3])
    np.testing.assert_array_equal(S["E_PRE"][0][rows, -1], S["E_PRE"][0][rows, -4])
    np.testing.assert_array_equal(S["E_PRE"][0][rows, 1], S["E_PRE"][0][rows, 2])
    # the density's top side is cleaned, the first physical row included
    S["E_DEN"][:] = 1.0
    o.ghost_fill("E_DEN", 5)
    assert np.all(S["E_DEN"][0][-3:, 2:-2] == 0.0) and np.all(S["E_DEN"][0][-4, 2:-2] == 1.0)


def test_vertex_restriction_boundary_conditions():
    """Axis and free-slip sides: no normal velocity on the boundary vertex
    (its other component taken from the inward neighbour, which the ghost
    vertex copies too); fixed bottom: the boundary vertex is at rest."""
    h, o = fresh("c1")                              # axis left, free-slip right, fixed bottom
    S = o.state
    v = np.random.default_rng(5).standard_normal(S["V_VEL"].shape)
    S["V_VEL"][:] = v
    o.restrict("V_VEL")
    V = S["V_VEL"]
    rows = slice(3, -3)
    assert np.all(V[0][rows, 2] == 0.0) and np.array_equal(V[1][rows, 2], v[1][rows, 3])     # axis
    assert np.array_equal(V[:, rows, 1], v[:, rows, 3])
    assert np.all(V[0][rows, -3] == 0.0) and np.array_equal(V[1][rows, -3], v[1][rows, -4])  # right
    cols = slice(3, -3)
    assert np.all(V[:, 2, cols] == 0.0)                                                       # bottom
    assert np.array_equal(V[:, 10, 10], v[:, 10, 10])                                          # interior untouched


# ----------------------------------------------------------------- remap --
def half_space(axis, sign=1.0):
    """c1 mesh, regolith where the cell centre has coordinate < 0 along axis
    (z, or r - R for the radial case), vacuum elsewhere; geometry kept."""
    h, o = fresh("c1", **{"numerical.damping": "off"})
    S = o.state
    ep = S["E_POS"]
    x = ep[1] if axis == "z" else ep[0] - 80.0
    full = x < 0
    S["M_VOF"][:] = 0.0
    S["M_VOF"][0][~full] = 1.0
    S["M_VOF"][1][full] = 1.0
    S["M_DEN"][:] = 0.0
    S["M_DEN"][1][full] = 1800.0
    S["M_ENG"][:] = 0.0
    S["M_ENG"][1][full] = 1800.0 * 1.0e5
    S["E_DEN"][0] = np.where(full, 1800.0, 0.0)
    o.phase("V_VOF")
    return h, o, S, x


def test_remap_transports_a_flat_interface():
    """A flat interface displaced normally by d: the first empty row takes
    the material that crossed its face over its own volume, d / dz; in r the
    material that crossed a face at R came from [R - d, R], d (R - d/2) dz
    per radian (Pappus), over the cell's r_c dr dz."""
    dt, u = 1.0e-3, 100.0
    h, o, S, z = half_space("z")
    S["V_DIS"][0] = 0.0
    S["V_DIS"][1] = u * dt
    o.phase("FLUX", dt=dt)
    o.phase("ADVECT", dt=dt)
    j0 = int(np.argmax(z[:, 5] > 0))
    np.testing.assert_allclose(S["M_VOF"][1][j0, 2:-2], u * dt / 10.0, rtol=1e-12)
    np.testing.assert_allclose(S["M_VOF"][1][j0 - 1, 2:-2], 1.0, rtol=0, atol=1e-14)
    assert np.all(S["M_VOF"][1][j0 + 1, 2:-2] == 0.0)
    np.testing.assert_allclose(S["M_DEN"][1][j0, 2:-2], 1800.0, rtol=1e-14)
    np.testing.assert_allclose(S["M_ENG"][1][j0, 2:-2], 1800.0 * 1.0e5, rtol=1e-14)
    # radial: the interface at r = 80 m moves out by d
    h, o, S, x = half_space("r")
    d = u * dt
    S["V_DIS"][0] = d
    S["V_DIS"][1] = 0.0
    o.phase("FLUX", dt=dt)
    o.phase("ADVECT", dt=dt)
    k0 = int(np.argmax(x[5, :] > 0))
    R, dx = 80.0, 10.0
    expect = d * (R - d / 2) / ((R + dx / 2) * dx)
    np.testing.assert_allclose(S["M_VOF"][1][4:-4, k0], expect, rtol=1e-12)


def test_remap_of_a_uniform_state():
    """Uniform material displaced by (d_r, d_z): the axial part moves mass
    through every horizontal face alike and changes nothing; the radial part
    moves d (r - d/2) dz through a vertical face at r (Pappus), so a cell
    keeps rho (1 - d_r / r_c)."""
    dt = 1.0e-3
    h, o, S, z = half_space("z")
    S["M_VOF"][:] = 0.0
    S["M_VOF"][1] = 1.0
    S["M_DEN"][1] = 1800.0
    S["M_ENG"][1] = 1800.0 * 1.0e5
    S["E_DEN"][0] = 1800.0
    o.phase("V_VOF")
    S["V_DIS"][0] = 0.03
    S["V_DIS"][1] = -0.05
    o.phase("FLUX", dt=dt)
    o.phase("ADVECT", dt=dt)
    inner = (slice(4, -4), slice(4, -4))
    np.testing.assert_allclose(S["M_VOF"][1][inner], 1.0, rtol=0, atol=1e-14)
    rc = S["E_POS"][0][inner]
    np.testing.assert_allclose(S["M_DEN"][1][inner], 1800.0 * (1.0 - 0.03 / rc), rtol=1e-12)
    np.testing.assert_allclose(S["M_ENG"][1][inner], 1.8e8 * (1.0 - 0.03 / rc), rtol=1e-12)


def test_void_vertex_velocity():
    """check_v_vel: a vertex every one of whose cells marks that corner void
    takes the density-weighted mean of the cell velocities (vacuum cells
    excluded)."""
    h, o = fresh()
    S = o.state
    S["E_WPT"][:] = 1.0
    S["E_DEN"][0] = 1000.0
    rng = np.random.default_rng(6)
    S["E_VEL"][:] = rng.standard_normal(S["E_VEL"].shape)
    ev = S["E_VEL"].copy()
    o.phase("CHECK_V_VEL")
    mean = 0.25 * (ev[:, :-1, :-1] + ev[:, :-1, 1:] + ev[:, 1:, 1:] + ev[:, 1:, :-1])
    np.testing.assert_allclose(S["V_VEL"][:, 2:-2, 2:-2], mean[:, 1:-1, 1:-1], rtol=1e-13, atol=1e-15)
    # one cell of each vertex at twice the density, one empty
    S["E_DEN"][0] = np.where(np.arange(S["E_DEN"].shape[2]) % 2 == 0, 2000.0, 1000.0)[None, :]
    o.phase("CHECK_V_VEL")
    d = S["E_DEN"][0]
    w = [d[:-1, :-1], d[:-1, 1:], d[1:, 1:], d[1:, :-1]]
    e = [ev[:, :-1, :-1], ev[:, :-1, 1:], ev[:, 1:, 1:], ev[:, 1:, :-1]]
    wm = sum(wi * ei for wi, ei in zip(w, e)) / sum(w)
    np.testing.assert_allclose(S["V_VEL"][:, 2:-2, 2:-2], wm[:, 1:-1, 1:-1], rtol=1e-13, atol=1e-15)


def test_pml_auxiliary_relaxation():
    """pml2 with no velocity gradient: the cell auxiliaries relax at the rate
    alpha_x + alpha_y (e_alpha_0 exactly; e_alpha_1 with the stress term
    gone)."""
    h, o = fresh(**{"numerical.damping": "pml2", "numerical.LambdaX": "7.0", "numerical.LambdaY": "3.0"})
    S = o.state
    S["M_VOF"][0] = 0.0
    S["M_VOF"][1] = 1.0
    S["E_GRAD_VEL"][:] = 0.0
    S["E_STE"][:] = 0.0
    S["E_ALPHA"][0] = 2.0
    S["E_ALPHA"][1] = -5.0
    h.cfg.damp_time = 1.0e9
    dt = 1e-3
    o.phase("DAMP", dt=dt, t=0.0)
    rate = h.cfg.alpha_damp_x + h.cfg.alpha_damp_y
    assert rate > 0
    lay = ((S["E_SIGMA"][0] + S["E_SIGMA"][1]) > 0)[IN]
    a0 = S["E_ALPHA"][0][IN][lay]
    a1 = S["E_ALPHA"][1][IN][lay]
    np.testing.assert_allclose(a0, 2.0 * (1.0 - rate * dt), rtol=1e-14)
    np.testing.assert_allclose(a1, -5.0 * (1.0 - rate * dt), rtol=1e-14)


@pytest.mark.parametrize("f", [0.3, 0.5, 0.8])
def test_remap_of_a_mixed_row(f):
    """A horizontal interface inside a row of mixed cells (material below,
    volume fraction f) moved up by d < (1 - f) dz: the reconstructed
    interface keeps the top face empty, so the row only gains the material
    entering from below, f + d/dz, and the row above stays empty."""
    dt, d = 1.0e-3, 1.0
    h, o, S, z = half_space("z")
    j0 = int(np.argmax(z[:, 5] > 0))
    S["M_VOF"][1][j0] = f
    S["M_VOF"][0][j0] = 1.0 - f
    S["M_DEN"][1][j0] = 1800.0
    S["M_ENG"][1][j0] = 1800.0 * 1.0e5
    S["E_DEN"][0][j0] = 1800.0 * f
    o.phase("V_VOF")
    S["V_DIS"][0] = 0.0
    S["V_DIS"][1] = d
    o.phase("FLUX", dt=dt)
    o.phase("ADVECT", dt=dt)
    np.testing.assert_allclose(S["M_VOF"][1][j0, 3:-3], f + d / 10.0, rtol=1e-12)
    assert np.all(S["M_VOF"][1][j0 + 1, 3:-3] == 0.0)
    np.testing.assert_allclose(S["M_DEN"][1][j0, 3:-3], 1800.0, rtol=1e-12)


def test_pml_first_form_relaxes_each_direction_on_its_own():
    """pml: sigma_y relaxes the axial velocity and alpha_y the axial
    auxiliary, sigma_x and alpha_x the radial ones (no cross terms)."""
    h, o = fresh(**{"numerical.damping": "pml", "numerical.LambdaX": "7.0", "numerical.LambdaY": "3.0"})
    S = o.state
    S["M_VOF"][0] = 0.0
    S["M_VOF"][1] = 0.5
    S["M_VOF"][2 if h.nmat > 2 else 1] = 0.5 if h.nmat > 2 else 1.0
    S["E_GRAD_VEL"][:] = 0.0
    S["E_DVEL"][:] = 0.0
    S["E_VEL"][0], S["E_VEL"][1] = 2.0, -3.0
    S["E_ALPHA"][0], S["E_ALPHA"][1] = 1.5, -4.0
    h.cfg.damp_time = 1.0e9
    dt = 1e-3
    o.phase("DAMP", dt=dt, t=0.0)
    sx, sy = S["E_SIGMA"][0][IN], S["E_SIGMA"][1][IN]
    ax, ay = h.cfg.alpha_damp_x, h.cfg.alpha_damp_y
    dv, al = S["E_DVEL"][:, IN[0], IN[1]], S["E_ALPHA"][:, IN[0], IN[1]]
    assert (sx > 0).sum() > 10 and (sy > 0).sum() > 10
    np.testing.assert_allclose(dv[1][sy > 0], 3.0 * sy[sy > 0] * dt, rtol=1e-14)
    np.testing.assert_allclose(dv[0][sx > 0], -2.0 * sx[sx > 0] * dt, rtol=1e-14)
    assert np.all(dv[1][sy == 0] == 0.0) and np.all(dv[0][sx == 0] == 0.0)
    np.testing.assert_allclose(al[1][sy > 0], -4.0 * (1.0 - ay * dt), rtol=1e-14)
    np.testing.assert_allclose(al[0][sx > 0], 1.5 * (1.0 - ax * dt), rtol=1e-14)


def test_sponge_layer_relaxes_toward_the_reference_state():
    """sbc: in a radial sponge cell (factor s_x < 1) the density and energy
    are pulled to the reference state: rho_ref + (rho - rho_ref) s_x and
    E_ref + (E - E_ref) s_x^2."""
    h, o = fresh(**{"numerical.damping": "sbc", "numerical.LambdaX": "0.3", "numerical.LambdaY": "0.3"})
    S = o.state
    S["M_VOF"][:] = 0.0
    S["M_VOF"][1] = 1.0
    S["E_DAMPREF"][0] = 2000.0
    S["E_DAMPREF"][1] = 3.0e8
    S["M_DEN"][1] = 2100.0
    S["M_ENG"][1] = 3.5e8
    h.cfg.damp_time = 1.0e9
    o.phase("DAMP", dt=1e-3, t=0.0)
    sx, sy = S["E_SIGMA"][0][IN], S["E_SIGMA"][1][IN]
    only_x = (sx < 1.0) & (sy >= 1.0)
    assert only_x.sum() > 5
    np.testing.assert_allclose(S["M_DEN"][1][IN][only_x], 2000.0 + 100.0 * sx[only_x], rtol=1e-15)
    np.testing.assert_allclose(S["M_ENG"][1][IN][only_x], 3.0e8 + 0.5e8 * sx[only_x] ** 2, rtol=1e-15)
    untouched = (sx >= 1.0) & (sy >= 1.0)
    assert np.all(S["M_DEN"][1][IN][untouched] == 2100.0)


def test_vertex_pml_auxiliary_relaxation():
    """pml2 vertices in the layer: with no stress beyond the reference
    pressure, v_alpha relaxes at alpha_x (radial) and alpha_y (axial) and
    enters the acceleration as -v_alpha / m."""
    h, o = fresh(**{"numerical.damping": "pml2", "numerical.LambdaX": "7.0", "numerical.LambdaY": "3.0",
                    "numerical.ANC": "0.0"})
    S = o.state
    S["E_STE"][:] = 0.0
    S["E_Q"][:] = 0.0
    S["E_PRE"][0] = 5.0e7
    S["E_DAMPREF"][2] = 5.0e7
    S["E_DEN"][0] = 2000.0
    S["V_MASS"][:] = 0.0
    S["V_ALPHA"][0], S["V_ALPHA"][1] = 4.0e8, -1.0e8     # well above the acceleration floor
    dt = 1e-3
    o.phase("ACC", dt=dt)
    es = S["E_SIGMA"]
    vs = 0.25 * (es[:, :-1, :-1] + es[:, :-1, 1:] + es[:, 1:, 1:] + es[:, 1:, :-1])   # vertex sigma (inner vertices)
    lay = ((vs[0] > 0) | (vs[1] > 0))[2:-2, 2:-2]
    al = S["V_ALPHA"][:, 3:-3, 3:-3]
    assert lay.sum() > 20
    np.testing.assert_allclose(al[0][lay], 4.0e8 * (1.0 - h.cfg.alpha_damp_x * dt), rtol=1e-14)
    np.testing.assert_allclose(al[1][lay], -1.0e8 * (1.0 - h.cfg.alpha_damp_y * dt), rtol=1e-14)
    vm = 2000.0 * S["V_VOL"][1:5].sum(axis=0)[3:-3, 3:-3]
    acc, bf = S["V_ACC"][:, 3:-3, 3:-3], S["V_BF"][:, 3:-3, 3:-3]
    np.testing.assert_allclose(acc[0][lay], bf[0][lay] - al[0][lay] / vm[lay], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(acc[1][lay], bf[1][lay] - al[1][lay] / vm[lay], rtol=1e-9, atol=1e-12)


def split_power(T, G, sx, sy):
    """Sum over velocity components i (r, z) and directions j (r, z) of
    T_ij dv_i/dx_j, the terms of v_r weighted by sigma_y and those of v_z by
    sigma_x, plus T_tt v_r / r weighted by sigma_x + sigma_y.  T = (T_rr,
    T_rz, T_zz, T_tt); G = (dv_r/dr, dv_r/dz, dv_z/dr, dv_z/dz, v_r/r)."""
    Tm = ((T[0], T[1]), (T[1], T[2]))
    grad = ((G[0], G[1]), (G[2], G[3]))
    w = (sy, sx)
    p = sum(w[i] * Tm[i][j] * grad[i][j] for i in range(2) for j in range(2))
    return p + (sx + sy) * T[3] * G[4]


@pytest.mark.parametrize("form,pty", [("pml2", 1.0), ("pml2", 1.25), ("pml3", 1.25)])
def test_pml_auxiliary_stress_terms(form, pty):
    """pml2/pml3 cells full of material, auxiliaries starting at zero
    (damp_on_cell_pml2, src/velocity.c:556-580; pml3 :658-682): e_alpha_0
    gains the split divergence (the split power of the identity) and
    e_alpha_1 loses the split stress power of T = S - (p + q - p_ref) I, whose
    hoop entry carries the deviator alone, as src/velocity.c:564-567 writes
    it.  pml2 then moves the density by -rho_ref e_alpha_0 dt and the energy
    by -e_alpha_1 dt, both over the distension; pml3 relaxes the density
    toward rho_ref with no auxiliary term and no distension."""
    h, o = fresh(**{"numerical.damping": form, "numerical.LambdaX": "7.0", "numerical.LambdaY": "3.0"})
    S = o.state
    S["M_VOF"][:] = 0.0
    S["M_VOF"][1] = 1.0
    S["M_PTY"][1] = pty
    G = (0.7, -1.3, 2.1, -0.4, 0.25)
    Sd = (3.0e6, -2.0e6, 1.5e6, -4.5e6)
    p, q, pref = 2.0e7, 1.0e6, 1.8e7
    for i, g in enumerate(G):
        S["E_GRAD_VEL"][i] = g
    for i, s in enumerate(Sd):
        S["E_STE"][i] = s
    S["E_PRE"][0], S["E_Q"][0], S["E_DAMPREF"][2] = p, q, pref
    S["E_ALPHA"][:] = 0.0
    rho_ref, e_ref = 2000.0, 3.0e5
    S["E_DAMPREF"][0], S["E_DAMPREF"][1] = rho_ref, e_ref
    rho0 = rho_ref if form == "pml2" else 2100.0
    S["M_DEN"][1] = rho0
    S["M_ENG"][1] = e_ref
    h.cfg.damp_time = 1.0e9
    dt = 1e-3
    o.phase("DAMP", dt=dt, t=0.0)
    sx, sy = S["E_SIGMA"][0][IN], S["E_SIGMA"][1][IN]
    lay = (sx > 0) | (sy > 0)
    assert ((sx > 0) & (sy == 0)).sum() > 10 and ((sy > 0) & (sx == 0)).sum() > 10 and ((sx > 0) & (sy > 0)).sum() > 0
    sx, sy = sx[lay], sy[lay]
    lp = p + q - pref
    a0 = split_power((1.0, 0.0, 1.0, 1.0), G, sx, sy) * dt
    a1 = -split_power((Sd[0] - lp, Sd[1], Sd[2] - lp, Sd[3]), G, sx, sy) * dt
    np.testing.assert_allclose(S["E_ALPHA"][0][IN][lay], a0, rtol=1e-13, atol=1e-18)
    np.testing.assert_allclose(S["E_ALPHA"][1][IN][lay], a1, rtol=1e-12, atol=1e-6)
    den, eng = S["M_DEN"][1][IN][lay], S["M_ENG"][1][IN][lay]
    if form == "pml2":
        np.testing.assert_allclose(den, rho_ref - rho_ref * a0 * dt / pty, rtol=1e-14)
        np.testing.assert_allclose(eng, e_ref - a1 * dt / pty, rtol=1e-13)
    else:
        np.testing.assert_allclose(den, rho0 - (sx + sy) * (rho0 - rho_ref) * dt, rtol=1e-14)
        np.testing.assert_allclose(eng, e_ref - a1 * dt, rtol=1e-13)
    out = ~lay
    assert np.all(S["E_ALPHA"][:, IN[0], IN[1]][:, out] == 0.0) and np.all(S["M_DEN"][1][IN][out] == rho0)


@pytest.mark.parametrize("kind", ["normal", "shear", "hoop"])
def test_vertex_pml_stress_terms(kind):
    """pml2 vertices (calculate_acc_pml, src/velocity.c:158-340): the vertex
    auxiliary integrates the stress force split by face direction, the force
    through radial-facing faces weighted by the vertex sigma_y, through
    axial-facing faces by sigma_x, the hoop force by sigma_x + sigma_y.  So
    a normal stress's radial force is weighted by sigma_y and its axial one
    by sigma_x, a shear's the other way round.  The force itself is read off
    the same state stepped with sigma = 0, and the auxiliary enters the
    acceleration as -v_alpha / m."""
    h, o = fresh(**{"numerical.damping": "pml2", "numerical.LambdaX": "7.0", "numerical.LambdaY": "3.0",
                    "numerical.ANC": "0.0"})
    h.cfg.acceleration_min = 0.0
    S = o.state
    rng = np.random.default_rng({"normal": 1, "shear": 2, "hoop": 3}[kind])
    S["E_STE"][:] = 0.0
    S["E_Q"][:] = 0.0
    S["E_PRE"][0] = 0.0
    S["E_DAMPREF"][2] = 0.0
    for c in {"normal": (0, 2), "shear": (1,), "hoop": (3,)}[kind]:
        S["E_STE"][c] = rng.uniform(-1.0e7, 1.0e7, S["E_STE"][c].shape)
    S["E_DEN"][0] = 2000.0
    S["V_ALPHA"][:] = 0.0
    dt = 1e-3
    sig = S["E_SIGMA"].copy()
    S["E_SIGMA"][:] = 0.0
    o.phase("ACC", dt=dt)
    acc_u = S["V_ACC"].copy()
    assert np.all(S["V_ALPHA"] == 0.0)
    S["E_SIGMA"][:] = sig
    o.phase("ACC", dt=dt)
    vs = 0.25 * (sig[:, :-1, :-1] + sig[:, :-1, 1:] + sig[:, 1:, 1:] + sig[:, 1:, :-1])[:, 2:-2, 2:-2]
    lay = (vs[0] > 0) | (vs[1] > 0)
    assert lay.sum() > 20
    V = (slice(None), slice(3, -3), slice(3, -3))
    vm = 2000.0 * S["V_VOL"][1:5].sum(axis=0)[V[1:]]
    bf = S["V_BF"][V]
    f = (bf - acc_u[V]) * vm                               # stress force of each vertex
    w = {"normal": (vs[1], vs[0]), "shear": (vs[0], vs[1]), "hoop": (vs[0] + vs[1], vs[1])}[kind]
    decay = (1.0 - h.cfg.alpha_damp_x * dt, 1.0 - h.cfg.alpha_damp_y * dt)
    al, acc = S["V_ALPHA"][V], S["V_ACC"][V]
    for c in range(2):
        exp = w[c][lay] * f[c][lay] * dt * decay[c]
        scale = np.abs(f[c][lay]).max() * np.max(vs) * dt
        if kind == "hoop" and c == 1:
            assert np.abs(f[1][lay]).max() < 1e-6 * np.abs(f[0][lay]).max()
        else:
            assert scale > 0
        np.testing.assert_allclose(al[c][lay], exp, rtol=1e-8, atol=1e-9 * scale)
        np.testing.assert_allclose(acc[c][lay], bf[c][lay] - (f[c][lay] + al[c][lay]) / vm[lay], rtol=1e-9,
                                   atol=1e-9 * np.abs(acc[c][lay]).max())
    assert np.all(al[:, ~lay] == 0.0)


def test_pml3_flux_auxiliaries():
    """pml3 remap (advect_flux_pml3, src/advect.c:366-383 and :453-480): the
    flux auxiliaries e_beta take the mass, energy and momentum that cross the
    cell's faces, the flux through radial-facing faces weighted by -sigma_y
    and through axial-facing faces by -sigma_x, relax at alpha_x + alpha_y,
    and are added, times dt, to the first material's mass and energy and to
    the momentum.

    Uniform material displaced radially by d gains -rho V d / r_c in every
    cell (Pappus, as in the plain remap), through radial-facing faces only;
    a flat surface lifted by d fills the first empty row with rho V d / dz
    through axial-facing faces only."""
    dt, d = 1.0e-4, 0.05
    h, o = fresh(**{"numerical.damping": "pml3", "numerical.LambdaX": "7.0", "numerical.LambdaY": "3.0"})
    S = o.state
    S["M_VOF"][:] = 0.0
    S["M_VOF"][1] = 1.0
    rho, eng, u = 1800.0, 1800.0 * 1.0e5, (3.0, -5.0)
    S["M_DEN"][:] = 0.0
    S["M_DEN"][1] = rho
    S["M_ENG"][:] = 0.0
    S["M_ENG"][1] = eng
    S["E_DEN"][0] = rho
    S["E_VEL"][0], S["E_VEL"][1] = u
    S["E_BETA"][:] = 0.0
    o.phase("V_VOF")
    S["V_DIS"][0], S["V_DIS"][1] = d, 0.0
    o.phase("FLUX", dt=dt)
    o.phase("ADVECT", dt=dt)
    inner = (slice(2, -2), slice(3, -2))              # physical cells off the axis
    sx, sy = S["E_SIGMA"][0][inner], S["E_SIGMA"][1][inner]
    rc, V = S["E_POS"][0][inner], S["E_VOL"][0][inner]
    decay = 1.0 - (h.cfg.alpha_damp_x + h.cfg.alpha_damp_y) * dt
    gain = -V * d / rc                                    # net volume entering each cell
    lay = sy > 0
    assert lay.sum() > 10 and (~lay & (sx > 0)).sum() > 10
    beta = S["E_BETA"][:, inner[0], inner[1]]
    for q, per_vol in enumerate((rho, eng, rho * u[0], rho * u[1])):
        exp = -sy * per_vol * gain * decay
        np.testing.assert_allclose(beta[q][lay], exp[lay], rtol=1e-9, atol=0)
        np.testing.assert_allclose(beta[q][~lay], 0.0, rtol=0, atol=1e-12 * np.abs(exp).max())
    rho_new = rho * (1.0 - d / rc) + beta[0] * dt / V
    np.testing.assert_allclose(S["M_DEN"][1][inner], rho_new, rtol=1e-12)
    np.testing.assert_allclose(S["M_ENG"][1][inner], eng * (1.0 - d / rc) + beta[1] * dt / V, rtol=1e-12)
    for c in range(2):
        dv = (rho * u[c] * gain + beta[2 + c] * dt) / (rho_new * V)
        np.testing.assert_allclose(S["E_DVEL"][c][inner], dv, rtol=1e-8, atol=1e-14)
    assert np.any(beta[0][lay] * dt / V[lay] > 1e-3 * rho * d / rc[lay])  # the coupling shows

    # a flat surface lifted by d: the first empty row gains through its bottom face
    h, o = fresh(**{"numerical.damping": "pml3", "numerical.LambdaX": "7.0", "numerical.LambdaY": "3.0"})
    S = o.state
    z = S["E_POS"][1]
    full = z < 0
    S["M_VOF"][:] = 0.0
    S["M_VOF"][0][~full] = 1.0
    S["M_VOF"][1][full] = 1.0
    S["M_DEN"][:] = 0.0
    S["M_DEN"][1][full] = rho
    S["M_ENG"][:] = 0.0
    S["M_ENG"][1][full] = eng
    S["E_DEN"][0] = np.where(full, rho, 0.0)
    S["E_VEL"][:] = 0.0
    S["E_BETA"][:] = 0.0
    o.phase("V_VOF")
    S["V_DIS"][0], S["V_DIS"][1] = 0.0, d
    o.phase("FLUX", dt=dt)
    o.phase("ADVECT", dt=dt)
    j0 = int(np.argmax(z[:, 5] > 0))
    cols = slice(2, -2)
    sx, sy = S["E_SIGMA"][0][j0, cols], S["E_SIGMA"][1][j0, cols]
    V = S["E_VOL"][0][j0, cols]
    assert (sx > 0).sum() >= 2 and (sx == 0).sum() > 3
    b0 = S["E_BETA"][0][j0, cols]
    np.testing.assert_allclose(b0, -sx * rho * V * d / 10.0 * decay, rtol=1e-9, atol=0)
    np.testing.assert_allclose(S["M_DEN"][1][j0, cols] * S["M_VOF"][1][j0, cols] * V,
                               rho * V * d / 10.0 + b0 * dt, rtol=1e-11)
    below = S["E_BETA"][0][j0 - 1, cols]
    assert np.abs(below).max() <= 1e-9 * np.abs(b0).max()   # as much leaves through the top as enters


def test_axial_sponge_and_corner_order():
    """sbc (damp_on_cell_sbc, src/velocity.c:737-762): the axial sponge
    (factor s_y < 1) adds rho_ref + (rho - rho_ref) s_y to the density and
    E_ref + (E - E_ref) s_y to the energy, where the radial one assigns
    rho_ref + (rho - rho_ref) s_x and E_ref + (E - E_ref) s_x^2; a corner
    cell (every axial-sponge cell of this set-up) takes the axial update
    first and relaxes its result radially."""
    h, o = fresh(**{"numerical.damping": "sbc", "numerical.LambdaX": "0.3", "numerical.LambdaY": "0.3"})
    S = o.state
    S["M_VOF"][:] = 0.0
    S["M_VOF"][1] = 1.0
    rref, eref, rho, e = 2000.0, 3.0e8, 2100.0, 3.5e8
    S["E_DAMPREF"][0], S["E_DAMPREF"][1] = rref, eref
    S["M_DEN"][1], S["M_ENG"][1] = rho, e
    h.cfg.damp_time = 1.0e9
    o.phase("DAMP", dt=1e-3, t=0.0)
    sx, sy = S["E_SIGMA"][0][IN], S["E_SIGMA"][1][IN]
    den, eng = S["M_DEN"][1][IN], S["M_ENG"][1][IN]
    corner = (sy < 1.0) & (sx < 1.0)
    # the set-up gives every axial-sponge cell a radial factor as well
    assert corner.sum() > 5 and np.all(sx[sy < 1.0] < 1.0)
    r1 = rho + rref + (rho - rref) * sy[corner]
    e1 = e + eref + (e - eref) * sy[corner]
    np.testing.assert_allclose(den[corner], rref + (r1 - rref) * sx[corner], rtol=1e-15)
    np.testing.assert_allclose(eng[corner], eref + (e1 - eref) * sx[corner] ** 2, rtol=1e-15)
    # a cell with vacuum in it is left alone
    h, o = fresh(**{"numerical.damping": "sbc", "numerical.LambdaX": "0.3", "numerical.LambdaY": "0.3"})
    S = o.state
    S["M_VOF"][0], S["M_VOF"][1] = 0.5, 0.5
    S["M_DEN"][1] = rho
    h.cfg.damp_time = 1.0e9
    o.phase("DAMP", dt=1e-3, t=0.0)
    assert np.all(S["M_DEN"][1] == rho)


def _clip_area(rect, n, c):
    """Area of {x in the rectangle (r0, r1, z0, z1): n.x >= c} (half-plane
    clip of the polygon, then the shoelace formula)."""
    r0, r1, z0, z1 = rect
    poly = [(r0, z0), (r1, z0), (r1, z1), (r0, z1)]
    out = []
    for i in range(4):
        p, q = poly[i], poly[(i + 1) % 4]
        fp, fq = n[0] * p[0] + n[1] * p[1] - c, n[0] * q[0] + n[1] * q[1] - c
        if fp >= 0:
            out.append(p)
        if (fp >= 0) != (fq >= 0):
            t = fp / (fp - fq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    return 0.5 * abs(sum(out[i][0] * out[(i + 1) % len(out)][1] - out[(i + 1) % len(out)][0] * out[i][1]
                         for i in range(len(out)))) if len(out) > 2 else 0.0


def _edge_share(a, b, n, c):
    """Share of the segment a-b on which n.x >= c."""
    fa, fb = n[0] * a[0] + n[1] * a[1] - c, n[0] * b[0] + n[1] * b[1] - c
    if fa >= 0 and fb >= 0:
        return 1.0
    if fa < 0 and fb < 0:
        return 0.0
    return max(fa, fb) / abs(fa - fb)


@pytest.mark.parametrize("theta", [30.0, 125.0, 210.0, 290.0])
@pytest.mark.parametrize("f", [0.35, 0.8])
@pytest.mark.parametrize("dz", [None, 6.0])
def test_inclined_interface_flux(theta, f, dz):
    """A mixed cell (material fraction f, vacuum 1 - f) in empty space whose
    vertex vacuum fractions rise linearly along n = (cos theta, sin theta):
    calculate_flux_mixed (src/advect.c:700-905) cuts the cell with the line
    normal to n that leaves the vacuum a share 1 - f of the cell's (r, z)
    area on the side n points to, and each face passes vacuum in the share
    of its length on that side.  Displaced by d through one face at a time,
    the empty neighbour behind that face receives material
    V_swept (1 - vacuum share of the face) / V_neighbour.  The line and the
    shares are found here by clipping the rectangle and bisection, on square
    cells and on 10 m x 6 m ones."""
    d = 0.2
    n = (math.cos(math.radians(theta)), math.sin(math.radians(theta)))
    cut = 0
    for face in ("left", "bottom", "right", "top"):
        h, o = fresh(**({"mesh.dy": repr(dz)} if dz else {}))
        S = o.state
        j0, k0 = 18, 12
        S["M_VOF"][:] = 0.0
        S["M_VOF"][0] = 1.0
        S["M_VOF"][0][j0, k0], S["M_VOF"][1][j0, k0] = 1.0 - f, f
        S["M_DEN"][:] = 0.0
        S["M_ENG"][:] = 0.0
        S["M_DEN"][1][j0, k0], S["M_ENG"][1][j0, k0] = 1800.0, 1.8e8
        S["E_DEN"][0] = 0.0
        S["E_DEN"][0][j0, k0] = 1800.0 * f
        S["E_VEL"][:] = 0.0
        o.phase("V_VOF")
        vp = S["V_POS"]
        corners = [(j0, k0), (j0, k0 + 1), (j0 + 1, k0 + 1), (j0 + 1, k0)]
        pc = S["E_POS"][0][j0, k0], S["E_POS"][1][j0, k0]
        for (jj, kk) in corners:
            v0 = 0.5 + 0.01 * (n[0] * (vp[0][jj, kk] - pc[0]) + n[1] * (vp[1][jj, kk] - pc[1]))
            S["V_VOF"][0][jj, kk], S["V_VOF"][1][jj, kk] = v0, 1.0 - v0
        r0, r1 = vp[0][j0, k0], vp[0][j0, k0 + 1]
        z0, z1 = vp[1][j0, k0], vp[1][j0 + 1, k0]
        rect = (r0, r1, z0, z1)
        A = (r1 - r0) * (z1 - z0)
        assert abs((z1 - z0) - (dz or 10.0)) < 1e-9 and abs((r1 - r0) - 10.0) < 1e-9
        lo, hi = min(n[0] * x + n[1] * y for x in (r0, r1) for y in (z0, z1)), \
            max(n[0] * x + n[1] * y for x in (r0, r1) for y in (z0, z1))
        for _ in range(200):                                  # vacuum share 1 - f on the n side
            c = 0.5 * (lo + hi)
            if _clip_area(rect, n, c) > (1.0 - f) * A:
                lo = c
            else:
                hi = c
        c = 0.5 * (lo + hi)
        edge = {"left": ((r0, z0), (r0, z1)), "right": ((r1, z0), (r1, z1)),
                "bottom": ((r0, z0), (r1, z0)), "top": ((r0, z1), (r1, z1))}[face]
        vac = _edge_share(*edge, n, c)
        cut += 0.0 < vac < 1.0
        disp, nb, swept = {
            "right": ((d, 0.0), (j0, k0 + 1), d * (r1 - d / 2) * (z1 - z0)),
            "left": ((-d, 0.0), (j0, k0 - 1), d * (r0 + d / 2) * (z1 - z0)),
            "top": ((0.0, d), (j0 + 1, k0), d * (r1 * r1 - r0 * r0) / 2),
            "bottom": ((0.0, -d), (j0 - 1, k0), d * (r1 * r1 - r0 * r0) / 2)}[face]
        S["V_DIS"][0], S["V_DIS"][1] = disp
        o.phase("FLUX", dt=1e-3)
        o.phase("ADVECT", dt=1e-3)
        Vn = S["E_VOL"][0][nb]
        np.testing.assert_allclose(Vn, S["E_POS"][0][nb] * (r1 - r0) * (z1 - z0), rtol=1e-12)
        np.testing.assert_allclose(S["M_VOF"][1][nb], swept * (1.0 - vac) / Vn, rtol=1e-9, atol=1e-15,
                                   err_msg=face)
    assert cut >= 1


def _hyper_c(nd, c):
    """Hyper-C (Leonard 1991), the compressive bound of the normalised face
    value, with the bound 3 nD for Courant numbers of 1/3 and above."""
    return min(nd / c, 1.0) if c < 1.0 / 3.0 else min(3.0 * nd, 1.0)


def _stoic(nd):
    """STOIC (Darwish 1993) on 0 <= nD <= 1."""
    if nd < 0.2:
        return 3.0 * nd
    if nd < 0.5:
        return 0.5 * (1.0 + nd)
    return min((6.0 * nd + 3.0) / 8.0, 1.0)


@pytest.mark.parametrize("phi", [0.0, 40.0, 75.0, 90.0])
def test_mixed_cell_shares_among_materials(phi):
    """A cell of vacuum 0.4 (the primary material: its cut line normal to n
    leaves the vacuum 0.4 of the area), material 1 at 0.36 and material 2
    at 0.24, vacuum to its left, displaced radially by d into a right
    neighbour full of material 2 (algebraic_flux_ratio and vofMSTACS,
    src/advect.c:1226-1274, inside calculate_flux_mixed).  The non-vacuum
    flux through the right face, swept (1 - vacuum share), is shared in
    proportion to each material's face value:
    * material 1, empty on both sides: its own fraction 0.36;
    * material 2, full across the face (U = 1) and empty opposite (A = 0):
      nD = (D - U) / (A - U) = 0.76 and the face value U + f (A - U), with f
      the blend g Hyper-C + (1 - g) STOIC, g = cos^4 of the angle between
      material 2's interface normal (along phi) and the face normal, at the
      cell's Courant number swept / V.
    The neighbour's material 1 over material 2 then equals
    F_1 / (V + F_2 - outflow through its own right face)."""
    d, n = 0.2, (math.cos(math.radians(300.0)), math.sin(math.radians(300.0)))
    m2 = (math.cos(math.radians(phi)), math.sin(math.radians(phi)))
    h, o = fresh("c1")
    S = o.state
    assert h.nmat == 3
    j0, k0 = 18, 12
    f0, f1, f2 = 0.4, 0.36, 0.24
    S["M_VOF"][:] = 0.0
    S["M_VOF"][0] = 1.0
    S["M_VOF"][:, j0, k0] = (f0, f1, f2)
    S["M_VOF"][:, j0, k0 + 1] = (0.0, 0.0, 1.0)
    S["M_DEN"][:] = 0.0
    S["M_ENG"][:] = 0.0
    S["M_DEN"][1:, j0, k0], S["M_DEN"][2, j0, k0 + 1] = 1800.0, 2800.0
    S["M_ENG"][1:, j0, k0], S["M_ENG"][2, j0, k0 + 1] = 1.8e8, 2.8e8
    S["E_DEN"][0] = 0.0
    S["E_DEN"][0][j0, k0], S["E_DEN"][0][j0, k0 + 1] = 1800.0 * (f1 + f2), 2800.0
    S["E_VEL"][:] = 0.0
    o.phase("V_VOF")
    vp = S["V_POS"]
    pc = S["E_POS"][0][j0, k0], S["E_POS"][1][j0, k0]
    for (jj, kk) in [(j0, k0), (j0, k0 + 1), (j0 + 1, k0 + 1), (j0 + 1, k0)]:
        x, y = vp[0][jj, kk] - pc[0], vp[1][jj, kk] - pc[1]
        v0 = 0.4 + 0.01 * (n[0] * x + n[1] * y)
        v2 = 0.3 + 0.01 * (m2[0] * x + m2[1] * y)
        S["V_VOF"][0][jj, kk], S["V_VOF"][2][jj, kk], S["V_VOF"][1][jj, kk] = v0, v2, 1.0 - v0 - v2
    r0, r1, r2 = vp[0][j0, k0], vp[0][j0, k0 + 1], vp[0][j0, k0 + 2]
    z0, z1 = vp[1][j0, k0], vp[1][j0 + 1, k0]
    assert abs(r1 - r0 - 10.0) < 1e-9 and abs(z1 - z0 - 10.0) < 1e-9
    rect, A = (r0, r1, z0, z1), (r1 - r0) * (z1 - z0)
    lo, hi = min(n[0] * x + n[1] * y for x in (r0, r1) for y in (z0, z1)), \
        max(n[0] * x + n[1] * y for x in (r0, r1) for y in (z0, z1))
    for _ in range(200):
        c = 0.5 * (lo + hi)
        lo, hi = (c, hi) if _clip_area(rect, n, c) > f0 * A else (lo, c)
    vac = _edge_share((r1, z0), (r1, z1), n, 0.5 * (lo + hi))
    assert 0.0 < vac < 1.0
    bn = S["BX_NORMAL"][:, j0, k0 + 1]
    assert abs(abs(bn[0]) - 1.0) < 1e-12 and abs(bn[1]) < 1e-12
    S["V_DIS"][0], S["V_DIS"][1] = d, 0.0
    o.phase("FLUX", dt=1e-3)
    o.phase("ADVECT", dt=1e-3)
    V = S["E_VOL"][0][j0, k0]
    swept = d * (r1 - d / 2) * (z1 - z0)
    nd = (f2 - 1.0) / (0.0 - 1.0)
    g = min(m2[0] ** 4, 1.0)
    nf = g * _hyper_c(nd, swept / V) + (1.0 - g) * _stoic(nd)
    b1, b2 = f1, 1.0 - nf
    F1 = swept * (1.0 - vac) * b1 / (b1 + b2)
    F2 = swept * (1.0 - vac) * b2 / (b1 + b2)
    out = d * (r2 - d / 2) * (z1 - z0)
    V2 = S["E_VOL"][0][j0, k0 + 1]
    ratio = S["M_VOF"][1][j0, k0 + 1] / S["M_VOF"][2][j0, k0 + 1]
    np.testing.assert_allclose(ratio, F1 / (V2 + F2 - out), rtol=1e-10)
