"""Tracer lubrication (tracer.lubrication, src/strength.c:265-275 and
:351-376, src/construct.c:76-92) in the host set-up and the oracle.

Pins:
* on a uniform 10 m mesh whose surface lies on a vertex row, the
  lubricating vertex rows [lub_gy0, lub_gy1) are exactly the L rows below
  the surface row;
* after a step, e_cnd of every cell equals the number of live tracers
  created in those rows that are now in the cell, counted here
  independently from the tracer arrays;
* in the next step every cell with such a tracer has its shear strength
  capped at min(p lub_frac + lub_y0, 0): with lub_frac = lub_y0 = 0 the cap
  is 0, so those cells end the step with zero strength, while the same
  cells in a run without lubrication hold positive strength;
* with lub_y0 < 0 and lub_frac = 0 the cap is the constant lub_y0;
* lubrication off leaves e_cnd at zero and the run unchanged.
"""
import os

import numpy as np

import parity as P
import pyoracle as po
import pysalec as ps

NX, NY, L = 32, 48, 4


def run(tmp_path, name, lub, steps_ahead=20, y0=0.0):
    over = {"tracer.lubrication": str(lub), "tracer.lub_frac": "0.0", "tracer.lub_y0": repr(y0)}
    h = ps.HostRun(*ps.write_config(str(tmp_path / name), "c1", nx=NX, ny=NY, sections_over=over))
    o = po.Oracle(h.cfg, h.state(), t=h.time0, dt=h.dt0)
    h.cfg.damp_time = o.init()
    tr = h.tracers(o.state["V_VEL"])
    o = po.Oracle(h.cfg, o.state, tracers=tr, t=h.time0, dt=h.dt0)
    for _ in range(steps_ahead):
        o.step()
    return h, o


def test_lubricating_rows_on_uniform_mesh(tmp_path):
    h = ps.HostRun(*ps.write_config(str(tmp_path / "r"), "c1", nx=NX, ny=NY,
                                    sections_over={"tracer.lubrication": str(L)}))
    vy = h.field("V_POS")[1][:, 0]
    js = int(np.flatnonzero(vy == 0.0)[0])            # the surface vertex row
    assert (h.cfg.lub_gy0, h.cfg.lub_gy1) == (js - L, js)
    h2 = ps.HostRun(*ps.write_config(str(tmp_path / "r0"), "c1", nx=NX, ny=NY))
    assert h2.cfg.tracer_lubrication == 0


def test_condition_counts_and_strength_cap(tmp_path):
    h, o = run(tmp_path, "lub", L)
    gy0, gy1 = h.cfg.lub_gy0, h.cfg.lub_gy1
    o.step()
    tr = o.tr
    cnt = np.zeros_like(o.state["E_CND"][0])
    for i in range(tr.n):
        if tr.alive[i] and gy0 <= tr.gy[i] < gy1:
            cnt[tr.idy[i], tr.idx[i]] += 1.0
    assert np.array_equal(o.state["E_CND"][0], cnt)
    lub = cnt >= 1.0
    assert lub.sum() > 10
    # the same state stepped without lubrication, for comparison
    h0, o0 = run(tmp_path, "nolub", 0)
    o0.step()
    o.step()
    o0.step()
    s_lub, s_ref = o.state["E_STRENGTH"][0], o0.state["E_STRENGTH"][0]
    assert np.all(s_lub[lub] == 0.0)
    assert (s_ref[lub] > 0.0).sum() > 10               # the cap changed these cells
    assert np.all(o0.state["E_CND"] == 0.0)


def test_cap_below_zero(tmp_path):
    """lub_y0 < 0 with lub_frac = 0: the cap is lub_y0, which every
    strength-bearing lubricated cell takes (cells mixed with vacuum, melted
    or vapourised keep zero strength).  A negative strength flips the
    deviatoric stress, and the run soon fails, as the reference's would:
    this checks the first step the cap acts on."""
    h, o = run(tmp_path, "neg", L, steps_ahead=0, y0=-5.0e6)
    o.step()
    lub = o.state["E_CND"][0] >= 1.0
    o.step()
    s = o.state["E_STRENGTH"][0][lub]
    assert np.all((s == -5.0e6) | (s == 0.0))
    assert (s == -5.0e6).sum() > 10


def test_lubrication_off_is_the_plain_run(tmp_path):
    """tracer.lubrication = 0 (the default): the run equals the plain one."""
    h, o = run(tmp_path, "lubL", 0, steps_ahead=3)
    h1 = ps.HostRun(*ps.write_config(str(tmp_path / "same"), "c1", nx=NX, ny=NY))
    o1 = po.Oracle(h1.cfg, h1.state(), t=h1.time0, dt=h1.dt0)
    h1.cfg.damp_time = o1.init()
    o1 = po.Oracle(h1.cfg, o1.state, tracers=h1.tracers(o1.state["V_VEL"]), t=h1.time0, dt=h1.dt0)
    for _ in range(3):
        o1.step()
    for k in P.COMPARED:
        assert np.array_equal(o.state[k], o1.state[k], equal_nan=True), k
