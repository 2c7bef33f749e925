"""Helpers shared by the parity tests and the smoke check: build a run from a
benchmark configuration, step the oracle and the HIP path from the same state
and compare them field by field."""
import os
import tempfile

import numpy as np

import pysalec as ps
import pyoracle as po

TOLVOF = 1.0e-6
TOLRHO = 1.0e-4
# per-step relative L-infinity bound for FP64 fields (the parity bar)
FIELD_RTOL = 1.0e-11

COMPARED = ps.PERSISTENT + ["E_VEL", "E_GRAD_VEL", "E_DIV_VEL", "E_Q", "E_STA2S", "E_STRENGTH", "V_VOF",
                            "V_ACC", "V_DIS", "Y_PROFILE"]


def make_run(which, nx, ny, workdir=None, **kw):
    """Host setup + oracle initialisation (src/salec2d.c:76-94)."""
    d = workdir or tempfile.mkdtemp(prefix="salec_%s_" % which)
    inp, eos = ps.write_config(os.path.join(d, which), which, nx=nx, ny=ny, **kw)
    h = ps.HostRun(inp, eos)
    o = po.Oracle(h.cfg, h.state(), t=h.time0, dt=h.dt0)
    damp = o.init()
    h.cfg.damp_time = damp
    return h, o


def rel_linf(a, b):
    """max |a-b| / max |b| (absolute if b is all zero); nan-aware."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    fa, fb = np.isfinite(a), np.isfinite(b)
    if not np.array_equal(fa, fb):
        return np.inf
    if not fb.any():
        return 0.0
    d = np.abs(a[fb] - b[fb]).max()
    s = np.abs(b[fb]).max()
    return d / s if s > 0 else d


def int_state(S, nmat):
    """Integer-valued state that must match exactly: material-present masks,
    vacuum flags, void-corner counts, damage-saturated and vibration flags."""
    v = S["M_VOF"]
    return {
        "present": (v > TOLVOF),
        "full_vacuum": (v[0] > 1.0 - TOLVOF),
        "void_corner": S["V_MASS"][1],
        "dam_one": (S["E_DAM"][0] == 1.0),
        "vib_pos": (S["E_VIB"][0] > 0.0),
        "den_pos": (S["E_DEN"][0] >= TOLRHO),
    }


def compare(gpu, ora, nmat, fields=None, rtol=FIELD_RTOL):
    """Return (worst_field, worst_rel, int_mismatches dict)."""
    worst, wf = 0.0, None
    for n in fields or COMPARED:
        r = rel_linf(gpu[n], ora[n])
        if r > worst:
            worst, wf = r, n
    ig, io = int_state(gpu, nmat), int_state(ora, nmat)
    mism = {k: int(np.count_nonzero(ig[k] != io[k])) for k in ig}
    return wf, worst, mism


def tracer_compare(tg, to):
    n = to.n
    if tg.n != n:
        return {"count": (tg.n, n)}
    out = {}
    for k in ("idx", "idy", "alive", "matid"):
        out[k] = int(np.count_nonzero(getattr(tg, k)[:n] != getattr(to, k)[:n]))
    alive = to.alive[:n] == 1
    out["pos"] = rel_linf(tg.pos[:, :n][:, alive], to.pos[:, :n][:, alive]) if alive.any() else 0.0
    out["vel"] = rel_linf(tg.vel[:, :n][:, alive], to.vel[:, :n][:, alive]) if alive.any() else 0.0
    return out
