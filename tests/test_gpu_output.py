"""End-to-end output of the C driver (salec2d) on the GPU.

The salec2d program runs a small strength deck with output every 5 steps (and at the
first step, as the reference does) and writes the state and tracer files at
the mid-step output point.  The same run is replayed here step by step through
the C ABI, stopping at the same output points (SALEC_STEP_TO_OUTPUT) and
writing the same files with the library's writers; the files must be equal
byte for byte, which pins the output schedule (update_cycle_control,
src/sale2d.c:96-117), the output point (src/salec2d.c:241-245) and the data
each file takes.  The writers themselves are pinned against the reference's
encoders in tests/test_output.py.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import pysalec as ps

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "salec-2d-public_amd", "salec2d")
OUT_FIELDS = ["V_VEL", "V_ACC", "V_BF", "V_VOF", "V_POS", "E_DEN", "E_PRE", "E_WPT", "M_VOF", "E_CSD", "E_DAM",
              "E_TPS", "E_Q", "E_TEM", "M_ENG", "M_DEN", "E_STE2S", "E_STA2S", "E_STE", "E_STRENGTH", "E_VIB",
              "E_EJT", "E_CVS", "M_PTY"]


def say(msg):
    print("[gpu-output] " + msg, flush=True)
    sys.stdout.flush()


def test_driver_files_equal_replayed_run():
    if not os.path.exists(DRIVER):
        pytest.skip("salec2d is not built")
    d = tempfile.mkdtemp(prefix="out_")
    nx, ny, steps = 32, 48, 12
    inp, eos = ps.write_config(os.path.join(d, "run"), "c1", nx=nx, ny=ny,
                               sections_over={"output.prefix": "T", "output.time": "1000.0", "output.out_cycle": "5"})
    out = os.path.join(d, "drv")
    say("driver %dx%d, %d steps, output every 5 steps" % (nx, ny, steps))
    r = subprocess.run([DRIVER, inp, eos, "--steps", str(steps), "--out", out, "--log", "100"], capture_output=True,
                       text=True, timeout=120)
    say(r.stdout.strip().splitlines()[-1] if r.stdout.strip() else r.stderr)
    assert r.returncode == 0, r.stderr
    vts = sorted(os.listdir(os.path.join(out, "vts")))
    assert vts == ["T.proc0000.%04d.vts" % c for c in range(3)]        # steps 1, 6 and 11
    for name in ("T.0002.pvts", "T.0002.vtm", "T.tracer.0002.vtm", "T.tracer.0002.pvtp"):
        assert os.path.exists(os.path.join(out, name)), name

    say("replay through the C ABI")
    h = ps.HostRun(inp, eos)
    ntr = ps.lib().salec_host_count_tracers(h.h)
    dev = ps.Device(h.cfg, max_tracers=ntr)
    dev.upload_state(h.state())
    h.cfg.damp_time = dev.init(h.damp_t_in)
    tr = h.tracers(dev.download("V_VEL"))
    dev.upload_tracers(tr)
    dev.set_time(h.time0, h.dt0)
    mine = os.path.join(d, "mine")
    os.makedirs(mine)
    ps.write_mass_tracers(os.path.join(mine, "mass.txt"), tr, nx, ny, dev.download("E_DEN"), h.field("E_VOL"))
    cyc = 0
    for k in range(1, steps + 1):
        if k in (1, 6, 11):
            dev.step_to_output()
            st = {f: dev.download(f) for f in OUT_FIELDS}
            ps.write_vts(os.path.join(mine, "s%d.vts" % cyc), st, nx, ny, h.cfg.nmat)
            t2 = dev.download_tracers(tr.n)
            ps.write_vtp(os.path.join(mine, "t%d.vtp" % cyc), t2, nx, ny, st["E_DEN"], st["E_TEM"], st["E_PRE"])
            dev.step_from_output()
            cyc += 1
        else:
            dev.step()
    dev.close()
    for c in range(3):
        a = open(os.path.join(out, "vts", "T.proc0000.%04d.vts" % c), "rb").read()
        b = open(os.path.join(mine, "s%d.vts" % c), "rb").read()
        assert a == b, c
        a = open(os.path.join(out, "vtp", "T.tracer.proc0000.%04d.vtp" % c), "rb").read()
        b = open(os.path.join(mine, "t%d.vtp" % c), "rb").read()
        assert a == b, c
    a = open(os.path.join(out, "txt", "mass.tracer.proc0000.txt")).read()
    assert a == open(os.path.join(mine, "mass.txt")).read()
    assert int(a.splitlines()[0]) == tr.n > 0
