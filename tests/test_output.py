"""Output writers (include/salec_output.h) against the reference's own VTK
encoders (src/io/output.c, compiled into oracle/_ref with ref_out_shim.c,
which calls them in the order of src/write_vtk.c).

* state files (.vts) and tracer files (.vtp) are compared byte for byte with
  the files the reference's encoders write from the same data laid out as the
  reference stores it (per node, components together);
* independently of the reference build, every binary array is decoded
  (base64 header + body, zlib) and compared with the Float32 values of the
  input planes, and the XML parses;
* the parallel headers (.pvts, .pvtp) and multiblock indexes (.vtm) are
  checked for the extents, array names, component counts and piece sources
  the reference's writers put there.
"""
import base64
import ctypes as C
import os
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np
import pytest

import pysalec as ps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_OUT = os.path.join(ROOT, "oracle", "_ref", "libref_out.so")

POINT = [("v_vel", "V_VEL"), ("v_acc", "V_ACC"), ("v_gra", "V_BF"), ("v_vof", "V_VOF")]
CELL = [("e_den", "E_DEN"), ("e_pre", "E_PRE"), ("e_wpt", "E_WPT"), ("e_vof", "M_VOF"), ("e_csd", "E_CSD"),
        ("e_dam", "E_DAM"), ("e_tps", "E_TPS"), ("e_q", "E_Q"), ("e_tem", "E_TEM"), ("e_eng", "M_ENG"),
        ("m_den", "M_DEN"), ("e_ste2s", "E_STE2S"), ("e_sta2s", "E_STA2S"), ("e_ste", "E_STE"),
        ("e_strength", "E_STRENGTH"), ("e_vib", "E_VIB"), ("e_ejt", "E_EJT"), ("e_cvs", "E_CVS"), ("m_pty", "M_PTY")]
NX, NY, NMAT = 23, 31, 3


def random_state(seed=5):
    rng = np.random.default_rng(seed)
    st = {}
    for name, _, _ in ps.FIELDS:
        shp = ps.field_shape(name, NMAT, NX, NY)
        st[name] = rng.standard_normal(shp) * 10.0 ** rng.uniform(-3, 9)
    # a realistic mesh and some exactly repeated values, which zlib packs
    VW, VH = NX + 5, NY + 5
    x = (np.arange(VW) - 2.0) * 10.0
    y = (np.arange(VH) - 2.0 - 0.7 * NY) * 10.0
    st["V_POS"] = np.stack([np.tile(x, (VH, 1)), np.tile(y[:, None], (1, VW))])
    st["M_VOF"][:, : NY // 2] = 0.0
    st["M_VOF"][1, : NY // 2] = 1.0
    return st


def aos(a):
    """planes [ncomp, H, W] -> per-node records, the reference's storage"""
    a = np.asarray(a, np.float64)
    if a.ndim == 2:
        a = a[None]
    return np.ascontiguousarray(a.reshape(a.shape[0], -1).T)


def ref_lib():
    if not os.path.exists(REF_OUT):
        pytest.skip("oracle/_ref/libref_out.so not built (reference sources absent)")
    return C.CDLL(REF_OUT)


def ref_vts(path, st, coord_format):
    R = ref_lib()
    VW, VH, CW, CH = NX + 5, NY + 5, NX + 4, NY + 4
    keep = []

    def arrays(spec):
        names = (C.c_char_p * len(spec))(*[n.encode() for n, _ in spec])
        ptrs = (C.POINTER(C.c_double) * len(spec))()
        ncs = (C.c_int * len(spec))()
        for i, (_, f) in enumerate(spec):
            a = aos(st[f])
            keep.append(a)
            ptrs[i] = a.ctypes.data_as(C.POINTER(C.c_double))
            ncs[i] = a.shape[1]
        return names, ptrs, ncs

    pn, pa, pc = arrays(POINT)
    cn, ca, cc = arrays(CELL)
    coord = aos(st["V_POS"])
    R.ref_out_vts.restype = C.c_int
    rc = R.ref_out_vts(path.encode(), b"0 %d 0 %d 0 0" % (VW - 1, VH - 1), b"0 %d 0 %d 0 0" % (NX + 2, NY + 2),
                       len(POINT), pn, pa, pc, C.c_long(VW * VH), len(CELL), cn, ca, cc, C.c_long(CW * CH),
                       coord.ctypes.data_as(C.POINTER(C.c_double)), coord_format.encode())
    assert rc == 0


def decode_binary(text):
    """(header ints, Float32 values) of one vtkZLibDataCompressor array"""
    text = text.strip()
    head = struct.unpack("<4i", base64.b64decode(text[:24]))
    raw = zlib.decompress(base64.b64decode(text[24:]))
    assert head[0] == 1 and head[1] == head[2] == len(raw)
    return head, np.frombuffer(raw, "<f4")


@pytest.mark.parametrize("coord_format", ["binary", "ascii"])
def test_vts_bytes_equal_reference_encoders(tmp_path, coord_format):
    st = random_state()
    ours, ref = str(tmp_path / "ours.vts"), str(tmp_path / "ref.vts")
    ps.write_vts(ours, st, NX, NY, NMAT, coord_format)
    ref_vts(ref, st, coord_format)
    a, b = open(ours, "rb").read(), open(ref, "rb").read()
    assert len(a) > 10000
    assert a == b


def test_vts_content_decodes_to_the_planes(tmp_path):
    st = random_state(seed=9)
    path = str(tmp_path / "s.vts")
    ps.write_vts(path, st, NX, NY, NMAT)
    root = ET.parse(path).getroot()
    piece = root.find("StructuredGrid/Piece")
    assert piece.get("Extent") == "0 %d 0 %d 0 0" % (NX + 4, NY + 4)
    assert root.find("StructuredGrid").get("WholeExtent") == "0 %d 0 %d 0 0" % (NX + 2, NY + 2)
    got = {}
    for sec in ("PointData", "CellData", "Points"):
        for da in piece.find(sec):
            _, v = decode_binary(da.text)
            got[da.get("Name")] = (int(da.get("NumberOfComponents")), v)
    assert [n for n, _ in POINT] == [da.get("Name") for da in piece.find("PointData")]
    assert [n for n, _ in CELL] == [da.get("Name") for da in piece.find("CellData")]
    for name, f in POINT + CELL:
        nc, v = got[name]
        want = aos(st[f]).astype(np.float32).reshape(-1)
        assert nc == aos(st[f]).shape[1]
        assert np.array_equal(v, want), name
    nc, v = got["coordinate"]
    xy = aos(st["V_POS"]).astype(np.float32)
    assert nc == 3 and np.array_equal(v.reshape(-1, 3)[:, :2], xy) and not v.reshape(-1, 3)[:, 2].any()


def random_tracers(n, seed=3, dead_every=7):
    rng = np.random.default_rng(seed)
    tr = ps.TracerSet(n)
    tr.pos[...] = rng.uniform(-100, 200, tr.pos.shape)
    tr.vel[...] = rng.standard_normal(tr.vel.shape) * 1e3
    tr.mpre[...] = rng.uniform(0, 1e9, tr.mpre.shape).astype(np.float32)
    tr.mtem[...] = rng.uniform(200, 2000, tr.mtem.shape).astype(np.float32)
    tr.gx[...] = rng.integers(0, NX, n)
    tr.gy[...] = rng.integers(0, NY, n)
    tr.tag[...] = np.arange(n)
    tr.matid[...] = rng.choice([-2, -1, 1, 2], n)
    tr.idx[...] = rng.integers(2, NX + 2, n)
    tr.idy[...] = rng.integers(2, NY + 2, n)
    tr.alive[...] = (np.arange(n) % dead_every) != 0
    return tr


def ref_vtp(path, tr, den, tem, pre):
    """the reference writer's per-tracer arrays (src/write_vtk.c:396-459) for
    the live tracers in order, or its dummy point"""
    R = ref_lib()
    CW = NX + 4
    live = np.nonzero(tr.alive[: tr.n])[0]
    m = max(len(live), 1)
    f = lambda: np.zeros(m, np.float32)
    mat, tag, gx, gy, dd, tt, pp, mp, mt = (f() for _ in range(9))
    vel, pos = np.zeros(2 * m, np.float32), np.zeros(3 * m, np.float32)
    if len(live) == 0:
        mat[0] = tag[0] = gx[0] = gy[0] = -1.0
    else:
        c = tr.idx[live] + CW * tr.idy[live]
        mat[:], tag[:] = tr.matid[live], tr.tag[live]
        gx[:], gy[:] = tr.gx[live], tr.gy[live]
        vel[0::2], vel[1::2] = tr.vel[0, live], tr.vel[1, live]
        pos[0::3], pos[1::3] = tr.pos[0, live], tr.pos[1, live]
        mp[:], mt[:] = tr.mpre[live], tr.mtem[live]
        dd[:], tt[:], pp[:] = den.reshape(-1)[c], tem.reshape(-1)[c], pre.reshape(-1)[c]
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    assert R.ref_out_vtp(path.encode(), m, fp(mat), fp(tag), fp(gx), fp(gy), fp(vel), fp(dd), fp(tt), fp(pp),
                         fp(mp), fp(mt), fp(pos)) == 0


@pytest.mark.parametrize("n,dead_every", [(5000, 7), (40, 1)], ids=["tracers", "all_deleted"])
def test_vtp_bytes_equal_reference_encoders(tmp_path, n, dead_every):
    tr = random_tracers(n, dead_every=dead_every)
    rng = np.random.default_rng(1)
    den, tem, pre = (rng.uniform(0, 3000, (NY + 4, NX + 4)) for _ in range(3))
    ours, ref = str(tmp_path / "ours.vtp"), str(tmp_path / "ref.vtp")
    txt = str(tmp_path / "ours.txt")
    ps.write_vtp(ours, tr, NX, NY, den, tem, pre, txt_path=txt)
    ref_vtp(ref, tr, den, tem, pre)
    assert open(ours, "rb").read() == open(ref, "rb").read()
    live = int(tr.alive[:n].sum())
    lines = open(txt).read().splitlines()
    assert int(lines[0]) == live and len(lines) == 1 + max(live, 1)
    root = ET.parse(ours).getroot()
    assert int(root.find("PolyData/Piece").get("NumberOfPoints")) == max(live, 1)


def test_parallel_headers_and_indexes(tmp_path):
    d = str(tmp_path)
    ps.write_pvts(os.path.join(d, "P.0003.pvts"), "P", 3, NX, NY, NMAT)
    root = ET.parse(os.path.join(d, "P.0003.pvts")).getroot()
    g = root.find("PStructuredGrid")
    assert g.get("WholeExtent") == "2 %d 0 %d 0 0" % (NX + 4, NY + 2) and g.get("GhostLevel") == "0"
    pd = [(a.get("Name"), int(a.get("NumberOfComponents"))) for a in g.find("PPointData")]
    assert pd == [("v_vel", 2), ("v_acc", 2), ("v_gra", 2), ("v_vof", NMAT)]
    cd = [a.get("Name") for a in g.find("PCellData")]
    assert cd == [n for n, _ in CELL][:17]
    pc = g.find("Piece")
    assert pc.get("Source") == "./vts/P.proc0000.0003.vts" and pc.get("Extent") == "0 %d 0 %d 0 0" % (NX + 2, NY + 2)
    ps.write_vtm(d, "P", 3)
    for name, src in (("P.0003.vtm", "./vts/P.proc0000.0003.vts"), ("P.tracer.0003.vtm", "./vtp/P.tracer.proc0000.0003.vtp")):
        r = ET.parse(os.path.join(d, name)).getroot()
        ds = r.find("vtkMultiBlockDataSet/DataSet")
        assert ds.get("index") == "0" and ds.get("file") == src
    ps.write_pvtp(os.path.join(d, "P.tracer.0003.pvtp"), "P", 3)
    r = ET.parse(os.path.join(d, "P.tracer.0003.pvtp")).getroot().find("PPolyData")
    assert [a.get("Name") for a in r.find("PPointData")] == ["matid", "tag", "gx", "gy", "vel", "den", "tem", "pre",
                                                              "mpre", "mtem"]
    assert r.find("Piece").get("Source") == "./vtp/P.tracer.proc0000.0003.vtp"


def test_mass_tracer_list(tmp_path):
    tr = random_tracers(300)
    rng = np.random.default_rng(2)
    den, vol = rng.uniform(1, 3000, (NY + 4, NX + 4)), rng.uniform(1, 1e3, (NY + 4, NX + 4))
    path = str(tmp_path / "mass.tracer.proc0000.txt")
    ps.write_mass_tracers(path, tr, NX, NY, den, vol)
    lines = open(path).read().splitlines()
    live = np.nonzero(tr.alive[:300])[0]
    assert int(lines[0]) == len(live) and len(lines) == 1 + len(live)
    i = live[5]
    c = tr.idx[i] + (NX + 4) * tr.idy[i]
    want = "%d, %d, %d, %10.5e,%10.5e, %10.5e, %10.5e" % (tr.tag[i], tr.gx[i], tr.gy[i], tr.pos[0, i], tr.pos[1, i],
                                                          den.reshape(-1)[c], vol.reshape(-1)[c])
    assert lines[6] == want


def test_writers_reject_bad_arguments(tmp_path):
    st = random_state()
    with pytest.raises(RuntimeError):
        ps.write_vts(str(tmp_path / "x.vts"), st, NX, NY, NMAT, "hex")
    with pytest.raises(RuntimeError):
        ps.write_vts(str(tmp_path / "no_such_dir" / "x.vts"), st, NX, NY, NMAT)
    tr = random_tracers(10)
    tr.idx[3] = 10 ** 6
    den = np.zeros((NY + 4, NX + 4))
    with pytest.raises(RuntimeError):
        ps.write_vtp(str(tmp_path / "x.vtp"), tr, NX, NY, den, den, den)
