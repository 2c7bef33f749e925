"""Tooling: the iSALE-2D input converter and the .vts packer.

* The converter (salec_isale_to_inp, salec_generate_input) must print what
  the reference's own converter prints, byte for byte: against the committed
  fixtures made by the reference build (tests/golden/isale/, written by
  make_isale_golden.py), and, where oracle/_ref holds that build, against it
  live on seeded random input pairs.
* The packer (salec_pack_vts / salec_unpack_vts / salec_merge_vts_blocks,
  salec_merge_vts) has no buildable reference (src/other/merge_vts.c needs
  the MPI headers): it is pinned by its format (the reference's
  "#BEGINFILE:<name>" ... "#ENDFILE" sections, the cycle split of its
  threads), by the round trip pack -> unpack restoring every file byte for
  byte, and by the multiblock files parsing as XML with one DataSet per
  domain that holds that domain's grid.
"""
import os
import shutil
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import pysalec as ps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ISALE = os.path.join(ROOT, "tests", "golden", "isale")
PKG = os.path.join(ROOT, "salec-2d-public_amd")
GEN = os.path.join(PKG, "salec_generate_input")
MERGE = os.path.join(PKG, "salec_merge_vts")
REF_GEN = os.path.join(ROOT, "oracle", "_ref", "ref_generate_input")
CASES = sorted(os.listdir(ISALE)) if os.path.isdir(ISALE) else []


@pytest.mark.parametrize("case", CASES)
def test_isale_conversion_golden(case, tmp_path):
    d = os.path.join(ISALE, case)
    exp = open(os.path.join(d, "expected.txt"), "rb").read()
    got = ps.isale_to_inp(os.path.join(d, "asteroid.inp"), os.path.join(d, "material.inp"))
    assert got.encode() == exp
    # the program, run as the reference's is: ./asteroid.inp and ./material.inp
    if os.path.exists(GEN):
        shutil.copy(os.path.join(d, "asteroid.inp"), tmp_path)
        shutil.copy(os.path.join(d, "material.inp"), tmp_path)
        r = subprocess.run([GEN], cwd=tmp_path, capture_output=True, check=True)
        assert r.stdout == exp


def _random_pair(rng):
    """A valid iSALE input pair within the reference's fixed field sizes."""
    def ln(name, *vals):
        return "%-14s%-30s: %s\n" % (name, name.lower() + " desc", "  : ".join(vals))

    def num(v):
        s = "%.4E" % v
        return s.replace("E", rng.choice(["D", "E", "d", "e"]))

    ex = [int(rng.integers(0, 40)) for _ in range(4)]
    n = [int(rng.integers(8, 200)) for _ in range(2)]
    a = "#ISINP\n" + "-" * 40 + "\n"
    a += ln("GRIDH", str(ex[0]), str(n[0]), str(ex[1]))
    a += ln("GRIDV", str(ex[2]), str(n[1]), str(ex[3]))
    a += ln("GRIDEXT", num(rng.uniform(1.0, 1.15)))
    a += ln("GRIDSPC", num(rng.uniform(1.0, 100.0)))
    spcm = [num(rng.choice([-1.0, 1.0]) * rng.uniform(1.0, 300.0)) for _ in range(int(rng.integers(1, 3)))]
    a += ln("GRIDSPCM", *spcm)
    nl = int(rng.integers(1, 4))
    Ny = ex[2] + n[1] + ex[3]
    a += ln("LAYNUM", str(nl))
    a += ln("LAYPOS", *[str(int(rng.integers(0, Ny + 1))) for _ in range(nl)])
    if rng.random() < 0.5:
        a += "\n-- a comment\n\n"
    a += "<<END\n"
    nm = int(rng.integers(1, 4))
    m = "#ISMAT\n" + "-" * 40 + "\n"
    m += ln("MATNAME", *["mat%d_" % k for k in range(nm)])
    m += ln("EOSTYPE", *[str(rng.choice(["aneos", "tillo", "ANEOS", "other"])) for _ in range(nm)])
    for name in ("POIS", "TMELT0", "TFRAC", "ASIMON", "YINT0", "FRICINT", "YLIMINT", "IVANOV_B", "JC_TREF",
                 "GAMBETA"):
        if rng.random() < 0.7:
            m += ln(name, *[num(rng.uniform(0.1, 1.0e9)) for _ in range(nm)])
    m += "<<END\n"
    return a, m


@pytest.mark.skipif(not os.path.exists(REF_GEN), reason="oracle/_ref/ref_generate_input is not built here")
def test_isale_conversion_random_against_reference_build(tmp_path):
    rng = np.random.default_rng(7)
    for i in range(40):
        a, m = _random_pair(rng)
        d = tmp_path / ("r%02d" % i)
        d.mkdir()
        (d / "asteroid.inp").write_text(a)
        (d / "material.inp").write_text(m)
        ref = subprocess.run([REF_GEN], cwd=d, capture_output=True, check=True).stdout
        got = ps.isale_to_inp(str(d / "asteroid.inp"), str(d / "material.inp"))
        assert got.encode() == ref, (i, a, m)


def test_isale_conversion_errors(tmp_path):
    with pytest.raises(RuntimeError):
        ps.isale_to_inp(str(tmp_path / "none.inp"), str(tmp_path / "none2.inp"))
    (tmp_path / "a.inp").write_text("#ISINP\nGRIDH  h : 0 : 0 : 0\nGRIDV v : 0 : 10 : 0\n")
    (tmp_path / "m.inp").write_text("#ISMAT\nMATNAME n : a_\n")
    with pytest.raises(RuntimeError):          # no horizontal cell at all
        ps.isale_to_inp(str(tmp_path / "a.inp"), str(tmp_path / "m.inp"))


def _fake_runs(d, prefix, nproc, ncycle, rng):
    """Per-domain state files with text lines (one holding the VTKFile tags)."""
    os.makedirs(d, exist_ok=True)
    files = {}
    for c in range(ncycle):
        for r in range(nproc):
            body = "".join("line %d %d %s\n" % (c, r, "x" * int(rng.integers(0, 80))) for _ in range(int(rng.integers(1, 30))))
            txt = '<?xml version="1.0"?>\n<VTKFile type="StructuredGrid">\n' + body + "</VTKFile>\n"
            p = os.path.join(d, "%s.proc%04d.%04d.vts" % (prefix, r, c))
            open(p, "w").write(txt)
            files[p] = txt.encode()
    return files


def test_pack_split_format_and_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    vts = str(tmp_path / "vts")
    files = _fake_runs(vts, "T", 2, 9, rng)
    out = str(tmp_path / "mvts")
    ps.pack_vts(vts, "T", 2, 0, 9, 4, out)
    per = 9 // 4 + 1                                   # cycles per pack, as the reference splits them
    for t in range(4):
        pk = open(os.path.join(out, "T.vts.pack%04d.0000" % t), "rb").read()
        exp = b""
        for c in range(t * per, min((t + 1) * per, 9)):
            for r in range(2):
                p = os.path.join(vts, "T.proc%04d.%04d.vts" % (r, c))
                exp += b"#BEGINFILE:" + p.encode() + b"\n" + files[p] + b"#ENDFILE\n"
        assert pk == exp, t
    # unpack into a fresh tree: every file back, byte for byte
    shutil.move(vts, str(tmp_path / "orig"))
    n = sum(ps.unpack_vts(os.path.join(out, "T.vts.pack%04d.0000" % t)) for t in range(4))
    assert n == 18
    for p, b in files.items():
        assert open(p, "rb").read() == b
    # default range: from MIN up to the first cycle with a domain missing
    os.remove(os.path.join(vts, "T.proc0001.0005.vts"))
    out2 = str(tmp_path / "m2")
    ps.pack_vts(vts, "T", 2, 0, -1, 1, out2)
    assert open(os.path.join(out2, "T.vts.pack0000.0000"), "rb").read().count(b"#ENDFILE") == 10
    with pytest.raises(RuntimeError):                 # an explicit range with a file missing
        ps.pack_vts(vts, "T", 2, 0, 9, 2, str(tmp_path / "m3"))


def test_merge_blocks_of_real_state_files(tmp_path):
    """Two domains' state files (written by the library's writer) merged per
    cycle into one multiblock file that holds both grids."""
    nx, ny, nmat = 6, 5, 3
    rng = np.random.default_rng(5)
    vts = str(tmp_path / "vts")
    os.makedirs(vts)
    for c in range(2):
        for r in range(2):
            fields = {}
            for name, _, _ in ps.FIELDS:
                try:
                    shape = ps.field_shape(name, nmat, nx, ny)
                except KeyError:
                    continue
                fields[name] = rng.standard_normal(shape)
            ps.write_vts(os.path.join(vts, "T.proc%04d.%04d.vts" % (r, c)), fields, nx, ny, nmat)
    out = str(tmp_path / "mvts")
    ps.merge_vts_blocks(vts, "T", 2, 0, -1, out)
    for c in range(2):
        root = ET.parse(os.path.join(out, "T.%04d.vts" % c)).getroot()
        assert root.tag == "VTKFile" and root.get("type") == "vtkMultiBlockDataSet"
        ds = root.find("vtkMultiBlockDataSet").findall("DataSet")
        assert [d.get("index") for d in ds] == ["0", "1"]
        for r, d in enumerate(ds):
            src = ET.parse(os.path.join(vts, "T.proc%04d.%04d.vts" % (r, c))).getroot()
            got, exp = d.find("StructuredGrid"), src.find("StructuredGrid")
            got.tail = exp.tail = None         # the white space after the element differs
            assert ET.tostring(got) == ET.tostring(exp)


@pytest.mark.skipif(not os.path.exists(MERGE), reason="salec_merge_vts is not built")
def test_merge_program(tmp_path):
    rng = np.random.default_rng(11)
    files = _fake_runs(str(tmp_path / "run" / "vts"), "Crater", 1, 5, rng)
    (tmp_path / "run" / "c.inp").write_text("[output]\n    prefix = Crater\n")
    r = subprocess.run([MERGE, "c.inp"], cwd=tmp_path / "run", capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "packed 1 domains" in r.stdout
    packs = sorted(os.listdir(tmp_path / "run" / "mvts"))
    assert packs == ["Crater.vts.pack%04d.0000" % t for t in range(4)]
    shutil.rmtree(tmp_path / "run" / "vts")
    os.makedirs(tmp_path / "run" / "vts")
    for p in packs:
        r = subprocess.run([MERGE, "--unpack", os.path.join("mvts", p)], cwd=tmp_path / "run", capture_output=True,
                           text=True)
        assert r.returncode == 0, r.stderr
    for p, b in files.items():
        assert open(p, "rb").read() == b
