#!/usr/bin/env python3
"""Writes the iSALE-conversion fixtures under tests/golden/isale/.

Each case directory holds an iSALE-2D input pair (asteroid.inp,
material.inp) and expected.txt, the output of the reference's own converter
(src/io/generate_input.c with src/io/iSALEInputParser.c and
src/io/InputParser.c, built unchanged as oracle/_ref/ref_generate_input by
`make -C oracle ref`) run in that directory.  Cases:

  example    the input pair the reference ships as an example
             (example/sup_asteroid.inp, example/sup_material.inp);
  layered    extension zones on all four sides, positive GRIDSPCM in metres
             (two values), three layers, three materials with Tillotson,
             ANEOS and an unknown EOS type, Johnson-Cook and acoustic
             fluidisation parameters;
  plain      no extension zone (no damping layer), one material;
  quirks     what the reader does with odd text: values starting with '.'
             (dropped), a '"' cutting a value, blank and CRLF lines, a
             repeated key, GRIDSPCM with a second value <= 1e-2 (the first
             is used), MATNAME missing (the built-in default name).

Run from the repository root after `make -C oracle ref`.
"""
import os
import shutil
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "isale")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_generate_input")
REF_EXAMPLE = "/root/reference/example"


def line(name, desc, *vals):
    return "%-14s%-30s: %s\n" % (name, desc, "  : ".join(str(v) for v in vals))


def asteroid(gridh, gridv, ext, spc, spcm, laypos, extra=""):
    s = "#ISINP\n" + "-" * 72 + "\n"
    s += line("VERSION", "__DO NOT MODIFY__", "4.1")
    s += line("GRIDH", "horizontal cells", *gridh)
    s += line("GRIDV", "vertical cells", *gridv)
    s += line("GRIDEXT", "ext. factor", ext)
    s += line("GRIDSPC", "grid spacing", spc)
    s += line("GRIDSPCM", "max. grid spacing", *spcm)
    s += line("OBJNUM", "number of objects", 1)
    s += line("LAYNUM", "layers number", len(laypos))
    s += line("LAYPOS", "layer position", *laypos)
    s += extra
    return s + "<<END\n"


def material(names, eos, params):
    s = "#ISMAT\n" + "-" * 77 + "\n"
    if names is not None:
        s += line("MATNAME", "Material name", *names)
    s += line("EOSTYPE", "EOS type", *eos)
    for k, v in params:
        s += line(k, k.lower(), *v)
    return s + "<<END\n"


def cases():
    c = {}
    with open(os.path.join(REF_EXAMPLE, "sup_asteroid.inp")) as f:
        a = f.read()
    with open(os.path.join(REF_EXAMPLE, "sup_material.inp")) as f:
        m = f.read()
    c["example"] = (a, m)
    c["layered"] = (
        asteroid((20, 100, 30), (40, 120, 10), "1.1D0", "25.0", ("100.0", "50.0"), (80, 150, 160),
                 extra=line("T_SURF", "Surface temp", "293.D0")),
        material(("granite_", "basalt_", "ice_"), ("tillo", "aneos", "other"),
                 [("POIS", ("2.5D-01", "2.8D-01", "3.3D-01")), ("TMELT0", ("1.3730D+03", "1.3600D+03", "2.73D+02")),
                  ("YINT0", ("1.0D+07", "2.0D+07", "1.0D+06")), ("FRICINT", ("1.58D+00", "1.4D+00", "0.5D+00")),
                  ("JC_A", ("1.0E+08", "2.0E+08", "3.0E+08")), ("JC_N", ("0.3", "0.4", "0.5")),
                  ("IVANOV_A", ("1.0D-04", "1.0D-04", "1.0D-04")),
                  ("GAMETA", ("8.0D-03", "8.0D-03", "8.0D-03")), ("GAMBETA", ("115.", "115.", "115."))]))
    c["plain"] = (
        asteroid((0, 64, 0), (0, 96, 0), "1.05d0", "10.0", ("-20.D0",), (60,)),
        material(("dunite_",), ("aneos",), [("POIS", ("2.5000D-01",)), ("YLIMINT", ("3.5000D+09",))]))
    q = asteroid((10, 50, 10), (10, 60, 5), "1.2D0", ".5", ("-4.D0", "1.0D-3"), (40,),
                 extra=line("GRAV_V", "gravity", '-1.622D0 "comment') + "\r\n\n" +
                 line("GRIDSPC", "grid spacing again", "2.0"))
    q = q.replace("\n", "\r\n", 3)
    c["quirks"] = (q, material(None, ("tillo",), [("TFRAC", (".7D+00",)), ("ASIMON", ("1.52D+09",))]))
    return c


def main():
    os.makedirs(OUT, exist_ok=True)
    for name, (a, m) in cases().items():
        d = os.path.join(OUT, name)
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "asteroid.inp"), "w", newline="") as f:
            f.write(a)
        with open(os.path.join(d, "material.inp"), "w", newline="") as f:
            f.write(m)
        w = tempfile.mkdtemp(prefix="isale_")
        shutil.copy(os.path.join(d, "asteroid.inp"), w)
        shutil.copy(os.path.join(d, "material.inp"), w)
        r = subprocess.run([REF_BIN], cwd=w, capture_output=True, check=True)
        with open(os.path.join(d, "expected.txt"), "wb") as f:
            f.write(r.stdout)
        shutil.rmtree(w)
        print(name, len(r.stdout), "bytes")


if __name__ == "__main__":
    main()
