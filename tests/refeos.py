"""ctypes access to oracle/_ref/libref_eos.so: the reference's own EOS and
geometry sources (src/eos/eos_state.c, src/eos/aneos_sieden.c,
src/geometry/*.c, src/io/InputParser.c) compiled as they are, behind the flat
entry points of oracle/ref_shim.c.  Used only to write the golden vectors
under tests/golden/ and, where the library is present, to check the oracle
against it directly."""
import ctypes as C
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libref_eos.so")

EOS_ANEOS, EOS_TILLOTSON, EOS_SIEDEN = 0, 1, 2
_R = None


def available():
    return os.path.exists(REF_PATH)


def rlib():
    global _R
    if _R is not None:
        return _R
    R = C.CDLL(REF_PATH)
    d, dp, vp = C.c_double, C.POINTER(C.c_double), C.c_void_p
    R.ref_eos_load.restype = vp
    R.ref_eos_load.argtypes = [C.c_char_p, C.c_int]
    R.ref_eos_free.argtypes = [vp]
    R.ref_calculate_state.argtypes = [vp, d, d, dp, dp, dp]
    R.ref_calculate_state_correct.argtypes = [vp, d, d, d, dp, dp, dp, d]
    R.ref_calculate_state_correct.restype = d
    R.ref_state_reference.argtypes = [vp, dp]
    R.ref_till_cold.argtypes = [vp, C.POINTER(dp), C.POINTER(dp), dp]
    R.ref_till_cold.restype = C.c_int
    R.ref_till_pres.argtypes = [vp, d, d, dp, dp]
    R.ref_till_pres.restype = d
    R.ref_till_temp.argtypes = [vp, d, d]
    R.ref_till_temp.restype = d
    R.ref_sd_dnorm.argtypes = [vp]
    R.ref_sd_dnorm.restype = d
    R.ref_bisearch.argtypes = [d, dp, C.c_int]
    R.ref_bisearch.restype = C.c_int
    R.ref_interpolate_eos_tp.argtypes = [vp, d, d, C.c_int]
    R.ref_interpolate_eos_tp.restype = d
    R.ref_laplace2d9.argtypes = [dp, dp, d]
    R.ref_laplace2d9.restype = d
    _R = R
    return R


class RefEOS:
    """One EOS file loaded by the reference loader (load_eos_table)."""

    def __init__(self, path, eos_type):
        self.R = rlib()
        self.p = self.R.ref_eos_load(path.encode(), eos_type)
        if not self.p:
            raise IOError(path)
        self.type = eos_type

    def state(self, den, eng):
        a = [C.c_double() for _ in range(3)]
        self.R.ref_calculate_state(self.p, den, eng, *[C.byref(x) for x in a])
        return a[0].value, a[1].value, a[2].value

    def state_correct(self, den, eng, plim, pty=1.0):
        a = [C.c_double() for _ in range(3)]
        self.R.ref_calculate_state_correct(self.p, den, eng, plim, *[C.byref(x) for x in a], pty)
        return a[0].value, a[1].value, a[2].value

    def state_reference(self):
        out = (C.c_double * 7)()
        self.R.ref_state_reference(self.p, out)
        return dict(zip(["MeltDen", "VaporDen", "c_heat", "thermexp", "den0", "csd0", "gamfrac"], list(out)))

    def cold_table(self):
        xd, ye, sd = C.POINTER(C.c_double)(), C.POINTER(C.c_double)(), C.c_double()
        n = self.R.ref_till_cold(self.p, C.byref(xd), C.byref(ye), C.byref(sd))
        return [xd[i] for i in range(n)], [ye[i] for i in range(n)], sd.value

    def till_pres(self, e, rho):
        p, c = C.c_double(), C.c_double()
        self.R.ref_till_pres(self.p, e, rho, C.byref(p), C.byref(c))
        return p.value, c.value

    def eos_tp(self, tem, pre):
        """(density, energy density) of the state (T, P), interpolate_eos_tp."""
        return (self.R.ref_interpolate_eos_tp(self.p, tem, pre, -1), self.R.ref_interpolate_eos_tp(self.p, tem, pre, 0))

    def till_temp(self, e, rho):
        return self.R.ref_till_temp(self.p, e, rho)
