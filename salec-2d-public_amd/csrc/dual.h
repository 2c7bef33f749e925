/*
 * dual.h -- forward-mode (tangent-linear) arithmetic for the step kernels.
 *
 * A Dual carries a value and its derivatives with respect to NTAN scalar
 * parameters.  The step kernels are templates on their scalar type S: with
 * S = double they are the solver; with S = Dual every value is computed by
 * the same operations in the same order, so the value parts are bit-equal to
 * the S = double run, and the derivative parts are the exact derivatives of
 * that computation with every branch (limiter, yield, cutoff, table cell)
 * taken as the values decide -- the frozen-branch linearisation of the step.
 *
 * Field access goes through fa<S, FIELD>(): with S = double it is the plain
 * field element; with S = Dual it is a DRef that reads and writes the value
 * plane and the NTAN tangent planes of the field together.  Fields whose
 * values never depend on the parameters (mesh geometry, PML profiles,
 * the ejecta flag, the crater profile), or whose values enter no other
 * value (the face weights e_wpt, read only by comparisons, and the
 * output-only e_strength), have no tangent planes and
 * are read as plain doubles in both modes.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include "../../include/salec.h"

namespace salec {

/* the double overloads stay visible beside the Dual ones below */
using ::sqrt; using ::fabs; using ::exp; using ::log; using ::tanh; using ::pow; using ::floor; using ::isnan;
using ::fmin; using ::fmax;

#define NTAN 3

struct Dual {
    double v;
    double d[NTAN];
    Dual() = default;
    __host__ __device__ constexpr Dual(double x) : v(x), d{0., 0., 0.} {}
};

__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(const Dual &x) { return x.v; }

/* ---- arithmetic: the value part is exactly the double operation -------- */
__device__ __forceinline__ Dual operator+(const Dual &a, const Dual &b)
{ Dual r; r.v = a.v + b.v; for (int k = 0; k < NTAN; ++k) r.d[k] = a.d[k] + b.d[k]; return r; }
__device__ __forceinline__ Dual operator+(const Dual &a, double b)
{ Dual r; r.v = a.v + b; for (int k = 0; k < NTAN; ++k) r.d[k] = a.d[k]; return r; }
__device__ __forceinline__ Dual operator+(double a, const Dual &b)
{ Dual r; r.v = a + b.v; for (int k = 0; k < NTAN; ++k) r.d[k] = b.d[k]; return r; }
__device__ __forceinline__ Dual operator-(const Dual &a, const Dual &b)
{ Dual r; r.v = a.v - b.v; for (int k = 0; k < NTAN; ++k) r.d[k] = a.d[k] - b.d[k]; return r; }
__device__ __forceinline__ Dual operator-(const Dual &a, double b)
{ Dual r; r.v = a.v - b; for (int k = 0; k < NTAN; ++k) r.d[k] = a.d[k]; return r; }
__device__ __forceinline__ Dual operator-(double a, const Dual &b)
{ Dual r; r.v = a - b.v; for (int k = 0; k < NTAN; ++k) r.d[k] = -b.d[k]; return r; }
__device__ __forceinline__ Dual operator-(const Dual &a)
{ Dual r; r.v = -a.v; for (int k = 0; k < NTAN; ++k) r.d[k] = -a.d[k]; return r; }
__device__ __forceinline__ Dual operator*(const Dual &a, const Dual &b)
{ Dual r; r.v = a.v * b.v; for (int k = 0; k < NTAN; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k]; return r; }
__device__ __forceinline__ Dual operator*(const Dual &a, double b)
{ Dual r; r.v = a.v * b; for (int k = 0; k < NTAN; ++k) r.d[k] = a.d[k] * b; return r; }
__device__ __forceinline__ Dual operator*(double a, const Dual &b)
{ Dual r; r.v = a * b.v; for (int k = 0; k < NTAN; ++k) r.d[k] = a * b.d[k]; return r; }
__device__ __forceinline__ Dual operator/(const Dual &a, const Dual &b)
{
    Dual r; r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
    for (int k = 0; k < NTAN; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
__device__ __forceinline__ Dual operator/(const Dual &a, double b)
{ Dual r; r.v = a.v / b; for (int k = 0; k < NTAN; ++k) r.d[k] = a.d[k] / b; return r; }
__device__ __forceinline__ Dual operator/(double a, const Dual &b)
{
    Dual r; r.v = a / b.v;
    const double f = -r.v / b.v;
    for (int k = 0; k < NTAN; ++k) r.d[k] = f * b.d[k];
    return r;
}
__device__ __forceinline__ Dual &operator+=(Dual &a, const Dual &b) { a = a + b; return a; }
__device__ __forceinline__ Dual &operator-=(Dual &a, const Dual &b) { a = a - b; return a; }
__device__ __forceinline__ Dual &operator*=(Dual &a, const Dual &b) { a = a * b; return a; }
__device__ __forceinline__ Dual &operator/=(Dual &a, const Dual &b) { a = a / b; return a; }
__device__ __forceinline__ Dual &operator+=(Dual &a, double b) { a = a + b; return a; }
__device__ __forceinline__ Dual &operator-=(Dual &a, double b) { a = a - b; return a; }
__device__ __forceinline__ Dual &operator*=(Dual &a, double b) { a = a * b; return a; }
__device__ __forceinline__ Dual &operator/=(Dual &a, double b) { a = a / b; return a; }

/* ---- comparisons decide on values only (frozen branches) --------------- */
#define SALEC_DUAL_CMP_(op) \
    __device__ __forceinline__ bool operator op(const Dual &a, const Dual &b) { return a.v op b.v; } \
    __device__ __forceinline__ bool operator op(const Dual &a, double b) { return a.v op b; } \
    __device__ __forceinline__ bool operator op(double a, const Dual &b) { return a op b.v; }
SALEC_DUAL_CMP_(<) SALEC_DUAL_CMP_(>) SALEC_DUAL_CMP_(<=) SALEC_DUAL_CMP_(>=)
SALEC_DUAL_CMP_(==) SALEC_DUAL_CMP_(!=)
#undef SALEC_DUAL_CMP_

/* ---- elementary functions: value by the double function, derivative by
 *      its closed form; where the derivative is unbounded (sqrt and
 *      pow(x, p < 1) at x = 0) the tangent is taken as zero ---------------- */
__device__ __forceinline__ Dual sqrt(const Dual &a)
{
    Dual r; r.v = ::sqrt(a.v);
    const double f = r.v > 0. ? 0.5 / r.v : 0.;
    for (int k = 0; k < NTAN; ++k) r.d[k] = f * a.d[k];
    return r;
}
__device__ __forceinline__ Dual fabs(const Dual &a)
{
    Dual r; r.v = ::fabs(a.v);
    const double s = a.v < 0. ? -1.0 : 1.0;
    for (int k = 0; k < NTAN; ++k) r.d[k] = s * a.d[k];
    return r;
}
__device__ __forceinline__ Dual exp(const Dual &a)
{ Dual r; r.v = ::exp(a.v); for (int k = 0; k < NTAN; ++k) r.d[k] = r.v * a.d[k]; return r; }
__device__ __forceinline__ Dual log(const Dual &a)
{ Dual r; r.v = ::log(a.v); for (int k = 0; k < NTAN; ++k) r.d[k] = a.d[k] / a.v; return r; }
__device__ __forceinline__ Dual tanh(const Dual &a)
{
    Dual r; r.v = ::tanh(a.v);
    const double f = 1.0 - r.v * r.v;
    for (int k = 0; k < NTAN; ++k) r.d[k] = f * a.d[k];
    return r;
}
__device__ __forceinline__ Dual pow(const Dual &a, double p)
{
    Dual r; r.v = ::pow(a.v, p);
    const double f = a.v != 0. ? p * ::pow(a.v, p - 1.0) : (p == 1.0 ? 1.0 : 0.);
    for (int k = 0; k < NTAN; ++k) r.d[k] = f * a.d[k];
    return r;
}
__device__ __forceinline__ double floor(const Dual &a) { return ::floor(a.v); }
/* fmin/fmax against a constant: the value is the double function's (NaN
 * handling included), the derivative that of the operand it returns */
__device__ __forceinline__ Dual fmin(double a, const Dual &b)
{
    Dual r(::fmin(a, b.v));
    if (r.v == b.v) for (int k = 0; k < NTAN; ++k) r.d[k] = b.d[k];
    return r;
}
__device__ __forceinline__ Dual fmax(double a, const Dual &b)
{
    Dual r(::fmax(a, b.v));
    if (r.v == b.v) for (int k = 0; k < NTAN; ++k) r.d[k] = b.d[k];
    return r;
}
__device__ __forceinline__ bool isnan(const Dual &a) { return ::isnan(a.v); }

/* Selection of one of two Duals word by word.  A conditional expression on
 * two Dual lvalues selects an address, which keeps both operands in scratch
 * memory; selecting each word keeps them in registers. */
__device__ __forceinline__ Dual pick(bool c, const Dual &a, const Dual &b)
{ Dual r; r.v = c ? a.v : b.v; for (int k = 0; k < NTAN; ++k) r.d[k] = c ? a.d[k] : b.d[k]; return r; }
__device__ __forceinline__ double pick(bool c, double a, double b) { return c ? a : b; }

/* selection helpers of the reference (Max/Min/Wind), both scalar types */
__device__ __forceinline__ Dual dMax(const Dual &a, const Dual &b) { return pick(a > b, a, b); }
__device__ __forceinline__ Dual dMax(const Dual &a, double b) { return pick(a > b, a, Dual(b)); }
__device__ __forceinline__ Dual dMax(double a, const Dual &b) { return pick(a > b, Dual(a), b); }
__device__ __forceinline__ Dual dMin(const Dual &a, const Dual &b) { return pick(a < b, a, b); }
__device__ __forceinline__ Dual dMin(const Dual &a, double b) { return pick(a < b, a, Dual(b)); }
__device__ __forceinline__ Dual dMin(double a, const Dual &b) { return pick(a < b, Dual(a), b); }
__device__ __forceinline__ Dual dWind(const Dual &x, double l, double r)
{ return pick(x < l, Dual(l), pick(x > r, Dual(r), x)); }

/* wavefront exchange of a Dual (all four words) */
__device__ __forceinline__ Dual shfl_xor(const Dual &a, int off)
{
    Dual r; r.v = __shfl_xor(a.v, off, 64);
    for (int k = 0; k < NTAN; ++k) r.d[k] = __shfl_xor(a.d[k], off, 64);
    return r;
}
__device__ __forceinline__ double shfl_xor(double a, int off) { return __shfl_xor(a, off, 64); }

/* ---- tangent planes of the fields -------------------------------------- */
/* Tangent k of a field element at primal offset o (component * points +
 * index) lives at t[field][k * st[field] + o], st = ncomp * points. */
struct TFields {
    double *t[SALEC_NFIELDS];
    long st[SALEC_NFIELDS];
};
struct NoTan {};
template <class S> using TF = typename std::conditional<std::is_same<S, double>::value, NoTan, TFields>::type;

/* fields whose values never depend on the parameters */
__host__ __device__ constexpr bool tangent_field(int f)
{
    return !(f == SALEC_E_EJT || f == SALEC_E_WPT || f == SALEC_E_STRENGTH || f == SALEC_E_POS || f == SALEC_E_VOL || f == SALEC_E_GRAD ||
             f == SALEC_E_SIGMA || f == SALEC_E_DAMPREF || f == SALEC_V_POS || f == SALEC_V_BF ||
             f == SALEC_V_VOL || f == SALEC_V_SUBVOL || f == SALEC_V_SUF || f == SALEC_V_LAPLACE ||
             f == SALEC_BX_NORMAL || f == SALEC_BY_NORMAL || f == SALEC_Y_PROFILE);
}

/* handle on one field element and its tangents */
struct DRef {
    double *p, *t;
    long st;
    __device__ operator Dual() const
    {
        Dual r; r.v = *p;
        for (int k = 0; k < NTAN; ++k) r.d[k] = t[k * st];
        return r;
    }
    __device__ const DRef &operator=(const Dual &x) const
    {
        *p = x.v;
        for (int k = 0; k < NTAN; ++k) t[k * st] = x.d[k];
        return *this;
    }
    __device__ const DRef &operator=(const DRef &x) const { return *this = Dual(x); }
    __device__ const DRef &operator+=(const Dual &x) const { return *this = Dual(*this) + x; }
    __device__ const DRef &operator-=(const Dual &x) const { return *this = Dual(*this) - x; }
    __device__ const DRef &operator*=(const Dual &x) const { return *this = Dual(*this) * x; }
};

template <class S, int FLD, class Tn>
__device__ __forceinline__ decltype(auto) fa(double *const *f, const Tn &T, long o)
{
    if constexpr (std::is_same<S, double>::value || !tangent_field(FLD)) {
        return (f[FLD][o]);
    } else {
        return DRef{f[FLD] + o, T.t[FLD] + o, T.st[FLD]};
    }
}

/* a material parameter, seeded as direction dir1 - 1 when dir1 > 0 */
template <class S> __device__ __forceinline__ S mparam(double x, int dir1);
template <> __device__ __forceinline__ double mparam<double>(double x, int) { return x; }
template <> __device__ __forceinline__ Dual mparam<Dual>(double x, int dir1)
{
    Dual r(x);
    if (dir1 > 0 && dir1 <= NTAN) r.d[dir1 - 1] = 1.0;
    return r;
}

} // namespace salec
