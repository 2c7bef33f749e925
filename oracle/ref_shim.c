/*
 * ref_shim.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Flat C entry points over the reference's own EOS and geometry sources
 * (src/eos/eos_state.c, src/eos/aneos_sieden.c, src/geometry/*.c,
 * src/io/InputParser.c), which compile on their own with gcc.  The Makefile
 * target "ref" builds this file together with those sources, where they lie
 * under the reference checkout, into _ref/libref_eos.so; tests use it to pin
 * the oracle's EOS functions and the host's Laplacian weights against the
 * reference's own arithmetic.  Nothing here is part of the product.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "eos_state.h"
#include "interpolate_2d9.h"

/* Load an EOS file with the reference loader (load_eos_table,
 * src/eos/eos_state.c:852-900); type is the eos_type value
 * (0 ANEOS, 1 Tillotson, 2 ANEOSSieDen). */
void *ref_eos_load(const char *path, int type)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return NULL;
    eos_table *e = (eos_table *)calloc(1, sizeof(eos_table));
    e->type = (eos_type)type;
    load_eos_table(e, fp); /* every loader closes fp itself */
    return e;
}

void ref_eos_free(void *p)
{
    free(p); /* the tables stay allocated: the reference never frees them either */
}

void ref_calculate_state(void *p, double den, double eng, double *tem, double *pre, double *csd)
{
    calculate_state((eos_table *)p, den, eng, tem, pre, csd);
}

double ref_calculate_state_correct(void *p, double den, double eng, double plim,
                                   double *tem, double *pre, double *csd, double pty)
{
    double a = pty;
    return calculate_state_correct((eos_table *)p, den, eng, plim, tem, pre, csd, &a);
}

/* state_reference values set by the loader's *InitStateRef:
 * out = {MeltDen, VaporDen, c_heat, thermexp, den0, csd0, gamfrac}. */
void ref_state_reference(void *p, double out[7])
{
    state_reference *s = ((eos_table *)p)->ref;
    out[0] = s->MeltDen; out[1] = s->VaporDen; out[2] = s->c_heat; out[3] = s->thermexp;
    out[4] = s->den0; out[5] = s->csd0; out[6] = s->gamfrac;
}

/* Tillotson table only: the cold-energy table of TillColdEnergy. */
int ref_till_cold(void *p, const double **xden, const double **yeng, double *stepden)
{
    tillotson_table *t = (tillotson_table *)((eos_table *)p)->table;
    *xden = t->xDen; *yeng = t->yEng; *stepden = t->StepDen;
    return t->nDen;
}

double ref_till_pres(void *p, double e, double rho, double *pre, double *cs)
{
    return TillPres((tillotson_table *)((eos_table *)p)->table, e, rho, pre, cs);
}

double ref_till_temp(void *p, double e, double rho)
{
    return TillTemp((tillotson_table *)((eos_table *)p)->table, e, rho);
}

/* SieDen table only: Dnorm found by ANEOSSieDenInitStateRef. */
double ref_sd_dnorm(void *p)
{
    return ((aneos_table *)((eos_table *)p)->table)->Dnorm;
}

int ref_bisearch(double x, double *xl, int n)
{
    return bisearch(x, xl, n);
}

/* 9-point Laplacian weight of Laplace2d9 (src/geometry/interpolate_2d9.c:398). */
double ref_laplace2d9(const double *xi, const double *yi, double gamma)
{
    return Laplace2d9(xi, yi, gamma);
}

/* interpolate_eos_tp (src/eos/eos_state.c:902-921); DataId -1 density, 0 energy density */
double ref_interpolate_eos_tp(void *p, double tem, double pre, int data_id)
{
    return interpolate_eos_tp((eos_table *)p, tem, pre, data_id);
}
