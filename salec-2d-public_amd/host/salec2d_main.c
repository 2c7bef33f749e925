/*
 * salec2d_main.c -- command-line driver of a whole run, the counterpart of
 * the reference's main() (src/salec2d.c:11-295) on one GPU:
 *
 *   salec2d <file.inp> <eos_dir> [--steps N] [--log K] [--state out.bin]
 *
 * Setup runs on the host (salec_host_*), the initialisation tail and the
 * time loop on the device (salec_init, salec_step).  The loop stops under the
 * rules of update_cycle_control (src/sale2d.c:71-136): after maxstep+1 steps
 * or once t > time1 + 2 maxdt (or after --steps N steps).  Output writers are
 * not part of this driver: the crater profile (y_profile) is downloaded every
 * K steps to report the transient crater depth and diameter, and --state
 * writes the final fields as raw FP64 planes in field-id order.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/salec.h"
#include "setup.h"

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int die(const char *what, int rc)
{
    fprintf(stderr, "salec2d: %s: %s (%d)\n", what, salec_strerror(rc), rc);
    return 1;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s <file.inp> <eos_dir> [--steps N] [--log K] [--state out.bin]\n", argv[0]);
        return 2;
    }
    long max_steps = -1, log_every = 100;
    const char *state_out = NULL;
    for (int a = 3; a < argc; ++a) {
        if (!strcmp(argv[a], "--steps") && a + 1 < argc) max_steps = atol(argv[++a]);
        else if (!strcmp(argv[a], "--log") && a + 1 < argc) log_every = atol(argv[++a]);
        else if (!strcmp(argv[a], "--state") && a + 1 < argc) state_out = argv[++a];
        else { fprintf(stderr, "salec2d: unknown argument %s\n", argv[a]); return 2; }
    }
    if (log_every < 1) log_every = 1;

    double t_start = now_s();
    char err[512] = {0};
    salec_host *h = NULL;
    int rc = salec_host_load(argv[1], argv[2], &h, err, sizeof err);
    if (rc != SALEC_OK) { fprintf(stderr, "salec2d: %s\n", err); return 1; }
    if ((rc = salec_host_setup(h)) != SALEC_OK) return die("host setup", rc);
    salec_config cfg = *salec_host_config(h);
    double cyc[6];
    salec_host_cycle(h, cyc);
    const double time0 = cyc[0], time1 = cyc[1], maxdt = cyc[2], dt0 = cyc[3];
    const long maxstep = (long)cyc[4];
    const int nx = cfg.nx, ny = cfg.ny, nmat = cfg.nmat;
    cfg.max_tracers = salec_host_count_tracers(h);

    salec_ctx *ctx = NULL;
    if ((rc = salec_create(&cfg, &ctx)) != SALEC_OK) return die("salec_create", rc);
    for (int f = 0; f < SALEC_NFIELDS; ++f)
        if ((rc = salec_upload(ctx, f, salec_host_field(h, f))) != SALEC_OK) return die("salec_upload", rc);
    double damp_t = cyc[5];
    if ((rc = salec_init(ctx, &damp_t)) != SALEC_OK) return die("salec_init", rc);
    salec_host_set_damp_time(h, damp_t);

    /* tracers are placed after the initial vertex velocity is known */
    long ntr = cfg.max_tracers;
    salec_tracers tr = {0};
    if (ntr > 0) {
        if ((rc = salec_download(ctx, SALEC_V_VEL, salec_host_field(h, SALEC_V_VEL))) != SALEC_OK)
            return die("salec_download", rc);
        tr.n = ntr;
        tr.pos = calloc(2 * ntr, sizeof(double));
        tr.vel = calloc(2 * ntr, sizeof(double));
        tr.mpre = calloc(ntr, sizeof(float));
        tr.mtem = calloc(ntr, sizeof(float));
        int **ip[] = {&tr.gx, &tr.gy, &tr.tag, &tr.matid, &tr.idx, &tr.idy, &tr.alive};
        for (unsigned k = 0; k < sizeof ip / sizeof ip[0]; ++k) *ip[k] = calloc(ntr, sizeof(int));
        if ((rc = salec_host_init_tracers(h, &tr)) != SALEC_OK) return die("tracer init", rc);
        if ((rc = salec_upload_tracers(ctx, &tr)) != SALEC_OK) return die("salec_upload_tracers", rc);
    }
    if ((rc = salec_set_time(ctx, time0, dt0)) != SALEC_OK) return die("salec_set_time", rc);

    /* cell-centre radius of each padded column, for the crater diameter */
    const double *epos = salec_host_field(h, SALEC_E_POS);
    long cw = nx + 4;
    double *ecx = malloc(sizeof(double) * cw);
    for (long i = 0; i < cw; ++i) ecx[i] = epos[2 * cw + i];  /* centre row, component 0 */
    double *yprof = malloc(sizeof(double) * salec_grid_points(SALEC_GRID_COLUMN, nx, ny));

    double t_loop = now_s();
    printf("salec2d: %dx%d cells, %d materials, %ld tracers, damp time %.4e s, setup %.2f s\n",
           nx, ny, nmat, ntr, damp_t, t_loop - t_start);
    salec_stats st = {0};
    double best_depth = -INFINITY, best_diam = 0.0, best_t = 0.0;
    long k = 0;
    for (;;) {
        /* the stop rule needs t after every step, so each step returns its stats */
        rc = salec_step(ctx, 0u, &st);
        if (rc != SALEC_OK) return die("salec_step", rc);
        ++k;
        int stop = (k > maxstep) || (st.t > time1 + 2.0 * maxdt) || (max_steps >= 0 && k >= max_steps);
        if (k % log_every == 0 || stop) {
            if ((rc = salec_download(ctx, SALEC_Y_PROFILE, yprof)) != SALEC_OK) return die("salec_download", rc);
            double depth, diam;
            salec_crater_metrics(yprof, ecx, nx, &depth, &diam);
            if (depth > best_depth) { best_depth = depth; best_diam = diam; best_t = st.t; }
            printf("step %8ld  t %.6e  dt %.4e  depth %.2f m  diameter %.2f m\n", k, st.t, st.dt, depth, diam);
            fflush(stdout);
        }
        if (stop) break;
    }
    salec_sync(ctx);
    double t_end = now_s();
    int bf; long bi;
    rc = salec_check(ctx, &bf, &bi);
    if (rc != SALEC_OK) fprintf(stderr, "salec2d: non-finite %s at %ld\n", salec_field_name(bf), bi);
    printf("salec2d: %ld steps in %.3f s (%.4e cell-steps/s); max transient depth %.2f m, diameter %.2f m at t %.4e\n",
           k, t_end - t_loop, (double)nx * ny * k / (t_end - t_loop), best_depth, best_diam, best_t);

    if (state_out) {
        FILE *fp = fopen(state_out, "wb");
        if (!fp) { fprintf(stderr, "salec2d: cannot write %s\n", state_out); return 1; }
        for (int f = 0; f < SALEC_NFIELDS; ++f) {
            int grid, nc;
            salec_field_info(f, nmat, &grid, &nc);
            long n = salec_grid_points(grid, nx, ny) * nc;
            double *buf = malloc(sizeof(double) * n);
            salec_download(ctx, f, buf);
            fwrite(buf, sizeof(double), n, fp);
            free(buf);
        }
        fclose(fp);
    }
    salec_destroy(ctx);
    salec_host_free(h);
    return rc == SALEC_OK ? 0 : 1;
}
