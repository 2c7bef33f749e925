/*
 * salec2d_main.c -- command-line driver of a whole run, the counterpart of
 * the reference's main() (src/salec2d.c:11-295) on one GPU:
 *
 *   salec2d <file.inp> <eos_dir> [--steps N] [--log K] [--state out.bin]
 *           [--out DIR] [--no-output]
 *
 * Setup runs on the host (salec_host_*), the initialisation tail and the
 * time loop on the device (salec_init, salec_step).  The loop stops under the
 * rules of update_cycle_control (src/sale2d.c:71-136): after maxstep+1 steps
 * or once t > time1 + 2 maxdt (or after --steps N steps).
 *
 * Output follows the reference's schedule and files: the first step and then
 * every output.time of simulated time (and the first output.init steps, and
 * every output.out_cycle steps), at the mid-step output point
 * (src/salec2d.c:241-245), it writes DIR/vts/<prefix>.proc0000.NNNN.vts and
 * DIR/vtp/<prefix>.tracer.proc0000.NNNN.vtp with their .pvts, .pvtp and .vtm
 * indexes in DIR, and once after tracer placement DIR/txt/mass.tracer.
 * proc0000.txt.  Directories are created when missing, never emptied.
 * The crater profile (y_profile) is downloaded every K steps to report the
 * transient crater depth and diameter, and --state writes the final fields as
 * raw FP64 planes in field-id order.
 */
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "../../include/salec.h"
#include "../../include/salec_output.h"
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

static int ensure_dir(const char *d)
{
    if (mkdir(d, 0755) == 0 || errno == EEXIST) return SALEC_OK;
    fprintf(stderr, "salec2d: cannot create %s\n", d);
    return SALEC_EINVAL;
}

/* fields a state file reads (point data, cell data, coordinates) */
static const int kOutFields[] = {
    SALEC_V_VEL, SALEC_V_ACC, SALEC_V_BF, SALEC_V_VOF, SALEC_V_POS, SALEC_E_DEN, SALEC_E_PRE, SALEC_E_WPT,
    SALEC_M_VOF, SALEC_E_CSD, SALEC_E_DAM, SALEC_E_TPS, SALEC_E_Q, SALEC_E_TEM, SALEC_M_ENG, SALEC_M_DEN,
    SALEC_E_STE2S, SALEC_E_STA2S, SALEC_E_STE, SALEC_E_STRENGTH, SALEC_E_VIB, SALEC_E_EJT, SALEC_E_CVS, SALEC_M_PTY};

/* one output (sale2d_write_vtk + sale2d_write_vtp and their indexes) of the
 * current device state into dir; host fields of h serve as buffers */
static int write_output(salec_ctx *ctx, salec_host *h, salec_tracers *tr, const char *dir, const char *prefix,
                        const char *format, int cycle, int nx, int ny, int nmat)
{
    int rc;
    const double *tab[SALEC_NFIELDS] = {0};
    for (unsigned i = 0; i < sizeof kOutFields / sizeof kOutFields[0]; ++i) {
        int f = kOutFields[i];
        if ((rc = salec_download(ctx, f, salec_host_field(h, f))) != SALEC_OK) return rc;
        tab[f] = salec_host_field(h, f);
    }
    size_t need = strlen(dir) + 2 * strlen(prefix) + 64;
    char *p = malloc(need);
    if (!p) return SALEC_ENOMEM;
    snprintf(p, need, "%s/vts/%s.proc%04d.%04d.vts", dir, prefix, 0, cycle);
    rc = salec_write_vts(p, nx, ny, nmat, tab, format);
    if (rc == SALEC_OK) {
        snprintf(p, need, "%s/%s.%04d.pvts", dir, prefix, cycle);
        rc = salec_write_pvts(p, prefix, cycle, nx, ny, nmat);
    }
    if (rc == SALEC_OK) rc = salec_write_vtm(dir, prefix, cycle, 1);
    if (rc == SALEC_OK && tr->n > 0) {
        long cap = tr->n;
        if ((rc = salec_download_tracers(ctx, tr)) != SALEC_OK) { free(p); return rc; }
        snprintf(p, need, "%s/vtp/%s.tracer.proc%04d.%04d.vtp", dir, prefix, 0, cycle);
        rc = salec_write_vtp(p, NULL, nx, ny, tr, tab[SALEC_E_DEN], tab[SALEC_E_TEM], tab[SALEC_E_PRE]);
        if (rc == SALEC_OK) {
            snprintf(p, need, "%s/%s.tracer.%04d.pvtp", dir, prefix, cycle);
            rc = salec_write_pvtp(p, prefix, cycle, 1);
        }
        tr->n = cap;
    }
    free(p);
    return rc;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s <file.inp> <eos_dir> [--steps N] [--log K] [--state out.bin]\n", argv[0]);
        return 2;
    }
    long max_steps = -1, log_every = 100;
    const char *state_out = NULL, *out_dir = ".";
    int output = 1;
    for (int a = 3; a < argc; ++a) {
        if (!strcmp(argv[a], "--steps") && a + 1 < argc) max_steps = atol(argv[++a]);
        else if (!strcmp(argv[a], "--log") && a + 1 < argc) log_every = atol(argv[++a]);
        else if (!strcmp(argv[a], "--state") && a + 1 < argc) state_out = argv[++a];
        else if (!strcmp(argv[a], "--out") && a + 1 < argc) out_dir = argv[++a];
        else if (!strcmp(argv[a], "--no-output")) output = 0;
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
    const char *prefix, *format;
    double out_time;
    int init_step, out_cycle;
    salec_host_output(h, &prefix, &format, &out_time, &init_step, &out_cycle);
    if (output) {
        char d[4096];
        const char *sub[] = {"", "/vts", "/vtp", "/txt"};
        for (int i = 0; i < 4; ++i) {
            snprintf(d, sizeof d, "%s%s", out_dir, sub[i]);
            if (ensure_dir(d) != SALEC_OK) return 1;
        }
    }
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
        if (output) {                    /* export_init_tracers, src/sale2d.c:2093-2126 */
            double *eden = salec_host_field(h, SALEC_E_DEN);
            if ((rc = salec_download(ctx, SALEC_E_DEN, eden)) != SALEC_OK) return die("salec_download", rc);
            char p[4096];
            snprintf(p, sizeof p, "%s/txt/mass.tracer.proc%04d.txt", out_dir, 0);
            if ((rc = salec_write_mass_tracers(p, nx, ny, &tr, eden, salec_host_field(h, SALEC_E_VOL))) != SALEC_OK)
                return die("tracer export", rc);
        }
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
    int out_vts = 1, k_out = 1, i_out = 0, n_out = 0;   /* src/salec2d.c:98-99 */
    double t_out = 0.0;
    for (;;) {
        /* the stop rule needs t after every step, so each step returns its stats */
        if (output && out_vts) {
            if ((rc = salec_step(ctx, SALEC_STEP_TO_OUTPUT, &st)) != SALEC_OK) return die("salec_step", rc);
            double t0w = now_s();
            rc = write_output(ctx, h, &tr, out_dir, prefix, format, k_out - 1, nx, ny, nmat);
            if (rc != SALEC_OK) return die("output", rc);
            t_out += now_s() - t0w;
            ++n_out;
            if ((rc = salec_step(ctx, SALEC_STEP_FROM_OUTPUT, &st)) != SALEC_OK) return die("salec_step", rc);
        } else {
            if ((rc = salec_step(ctx, 0u, &st)) != SALEC_OK) return die("salec_step", rc);
        }
        ++k;
        /* output state of the next step, update_cycle_control (src/sale2d.c:96-117) */
        if (st.t >= out_time * (k_out - i_out)) { ++k_out; out_vts = 1; }
        else if (k <= init_step) { ++k_out; ++i_out; out_vts = 1; }
        else if (out_cycle > 0 && k % out_cycle == 0) { ++k_out; ++i_out; out_vts = 1; }
        else out_vts = 0;
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
    if (output) printf("salec2d: %d outputs (%s) in %s, %.3f s of writing\n", n_out, prefix, out_dir, t_out);

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
