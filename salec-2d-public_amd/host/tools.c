/*
 * tools.c -- iSALE-2D input conversion and .vts packing
 * (include/salec_tools.h).
 *
 * iSALE input files hold one "NAME  description : v1 : v2 ..." line per
 * variable.  They are read into a key/value set sorted by key (case
 * insensitive, stable), list variables as "[v1, v2, ...]", as
 * OpeniSALEInputFile (src/io/iSALEInputParser.c:44-119) reads them; the
 * conversion then follows GenerateMesh and GenerateMaterial (:166-475) and
 * prints the resulting keys as generate_input does.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <sys/stat.h>
#include <sys/types.h>

#include "../../include/salec_tools.h"

/* Keys and values live in 100-byte fields as in the reference's InputFile
 * (MaxStrLen and MaxKeyId, src/io/InputParser.h:16-23): longer text is cut
 * at 99 characters (the reference writes past the field there). */
#define FLD 100
#define MAXKV 200
#define MAXLAY 20   /* MaxObjNum, src/io/iSALEInputParser.h:9 */

typedef struct { char key[FLD], val[FLD]; } kv;
typedef struct { kv e[MAXKV]; int n; } kvset;

/* rtrim drops white space; ltrim drops white space and '.'
 * (src/io/InputParser.c:7-47) */
static void trim_rl(char *s)
{
    size_t n = strlen(s);
    while (n && isspace((unsigned char)s[n - 1])) s[--n] = 0;
    size_t i = 0;
    while (s[i] && (isspace((unsigned char)s[i]) || s[i] == '.')) ++i;
    memmove(s, s + i, strlen(s + i) + 1);
}

/* Strok (src/io/InputParser.c:60-85): the first run of characters outside
 * delim, leading delimiters skipped, copied to out (cut at cap - 1).
 * Returns the run's length plus the delimiters following it (the skipped
 * leading delimiters are not counted). */
static int token(const char *s, const char *delim, char *out, int cap)
{
    int i = 0, k = 0;
    const char *p = s;
    for (; *p; ++p) {
        if (strchr(delim, *p)) {
            if (k == 0) continue;
            break;
        }
        if (k < cap - 1) out[k] = *p;
        ++k;
        ++i;
    }
    for (; *p && strchr(delim, *p); ++p) ++i;
    out[k < cap - 1 ? k : cap - 1] = 0;
    return i;
}

/* copy into a FLD field, cut at FLD - 1 characters */
static void bcopy_fld(char *dst, const char *src)
{
    size_t n = strlen(src);
    if (n > FLD - 1) n = FLD - 1;
    memcpy(dst, src, n);
    dst[n] = 0;
}

/* bounded strcat into a FLD field */
static void bcat(char *dst, const char *src)
{
    size_t n = strlen(dst);
    if (n < FLD - 1) snprintf(dst + n, FLD - n, "%s", src);
}

static int is_list_name(const char *k)
{
    static const char *L[] = {"GRIDH", "GRIDV", "GRIDSPCM", "LAYMAT", "LAYPOS", "LAYTPROF", "TR_SPCH", "TR_SPCV"};
    for (size_t i = 0; i < sizeof L / sizeof L[0]; ++i)
        if (!strcasecmp(k, L[i])) return 1;
    return 0;
}

/* stable sort by key, case insensitive (SortInputFile's merge sort) */
static void sort_set(kvset *S)
{
    for (int i = 1; i < S->n; ++i) {
        kv t = S->e[i];
        int j = i - 1;
        while (j >= 0 && strcasecmp(S->e[j].key, t.key) > 0) { S->e[j + 1] = S->e[j]; --j; }
        S->e[j + 1] = t;
    }
}

static int read_isale(const char *path, kvset *S)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return SALEC_EINVAL;
    S->n = 0;
    int is_asteroid = 0, rc = SALEC_OK;
    char *line = NULL;
    size_t cap = 0;
    while (getline(&line, &cap, fp) != -1) {
        trim_rl(line);
        if (!strcasecmp(line, "#ISINP")) { is_asteroid = 1; continue; }
        if (!line[0] || strchr("-#<", line[0])) continue;          /* comment lines */
        size_t L = strlen(line) + 1;
        char *tkey = (char *)malloc(L), *tval = (char *)malloc(L), tshort[FLD];
        if (!tkey || !tval) { free(tkey); free(tval); rc = SALEC_ENOMEM; break; }
        int r = token(line, ":", tkey, (int)L);
        token(line + r, "\"", tval, (int)L);
        if (tkey[0] && tval[0]) {
            if (S->n >= MAXKV - 1) { free(tkey); free(tval); rc = SALEC_EINVAL; break; }
            token(tkey, " ", tshort, FLD);                            /* the name before the description */
            kv *e = &S->e[S->n++];
            bcopy_fld(e->key, tshort);
            if (is_list_name(tshort) || !is_asteroid) snprintf(e->val, FLD, "[%s]", tval);   /* material.inp: all lists */
            else bcopy_fld(e->val, tval);
            trim_rl(e->key);
            trim_rl(e->val);
            int ncomp = 0;
            for (char *p = e->val; *p; ++p)
                if (*p == ':') { *p = ','; ++ncomp; }
            if (!strcasecmp("MATNAME", tshort)) {                    /* number of materials */
                kv *m = &S->e[S->n++];
                bcopy_fld(m->key, "NM");
                snprintf(m->val, FLD, "%d", ncomp + 1);
            }
        }
        free(tkey);
        free(tval);
    }
    free(line);
    fclose(fp);
    if (rc == SALEC_OK) sort_set(S);
    return rc;
}

/* SearchInput (src/io/InputParser.c:345-371): binary search over the
 * sorted keys; with repeated keys the probe order decides which is found */
static int find(const kvset *S, const char *key)
{
    if (S->n == 0) return -1;
    int il = 0, ir = S->n - 1;
    if (strcasecmp(S->e[il].key, key) > 0 || strcasecmp(S->e[ir].key, key) < 0) return -1;
    while (il + 1 < ir) {
        int mid = (il + ir) / 2, c = strcasecmp(S->e[mid].key, key);
        if (c > 0) ir = mid;
        else if (c < 0) il = mid;
        else return mid;
    }
    if (!strcasecmp(S->e[il].key, key)) return il;
    if (!strcasecmp(S->e[ir].key, key)) return ir;
    return -1;
}

/* element pos of "[a,b,...]": its length (0 when absent or empty), -1 when
 * the value is no bracketed list (GetStrk, src/io/InputParser.c:400-423) */
static int elem(const char *v, int pos, char *out)
{
    int len = (int)strlen(v), lpos = 0, k = 0;
    out[0] = 0;
    if (len == 0 || v[0] != '[' || v[len - 1] != ']') return -1;
    for (int i = 1; i < len - 1; ++i) {
        if (v[i] == ',') { ++lpos; continue; }
        if (lpos > pos) break;
        if (lpos == pos) out[k++] = v[i];
    }
    out[k] = 0;
    return k;
}

static void fortran_exp(char *s)
{
    for (; *s; ++s)
        if (*s == 'D' || *s == 'E') *s = 'e';
}

static int get_i(const kvset *S, const char *key, const char *dflt)
{
    int r = find(S, key);
    return (int)strtol(r < 0 ? dflt : S->e[r].val, NULL, 10);
}

static int get_ik(const kvset *S, const char *key, int k, const char *dflt)
{
    int r = find(S, key);
    if (r < 0) return (int)strtod(dflt, NULL);
    char t[FLD];
    return elem(S->e[r].val, k, t) > 0 ? (int)strtol(t, NULL, 10) : (int)strtol(dflt, NULL, 10);
}

static double get_df90(const kvset *S, const char *key, const char *dflt)
{
    int r = find(S, key);
    if (r < 0) return strtod(dflt, NULL);
    char t[FLD];
    bcopy_fld(t, S->e[r].val);
    fortran_exp(t);
    return strtod(t, NULL);
}

static double get_dkf90(const kvset *S, const char *key, int k, const char *dflt)
{
    int r = find(S, key);
    if (r < 0) return strtod(dflt, NULL);
    char t[FLD];
    if (elem(S->e[r].val, k, t) > 0) { fortran_exp(t); return strtod(t, NULL); }
    return strtod(dflt, NULL);
}

static void get_s(const kvset *S, const char *key, char *out, const char *dflt)
{
    int r = find(S, key);
    bcopy_fld(out, r < 0 ? dflt : S->e[r].val);
}

static void get_sk(const kvset *S, const char *key, int k, char *out, const char *dflt)
{
    int r = find(S, key);
    if (r < 0) { bcopy_fld(out, dflt); return; }
    char t[FLD];
    elem(S->e[r].val, k, t);
    trim_rl(t);
    bcopy_fld(out, t[0] ? t : dflt);
}

static int add(kvset *O, const char *key, const char *val)
{
    if (O->n >= MAXKV - 1) return SALEC_EINVAL;
    bcopy_fld(O->e[O->n].key, key);
    bcopy_fld(O->e[O->n].val, val);
    O->n++;
    return SALEC_OK;
}

/* cell widths of one direction, in units of the fine cell: ext_lo cells
 * growing by the factor f towards the low end, n fine cells, ext_hi cells
 * growing towards the high end; widths above wmax are clamped to it and
 * counted in npml.  pos[k] is the cumulative width (GenerateMesh,
 * src/io/iSALEInputParser.c:198-262). */
static int stretch(int ext_lo, int n, int ext_hi, double f, double wmax, int npml[2], double *pos)
{
    int N = n + ext_lo + ext_hi;
    double *d = (double *)malloc(sizeof(double) * (size_t)N);
    if (!d) return SALEC_ENOMEM;
    for (int k = 0; k < N; ++k) d[k] = 1.0;
    npml[0] = npml[1] = 0;
    for (int k = ext_lo - 1; k >= 0; --k) {
        d[k] = f * d[k + 1];
        if (d[k] > wmax) { d[k] = wmax; npml[0]++; }
    }
    for (int k = N - ext_hi; k < N; ++k) {
        d[k] = f * d[k - 1];
        if (d[k] > wmax) { d[k] = wmax; npml[1]++; }
    }
    pos[0] = 0.;
    for (int k = 0; k < N; ++k) pos[k + 1] = pos[k] + d[k];
    free(d);
    return SALEC_OK;
}

static int gen_mesh(const kvset *A, kvset *O)
{
    int nx = get_ik(A, "GRIDH", 1, "0"), ex0 = get_ik(A, "GRIDH", 0, "0"), ex1 = get_ik(A, "GRIDH", 2, "0");
    int ny = get_ik(A, "GRIDV", 1, "0"), ey0 = get_ik(A, "GRIDV", 0, "0"), ey1 = get_ik(A, "GRIDV", 2, "0");
    /* the growth loops read the neighbouring width: each side needs a cell
     * beside its extension zone */
    if (nx < 0 || ex0 < 0 || ex1 < 0 || ny < 0 || ey0 < 0 || ey1 < 0 || nx + ex1 < 1 || nx + ex0 < 1 ||
        ny + ey1 < 1 || ny + ey0 < 1)
        return SALEC_EINVAL;
    int Nx = nx + ex0 + ex1, Ny = ny + ey0 + ey1;
    double f = get_df90(A, "GRIDEXT", "1."), spc = get_df90(A, "GRIDSPC", "1.");
    double wm0 = get_dkf90(A, "GRIDSPCM", 0, "0.0"), wm1 = get_dkf90(A, "GRIDSPCM", 1, "0.0");
    if (wm1 >= 0. && wm1 <= 1.0e-2) wm1 = wm0;
    wm0 = wm0 > 0 ? wm0 / spc : -1.0 * wm0;      /* positive: metres; negative: cells */
    wm1 = wm1 > 0 ? wm1 / spc : -1.0 * wm1;
    int lay_num = get_i(A, "LAYNUM", "1");
    if (lay_num < 0 || lay_num > MAXLAY) return SALEC_EINVAL;
    int lay_pos[MAXLAY];
    for (int k = 0; k < lay_num; ++k) {
        lay_pos[k] = get_ik(A, "LAYPOS", k, "1");
        if (lay_pos[k] < 0 || lay_pos[k] > Ny) return SALEC_EINVAL;
    }
    double *xpos = (double *)malloc(sizeof(double) * (size_t)(Nx + 1));
    double *ypos = (double *)malloc(sizeof(double) * (size_t)(Ny + 1));
    int xpml[2], ypml[2], rc = SALEC_ENOMEM;
    if (!xpos || !ypos) goto out;
    if ((rc = stretch(ex0, nx, ex1, f, wm0, xpml, xpos)) != SALEC_OK) goto out;
    if ((rc = stretch(ey0, ny, ey1, f, wm1, ypml, ypos)) != SALEC_OK) goto out;
    double Oy = 0., Ox = 0.;
    for (int k = 0; k < lay_num; ++k) {
        double v = ypos[lay_pos[k]] / ypos[Ny];
        Oy = Oy > v ? Oy : v;
    }
    char v[FLD];
    snprintf(v, FLD, "%f", spc);
    add(O, "mesh.dx", v);
    add(O, "mesh.dy", v);
    snprintf(v, FLD, "%d", Nx); add(O, "mesh.npx", v);
    snprintf(v, FLD, "%d", Ny); add(O, "mesh.npy", v);
    snprintf(v, FLD, "%f", Ox); add(O, "mesh.Ox", v);
    snprintf(v, FLD, "%f", Oy); add(O, "mesh.Oy", v);
    add(O, "mesh.ext", "on");
    snprintf(v, FLD, "[%f,%d,%d]", f, ex0 - xpml[0], ex1 - xpml[1]); add(O, "mesh.ex", v);
    snprintf(v, FLD, "[%f,%d,%d]", f, ey0 - ypml[0], ey1 - ypml[1]); add(O, "mesh.ey", v);
    /* the reference's test counts the top layer twice and the bottom one not
     * (src/io/iSALEInputParser.c:298) */
    if (xpml[0] + xpml[1] + ypml[1] + ypml[1] > 0) {
        add(O, "numerical.damping", "ctlg");
        snprintf(v, FLD, "%d", xpml[1]); add(O, "numerical.dampX", v);
        snprintf(v, FLD, "%d", ypml[0]); add(O, "numerical.dampY", v);
        snprintf(v, FLD, "%d", xpml[0]); add(O, "numerical.dampX_LEFT", v);
        snprintf(v, FLD, "%d", ypml[1]); add(O, "numerical.dampY_TOP", v);
    } else {
        add(O, "numerical.damping", "off");
    }
    snprintf(v, FLD, "%d", lay_num);
    add(O, "target.number", v);
    /* layer thicknesses in fine-cell units; every layer but the last is
     * measured from its own top (zero), as the reference computes it */
    bcopy_fld(v, "[");
    for (int k = 0; k < lay_num; ++k) {
        double top = ypos[lay_pos[k]], bottom = 0.;
        if (k + 1 < lay_num) bottom = ypos[lay_pos[k]];
        char t[FLD];
        snprintf(t, FLD, "%.8e,", top - bottom);
        bcat(v, t);
    }
    v[strlen(v) - 1] = ']';
    add(O, "target.depth", v);
    rc = SALEC_OK;
out:
    free(xpos);
    free(ypos);
    return rc;
}

/* ConvertInputVAR: "[v1, v2]" of material.inp becomes "[, v1, v2]" (the
 * vacuum slot first) with Fortran exponents D/E as e */
static void convert_var(const kvset *M, kvset *O, const char *name, const char *key)
{
    char t[FLD], v[FLD], w[2 * FLD];
    get_s(M, name, t, "[ ]");
    snprintf(w, sizeof w, "[,%s]", t + 1);
    bcopy_fld(v, w);                       /* cut at 99 characters, then drop the last one */
    v[strlen(v) - 1] = 0;
    fortran_exp(v);
    add(O, key, v);
}

static void gen_material(const kvset *M, kvset *O)
{
    char t[FLD], v[FLD];
    int nm = get_i(M, "NM", "1");
    snprintf(v, FLD, "%d", nm + 1);
    add(O, "material.nm", v);
    get_s(M, "MATNAME", t, "dunite_");
    char w[2 * FLD];
    snprintf(w, sizeof w, "[vacuum_,%s]", t + 1);
    bcopy_fld(v, w);
    v[strlen(v) - 1] = 0;
    add(O, "material.name", v);
    bcopy_fld(v, "[null   ");
    for (int k = 0; k < nm; ++k) {
        bcat(v, ",");
        get_sk(M, "EOSTYPE", k, t, "unknown");
        bcat(v, !strcasecmp(t, "aneos") ? "aneos" : (!strcasecmp(t, "tillo") ? "tillotson" : "unknown"));
    }
    bcat(v, "]");
    add(O, "material.postfix", v);
    static const char *map[][2] = {
        {"POIS", "material.poisson"}, {"TMELT0", "material.SimonT0"}, {"TFRAC", "material.OhnakaXi"},
        {"ASIMON", "material.SimonA"}, {"CSIMON", "material.SimonC"}, {"YDAM0", "material.ydam0"},
        {"FRICDAM", "material.ydamfri"}, {"YLIMDAM", "material.ydamlim"}, {"YINT0", "material.yint0"},
        {"FRICINT", "material.yintfri"}, {"YLIMINT", "material.yintlim"}, {"JC_A", "material.JcA"},
        {"JC_B", "material.JcB"}, {"JC_N", "material.JcN"}, {"JC_C", "material.JcC"}, {"JC_M", "material.JcM"},
        {"JC_TREF", "material.JcTref"}, {"IVANOV_A", "material.IvanA"}, {"IVANOV_B", "material.IvanB"},
        {"IVANOV_C", "material.IvanC"}, {"GAMETA", "material.GammaEta"}, {"GAMBETA", "material.GammaBeta"}};
    for (size_t i = 0; i < sizeof map / sizeof map[0]; ++i) convert_var(M, O, map[i][0], map[i][1]);
}

int salec_isale_to_inp(const char *asteroid_inp, const char *material_inp, FILE *out)
{
    if (!asteroid_inp || !material_inp) return SALEC_EINVAL;
    if (!out) out = stdout;
    kvset *A = (kvset *)malloc(sizeof(kvset)), *M = (kvset *)malloc(sizeof(kvset)),
          *O = (kvset *)calloc(1, sizeof(kvset));
    int rc = SALEC_ENOMEM;
    if (!A || !M || !O) goto done;
    if ((rc = read_isale(asteroid_inp, A)) != SALEC_OK) goto done;
    if ((rc = gen_mesh(A, O)) != SALEC_OK) goto done;
    if ((rc = read_isale(material_inp, M)) != SALEC_OK) goto done;
    gen_material(M, O);
    fprintf(out, "####Read asteroid.inp&material.inp from iSALE2D:\n");
    fprintf(out, "####Following LINES can be put int to *.inp for salec2d\n");
    fprintf(out, "#### %d variables IN InputFile ####\n", O->n);
    for (int k = 0; k < O->n; ++k) fprintf(out, "%-25s = %s\n", O->e[k].key, O->e[k].val);
    fprintf(out, "#### End of InputFile ####\n");
    fflush(out);
    rc = SALEC_OK;
done:
    free(A);
    free(M);
    free(O);
    return rc;
}

/* ------------------------------------------------------------- .vts packs */

static int ensure_dir(const char *d)
{
    if (mkdir(d, 0755) == 0 || errno == EEXIST) return SALEC_OK;
    return SALEC_EINVAL;
}

static void vts_name(char *buf, size_t n, const char *dir, const char *prefix, int rank, int cycle)
{
    snprintf(buf, n, "%s/%s.proc%04d.%04d.vts", dir, prefix, rank, cycle);
}

static int file_exists(const char *p)
{
    struct stat st;
    return stat(p, &st) == 0 && S_ISREG(st.st_mode);
}

/* the cycles present from min_cycle on, up to the first one missing a domain */
static int last_cycle(const char *vts_dir, const char *prefix, int nproc, int min_cycle)
{
    char name[4096];
    int c = min_cycle;
    for (;; ++c) {
        for (int r = 0; r < nproc; ++r) {
            vts_name(name, sizeof name, vts_dir, prefix, r, c);
            if (!file_exists(name)) return c;
        }
    }
}

/* add_file: "#BEGINFILE:<name>", the file's bytes, "#ENDFILE" */
static int add_file(FILE *dest, const char *name)
{
    FILE *fp = fopen(name, "rb");
    if (!fp) return SALEC_EINVAL;
    fprintf(dest, "#BEGINFILE:%s\n", name);
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, fp)) > 0)
        if (fwrite(buf, 1, n, dest) != n) { fclose(fp); return SALEC_EINVAL; }
    fclose(fp);
    fputs("#ENDFILE\n", dest);
    return SALEC_OK;
}

int salec_pack_vts(const char *vts_dir, const char *prefix, int nproc, int min_cycle, int max_cycle, int nthreads,
                   const char *out_dir)
{
    if (!vts_dir || !prefix || !out_dir || nproc < 1 || min_cycle < 0 || nthreads < 1) return SALEC_EINVAL;
    if (max_cycle < 0) max_cycle = last_cycle(vts_dir, prefix, nproc, min_cycle);
    if (ensure_dir(out_dir) != SALEC_OK) return SALEC_EINVAL;
    int per = (max_cycle - min_cycle) / nthreads + 1, rc = SALEC_OK;
    char name[4096];
    for (int t = 0; t < nthreads && rc == SALEC_OK; ++t) {
        snprintf(name, sizeof name, "%s/%s.vts.pack%04d.%04d", out_dir, prefix, t, 0);
        FILE *pk = fopen(name, "wb");
        if (!pk) return SALEC_EINVAL;
        for (int c = min_cycle + t * per; c < min_cycle + (t + 1) * per && c < max_cycle && rc == SALEC_OK; ++c)
            for (int r = 0; r < nproc && rc == SALEC_OK; ++r) {
                vts_name(name, sizeof name, vts_dir, prefix, r, c);
                rc = add_file(pk, name);
            }
        if (fclose(pk) != 0) rc = SALEC_EINVAL;
    }
    return rc;
}

/* create the missing directories on the way to a file */
static void make_parents(const char *path)
{
    char buf[4096];
    snprintf(buf, sizeof buf, "%s", path);
    for (char *p = buf + 1; *p; ++p)
        if (*p == '/') {
            *p = 0;
            mkdir(buf, 0755);
            *p = '/';
        }
}

int salec_unpack_vts(const char *pack, const char *dir, int *nfiles)
{
    if (!pack) return SALEC_EINVAL;
    FILE *fp = fopen(pack, "rb");
    if (!fp) return SALEC_EINVAL;
    char *line = NULL;
    size_t cap = 0;
    ssize_t len;
    FILE *dest = NULL;
    int count = 0, rc = SALEC_OK;
    while ((len = getline(&line, &cap, fp)) != -1) {
        if (!dest && len >= 11 && !strncasecmp(line, "#BEGINFILE:", 11)) {
            if (line[len - 1] == '\n') line[--len] = 0;
            char path[4096];
            if (dir) snprintf(path, sizeof path, "%s/%s", dir, line + 11);
            else snprintf(path, sizeof path, "%s", line + 11);
            make_parents(path);
            dest = fopen(path, "wb");
            if (!dest) { rc = SALEC_EINVAL; break; }
            continue;
        }
        if (dest && !strncasecmp(line, "#ENDFILE", 8)) {
            if (fclose(dest) != 0) rc = SALEC_EINVAL;
            dest = NULL;
            ++count;
            continue;
        }
        if (dest && fwrite(line, 1, (size_t)len, dest) != (size_t)len) { rc = SALEC_EINVAL; break; }
    }
    if (dest) { fclose(dest); rc = SALEC_EINVAL; }   /* a section without its end marker */
    free(line);
    fclose(fp);
    if (nfiles) *nfiles = count;
    return rc;
}

/* copy_vtk_file: the lines strictly between the first line holding
 * "VTKFile" within its first 10 characters and the next such line */
static int copy_inner(FILE *src, FILE *dest)
{
    char *line = NULL;
    size_t cap = 0;
    ssize_t len;
    int found = 0;
    while ((len = getline(&line, &cap, src)) != -1) {
        char head[11];
        snprintf(head, sizeof head, "%s", line);
        if (strstr(head, "VTKFile")) {
            if (found) break;
            found = 1;
            continue;
        }
        if (found) fwrite(line, 1, (size_t)len, dest);
    }
    free(line);
    return found ? SALEC_OK : SALEC_EINVAL;
}

int salec_merge_vts_blocks(const char *vts_dir, const char *prefix, int nproc, int min_cycle, int max_cycle,
                           const char *out_dir)
{
    if (!vts_dir || !prefix || !out_dir || nproc < 1 || min_cycle < 0) return SALEC_EINVAL;
    if (max_cycle < 0) max_cycle = last_cycle(vts_dir, prefix, nproc, min_cycle);
    if (ensure_dir(out_dir) != SALEC_OK) return SALEC_EINVAL;
    char name[4096];
    for (int c = min_cycle; c < max_cycle; ++c) {
        snprintf(name, sizeof name, "%s/%s.%04d.vts", out_dir, prefix, c);
        FILE *m = fopen(name, "wb");
        if (!m) return SALEC_EINVAL;
        fputs("<?xml version=\"1.0\"?>\n", m);
        fputs("<VTKFile type=\"vtkMultiBlockDataSet\" version=\"0.1\" compressor=\"vtkZLibDataCompressor\" "
              "byte_order=\"LittleEndian\">\n", m);
        fputs("  <vtkMultiBlockDataSet>\n", m);
        int rc = SALEC_OK;
        for (int r = 0; r < nproc && rc == SALEC_OK; ++r) {
            vts_name(name, sizeof name, vts_dir, prefix, r, c);
            FILE *v = fopen(name, "rb");
            if (!v) { rc = SALEC_EINVAL; break; }
            fprintf(m, "    <DataSet index=\"%d\">\n", r);
            rc = copy_inner(v, m);
            fputs("    </DataSet>\n", m);
            fclose(v);
        }
        fputs("  </vtkMultiBlockDataSet>\n", m);
        fputs("</VTKFile>\n", m);
        if (fclose(m) != 0) rc = SALEC_EINVAL;
        if (rc != SALEC_OK) return rc;
    }
    return SALEC_OK;
}
