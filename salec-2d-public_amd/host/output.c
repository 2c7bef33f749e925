/*
 * output.c -- VTK XML writers for the solver state and the tracers
 * (include/salec_output.h).  The file structure, array order, names and
 * encodings are those of the reference's writers (src/write_vtk.c,
 * src/io/output.c), so the files are byte-for-byte what the reference
 * writes for the same state on one domain.
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include "../../include/salec_output.h"

/* ---------------------------------------------------------- encoders -- */
static const char B64[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

/* standard base64 with '=' padding, written straight to fp */
static void put_b64(FILE *fp, const unsigned char *p, size_t n)
{
    char q[4];
    for (size_t i = 0; i < n; i += 3) {
        size_t r = n - i;
        unsigned v = (unsigned)p[i] << 16;
        if (r > 1) v |= (unsigned)p[i + 1] << 8;
        if (r > 2) v |= p[i + 2];
        q[0] = B64[(v >> 18) & 63];
        q[1] = B64[(v >> 12) & 63];
        q[2] = r > 1 ? B64[(v >> 6) & 63] : '=';
        q[3] = r > 2 ? B64[v & 63] : '=';
        fwrite(q, 1, 4, fp);
    }
}

static void put_le32(unsigned char *b, unsigned v)
{
    b[0] = v & 255; b[1] = (v >> 8) & 255; b[2] = (v >> 16) & 255; b[3] = (v >> 24) & 255;
}

/* one Float32 array in the "binary" encoding of vtkZLibDataCompressor:
 * base64(header) base64(zlib(raw)) newline, header = {1, raw, raw, packed} */
static int put_binary(FILE *fp, const float *a, size_t n)
{
    size_t raw = 4 * n;
    unsigned char *in = malloc(raw ? raw : 1);
    if (!in) return SALEC_EINVAL;
    for (size_t i = 0; i < n; ++i) {
        unsigned u;
        memcpy(&u, a + i, 4);
        put_le32(in + 4 * i, u);
    }
    z_stream z;
    memset(&z, 0, sizeof z);
    if (deflateInit(&z, Z_DEFAULT_COMPRESSION) != Z_OK) { free(in); return SALEC_EINVAL; }
    uLong cap = deflateBound(&z, (uLong)raw);
    unsigned char *out = malloc(cap ? cap : 1);
    if (!out) { deflateEnd(&z); free(in); return SALEC_EINVAL; }
    z.next_in = in; z.avail_in = (uInt)raw;
    z.next_out = out; z.avail_out = (uInt)cap;
    int rc = deflate(&z, Z_FINISH);
    size_t packed = cap - z.avail_out;
    deflateEnd(&z);
    free(in);
    if (rc != Z_STREAM_END) { free(out); return SALEC_EINVAL; }
    unsigned char head[16];
    put_le32(head, 1u); put_le32(head + 4, (unsigned)raw); put_le32(head + 8, (unsigned)raw);
    put_le32(head + 12, (unsigned)packed);
    put_b64(fp, head, 16);
    put_b64(fp, out, packed);
    fputc('\n', fp);
    free(out);
    return SALEC_OK;
}

/* the ascii encoding: rows of ncomp values (1, 2 or 3 per row with %.4e,
 * 6 with %.6e), the only row shapes the reference writes */
static int put_ascii(FILE *fp, const float *a, size_t n, int ncomp)
{
    const char *f1 = ncomp == 6 ? "%.6e" : "%.4e";
    if (ncomp != 1 && ncomp != 2 && ncomp != 3 && ncomp != 6) return SALEC_EINVAL;
    for (size_t i = 0; i + ncomp <= n; i += ncomp)
        for (int c = 0; c < ncomp; ++c) {
            fprintf(fp, f1, (double)a[i + c]);
            fputc(c + 1 < ncomp ? ' ' : '\n', fp);
        }
    return SALEC_OK;
}

/* <DataArray> of npts points with ncomp Float32 components each */
static int put_array(FILE *fp, const char *name, const char *format, const float *a, long npts, int ncomp)
{
    int binary = !strcmp(format, "binary");
    fprintf(fp, "        <DataArray type=\"Float32\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"%s\">\n", name, ncomp,
            format);
    int rc = binary ? put_binary(fp, a, (size_t)npts * ncomp) : put_ascii(fp, a, (size_t)npts * ncomp, ncomp);
    fputs("        </DataArray>\n", fp);
    return rc;
}

/* a field's planes as per-point Float32 records (AoS) */
static float *aos(const double *f, long npts, int ncomp)
{
    float *a = malloc(sizeof(float) * (size_t)(npts * ncomp > 0 ? npts * ncomp : 1));
    if (!a) return NULL;
    for (long i = 0; i < npts; ++i)
        for (int c = 0; c < ncomp; ++c) a[i * ncomp + c] = (float)f[(long)c * npts + i];
    return a;
}

static int put_field(FILE *fp, const char *name, const double *const *fields, int field, int nx, int ny, int nmat)
{
    int grid, nc;
    if (salec_field_info(field, nmat, &grid, &nc) || !fields[field]) return SALEC_EINVAL;
    long npts = salec_grid_points(grid, nx, ny);
    float *a = aos(fields[field], npts, nc);
    if (!a) return SALEC_EINVAL;
    int rc = put_array(fp, name, "binary", a, npts, nc);
    free(a);
    return rc;
}

/* arrays of a state file, src/write_vtk.c:242-283 */
struct named { const char *name; int field; };
static const struct named kPoint[] = {
    {"v_vel", SALEC_V_VEL}, {"v_acc", SALEC_V_ACC}, {"v_gra", SALEC_V_BF}, {"v_vof", SALEC_V_VOF}};
static const struct named kCell[] = {
    {"e_den", SALEC_E_DEN}, {"e_pre", SALEC_E_PRE}, {"e_wpt", SALEC_E_WPT}, {"e_vof", SALEC_M_VOF},
    {"e_csd", SALEC_E_CSD}, {"e_dam", SALEC_E_DAM}, {"e_tps", SALEC_E_TPS}, {"e_q", SALEC_E_Q},
    {"e_tem", SALEC_E_TEM}, {"e_eng", SALEC_M_ENG}, {"m_den", SALEC_M_DEN}, {"e_ste2s", SALEC_E_STE2S},
    {"e_sta2s", SALEC_E_STA2S}, {"e_ste", SALEC_E_STE}, {"e_strength", SALEC_E_STRENGTH}, {"e_vib", SALEC_E_VIB},
    {"e_ejt", SALEC_E_EJT}, {"e_cvs", SALEC_E_CVS}, {"m_pty", SALEC_M_PTY}};
/* the parallel header lists the cell arrays up to e_ejt (src/write_vtk.c:98-114) */
#define N_PCELL 17

/* ------------------------------------------------------------- state -- */
int salec_write_vts(const char *path, int nx, int ny, int nmat, const double *const *fields, const char *coord_format)
{
    if (!path || !fields || nx < 1 || ny < 1 || nmat < 1 || nmat > SALEC_MAXMAT) return SALEC_EINVAL;
    const char *cf = coord_format ? coord_format : "binary";
    if (strcmp(cf, "binary") && strcmp(cf, "ascii")) return SALEC_EINVAL;
    if (!fields[SALEC_V_POS]) return SALEC_EINVAL;
    FILE *fp = fopen(path, "w");
    if (!fp) return SALEC_EINVAL;
    const int VW = nx + 5, VH = ny + 5;
    int rc = SALEC_OK;
    /* piece: the whole padded vertex grid; whole extent as the reference
     * computes it on a single domain with ghost output (x and y end at
     * n + 2) */
    fprintf(fp, "<?xml version=\"1.0\"?>\n"
                "<VTKFile type=\"StructuredGrid\" version=\"0.1\" compressor=\"vtkZLibDataCompressor\" "
                "byte_order=\"LittleEndian\">\n"
                "  <StructuredGrid WholeExtent=\"%d %d %d %d 0 0\">\n"
                "    <Piece Extent=\"%d %d %d %d 0 0\">\n",
            0, nx + 2, 0, ny + 2, 0, VW - 1, 0, VH - 1);
    fputs("      <PointData>\n", fp);
    for (size_t i = 0; i < sizeof kPoint / sizeof kPoint[0] && rc == SALEC_OK; ++i)
        rc = put_field(fp, kPoint[i].name, fields, kPoint[i].field, nx, ny, nmat);
    fputs("      </PointData>\n", fp);
    fputs("      <CellData>\n", fp);
    for (size_t i = 0; i < sizeof kCell / sizeof kCell[0] && rc == SALEC_OK; ++i)
        rc = put_field(fp, kCell[i].name, fields, kCell[i].field, nx, ny, nmat);
    fputs("      </CellData>\n", fp);
    if (rc == SALEC_OK) {
        long nv = (long)VW * VH;
        float *p = malloc(sizeof(float) * 3 * nv);
        if (!p) rc = SALEC_EINVAL;
        else {
            const double *vp = fields[SALEC_V_POS];
            for (long i = 0; i < nv; ++i) { p[3 * i] = (float)vp[i]; p[3 * i + 1] = (float)vp[nv + i]; p[3 * i + 2] = 0.0f; }
            fputs("      <Points>\n", fp);
            rc = put_array(fp, "coordinate", cf, p, nv, 3);
            fputs("      </Points>\n", fp);
            free(p);
        }
    }
    fputs("    </Piece>\n  </StructuredGrid>\n</VTKFile>\n", fp);
    if (fclose(fp) != 0 && rc == SALEC_OK) rc = SALEC_EINVAL;
    return rc;
}

static void pdesc(FILE *fp, const char *name, int nc)
{
    fprintf(fp, "       <PDataArray type=\"Float32\" Name=\"%s\" NumberOfComponents=\"%d\"/>\n", name, nc);
}

int salec_write_pvts(const char *path, const char *prefix, int cycle, int nx, int ny, int nmat)
{
    if (!path || !prefix || nx < 1 || ny < 1 || nmat < 1 || nmat > SALEC_MAXMAT) return SALEC_EINVAL;
    FILE *fp = fopen(path, "w");
    if (!fp) return SALEC_EINVAL;
    fputs("<?xml version=\"1.0\"?>\n"
          "<VTKFile type=\"PStructuredGrid\" version=\"1.0\" compressor=\"vtkZLibDataCompressor\" "
          "byte_order=\"LittleEndian\">\n", fp);
    /* WholeExtent of the ghost-output build: x from the pad to n + 2 pads,
     * y from 0 to n + 1 pad (src/write_vtk.c:74-76) */
    fprintf(fp, "<PStructuredGrid WholeExtent=\"%d %d %d %d 0 0\" GhostLevel=\"%d\">\n", 2, nx + 4, 0, ny + 2, 0);
    fputs("      <PPointData>\n", fp);
    for (size_t i = 0; i < sizeof kPoint / sizeof kPoint[0]; ++i) {
        int grid, nc;
        salec_field_info(kPoint[i].field, nmat, &grid, &nc);
        pdesc(fp, kPoint[i].name, nc);
    }
    fputs("      </PPointData>\n", fp);
    fputs("      <PCellData>\n", fp);
    for (int i = 0; i < N_PCELL; ++i) {
        int grid, nc;
        salec_field_info(kCell[i].field, nmat, &grid, &nc);
        pdesc(fp, kCell[i].name, nc);
    }
    fputs("      </PCellData>\n", fp);
    fputs("    <PPoints>\n"
          "      <PDataArray type=\"Float32\" Name=\"coordinate\" NumberOfComponents=\"3\"/>\n"
          "    </PPoints>\n", fp);
    fprintf(fp, "      <Piece Extent=\"%d %d %d %d 0 0\" Source=\"./vts/%s.proc%04d.%04d.vts\"/>\n", 0, nx + 2, 0, ny + 2,
            prefix, 0, cycle);
    fputs("</PStructuredGrid>\n</VTKFile>", fp);
    return fclose(fp) == 0 ? SALEC_OK : SALEC_EINVAL;
}

static int write_vtm_one(const char *path, const char *prefix, int cycle, int nproc, int tracer)
{
    FILE *fp = fopen(path, "w");
    if (!fp) return SALEC_EINVAL;
    fputs("<?xml version=\"1.0\"?>\n"
          "<VTKFile type=\"vtkMultiBlockDataSet\" version=\"1.0\" compressor=\"vtkZLibDataCompressor\" "
          "byte_order=\"LittleEndian\">\n"
          "  <vtkMultiBlockDataSet>\n", fp);
    for (int k = 0; k < nproc; ++k) {
        if (tracer)
            fprintf(fp, "    <DataSet index=\"%d\" file=\"./vtp/%s.tracer.proc%04d.%04d.vtp\"/>\n", k, prefix, k, cycle);
        else
            fprintf(fp, "    <DataSet index=\"%d\" file=\"./vts/%s.proc%04d.%04d.vts\"/>\n", k, prefix, k, cycle);
    }
    fputs("  </vtkMultiBlockDataSet>\n</VTKFile>", fp);
    return fclose(fp) == 0 ? SALEC_OK : SALEC_EINVAL;
}

int salec_write_vtm(const char *dir, const char *prefix, int cycle, int nproc)
{
    if (!dir || !prefix || nproc < 1) return SALEC_EINVAL;
    size_t need = strlen(dir) + strlen(prefix) + 40;
    char *p = malloc(need);
    if (!p) return SALEC_EINVAL;
    snprintf(p, need, "%s/%s.%04d.vtm", dir, prefix, cycle);
    int rc = write_vtm_one(p, prefix, cycle, nproc, 0);
    if (rc == SALEC_OK) {
        snprintf(p, need, "%s/%s.tracer.%04d.vtm", dir, prefix, cycle);
        rc = write_vtm_one(p, prefix, cycle, nproc, 1);
    }
    free(p);
    return rc;
}

/* ----------------------------------------------------------- tracers -- */
int salec_write_vtp(const char *path, const char *txt_path, int nx, int ny, const salec_tracers *tr,
                    const double *e_den, const double *e_tem, const double *e_pre)
{
    if (!path || !tr || !e_den || !e_tem || !e_pre || nx < 1 || ny < 1 || tr->n < 0) return SALEC_EINVAL;
    const long CW = nx + 4, CH = ny + 4, n = tr->n;
    long live = 0;
    for (long i = 0; i < n; ++i) live += tr->alive[i] != 0;
    const int dummy = live == 0;
    const long m = dummy ? 1 : live;
    float *buf = calloc((size_t)m * 15, sizeof(float));
    if (!buf) return SALEC_EINVAL;
    float *pos = buf, *mat = buf + 3 * m, *tag = mat + m, *gx = tag + m, *gy = gx + m, *vel = gy + m;
    float *den = vel + 2 * m, *tem = den + m, *pre = tem + m, *mtem = pre + m, *mpre = mtem + m;
    if (dummy) {
        mat[0] = tag[0] = gx[0] = gy[0] = -1.0f;
    } else {
        long q = 0;
        for (long i = 0; i < n; ++i) {
            if (!tr->alive[i]) continue;
            long k = tr->idx[i], j = tr->idy[i];
            if (k < 0 || k >= CW || j < 0 || j >= CH) { free(buf); return SALEC_EINVAL; }
            long c = k + CW * j;
            pos[3 * q] = (float)tr->pos[i];
            pos[3 * q + 1] = (float)tr->pos[n + i];
            pos[3 * q + 2] = 0.0f;
            mat[q] = 1.0f * tr->matid[i];
            tag[q] = 1.0f * tr->tag[i];
            gx[q] = 1.0f * tr->gx[i];
            gy[q] = 1.0f * tr->gy[i];
            vel[2 * q] = (float)tr->vel[i];
            vel[2 * q + 1] = (float)tr->vel[n + i];
            mtem[q] = tr->mtem[i];
            mpre[q] = tr->mpre[i];
            den[q] = (float)e_den[c];
            tem[q] = (float)e_tem[c];
            pre[q] = (float)e_pre[c];
            ++q;
        }
    }
    FILE *fp = fopen(path, "w");
    if (!fp) { free(buf); return SALEC_EINVAL; }
    fprintf(fp, "<?xml version=\"1.0\"?>\n"
                "<VTKFile type=\"PolyData\" version=\"1.0\" compressor=\"vtkZLibDataCompressor\" "
                "byte_order=\"LittleEndian\">\n"
                "  <PolyData>\n"
                "    <Piece NumberOfPoints=\"%ld\" NumberOfVerts=\"0\" NumberOfLines=\"0\" NumberOfStrips=\"0\" "
                "NumberOfPolys=\"0\">\n", m);
    fputs("      <PointData>\n", fp);
    struct { const char *name; const float *a; int nc; } arr[] = {
        {"matid", mat, 1}, {"tag", tag, 1}, {"gx", gx, 1}, {"gy", gy, 1}, {"vel", vel, 2},
        {"den", den, 1}, {"tem", tem, 1}, {"pre", pre, 1}, {"mpre", mpre, 1}, {"mtem", mtem, 1}};
    int rc = SALEC_OK;
    for (size_t i = 0; i < sizeof arr / sizeof arr[0] && rc == SALEC_OK; ++i)
        rc = put_array(fp, arr[i].name, "binary", arr[i].a, m, arr[i].nc);
    fputs("      </PointData>\n", fp);
    fputs("      <Points>\n", fp);
    if (rc == SALEC_OK) rc = put_array(fp, "coordinate", "binary", pos, m, 3);
    fputs("      </Points>\n", fp);
    fputs("    </Piece>\n  </PolyData>\n</VTKFile>\n", fp);
    if (fclose(fp) != 0 && rc == SALEC_OK) rc = SALEC_EINVAL;
    if (rc == SALEC_OK && txt_path) {
        FILE *tp = fopen(txt_path, "w");
        if (!tp) rc = SALEC_EINVAL;
        else {
            fprintf(tp, "   %ld\n", m - dummy);
            for (long q = 0; q < m; ++q)
                fprintf(tp, "%10f, %5f, %5f, %5f, %10.5e, %10.5e, %10.5e, %10.5e, %10.5e, %10.5e, %10.5e, %10.5e\n",
                        (double)tag[q], (double)mat[q], (double)gx[q], (double)gy[q], (double)pos[3 * q],
                        (double)pos[3 * q + 1], (double)vel[2 * q], (double)vel[2 * q + 1], (double)mpre[q],
                        (double)mtem[q], (double)pre[q], (double)tem[q]);
            if (fclose(tp) != 0) rc = SALEC_EINVAL;
        }
    }
    free(buf);
    return rc;
}

int salec_write_pvtp(const char *path, const char *prefix, int cycle, int nproc)
{
    if (!path || !prefix || nproc < 1) return SALEC_EINVAL;
    FILE *fp = fopen(path, "w");
    if (!fp) return SALEC_EINVAL;
    fputs("<?xml version=\"1.0\"?>\n<VTKFile type=\"PPolyData\" version=\"1.0\" byte_order=\"LittleEndian\">\n", fp);
    fprintf(fp, "<PPolyData GhostLevel=\"%d\">\n", 0);
    fputs("      <PPointData>\n", fp);
    const char *names[] = {"matid", "tag", "gx", "gy", "vel", "den", "tem", "pre", "mpre", "mtem"};
    for (int i = 0; i < 10; ++i) pdesc(fp, names[i], i == 4 ? 2 : 1);
    fputs("      </PPointData>\n", fp);
    fputs("    <PPoints>\n"
          "      <PDataArray type=\"Float32\" Name=\"coordinate\" NumberOfComponents=\"3\"/>\n"
          "    </PPoints>\n", fp);
    for (int k = 0; k < nproc; ++k)
        fprintf(fp, "      <Piece Source=\"./vtp/%s.tracer.proc%04d.%04d.vtp\"/>\n", prefix, k, cycle);
    fputs("</PPolyData>\n</VTKFile>", fp);
    return fclose(fp) == 0 ? SALEC_OK : SALEC_EINVAL;
}

int salec_write_mass_tracers(const char *path, int nx, int ny, const salec_tracers *tr, const double *e_den,
                             const double *e_vol)
{
    if (!path || !tr || !e_den || !e_vol || nx < 1 || ny < 1 || tr->n < 0) return SALEC_EINVAL;
    const long CW = nx + 4, CH = ny + 4, n = tr->n;
    long live = 0;
    for (long i = 0; i < n; ++i) live += tr->alive[i] != 0;
    FILE *fp = fopen(path, "w");
    if (!fp) return SALEC_EINVAL;
    fprintf(fp, "%ld\n", live);
    int rc = SALEC_OK;
    for (long i = 0; i < n; ++i) {
        if (!tr->alive[i]) continue;
        long k = tr->idx[i], j = tr->idy[i];
        if (k < 0 || k >= CW || j < 0 || j >= CH) { rc = SALEC_EINVAL; break; }
        long c = k + CW * j;
        fprintf(fp, "%d, %d, %d, %10.5e,%10.5e, %10.5e, %10.5e\n", tr->tag[i], tr->gx[i], tr->gy[i], tr->pos[i],
                tr->pos[n + i], e_den[c], e_vol[c]);
    }
    if (fclose(fp) != 0 && rc == SALEC_OK) rc = SALEC_EINVAL;
    return rc;
}
