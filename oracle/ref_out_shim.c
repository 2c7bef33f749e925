/*
 * ref_out_shim.c -- TEST INFRASTRUCTURE ONLY.  Flat entry points around the
 * reference's own VTK array encoders (src/io/output.c, compiled where it
 * lies): they call those encoders in the order the reference's writers do
 * (src/write_vtk.c:190-300 for a state file, :343-529 for a tracer file),
 * from data laid out as the reference stores it (per node, components
 * together), so the files can be compared with the build's writers byte for
 * byte.  Used by tests only.
 */
#include <stdio.h>
#include <stdlib.h>
#include "output.h"

/* one state file: npt point arrays then ncl cell arrays (AoS doubles with
 * the given component counts), then coordinates (2 per vertex) */
int ref_out_vts(const char *path, const char *pextent, const char *wextent, int npt, const char **pnames,
                const double **parr, const int *pnc, long nv, int ncl, const char **cnames, const double **carr,
                const int *cnc, long nc, const double *coord, const char *coord_format)
{
    FILE *fp = fopen(path, "w");
    if (!fp) return -1;
    vts_file_header(fp, pextent, wextent);
    vtk_point_data_header(fp);
    for (int i = 0; i < npt; ++i) vtk_dataarray_vec(fp, pnames[i], "binary", parr[i], (int)nv, pnc[i]);
    vtk_point_data_trailer(fp);
    vtk_cell_data_header(fp);
    for (int i = 0; i < ncl; ++i) vtk_dataarray_vec(fp, cnames[i], "binary", carr[i], (int)nc, cnc[i]);
    vtk_cell_data_trailer(fp);
    vtk_output_coord(fp, coord_format, coord, (int)nv);
    vts_file_trailer(fp);
    fclose(fp);
    return 0;
}

/* one tracer file from the per-tracer float arrays (dummy point included
 * by the caller when there is no live tracer) */
int ref_out_vtp(const char *path, int n, const float *mat, const float *tag, const float *gx, const float *gy,
                const float *vel, const float *den, const float *tem, const float *pre, const float *mpre,
                const float *mtem, const float *pos)
{
    FILE *fp = fopen(path, "w");
    if (!fp) return -1;
    vtp_file_header(fp, n);
    vtk_point_data_header(fp);
    vtk_dataarray_vec_f(fp, "matid", "binary", mat, n, 1);
    vtk_dataarray_vec_f(fp, "tag", "binary", tag, n, 1);
    vtk_dataarray_vec_f(fp, "gx", "binary", gx, n, 1);
    vtk_dataarray_vec_f(fp, "gy", "binary", gy, n, 1);
    vtk_dataarray_vec_f(fp, "vel", "binary", vel, n, 2);
    vtk_dataarray_vec_f(fp, "den", "binary", den, n, 1);
    vtk_dataarray_vec_f(fp, "tem", "binary", tem, n, 1);
    vtk_dataarray_vec_f(fp, "pre", "binary", pre, n, 1);
    vtk_dataarray_vec_f(fp, "mpre", "binary", mpre, n, 1);
    vtk_dataarray_vec_f(fp, "mtem", "binary", mtem, n, 1);
    vtk_point_data_trailer(fp);
    vtk_point_header(fp);
    vtk_dataarray_vec_f(fp, "coordinate", "binary", pos, n, 3);
    vtk_point_trailer(fp);
    vtp_file_trailer(fp);
    fclose(fp);
    return 0;
}
