/*
 * salec_output.h -- output writers of the step path's host side: VTK XML
 * files of the solver state and the tracers, in the reference's layout and
 * encoding, so that existing ParaView workflows read them unchanged.
 *
 * Arrays are written as the reference writes them: per node, all components
 * together (AoS), over the whole padded grid, converted to Float32; binary
 * arrays are zlib-compressed (default level) and base64-encoded after a
 * base64 header of four int32 {1, raw bytes, raw bytes, compressed bytes}.
 *
 * Host fields are the planes of include/salec.h (f[c*N + idx], x fastest),
 * indexed by field id; a writer reads only the fields it lists.
 * All functions return SALEC_OK or SALEC_EINVAL (bad argument, cannot open
 * the file, compression failure).
 */
#ifndef SALEC_OUTPUT_H
#define SALEC_OUTPUT_H
#include "salec.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One state file (.vts): point data v_vel, v_acc, v_gra, v_vof; cell data
 * e_den ... m_pty; coordinates from v_pos in coord_format ("binary" or
 * "ascii").  Replaces sale2d_write_vtk (src/write_vtk.c:190-300) for one
 * domain, built on the array encoders of src/io/output.c:1-482. */
int salec_write_vts(const char *path, int nx, int ny, int nmat, const double *const *fields, const char *coord_format);

/* The parallel header of one output (.pvts) for a single domain, as
 * sale2d_write_pvtk (src/write_vtk.c:61-187) writes it. */
int salec_write_pvts(const char *path, const char *prefix, int cycle, int nx, int ny, int nmat);

/* The two multiblock indexes <prefix>.%04d.vtm and <prefix>.tracer.%04d.vtm
 * in directory dir, as sale2d_write_vtm (src/write_vtk.c:26-59) writes them
 * for nproc domains. */
int salec_write_vtm(const char *dir, const char *prefix, int cycle, int nproc);

/* Tracer file (.vtp): matid, tag, gx, gy, vel, den, tem, pre, mpre, mtem and
 * coordinates of the live tracers in array order; den/tem/pre are the
 * fields of each tracer's cell.  A run with no live tracer writes one dummy
 * point.  Replaces sale2d_write_vtp (src/write_vtk.c:343-529) for one domain;
 * txt_path, when not NULL, also writes its text export (:492-508). */
int salec_write_vtp(const char *path, const char *txt_path, int nx, int ny, const salec_tracers *tr,
                    const double *e_den, const double *e_tem, const double *e_pre);

/* The parallel header of one tracer output (.pvtp), sale2d_write_pvtp
 * (src/write_vtk.c:303-341). */
int salec_write_pvtp(const char *path, const char *prefix, int cycle, int nproc);

/* The initial tracer list mass.tracer.proc%04d.txt: tag, gx, gy, position and
 * the density and volume of each tracer's cell, export_init_tracers
 * (src/sale2d.c:2093-2126). */
int salec_write_mass_tracers(const char *path, int nx, int ny, const salec_tracers *tr, const double *e_den,
                             const double *e_vol);

#ifdef __cplusplus
}
#endif
#endif /* SALEC_OUTPUT_H */
