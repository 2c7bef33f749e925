/*
 * salec_tools.h -- the solver's command-line tooling as library calls:
 * conversion of iSALE-2D input files into .inp keys, and packing of the
 * per-domain .vts output files.
 *
 * All functions return SALEC_OK or SALEC_EINVAL (bad argument, a file that
 * cannot be opened or created, a malformed pack).
 */
#ifndef SALEC_TOOLS_H
#define SALEC_TOOLS_H
#include <stdio.h>
#include "salec.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Read an iSALE-2D asteroid.inp (mesh, layers) and material.inp (materials,
 * strength, damage, melt and softening parameters) and write the .inp lines
 * they translate to, as the reference's generate_input prints them (the two
 * banner lines, "#### N variables IN InputFile ####", one "key = value" line
 * per key, "#### End of InputFile ####").  Replaces generate_input
 * (src/io/generate_input.c:13-30) with GenerateMesh and GenerateMaterial
 * (src/io/iSALEInputParser.c:166-475) on files read as OpeniSALEInputFile
 * (:44-119) reads them.  out == NULL writes to stdout. */
int salec_isale_to_inp(const char *asteroid_inp, const char *material_inp, FILE *out);

/* Pack the per-domain state files <vts_dir>/<prefix>.proc%04d.%04d.vts of
 * cycles [min_cycle, max_cycle) and domains [0, nproc) into nthreads pack
 * files <out_dir>/<prefix>.vts.pack%04d.0000: pack t holds the cycles
 * [min + t*c, min + (t+1)*c) with c = (max - min)/nthreads + 1, each file as
 * the line "#BEGINFILE:<vts_dir>/<name>", its bytes and the line "#ENDFILE".
 * The reference's merge_vts main (src/other/merge_vts.c:209-296, with
 * add_file :90-107) packs cycles [0, 387) with 4 threads from ./vts into
 * ./mvts.  max_cycle < 0 packs the cycles present from min_cycle up to the
 * first missing one; a missing file inside an explicit range is an error
 * (the reference reads through a NULL FILE there). */
int salec_pack_vts(const char *vts_dir, const char *prefix, int nproc, int min_cycle, int max_cycle,
                   int nthreads, const char *out_dir);

/* Unpack a pack file: every "#BEGINFILE:<name>" ... "#ENDFILE" section is
 * written back to dir/<name> (dir == NULL: <name> as stored), creating
 * missing directories.  Replaces extract_file
 * (src/other/merge_vts.c:109-141), which keeps the line's newline in the
 * file name.  *nfiles (may be NULL) receives the count. */
int salec_unpack_vts(const char *pack, const char *dir, int *nfiles);

/* One multiblock file per cycle, <out_dir>/<prefix>.%04d.vts, holding the
 * nproc domain files of that cycle as DataSet blocks (the lines between each
 * file's VTKFile tags), as create_vtm_block (src/other/merge_vts.c:143-207)
 * writes it for its hard-coded cycles 8 and 9. */
int salec_merge_vts_blocks(const char *vts_dir, const char *prefix, int nproc, int min_cycle, int max_cycle,
                           const char *out_dir);

#ifdef __cplusplus
}
#endif
#endif /* SALEC_TOOLS_H */
