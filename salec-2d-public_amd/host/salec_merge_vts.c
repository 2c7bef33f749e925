/*
 * salec_merge_vts -- reduce the number of output files of a run, the
 * counterpart of the reference's merge_vts (src/other/merge_vts.c:209-296):
 *
 *   salec_merge_vts [INP] [--nproc N] [--cycles MIN MAX] [--threads T]
 *                   [--vts DIR] [--out DIR] [--blocks]
 *   salec_merge_vts --unpack PACK [DIR]
 *
 * INP (default ../sale2d.inp, as in the reference) gives output.prefix.
 * The per-domain state files DIR/<prefix>.proc%04d.%04d.vts (default
 * ./vts) of cycles [MIN, MAX) are packed into T pack files (default 4)
 * OUT/<prefix>.vts.pack%04d.0000 (default ./mvts); --blocks writes one
 * multiblock file per cycle instead; --unpack restores the files of a pack.
 * Without --cycles every cycle from 0 up to the first missing one is taken
 * (the reference packs the fixed range [0, 387)); without --nproc the
 * domains present for the first cycle are counted (the reference reads
 * processor.npgx x npgy; salec2d writes one domain).  The output directory
 * is created when missing and never emptied.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "../../include/salec_tools.h"
#include "inp.h"

static int usage(const char *p)
{
    fprintf(stderr, "usage: %s [INP] [--nproc N] [--cycles MIN MAX] [--threads T] [--vts DIR] [--out DIR] [--blocks]\n"
                    "       %s --unpack PACK [DIR]\n", p, p);
    return 2;
}

int main(int argc, char **argv)
{
    if (argc >= 2 && !strcmp(argv[1], "--unpack")) {
        if (argc < 3 || argc > 4) return usage(argv[0]);
        int n = 0, rc = salec_unpack_vts(argv[2], argc == 4 ? argv[3] : NULL, &n);
        if (rc != SALEC_OK) { fprintf(stderr, "%s: cannot unpack %s\n", argv[0], argv[2]); return 1; }
        fprintf(stdout, "extracted %d files from %s\n", n, argv[2]);
        return 0;
    }
    const char *inp_path = "../sale2d.inp", *vts = "./vts", *out = "./mvts";
    int nproc = -1, cmin = 0, cmax = -1, threads = 4, blocks = 0, have_inp = 0;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--nproc") && i + 1 < argc) nproc = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--cycles") && i + 2 < argc) { cmin = atoi(argv[++i]); cmax = atoi(argv[++i]); }
        else if (!strcmp(argv[i], "--threads") && i + 1 < argc) threads = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--vts") && i + 1 < argc) vts = argv[++i];
        else if (!strcmp(argv[i], "--out") && i + 1 < argc) out = argv[++i];
        else if (!strcmp(argv[i], "--blocks")) blocks = 1;
        else if (argv[i][0] != '-' && !have_inp) { inp_path = argv[i]; have_inp = 1; }
        else return usage(argv[0]);
    }
    inp_file *f = inp_open(inp_path, 0);
    if (!f) { fprintf(stderr, "%s: cannot open %s\n", argv[0], inp_path); return 1; }
    char prefix[100];
    inp_s(f, "output.prefix", prefix, sizeof prefix, "ParaTest");
    inp_close(f);
    if (nproc < 0) {
        char name[4096];
        struct stat st;
        for (nproc = 0;; ++nproc) {
            snprintf(name, sizeof name, "%s/%s.proc%04d.%04d.vts", vts, prefix, nproc, cmin);
            if (stat(name, &st) != 0) break;
        }
        if (nproc == 0) { fprintf(stderr, "%s: no %s/%s.proc0000.%04d.vts\n", argv[0], vts, prefix, cmin); return 1; }
    }
    int rc = blocks ? salec_merge_vts_blocks(vts, prefix, nproc, cmin, cmax, out)
                    : salec_pack_vts(vts, prefix, nproc, cmin, cmax, threads, out);
    if (rc != SALEC_OK) { fprintf(stderr, "%s: a file could not be read or written\n", argv[0]); return 1; }
    fprintf(stdout, "%s %d domains of %s from %s into %s\n", blocks ? "merged" : "packed", nproc, prefix, vts, out);
    return 0;
}
