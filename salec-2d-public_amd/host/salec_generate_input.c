/*
 * salec_generate_input -- print the .inp lines an iSALE-2D model translates
 * to, the counterpart of the reference's generate_input
 * (src/io/generate_input.c:13-30):
 *
 *   salec_generate_input [asteroid.inp material.inp]
 *
 * Without arguments it reads ./asteroid.inp and ./material.inp, as the
 * reference does.
 */
#include <stdio.h>

#include "../../include/salec_tools.h"

int main(int argc, char **argv)
{
    if (argc != 1 && argc != 3) {
        fprintf(stderr, "usage: %s [asteroid.inp material.inp]\n", argv[0]);
        return 2;
    }
    const char *a = argc == 3 ? argv[1] : "./asteroid.inp";
    const char *m = argc == 3 ? argv[2] : "./material.inp";
    int rc = salec_isale_to_inp(a, m, stdout);
    if (rc != SALEC_OK) {
        fprintf(stderr, "%s: cannot convert %s and %s: %s\n", argv[0], a, m, salec_strerror(rc));
        return 1;
    }
    return 0;
}
