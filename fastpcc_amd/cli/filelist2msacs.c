/*
 * filelist2msacs — gathers the SAC files of a list into one MSACS1 file (the
 * reference's Filelist2msacs, src/Filelist2msacs.c:18-51), through libfastpcc's
 * format entry points:
 *
 *   filelist2msacs filelist.txt out.msacs [Nmax=samples]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fastpcc_io.h"

int main(int argc, char **argv)
{
    if (argc < 3) {
        puts("\nUSAGE: filelist2msacs \"List of sac files\" \"Output file name\" [Nmax=\"maximum number of samples per sequence\"]");
        return 0;
    }
    unsigned N = 0, n = 0;
    if (argc > 3 && !strncmp(argv[3], "Nmax=", 5)) {
        char *e;
        long v = strtol(argv[3] + 5, &e, 10);
        if (e == argv[3] + 5 || v < 0) puts("Error when reading Nmax.\n");
        else N = (unsigned)v;
    }
    float dt = 0;
    int r = fpcc_sac_read_list(argv[1], &N, 0, NULL, NULL, &n, &dt);
    fpcc_trace_header *h = NULL;
    float *x = NULL;
    if (!r) {
        h = calloc(n ? n : 1, sizeof *h);
        x = calloc((size_t)(n ? n : 1) * N, sizeof(float));
        r = (h && x) ? fpcc_sac_read_list(argv[1], &N, n, h, x, &n, &dt) : 4;
    }
    if (r) printf("Error %d when reading %s.\n", r, argv[1]);
    else if ((r = fpcc_msacs_write(argv[2], h, x, n, N))) printf("Error %d when writing to %s.\n", r, argv[2]);
    free(h);
    free(x);
    return r;
}
