/*
 * fetd_oracle_main.c -- command line of the CPU restatement (TEST
 * INFRASTRUCTURE ONLY; the checker and the CPU baseline, never the product).
 *
 * Mirrors the reference driver (src/FETD_MAIN.f90:26-98,110-165):
 *   fetd_oracle -f ROOT [-v full|none] [--max-steps N] [--threads N]
 * reads ROOT.mesh and ROOT.cntl and writes ROOT.field and ROOT.field2.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fetd_oracle.h"

int main(int argc, char **argv) {
    const char *root = NULL;
    int verbose = 1, max_steps = 0, threads = 0;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "-f") && i + 1 < argc) root = argv[++i];
        else if (!strcmp(argv[i], "-v") && i + 1 < argc) verbose = strcmp(argv[++i], "none") != 0;
        else if (!strcmp(argv[i], "--max-steps") && i + 1 < argc) max_steps = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--threads") && i + 1 < argc) threads = atoi(argv[++i]);
    }
    if (!root) {
        printf(" Usage: ./fetd_oracle -[job] FILE_ROOT\n");
        printf(" [FORWARD]\t\t-f\tCalculates the predicted data and saves the EM solution\n");
        printf(" Optional\t\t-v\t[full|none]\n");
        return 1;
    }
    if (threads > 0) fo_set_threads(threads);
    printf(" TEMF3DT oracle: Run job -f with root name: %s\n", root);
    fo_model *m = fo_load(root, verbose);
    if (!m) { fprintf(stderr, "%s\n", fo_error()); return 2; }
    if (fo_assemble(m)) { fprintf(stderr, "%s\n", fo_error()); return 3; }
    char f1[1024], f2[1024];
    snprintf(f1, sizeof f1, "%s.field", root);
    snprintf(f2, sizeof f2, "%s.field2", root);
    if (fo_run(m, f1, f2, max_steps)) { fprintf(stderr, "%s\n", fo_error()); return 4; }
    double t[6];
    fo_get_timers(m, t);
    printf(" total rows: %d  factorisations: %d  loop wall: %.3f s\n", fo_num_rows(m),
           fo_num_factorizations(m), t[5]);
    fo_free(m);
    return 0;
}
