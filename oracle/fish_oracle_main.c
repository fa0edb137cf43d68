/*
 * CPU ORACLE -- test infrastructure only (see oracle.h).
 * Command-line front end: `fish_oracle <fish options>` prints what fish
 * prints (c/ch6/fish.c:277-280 and the PETSc monitors it enables).
 */
#include <stdio.h>
#include <stdlib.h>
#include "oracle.h"

int main(int argc, char **argv) {
    static char out[1 << 16];
    fish_result r;
    int e = oracle_fish_run(argc - 1, (const char **)(argv + 1), &r, NULL, out, sizeof out);
    fputs(out, stdout);
    return e;
}
