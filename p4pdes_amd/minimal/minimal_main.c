/*
 * minimal -- command-line front end: `minimal <options>` as c/ch7/minimal.c
 * is run.
 */
#include "minimal.h"

int main(int argc, char **argv) {
    minimal_result r;
    return minimal_run(argc - 1, (const char **)(argv + 1), &r, NULL, NULL, 0);
}
