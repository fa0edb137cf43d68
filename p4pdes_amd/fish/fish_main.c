/*
 * fish -- command-line front end: `fish <options>` as c/ch6/fish.c is run.
 */
#include <stdio.h>
#include "fish.h"

int main(int argc, char **argv) {
    fish_result r;
    return fish_run(argc - 1, (const char **)(argv + 1), &r, NULL, NULL, 0);
}
