/*
 * obstacle -- command-line front end: `obstacle <options>` as
 * c/ch12/obstacle.c is run.
 */
#include "obstacle.h"

int main(int argc, char **argv) {
    obstacle_result r;
    return obstacle_run(argc - 1, (const char **)(argv + 1), &r, NULL, NULL, 0);
}
