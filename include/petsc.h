/*
 * petsc.h -- lets residual/Jacobian callbacks that `#include <petsc.h>`
 * (c/ch6/poissonfunctions.c:1) compile unchanged against the p4g host
 * surface.  Only the subset declared in p4g_petsc.h exists.
 */
#ifndef P4G_PETSC_FORWARD_H
#define P4G_PETSC_FORWARD_H
#include "p4g_petsc.h"
#endif
