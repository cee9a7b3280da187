/*
 * a12_oracle_atlas.c -- ATLAS12 parts of the CPU oracle (test
 * infrastructure).  Filled in with the LINOP1 / RT-pass restatements.
 */
#include "a12_oracle.h"
