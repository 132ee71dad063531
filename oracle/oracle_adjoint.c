/* oracle_adjoint.c -- TEST INFRASTRUCTURE (filled in below) */
#include "oracle.h"
