/*
 * ORACLE — test infrastructure only (see oracle.h): the C restatement of the reference's
 * PISO kernels and Krylov recurrences, instantiated for double (_f64) and float (_f32).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define REAL double
#define SFX _f64
#include "oracle_impl.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX _f32
#include "oracle_impl.h"
#undef REAL
#undef SFX
