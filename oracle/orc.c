/*
 * Oracle library: the CPU restatement of the reference's forward, adjoint and
 * imaging, instantiated for float, double and double-complex.
 *
 * TEST INFRASTRUCTURE ONLY (tests/, the smoke check and the CPU-baseline leg
 * of the benchmark load it; the product never does).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int nz_dom, nx_dom, pad, free_surface, fd_order, nt;
    double dx, dz, dt, fc;
} orc_geom;

int orc_version(void) { return 1; }

#define T float
#define SFX f32
#define RE(x) ((double)(x))
#include "orc_impl.h"
#undef T
#undef SFX
#undef RE

#define T double
#define SFX f64
#define RE(x) ((double)(x))
#include "orc_impl.h"
#undef T
#undef SFX
#undef RE

#define T double _Complex
#define SFX c128
#define RE(x) creal(x)
#include "orc_impl.h"
#undef T
#undef SFX
#undef RE
