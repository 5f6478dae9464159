/*
 * CPU oracle for the elastic velocity-stress staggered-grid propagator and its
 * adjoint-state gradient.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and the CPU-baseline
 * leg of the benchmark load this library, and only as the checker or as the
 * CPU timing baseline; the product (pyfwi_amd) never links or calls it.
 *
 * It restates, in plain C, the per-step kernel sequence that PyFWI v0.1.10
 * launches from Python:
 *   forward   src/PyFWI/wave_propagation.py:625-733 (kernels elastic.cl:155-308,
 *             elastic_surface.cl:7-41)
 *   gradient  src/PyFWI/wave_propagation.py:735-856 (kernels elastic.cl:372-526,
 *             elastic_surface.cl:48-72)
 * and adds an exact discrete adjoint of the same forward (orc_exact_gradient),
 * which the reference does not have, for finite-difference checks.
 *
 * Two instantiations: *_f (float storage with double intermediates, the
 * reference's arithmetic on an fp64-capable OpenCL device) and *_d (FP64).
 *
 * Pinning (what checks each part, see tests/test_oracle.py):
 *   derivative operators  adjointness <x, D+ y> = -<D- x, y>
 *                         (src/PyFWI/fwi_tools.py:621-632), exact polynomial
 *                         derivatives, coefficient values
 *   forward step          exact time reversibility without damping and source,
 *                         P-wave arrival time in a homogeneous medium, energy
 *                         conservation without damping, source-receiver
 *                         reciprocity
 *   pyfwi adjoint step    interior transposition: with zero damping it equals
 *                         the transpose of the stress-then-velocity step
 *   exact gradient        central finite differences in FP64
 *   pyfwi damping terms of the adjoint step and the imaging-condition formulas
 *                         (elastic.cl:395-408, 465-466, 485-524): pinned by
 *                         restatement and by agreement in direction with the
 *                         exact gradient only -- parity unpinned beyond that.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef struct orc_grid {
    int nz, nx;          /* grid including the absorbing layer (tnz, tnx)  */
    int sdo;             /* half stencil width: 2 (O4) or 4 (O8)           */
    int npml;            /* absorbing-layer thickness in cells             */
    double c1, c2, c3, c4;
    double dt, dx, dz;   /* kernel constants as the reference prints them  */
} orc_grid;

typedef struct orc_acq {
    int ns;              /* shots                                          */
    int nr;              /* receivers per shot (row length of a record)    */
    int n_main_rec;      /* surface receivers                              */
    int dxr;             /* receiver spacing in cells                      */
    int rec_row;         /* receiver row incl. absorbing layer             */
    int rec_col0;        /* first receiver column incl. absorbing layer    */
    const int *srcx;     /* [ns] source column incl. absorbing layer       */
    const int *srcz;     /* [ns] source row incl. absorbing layer          */
    int nt;              /* internal time steps                            */
    int dt_scale;        /* internal steps per recorded sample             */
    int nt_ext;          /* recorded samples                               */
} orc_acq;

enum { TAP_F = 0, TAP_FA = 1, TAP_FM = 2 };

#define R float
#define SFX(x) x##_f
#include "elastic_oracle_impl.h"
#undef R
#undef SFX

#define R double
#define SFX(x) x##_d
#include "elastic_oracle_impl.h"
#undef R
#undef SFX

int orc_abi_version(void) { return 1; }
