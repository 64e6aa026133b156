/*
 * ORACLE -- test infrastructure, not product code.
 *
 * Plain-C restatement of the reference solver's hot path (fd2dmpi/fdcore.f90,
 * pml.f90, io.f90:231-300, math.f90:342-361, solver.f90:120-281), compiled once in
 * FP32 (the reference's own arithmetic, operation order kept, no FMA contraction) and
 * once in FP64 (the error reference for every GPU result).  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, and only as
 * the checker.
 *
 * How each function is pinned (tests/test_oracle_pins.py):
 *   orc_forward_shot   reference dot-product known answer a = 1.25501e35
 *                      (dev/test/02_adjoint_test.ipynb:276), exact 2-D Green's
 *                      function solution in a homogeneous medium (waveform and
 *                      amplitude), reciprocity in a heterogeneous medium with and
 *                      without damping, one-step impulse response against the
 *                      textbook 4th-order staggered stencil, source/receiver cell rules
 *   orc_adjoint_shot   known answer ratio = 1.00000 (same notebook), reconstruction
 *                      invariant (reconstructed p == forward p at every step),
 *                      forward illumination == sum of forward p^2, kernel against the
 *                      finite-difference derivative of the L2 misfit (sign, dt scaling,
 *                      and <g, dc> = dt dJ/dc in absolute terms, which fixes the -2/c),
 *                      against the exact discrete gradient (correlation), and the
 *                      coefficient swap: in a heterogeneous medium the adjoint record
 *                      equals the forward record of the medium (c, 1/(den c^2))
 *   orc_setup_pml      closed form of the quadratic profile, corner split, symmetry
 *   orc_pad, orc_vmin  brute force on small arrays
 * Parity unpinned: the density node used by the velocity updates in a heterogeneous
 * medium (den(iz,ix), not an average; fdcore.f90:99,108) is taken from the reference
 * text; reciprocity and the known answers do not distinguish it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define R float
#define FN(name) name##_f32
#define LOGR logf
#include "fd2d_impl.h"
#undef R
#undef FN
#undef LOGR

#define R double
#define FN(name) name##_f64
#define LOGR log
#include "fd2d_impl.h"
#undef R
#undef FN
#undef LOGR

#undef IDX
