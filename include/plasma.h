/*
 * plasma.h -- umbrella header of the MI355X implementation of PLASMA's
 * Cholesky solve path (reference include/plasma.h).  Only the routines of
 * that path are provided: plasma_{d,s,z,c}{potrf,potrs,posv},
 * plasma_dsposv and plasma_zcposv, their plasma_omp_* tile forms, the layout
 * translations and the precision conversions, with the inverses, norms,
 * BLAS-3 and band routines of the same path.
 */
#ifndef PLASMA_H
#define PLASMA_H

#include "plasma_types.h"
#include "plasma_error.h"
#include "plasma_async.h"
#include "plasma_descriptor.h"
#include "plasma_context.h"
#include "plasma_d.h"
#include "plasma_s.h"
#include "plasma_ds.h"
#include "plasma_z.h"
#include "plasma_c.h"
#include "plasma_zc.h"
#include "plasma_mi355x.h"
#include "plasma_batched.h"
#include "plasma_adjoint.h"

#endif /* PLASMA_H */
