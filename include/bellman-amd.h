/*
 * libbellman extensions for MI355X.  These entry points have no counterpart
 * in the reference's src/bellman-cuda.h; they serve callers that need
 * fixed-base generation of MSM bases and batched / served execution of many
 * MSM, NTT and permutation jobs.
 * Conventions are those of bellman-cuda.h (bc_error codes, Montgomery
 * elements, asynchronous on the given stream).
 */
#pragma once

#include "bellman-cuda.h"

#ifdef __cplusplus
extern "C" {
#endif

/* bases[i] = [scalars[i]] G as affine Montgomery Fq points (64 bytes each),
 * G = (1, 2).  scalars are canonical (not Montgomery) Fr values, device memory.
 * Requires ff_set_up.  A zero scalar yields the affine encoding (0, 0). */
bc_error bc_amd_generator_multiples(const void *scalars, void *bases, unsigned count, bc_stream stream);

#ifdef __cplusplus
} // extern "C"
#endif
