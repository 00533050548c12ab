// Batched mode (internal): per-case reductions of many independent cases
// stored as one mesh.
#pragma once

#include <hip/hip_runtime.h>

#include "fv_device.h"

namespace fvh {
namespace batch {

constexpr int kChunkCells = 2048; // cells per reduction chunk

// partials[chunk][8]: sum |R| V (R may be null: zeros), min V/ds1,
// min ds2/(rho V) over the chunk's cells (int3 = case, begin, end).
void chunk_partials(hipStream_t s, const dev::DevMesh &m, const dev::PhysDev &P,
                    int n_chunks, const int3 *chunks, const double *U,
                    const double *Uprim, const double *R, double *partials);

// Per case k (its chunks chunk_begin[k] .. chunk_end[k] - 1): mode 1 writes
// the residual norms with the step's dt and advances case_t; both modes
// write the next dt into case_dt.
void case_fold(hipStream_t s, int mode, const dev::PhysDev &P, int n_cases,
               const int *chunk_begin, const int *chunk_end, const double *partials,
               const double *case_mu, const double *case_vol, double *case_dt, double *case_t,
               double *case_res);

// counters[0 .. n) = 0 (a kernel, so that captured graphs hold no memset)
void reset_counters(hipStream_t s, unsigned int *counters, int n);

} // namespace batch
} // namespace fvh
