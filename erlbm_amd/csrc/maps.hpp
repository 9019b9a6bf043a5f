// maps.hpp -- job / result records of the batched tiny-lattice engine.
// The layouts are those of erlbm_job / erlbm_job_result in include/erlbm.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "../../include/erlbm.h"

namespace erlbm {

using MapJob = erlbm_job;
using MapResult = erlbm_job_result;

enum { INIT_REST = ERLBM_INIT_REST, INIT_TAYLOR_GREEN = ERLBM_INIT_TAYLOR_GREEN };

// Chunked launches of the persistent tiny-lattice kernel (maps.hip), for the
// serving loop: jobs[job_idx[k]] for k < n (n read from n_jobs_dev when set,
// else n_jobs_max), each advanced by at most max_steps steps, state kept in
// `saved` (state_bytes() per job, job-indexed) between launches.  Graph
// capturable (the queue head is cleared by a kernel).
namespace tiny {
size_t state_bytes(const erlbm_batch *b, int precision);
int fluid_sites(const erlbm_batch *b);
int launch(erlbm_batch *b, int precision, int collision, const MapJob *jobs, MapResult *results, const int *job_idx,
           const int *n_jobs_dev, int n_jobs_max, unsigned int *counter, int max_steps, void *saved,
           hipStream_t stream);
}  // namespace tiny

}  // namespace erlbm
