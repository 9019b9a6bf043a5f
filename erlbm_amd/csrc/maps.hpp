// maps.hpp -- job / result records of the batched tiny-lattice engine.
// The layouts are those of erlbm_job / erlbm_job_result in include/erlbm.h.
#pragma once

#include "../../include/erlbm.h"

namespace erlbm {

using MapJob = erlbm_job;
using MapResult = erlbm_job_result;

enum { INIT_REST = ERLBM_INIT_REST, INIT_TAYLOR_GREEN = ERLBM_INIT_TAYLOR_GREEN };

}  // namespace erlbm
