// The UV map's interior solve on the GPU (uvsolve.hip): the host's iteration of
// cut_and_parameterize, operation for operation, so the solution is the host's bit for bit.
#pragma once

#include <vector>

#include "mesh.h"

namespace tissue {

// whether a GPU is visible to this process
bool uv_solve_gpu_available();
// solves s on the current GPU into sol[0], sol[1]; throws when it does not converge within
// s.max_it iterations (as the host's solve would) or on a HIP error
bool uv_solve_gpu(const UvSystem& s, std::vector<double>* sol);

}  // namespace tissue
