// The distance table's local marches on the current GPU (mesh_gpu.hip): fills rows[s] and sets
// done[s] for every source whose march fits the kernel's tables (the host marches the rest);
// the rows equal the host march's. Throws std::runtime_error on a HIP error.
#pragma once
#include "mesh.h"

namespace tissue {

void march_on_gpu(const UVMesh& m, const MarchMesh& mm, double lim, MarchRows& rows, std::vector<char>& done);

}  // namespace tissue
