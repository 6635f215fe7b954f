// Registration of the "hip" linear-solver module, the way the reference
// registers "conjugate" and "jacobi" (src/linear/linear.cpp:16-23): a
// maintainer adds this file to the linear library and links libaphpois.
#include "linear_hip.h"

DECLARE_FORCE_LINK_TARGET(linear_hip);

namespace linear {

template class SolverHip<MeshCartesian<double, 3>>;

bool kReg_hip[] = {RegisterModule<ModuleLinearHip<MeshCartesian<double, 3>>>()};

}  // namespace linear
