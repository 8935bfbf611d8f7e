// worklist_2048.hip -- the persistent work-list kernels for 2048-bit keys.
#include "paillier_device.hpp"

namespace heu_hip {

WorklistShape worklist_n2048() { return Launch<ShapeN2048>::worklist_shape<false>(); }
WorklistShape worklist_p2048() { return Launch<ShapeP2048>::worklist_shape<true>(); }

}  // namespace heu_hip
