// worklist_2048.hip -- the persistent work-list kernel of the n^2 ops for 2048-bit keys.
#include "paillier_device.hpp"

namespace heu_hip {

WorklistShape worklist_n2048() { return Launch<ShapeN2048>::worklist_shape<false>(); }

}  // namespace heu_hip
