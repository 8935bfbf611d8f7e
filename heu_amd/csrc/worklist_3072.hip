// worklist_3072.hip -- the persistent work-list kernel of the n^2 ops for 3072-bit keys.
#include "paillier_device.hpp"

namespace heu_hip {

WorklistShape worklist_n3072() { return Launch<ShapeN3072>::worklist_shape<false>(); }

}  // namespace heu_hip
