// worklist_3072.hip -- the persistent work-list kernels for 3072-bit keys.
#include "paillier_device.hpp"

namespace heu_hip {

WorklistShape worklist_n3072() { return Launch<ShapeN3072>::worklist_shape<false>(); }
WorklistShape worklist_p3072() { return Launch<ShapeP3072>::worklist_shape<true>(); }

}  // namespace heu_hip
