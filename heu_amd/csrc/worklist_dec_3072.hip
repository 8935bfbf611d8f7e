// worklist_dec_3072.hip -- the persistent decryption work-list kernels for
// 3072-bit keys, one per decryption shape.
#include "paillier_device.hpp"

namespace heu_hip {

WorklistShape worklist_p3072(int shape) {
  return shape ? Launch<ShapeP3072w>::worklist_shape<true>() : Launch<ShapeP3072>::worklist_shape<true>();
}

}  // namespace heu_hip
