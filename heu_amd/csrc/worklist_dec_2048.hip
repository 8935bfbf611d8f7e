// worklist_dec_2048.hip -- the persistent decryption work-list kernels for
// 2048-bit keys, one per decryption shape.
#include "paillier_device.hpp"

namespace heu_hip {

WorklistShape worklist_p2048(int shape) {
  return shape ? Launch<ShapeP2048w>::worklist_shape<true>() : Launch<ShapeP2048>::worklist_shape<true>();
}

}  // namespace heu_hip
