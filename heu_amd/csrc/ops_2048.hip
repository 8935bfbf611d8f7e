// ops_2048.hip -- launch table for 2048-bit keys (n^2 over 4 lanes x 36 limbs;
// p^2 over 2 lanes x 36 or 4 lanes x 18 limbs) and the key-size dispatch.
#include "paillier_device.hpp"

namespace heu_hip {

WorklistShape worklist_n2048();
WorklistShape worklist_p2048(int shape);
const Ops* ops_3072();

const Ops* ops_2048() {
  static const Ops o = make_ops<ShapeN2048, 32>(worklist_n2048(), Launch<ShapeP2048>::dec_shape(worklist_p2048(0)),
                                                Launch<ShapeP2048w>::dec_shape(worklist_p2048(1)));
  return &o;
}

const Ops* ops_for_key_bits(int key_bits) {
  if (key_bits == 2048) return ops_2048();
  if (key_bits == 3072) return ops_3072();
  return nullptr;
}

}  // namespace heu_hip
