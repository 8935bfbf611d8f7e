// ops_3072.hip -- launch table for 3072-bit keys (n^2 over 8 lanes x 27 limbs;
// p^2 over 4 lanes x 27 or 8 lanes x 14 limbs).
#include "paillier_device.hpp"

namespace heu_hip {

WorklistShape worklist_n3072();
WorklistShape worklist_p3072(int shape);

const Ops* ops_3072() {
  static const Ops o = make_ops<ShapeN3072, 48>(worklist_n3072(), Launch<ShapeP3072>::dec_shape(worklist_p3072(0)),
                                                Launch<ShapeP3072w>::dec_shape(worklist_p3072(1)));
  return &o;
}

}  // namespace heu_hip
