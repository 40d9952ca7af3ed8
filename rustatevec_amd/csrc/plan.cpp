// plan.cpp -- launch-list passes of a program.
#include "program.hpp"

#include "../../include/qsvh.h"

namespace qsvh {

int plan_tiles(Program* p, int tile_qubits) {
  if (tile_qubits > 0)
    return set_err(p->st->ctx, kInvalidArg, "tile passes are not available in this build");
  return kOk;
}

int plan_mfma(Program* p) {
  if (p->precision != QSVH_PREC_EXACT)
    return set_err(p->st->ctx, kInvalidArg, "MFMA precisions are not available in this build");
  return kOk;
}

void release_extra(Program* p) {
  for (void* b : p->owned) hipFree(b);
  p->owned.clear();
}

}  // namespace qsvh
