// Pieces of the fused step shared by the plain and the learned-closure kernels.
#pragma once

#include "clbm_collide.h"

namespace clbm {

template <typename R, bool FORCE, bool BOX>
__device__ __forceinline__ void force_at(int c, int x, int y, int z, const float* __restrict__ fx,
                                         const float* __restrict__ fy,
                                         const float* __restrict__ fz,
                                         const int32_t* __restrict__ box, R& Fx, R& Fy, R& Fz) {
  Fx = Fy = Fz = R(0);
  if (!FORCE) return;
  if (BOX) {
    const bool in = (x >= box[0]) & (y >= box[1]) & (z >= box[2]) & (x <= box[3]) &
                    (y <= box[4]) & (z <= box[5]);
    if (!in) return;
  }
  if (fx) Fx = R(fx[c]);
  if (fy) Fy = R(fy[c]);
  if (fz) Fz = R(fz[c]);
}

// Everything one launch of the fused step needs for one lattice.
struct StepArgs {
  Geo g;
  double s_v, s_e;
  Planes pl;
  void* f;
  const float *fx, *fy, *fz;
  const int32_t* box;
  Region rg;
};


}  // namespace clbm
