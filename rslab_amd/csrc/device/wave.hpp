// Wave-wide (64-lane) reductions on DPP lane moves (no LDS round trips):
// row_shr 1/2/4/8 inside each row of 16 lanes, then row_bcast 15/31 across
// rows; lane 63 ends with the result, which readlane makes wave-uniform.
#pragma once
#include <hip/hip_runtime.h>

namespace rsl {

template <int CTRL, int RM, int BM>
__device__ inline double dpp_move(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, RM, BM, false);
  hi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, RM, BM, false);
  return __hiloint2double(hi, lo);
}

// max over the 64 lanes of a wave, returned to every lane
__device__ inline double wave_max_dpp(double v) {
  v = fmax(v, dpp_move<0x111, 0xF, 0xF>(v));
  v = fmax(v, dpp_move<0x112, 0xF, 0xF>(v));
  v = fmax(v, dpp_move<0x114, 0xF, 0xF>(v));
  v = fmax(v, dpp_move<0x118, 0xF, 0xF>(v));
  v = fmax(v, dpp_move<0x142, 0xA, 0xF>(v));
  v = fmax(v, dpp_move<0x143, 0xC, 0xF>(v));
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// Column max and its first row over candidates i = i0 + lane + 64 q: each lane
// enters with its own best (m, idx) found scanning its rows upward with a
// strict '>' (idx = i0 - 1 when it saw none above 0). On return every lane
// holds the wave max in m and the smallest row reaching it in idx (i0 - 1
// when the max is 0). Rows are below i0 + 128 (at most two per lane).
__device__ inline void wave_argmax_dpp(double& m, int& idx, int i0) {
  const int lane = threadIdx.x & 63;
  const double wm = wave_max_dpp(m);
  if (wm == 0.0) {
    m = 0.0;
    idx = i0 - 1;
    return;
  }
  const unsigned long long first = __ballot(m == wm && idx == i0 + lane);
  if (first) {
    idx = i0 + (__ffsll((long long)first) - 1);
  } else {
    // every lane reaching the max found it on its second row: the lowest
    // such lane holds the smallest row among them
    const unsigned long long any = __ballot(m == wm);
    const int l = __ffsll((long long)any) - 1;
    idx = __builtin_amdgcn_readlane(idx, l);
  }
  m = wm;
}

}  // namespace rsl
