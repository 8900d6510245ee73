// Move kernels instantiated for FP32 hash sums (mixed precision).
#pragma clang fp contract(off)
#include "mgl_move.h"

int move_bin_f32(int mode, int rec32, bool cnt, const BinSpec& bs, const MoveArgs& A,
                const vid* perm, int64_t nb, hipStream_t st, char* name, int len) {
  if (mode != MODE_F32) return -1;
  return move_bin_mode<ModeF32, double>(cnt, bs, A, perm, nb, st, name, len);
}

void move_allow_lds_f32(int bytes) {
  move_allow_lds_t<ModeF32, double>(bytes);
}
