// Move kernels instantiated for integer weights below 2^31 (coarse levels of unit graphs).
#pragma clang fp contract(off)
#include "mgl_move.h"

int move_bin_i32(int mode, int rec32, bool cnt, const BinSpec& bs, const MoveArgs& A,
                const vid* perm, int64_t nb, hipStream_t st, char* name, int len) {
  if (mode != MODE_I32) return -1;
  return rec32 ? move_bin_mode<ModeI32, int32_t>(cnt, bs, A, perm, nb, st, name, len)
               : move_bin_mode<ModeI32, int64_t>(cnt, bs, A, perm, nb, st, name, len);
}

void move_allow_lds_i32(int bytes) {
  move_allow_lds_t<ModeI32, int32_t>(bytes);
  move_allow_lds_t<ModeI32, int64_t>(bytes);
}
