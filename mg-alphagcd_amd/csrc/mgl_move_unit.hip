// Move kernels instantiated for unit-weight level-0 graphs (int32 hash counts).
#pragma clang fp contract(off)
#include "mgl_move.h"

int move_bin_unit(int mode, int rec32, bool cnt, const BinSpec& bs, const MoveArgs& A,
                 const vid* perm, int64_t nb, hipStream_t st, char* name, int len) {
  if (mode != MODE_UNIT) return -1;
  return rec32 ? move_bin_mode<ModeUnit, int32_t>(cnt, bs, A, perm, nb, st, name, len)
               : move_bin_mode<ModeUnit, int64_t>(cnt, bs, A, perm, nb, st, name, len);
}

void move_allow_lds_unit(int bytes) {
  move_allow_lds_t<ModeUnit, int32_t>(bytes);
  move_allow_lds_t<ModeUnit, int64_t>(bytes);
}
