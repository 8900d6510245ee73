// Move kernels instantiated for int64 fixed-point weights and the all-FP64 mode.
#pragma clang fp contract(off)
#include "mgl_move.h"

int move_bin_int(int mode, int rec32, bool cnt, const BinSpec& bs, const MoveArgs& A,
                const vid* perm, int64_t nb, hipStream_t st, char* name, int len) {
  if (mode == MODE_INT) return move_bin_mode<ModeInt, int64_t>(cnt, bs, A, perm, nb, st, name, len);
  if (mode == MODE_F64) return move_bin_mode<ModeF64, double>(cnt, bs, A, perm, nb, st, name, len);
  return -1;
}

void move_allow_lds_int(int bytes) {
  move_allow_lds_t<ModeInt, int64_t>(bytes);
  move_allow_lds_t<ModeF64, double>(bytes);
}
