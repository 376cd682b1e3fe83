// bscf_layout.h -- the layout of one server slot's matrices in its single
// device buffer, shared by the size computation and the carving so that the
// two cannot disagree.  Plain C++ (the CPU tests check it).
#pragma once
#include <cstddef>

namespace hfg {

// Matrices of a slot, in buffer order.
enum SlotMat {
  SM_S, SM_H, SM_X, SM_F, SM_P, SM_PPREV, SM_G, SM_PB, SM_T1, SM_T2, SM_E, SM_FU,  // n x n (X: n x m)
  SM_FO, SM_XS, SM_X2,                                                           // m x m
  SM_CF, SM_AF, SM_AP,                                                           // BMAXS n x n each
  SM_CE,                                                                         // BMAXS x elen
  SM_COUNT
};

constexpr int SLOT_DIIS = 8;       // DIIS subspace of a slot
constexpr size_t SLOT_ALIGN = 32;  // doubles (256 bytes) between sub-buffers

// Offsets (in doubles) of every matrix of a slot of n functions and m
// orthogonal functions into off[SM_COUNT]; returns the doubles the buffer
// needs.  elen: the DIIS error length (m*m when m < n, else n*n).
inline size_t slot_layout(size_t n, size_t m, size_t* off) {
  const size_t nn = n * n, mm = m * m, elen = m < n ? mm : nn;
  const size_t len[SM_COUNT] = {nn, nn, n * m, nn, nn, nn, nn, nn, nn, nn, nn, nn, mm, mm, mm,
                                SLOT_DIIS * nn, SLOT_DIIS * nn, SLOT_DIIS * nn, SLOT_DIIS * elen};
  size_t p = 0;
  for (int i = 0; i < SM_COUNT; ++i) {
    if (off) off[i] = p;
    p += (len[i] + SLOT_ALIGN - 1) / SLOT_ALIGN * SLOT_ALIGN;
  }
  return p;
}

}  // namespace hfg
