// classes.h -- the 21 shell-quartet classes up to (dd|dd), in kernel
// orientation (la>=lb, lc>=ld, pair class of bra >= ket), ordered by
// quartet_class() index.
#pragma once
#define HFG_FOR_EACH_CLASS(X)                                                                        \
  X(0, 0, 0, 0) X(1, 0, 0, 0) X(1, 0, 1, 0) X(1, 1, 0, 0) X(1, 1, 1, 0) X(1, 1, 1, 1) X(2, 0, 0, 0)     \
  X(2, 0, 1, 0) X(2, 0, 1, 1) X(2, 0, 2, 0) X(2, 1, 0, 0) X(2, 1, 1, 0) X(2, 1, 1, 1) X(2, 1, 2, 0)     \
  X(2, 1, 2, 1) X(2, 2, 0, 0) X(2, 2, 1, 0) X(2, 2, 1, 1) X(2, 2, 2, 0) X(2, 2, 2, 1) X(2, 2, 2, 2)
#define HFG_NCLASS 21

// All 36 classes, including the auxiliary f-s and g-s pair classes used by the
// density-fitting integrals (block kernels only).
#include "gen/class_list.inc"
