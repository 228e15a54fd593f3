// Host FP64 dense block.  Drop-in for matrix/Dense.h:6-18 of the reference:
// m rows, n columns, leading dimension lda = max(m, lda), storage lda*n
// doubles owned by the object.  Copying is disabled (the reference's implicit
// copy would free `a` twice).
#pragma once

#include <cstdint>

class Dense {
public:
  int64_t m;
  int64_t n;
  int64_t lda;

  double* a;

  Dense(int64_t m, int64_t n, int64_t lda);
  ~Dense();

  Dense(const Dense&) = delete;
  Dense& operator=(const Dense&) = delete;
};
