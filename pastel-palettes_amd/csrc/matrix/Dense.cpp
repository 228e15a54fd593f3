// Host dense block (reference matrix/Dense.cpp:8-16).
#include "Dense.h"

#include <algorithm>
#include <cstdlib>
#include <new>

Dense::Dense(int64_t m_, int64_t n_, int64_t lda_) : m(m_), n(n_), lda(std::max(m_, lda_)), a(nullptr) {
  const size_t count = (size_t)lda * (size_t)(n > 0 ? n : 0);
  if (count > 0) {
    a = static_cast<double*>(std::calloc(count, sizeof(double)));
    if (a == nullptr) throw std::bad_alloc();
  }
}

Dense::~Dense() { std::free(a); }
