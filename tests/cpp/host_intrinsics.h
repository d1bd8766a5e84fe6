// Host definitions of the two device intrinsics the library's headers use, for compiling the
// host-side scheduler (fh_serve.cpp) with a plain C++ compiler in tests/cpp/serve_sched_check.cpp.
#pragma once
#include <cstdint>
static inline unsigned long long __umul64hi(unsigned long long a, unsigned long long b) {
  return (unsigned long long)(((unsigned __int128)a * b) >> 64);
}
static inline unsigned int __brev(unsigned int x) {
  unsigned int r = 0;
  for (int i = 0; i < 32; ++i)
    if (x >> i & 1u) r |= 1u << (31 - i);
  return r;
}
