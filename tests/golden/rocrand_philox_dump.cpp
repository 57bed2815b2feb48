// Prints rocRAND's Philox4x32-10 engine output (host build) for a few keys and
// counters, as "key counter w0 w1 w2 w3" lines.  Used by
// tests/test_oracle_philox.py as an independent implementation to pin the
// oracle's and the kernels' Philox.  Build: hipcc -O2 this.cpp -o dump
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

int main() {
  const unsigned long long keys[] = {0ull, 1ull, 4095ull, 0xffffffffffffffffull,
                                     0x299f31d0a4093822ull};
  for (unsigned long long k : keys) {
    // offset o advances the counter by o 4-word blocks (counter = o, subsequence 0)
    for (unsigned long long o : {0ull, 1ull, 7ull, 100000ull, 4294967296ull}) {
      rocrand_device::philox4x32_10_engine e(k, 0, o * 4);
      unsigned w0 = e(), w1 = e(), w2 = e(), w3 = e();
      printf("%llu %llu %u %u %u %u\n", k, o, w0, w1, w2, w3);
    }
  }
  return 0;
}
