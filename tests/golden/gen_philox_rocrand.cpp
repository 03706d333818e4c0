// Writes Philox4x32-10 outputs of rocRAND's host-side engine
// (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h) for a set of (counter,
// key) pairs, as JSON, to pin the project's own Philox implementation.
// rocRAND's engine(seed, subsequence, offset) uses key = (seed_lo, seed_hi)
// and counter = (offset/4 as 64 bits, subsequence as 64 bits).
#include <stdio.h>
#include <rocrand/rocrand_philox4x32_10.h>

int main() {
  const unsigned long long cases[][3] = {
      // seed, subsequence, offset (offset multiple of 4 -> counter words 0,1)
      {0ull, 0ull, 0ull},
      {1ull, 0ull, 0ull},
      {0xffffffffffffffffull, 0xffffffffffffffffull, 0x3fffffffffffffffull * 4ull},
      {0x299f31d0a4093822ull, 0x0370734413198a2eull, 0x85a308d3243f6a88ull * 4ull},
      {1ull, 7ull, 12ull * 4ull},
      {1ull, (32ull << 32) | 0ull, 123ull * 4ull},
      {0xdeadbeefull, 5ull, 99ull * 4ull},
  };
  printf("[\n");
  const int n = sizeof(cases) / sizeof(cases[0]);
  for (int i = 0; i < n; ++i) {
    rocrand_device::philox4x32_10_engine e(cases[i][0], cases[i][1], cases[i][2]);
    unsigned int r[4];
    for (int j = 0; j < 4; ++j) r[j] = e();
    unsigned long long c01 = cases[i][2] / 4ull, c23 = cases[i][1];
    printf("  {\"counter\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
           (unsigned)(c01 & 0xffffffffu), (unsigned)(c01 >> 32), (unsigned)(c23 & 0xffffffffu),
           (unsigned)(c23 >> 32), (unsigned)(cases[i][0] & 0xffffffffu), (unsigned)(cases[i][0] >> 32), r[0], r[1],
           r[2], r[3], i + 1 < n ? "," : "");
  }
  printf("]\n");
  return 0;
}
