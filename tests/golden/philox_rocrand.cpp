// Emits Philox4x32-10 outputs from rocRAND's engine (an independent implementation) for
// the golden fixture philox_vectors.json. Built and run by make_philox_vectors.py.
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
int main() {
  const unsigned long long seeds[] = {0ull, 1ull, 0xffffffffffffffffull, 0x299f31d0a4093822ull, 123456789ull};
  const unsigned long long subseqs[] = {0ull, 1ull, 0x20000000ull, 0x0000000700000003ull};
  const unsigned long long blocks[] = {0ull, 1ull, 12345ull, 0xfffffffeull};
  for (auto s : seeds)
    for (auto q : subseqs)
      for (auto b : blocks) {
        rocrand_state_philox4x32_10 st;
        rocrand_init(s, q, b * 4ull, &st);  // offset in 32-bit outputs: 4 per counter block
        uint4 r = rocrand4(&st);
        printf("%llu %llu %llu %u %u %u %u\n", s, q, b, r.x, r.y, r.z, r.w);
      }
  return 0;
}
