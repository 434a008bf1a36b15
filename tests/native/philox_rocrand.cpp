// Second source for Philox4x32-10: rocRAND's engine (header-only, host side).
// Reads lines "c0 c1 c2 c3 k0 k1" (hex) on stdin, prints the 4 output words.
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

// ten_rounds is a protected member; a subclass makes it callable.
struct Engine : rocrand_device::philox4x32_10_engine {
  using rocrand_device::philox4x32_10_engine::ten_rounds;
};

int main() {
  unsigned c[4], k[2];
  Engine eng;
  while (std::scanf("%x %x %x %x %x %x", &c[0], &c[1], &c[2], &c[3], &k[0], &k[1]) == 6) {
    uint4 o = eng.ten_rounds(uint4{c[0], c[1], c[2], c[3]}, uint2{k[0], k[1]});
    std::printf("%08x %08x %08x %08x\n", o.x, o.y, o.z, o.w);
  }
  return 0;
}
