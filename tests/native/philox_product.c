/* The product's Philox (transportoptimiser_amd/csrc/rng.h) built as host C.
 * Reads lines "c0 c1 c2 c3 k0 k1" (hex), prints the 4 output words and the
 * binary32 uniforms of the first two words. */
#include <stdio.h>

#include "rng.h"

int main(void) {
  unsigned c[4], k[2];
  while (scanf("%x %x %x %x %x %x", &c[0], &c[1], &c[2], &c[3], &k[0], &k[1]) == 6) {
    rmc_u4 o = rmc_philox(c[0], c[1], c[2], c[3], k[0], k[1]);
    printf("%08x %08x %08x %08x %.9g %.9g\n", o.x, o.y, o.z, o.w, (double)rmc_u01f(o.x),
           (double)rmc_u01f(o.y));
  }
  return 0;
}
