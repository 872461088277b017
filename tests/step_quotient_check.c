/* Exhaustive check of the quotient the count draws use in place of a division.

   The inversion loops of the count draw divide by the step number s = 1, 2, ...; on the
   GPU, steps s <= 64 compute t / s as

       y = RN(1/s),  q0 = RN(t * y),  r = fma(-s, q0, t),  q = fma(r, y, q0)

   (Markstein's correction of t * y by its exact residual).  This program compares q with
   the correctly rounded IEEE quotient RN(t / s) for every significand of t in [1, 2) and
   every s in 1 .. S (the first argument).  Scaling t by a power of two scales every step
   exactly while all values stay normal, so one binade covers every exponent at which the
   residual is normal.

   A second argument plants a plausible mistake, which must then show as mismatches:
   "nocorr" drops the residual correction (q = q0), "lane" reads the reciprocal of the
   next step (y = RN(1/(s+1)), an off-by-one in the lane index).  A reciprocal one ulp off
   is not such a mistake: the correction absorbs it.

   Prints "mismatches N" and exits 0 on success of the comparison it was asked for. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char** argv) {
  const int smax = argc > 1 ? atoi(argv[1]) : 64;
  const int nocorr = argc > 2 && strcmp(argv[2], "nocorr") == 0;
  const int lane = argc > 2 && strcmp(argv[2], "lane") == 0;
  const int plant = nocorr || lane;
  long bad = 0;
  for (int s = 1; s <= smax; ++s) {
    volatile float sf = (float)s;
    volatile float sn = (float)(s + 1);
    const float y = lane ? 1.0f / sn : 1.0f / sf;
    for (uint32_t m = 0; m < (1u << 23); ++m) {
      const uint32_t bits = (127u << 23) | m;
      float t;
      memcpy(&t, &bits, sizeof t);
      const float want = t / sf;
      const float q0 = t * y;
      const float r = fmaf(-sf, q0, t);
      const float q = nocorr ? q0 : fmaf(r, y, q0);
      if (q != want) ++bad;
    }
  }
  printf("mismatches %ld\n", bad);
  return plant ? (bad > 0 ? 0 : 1) : (bad == 0 ? 0 : 1);
}
