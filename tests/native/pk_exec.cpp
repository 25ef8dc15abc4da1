// Host build of the packed two-state interpreter (sx_pk.h) for the CPU test
// that checks it against the oracle's executor: pk_exec runs word[i] on the
// state pair (lo[i], hi[i]) (8 bytes each, A X Y P S M S0 S1) in place.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../6502-optimizer_amd/csrc/sx_pk.h"

extern "C" int pk_exec(uint8_t *lo, uint8_t *hi, const uint32_t *word, uint64_t n) {
  using namespace sx;
  for (uint64_t i = 0; i < n; i++) {
    uint8_t *a = lo + 8 * i, *b = hi + 8 * i;
    PState ps[1];
    pk16 *f[8] = {&ps[0].a, &ps[0].x, &ps[0].y, &ps[0].p, &ps[0].s, &ps[0].m, &ps[0].s0, &ps[0].s1};
    for (int r = 0; r < 8; r++) {
      f[r]->x = a[r];
      f[r]->y = b[r];
    }
    if (!step_pk<1>(ps, word[i] & 0xFFFFu, word[i] >> 16)) return -1;
    for (int r = 0; r < 8; r++) {
      a[r] = (uint8_t)f[r]->x;
      b[r] = (uint8_t)f[r]->y;
    }
  }
  return 0;
}
