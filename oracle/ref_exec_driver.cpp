// Cross-check driver (test infrastructure): links the reference CUDA tool's
// host-side executor header (cuda/6502_common.h, compiled where it lies) and
// runs it over a stream of (state, op, imm) records.
//
// stdin : n records of 12 bytes: 8 state bytes A,X,Y,P,S,M,S0,S1, u16 op, u16 imm
// stdout: n records of 8 bytes: the state after h_exec_instruction
#include <cstdint>
#include <cstdio>
#include <vector>

#include "6502_common.h"

int main() {
  init_tables();
  std::vector<uint8_t> in;
  uint8_t buf[1 << 16];
  size_t got;
  while ((got = fread(buf, 1, sizeof(buf), stdin)) > 0) in.insert(in.end(), buf, buf + got);
  size_t n = in.size() / 12;
  std::vector<uint8_t> out(n * 8);
  for (size_t i = 0; i < n; i++) {
    const uint8_t *r = &in[i * 12];
    State6502 s{r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7]};
    uint16_t op = (uint16_t)(r[8] | (r[9] << 8));
    uint16_t imm = (uint16_t)(r[10] | (r[11] << 8));
    h_exec_instruction(s, op, imm);
    uint8_t *o = &out[i * 8];
    o[0] = s.A; o[1] = s.X; o[2] = s.Y; o[3] = s.P;
    o[4] = s.S; o[5] = s.M; o[6] = s.S0; o[7] = s.S1;
  }
  fwrite(out.data(), 1, out.size(), stdout);
  return 0;
}
