// 6502opt -- command-line front end over lib6502core (include/sx6502.h).
//
// Keeps the subcommands, flags and output schemas of the reference CLI
// (cmd/6502opt/main.go): enumerate (:33-83), target (:88-132), verify-jsonl
// (:136-145, :386-463), stoke (:196-275) and the parsers parseDeadFlags / parseAssembly /
// parseSingleInstruction / parseImmediate (:285-384, :465-494).  The search
// itself runs on the GPU through the C ABI.
//
// Additions: --jsonl PATH writes the CUDA tool's JSONL schema
// (cuda/6502_search_v2.cu:912-930); --bytes-model go|cuda, --hits first|all,
// --exh projected|full, --no-mid and --device select the device semantics;
// --first-target/--end-target select a range of the target index space (the
// multi-process split of the CUDA tool, cu:775-777); --max-target 3 runs the
// length-2 targets and then the length-3 ones (raw Go-order triple indices),
// both against candidates of length 1 and 2 (search.go:26-50);
// --checkpoint FILE [--resume] [--chunk N] appends JSONL chunk by chunk and
// records the next target index and the committed JSONL size; stoke
// takes --seed (the chains are reproducible) and reports the first verified
// rewrite.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include <unistd.h>

#include "sx6502.h"

namespace {

const char *kMnem[58] = {
    "TAX", "TAY", "TXA", "TYA", "TSX", "TXS", "INX", "INY", "DEX", "DEY",
    "ASL A", "LSR A", "ROL A", "ROR A", "CLC", "SEC", "CLD", "SED", "CLI", "SEI", "CLV",
    "PHA", "PLA", "PHP", "PLP", "NOP",
    "LDA #n", "LDX #n", "LDY #n", "ADC #n", "SBC #n", "AND #n", "ORA #n", "EOR #n",
    "CMP #n", "CPX #n", "CPY #n",
    "LDA M", "LDX M", "LDY M", "STA M", "STX M", "STY M", "ADC M", "SBC M", "AND M",
    "ORA M", "EOR M", "CMP M", "CPX M", "CPY M", "INC M", "DEC M", "ASL M", "LSR M",
    "ROL M", "ROR M", "BIT M"};

bool has_imm(int op) { return op >= 26 && op <= 36; }
int byte_size(int op, int model) { return model == 1 ? (op <= 25 ? 1 : 2) : (has_imm(op) ? 2 : 1); }
int cycles(int op) {
  if (op == 21 || op == 23) return 3;
  if (op == 22 || op == 24) return 4;
  if (op >= 51 && op <= 56) return 5;
  if (op >= 37) return 3;
  return 2;
}

std::string upper(std::string s) {
  for (auto &c : s) c = (char)toupper((unsigned char)c);
  return s;
}
std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

std::string disasm(uint32_t w) {
  int op = w & 0xFFFF;
  if (op >= 58) return "OP(" + std::to_string(op) + ")";
  std::string m = kMnem[op];
  if (!has_imm(op)) return m;
  char hex[8];
  snprintf(hex, sizeof(hex), "$%02X", (w >> 16) & 0xFF);
  return m.substr(0, m.size() - 1) + hex;
}

std::string disasm_seq(const uint32_t *w, int n) {
  std::string s;
  for (int i = 0; i < n; i++) {
    if (i) s += " : ";
    s += disasm(w[i]);
  }
  return s;
}

// main.go:285-299
bool parse_dead_flags(const std::string &in, uint32_t *out) {
  std::string s = in;
  for (auto &c : s) c = (char)tolower((unsigned char)c);
  if (s == "none" || s.empty()) return *out = 0, true;
  if (s == "all") return *out = 0xFF, true;
  if (s.rfind("0x", 0) == 0) s = s.substr(2);
  if (s.empty() || s.size() > 2) return false;
  char *end;
  unsigned long v = strtoul(s.c_str(), &end, 16);
  if (*end) return false;
  *out = (uint32_t)v;
  return true;
}

// main.go:468-494
bool parse_immediate(std::string s, int *out) {
  s = trim(s);
  if (s.empty()) return false;
  int base = 10;
  if (s[0] == '$') s = s.substr(1), base = 16;
  else if (s.rfind("0X", 0) == 0 || s.rfind("0x", 0) == 0) s = s.substr(2), base = 16;
  else if (s.back() == 'H') s.pop_back(), base = 16;
  if (s.empty()) return false;
  char *end;
  unsigned long v = strtoul(s.c_str(), &end, base);
  if (*end || v > 0xFFFF) return false;
  *out = (int)v;
  return true;
}

// main.go:343-384
bool parse_single(const std::string &text_in, uint32_t *out) {
  std::string text = trim(text_in), up = upper(text);
  for (int op = 0; op < 58; op++) {
    std::string m = kMnem[op];
    if (!has_imm(op)) {
      if (upper(m) == up) return *out = (uint32_t)op, true;
      continue;
    }
    std::string pat = upper(m);
    size_t n = pat.rfind('N');
    std::string prefix = pat.substr(0, n), suffix = pat.substr(n + 1);
    if (up.rfind(prefix, 0) != 0) continue;
    if (!suffix.empty() && (up.size() < suffix.size() || up.compare(up.size() - suffix.size(), suffix.size(), suffix)))
      continue;
    std::string val = up.substr(prefix.size());
    if (!suffix.empty()) val = val.substr(0, val.size() - suffix.size());
    int v;
    if (!parse_immediate(val, &v)) continue;
    *out = (uint32_t)op | ((uint32_t)v << 16);
    return true;
  }
  return false;
}

// main.go:302-323
bool parse_assembly(const std::string &text, std::vector<uint32_t> *seq, std::string *err) {
  seq->clear();
  std::stringstream ss(text);
  std::string part;
  while (std::getline(ss, part, ':')) {
    part = trim(part);
    if (part.empty()) continue;
    uint32_t w;
    if (!parse_single(part, &w)) {
      *err = "cannot parse \"" + part + "\": unknown instruction: " + part;
      return false;
    }
    seq->push_back(w);
  }
  if (seq->empty()) {
    *err = "no instructions parsed from \"" + text + "\"";
    return false;
  }
  return true;
}

// pkg/result/output.go:77-101
std::string dead_flag_desc(int mask) {
  if (mask == 0) return "";
  if (mask == 0xFF) return "all";
  const struct { int bit; const char *n; } f[] = {{1, "C"}, {2, "Z"}, {4, "I"}, {8, "D"}, {0x40, "V"}, {0x80, "N"}};
  std::string s;
  for (auto &x : f)
    if (mask & x.bit) s += (s.empty() ? "" : ",") + std::string(x.n);
  return s;
}

std::string json_str(const std::string &s) { return "\"" + s + "\""; }

// pkg/result/output.go:21-35 with table.go:38-51 ordering.
// A rule for the JSON writer: sequences of any length (STOKE's rewrites run
// to 10 instructions; sx_rule holds 4).
struct JsonRule {
  std::vector<uint32_t> src, rep;
  int bytes_saved, cycles_saved;
  uint32_t dead_flags;
};

// result.WriteJSON (pkg/result/output.go:20-35): the rules in the order given.
void write_json(FILE *f, const std::vector<JsonRule> &rules) {
  if (rules.empty()) {
    fputs("[]\n", f);
    return;
  }
  fputs("[\n", f);
  for (size_t i = 0; i < rules.size(); i++) {
    const JsonRule &r = rules[i];
    fputs("  {\n    \"source\": [\n", f);
    for (size_t k = 0; k < r.src.size(); k++)
      fprintf(f, "      %s%s\n", json_str(disasm(r.src[k])).c_str(), k + 1 < r.src.size() ? "," : "");
    fputs("    ],\n    \"replacement\": [\n", f);
    for (size_t k = 0; k < r.rep.size(); k++)
      fprintf(f, "      %s%s\n", json_str(disasm(r.rep[k])).c_str(), k + 1 < r.rep.size() ? "," : "");
    fprintf(f, "    ],\n    \"bytes_saved\": %d,\n    \"cycles_saved\": %d", r.bytes_saved, r.cycles_saved);
    std::string d = dead_flag_desc(r.dead_flags);
    if (!d.empty()) fprintf(f, ",\n    \"dead_flags\": \"%s\"", d.c_str());
    fprintf(f, "\n  }%s\n", i + 1 < rules.size() ? "," : "");
  }
  fputs("]\n", f);
}

// An enumeration's rules as table.Rules() hands them to WriteJSON
// (pkg/result/table.go:38-52): by bytes saved, then cycles saved, descending.
void write_json(FILE *f, const std::vector<sx_rule> &rules) {
  std::vector<JsonRule> j;
  j.reserve(rules.size());
  for (const sx_rule &r : rules)
    j.push_back({std::vector<uint32_t>(r.src, r.src + r.src_len), std::vector<uint32_t>(r.rep, r.rep + r.rep_len),
                 r.bytes_saved, r.cycles_saved, r.dead_flags});
  std::stable_sort(j.begin(), j.end(), [](const JsonRule &a, const JsonRule &b) {
    if (a.bytes_saved != b.bytes_saved) return a.bytes_saved > b.bytes_saved;
    return a.cycles_saved > b.cycles_saved;
  });
  write_json(f, j);
}
// cu:912-930: one JSON object per line.  A million-line file is formatted
// on up to 16 threads into per-thread buffers (disassembly from a table of
// every opcode and immediate, integers by hand), written in order.
struct JsonlFormatter {
  char dis[58][256][8];  // NUL-terminated disassembly of (opcode, immediate)
  uint8_t dlen[58][256];
  JsonlFormatter() {
    for (int op = 0; op < 58; op++)
      for (int imm = 0; imm < 256; imm++) {
        const std::string d = disasm((uint32_t)op | ((uint32_t)imm << 16));
        memcpy(dis[op][imm], d.c_str(), d.size() + 1);
        dlen[op][imm] = (uint8_t)d.size();
      }
  }
  static char *put_int(char *p, int v) {
    if (v < 0) {
      *p++ = '-';
      v = -v;
    }
    char t[12];
    int n = 0;
    do {
      t[n++] = (char)('0' + v % 10);
      v /= 10;
    } while (v);
    while (n) *p++ = t[--n];
    return p;
  }
  static char *put(char *p, const char *s) {
    while (*s) *p++ = *s++;
    return p;
  }
  char *put_seq(char *p, const uint32_t *w, int n) const {
    for (int i = 0; i < n; i++) {
      if (i) p = put(p, " : ");
      const uint32_t op = w[i] & 0xFFFFu;
      if (op >= 58) {
        const std::string d = disasm(w[i]);
        memcpy(p, d.data(), d.size());
        p += d.size();
        continue;
      }
      const uint32_t imm = has_imm((int)op) ? (w[i] >> 16) & 0xFFu : 0u;
      memcpy(p, dis[op][imm], dlen[op][imm]);
      p += dlen[op][imm];
    }
    return p;
  }
  // one line, at most 8 x 11 + 120 characters
  char *line(char *p, const sx_rule &r, int model) const {
    int sb = 0, rb = 0;
    for (int k = 0; k < r.src_len; k++) sb += byte_size(r.src[k] & 0xFFFF, model);
    for (int k = 0; k < r.rep_len; k++) rb += byte_size(r.rep[k] & 0xFFFF, model);
    p = put(p, "{\"source_asm\":\"");
    p = put_seq(p, r.src, r.src_len);
    p = put(p, "\",\"replacement_asm\":\"");
    p = put_seq(p, r.rep, r.rep_len);
    p = put(p, "\",\"source_bytes\":");
    p = put_int(p, sb);
    p = put(p, ",\"replacement_bytes\":");
    p = put_int(p, rb);
    p = put(p, ",\"bytes_saved\":");
    p = put_int(p, r.bytes_saved);
    p = put(p, ",\"cycles_saved\":");
    p = put_int(p, r.cycles_saved);
    if (r.dead_flags) {
      static const char hx[] = "0123456789ABCDEF";
      p = put(p, ",\"dead_flags\":\"0x");
      *p++ = hx[(r.dead_flags >> 4) & 15];
      *p++ = hx[r.dead_flags & 15];
      *p++ = '"';
    }
    p = put(p, "}\n");
    return p;
  }
};

bool write_jsonl(FILE *f, const sx_rule *rules, size_t n, int model) {
  static const JsonlFormatter fmt;
  constexpr size_t kLineMax = 256;
  const unsigned nth = n < 65536 ? 1u : std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  std::vector<std::vector<char>> bufs(nth);
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nth; t++)
    pool.emplace_back([&, t] {
      const size_t a = n * t / nth, b = n * (t + 1) / nth;
      std::vector<char> &buf = bufs[t];
      buf.resize((b - a) * kLineMax);
      char *p = buf.data();
      for (size_t i = a; i < b; i++) p = fmt.line(p, rules[i], model);
      buf.resize((size_t)(p - buf.data()));
    });
  for (auto &th : pool) th.join();
  for (const auto &buf : bufs)
    if (!buf.empty() && fwrite(buf.data(), 1, buf.size(), f) != buf.size()) return false;
  return true;
}

struct Args {
  std::vector<std::string> pos;
  std::vector<std::pair<std::string, std::string>> kv;
  bool has(const std::string &k) const {
    for (auto &p : kv)
      if (p.first == k) return true;
    return false;
  }
  std::string get(const std::string &k, const std::string &def) const {
    for (auto &p : kv)
      if (p.first == k) return p.second;
    return def;
  }
};

Args parse_args(int argc, char **argv, int from, const std::vector<std::string> &bool_flags) {
  Args a;
  for (int i = from; i < argc; i++) {
    std::string s = argv[i];
    if (s.rfind("-", 0) == 0 && s.size() > 1) {
      std::string k = s.substr(s.find_first_not_of('-'));
      std::string v;
      size_t eq = k.find('=');
      if (eq != std::string::npos) {
        v = k.substr(eq + 1);
        k = k.substr(0, eq);
      } else if (std::find(bool_flags.begin(), bool_flags.end(), k) != bool_flags.end()) {
        v = "true";
      } else if (i + 1 < argc) {
        v = argv[++i];
      }
      if (k == "v") k = "verbose";
      a.kv.push_back({k, v});
    } else {
      a.pos.push_back(s);
    }
  }
  return a;
}

int die(const std::string &msg) {
  fprintf(stderr, "Error: %s\n", msg.c_str());
  return 1;
}

int check(sx_handle *h, int rc, const char *what) {
  if (rc == SX_OK) return 0;
  char buf[512] = "";
  if (h) sx_last_error(h, buf, sizeof(buf));
  fprintf(stderr, "Error: %s failed (%d): %s\n", what, rc, buf);
  return 1;
}

// Checkpoint of a chunked enumeration (the role of pkg/result/checkpoint.go:
// 9-37, as JSON text): the next unprocessed index of the enumeration's
// target space, the range end, the rules and the JSONL bytes written and
// committed so far, and the whole configuration.  Per chunk the JSONL is
// written and flushed first, then the checkpoint is replaced atomically; a
// resume truncates the JSONL to the committed offset, so a crash between the
// two writes (or in the middle of either) neither duplicates nor loses rules.
struct Checkpoint {
  uint64_t next = 0, end = 0, rules = 0, offset = 0;
  int max_target = 2, bytes_model = 0, hits = 0, exh = 0, use_mid = 1;
  uint32_t dead = 0;
  bool same_config(const Checkpoint &o) const {
    return max_target == o.max_target && dead == o.dead && bytes_model == o.bytes_model && hits == o.hits &&
           exh == o.exh && use_mid == o.use_mid && end == o.end;
  }
};

bool read_checkpoint(const std::string &path, Checkpoint *c) {
  std::ifstream in(path);
  if (!in) return false;
  std::stringstream ss;
  ss << in.rdbuf();
  std::string t = ss.str();
  auto num = [&](const char *key, uint64_t *v) {
    size_t p = t.find(std::string("\"") + key + "\"");
    if (p == std::string::npos) return false;
    p = t.find(':', p);
    if (p == std::string::npos) return false;
    *v = strtoull(t.c_str() + p + 1, nullptr, 10);
    return true;
  };
  uint64_t mt = 0, dd = 0, bm = 0, hi = 0, ex = 0, um = 0;
  if (!num("next", &c->next) || !num("end", &c->end) || !num("rules", &c->rules) || !num("offset", &c->offset) ||
      !num("max_target", &mt) || !num("dead_flags", &dd) || !num("bytes_model", &bm) || !num("hits", &hi) ||
      !num("exh", &ex) || !num("use_mid", &um))
    return false;
  c->max_target = (int)mt;
  c->dead = (uint32_t)dd;
  c->bytes_model = (int)bm;
  c->hits = (int)hi;
  c->exh = (int)ex;
  c->use_mid = (int)um;
  return true;
}

bool write_checkpoint(const std::string &path, const Checkpoint &c) {
  std::string tmp = path + ".tmp";
  FILE *f = fopen(tmp.c_str(), "w");
  if (!f) return false;
  fprintf(f,
          "{\"next\": %llu, \"end\": %llu, \"rules\": %llu, \"offset\": %llu, \"max_target\": %d, "
          "\"dead_flags\": %u, \"bytes_model\": %d, \"hits\": %d, \"exh\": %d, \"use_mid\": %d}\n",
          (unsigned long long)c.next, (unsigned long long)c.end, (unsigned long long)c.rules,
          (unsigned long long)c.offset, c.max_target, c.dead, c.bytes_model, c.hits, c.exh, c.use_mid);
  if (fflush(f) != 0 || fsync(fileno(f)) != 0) {
    fclose(f);
    return false;
  }
  fclose(f);
  return rename(tmp.c_str(), path.c_str()) == 0;  // the checkpoint is replaced atomically
}

// search.go:26-50: --max-target L enumerates every target length 2..L, each
// against candidates of length 1..L-1.  The lengths' target spaces are laid
// end to end as one index space: the pruned length-2 targets (positions in Go
// order), then for L = 3 the raw Go-order triple indices.
struct Phase {
  int target_len;
  uint64_t base, size;  // global indices [base, base + size)
};

int cmd_enumerate(int argc, char **argv) {
  Args a = parse_args(argc, argv, 2, {"verbose", "no-mid", "resume"});
  int max_target = std::atoi(a.get("max-target", "2").c_str());
  uint32_t dead;
  if (!parse_dead_flags(a.get("dead-flags", "none"), &dead))
    return die("invalid --dead-flags value: use none, all, or hex (e.g. 0xFF)");
  int model = a.get("bytes-model", "go") == "cuda" ? SX_BYTES_CUDA : SX_BYTES_GO;
  int hits = a.get("hits", "first") == "all" ? SX_HITS_ALL : SX_HITS_FIRST;
  int exh = a.get("exh", "projected") == "full" ? SX_EXH_FULL : SX_EXH_PROJECTED;
  int device = std::atoi(a.get("device", "0").c_str());
  if (max_target != 2 && max_target != 3) return die("this build enumerates targets of length 2 or 3 (--max-target)");
  std::string out = a.get("output", ""), jl = a.get("jsonl", ""), ckpt = a.get("checkpoint", "");
  const bool resume = a.has("resume");
  if (!ckpt.empty() && (jl.empty() || jl == "-" || !out.empty()))
    return die("--checkpoint appends to a --jsonl file (and takes no --output)");
  printf("6502 Superoptimizer\n");
  printf("  Max target length: %d\n", max_target);
  printf("  Instructions: %d per position\n", 2863);
  printf("  Device: %d (MI355X search core)\n", device);
  if (dead) printf("  Dead flags: 0x%02X (%s)\n", dead, dead_flag_desc((int)dead).c_str());
  printf("\n");
  sx_handle *h = nullptr;
  if (check(nullptr, sx_open(device, &h), "sx_open")) return 1;
  sx_enum_cfg cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.struct_size = sizeof(cfg);
  cfg.max_cand_len = max_target - 1;  // enumerator.go: candidates shorter than the longest target
  cfg.bytes_model = model;
  cfg.hits = hits;
  cfg.exh_mode = exh;
  cfg.use_mid = a.has("no-mid") ? 0 : 1;
  cfg.dead_flags = dead;
  uint64_t n2 = 0;
  if (check(h, sx_pruned_targets(h, 0, 0, nullptr, &n2), "sx_pruned_targets")) return 1;
  std::vector<Phase> phases{{2, 0, n2}};
  if (max_target == 3) phases.push_back({3, n2, 2863ull * 2863ull * 2863ull});
  const uint64_t total = phases.back().base + phases.back().size;
  uint64_t first = strtoull(a.get("first-target", "0").c_str(), nullptr, 10);
  uint64_t end = strtoull(a.get("end-target", "0").c_str(), nullptr, 10);
  if (end == 0) end = total;
  if (first > end || end > total) return die("target range out of bounds");
  Checkpoint ck;
  ck.next = first;
  ck.end = end;
  ck.max_target = max_target;
  ck.dead = dead;
  ck.bytes_model = model;
  ck.hits = hits;
  ck.exh = exh;
  ck.use_mid = cfg.use_mid;
  FILE *jf = nullptr;
  if (resume) {
    if (ckpt.empty()) return die("--resume needs --checkpoint");
    Checkpoint got;
    if (!read_checkpoint(ckpt, &got)) return die("cannot read checkpoint " + ckpt);
    if (!got.same_config(ck)) return die("checkpoint is for another configuration");
    ck = got;
    // drop whatever follows the last committed chunk, then append
    jf = fopen(jl.c_str(), "r+");
    if (!jf || ftruncate(fileno(jf), (off_t)ck.offset) != 0 || fseek(jf, 0, SEEK_END) != 0)
      return die("cannot reopen " + jl + " at its committed size");
    printf("  Resuming at target %llu of [.., %llu), %llu rules so far\n", (unsigned long long)ck.next,
           (unsigned long long)ck.end, (unsigned long long)ck.rules);
  } else if (!jl.empty()) {
    jf = jl == "-" ? stdout : fopen(jl.c_str(), "w");
    if (!jf) return die("cannot create " + jl);
    if (!ckpt.empty() && !write_checkpoint(ckpt, ck)) return die("cannot write checkpoint " + ckpt);
  }
  // chunks: one per phase unless checkpointing
  const uint64_t chunk = ckpt.empty() ? total
                                      : strtoull(a.get("chunk", max_target == 2 ? "1048576" : "16777216").c_str(),
                                                 nullptr, 10);
  // test hook: exit after this many chunks' JSONL writes, before their checkpoint
  const char *crash_env = getenv("SX_TEST_EXIT_BEFORE_CHECKPOINT");
  const long crash_after = crash_env ? atol(crash_env) : 0;
  long chunks_done = 0;
  std::vector<sx_rule> all;
  sx_stats tot;
  memset(&tot, 0, sizeof(tot));
  auto t0 = std::chrono::steady_clock::now();
  while (ck.next < ck.end) {
    const Phase *ph = &phases[0];
    for (const Phase &p : phases)
      if (ck.next >= p.base && ck.next < p.base + p.size) ph = &p;
    const uint64_t stop = std::min<uint64_t>({ck.end, ph->base + ph->size, ck.next + std::max<uint64_t>(chunk, 1)});
    cfg.target_len = ph->target_len;
    cfg.first_target = ck.next - ph->base;
    cfg.end_target = stop - ph->base;
    uint64_t n = 0;
    sx_stats st;
    if (check(h, sx_enumerate(h, &cfg, nullptr, 0, &n, &st), "sx_enumerate")) return 1;
    // (left uninitialised: sx_fetch_rules writes every byte of every record)
    std::unique_ptr<sx_rule[]> rules(new sx_rule[std::max<uint64_t>(n, 1)]);
    uint64_t got = 0;
    if (check(h, sx_fetch_rules(h, rules.get(), 0, n, &got), "sx_fetch_rules")) return 1;
    tot.targets += st.targets;
    tot.qc_pairs += st.qc_pairs;
    tot.qc_hits += st.qc_hits;
    tot.mid_pass += st.mid_pass;
    tot.exh_pairs += st.exh_pairs;
    if (jf) {
      if (!write_jsonl(jf, rules.get(), got, model) || fflush(jf) != 0) return die("cannot write " + jl);
    }
    ck.rules += n;
    ck.next = stop;
    if (!ckpt.empty()) {
      if (fsync(fileno(jf)) != 0) return die("cannot sync " + jl);
      ck.offset = (uint64_t)ftello(jf);
      if (crash_after > 0 && ++chunks_done >= crash_after) exit(3);
      if (!write_checkpoint(ckpt, ck)) return die("cannot write checkpoint " + ckpt);
      double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      fprintf(stderr, "  [%.0fs] %llu/%llu (%.2f%%) | %llu rules\n", secs, (unsigned long long)ck.next,
              (unsigned long long)ck.end, 100.0 * (double)ck.next / (double)std::max<uint64_t>(ck.end, 1),
              (unsigned long long)ck.rules);
    }
    if (out.size() || a.has("verbose")) all.insert(all.end(), rules.get(), rules.get() + got);
  }
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  printf("  %llu targets | %llu QC pairs | QC hits %llu | Mid %llu | Exh jobs %llu | %.2f s (%.3g pairs/s)\n",
         (unsigned long long)tot.targets, (unsigned long long)tot.qc_pairs, (unsigned long long)tot.qc_hits,
         (unsigned long long)tot.mid_pass, (unsigned long long)tot.exh_pairs, secs, tot.qc_pairs / secs);
  if (a.has("verbose"))
    for (auto &r : all)
      printf("  FOUND%s: %s -> %s (-%d bytes, -%d cycles)\n", r.pass == 2 ? " (dead flags)" : "",
             disasm_seq(r.src, r.src_len).c_str(), disasm_seq(r.rep, r.rep_len).c_str(), r.bytes_saved,
             r.cycles_saved);
  printf("\nFound %llu optimizations\n", (unsigned long long)ck.rules);
  if (!out.empty()) {
    FILE *f = fopen(out.c_str(), "w");
    if (!f) return die("cannot create " + out);
    write_json(f, all);
    fclose(f);
    printf("Written to %s\n", out.c_str());
  }
  if (jf && jf != stdout) fclose(jf);
  sx_close(h);
  return 0;
}

int cmd_target(int argc, char **argv) {
  Args a = parse_args(argc, argv, 2, {"verbose"});
  std::string input;
  for (auto &p : a.pos) input += (input.empty() ? "" : " ") + p;
  if (input.empty()) return die("requires at least 1 arg(s)");
  uint32_t dead;
  if (!parse_dead_flags(a.get("dead-flags", "none"), &dead)) return die("invalid --dead-flags value");
  std::vector<uint32_t> seq;
  std::string err;
  if (!parse_assembly(input, &seq, &err)) return die("failed to parse: " + err);
  int tb = 0, tc = 0;
  for (uint32_t w : seq) tb += byte_size(w & 0xFFFF, 0), tc += cycles(w & 0xFFFF);
  printf("Target: %s (%d bytes, %d cycles)\n", input.c_str(), tb, tc);
  int max_cand = std::atoi(a.get("max-candidate", "4").c_str());
  if (seq.empty() || seq.size() > 8) return die("target must have 1..8 instructions");
  if (max_cand < 1 || max_cand > 4) return die("--max-candidate must be 1..4");
  sx_handle *h = nullptr;
  if (check(nullptr, sx_open(std::atoi(a.get("device", "0").c_str()), &h), "sx_open")) return 1;
  sx_target_result res[2];
  uint64_t scanned = 0;
  auto t0 = std::chrono::steady_clock::now();
  if (check(h, sx_target_search(h, seq.data(), (int)seq.size(), max_cand, dead, 0, res, &scanned),
            "sx_target_search"))
    return 1;
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  sx_close(h);
  if (a.has("verbose")) {
    for (const sx_target_result &r : res)
      if (r.found) {
        if (r.pass == 2)
          printf("  FOUND (dead flags 0x%02X): %s -> %s (-%d bytes, -%d cycles)\n", r.dead_flags, input.c_str(),
                 disasm_seq(r.rep, r.rep_len).c_str(), r.bytes_saved, r.cycles_saved);
        else
          printf("  FOUND: %s -> %s (-%d bytes, -%d cycles)\n", input.c_str(), disasm_seq(r.rep, r.rep_len).c_str(),
                 r.bytes_saved, r.cycles_saved);
      }
    fprintf(stderr, "  %llu candidates in %.3f s\n", (unsigned long long)scanned, secs);
  }
  // Table.Rules() order (pkg/result/table.go:40-50): bytes saved, then cycles saved, descending
  const sx_target_result *best = nullptr;
  for (const sx_target_result &r : res) {
    if (!r.found) continue;
    if (!best || r.bytes_saved > best->bytes_saved ||
        (r.bytes_saved == best->bytes_saved && r.cycles_saved > best->cycles_saved))
      best = &r;
  }
  if (!best) {
    printf("No shorter replacement found.\n");
    return 0;
  }
  printf("Replacement: %s (-%d bytes, -%d cycles)\n", disasm_seq(best->rep, best->rep_len).c_str(),
         best->bytes_saved, best->cycles_saved);
  return 0;
}

// stoke (main.go:196-275): chains on the device, every rewrite verified by
// the complete ExhaustiveCheck before it is reported.
int cmd_stoke(int argc, char **argv) {
  Args a = parse_args(argc, argv, 2, {"verbose"});
  std::string target = a.get("target", "");
  if (target.empty()) return die("--target is required");
  std::vector<uint32_t> seq;
  std::string err;
  if (!parse_assembly(target, &seq, &err)) return die("failed to parse target: " + err);
  if (seq.size() > 8) return die("target must have at most 8 instructions");
  uint32_t dead;
  if (!parse_dead_flags(a.get("dead-flags", "none"), &dead)) return die("invalid --dead-flags value");
  sx_stoke_cfg cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.struct_size = sizeof(cfg);
  cfg.chains_per_target = (uint32_t)std::atoi(a.get("chains", "16").c_str());
  cfg.steps = strtoull(a.get("iterations", "10000000").c_str(), nullptr, 10);
  cfg.decay = std::atof(a.get("decay", "0.9999").c_str());
  cfg.t0 = 1.0;
  cfg.seed = strtoull(a.get("seed", "0").c_str(), nullptr, 10);
  cfg.dead_flags = dead;
  uint32_t tw[8] = {0};
  for (size_t k = 0; k < seq.size(); k++) tw[k] = seq[k];
  uint8_t tl = (uint8_t)seq.size();
  sx_handle *h = nullptr;
  if (check(nullptr, sx_open(std::atoi(a.get("device", "0").c_str()), &h), "sx_open")) return 1;
  // stoke.Run then stoke.Deduplicate (main.go:231-232)
  std::vector<sx_stoke_found> found(4096);
  uint64_t nf = 0;
  auto t0 = std::chrono::steady_clock::now();
  if (check(h, sx_stoke_collect(h, &cfg, tw, &tl, 1, found.data(), found.size(), &nf, nullptr), "sx_stoke_collect"))
    return 1;
  if (nf > found.size()) {
    found.resize(nf);
    if (check(h, sx_stoke_collect(h, &cfg, tw, &tl, 1, found.data(), found.size(), &nf, nullptr),
              "sx_stoke_collect"))
      return 1;
  }
  found.resize(std::min<uint64_t>(nf, found.size()));
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  sx_close(h);
  if (a.has("verbose")) {
    fprintf(stderr, "  %u chains x %llu iterations, %.3f s\n", cfg.chains_per_target,
            (unsigned long long)cfg.steps, secs);
    for (const sx_stoke_found &r : found)
      fprintf(stderr, "  Chain %u @ iter %llu: %s VERIFIED\n", r.chain, (unsigned long long)r.step,
              disasm_seq(r.rep, r.rep_len).c_str());
  }
  printf("\n%zu unique optimizations found\n", found.size());
  std::vector<JsonRule> rules;
  for (size_t i = 0; i < found.size(); i++) {
    const sx_stoke_found &r = found[i];
    printf("  %zu. %s", i + 1, disasm_seq(r.rep, r.rep_len).c_str());
    if (r.dead_flags)
      printf(" (-%d bytes, -%d cycles, dead flags: %s)\n", r.bytes_saved, r.cycles_saved,
             dead_flag_desc(r.dead_flags).c_str());
    else
      printf(" (-%d bytes, -%d cycles)\n", r.bytes_saved, r.cycles_saved);
    rules.push_back({seq, std::vector<uint32_t>(r.rep, r.rep + r.rep_len), r.bytes_saved, r.cycles_saved,
                     r.dead_flags});
  }
  std::string outp = a.get("output", "");
  if (!outp.empty() && !rules.empty()) {
    FILE *f = fopen(outp.c_str(), "w");
    if (!f) return die("cannot create " + outp);
    write_json(f, rules);
    fclose(f);
    printf("Written to %s\n", outp.c_str());
  }
  return 0;
}

// Extract a JSON string field from one flat JSONL object.
bool json_field(const std::string &line, const std::string &key, std::string *out) {
  std::string pat = "\"" + key + "\"";
  size_t p = line.find(pat);
  if (p == std::string::npos) return false;
  p = line.find(':', p + pat.size());
  if (p == std::string::npos) return false;
  p = line.find('"', p);
  if (p == std::string::npos) return false;
  std::string v;
  for (size_t i = p + 1; i < line.size(); i++) {
    if (line[i] == '\\' && i + 1 < line.size()) {
      const char e = line[++i];
      if (e == 'u' && i + 4 < line.size()) {  // \uXXXX (Go's encoder writes <, > and & so)
        const unsigned long c = strtoul(line.substr(i + 1, 4).c_str(), nullptr, 16);
        v += c < 0x80 ? (char)c : '?';
        i += 4;
      } else {
        v += e == 'n' ? '\n' : e == 't' ? '\t' : e == 'r' ? '\r' : e == 'b' ? '\b' : e == 'f' ? '\f' : e;
      }
      continue;
    }
    if (line[i] == '"') {
      *out = v;
      return true;
    }
    v += line[i];
  }
  return false;
}

// main.go:386-463, with the checks batched into one device call.
int cmd_verify_jsonl(int argc, char **argv) {
  Args a = parse_args(argc, argv, 2, {"verbose"});
  if (a.pos.size() != 1) return die("accepts 1 arg(s)");
  uint32_t dead;
  if (!parse_dead_flags(a.get("dead-flags", "none"), &dead)) return die("invalid --dead-flags value");
  std::ifstream in(a.pos[0]);
  if (!in) return die("open " + a.pos[0] + ": no such file or directory");
  std::vector<uint32_t> tw, cw;
  std::vector<uint8_t> tl, cl;
  std::vector<std::string> desc;
  std::vector<int> lineno;
  std::string line;
  int total = 0, skipped = 0;
  while (std::getline(in, line)) {
    line = trim(line);
    if (line.empty()) continue;
    total++;
    std::string s, r, err;
    // an object without a field decodes it as "" (Go's zero value), which then fails to parse
    if (line.front() != '{' || line.back() != '}') {
      fprintf(stderr, "  [%d] JSON parse error\n", total);
      skipped++;
      continue;
    }
    if (!json_field(line, "source_asm", &s)) s.clear();
    if (!json_field(line, "replacement_asm", &r)) r.clear();
    std::vector<uint32_t> src, rep;
    if (!parse_assembly(s, &src, &err)) {
      fprintf(stderr, "  [%d] Cannot parse source \"%s\": %s\n", total, s.c_str(), err.c_str());
      skipped++;
      continue;
    }
    if (!parse_assembly(r, &rep, &err)) {
      fprintf(stderr, "  [%d] Cannot parse replacement \"%s\": %s\n", total, r.c_str(), err.c_str());
      skipped++;
      continue;
    }
    if (src.size() > 4 || rep.size() > 4) {
      fprintf(stderr, "  [%d] sequence longer than 4 instructions\n", total);
      skipped++;
      continue;
    }
    for (int k = 0; k < 4; k++) {
      tw.push_back(k < (int)src.size() ? src[k] : 0);
      cw.push_back(k < (int)rep.size() ? rep[k] : 0);
    }
    tl.push_back((uint8_t)src.size());
    cl.push_back((uint8_t)rep.size());
    desc.push_back(s + " -> " + r);
    lineno.push_back(total);
  }
  std::vector<uint8_t> ok(tl.size(), 0);
  if (!tl.empty()) {
    sx_handle *h = nullptr;
    if (check(nullptr, sx_open(std::atoi(a.get("device", "0").c_str()), &h), "sx_open")) return 1;
    if (check(h, sx_check_pairs(h, tw.data(), tl.data(), cw.data(), cl.data(), tl.size(), SX_STAGE_EXH, dead,
                                SX_EXH_PROJECTED, ok.data()),
              "sx_check_pairs"))
      return 1;
    sx_close(h);
  }
  int passed = 0, failed = 0;
  for (size_t i = 0; i < tl.size(); i++) {
    if (ok[i] & SX_STAGE_EXH) {
      passed++;
      if (a.has("verbose")) printf("  [%d] PASS: %s\n", lineno[i], desc[i].c_str());
    } else {
      failed++;
      printf("  [%d] FAIL: %s\n", lineno[i], desc[i].c_str());
    }
  }
  printf("\nVerification complete: %d total, %d passed, %d failed, %d skipped\n", total, passed, failed, skipped);
  if (failed) {
    fprintf(stderr, "Error: %d rules failed verification\n", failed);
    return 1;
  }
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: 6502opt enumerate|target|verify-jsonl|stoke [flags]\n");
    return 1;
  }
  std::string cmd = argv[1];
  if (cmd == "enumerate") return cmd_enumerate(argc, argv);
  if (cmd == "target") return cmd_target(argc, argv);
  if (cmd == "verify-jsonl") return cmd_verify_jsonl(argc, argv);
  if (cmd == "stoke") return cmd_stoke(argc, argv);
  fprintf(stderr, "Error: unknown command \"%s\" for \"6502opt\"\n", cmd.c_str());
  return 1;
}
