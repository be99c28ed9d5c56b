// Test driver for the drop-in layer's S/BLOW5 reader and writer
// (include/odba/odba_slow5.hpp); it needs no GPU.
//
//   slow5_tool dump <file>                 one line per read: id, length,
//                                          FNV-1a hash of the int16 signal,
//                                          first eight samples
//   slow5_tool rewrite <in> <out> <k>      writeSlow5Output of every read of
//                                          <in> with sample j replaced by
//                                          sample j + (j % k) - k / 2, then
//                                          reads <out> back and checks it;
//                                          prints "ok" and the read count
//   slow5_tool rewrite_named <in> <out> <name> <length>
//                                          writeSlow5Output of one named read
//                                          with a zero signal of the given
//                                          length; prints the return code
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/odba/odba_slow5.hpp"

static uint64_t fnv(const std::vector<int16_t>& s) {
  uint64_t h = 1469598103934665603ull;
  for (int16_t v : s) {
    const uint16_t u = (uint16_t)v;
    for (int k = 0; k < 2; k++) {
      h ^= (uint8_t)(u >> (8 * k));
      h *= 1099511628211ull;
    }
  }
  return h;
}

int main(int argc, char** argv) {
  if (argc >= 3 && !std::strcmp(argv[1], "dump")) {
    odba_slow5::File f;
    if (odba_slow5::load(argv[2], f)) {
      std::fprintf(stderr, "unreadable: %s\n", f.err.c_str());
      return SLOW5_FILE_UNREADABLE;
    }
    for (const odba_slow5::Record& r : f.recs) {
      std::printf("%s\t%zu\t%016llx", r.id.c_str(), r.sig.size(), (unsigned long long)fnv(r.sig));
      for (size_t j = 0; j < r.sig.size() && j < 8; j++) std::printf("%c%d", j ? ',' : '\t', r.sig[j]);
      std::printf("\n");
    }
    return 0;
  }
  if (argc >= 5 && !std::strcmp(argv[1], "rewrite")) {
    const int k = std::atoi(argv[4]);
    odba_slow5::File f;
    if (odba_slow5::load(argv[2], f) || k < 1) return 2;
    const int n = (int)f.recs.size();
    std::vector<std::vector<short>> seqs(n);
    std::vector<short*> sp(n);
    std::vector<char*> names(n);
    std::vector<size_t> lens(n);
    for (int i = 0; i < n; i++) {
      const odba_slow5::Record& r = f.recs[i];
      for (size_t j = 0; j < r.sig.size(); j++) seqs[i].push_back((short)(r.sig[j] + (int)(j % k) - k / 2));
      sp[i] = seqs[i].data();
      names[i] = const_cast<char*>(r.id.c_str());
      lens[i] = r.sig.size();
    }
    // every read in reverse order, so the copy's order follows the list
    std::vector<char*> rn(names.rbegin(), names.rend());
    std::vector<short*> rs(sp.rbegin(), sp.rend());
    std::vector<size_t> rl(lens.rbegin(), lens.rend());
    if (writeSlow5Output(argv[2], argv[3], rn.data(), rs.data(), rl.data(), n) != 0) return 3;
    odba_slow5::File g;
    if (odba_slow5::load(argv[3], g) || (int)g.recs.size() != n) return 4;
    if (g.header != f.header || g.rec_press != f.rec_press || g.sig_press != f.sig_press) return 5;
    for (int i = 0; i < n; i++) {
      const odba_slow5::Record& a = f.recs[n - 1 - i];
      const odba_slow5::Record& b = g.recs[i];
      if (a.id != b.id || a.head != b.head || a.aux != b.aux) return 6;
      if (b.sig.size() != seqs[n - 1 - i].size()) return 7;
      for (size_t j = 0; j < b.sig.size(); j++)
        if (b.sig[j] != seqs[n - 1 - i][j]) return 8;
    }
    std::printf("ok %d\n", n);
    return 0;
  }
  if (argc >= 6 && !std::strcmp(argv[1], "rewrite_named")) {
    char* name = argv[4];
    size_t len = (size_t)std::strtoull(argv[5], nullptr, 10);
    std::vector<short> z(len ? len : 1, 0);
    short* zp = z.data();
    std::printf("%d\n", writeSlow5Output(argv[2], argv[3], &name, &zp, &len, 1));
    return 0;
  }
  std::fprintf(stderr, "usage: slow5_tool dump <file> | rewrite <in> <out> <k> | rewrite_named <in> <out> <name> <len>\n");
  return 1;
}
