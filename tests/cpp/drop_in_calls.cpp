// A reference-style caller of the drop-in header: every template entry
// point called with the reference's own argument lists (the setupAndRun
// call of tests/openDBA_test.cu:79, and the calls performDBA and
// setupAndRun make: dba.hpp:909,928-929,1087-1088, openDBA.cuh:80,87,
// gpu_utils.hpp:133,173, segmentation.hpp:342, io_utils.hpp:160,244).
// It compiles unchanged against include/odba/opendba.hpp; the tests run it
// on a device and check what it writes against the oracle.
//
//   drop_in_calls kat <output prefix> <open_end> <text files...>
//   drop_in_calls calls <output dir> <text files...>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/odba/opendba.hpp"

namespace {

char* dup(const std::string& s) {
  char* d = (char*)std::malloc(s.size() + 1);
  std::strcpy(d, s.c_str());
  return d;
}

char** deep_copy(char** a, int n) {
  char** b = (char**)std::malloc(sizeof(char*) * n);
  for (int i = 0; i < n; i++) b[i] = dup(a[i]);
  return b;
}

template <typename T> void dump(const std::string& path, const T* v, size_t n) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) std::exit(2);
  if (n) std::fwrite(v, sizeof(T), n, f);
  std::fclose(f);
}

int kat(char* prefix, int open_end, char** files, int n) {
  char* seqprefix_filename = 0;
  char min_segment_length[] = "0";
  int use_open_start = 0, norm_sequences = 0, read_mode = TEXT_READ_MODE;
  setupAndRun<float>(seqprefix_filename, deep_copy(files, n), n, prefix, read_mode, use_open_start, open_end,
                     min_segment_length, norm_sequences, 1.0);
  return 0;
}

int calls(const std::string& dir, char** files, int n) {
  hipStream_t stream = 0;
  float** seqs = nullptr;
  char** names = nullptr;
  size_t* lens = nullptr;
  int got = readSequenceTextFiles<float>(files, n, &seqs, &names, &lens);
  if (got != n) return 3;
  // normalizeSequences, both overloads (gpu_utils.hpp:133, 173)
  std::vector<double> mu(n), sigma(n);
  normalizeSequences<float>(seqs, (size_t)n, lens, -1, mu.data(), sigma.data(), stream);
  dump(dir + "/znorm_mu.bin", mu.data(), n);
  dump(dir + "/znorm_sigma.bin", sigma.data(), n);
  for (int i = 0; i < n; i++) dump(dir + "/znorm_" + std::to_string(i) + ".bin", seqs[i], lens[i]);
  float** again = nullptr;
  char** again_names = nullptr;
  size_t* again_lens = nullptr;
  readSequenceTextFiles<float>(files, n, &again, &again_names, &again_lens);
  normalizeSequences<float>(again, (size_t)n, again_lens, -1, stream);
  for (int i = 0; i < n; i++) dump(dir + "/znorm5_" + std::to_string(i) + ".bin", again[i], again_lens[i]);
  // approximateMedoidIndices on the evenly spaced copy performDBA makes
  // (dba.hpp:919-929); the inputs are in file order, which the test sorts
  size_t maxLength = 0;
  for (int i = 0; i < n; i++) maxLength = std::max(maxLength, lens[i]);
  std::vector<float> gpu_sequences((size_t)n * maxLength, 0.0f);
  for (int i = 0; i < n; i++) std::memcpy(&gpu_sequences[(size_t)i * maxLength], seqs[i], lens[i] * sizeof(float));
  double cdist = 1.0;
  std::vector<int> memberships(n);
  char* out_prefix = dup(dir + "/medoid");
  int* medoidIndices = approximateMedoidIndices<float>(gpu_sequences.data(), maxLength, (size_t)n, lens, names, 0, 0,
                                                       out_prefix, &cdist, memberships.data(), stream);
  dump(dir + "/medoid.bin", medoidIndices, 1);
  // DBAUpdate with the medoid as the centroid (dba.hpp:1087-1088)
  const int m = medoidIndices[0];
  std::vector<float> updated(lens[m]);
  double delta = DBAUpdate<float>(seqs[m], lens[m], seqs, names, (size_t)n, lens, 0, 0, updated.data(),
                                  dir + "/round", stream);
  dump(dir + "/dba_update.bin", updated.data(), updated.size());
  dump(dir + "/dba_delta.bin", &delta, 1);
  // adaptive_segmentation (openDBA.cuh:87)
  float** seg = nullptr;
  size_t* seglens = nullptr;
  adaptive_segmentation<float>(seqs, lens, n, 8, &seg, &seglens, 0);
  adaptive_segmentation<float>(seqs, lens, n, 8, &seg, &seglens, 0, stream);
  for (int i = 0; i < n; i++) dump(dir + "/kseg_" + std::to_string(i) + ".bin", seg[i], seglens[i]);
  // writeSequences (io_utils.hpp:244)
  writeSequences(seg, seglens, names, n, (dir + "/segmented_seqs.txt").c_str());
  // chopPrefixFromSequences with the first 40 values of sequence 0 as the
  // prefix (openDBA.cuh:80)
  std::vector<float> prefix(seqs[0], seqs[0] + 40);
  int num = n;
  char* chop_prefix = dup(dir + "/chop");
  chopPrefixFromSequences<float>(prefix.data(), prefix.size(), seqs, &num, lens, names, chop_prefix, 0, stream);
  // performDBA: clusters with cdist 2, then a consensus-only call reading
  // the memberships back (dba.hpp:928-936)
  float** pseqs = nullptr;
  char** pnames = nullptr;
  size_t* plens = nullptr;
  readSequenceTextFiles<float>(files, n, &pseqs, &pnames, &plens);
  for (int i = 0; i < n; i++) { char* z = std::strchr(pnames[i], '.'); if (z) *z = '\0'; }
  char* perf_prefix = dup(dir + "/perf");
  performDBA<float>(pseqs, n, plens, pnames, 0, 0, perf_prefix, 1, 2.0, files, n, TEXT_READ_MODE, false, CLUSTER_ONLY,
                    stream);
  performDBA<float>(pseqs, n, plens, pnames, 0, 0, perf_prefix, 1, 2.0, files, n, TEXT_READ_MODE, false,
                    CONSENSUS_ONLY);
  // readMedoidIndices on the membership file performDBA wrote (io_utils.hpp:160)
  std::vector<int> memb(n);
  int* med2 = readMedoidIndices((dir + "/perf.cluster_membership.txt").c_str(), n, pnames, memb.data());
  dump(dir + "/read_membership.bin", memb.data(), n);
  dump(dir + "/read_medoids.bin", med2, 2);
  delete[] med2;
  delete[] medoidIndices;
  std::printf("calls ok\n");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc >= 5 && !std::strcmp(argv[1], "kat")) return kat(argv[2], std::atoi(argv[3]), argv + 4, argc - 4);
  if (argc >= 4 && !std::strcmp(argv[1], "calls")) return calls(argv[2], argv + 3, argc - 3);
  std::fprintf(stderr, "usage: drop_in_calls kat <prefix> <open_end> files... | calls <dir> files...\n");
  return 1;
}
