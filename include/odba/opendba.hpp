/*
 * Drop-in C++ host layer over libodba_hip: the reference's templated entry
 * points (same names, argument meaning, output files and exit codes) with
 * the DTW / DBA work done by the HIP library through the C ABI (odba.h).
 *
 *   readSequenceTextFiles / readSequenceBinaryFiles / readSequenceTSVFiles
 *                                     cpu_utils.hpp:60-134, 381-456, 553-624
 *   normalizeSequences<T>             gpu_utils.hpp:131-175
 *   approximateMedoidIndices<T>       dba.hpp:36-365
 *   DBAUpdate<T>                      dba.hpp:432-848
 *   performDBA<T>                     dba.hpp:863-1189
 *   chopPrefixFromSequences<T>        dba.hpp:1192-1395
 *   adaptive_segmentation<T>          segmentation.hpp:339-583
 *   setupAndRun<T>                    openDBA.cuh:13-200
 *
 * Differences of form: sequences live in host memory (malloc) instead of
 * CUDA managed memory, and the device is the library's (device 0 unless
 * ODBA_DEVICE is set). Errors exit() with the reference's codes, as the
 * reference does; ODBA_E_* codes (>= 100) come from the library.
 */
#ifndef ODBA_OPENDBA_HPP
#define ODBA_OPENDBA_HPP

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "odba.h"

#define TEXT_READ_MODE 0
#define BINARY_READ_MODE 1
#define TSV_READ_MODE 2
#define CLUSTER_ONLY ODBA_CLUSTER_ONLY
#define CONSENSUS_ONLY ODBA_CONSENSUS_ONLY
#define CLUSTER_AND_CONSENSUS ODBA_CLUSTER_AND_CONSENSUS
#define NOT_ENOUGH_SEQUENCES 32
#define CANNOT_WRITE_DBA_AVG 23
#define CANNOT_WRITE_MEMBERSHIP 33
#define MEDOID_FINDING_ERROR 36
#define UNKNOWN_ALGO 42
#define CANNOT_READ_SEQUENCE_PREFIX_FILE 24
#define CANNOT_COPY_PREFIX_CHOPPED_SEQ 25
#define CANNOT_WRITE_DTW_PATH_MATRIX 26

namespace odba_host {

template <typename T> struct DType;
template <> struct DType<int> { static const int v = ODBA_I32; };
template <> struct DType<unsigned int> { static const int v = ODBA_U32; };
template <> struct DType<unsigned long long> { static const int v = ODBA_U64; };
template <> struct DType<float> { static const int v = ODBA_F32; };
template <> struct DType<double> { static const int v = ODBA_F64; };

inline odba_ctx* ctx() {
  static odba_ctx* c = nullptr;
  if (!c) {
    const char* dev = std::getenv("ODBA_DEVICE");
    int rc = odba_ctx_create(dev ? std::atoi(dev) : 0, &c);
    if (rc) {
      std::cerr << "Cannot open a HIP device for OpenDBA (code " << rc << ")" << std::endl;
      std::exit(rc);
    }
  }
  return c;
}

inline void check(int rc, const char* what) {
  if (rc) {
    std::cerr << "OpenDBA error in " << what << ": " << odba_last_error(ctx()) << " (code " << rc << ")" << std::endl;
    std::exit(rc);
  }
}

inline char* dup_name(const char* s) {
  char* d = (char*)std::malloc(std::strlen(s) + 1);
  std::strcpy(d, s);
  return d;
}

// Number of '\n' bytes in a file (the reference's value count, cpu_utils.hpp:31-47, 100-108).
inline long count_newlines(const char* path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return -1;
  long n = 0;
  std::vector<char> buf(1 << 20);
  while (f) {
    f.read(buf.data(), buf.size());
    std::streamsize got = f.gcount();
    for (std::streamsize i = 0; i < got; i++) n += buf[i] == '\n';
  }
  return n;
}

}  // namespace odba_host

// ---------------------------------------------------------------------------
// Readers
// ---------------------------------------------------------------------------
template <typename T>
int read_text_data(const char* text_file_name, T** output_vals, size_t* num_output_vals) {
  long n = odba_host::count_newlines(text_file_name);
  if (n < 0) {
    std::cerr << "Error reading in file " << text_file_name << ", skipping" << std::endl;
    return 1;
  }
  *num_output_vals = (size_t)n;
  if (n == 0) {
    std::cerr << "File " << text_file_name << " is empty or not properly formatted. Skipping." << std::endl;
    return 1;
  }
  T* out = (T*)std::calloc((size_t)n, sizeof(T));
  std::ifstream ifs(text_file_name);
  std::string line;
  for (long i = 0; i < n; i++) {
    std::getline(ifs, line);
    std::istringstream in(line);
    in >> out[i];  // first whitespace-separated token
  }
  *output_vals = out;
  return 0;
}

template <typename T>
int read_binary_data(const char* binary_file_name, T** output_vals, size_t* num_output_vals, bool is_short = false) {
  std::ifstream ifs(binary_file_name, std::ios::binary);
  if (!ifs) return 1;
  ifs.seekg(0, ifs.end);
  const std::streamoff bytes = ifs.tellg();
  ifs.seekg(0, ifs.beg);
  const size_t n = (size_t)(bytes / (is_short ? 2 : (std::streamoff)sizeof(T)));
  *num_output_vals = n;
  T* out = (T*)std::calloc(n ? n : 1, sizeof(T));
  if (is_short) {  // int16 expanded to the value type (cpu_utils.hpp:81-85)
    std::vector<short> tmp(n);
    ifs.read((char*)tmp.data(), (std::streamsize)(n * 2));
    for (size_t i = 0; i < n; i++) out[i] = (T)tmp[i];
  } else {
    ifs.read((char*)out, (std::streamsize)(n * sizeof(T)));
  }
  *output_vals = out;
  return 0;
}

template <typename T>
int readSequenceTextFiles(char** filenames, int num_files, T*** sequences, char*** sequence_names,
                          size_t** sequence_lengths) {
  *sequences = (T**)std::calloc(num_files, sizeof(T*));
  *sequence_names = (char**)std::calloc(num_files, sizeof(char*));
  *sequence_lengths = (size_t*)std::calloc(num_files, sizeof(size_t));
  std::cerr << "Step 1 of 3: Loading " << num_files << (num_files == 1 ? " text data file" : " text data files")
            << ", total sequence count " << num_files << std::endl;
  int actual = 0;
  for (int i = 0; i < num_files; ++i) {
    if (read_text_data<T>(filenames[i], (*sequences) + actual, (*sequence_lengths) + actual))
      std::cerr << "Error reading in text file " << filenames[i] << ", skipping" << std::endl;
    else
      actual++;
    (*sequence_names)[i] = odba_host::dup_name(filenames[i]);
  }
  return actual;
}

template <typename T>
int readSequenceBinaryFiles(char** filenames, int num_files, T*** sequences, char*** sequence_names,
                            size_t** sequence_lengths, bool is_short = false) {
  *sequences = (T**)std::calloc(num_files, sizeof(T*));
  *sequence_names = (char**)std::calloc(num_files, sizeof(char*));
  *sequence_lengths = (size_t*)std::calloc(num_files, sizeof(size_t));
  std::cerr << "Step 1 of 3: Loading " << num_files << (num_files == 1 ? " binary data file" : " binary data files")
            << ", total sequence count " << num_files << std::endl;
  int actual = 0;
  for (int i = 0; i < num_files; ++i) {
    if (read_binary_data<T>(filenames[i], (*sequences) + actual, (*sequence_lengths) + actual, is_short))
      std::cerr << "Error reading in binary file " << filenames[i] << ", skipping" << std::endl;
    else
      actual++;
    (*sequence_names)[i] = odba_host::dup_name(filenames[i]);
  }
  return actual;
}

template <typename T>
int readSequenceTSVFiles(char** filenames, int num_files, T*** sequences, char*** sequence_names,
                         size_t** sequence_lengths) {
  size_t total = 0;
  for (int i = 0; i < num_files; ++i) {
    long c = odba_host::count_newlines(filenames[i]);
    if (c > 0) total += (size_t)c;
  }
  *sequences = (T**)std::calloc(total ? total : 1, sizeof(T*));
  *sequence_names = (char**)std::calloc(total ? total : 1, sizeof(char*));
  *sequence_lengths = (size_t*)std::calloc(total ? total : 1, sizeof(size_t));
  int actual = 0;
  for (int i = 0; i < num_files; ++i) {
    std::ifstream ifs(filenames[i]);
    if (!ifs) { std::cerr << "Error reading in TSV file " << filenames[i] << ", skipping" << std::endl; continue; }
    for (std::string line; std::getline(ifs, line);) {
      if ((size_t)actual >= total) break;
      const int ncols = (int)std::count(line.begin(), line.end(), '\t');  // cpu_utils.hpp:392
      T* seq = (T*)std::calloc(ncols ? ncols : 1, sizeof(T));
      std::istringstream iss(line);
      std::string nm;
      iss >> nm;
      int k = 0;
      while (iss.good() && k < ncols) iss >> seq[k++];
      (*sequences)[actual] = seq;
      (*sequence_lengths)[actual] = (size_t)ncols;
      (*sequence_names)[actual] = odba_host::dup_name(nm.c_str());
      actual++;
    }
  }
  return actual;
}

// ---------------------------------------------------------------------------
// Hot-path entry points
// ---------------------------------------------------------------------------
template <typename T>
void normalizeSequences(T** sequences, size_t num_sequences, size_t* sequence_lengths, int refSequenceIndex,
                        double* sequence_means, double* sequence_sigmas) {
  (void)refSequenceIndex;  // the reference's pipeline always z-normalises (-1, gpu_utils.hpp:154-157)
  std::vector<uint64_t> l(sequence_lengths, sequence_lengths + num_sequences);
  odba_host::check(odba_znorm(odba_host::ctx(), odba_host::DType<T>::v, (void* const*)sequences, l.data(),
                              (int)num_sequences, sequence_means, sequence_sigmas),
                   "normalizeSequences");
}

/* chopPrefixFromSequences<T> (dba.hpp:1192-1395): locates the prefix in
 * every sequence (open-end DTW, prefix on Y), writes
 * <output_prefix>.prefix_chop.txt (name, chop position, original length,
 * then the per-prefix-position path histogram, tab-separated;
 * dba.hpp:1349-1358) and replaces each sequence by its part from the chop
 * position on. With norm_sequences the sequences are z-normalised in place
 * first, as the reference does, and stay so. A sequence left empty would be
 * dropped and *num_sequences lowered (dba.hpp:1362-1374); the chop position
 * is at most length - 1, so none is. */
template <typename T>
void chopPrefixFromSequences(T* sequence_prefix, size_t sequence_prefix_length, T** sequences, int* num_sequences,
                             size_t* sequence_lengths, char** sequence_names, char* output_prefix,
                             int norm_sequences, const void* stream = nullptr) {
  (void)stream;
  const int n = *num_sequences;
  std::vector<uint64_t> l(sequence_lengths, sequence_lengths + n);
  std::vector<uint64_t> chop(n);
  std::vector<int32_t> hist((size_t)n * sequence_prefix_length);
  odba_host::check(odba_prefix_chop(odba_host::ctx(), odba_host::DType<T>::v, sequence_prefix, sequence_prefix_length,
                                    (void* const*)sequences, l.data(), n, norm_sequences, chop.data(), hist.data()),
                   "chopPrefixFromSequences");
  std::ofstream out((std::string(output_prefix) + ".prefix_chop.txt").c_str());
  int kept = 0;
  for (int i = 0; i < n; i++) {
    out << sequence_names[i] << "\t" << chop[i] << "\t" << sequence_lengths[i];
    for (size_t j = 0; j < sequence_prefix_length; j++) out << "\t" << hist[(size_t)i * sequence_prefix_length + j];
    out << std::endl;
    const size_t left = sequence_lengths[i] - (size_t)chop[i];
    if (left == 0) {
      std::cerr << "Skipping " << sequence_names[i] << " due to zero-length after prefix chopping" << std::endl;
      std::free(sequences[i]);
      continue;
    }
    T* cut = (T*)std::malloc(sizeof(T) * left);
    if (!cut) {
      std::cerr << "Running memcpy to copy prefix chopped sequence failed";
      std::exit(CANNOT_COPY_PREFIX_CHOPPED_SEQ);
    }
    std::memcpy(cut, sequences[i] + chop[i], sizeof(T) * left);
    std::free(sequences[i]);
    sequences[kept] = cut;
    sequence_lengths[kept] = left;
    sequence_names[kept] = sequence_names[i];
    kept++;
  }
  *num_sequences = kept;
}

namespace odba_host {

inline const char* move_name(int code) {  // io_utils.hpp:303-308
  switch (code) {
    case 1: return "DIAG";
    case 2: return "RIGHT";
    case 3: return "UP";
    case 4: return "OPEN_RIGHT";
    case 0: return "NIL";
    case 5: return "NIL_OPEN_RIGHT";
    default: return "?";
  }
}

/* The per-sequence path files of one DBA round, <prefix>.path<k>.txt
 * (dba.hpp:513-518, writeDTWPath io_utils.hpp:285-323): the sequence name,
 * then one line per traceback cell from the top-right corner,
 * "seq_index<TAB>seq_value<TAB>centroid_index<TAB>centroid_value<TAB>MOVE",
 * then the anchor cell, printed as row, seq[row], column, centroid[column]
 * whatever the orientation (as the reference prints it). With an empty
 * prefix the reference creates the files and writes nothing into them. */
template <typename T>
void write_dba_paths(odba_dba* dba, const std::string& prefix, T* const* seqs, char* const* names, const size_t* lens,
                     size_t n, const T* centroid, size_t lc, int use_open_end) {
  std::vector<uint64_t> st;
  for (size_t k = 0; k < n; k++) {
    const std::string fn = prefix + ".path" + std::to_string(k) + ".txt";
    std::ofstream f(fn.c_str());
    if (!f.is_open()) {
      std::cerr << "Cannot write to " << fn << std::endl;
      std::exit(CANNOT_WRITE_DTW_PATH_MATRIX);
    }
    if (prefix.empty()) continue;
    uint64_t ns = 0;
    check(odba_dba_get_path(dba, (int)k, nullptr, 0, &ns), "DTW path");
    st.resize(ns);
    check(odba_dba_get_path(dba, (int)k, st.data(), ns, &ns), "DTW path");
    const bool flip = use_open_end && lc < lens[k];
    const T* sq = seqs[k];
    f << names[k] << std::endl;
    for (uint64_t s = 0; s < ns; s++) {
      const size_t i = (size_t)(st[s] >> 33), j = (size_t)((st[s] >> 3) & 0x3fffffffu);
      const int code = (int)(st[s] & 7u);
      if (code == 0 || code == 5) {  // anchor (io_utils.hpp:316-318)
        f << i << "\t" << (i < lens[k] ? sq[i] : T(0)) << "\t" << j << "\t" << (j < lc ? centroid[j] : T(0)) << "\t"
          << move_name(code) << std::endl;
      } else if (flip) {
        f << j << "\t" << sq[j] << "\t" << i << "\t" << centroid[i] << "\t" << move_name(code) << std::endl;
      } else {
        f << i << "\t" << sq[i] << "\t" << j << "\t" << centroid[j] << "\t" << move_name(code) << std::endl;
      }
    }
  }
}

}  // namespace odba_host

/* Medoid search on sequences already sorted by length (dba.hpp:36-365):
 * all-vs-all distances, <prefix>.pair_dists.txt, sum of squares in T, one
 * cluster when *cdist == 1. Returns new int[num_clusters]. */
template <typename T>
int* approximateMedoidIndices(T** sequences, size_t num_sequences, size_t* sequence_lengths, char** sequence_names,
                              int use_open_start, int use_open_end, char* output_prefix, double* cdist,
                              int* memberships, const void* stream = nullptr) {
  (void)stream;
  const size_t n = num_sequences;
  std::vector<uint64_t> l(sequence_lengths, sequence_lengths + n);
  std::vector<T> d(n * (n - 1) / 2);
  odba_host::check(odba_dtw_all_pairs(odba_host::ctx(), odba_host::DType<T>::v, (const void* const*)sequences, l.data(),
                                      (int)n, use_open_start, use_open_end, ODBA_REF_W, d.data()),
                   "approximateMedoidIndices");
  std::vector<T> sos(n, T(0));
  std::ofstream mats((std::string(output_prefix) + ".pair_dists.txt").c_str());
  size_t off = 0;
  for (size_t r = 0; r + 1 < n; r++) {  // dba.hpp:188-207
    mats << sequence_names[r];
    for (size_t pad = 0; pad < r; ++pad) mats << "\t";
    mats << "\t0";
    for (size_t c = r + 1; c < n; ++c) {
      T v = d[off + c - r - 1];
      mats << "\t" << v;
      v *= v;
      sos[r] += v;
      sos[c] += v;
    }
    off += n - r - 1;
    mats << std::endl;
  }
  mats << sequence_names[n - 1];
  for (size_t pad = 0; pad < n; ++pad) mats << "\t";
  mats << "0" << std::endl;
  // clusters (cdist > 1: K, == 1: one, [0,1): height cut) and their medoids
  // by the reference's rules (dba.hpp:219-351)
  std::vector<int> med(n > 0 ? n : 1);
  int k = 0;
  odba_host::check(odba_cluster(odba_host::DType<T>::v, d.data(), (int)n, l.data(), *cdist, memberships, med.data(), &k),
                   "approximateMedoidIndices");
  std::cerr << "There are " << k << " clusters" << std::endl;
  int* medoids = new int[k];
  for (int c = 0; c < k; c++) {
    medoids[c] = med[c];
    std::cerr << "medoid is " << med[c] << std::endl;
  }
  return medoids;
}

/* One DBA round (dba.hpp:432-848): the new centroid in updatedMean, the
 * path files <output_prefix>.path<k>.txt; returns the max centroid
 * movement. */
template <typename T>
double DBAUpdate(T* C, size_t centerLength, T** sequences, char** sequence_names, size_t num_sequences,
                 size_t* sequence_lengths, int use_open_start, int use_open_end, T* updatedMean,
                 std::string output_prefix, const void* stream = nullptr) {
  (void)stream;
  std::vector<uint64_t> l(sequence_lengths, sequence_lengths + num_sequences);
  double delta = 0;
  odba_dba* dba = nullptr;
  odba_host::check(odba_dba_create(odba_host::ctx(), odba_host::DType<T>::v, (const void* const*)sequences, l.data(),
                                   (int)num_sequences, use_open_start, use_open_end, ODBA_PATH_FULL, 0, &dba),
                   "DBAUpdate");
  odba_host::check(odba_dba_set_centroid(dba, C, centerLength), "DBAUpdate");
  odba_host::check(odba_dba_record_paths(dba, 1), "DBAUpdate");
  odba_host::check(odba_dba_step(dba, updatedMean, nullptr, &delta), "DBAUpdate");
  odba_host::write_dba_paths<T>(dba, output_prefix, sequences, sequence_names, sequence_lengths, num_sequences, C,
                                centerLength, use_open_end);
  odba_dba_destroy(dba);
  return delta;
}

template <typename T>
void writeCentroidCheckpointToFile(const char* name, const T* c, size_t len) {  // io_utils.hpp:57-76
  std::ofstream f(name);
  if (!f.is_open()) return;
  f << c[0];
  for (size_t i = 1; i < len; i++) f << " " << c[i];
  f << std::endl;
}

template <typename T>
int readCentroidCheckpointFromFile(const char* name, T* c, size_t len) {  // io_utils.hpp:80-115
  std::ifstream f(name);
  if (!f.is_open()) return 0;
  std::string line;
  if (!std::getline(f, line)) return 0;
  std::vector<std::string> vals;
  std::stringstream ss0(line);
  for (std::string tok; std::getline(ss0, tok, ' ');) vals.push_back(tok);
  if (vals.size() != len) return 0;
  std::stringstream ss(line);
  for (size_t i = 0; i < len; i++) ss >> c[i];
  return 1;
}

/* performDBA<T> (dba.hpp:863-1189): medoids and clusters by cdist, then
 * every cluster's consensus (packed rounds on the device), written to
 * <prefix>.avg.txt with the membership file and per-cluster path files.
 * Reorders sequences / names / lengths in place by stable length sort and
 * z-normalises in place when norm_sequences, as the reference does. */
template <typename T>
void performDBA(T** sequences, int num_sequences, size_t* sequence_lengths, char** sequence_names, int use_open_start,
                int use_open_end, char* output_prefix, int norm_sequences, double cdist, char** series_file_names,
                int num_series, int read_mode, bool is_segmented, int algo_mode, const void* stream = nullptr) {
  (void)series_file_names; (void)num_series; (void)read_mode; (void)is_segmented; (void)stream;
  const int n = num_sequences;
  for (int i = 0; i < n; i++)  // dba.hpp:868-872
    if (sequence_lengths[i] && (double)sequences[i][sequence_lengths[i] - 1] >= std::sqrt((double)std::numeric_limits<T>::max()))
      sequence_lengths[i]--;
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return sequence_lengths[a] < sequence_lengths[b]; });
  std::vector<T*> s(n);
  std::vector<char*> nm(n);
  std::vector<size_t> l(n);
  for (int i = 0; i < n; i++) { s[i] = sequences[perm[i]]; nm[i] = sequence_names[perm[i]]; l[i] = sequence_lengths[perm[i]]; }
  for (int i = 0; i < n; i++) { sequences[i] = s[i]; sequence_names[i] = nm[i]; sequence_lengths[i] = l[i]; }
  std::vector<double> mu(n, 0.0), sg(n, 1.0);
  if (norm_sequences) normalizeSequences<T>(sequences, n, sequence_lengths, -1, mu.data(), sg.data());
  std::vector<int> memb(n, 0);
  int* medoids = nullptr;
  if (algo_mode == CLUSTER_AND_CONSENSUS || algo_mode == CLUSTER_ONLY) {
    medoids = approximateMedoidIndices<T>(sequences, n, sequence_lengths, sequence_names, use_open_start, use_open_end,
                                          output_prefix, &cdist, memb.data());
  } else {
    std::cerr << "Call to performDBA included an unsupported algorithm mode " << algo_mode << std::endl;
    std::exit(UNKNOWN_ALGO);
  }
  const std::string prefix(output_prefix);
  int nclusters = 1;
  for (int i = 0; i < n; i++) nclusters = std::max(nclusters, memb[i] + 1);
  if (cdist != 1) {  // dba.hpp:951-963
    std::ofstream mf((prefix + ".cluster_membership.txt").c_str());
    if (!mf.is_open()) {
      std::cerr << "Cannot open sequence cluster membership file " << prefix << ".cluster_membership.txt for writing"
                << std::endl;
      std::exit(CANNOT_WRITE_MEMBERSHIP);
    }
    mf << "## cluster distance threshold was " << cdist << std::endl;
    for (int i = 0; i < n; i++) mf << sequence_names[i] << "\t" << memb[i] << "\t" << sequence_names[medoids[memb[i]]] << std::endl;
    std::cerr << "Found " << nclusters << " clusters using complete linkage and cluster distance cutoff " << cdist << std::endl;
  }
  if (algo_mode == CLUSTER_ONLY) { delete[] medoids; return; }
  // resume after the clusters whose averages are already in avg.txt (dba.hpp:983-988)
  int first = 0;
  {
    std::ifstream prev((prefix + ".avg.txt").c_str());
    for (std::string line; std::getline(prev, line);)
      if (!line.empty()) first++;
    if (first) std::cerr << "Restarting convergence with cluster " << (first + 1) << "/" << nclusters << std::endl;
  }
  std::ofstream avgs((prefix + ".avg.txt").c_str(), std::ios::app);
  if (!avgs.is_open()) {
    std::cerr << "Cannot open sequence averages file " << prefix << ".avg.txt for writing" << std::endl;
    std::exit(CANNOT_WRITE_DBA_AVG);
  }
  for (int cc = first; cc < nclusters; cc++) {
    const int med = medoids[cc];
    const size_t lc = sequence_lengths[med];
    std::vector<T*> cs;
    std::vector<char*> cn;
    std::vector<size_t> cl;
    for (int i = 0; i < n; i++)
      if (memb[i] == cc) { cs.push_back(sequences[i]); cn.push_back(sequence_names[i]); cl.push_back(sequence_lengths[i]); }
    std::vector<T> nb(sequences[med], sequences[med] + lc);
    if (cs.size() > 1) {
      std::vector<T> cen(nb), prev(lc, T(0)), prev2(lc, T(0));
      const std::string ck = prefix + "." + std::to_string(cc) + ".evolving_centroid.txt";
      readCentroidCheckpointFromFile<T>(ck.c_str(), cen.data(), lc);
      std::vector<uint64_t> ul(cl.begin(), cl.end());
      odba_dba* dba = nullptr;
      odba_host::check(odba_dba_create(odba_host::ctx(), odba_host::DType<T>::v, (const void* const*)cs.data(), ul.data(),
                                       (int)cs.size(), use_open_start, use_open_end, ODBA_PATH_FULL, 0, &dba),
                       "DBA session");
      odba_host::check(odba_dba_set_centroid(dba, cen.data(), lc), "DBA centroid");
      odba_host::check(odba_dba_record_paths(dba, 1), "DBA paths");
      std::cerr << "Processing cluster " << (cc + 1) << " of " << nclusters << ", " << cs.size() << " members, medoid "
                << sequence_names[med] << " has length " << lc << std::endl;
      // Every round rewrites <prefix>.<cluster>.path<k>.txt against that
      // round's input centroid; only the last round's files remain, so they
      // are written once, from the last round's recorded tracebacks.
      std::vector<T> round_in(cen);
      for (int i = 0; i < 250; i++) {  // dba.hpp:1078-1110
        double delta = 0;
        if (i > 0) round_in = nb;
        odba_host::check(odba_dba_step(dba, nb.data(), nullptr, &delta), "DBAUpdate");
        std::cerr << "New delta is " << delta << std::endl;
        if (delta == 0) break;
        if (use_open_start || use_open_end) {
          if (i >= 1 && !std::memcmp(nb.data(), prev2.data(), sizeof(T) * lc)) {
            std::cerr << "Detected a flip-flop between two alternative converged centroids, keeping the first one calculated" << std::endl;
            break;
          }
          if (i > 1) prev2 = prev;
          if (i > 0) prev = nb;
        }
        writeCentroidCheckpointToFile<T>(ck.c_str(), nb.data(), lc);
      }
      odba_host::write_dba_paths<T>(dba, prefix + "." + std::to_string(cc), cs.data(), cn.data(), cl.data(), cs.size(),
                                    round_in.data(), lc, use_open_end);
      odba_dba_destroy(dba);
      std::remove(ck.c_str());
    } else {
      std::cerr << "Outputting singleton sequence " << sequence_names[med] << " as-is (a.k.a. cluster " << (cc + 1) << "/"
                << nclusters << ")." << std::endl;
    }
    if (norm_sequences)  // dba.hpp:1017-1023, 1117-1125
      for (size_t i = 0; i < lc; i++) nb[i] = (T)(mu[med] + nb[i] * sg[med]);
    avgs << sequence_names[med];
    for (size_t i = 0; i < lc; ++i) avgs << "\t" << nb[i];
    avgs << std::endl;
    avgs.flush();
  }
  delete[] medoids;
}

/* adaptive_segmentation<T> (segmentation.hpp:339-583); allocates the
 * outputs with malloc. lds_bytes defaults to the opt-in shared memory of
 * the GPUs the reference documents (64 KiB). */
template <typename T>
void adaptive_segmentation(T** sequences, size_t* seq_lengths, int num_seqs, int min_segment_length,
                           T*** segmented_sequences, size_t** segmented_seq_lengths, int prefix_length_to_skip,
                           int lds_bytes = 65536) {
  std::vector<uint64_t> l(seq_lengths, seq_lengths + num_seqs);
  const uint64_t cap = odba_kseg_capacity(l.data(), num_seqs, min_segment_length);
  std::vector<T> out(cap ? cap : 1);
  std::vector<uint64_t> ol(num_seqs), oo(num_seqs);
  odba_host::check(odba_kseg(odba_host::ctx(), odba_host::DType<T>::v, (const void* const*)sequences, l.data(),
                             num_seqs, min_segment_length, prefix_length_to_skip, lds_bytes, out.data(), cap, ol.data(),
                             oo.data()),
                   "adaptive_segmentation");
  *segmented_sequences = (T**)std::calloc(num_seqs, sizeof(T*));
  *segmented_seq_lengths = (size_t*)std::calloc(num_seqs, sizeof(size_t));
  for (int i = 0; i < num_seqs; i++) {
    (*segmented_seq_lengths)[i] = (size_t)ol[i];
    (*segmented_sequences)[i] = (T*)std::malloc((ol[i] ? ol[i] : 1) * sizeof(T));
    std::memcpy((*segmented_sequences)[i], out.data() + oo[i], ol[i] * sizeof(T));
  }
}

/* setupAndRun<T> (openDBA.cuh:13-200). */
template <typename T>
void setupAndRun(char* seqprefix_file_name, char** series_file_names, int num_series, char* output_prefix,
                 int read_mode, int use_open_start, int use_open_end, char* min_segment_length_string,
                 int norm_sequences, double cdist, const int prefix_start = 0, const int prefix_length = 0,
                 bool is_short = false) {
  size_t* lens = nullptr;
  T** seqs = nullptr;
  char** names = nullptr;
  int min2 = -1;
  char* pos = std::strchr(min_segment_length_string, ',');
  if (pos) { min2 = std::atoi(pos + 1); *pos = '\0'; }
  const int minseg = std::atoi(min_segment_length_string);
  int n;
  if (read_mode == BINARY_READ_MODE) n = readSequenceBinaryFiles<T>(series_file_names, num_series, &seqs, &names, &lens, is_short);
  else if (read_mode == TSV_READ_MODE) n = readSequenceTSVFiles<T>(series_file_names, num_series, &seqs, &names, &lens);
  else n = readSequenceTextFiles<T>(series_file_names, num_series, &seqs, &names, &lens);
  if (n < 2) {
    std::cerr << "At least two sequences must be provided to calculate an average, but found " << n << ", aborting" << std::endl;
    std::exit(NOT_ENOUGH_SEQUENCES);
  }
  for (int i = 0; i < n; i++) { char* z = std::strchr(names[i], '.'); if (z) *z = '\0'; }  // openDBA.cuh:59
  if (seqprefix_file_name != nullptr) {  // openDBA.cuh:62-83
    T** pre = nullptr;
    char** pre_name = nullptr;
    size_t* pre_len = nullptr;
    int got;
    if (read_mode == BINARY_READ_MODE) got = readSequenceBinaryFiles<T>(&seqprefix_file_name, 1, &pre, &pre_name, &pre_len);
    else got = readSequenceTextFiles<T>(&seqprefix_file_name, 1, &pre, &pre_name, &pre_len);
    if (got < 1 || pre_len[0] == 0) {
      std::cerr << "Cannot read prefix " << (read_mode == BINARY_READ_MODE ? "binary" : "text") << " data from "
                << seqprefix_file_name << ", aborting" << std::endl;
      std::exit(CANNOT_READ_SEQUENCE_PREFIX_FILE);
    }
    chopPrefixFromSequences<T>(pre[0], pre_len[0], seqs, &n, lens, names, output_prefix, norm_sequences);
    std::free(pre[0]);
    std::free(pre);
    std::free(pre_name[0]);
    std::free(pre_name);
    std::free(pre_len);
  }
  if (minseg > 0) {  // openDBA.cuh:85-177
    T** seg = nullptr;
    size_t* seglens = nullptr;
    adaptive_segmentation<T>(seqs, lens, n, minseg, &seg, &seglens, prefix_start);
    int kept = 0;
    for (int i = 0; i < n; i++) {
      std::free(seqs[i]);
      if (seglens[i] < 2 || (prefix_length > 0 && seglens[i] < (size_t)prefix_length)) { std::free(seg[i]); continue; }
      seg[kept] = seg[i]; seglens[kept] = seglens[i]; names[kept] = names[i]; kept++;
    }
    if (kept < n) {
      std::cerr << "Removing " << (n - kept) << " segmented sequences that are too short" << std::endl;
      n = kept;
      if (n < 2) { std::cerr << "At least two sequences must survive segmentation filters" << std::endl; std::exit(NOT_ENOUGH_SEQUENCES); }
    }
    if (prefix_length > 0) for (int i = 0; i < n; i++) seglens[i] = (size_t)prefix_length;
    std::ofstream out((std::string(output_prefix) + ".segmented_seqs.txt").c_str());  // io_utils.hpp:242-261
    for (int i = 0; i < n; i++) {
      out << names[i] << "\t" << seg[i][0];
      for (size_t j = 1; j < seglens[i]; j++) out << "\t" << seg[i][j];
      out << std::endl;
    }
    std::free(seqs);
    std::free(lens);
    seqs = seg;
    lens = seglens;
    (void)min2;
  }
  performDBA<T>(seqs, n, lens, names, use_open_start, use_open_end, output_prefix, norm_sequences, cdist,
                series_file_names, num_series, read_mode, minseg > 1, CLUSTER_AND_CONSENSUS);
  for (int i = 0; i < n; i++) { std::free(names[i]); std::free(seqs[i]); }
  std::free(names);
  std::free(seqs);
  std::free(lens);
}

#endif /* ODBA_OPENDBA_HPP */
