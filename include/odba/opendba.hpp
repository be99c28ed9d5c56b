/*
 * Drop-in C++ host layer over libodba_hip: the reference's templated entry
 * points (same names, argument meaning, output files and exit codes) with
 * the DTW / DBA work done by the HIP library through the C ABI (odba.h).
 *
 *   readSequenceTextFiles / readSequenceBinaryFiles / readSequenceTSVFiles
 *                                     cpu_utils.hpp:60-134, 381-456, 553-624
 *   readSequenceSLOW5Files / writeSlow5Output (odba_slow5.hpp)
 *                                     cpu_utils.hpp:458-500, io_utils.hpp:363-452
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
#include "odba_slow5.hpp"

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
#define CANNOT_READ_MEMBERSHIP 43
#define MEMBERSHIP_FILE_FORMAT_VIOLATION 44

/* HIP's stream handle (hip_runtime_api.h declares the same typedef). The
 * reference's entry points take a CUDA stream as their last argument; here
 * that slot takes a hipStream_t or 0. The library orders its work on its own
 * context's stream, so the value is accepted and not used. */
typedef struct ihipStream_t* hipStream_t;

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

/* The free-memory figure of the stripe-mode rule (dba.hpp:523-548): 0 lets
 * the library read the device's free memory; ODBA_REF_FREE_MEM (bytes)
 * states one instead (tests force stripe mode with it). */
inline uint64_t ref_free_mem() {
  const char* v = std::getenv("ODBA_REF_FREE_MEM");
  return (v && *v) ? (uint64_t)std::strtoull(v, nullptr, 10) : 0;
}

/* Shared-memory budget the k-segmentation tables are sized with
 * (segmentation.hpp:458-462: the CUDA device's opt-in maximum per block):
 * ODBA_KSEG_LDS_BYTES, else 64 KiB. */
inline int kseg_lds_bytes() {
  const char* v = std::getenv("ODBA_KSEG_LDS_BYTES");
  return (v && *v) ? std::atoi(v) : 65536;
}

inline bool file_exists(const std::string& name) {  // io_utils.hpp:34-37
  std::ifstream f(name.c_str());
  return f.good();
}

inline std::vector<std::string> split(const std::string& line, char delim) {  // io_utils.hpp:40-47
  std::vector<std::string> out;
  std::stringstream ss(line);
  for (std::string tok; std::getline(ss, tok, delim);) out.push_back(tok);
  return out;
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
/* normalizeSequences<T> (gpu_utils.hpp:131-175): z-normalises in place;
 * sequence_means / sequence_sigmas (optional) receive mu and sigma. */
template <typename T>
void normalizeSequences(T** sequences, size_t num_sequences, size_t* sequence_lengths, int refSequenceIndex,
                        double* sequence_means, double* sequence_sigmas, hipStream_t stream = 0) {
  (void)refSequenceIndex;  // the reference's pipeline always z-normalises (-1, gpu_utils.hpp:154-157)
  (void)stream;
  std::vector<uint64_t> l(sequence_lengths, sequence_lengths + num_sequences);
  odba_host::check(odba_znorm(odba_host::ctx(), odba_host::DType<T>::v, (void* const*)sequences, l.data(),
                              (int)num_sequences, sequence_means, sequence_sigmas),
                   "normalizeSequences");
}

/* The overload without mu / sigma (gpu_utils.hpp:173-175). */
template <typename T>
void normalizeSequences(T** sequences, size_t num_sequences, size_t* sequence_lengths, int refSequenceIndex,
                        hipStream_t stream = 0) {
  normalizeSequences<T>(sequences, num_sequences, sequence_lengths, refSequenceIndex, (double*)nullptr,
                        (double*)nullptr, stream);
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
                             int norm_sequences, hipStream_t stream = 0) {
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

/* The medoid search of approximateMedoidIndices (dba.hpp:36-365) on
 * sequences already sorted by length: all-vs-all distances,
 * <prefix>.pair_dists.txt, sum of squares in T, clusters by cdist (> 1: K,
 * == 1: one, [0, 1): height cut) and their medoids by the reference's rules.
 * Returns new int[num_clusters]. */
template <typename T>
int* medoids_of(T* const* sequences, size_t num_sequences, const size_t* sequence_lengths, char* const* sequence_names,
                int use_open_start, int use_open_end, const char* output_prefix, double* cdist, int* memberships) {
  const size_t n = num_sequences;
  std::vector<uint64_t> l(sequence_lengths, sequence_lengths + n);
  std::vector<T> d(n * (n - 1) / 2);
  check(odba_dtw_all_pairs(ctx(), DType<T>::v, (const void* const*)sequences, l.data(), (int)n, use_open_start,
                           use_open_end, ODBA_REF_W, d.data()),
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
  std::vector<int> med(n > 0 ? n : 1);
  int k = 0;
  check(odba_cluster(DType<T>::v, d.data(), (int)n, l.data(), *cdist, memberships, med.data(), &k),
        "approximateMedoidIndices");
  std::cerr << "There are " << k << " clusters" << std::endl;
  int* medoids = new int[k];
  for (int c = 0; c < k; c++) {
    medoids[c] = med[c];
    std::cerr << "medoid is " << med[c] << std::endl;
  }
  return medoids;
}

}  // namespace odba_host

/* approximateMedoidIndices<T> (dba.hpp:36-365): gpu_sequences holds the
 * num_sequences sequences (already sorted by length) evenly spaced,
 * sequence i at gpu_sequences + i * maxSeqLength, as performDBA lays them
 * out (dba.hpp:919-923). Writes <output_prefix>.pair_dists.txt, fills
 * memberships[num_sequences] and returns new int[num_clusters] medoids. */
template <typename T>
int* approximateMedoidIndices(T* gpu_sequences, size_t maxSeqLength, size_t num_sequences, size_t* sequence_lengths,
                              char** sequence_names, int use_open_start, int use_open_end, char* output_prefix,
                              double* cdist, int* memberships, hipStream_t stream = 0) {
  (void)stream;
  std::vector<T*> p(num_sequences);
  for (size_t i = 0; i < num_sequences; i++) p[i] = gpu_sequences + i * maxSeqLength;
  return odba_host::medoids_of<T>(p.data(), num_sequences, sequence_lengths, sequence_names, use_open_start,
                                  use_open_end, output_prefix, cdist, memberships);
}

/* One DBA round (dba.hpp:432-848): the new centroid in updatedMean, the
 * path files <output_prefix>.path<k>.txt; returns the max centroid
 * movement. Each sequence takes the full path matrix or stripe mode by the
 * reference's free-memory rule (dba.hpp:523-548). */
template <typename T>
double DBAUpdate(T* C, size_t centerLength, T** sequences, char** sequence_names, size_t num_sequences,
                 size_t* sequence_lengths, int use_open_start, int use_open_end, T* updatedMean,
                 std::string output_prefix, hipStream_t stream = 0) {
  (void)stream;
  std::vector<uint64_t> l(sequence_lengths, sequence_lengths + num_sequences);
  double delta = 0;
  odba_dba* dba = nullptr;
  odba_host::check(odba_dba_create(odba_host::ctx(), odba_host::DType<T>::v, (const void* const*)sequences, l.data(),
                                   (int)num_sequences, use_open_start, use_open_end, ODBA_PATH_AUTO,
                                   odba_host::ref_free_mem(), &dba),
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
  if (!odba_host::file_exists(name)) return 0;
  std::ifstream f(name);
  std::string line;
  if (!std::getline(f, line)) return 0;
  if (odba_host::split(line, ' ').size() != len) return 0;
  std::cerr << "Resuming convergence from partially converged centroid in file " << name << std::endl;
  std::stringstream ss(line);
  for (size_t i = 0; i < len; i++) ss >> c[i];
  return 1;
}

/* readMedoidIndices (io_utils.hpp:160-240): memberships by sequence name
 * from a membership file (a '#' comment line, then name, cluster, medoid
 * name per line); returns new int[] of each cluster's medoid index (the
 * sequence whose name is its own medoid's). Exits with the reference's
 * codes on a missing or malformed file. Sequences the file does not name
 * keep membership 0; a cluster whose medoid is not among sequence_names
 * exits with MEDOID_FINDING_ERROR (the reference reads an unset entry). */
inline int* readMedoidIndices(const char* membership_filename, int num_sequences, char** sequence_names,
                              int* sequences_membership) {
  std::ifstream f(membership_filename);
  if (!f.is_open()) {
    std::cerr << "Cannot open sequence cluster membership file " << membership_filename << " for reading" << std::endl;
    std::exit(CANNOT_READ_MEMBERSHIP);
  }
  std::string line;
  std::getline(f, line);
  if (line.compare(0, 1, "#")) {
    std::cerr << "The existing sequence cluster membership file " << membership_filename << " has a first line "
              << "without the expected '#' comment start" << std::endl;
    std::exit(MEMBERSHIP_FILE_FORMAT_VIOLATION);
  }
  auto index_of = [&](const std::string& nm) {
    for (int i = 0; i < num_sequences; i++)
      if (nm == sequence_names[i]) return i;
    return -1;
  };
  for (int i = 0; i < num_sequences; i++) sequences_membership[i] = 0;
  int max_cluster = 0, line_number = 1;
  std::vector<std::pair<int, std::string>> medoid_rows;
  while (std::getline(f, line)) {
    line_number++;
    const std::vector<std::string> v = odba_host::split(line, '\t');
    if (v.size() != 3) {
      std::cerr << "The existing sequence cluster membership file " << membership_filename << " has a line (#"
                << line_number << ") without the expected three columns (found " << v.size() << ")" << std::endl;
      std::exit(MEMBERSHIP_FILE_FORMAT_VIOLATION);
    }
    int cluster = 0;
    try {
      cluster = std::stoi(v[1]);
    } catch (const std::exception& ex) {
      std::cerr << "The existing sequence cluster membership file " << membership_filename << " has a line (#"
                << line_number << ") where the second tab-delimited column value (" << v[1]
                << ") is not an integer as expected: " << ex.what() << std::endl;
      std::exit(MEMBERSHIP_FILE_FORMAT_VIOLATION);
    }
    const int idx = index_of(v[0]);
    if (idx == -1) {
      std::cerr << "The existing sequence cluster membership file " << membership_filename << " has a line (#"
                << line_number << ") with a sequence name not found in the input (" << v[0]
                << " not in existing list of " << num_sequences << " names)" << std::endl;
      std::exit(MEMBERSHIP_FILE_FORMAT_VIOLATION);
    }
    max_cluster = std::max(max_cluster, cluster);
    sequences_membership[idx] = cluster;
    if (v[0] == v[2]) medoid_rows.push_back(std::make_pair(cluster, v[0]));
  }
  int* medoids = new int[max_cluster + 1];
  for (int c = 0; c <= max_cluster; c++) medoids[c] = -1;
  for (const auto& mr : medoid_rows) medoids[mr.first] = index_of(mr.second);
  for (int c = 0; c <= max_cluster; c++)
    if (medoids[c] < 0) {
      std::cerr << "The existing sequence cluster membership file " << membership_filename << " names no medoid for "
                << "cluster " << c << std::endl;
      std::exit(MEDOID_FINDING_ERROR);
    }
  return medoids;
}

/* writeSequences (io_utils.hpp:242-261): one line per sequence, its name
 * then its values, tab-separated. */
template <typename T>
int writeSequences(T** cpu_sequences, size_t* seq_lengths, char** seq_names, int num_seqs, const char* filename) {
  std::ofstream out(filename);
  if (!out.is_open()) {
    std::cerr << "Cannot write to " << filename << std::endl;
    return CANNOT_WRITE_DTW_PATH_MATRIX;
  }
  for (int i = 0; i < num_seqs; i++) {
    out << seq_names[i] << "\t" << cpu_sequences[i][0];
    for (size_t j = 1; j < seq_lengths[i]; j++) out << "\t" << cpu_sequences[i][j];
    out << std::endl;
  }
  return 0;
}

/* performDBA<T> (dba.hpp:863-1189): stable length sort of sequences /
 * names / lengths in place, z-norm in place when norm_sequences, then by
 * algo_mode:
 *   CLUSTER_AND_CONSENSUS, CLUSTER_ONLY: medoids and clusters by cdist
 *     (<prefix>.pair_dists.txt; <prefix>.cluster_membership.txt when
 *     cdist != 1); CLUSTER_ONLY stops there;
 *   CONSENSUS_ONLY: the memberships and medoids of a previous call, read
 *     by sequence name from <prefix>.cluster_membership.txt;
 * then every cluster's consensus (at most 250 rounds, delta == 0 or the
 * open-mode flip-flop stop), rescaled by its medoid's mu / sigma, appended
 * to <prefix>.avg.txt, with per-cluster path files and the evolving-centroid
 * checkpoint <prefix>.<cluster>.evolving_centroid.txt read back at the
 * cluster's start. Resume rule of the reference's default build
 * (dba.hpp:981-984, io_utils.hpp:121-123): when <prefix>.avg.txt already
 * exists, the run starts at cluster index 1. */
template <typename T>
void performDBA(T** sequences, int num_sequences, size_t* sequence_lengths, char** sequence_names, int use_open_start,
                int use_open_end, char* output_prefix, int norm_sequences, double cdist, char** series_file_names,
                int num_series, int read_mode, bool is_segmented, int algo_mode, hipStream_t stream = 0) {
  (void)stream;
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
  const std::string prefix(output_prefix);
  if (algo_mode == CLUSTER_AND_CONSENSUS || algo_mode == CLUSTER_ONLY) {
    medoids = odba_host::medoids_of<T>(sequences, n, sequence_lengths, sequence_names, use_open_start, use_open_end,
                                       output_prefix, &cdist, memb.data());
  } else if (algo_mode == CONSENSUS_ONLY) {
    std::cerr << "Reading previous clustering data" << std::endl;
    medoids = readMedoidIndices((prefix + ".cluster_membership.txt").c_str(), n, sequence_names, memb.data());
  } else {
    std::cerr << "Call to performDBA included an unrecognized algorithm mode " << algo_mode << std::endl;
    std::exit(UNKNOWN_ALGO);
  }
  int nclusters = 1;
  for (int i = 0; i < n; i++) nclusters = std::max(nclusters, memb[i] + 1);
  if (cdist != 1 && algo_mode != CONSENSUS_ONLY) {  // dba.hpp:951-963
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
  int first = 0;
  if (odba_host::file_exists(prefix + ".avg.txt")) {  // dba.hpp:981-985
    first = 1;
    std::cerr << "Restarting convergence with cluster " << (first + 1) << "/" << nclusters << " based on checkpoint in "
              << prefix << ".avg.txt" << std::endl;
  }
  std::ofstream avgs((prefix + ".avg.txt").c_str(), std::ios::app);
  if (!avgs.is_open()) {
    std::cerr << "Cannot open sequence averages file " << prefix << ".avg.txt for writing" << std::endl;
    std::exit(CANNOT_WRITE_DBA_AVG);
  }
  // A single signal file's consensus sequences are also written back into a
  // copy of its records, as <prefix>.avg.blow5 (dba.hpp:1170-1185).
  const bool want_blow5 = !is_segmented && read_mode == SLOW5_READ_MODE && num_series == 1;
  std::vector<std::vector<short>> avg_short(want_blow5 ? nclusters : 0);
  std::vector<char*> avg_names(want_blow5 ? nclusters : 0, nullptr);
  std::vector<size_t> avg_lens(want_blow5 ? nclusters : 0, 0);
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
                                       (int)cs.size(), use_open_start, use_open_end, ODBA_PATH_AUTO,
                                       odba_host::ref_free_mem(), &dba),
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
    if (want_blow5) {  // dba.hpp:1008-1037, 1133-1139
      avg_short[cc].resize(lc);
      avg_names[cc] = sequence_names[med];
      avg_lens[cc] = lc;
      if (cs.size() == 1 && norm_sequences)  // a singleton's short comes from the double rescale itself
        for (size_t i = 0; i < lc; i++) avg_short[cc][i] = (short)(mu[med] + nb[i] * sg[med]);
    }
    if (norm_sequences)  // dba.hpp:1017-1023, 1117-1125
      for (size_t i = 0; i < lc; i++) nb[i] = (T)(mu[med] + nb[i] * sg[med]);
    if (want_blow5 && !(cs.size() == 1 && norm_sequences))  // templateToShort of the (rescaled) T values
      for (size_t i = 0; i < lc; i++) avg_short[cc][i] = (short)nb[i];
    avgs << sequence_names[med];
    for (size_t i = 0; i < lc; ++i) avgs << "\t" << nb[i];
    avgs << std::endl;
    avgs.flush();
  }
  if (want_blow5 && first > 0) {
    // The reference's resume leaves the skipped cluster's buffers unset;
    // this layer writes no partial file.
    std::cerr << "Not writing " << prefix << ".avg.blow5 after a resumed run (cluster 1 came from the checkpoint)"
              << std::endl;
  } else if (want_blow5) {
    std::cerr << "Writing medoids to new slow5 file..." << std::endl;
    std::vector<short*> sp(nclusters);
    for (int cc = 0; cc < nclusters; cc++) sp[cc] = avg_short[cc].data();
    if (writeSlow5Output(series_file_names[0], (prefix + ".avg.blow5").c_str(), avg_names.data(), sp.data(),
                         avg_lens.data(), nclusters) == 1) {
      std::cerr << "Cannot write updated sequences to new Fast5 file " << prefix << ".avg.blow5, aborting." << std::endl;
      std::exit(CANNOT_WRITE_UPDATED_SLOW5);
    }
  }
  delete[] medoids;
}

/* adaptive_segmentation<T> (segmentation.hpp:339-583); allocates the
 * outputs with malloc. The shared-memory budget the reference sizes its
 * tables with comes from ODBA_KSEG_LDS_BYTES (64 KiB by default). */
template <typename T>
void adaptive_segmentation(T** sequences, size_t* seq_lengths, int num_seqs, int min_segment_length,
                           T*** segmented_sequences, size_t** segmented_seq_lengths, int prefix_length_to_skip,
                           hipStream_t stream = 0) {
  (void)stream;
  std::vector<uint64_t> l(seq_lengths, seq_lengths + num_seqs);
  const uint64_t cap = odba_kseg_capacity(l.data(), num_seqs, min_segment_length);
  std::vector<T> out(cap ? cap : 1);
  std::vector<uint64_t> ol(num_seqs), oo(num_seqs);
  odba_host::check(odba_kseg(odba_host::ctx(), odba_host::DType<T>::v, (const void* const*)sequences, l.data(),
                             num_seqs, min_segment_length, prefix_length_to_skip, odba_host::kseg_lds_bytes(),
                             out.data(), cap, ol.data(), oo.data()),
                   "adaptive_segmentation");
  *segmented_sequences = (T**)std::calloc(num_seqs, sizeof(T*));
  *segmented_seq_lengths = (size_t*)std::calloc(num_seqs, sizeof(size_t));
  for (int i = 0; i < num_seqs; i++) {
    (*segmented_seq_lengths)[i] = (size_t)ol[i];
    (*segmented_sequences)[i] = (T*)std::malloc((ol[i] ? ol[i] : 1) * sizeof(T));
    std::memcpy((*segmented_sequences)[i], out.data() + oo[i], ol[i] * sizeof(T));
  }
}

/* setupAndRun<T> (openDBA.cuh:13-200): read, optional prefix chop, optional
 * k-segmentation, then performDBA. min_segment_length_string is "m" or
 * "m,m2": with m2, the segmented sequences are clustered (CLUSTER_ONLY) and
 * the consensus is built on the unsegmented ones from the membership file
 * (CONSENSUS_ONLY), as the reference does. The reference keeps one names
 * array for both sets: the clustering call's length sort reorders it and
 * the segmentation filter shifts it, while the unsegmented sequences stay
 * in input order, so the consensus call pairs sequence i with the name at i
 * after those changes; this layer does the same. */
template <typename T>
void setupAndRun(char* seqprefix_file_name, char** series_file_names, int num_series, char* output_prefix,
                 int read_mode, int use_open_start, int use_open_end, char* min_segment_length_string,
                 int norm_sequences, double cdist, const int prefix_start = 0, const int prefix_length = 0,
                 bool is_short = false) {
  size_t* lens = nullptr;
  T** seqs = nullptr;
  char** names = nullptr;
  int min2 = -1;  // -1: not given (openDBA.cuh:26-33)
  char* pos = std::strchr(min_segment_length_string, ',');
  if (pos) { min2 = std::atoi(pos + 1); *pos = '\0'; }
  const int minseg = std::atoi(min_segment_length_string);
  int n;
  if (read_mode == BINARY_READ_MODE) n = readSequenceBinaryFiles<T>(series_file_names, num_series, &seqs, &names, &lens, is_short);
  else if (read_mode == TSV_READ_MODE) n = readSequenceTSVFiles<T>(series_file_names, num_series, &seqs, &names, &lens);
  else if (read_mode == SLOW5_READ_MODE) n = readSequenceSLOW5Files<T>(series_file_names, num_series, &seqs, &names, &lens);
  else n = readSequenceTextFiles<T>(series_file_names, num_series, &seqs, &names, &lens);
  if (n < 2) {
    std::cerr << "At least two sequences must be provided to calculate an average, but found " << n << ", aborting" << std::endl;
    std::exit(NOT_ENOUGH_SEQUENCES);
  }
  const int n_read = n;
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
  const int n_raw = n;  // unsegmented sequences held (the two-stage flow keeps them)
  int reloaded = -1;    // sequences read again for the consensus (signal files)
  if (minseg > 0) {  // openDBA.cuh:85-177
    T** seg = nullptr;
    size_t* seglens = nullptr;
    adaptive_segmentation<T>(seqs, lens, n, minseg, &seg, &seglens, prefix_start);
    int removed = 0;
    for (int i = 0; i < n; i++) {
      if (min2 == -1) std::free(seqs[i]);
      const int k = i - removed;
      if (seglens[k] < 2 || (prefix_length > 0 && seglens[k] < (size_t)prefix_length)) {
        std::free(seg[k]);
        for (int j = k + 1; j < n; j++) { seg[j - 1] = seg[j]; seglens[j - 1] = seglens[j]; names[j - 1] = names[j]; }
        removed++;
      }
    }
    if (removed) {
      std::cerr << "Removing " << removed << " segmented sequences that are too short, as they may unduly skew the "
                << "convergence process. To retain more sequences, consider setting a smaller minimum segment size "
                << "(currently " << minseg << ")" << std::endl;
      n -= removed;
      if (n < 2) {
        std::cerr << "At least two sequences must survive segmentation filters to calculate an average, but found " << n
                  << ", aborting" << std::endl;
        std::exit(NOT_ENOUGH_SEQUENCES);
      }
    }
    if (prefix_length > 0) for (int i = 0; i < n; i++) seglens[i] = (size_t)prefix_length;
    writeSequences<T>(seg, seglens, names, n, (std::string(output_prefix) + ".segmented_seqs.txt").c_str());
    if (min2 != -1) {
      std::cerr << "Performing cluster generation with segment size of " << minseg << std::endl;
      performDBA<T>(seg, n, seglens, names, use_open_start, use_open_end, output_prefix, norm_sequences, cdist,
                    series_file_names, num_series, read_mode, minseg > 1, CLUSTER_ONLY);
      for (int i = 0; i < n; i++) std::free(seg[i]);
      std::free(seg);
      std::free(seglens);
      // A prefix-chopped signal file is read again for a consensus without
      // segmentation, so the consensus keeps the leader as a raw signal
      // would (openDBA.cuh:135-167).
      if (seqprefix_file_name != nullptr && min2 == 0 && read_mode == SLOW5_READ_MODE) {
        std::cerr << "Restoring raw signals (no prefix chop) before FAST5/SLOW5 consensus generation without segmentation"
                  << std::endl;
        for (int i = 0; i < n_raw; i++) std::free(seqs[i]);
        for (int i = 0; i < n; i++) std::free(names[i]);
        std::free(seqs);
        std::free(names);
        std::free(lens);
        n = readSequenceSLOW5Files<T>(series_file_names, num_series, &seqs, &names, &lens);
        reloaded = n;
      }
    } else {
      std::free(seqs);
      std::free(lens);
      seqs = seg;
      lens = seglens;
    }
  }
  if (min2 != -1) {
    std::cerr << "Performing consensus generation with segment size of " << min2 << std::endl;
    performDBA<T>(seqs, n, lens, names, use_open_start, use_open_end, output_prefix, norm_sequences, cdist,
                  series_file_names, num_series, read_mode, minseg > 1, CONSENSUS_ONLY);
  } else {
    std::cerr << "Performing both clustering and consensus generation with segment size of " << minseg << std::endl;
    performDBA<T>(seqs, n, lens, names, use_open_start, use_open_end, output_prefix, norm_sequences, cdist,
                  series_file_names, num_series, read_mode, minseg > 1, CLUSTER_AND_CONSENSUS);
  }
  const int held = reloaded >= 0 ? reloaded : ((min2 != -1 || minseg <= 0) ? n_raw : n);
  for (int i = 0; i < held; i++) std::free(seqs[i]);
  for (int i = 0; i < n; i++) std::free(names[i]);
  (void)n_read;
  std::free(names);
  std::free(seqs);
  std::free(lens);
}

#endif /* ODBA_OPENDBA_HPP */
