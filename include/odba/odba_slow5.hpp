/*
 * S/BLOW5 nanopore signal files for the drop-in layer: the reference reads
 * them through slow5lib (cpu_utils.hpp:136-165, 235-278, 458-500) and writes
 * the cluster consensus sequences back as <prefix>.avg.blow5
 * (io_utils.hpp:363-452, dba.hpp:1170-1185). slow5lib is a submodule the
 * reference does not vendor, so this reader and writer follow the file
 * format itself:
 *
 *   SLOW5 (text): '@' header lines, a '#' line of column types, a
 *     '#read_id...' line of column names, then one tab-separated record per
 *     read: read_id, read_group, digitisation, offset, range, sampling_rate,
 *     len_raw_signal, raw_signal (comma-separated int16), auxiliary columns.
 *   BLOW5 (binary): "BLOW5\1", version (3 bytes), record compression (0
 *     none, 1 zlib; 2 zstd is refused), read-group count (uint32), signal
 *     compression (0 none, 1 svb-zd; version 0.2.0 and later), zero padding
 *     to 64 bytes, the text header's byte count (uint32) and the text header;
 *     then records, each its (compressed) byte count (uint64) and bytes; then
 *     "5WOLB". A record: read_id (uint16 count + chars), read_group
 *     (uint32), digitisation, offset, range, sampling_rate (double), then
 *     the signal - without signal compression len_raw_signal (uint64) and
 *     that many int16; with svb-zd the encoding's byte count (uint64) and
 *     the encoding: the sample count (uint32), one control byte per four
 *     samples (two bits each: 1-4 bytes), the little-endian bytes of the
 *     zigzag-coded differences of consecutive samples (the first against
 *     0) - then the auxiliary fields, kept as bytes.
 *
 * Reading needs only the read ids and signals; rewriting a file with new
 * signals of the same lengths keeps every other byte of each record and the
 * header, and re-encodes the signal and the record with the input file's
 * own compression. Exit and return codes follow the reference.
 */
#ifndef ODBA_SLOW5_HPP
#define ODBA_SLOW5_HPP

#include <zlib.h>

#include <cstdint>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#define FAST5_READ_MODE 3  /* read_mode_codes.h:8 */
#define SLOW5_READ_MODE 4  /* read_mode_codes.h:11 */
#define CANNOT_WRITE_UPDATED_SLOW5 40  /* exit_codes.hpp:26 */
#define SLOW5_FILE_UNREADABLE 41       /* exit_codes.hpp:27 */

namespace odba_slow5 {

struct Record {
  std::string id;
  std::vector<int16_t> sig;
  std::vector<uint8_t> head;  // binary: read_id .. sampling_rate, as stored
  std::vector<uint8_t> aux;   // binary: the auxiliary fields, as stored
};

struct File {
  bool binary = false;
  uint8_t rec_press = 0, sig_press = 0;
  std::vector<uint8_t> header;  // binary: everything before the first record
  std::vector<Record> recs;
  std::string err;
};

inline bool inflate_all(const uint8_t* in, size_t n, std::vector<uint8_t>& out) {
  z_stream zs;
  std::memset(&zs, 0, sizeof(zs));
  if (inflateInit(&zs) != Z_OK) return false;
  out.resize(n * 4 + 1024);
  zs.next_in = const_cast<Bytef*>(in);
  zs.avail_in = (uInt)n;
  int rc = Z_OK;
  while (rc != Z_STREAM_END) {
    if (zs.total_out >= out.size()) out.resize(out.size() * 2);
    zs.next_out = out.data() + zs.total_out;
    zs.avail_out = (uInt)(out.size() - zs.total_out);
    rc = inflate(&zs, Z_NO_FLUSH);
    if (rc != Z_OK && rc != Z_STREAM_END) {
      inflateEnd(&zs);
      return false;
    }
    if (rc == Z_OK && zs.avail_in == 0 && zs.avail_out != 0) {  // truncated stream
      inflateEnd(&zs);
      return false;
    }
  }
  out.resize(zs.total_out);
  inflateEnd(&zs);
  return true;
}

inline bool deflate_all(const std::vector<uint8_t>& in, std::vector<uint8_t>& out) {
  uLongf cap = compressBound((uLong)in.size());
  out.resize(cap);
  if (compress2(out.data(), &cap, in.data(), (uLong)in.size(), Z_DEFAULT_COMPRESSION) != Z_OK) return false;
  out.resize(cap);
  return true;
}

template <typename V> inline bool take(const std::vector<uint8_t>& b, size_t& p, V& v) {
  if (p + sizeof(V) > b.size()) return false;
  std::memcpy(&v, b.data() + p, sizeof(V));
  p += sizeof(V);
  return true;
}
template <typename V> inline void put(std::vector<uint8_t>& b, V v) {
  const uint8_t* s = (const uint8_t*)&v;
  b.insert(b.end(), s, s + sizeof(V));
}

// svb-zd: sample count, control bytes, then zigzag-coded differences.
inline bool svb_zd_decode(const uint8_t* p, size_t n, std::vector<int16_t>& sig) {
  if (n < 4) return false;
  uint32_t cnt;
  std::memcpy(&cnt, p, 4);
  const size_t ncb = ((size_t)cnt + 3) / 4;
  if (4 + ncb > n) return false;
  const uint8_t* ctrl = p + 4;
  const uint8_t* d = ctrl + ncb;
  const uint8_t* end = p + n;
  sig.resize(cnt);
  int32_t prev = 0;
  for (uint32_t i = 0; i < cnt; i++) {
    const int nb = ((ctrl[i >> 2] >> ((i & 3) * 2)) & 3) + 1;
    if (d + nb > end) return false;
    uint32_t u = 0;
    for (int k = 0; k < nb; k++) u |= (uint32_t)d[k] << (8 * k);
    d += nb;
    const int32_t delta = (int32_t)(u >> 1) ^ -(int32_t)(u & 1);
    prev += delta;
    sig[i] = (int16_t)prev;
  }
  return d == end;
}

inline void svb_zd_encode(const std::vector<int16_t>& sig, std::vector<uint8_t>& out) {
  const uint32_t cnt = (uint32_t)sig.size();
  const size_t ncb = ((size_t)cnt + 3) / 4;
  out.assign(4 + ncb, 0);
  std::memcpy(out.data(), &cnt, 4);
  int32_t prev = 0;
  for (uint32_t i = 0; i < cnt; i++) {
    const int32_t delta = (int32_t)sig[i] - prev;
    prev = sig[i];
    const uint32_t u = ((uint32_t)delta << 1) ^ (uint32_t)(delta >> 31);
    const int nb = u < (1u << 8) ? 1 : u < (1u << 16) ? 2 : u < (1u << 24) ? 3 : 4;
    out[4 + (i >> 2)] |= (uint8_t)((nb - 1) << ((i & 3) * 2));
    for (int k = 0; k < nb; k++) out.push_back((uint8_t)(u >> (8 * k)));
  }
}

inline bool parse_record(const std::vector<uint8_t>& r, uint8_t sig_press, Record& rec) {
  size_t p = 0;
  uint16_t idlen;
  if (!take(r, p, idlen) || p + idlen > r.size()) return false;
  rec.id.assign((const char*)r.data() + p, idlen);
  p += idlen;
  p += 4 + 4 * 8;  // read_group, digitisation, offset, range, sampling_rate
  if (p > r.size()) return false;
  rec.head.assign(r.begin(), r.begin() + p);
  uint64_t n;
  if (!take(r, p, n)) return false;
  if (sig_press == 0) {
    if (p + n * 2 > r.size()) return false;
    rec.sig.resize(n);
    if (n) std::memcpy(rec.sig.data(), r.data() + p, n * 2);
    p += n * 2;
  } else {
    if (p + n > r.size() || !svb_zd_decode(r.data() + p, n, rec.sig)) return false;
    p += n;
  }
  rec.aux.assign(r.begin() + p, r.end());
  return true;
}

inline bool load_binary(const std::vector<uint8_t>& d, File& f) {
  if (d.size() < 68 + 5) return false;
  const uint8_t vmaj = d[6], vmin = d[7];
  f.rec_press = d[9];
  f.sig_press = (vmaj > 0 || vmin >= 2) ? d[14] : 0;
  if (f.rec_press > 1 || f.sig_press > 1) {
    f.err = "record compression " + std::to_string(f.rec_press) + ", signal compression " + std::to_string(f.sig_press) +
            " not supported (none, zlib and svb-zd are)";
    return false;
  }
  uint32_t hs;
  std::memcpy(&hs, d.data() + 64, 4);
  size_t p = 68 + (size_t)hs;
  if (p > d.size()) return false;
  f.header.assign(d.begin(), d.begin() + p);
  std::vector<uint8_t> raw;
  while (p + 5 <= d.size() && std::memcmp(d.data() + p, "5WOLB", 5) != 0) {
    uint64_t sz;
    if (p + 8 > d.size()) return false;
    std::memcpy(&sz, d.data() + p, 8);
    p += 8;
    if (p + sz > d.size()) return false;
    if (f.rec_press == 1) {
      if (!inflate_all(d.data() + p, sz, raw)) return false;
    } else {
      raw.assign(d.begin() + p, d.begin() + p + sz);
    }
    p += sz;
    Record rec;
    if (!parse_record(raw, f.sig_press, rec)) return false;
    f.recs.push_back(std::move(rec));
  }
  return p + 5 <= d.size();
}

inline bool load_text(const std::string& txt, File& f) {
  std::istringstream in(txt);
  std::string line;
  int col_len = 6, col_sig = 7;
  while (std::getline(in, line)) {
    if (line.empty()) continue;
    if (line[0] == '@') continue;
    if (line[0] == '#') {
      if (line.compare(0, 8, "#read_id") == 0) {
        std::istringstream cs(line.substr(1));
        std::string c;
        int k = 0;
        while (std::getline(cs, c, '\t')) {
          if (c == "len_raw_signal") col_len = k;
          if (c == "raw_signal") col_sig = k;
          k++;
        }
      }
      continue;
    }
    std::vector<std::string> cols;
    std::istringstream ls(line);
    std::string c;
    while (std::getline(ls, c, '\t')) cols.push_back(c);
    if ((int)cols.size() <= std::max(col_len, col_sig)) return false;
    Record rec;
    rec.id = cols[0];
    const unsigned long long n = std::strtoull(cols[col_len].c_str(), nullptr, 10);
    std::istringstream ss(cols[col_sig]);
    std::string v;
    while (std::getline(ss, v, ',')) rec.sig.push_back((int16_t)std::strtol(v.c_str(), nullptr, 10));
    if (rec.sig.size() != n) return false;
    f.recs.push_back(std::move(rec));
  }
  return true;
}

// 0 on success, SLOW5_FILE_UNREADABLE otherwise.
inline int load(const char* path, File& f) {
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) {
    f.err = "cannot open";
    return SLOW5_FILE_UNREADABLE;
  }
  std::vector<uint8_t> d((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  f.binary = d.size() >= 6 && std::memcmp(d.data(), "BLOW5\1", 6) == 0;
  const bool ok = f.binary ? load_binary(d, f) : load_text(std::string(d.begin(), d.end()), f);
  if (!ok) {
    if (f.err.empty()) f.err = "malformed file";
    return SLOW5_FILE_UNREADABLE;
  }
  return 0;
}

// A BLOW5 file with the template's header and, per record, its stored
// fields with the given signal, in the template's compression.
inline bool save_binary(const char* path, const File& tmpl, const std::vector<const Record*>& recs,
                        const std::vector<const std::vector<int16_t>*>& sigs) {
  std::ofstream out(path, std::ios::binary);
  if (!out.is_open()) return false;
  out.write((const char*)tmpl.header.data(), (std::streamsize)tmpl.header.size());
  std::vector<uint8_t> r, enc, z;
  for (size_t k = 0; k < recs.size(); k++) {
    const Record& rec = *recs[k];
    const std::vector<int16_t>& s = *sigs[k];
    r = rec.head;
    if (tmpl.sig_press == 1) {
      svb_zd_encode(s, enc);
      put<uint64_t>(r, (uint64_t)enc.size());
      r.insert(r.end(), enc.begin(), enc.end());
    } else {
      put<uint64_t>(r, (uint64_t)s.size());
      const uint8_t* b = (const uint8_t*)s.data();
      r.insert(r.end(), b, b + s.size() * 2);
    }
    r.insert(r.end(), rec.aux.begin(), rec.aux.end());
    const std::vector<uint8_t>* blob = &r;
    if (tmpl.rec_press == 1) {
      if (!deflate_all(r, z)) return false;
      blob = &z;
    }
    const uint64_t sz = blob->size();
    out.write((const char*)&sz, 8);
    out.write((const char*)blob->data(), (std::streamsize)blob->size());
  }
  out.write("5WOLB", 5);
  return (bool)out;
}

}  // namespace odba_slow5

/* scan_slow5_data (cpu_utils.hpp:136-165): the number of reads in a file. */
inline int scan_slow5_data(const char* slow5_file_name, size_t* num_sequences) {
  odba_slow5::File f;
  if (odba_slow5::load(slow5_file_name, f)) {
    std::cerr << "Could not open SLOW5 file " << slow5_file_name << std::endl;
    return SLOW5_FILE_UNREADABLE;
  }
  *num_sequences = f.recs.size();
  return 0;
}

/* read_slow5_data<T> (cpu_utils.hpp:235-278): every read's raw signal as T,
 * its read id as the name; returns the count, 0 if the file cannot be
 * opened, SLOW5_FILE_UNREADABLE if it is malformed. */
template <typename T>
int read_slow5_data(const char* slow5_file_name, T** sequences, char** sequence_names, size_t* sequence_lengths) {
  odba_slow5::File f;
  const int rc = odba_slow5::load(slow5_file_name, f);
  if (rc) {
    if (f.err == "cannot open") return 0;  // the scan reported it
    std::cerr << "Error reading SLOW5 file " << slow5_file_name << ": " << f.err << std::endl;
    return SLOW5_FILE_UNREADABLE;
  }
  int k = 0;
  for (const odba_slow5::Record& r : f.recs) {
    T* s = (T*)std::malloc(sizeof(T) * (r.sig.size() ? r.sig.size() : 1));
    for (size_t j = 0; j < r.sig.size(); j++) s[j] = (T)r.sig[j];
    sequences[k] = s;
    sequence_lengths[k] = r.sig.size();
    sequence_names[k] = (char*)std::malloc(r.id.size() + 1);
    std::memcpy(sequence_names[k], r.id.c_str(), r.id.size() + 1);
    k++;
  }
  return k;
}

/* readSequenceSLOW5Files<T> (cpu_utils.hpp:458-500). */
template <typename T>
int readSequenceSLOW5Files(char** filenames, int num_files, T*** sequences, char*** sequence_names,
                           size_t** sequence_lengths) {
  std::cerr << "Step 1 of 3: Loading " << num_files << (num_files == 1 ? " S/BLOW5 file" : " S/BLOW5 files");
  size_t total = 0;
  for (int i = 0; i < num_files; ++i) {
    size_t c = 0;
    scan_slow5_data(filenames[i], &c);
    total += c;
  }
  std::cerr << ", total sequence count " << total << std::endl;
  *sequences = (T**)std::calloc(total ? total : 1, sizeof(T*));
  *sequence_names = (char**)std::calloc(total ? total : 1, sizeof(char*));
  *sequence_lengths = (size_t*)std::calloc(total ? total : 1, sizeof(size_t));
  int count = 0;
  for (int i = 0; i < num_files; ++i) {
    const int got = read_slow5_data<T>(filenames[i], (*sequences) + count, (*sequence_names) + count,
                                       (*sequence_lengths) + count);
    if (got < 1 || got == SLOW5_FILE_UNREADABLE)
      std::cerr << "Error reading in SLOW5 file " << filenames[i] << ", skipping" << std::endl;
    else
      count += got;
  }
  return count;
}

/* writeSlow5Output (io_utils.hpp:363-452): a copy of the named reads of
 * slow5_file_name, in the given order, each with its raw signal replaced by
 * the given sequence of the same length, written as new_slow5_file (BLOW5,
 * the input's compression). Returns 1 on a failure, 0 on success. */
inline int writeSlow5Output(const char* slow5_file_name, const char* new_slow5_file, char** sequence_names,
                            short** sequences, size_t* sequence_lengths, int num_sequences) {
  odba_slow5::File f;
  if (odba_slow5::load(slow5_file_name, f)) {
    std::cerr << "Error opening Slow5 file " << slow5_file_name << " Exiting." << std::endl;
    return 1;
  }
  if (!f.binary) {  // a text template has no binary header to copy
    std::cerr << "Error opening Slow5 file " << slow5_file_name << " as BLOW5 (binary) Exiting." << std::endl;
    return 1;
  }
  std::map<std::string, size_t> by_id;
  for (size_t k = 0; k < f.recs.size(); k++) by_id.emplace(f.recs[k].id, k);
  std::vector<std::vector<int16_t>> sigs(num_sequences);
  std::vector<const odba_slow5::Record*> recs(num_sequences);
  std::vector<const std::vector<int16_t>*> sp(num_sequences);
  for (int i = 0; i < num_sequences; i++) {
    auto it = by_id.find(sequence_names[i]);
    if (it == by_id.end()) {
      std::cerr << "Error. Sequence " << sequence_names[i] << " does not exist in Slow5 file " << slow5_file_name
                << " Exiting." << std::endl;
      return 1;
    }
    const odba_slow5::Record& r = f.recs[it->second];
    if (r.sig.size() != sequence_lengths[i]) {
      std::cerr << "Length of sequence " << sequence_names[i] << " in Slow5 file " << slow5_file_name << " ("
                << r.sig.size() << ") does not match length of sequence given (" << sequence_lengths[i]
                << ") Exiting." << std::endl;
      return 1;
    }
    sigs[i].assign(sequences[i], sequences[i] + sequence_lengths[i]);
    recs[i] = &r;
    sp[i] = &sigs[i];
  }
  if (!odba_slow5::save_binary(new_slow5_file, f, recs, sp)) {
    std::cerr << "Error writing new sequences to new Slow5 file " << new_slow5_file << " Exiting." << std::endl;
    return 1;
  }
  return 0;
}

#endif  // ODBA_SLOW5_HPP
