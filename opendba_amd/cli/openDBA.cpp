// openDBA command line (same contract as openDBA.cu:11-158):
//   openDBA [-n] <binary|text|tsv> <short|int|uint|ulong|float|double>
//           <global|open_start|open_end|open|open_prefix_S_L> <output prefix>
//           <minseg[,minseg2]> <prefix file|/dev/null> <cdist> <series...>
// Options are parsed with getopt before the positional arguments are read
// from argv[1...], as in the reference. GNU getopt permutes "-n" to argv[1],
// where it is then rejected as a read mode, so, as with the reference
// binary, a "-n" anywhere ends the run with exit code 1.
#include <getopt.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/odba/opendba.hpp"

int main(int argc, char** argv) {
  int norm_sequences = 1;
  int prefix_to_skip = 0;
  int prefix_length = 0;
  int c;
  while ((c = getopt(argc, argv, "n")) != -1) {
    if (c == 'n') norm_sequences = 0;
  }
  if (argc < 9) {
    std::cout << "Usage: " << argv[0] << " <binary|text|tsv|slow5> <short|int|uint|ulong|float|double> "
              << "<global|open_start|open_end|open> <output files prefix> <minimum unimodal segment length for "
                 "clustering[,for consensus generation]> <prefix sequence to remove|/dev/null> <clustering threshold> "
                 "<series.tsv|<series1> <series2> [series3...]>\n";
    std::exit(1);
  }
  const int num_series = argc - 8;
  char* min_segment_length = argv[5];
  int read_mode = TEXT_READ_MODE;
  if (!std::strcmp(argv[1], "binary")) read_mode = BINARY_READ_MODE;
  else if (!std::strcmp(argv[1], "tsv")) read_mode = TSV_READ_MODE;
  else if (!std::strcmp(argv[1], "slow5")) read_mode = SLOW5_READ_MODE;
  else if (std::strcmp(argv[1], "text")) {
    std::cerr << "First argument (" << argv[1] << ") is neither 'binary' nor 'text'" << std::endl;
    std::exit(1);
  }
  int use_open_start = 0, use_open_end = 0;
  if (!std::strcmp(argv[3], "global")) {
  } else if (!std::strcmp(argv[3], "open_start")) {
    use_open_start = 1;
  } else if (!std::strcmp(argv[3], "open_end")) {
    use_open_end = 1;
  } else if (!std::strncmp(argv[3], "open_prefix", 11)) {
    use_open_start = 0;
    use_open_end = 1;
    norm_sequences = 1;
    // "open_prefix_<skip>_<length>": exactly two numbers after the mode name
    const std::string spec(std::strlen(argv[3]) > 12 ? argv[3] + 12 : "");
    const size_t cut = spec.find('_');
    if (cut == std::string::npos || spec.find('_', cut + 1) != std::string::npos) {
      std::cerr << "Unexpected alignment type specified, expected open_prefix_##_## but did not find a second "
                   "underscore in " << argv[3] << std::endl;
      std::exit(1);
    }
    prefix_to_skip = std::stoi(spec.substr(0, cut));
    prefix_length = std::stoi(spec.substr(cut + 1));
  } else if (!std::strcmp(argv[3], "open")) {
    use_open_start = 1;
    use_open_end = 1;
  } else {
    std::cerr << "Third argument (" << argv[3] << ") is not one of the accept values 'global', 'open_start', "
                 "'open_end' or 'open'" << std::endl;
    std::exit(1);
  }
  char* output_prefix = argv[4];
  char* seqprefix = std::strcmp(argv[6], "/dev/null") ? argv[6] : nullptr;
  const double cdist = std::atof(argv[7]);
  char** files = &argv[8];
  if (!std::strcmp(argv[2], "int"))
    setupAndRun<int>(seqprefix, files, num_series, output_prefix, read_mode, use_open_start, use_open_end, min_segment_length, norm_sequences, cdist, prefix_to_skip, prefix_length);
  else if (!std::strcmp(argv[2], "uint"))
    setupAndRun<unsigned int>(seqprefix, files, num_series, output_prefix, read_mode, use_open_start, use_open_end, min_segment_length, norm_sequences, cdist, prefix_to_skip, prefix_length);
  else if (!std::strcmp(argv[2], "ulong"))
    setupAndRun<unsigned long long>(seqprefix, files, num_series, output_prefix, read_mode, use_open_start, use_open_end, min_segment_length, norm_sequences, cdist, prefix_to_skip, prefix_length);
  else if (!std::strcmp(argv[2], "float"))
    setupAndRun<float>(seqprefix, files, num_series, output_prefix, read_mode, use_open_start, use_open_end, min_segment_length, norm_sequences, cdist, prefix_to_skip, prefix_length);
  else if (!std::strcmp(argv[2], "double"))
    setupAndRun<double>(seqprefix, files, num_series, output_prefix, read_mode, use_open_start, use_open_end, min_segment_length, norm_sequences, cdist, prefix_to_skip, prefix_length);
  else if (!std::strcmp(argv[2], "short"))
    setupAndRun<float>(seqprefix, files, num_series, output_prefix, read_mode, use_open_start, use_open_end, min_segment_length, norm_sequences, cdist, prefix_to_skip, prefix_length, true);
  else {
    std::cerr << "Second argument (" << argv[2] << ") was not one of the accepted numerical representations: 'int', "
                 "'uint', 'ulong', 'float' or 'double'" << std::endl;
    std::exit(1);
  }
  return 0;
}
