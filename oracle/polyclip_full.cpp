// polyclip_full.cpp -- test infrastructure only (the oracle's checker).
//
// Builds the reference's sequential Foster clipper
// (lib/optimizedFostersAlgorithm/polyclip.cpp, compiled where it lies under
// the reference tree; its path comes in as REF_POLYCLIP_SOURCE) with every
// output file stream it opens set to 17 significant digits.  The reference
// writes its result rings through savePolygon (polyclip.cpp:897-905), an
// ofstream at the stream default of 6 digits; here that stream is a subclass
// whose constructor raises the precision, so the same program prints every
// coordinate so that it reads back as the identical double.  Nothing else of
// the program changes.
#include <fstream>
#include <iostream>
#include <string>

namespace segclip_oracle {
struct full_precision_ofstream : std::ofstream {
    explicit full_precision_ofstream(const std::string &path) : std::ofstream(path) { precision(17); }
};
} // namespace segclip_oracle
using segclip_oracle::full_precision_ofstream;

#define ofstream full_precision_ofstream
#include REF_POLYCLIP_SOURCE
