// Basis manifests (QMBED src/archive.rs:246-416): a versioned zip archive of
// ordered basis states with their logical width and scalar metadata.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "qmbed.hpp"

namespace qmbed {

// BasisArchive (src/archive.rs:246-331): states in canonical decimal, in row
// order, unique; metadata keys in sorted order.
struct BasisArchive {
  size_t width_bits = 0;
  std::vector<std::string> states;
  std::map<std::string, std::string> metadata;
  void insert_metadata(const std::string& key, const std::string& value);
};

// BasisArchive::new over states given as decimal text or as 64-bit words
BasisArchive basis_archive(size_t width_bits, const std::vector<std::string>& states);
BasisArchive basis_archive(size_t width_bits, const uint64_t* states, size_t n);
void save_basis_zip(const std::string& path, const BasisArchive& basis);
BasisArchive load_basis_zip(const std::string& path);
// whether a canonical decimal state fits one 64-bit word
bool fits_u64(const std::string& canonical);

}  // namespace qmbed
