// Light-curve input files (isakruas/fbls src/io/npz.rs:17-113,
// src/main.rs:50-132): NPZ archives (stored or deflated zip members of NPY
// arrays, zip64 records included) decoded to doubles, and the merge of
// several files into one observation set with per-(file, sector) groups.
#pragma once
#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace fbls {
namespace io {

struct Array {
  std::vector<size_t> shape;
  std::vector<double> data;
};

// npz.rs:17-96: one NPY member; dtypes f8 f4 i8 i4 i2 u1 b1, either byte
// order, C order only.
bool parse_npy(const unsigned char* bytes, size_t len, Array* out, std::string* error);
// npz.rs:98-113: every member, named without its ".npy" suffix.
bool read_npz(const std::string& path, std::map<std::string, Array>* out, std::string* error);

struct Observations {
  std::vector<double> time, flux, err;
  bool has_err = false;
  std::vector<int64_t> group;
};

// main.rs:50-132 for NPZ inputs: time_bjd and flux required, flux_err in
// every file or in none, sector a scalar or one value per observation;
// groups number the (file, sector) pairs in order of first appearance.
bool read_inputs(const std::vector<std::string>& paths, Observations* out, std::string* error);

}  // namespace io
}  // namespace fbls
