// compare.cpp -- thv_compare: element-wise diff of two values.bin files.
//
// Same flags and tolerances as the reference's comparer
// (comparer/compare.cpp:46-208): --input1 A --input2 B --vertices N
// --type 0 (u32, exact) | 1 (u64, exact) | 2 (f64, |a-b| <= 1e-3).
// Prints the first 20 mismatches.  Unlike the reference it reads exactly
// N elements of the requested width (compare.cpp:59-60 reads n*8 bytes into
// a u32 array), fails on short files, and exits 0 only when the files agree.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {

template <class T>
bool load(const std::string &path, uint64_t n, std::vector<T> &out) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) {
    std::cerr << "Could not open file: " << path << std::endl;
    return false;
  }
  out.resize(n);
  f.read(reinterpret_cast<char *>(out.data()), (std::streamsize)(n * sizeof(T)));
  if ((uint64_t)f.gcount() != n * sizeof(T)) {
    std::cerr << path << " holds fewer than " << n << " values" << std::endl;
    return false;
  }
  return true;
}

template <class T, class Eq>
int compare(const std::string &a, const std::string &b, uint64_t n, Eq eq) {
  std::vector<T> x, y;
  if (!load(a, n, x) || !load(b, n, y)) return 2;
  uint64_t bad = 0;
  for (uint64_t i = 0; i < n; i++) {
    if (eq(x[i], y[i])) continue;
    if (bad < 20) {
      std::cout << "Arrays do not match at: " << i << std::endl;
      std::cout << x[i] << " | " << y[i] << std::endl;
    }
    bad++;
  }
  std::cout << "Mismatches: " << bad << " of " << n << std::endl;
  return bad ? 1 : 0;
}

}  // namespace

int main(int argc, char **argv) {
  std::string a, b;
  uint64_t n = 0;
  int type = 0;
  for (int i = 1; i < argc - 1; i += 2) {
    if (!strcmp(argv[i], "--input1")) a = argv[i + 1];
    else if (!strcmp(argv[i], "--input2")) b = argv[i + 1];
    else if (!strcmp(argv[i], "--type")) type = atoi(argv[i + 1]);
    else if (!strcmp(argv[i], "--vertices")) n = strtoull(argv[i + 1], nullptr, 10);
  }
  if (a.empty() || b.empty()) {
    std::cout << "Usage: " << argv[0]
              << " --input1 A --input2 B --vertices N --type 0 (u32) | 1 (u64) | 2 (f64)"
              << std::endl;
    return 2;
  }
  std::cout << "Number of Vertices: " << n << std::endl;
  switch (type) {
    case 0:
      return compare<uint32_t>(a, b, n, [](uint32_t p, uint32_t q) { return p == q; });
    case 1:
      return compare<uint64_t>(a, b, n, [](uint64_t p, uint64_t q) { return p == q; });
    case 2:
      return compare<double>(a, b, n, [](double p, double q) { return std::abs(p - q) <= 1e-3; });
    default:
      std::cerr << "unknown --type " << type << std::endl;
      return 2;
  }
}
