// Test program of the product's host exact sums (autan_amd/csrc/device/exact.hpp), built with g++:
//   exact_check sum    < one double per line (C99 hex float)  -> the exact sum, rounded once
//   exact_check words  < lines of 71 integers (68 limbs, +inf, -inf, NaN counts) -> their value
// Each result is printed as a C99 hex float on its own line ("nan", "inf", "-inf" as printf gives).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>

#include "../../autan_amd/csrc/device/exact.hpp"

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: exact_check sum|words\n");
    return 2;
  }
  std::string line;
  if (std::strcmp(argv[1], "sum") == 0) {
    mxe::ExactSum s;
    while (std::getline(std::cin, line)) {
      if (!line.empty()) s.add(std::strtod(line.c_str(), nullptr));
    }
    std::printf("%a\n", s.value());
    return 0;
  }
  if (std::strcmp(argv[1], "words") == 0) {
    while (std::getline(std::cin, line)) {
      if (line.empty()) continue;
      std::istringstream in(line);
      long long w[mxe::kExactWords];
      for (long long& v : w) {
        if (!(in >> v)) {
          std::fprintf(stderr, "short line\n");
          return 2;
        }
      }
      std::printf("%a\n", mxe::exact_value(w));
    }
    return 0;
  }
  return 2;
}
