// Test infrastructure: prints what the reference runtime's JSON library (nlohmann::ordered_json,
// runtime/src/run.cpp:63) writes for history values and lines, so that the history writer's
// text can be checked against it. Built by oracle/build_ref.sh from the library's header where
// it lies in the reference tree; nothing in the product links it.
//
// stdin, one request per line:
//   n BITS                       -> dump() of the double with these 16 hex bits
//   s ITERATION PROV BITS*5      -> a steady history line (time null, no monitor)
//   u ITERATION TIMEBITS BITS*5  -> an unsteady history line (no provisional key)
// One output line per request.
#include <cstdint>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>

#include <nlohmann/json.hpp>

namespace {

double from_bits(const std::string& hex) {
  const std::uint64_t b = std::stoull(hex, nullptr, 16);
  double x;
  std::memcpy(&x, &b, sizeof x);
  return x;
}

using Json = nlohmann::ordered_json;

Json residuals(std::istringstream& in) {
  static const char* names[5] = {"density", "momentum-x", "momentum-y", "momentum-z", "energy"};
  Json r = Json::object();
  for (const char* name : names) {
    std::string h;
    in >> h;
    const double x = from_bits(h);
    r[name] = std::isfinite(x) ? Json(x) : Json(nullptr);
  }
  return r;
}

}  // namespace

int main() {
  std::string line;
  while (std::getline(std::cin, line)) {
    std::istringstream in(line);
    std::string kind;
    in >> kind;
    if (kind == "n") {
      std::string h;
      in >> h;
      std::cout << Json(from_bits(h)).dump() << "\n";
    } else if (kind == "s" || kind == "u") {
      long long it = 0;
      std::string second;
      in >> it >> second;
      const bool steady = kind == "s";
      Json record{{"iteration", it},
                  {"time", steady ? Json(nullptr) : Json(from_bits(second))},
                  {"residuals", residuals(in)},
                  {"monitors", Json::object()}};
      if (steady) record["provisional"] = second == "1";
      std::cout << record.dump() << "\n";
    } else {
      std::cout << "?\n";
    }
  }
  return 0;
}
