// Drives the C++ stepper adapter (include/mxeuler.hpp) the way the reference's runtime drives its
// DeviceStepper<P> (physics/src/euler_physics.cpp:1515-1571, physics/bench/euler_step_bench.cpp),
// for tests/test_adapter.py: the case's arrays come from raw little-endian files in a directory,
// the results go back into it.
//
//   adapter_check DIR
//
// DIR/case.txt holds "key value..." lines (policy, the scheme's fields, cfl, dt); the arrays are
// DIR/<name>.i32, .i64 or .f64, their sizes taken from the files. Writes DIR/res.f64 (R of u),
// DIR/u_out.f64 (the state after the steps) and DIR/results.txt, one line per call:
// "name r0 r1 r2 r3 r4 time_step invalid" (invalid -1 for none), doubles as %.17g.
// Exit 0 on success, 3 on std::invalid_argument, 4 on std::runtime_error, 5 on another exception,
// 6 when this program's own inputs are wrong; the exception's kind and message on stdout.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "mxeuler.hpp"

namespace {

std::string dir;

// A problem with the inputs of this program, not with the library.
struct HarnessError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <class T>
std::vector<T> read(const std::string& name, bool required = true) {
  std::ifstream f(dir + "/" + name, std::ios::binary | std::ios::ate);
  if (!f) {
    if (required) throw HarnessError("missing input " + name);
    return {};
  }
  const std::streamsize bytes = f.tellg();
  if (bytes % static_cast<std::streamsize>(sizeof(T)) != 0) throw HarnessError("ragged input " + name);
  std::vector<T> v(static_cast<std::size_t>(bytes) / sizeof(T));
  f.seekg(0);
  f.read(reinterpret_cast<char*>(v.data()), bytes);
  return v;
}

void write(const std::string& name, const std::vector<double>& v) {
  std::ofstream f(dir + "/" + name, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(double)));
}

std::map<std::string, std::vector<double>> read_case() {
  std::ifstream f(dir + "/case.txt");
  if (!f) throw HarnessError("missing case.txt");
  std::map<std::string, std::vector<double>> kv;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream s(line);
    std::string key;
    if (!(s >> key)) continue;
    double x;
    while (s >> x) kv[key].push_back(x);
  }
  return kv;
}

template <class R>
void report(std::FILE* out, const char* name, const R& r) {
  std::fprintf(out, "%s", name);
  for (double x : r.residuals) std::fprintf(out, " %.17g", x);
  std::fprintf(out, " %.17g %d\n", r.time_step, r.invalid ? static_cast<int>(*r.invalid) : -1);
}

template <class P>
void run(const std::map<std::string, std::vector<double>>& kv) {
  auto get = [&](const char* k) {
    const auto it = kv.find(k);
    if (it == kv.end() || it->second.empty()) throw HarnessError(std::string("case.txt lacks ") + k);
    return it->second;
  };
  mxeuler_scheme sc{};
  sc.flux = static_cast<int32_t>(get("flux")[0]);
  sc.second_order = static_cast<int32_t>(get("second_order")[0]);
  sc.limiter = static_cast<int32_t>(get("limiter")[0]);
  sc.planar = static_cast<int32_t>(get("planar")[0]);
  sc.gamma = get("gamma")[0];
  sc.gas_constant = get("gas_constant")[0];
  sc.mach_ref = get("mach_ref")[0];
  sc.venkat_k = get("venkat_k")[0];
  sc.length = get("length")[0];
  for (int k = 0; k < 5; ++k) {
    sc.scale[k] = get("scale").at(k);
    sc.reference[k] = get("reference").at(k);
  }

  const auto owner = read<int32_t>("owner.i32");
  const auto neighbour = read<int32_t>("neighbour.i32");
  const auto cfo = read<int64_t>("cell_face_offsets.i64");
  const auto cell_faces = read<int32_t>("cell_faces.i32");
  const auto normal = read<double>("face_normal.f64");
  const auto area = read<double>("face_area.f64");
  const auto r_left = read<double>("r_left.f64");
  const auto r_right = read<double>("r_right.f64");
  const auto moment = read<double>("face_moment.f64", false);
  const auto volume = read<double>("cell_volume.f64");
  const auto st_off = read<int64_t>("stencil_offsets.i64");
  const auto st_ent = read<int32_t>("stencil.i32");
  const auto st_w = read<double>("stencil_weights.f64");
  const auto bc_kind = read<double>("bc_kind.f64");
  const auto bc = read<double>("bc.f64");
  const auto source = read<double>("source.f64", false);
  const auto u = read<double>("u.f64");

  const int32_t n = static_cast<int32_t>(volume.size());
  mxeuler_topology t{n,
                     static_cast<int32_t>(area.size()),
                     static_cast<int32_t>(neighbour.size()),
                     owner.data(),
                     neighbour.data(),
                     cfo.data(),
                     cell_faces.data()};
  mxeuler_geometry g{normal.data(), area.data(), r_left.data(), r_right.data(),
                     moment.empty() ? nullptr : moment.data(), volume.data(),
                     static_cast<int32_t>(get("all_planar")[0])};
  mxeuler_stencil st{st_off.data(), st_ent.data(), st_w.data()};

  auto dev = mxeuler::make_device_stepper<P>(sc, t, g, st, bc_kind, bc, source);
  const double cfl = get("cfl")[0];
  const double dt = get("dt")[0];

  std::FILE* out = std::fopen((dir + "/results.txt").c_str(), "w");
  if (out == nullptr) throw HarnessError("cannot write results.txt");
  dev->upload(u);
  std::vector<double> res(u.size());
  dev->residual(u, res);
  write("res.f64", res);
  report(out, "evaluate", dev->evaluate());
  report(out, "step_global", dev->step(cfl, std::nullopt, false));
  report(out, "step_dt", dev->step(cfl, dt, false));
  dev->set_second_order(false);
  report(out, "step_local_first_order", dev->step(cfl, std::nullopt, true));
  dev->set_second_order(true);
  // A moved stepper keeps the device state; the moved-from one refuses further calls.
  mxeuler::DeviceStepper<P> moved = std::move(*dev);
  report(out, "step_local", moved.step(cfl, std::nullopt, true));
  try {
    dev->evaluate();
    std::fprintf(out, "moved_from callable\n");
  } catch (const std::logic_error&) {
    std::fprintf(out, "moved_from refused\n");
  }
  std::vector<double> u_out(u.size());
  moved.download(u_out);
  write("u_out.f64", u_out);
  std::fclose(out);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s DIR\n", argv[0]);
    return 2;
  }
  dir = argv[1];
  try {
    const auto kv = read_case();
    const auto it = kv.find("policy");
    const int policy = it == kv.end() || it->second.empty() ? -1 : static_cast<int>(it->second[0]);
    if (policy == MXEULER_FP64) {
      run<mxeuler::fp64>(kv);
    } else if (policy == MXEULER_FP32_ACC64) {
      run<mxeuler::fp32_acc64>(kv);
    } else {
      throw HarnessError("case.txt: policy 0 or 1");
    }
  } catch (const HarnessError& e) {
    std::printf("harness: %s\n", e.what());
    return 6;
  } catch (const std::invalid_argument& e) {
    std::printf("invalid_argument: %s\n", e.what());
    return 3;
  } catch (const std::runtime_error& e) {
    std::printf("runtime_error: %s\n", e.what());
    return 4;
  } catch (const std::exception& e) {
    std::printf("exception: %s\n", e.what());
    return 5;
  }
  std::printf("ok\n");
  return 0;
}
