// C++ face of libmxeuler (mxeuler.h) in the shape of the reference's device stepper:
// DeviceStepper<P> and make_device_stepper<P> (physics/include/autan/physics/euler_device.hpp:22-54).
// The same Result, the same six calls with std::span and std::optional arguments, the same
// exception kinds (std::invalid_argument for refused input, std::runtime_error for a device failure,
// its message starting E1101, as physics/src/device_stepper.cuh:179-193 throws them). Header only:
// a caller links libmxeuler.so and needs nothing else.
//
// Where the reference's factory reads a host Euler<P> and packs its geometry, this one takes the
// arrays themselves (mxeuler_topology, mxeuler_geometry, mxeuler_stencil, boundary codes and data,
// the source); ranks, ghosts and global keys have no counterpart on one GPU.
#ifndef MXEULER_HPP
#define MXEULER_HPP

#include <cstddef>
#include <cstdint>
#include <memory>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mxeuler.h"

namespace mxeuler {

// Precision policies (codegen/include/autan/kernels/runtime.hpp:73-99).
struct fp64 {
  static constexpr int32_t policy = MXEULER_FP64;
};
struct fp32_acc64 {
  static constexpr int32_t policy = MXEULER_FP32_ACC64;
};

namespace detail {

struct Message {
  char text[1024] = {};
};

// The library's return value back into the exception the reference's stepper throws.
inline void raise(int rc, const Message& m) {
  if (rc == 0) return;
  if (rc == 1) throw std::invalid_argument(m.text);
  throw std::runtime_error(m.text);
}

inline void expect(bool ok, const char* what) {
  if (!ok) throw std::invalid_argument(what);
}

}  // namespace detail

template <class P>
class DeviceStepper {
 public:
  struct Result {
    std::vector<double> residuals;   // RMS over the cells of R at the step's start state (5)
    double time_step = 0.0;          // the global time step; the smallest local one if local
    std::optional<int32_t> invalid;  // the least invalid cell after the step
  };

  // make_device_stepper's arguments: bc_kind holds one BoundaryKind code per boundary face (whole
  // numbers as FP64), bc 5 values per boundary face (component-major), source 5 x n_cells per unit
  // volume or empty for none; stream is a hipStream_t of the /opt/rocm runtime, or null for the
  // library's own stream.
  DeviceStepper(const mxeuler_scheme& scheme, const mxeuler_topology& topology, const mxeuler_geometry& geometry,
                const mxeuler_stencil& stencil, std::span<const double> bc_kind, std::span<const double> bc,
                std::span<const double> source, void* stream = nullptr)
      : n_(topology.n_cells) {
    detail::expect(topology.n_cells > 0 && topology.n_faces >= topology.n_interior && topology.n_interior >= 0,
                   "topology sizes");
    const std::size_t nb = static_cast<std::size_t>(topology.n_faces - topology.n_interior);
    detail::expect(bc_kind.size() == nb && bc.size() == 5 * nb, "boundary arrays do not match the boundary faces");
    detail::expect(source.empty() || source.size() == cells5(), "source array does not match the cells");
    detail::Message m;
    detail::raise(mxeuler_create(P::policy, &scheme, &topology, &geometry, &stencil, bc_kind.data(), bc.data(),
                                 source.empty() ? nullptr : source.data(), stream, &h_, m.text, sizeof m.text),
                  m);
  }
  ~DeviceStepper() {
    if (h_ != nullptr) mxeuler_destroy(h_);
  }
  DeviceStepper(const DeviceStepper&) = delete;
  DeviceStepper& operator=(const DeviceStepper&) = delete;
  DeviceStepper(DeviceStepper&& o) noexcept : h_(std::exchange(o.h_, nullptr)), n_(o.n_) {}
  DeviceStepper& operator=(DeviceStepper&& o) noexcept {
    if (this != &o) {
      if (h_ != nullptr) mxeuler_destroy(h_);
      h_ = std::exchange(o.h_, nullptr);
      n_ = o.n_;
    }
    return *this;
  }

  // One step from the device state: time step `dt` if given, else min CFL V / lambda over the
  // cells, or with `local` CFL V / lambda per cell (pseudo-time).
  Result step(double cfl, std::optional<double> dt, bool local) {
    Result r;
    r.residuals.resize(5);
    int32_t invalid = -1;
    detail::Message m;
    detail::raise(mxeuler_step(live(), cfl, dt.has_value(), dt.value_or(0.0), local, r.residuals.data(), &r.time_step,
                               &invalid, m.text, sizeof m.text),
                  m);
    if (invalid >= 0) r.invalid = invalid;
    return r;
  }

  // The least invalid cell of the device state, and the RMS residual of that state.
  Result evaluate() {
    Result r;
    r.residuals.resize(5);
    int32_t invalid = -1;
    detail::Message m;
    detail::raise(mxeuler_evaluate(live(), r.residuals.data(), &invalid, m.text, sizeof m.text), m);
    if (invalid >= 0) r.invalid = invalid;
    return r;
  }

  // R of `u` (5 x n_cells, FP64) into `res` (5 x n_cells); the device state is kept.
  void residual(std::span<const double> u, std::span<double> res) {
    detail::expect(u.size() == cells5() && res.size() == cells5(), "state arrays do not match the cells");
    detail::Message m;
    detail::raise(mxeuler_residual(live(), u.data(), res.data(), m.text, sizeof m.text), m);
  }

  // Switches the linear reconstruction on or off, as Euler<P>::set_second_order.
  void set_second_order(bool on) {
    detail::Message m;
    detail::raise(mxeuler_set_second_order(live(), on, m.text, sizeof m.text), m);
  }

  void upload(std::span<const double> u) {
    detail::expect(u.size() == cells5(), "state array does not match the cells");
    detail::Message m;
    detail::raise(mxeuler_upload(live(), u.data(), m.text, sizeof m.text), m);
  }

  void download(std::span<double> u) {
    detail::expect(u.size() == cells5(), "state array does not match the cells");
    detail::Message m;
    detail::raise(mxeuler_download(live(), u.data(), m.text, sizeof m.text), m);
  }

  int32_t n_cells() const { return n_; }
  // The C handle, for the library's calls beyond the reference's interface (loads, steady runs,
  // conservation, tuning); it stays owned by this object.
  mxeuler_stepper* handle() const { return h_; }

 private:
  std::size_t cells5() const { return 5 * static_cast<std::size_t>(n_); }
  mxeuler_stepper* live() const {
    if (h_ == nullptr) throw std::logic_error("the stepper was moved from");
    return h_;
  }

  mxeuler_stepper* h_ = nullptr;
  int32_t n_ = 0;
};

// make_device_stepper<P> (physics/include/autan/physics/euler_device.hpp:49-54) over the arrays.
template <class P>
std::unique_ptr<DeviceStepper<P>> make_device_stepper(const mxeuler_scheme& scheme, const mxeuler_topology& topology,
                                                      const mxeuler_geometry& geometry, const mxeuler_stencil& stencil,
                                                      std::span<const double> bc_kind, std::span<const double> bc,
                                                      std::span<const double> source, void* stream = nullptr) {
  return std::make_unique<DeviceStepper<P>>(scheme, topology, geometry, stencil, bc_kind, bc, source, stream);
}

}  // namespace mxeuler

#endif
