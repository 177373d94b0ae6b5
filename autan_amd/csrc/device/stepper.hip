// The device stepper behind include/mxeuler.h: device layout, launches and the C ABI.
//
// Mirrors CudaStepper<P> (physics/src/device_stepper.cuh:200-529): one residual is the chain
// primitives -> boundary states -> gradients + limiter -> sources + faces + boundaries; one step
// is R(u) and its exact norms, the spectral radius and time step, and three SSP-RK3 stages.
// Unlike the reference, a step makes no blocking read before its end: the least time step stays
// on the device for the stages, and the norms, the time step and the validity flags come back in
// one copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mxeuler.h"
#include "exact.hpp"
#include "implicit.hpp"
#include "jacobian.hpp"
#include "kernels.hpp"
#include "steady.hpp"
#include "stream.hpp"

namespace mxe {

namespace {

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("E1101: ") + what + ": " + hipGetErrorString(e));
}

int blocks(long long n) { return static_cast<int>((n + kBlock - 1) / kBlock); }
// Workgroups for n items under the XCD-aware block mapping.
int grid(long long n) { return xcd_grid(blocks(n)); }

template <class T>
class Buffer {
 public:
  Buffer() = default;
  explicit Buffer(std::size_t n) : n_(n) {
    if (n > 0) check(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), "hipMalloc");
  }
  Buffer(const Buffer&) = delete;
  Buffer& operator=(const Buffer&) = delete;
  Buffer(Buffer&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  Buffer& operator=(Buffer&& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    return *this;
  }
  ~Buffer() {
    if (p_ != nullptr) (void)hipFree(p_);
  }
  void upload(const T* h, std::size_t n, hipStream_t s) {
    if (n > n_) throw std::invalid_argument("upload larger than the buffer");
    if (n > 0) check(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
  }
  void upload(const std::vector<T>& v, hipStream_t s) {
    *this = Buffer<T>(v.size());
    upload(v.data(), v.size(), s);
  }
  void download(T* h, std::size_t n, hipStream_t s) const {
    if (n > 0) check(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
  }
  T* get() const { return p_; }
  std::size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  std::size_t n_ = 0;
};

// Whole-number boundary codes as the bodies' midpoint comparisons read them
// (kernels/euler.py:391-422).
int code_of(double kind) {
  if (kind < 0.5) return 0;
  if (kind < 1.5) return 1;
  if (kind < 2.5) return 2;
  if (kind < 3.5) return 3;
  if (kind < 4.5) return 4;
  return 5;
}

}  // namespace

// ---------------------------------------------------------------- the static face-force staging
//
// mxeuler_face_force_noalloc stages its face list and its forces here instead of in buffers of its
// own: at most kStaticForceFaces faces, the list (int32) first, the forces (3 FP64 per face, SoA by
// the call's face count) from byte offset kStaticForceOut. Each GPU holds its own copy for the life
// of the process. In host code `scratch` names a host-side stand-in: its device address on each GPU
// comes from hipGetSymbolAddress, taken once per device when a stepper is created there.
constexpr long long kStaticForceFaces = MXEULER_STATIC_FORCE_FACES;
constexpr std::size_t kStaticForceOut = std::size_t(kStaticForceFaces) * sizeof(int);
static __device__ __align__(256) unsigned char scratch[kStaticForceOut + 3 * std::size_t(kStaticForceFaces) * sizeof(double)];
static_assert(kStaticForceOut % 256 == 0, "the forces start on a 256-byte boundary");

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
void* scratch_address[kMaxDevices] = {};

// The calling thread's current device's address of `scratch`, looked up on the first call there.
void* static_scratch() {
  int dev = -1;
  check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("E1101: device ordinal out of range");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (scratch_address[dev] == nullptr) {
    void* ptr = nullptr;
    check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    scratch_address[dev] = ptr;
  }
  return scratch_address[dev];
}

class StepperBase {
 public:
  virtual ~StepperBase() = default;
  virtual void upload(const double* u) = 0;
  virtual void download(double* u) = 0;
  virtual void residual(const double* u, double* res) = 0;
  virtual void residual_and_jacobian(const double* u, double* res, double* dwdu, double* face_jac,
                                     double* bnd_jac) = 0;
  virtual void upload_device(const double* u, hipStream_t caller) = 0;
  virtual void download_device(double* u, hipStream_t caller) = 0;
  virtual void residual_device(const double* u, double* res, hipStream_t caller) = 0;
  virtual void step(double cfl, bool has_dt, double dt, bool local, double* residuals, double* time_step,
                    int* invalid) = 0;
  virtual void evaluate(double* residuals, int* invalid) = 0;
  virtual void set_second_order(bool on) = 0;
  virtual void set_boundary_data(const double* bc) = 0;
  virtual void set_source(const double* source) = 0;
  virtual void spectral_radius(double* lam) = 0;
  virtual void arrays(double* prim, double* grad, double* phi, double* prim_b) = 0;
  virtual void face_force(const int* faces, long long n, double* force) = 0;
  virtual void face_force_noalloc(const int* faces, long long n, double* force) = 0;
  virtual void implicit_evaluate(double* norms, int* invalid) = 0;
  virtual void implicit_solve(double cfl, double* report) = 0;
  virtual void implicit_colours(int* colours) = 0;
  virtual int implicit_worst_cell() = 0;
  virtual void conservation(double* out) = 0;
  virtual void tune(const std::string& key, int value) = 0;
  virtual double area_scale() const = 0;
  virtual int ncase() const = 0;
  virtual void batch_step(double cfl, bool local, double* residuals, double* time_steps, int* invalid) = 0;
  virtual void set_flags(const int* flags) = 0;
  virtual void batch_steady(const mxsteady_options& o, const double* case_scales, const double* case_references,
                            mxsteady_result* results, double* history) = 0;
  virtual std::string autotune(int steps, double cfl, bool local) = 0;
  virtual double gamma() const = 0;
  virtual int profile(int steps, double cfl, bool local, double* ms_step, double* ms_kernel, int* launches,
                      int max_kernels, std::string* names) = 0;
};

template <class C>
class Stepper final : public StepperBase {
 public:
  Stepper(const mxeuler_scheme& sc, const mxeuler_topology& t, const mxeuler_geometry& geo, const mxeuler_stencil& st,
          const double* bc_kind, const double* bc, const double* source, hipStream_t stream)
      : Stepper(&sc, 1, std::array<long long, 2>{0, t.n_cells}.data(), std::array<long long, 1>{t.n_cells}.data(), t, geo,
                st, bc_kind, bc, source, stream) {}

  // ncase cases in one index space: case k's cells are [case_first[k], case_first[k] +
  // case_cells[k]), the rest of its range up to case_first[k + 1] filler cells without faces or
  // stencil; every case_first a multiple of kBlock, the last n_cells. The cases share the flux,
  // order, limiter, planarity and gas; their reference states, scales and cut-off Mach numbers are
  // their own.
  Stepper(const mxeuler_scheme* schemes, int ncase, const long long* case_first, const long long* case_cells,
          const mxeuler_topology& t, const mxeuler_geometry& geo, const mxeuler_stencil& st, const double* bc_kind,
          const double* bc, const double* source, hipStream_t stream)
      : sc_(schemes[0]) {
    validate(schemes[0], t, geo, st);
    validate_cases(schemes, ncase, case_first, case_cells, t.n_cells);
    ncase_ = ncase;
    nslot_ = ncase > 1 ? kCaseSlots : kNormSlots;
    case_first_.assign(case_first, case_first + ncase + 1);
    case_cells_.assign(case_cells, case_cells + ncase);
    if (stream != nullptr) {
      stream_ = stream;
    } else {
      check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
      own_stream_ = true;
    }
    // This device's address of the static face-force staging (mxeuler_face_force_noalloc).
    scratch_ = static_scratch();
    nc_ = t.n_cells;
    nf_ = t.n_faces;
    ni_ = t.n_interior;
    nb_ = nf_ - ni_;
    second_ = sc_.second_order != 0;
    planar_ = sc_.planar != 0;
    for (int k = 0; k < ncase; ++k) ptab_host_.push_back(params(schemes[k], case_first[k], case_first[k] + case_cells[k]));
    ptab_.upload(ptab_host_, stream_);
    if (ncase > 1) {
      std::vector<int> bcase(std::size_t(grid(nc_)), ncase - 1);
      for (int k = 0; k < ncase; ++k) {
        for (long long c = case_first[k]; c < case_first[k + 1]; c += kBlock) bcase[std::size_t(c / kBlock)] = k;
      }
      bcase_.upload(bcase, stream_);
      cflags_host_.assign(std::size_t(ncase), 1 | (second_ ? 2 : 0));
      cflags_.upload(cflags_host_, stream_);
      eflags_.upload(cflags_host_, stream_);
      small_eval_ = Buffer<Small>(std::size_t(ncase));
    }
    build_topology(t);
    if (ncase > 1 && !fused_) throw std::invalid_argument("cases need the fused face kernel (blocks with too many faces)");
    build_stencil(t, st);
    build_geometry(t, geo);
    set_boundary(bc_kind, bc);
    set_source(source);
    // State and work arrays.
    u_ = Buffer<double>(5 * std::size_t(nc_));
    stage_ = Buffer<double>(5 * std::size_t(nc_));
    res_ = Buffer<double>(5 * std::size_t(nc_));
    prim_ = Buffer<C>(kPrimStride * std::size_t(nc_));
    prim2_ = Buffer<C>(kPrimStride * std::size_t(nc_));
    prim_b_ = Buffer<C>(kPrimStride * std::size_t(std::max(nb_, 1)));
    rec_ = Buffer<C>(20 * std::size_t(nc_));
    fflux_ = Buffer<C>(8 * std::size_t(nf_));
    lam_ = Buffer<C>(std::size_t(nc_));
    dt_ = Buffer<double>(std::size_t(nc_));
    small_ = Buffer<Small>(std::size_t(ncase_));
    partial_ = Buffer<unsigned long long>(std::size_t(std::max(grid(nc_), xcd_grid(face_blocks_))));
    nslots_ = Buffer<long long>(std::size_t(ncase_) * nslot_ * kSums * kWords);
    check(hipHostMalloc(reinterpret_cast<void**>(&host_small_), sizeof(Small) * std::size_t(ncase_), hipHostMallocDefault),
          "hipHostMalloc");
    check(hipMemsetAsync(rec_.get(), 0, rec_.size() * sizeof(C), stream_), "hipMemsetAsync");
    if (ncase_ > 1) {
      compact_ = true;
      cblk_ = Buffer<int>(std::size_t(grid(nc_)));
      fblk_ = Buffer<int>(std::size_t(xcd_grid(face_blocks_)));
      build_lists();
    }
    check(hipStreamSynchronize(stream_), "create");
  }

  ~Stepper() override {
    (void)hipStreamSynchronize(stream_);
    drop_graphs();
    if (host_small_ != nullptr) (void)hipHostFree(host_small_);
    if (host_scalars_ != nullptr) (void)hipHostFree(host_scalars_);
    if (ev_in_ != nullptr) (void)hipEventDestroy(ev_in_);
    if (ev_out_ != nullptr) (void)hipEventDestroy(ev_out_);
    if (own_stream_) (void)hipStreamDestroy(stream_);
  }

  void upload(const double* u) override {
    prim_fresh_ = false;
    u_.upload(u, 5 * std::size_t(nc_), stream_);
    check(hipStreamSynchronize(stream_), "upload");
  }

  void download(double* u) override {
    u_.download(u, 5 * std::size_t(nc_), stream_);
    check(hipStreamSynchronize(stream_), "download");
  }

  void residual(const double* u, double* res) override {
    prim_fresh_ = false;  // the primitive buffer now holds u's
    stage_.upload(u, 5 * std::size_t(nc_), stream_);
    launch_prim(stage_.get());
    residual_from_prim();
    res_.download(res, 5 * std::size_t(nc_), stream_);
    check(hipStreamSynchronize(stream_), "residual");
  }

  // Euler::residual_and_jacobian (physics/src/euler.cpp:983-991): R(u), then from the same
  // primitives, gradients and limiter the cells' dW/dq and the faces' dF/dW (jacobian.hpp).
  void residual_and_jacobian(const double* u, double* res, double* dwdu, double* face_jac,
                             double* bnd_jac) override {
    if (ncase() != 1) throw std::invalid_argument("residual_and_jacobian: one case");
    Buffer<double> dj(25 * std::size_t(nc_)), fj(50 * std::size_t(std::max(ni_, 1))), bj(25 * std::size_t(std::max(nb_, 1)));
    residual(u, res);
    const Dev<C> d = dev();
    launch("dwdu", [&] { k_dwdu<C><<<blocks(nc_), kBlock, 0, stream_>>>(d, stage_.get(), dj.get()); });
    launch("face_jacobian", [&] {
      switch (sc_.flux) {
        case MXEULER_ROE:
          launch_face_jacobian<0>(d, fj.get(), bj.get());
          break;
        case MXEULER_HLLC:
          launch_face_jacobian<1>(d, fj.get(), bj.get());
          break;
        default:
          launch_face_jacobian<2>(d, fj.get(), bj.get());
          break;
      }
    });
    dj.download(dwdu, 25 * std::size_t(nc_), stream_);
    if (ni_ > 0) fj.download(face_jac, 50 * std::size_t(ni_), stream_);
    if (nb_ > 0) bj.download(bnd_jac, 25 * std::size_t(nb_), stream_);
    check(hipStreamSynchronize(stream_), "residual_and_jacobian");
  }

  // ---- the steady implicit step (EulerPhysics::steady_step with linear_solver = "gauss-seidel",
  // physics/src/euler_physics.cpp:1056-1250), on the stepper's state u_; implicit.hpp has the kernels.

  // R(u_) at the order set, its norms and validity (as evaluate), the spectral radius, and the
  // approximate Jacobian's parts, all kept on the device for implicit_solve.
  void implicit_evaluate(double* norms, int* invalid) override {
    if (ncase_ != 1) throw std::invalid_argument("implicit step: one case");
    implicit_setup();
    evaluate(norms, invalid);
    const Dev<C> d = dev();
    launch("dwdu", [&] { k_dwdu<C, true><<<blocks(nc_), kBlock, 0, stream_>>>(d, u_.get(), jd_.get()); });
    launch("face_jacobian", [&] {
      switch (sc_.flux) {
        case MXEULER_ROE:
          launch_face_jacobian<0>(d, jf_.get(), jb_.get());
          break;
        case MXEULER_HLLC:
          launch_face_jacobian<1>(d, jf_.get(), jb_.get());
          break;
        default:
          launch_face_jacobian<2>(d, jf_.get(), jb_.get());
          break;
      }
    });
    check(hipStreamSynchronize(stream_), "implicit_evaluate");
    have_jacobian_ = true;
    jac_second_ = second_;
  }

  // Assembles lambda / cfl I + dR/dU from the last implicit_evaluate, solves it for du by GCR (at
  // most kImplDirections directions, to kImplTolerance) preconditioned by kImplSweeps symmetric
  // multicolour block Gauss-Seidel sweeps from zero, and, unless the solve failed, adds the
  // admissible fraction of du to u_. report: cycles, residual drop, alpha (0: no update), failed,
  // singular.
  void implicit_solve(double cfl, double* report) override {
    if (!have_jacobian_) throw std::invalid_argument("implicit_solve: no implicit_evaluate of this state");
    if (!(cfl > 0.0)) throw std::invalid_argument("implicit_solve: cfl must be > 0");
    have_jacobian_ = false;
    for (int q = 0; q < 5; ++q) report[q] = 0.0;
    const Dev<C> d = dev();
    const long long m = 5LL * nc_;
    const int gm = static_cast<int>((m + kBlock - 1) / kBlock);
    dot_next_ = 0;
    launch("impl_clear", [&] { k_impl_clear<<<1, 1024, 0, stream_>>>(words_.get(), kImplDots, bits_.get(), flag_.get()); });
    if (ni_ > 0) {
      launch("impl_off", [&] {
        k_impl_off<C><<<blocks(ni_), kBlock, 0, stream_>>>(d, jd_.get(), jf_.get(), rat_.get(), offd_.get());
      });
    }
    launch("impl_diag", [&] {
      k_impl_diag<C><<<impl_group_blocks(nc_), kBlock, 0, stream_>>>(d, jd_.get(), jb_.get(), offd_.get(), rat_.get(),
                                                                     cfl, rank_.get(), diag_.get());
    });
    launch("impl_invert", [&] {
      k_impl_invert<<<blocks(nc_), kBlock, 0, stream_>>>(diag_.get(), ccells_.get(), nc_, inv_.get(), flag_.get());
    });
    launch("impl_zero", [&] { k_impl_zero<<<gm, kBlock, 0, stream_>>>(du_.get(), m); });
    launch("impl_negate", [&] { k_impl_negate<<<gm, kBlock, 0, stream_>>>(res_.get(), kr_.get(), nc_); });
    int singular = INT32_MAX;
    check(hipMemcpyAsync(&singular, flag_.get(), sizeof(int), hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
    check(hipStreamSynchronize(stream_), "implicit_solve");
    if (singular != INT32_MAX) {
      // A singular diagonal block: the solve failed (EulerPhysics::steady_step's domain_error).
      report[1] = std::numeric_limits<double>::quiet_NaN();
      report[3] = 1.0;
      report[4] = 1.0;
      return;
    }
    int cycles = 0;
    const double drop = gcr(m, gm, cycles);
    report[0] = cycles;
    report[1] = drop;
    if (!(drop <= kImplFailure)) {
      report[3] = 1.0;
      return;
    }
    launch("impl_admissible", [&] {
      k_impl_admissible<<<blocks(nc_), kBlock, 0, stream_>>>(u_.get(), du_.get(), nc_, sc_.gamma, bits_.get());
    });
    unsigned long long abits = 0;
    check(hipMemcpyAsync(&abits, bits_.get(), sizeof(abits), hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
    check(hipStreamSynchronize(stream_), "implicit_solve");
    double alpha;
    std::memcpy(&alpha, &abits, sizeof alpha);
    report[2] = alpha;
    if (alpha < kImplMinStep) return;
    launch("impl_update", [&] { k_impl_update<<<gm, kBlock, 0, stream_>>>(u_.get(), du_.get(), nc_, alpha); });
    prim_fresh_ = false;
    check(hipStreamSynchronize(stream_), "implicit_solve");
  }

  // EulerPhysics::worst_cell: the cell whose density or pressure the full step u + du lowers the
  // most, relatively (ties: the least index), from the last solve's du.
  int implicit_worst_cell() override {
    if (!implicit_ready_) throw std::invalid_argument("implicit_worst_cell: no implicit solve");
    const std::size_t n = static_cast<std::size_t>(nc_);
    std::vector<double> u(5 * n), du(5 * n);
    u_.download(u.data(), u.size(), stream_);
    du_.download(du.data(), du.size(), stream_);
    check(hipStreamSynchronize(stream_), "implicit_worst_cell");
    auto pressure = [&](std::size_t c, double a) {
      double v[5];
      for (std::size_t k = 0; k < 5; ++k) v[k] = u[k * n + c] + a * du[5 * c + k];  // du cell-major
      return (sc_.gamma - 1.0) * (v[4] - 0.5 * (v[1] * v[1] + v[2] * v[2] + v[3] * v[3]) / v[0]);
    };
    std::size_t worst = 0;
    double lowest = std::numeric_limits<double>::infinity();
    for (std::size_t c = 0; c < n; ++c) {
      const double r = std::min((u[c] + du[5 * c]) / u[c], pressure(c, 1.0) / pressure(c, 0.0));
      if (r < lowest) {
        lowest = r;
        worst = c;
      }
    }
    return static_cast<int>(worst);
  }

  // The colour of every cell in the implicit step's sweeps.
  void implicit_colours(int* colours) override {
    implicit_setup();
    std::copy(colour_host_.begin(), colour_host_.end(), colours);
  }

 private:
  static constexpr int kImplSweeps = 4;
  static constexpr int kImplDirections = 10;
  static constexpr double kImplTolerance = 0.1;
  static constexpr double kImplFailure = 0.9;
  // The exact inner products one solve can take: the first, then per direction k its k
  // orthogonalisation coefficients, its norm, its step and the new residual's norm.
  static constexpr int kImplDots = 1 + kImplDirections * (kImplDirections - 1) / 2 + 3 * kImplDirections;

  // Buffers, the rows (each cell's interior faces in ascending face order: BlockMatrix's lists) and
  // the greedy colouring in cell order, made on the first implicit call.
  void implicit_setup() {
    if (implicit_ready_) return;
    const std::size_t n = static_cast<std::size_t>(nc_), ni = static_cast<std::size_t>(ni_);
    std::vector<int> off(n + 1, 0);
    for (std::size_t f = 0; f < ni; ++f) {
      ++off[static_cast<std::size_t>(owner_host_[f]) + 1];
      ++off[static_cast<std::size_t>(neighbour_host_[f]) + 1];
    }
    for (std::size_t c = 0; c < n; ++c) off[c + 1] += off[c];
    std::vector<int2> ent(std::max<std::size_t>(2 * ni, 1));
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (std::size_t f = 0; f < ni; ++f) {
      const int o = owner_host_[f], nb = neighbour_host_[f];
      // Row owner: the upper block (index f), other cell the neighbour; row neighbour: the lower
      // block (index ni + f), other cell the owner.
      ent[static_cast<std::size_t>(fill[static_cast<std::size_t>(o)]++)] = make_int2(static_cast<int>(f), nb);
      ent[static_cast<std::size_t>(fill[static_cast<std::size_t>(nb)]++)] = make_int2(static_cast<int>(ni + f), o);
    }
    colour_host_.assign(n, -1);
    int colours = 0;
    std::vector<char> seen;
    for (std::size_t c = 0; c < n; ++c) {
      seen.assign(static_cast<std::size_t>(colours) + 1, 0);
      for (int k = off[c]; k < off[c + 1]; ++k) {
        const int other = ent[static_cast<std::size_t>(k)].y;
        if (colour_host_[static_cast<std::size_t>(other)] >= 0) seen[static_cast<std::size_t>(colour_host_[static_cast<std::size_t>(other)])] = 1;
      }
      int k = 0;
      while (seen[static_cast<std::size_t>(k)] != 0) ++k;
      colour_host_[c] = k;
      colours = std::max(colours, k + 1);
    }
    std::vector<int> cells(std::max<std::size_t>(n, 1));
    colour_off_.assign(static_cast<std::size_t>(colours) + 1, 0);
    for (std::size_t c = 0; c < n; ++c) ++colour_off_[static_cast<std::size_t>(colour_host_[c]) + 1];
    for (int k = 0; k < colours; ++k) colour_off_[static_cast<std::size_t>(k) + 1] += colour_off_[static_cast<std::size_t>(k)];
    std::vector<int> at(colour_off_.begin(), colour_off_.end() - 1);
    for (std::size_t c = 0; c < n; ++c) cells[static_cast<std::size_t>(at[static_cast<std::size_t>(colour_host_[c])]++)] = static_cast<int>(c);
    std::vector<int> bo;
    for (std::size_t c = 0; c < n; ++c) {
      for (int k = cfo_host_[c]; k < cfo_host_[c + 1]; ++k) {
        if (cell_faces_host_[static_cast<std::size_t>(k)] >= ni_) {
          bo.push_back(static_cast<int>(c));
          break;
        }
      }
    }
    nbo_ = static_cast<int>(bo.size());
    if (nbo_ > 0) bocells_.upload(bo, stream_);
    // The rows in the colour-major order of `cells`: row p is cell cells[p]'s, its entries (the same
    // ascending faces) numbered consecutively, so that a colour's rows and their blocks are one
    // contiguous run; each interior face's two blocks sit at its entries in the two rows.
    std::vector<int> rank(std::max<std::size_t>(n, 1)), poff(n + 1, 0);
    for (std::size_t p = 0; p < n; ++p) {
      const auto c = static_cast<std::size_t>(cells[p]);
      rank[c] = static_cast<int>(p);
      poff[p + 1] = poff[p] + (off[c + 1] - off[c]);
    }
    std::vector<int> other(std::max<std::size_t>(2 * ni, 1));
    std::vector<int2> face_at(std::max<std::size_t>(ni, 1));
    std::vector<int> pfill(poff.begin(), poff.end() - 1);
    for (std::size_t f = 0; f < ni; ++f) {
      const int o = owner_host_[f], nb = neighbour_host_[f];
      const int ko = pfill[static_cast<std::size_t>(rank[static_cast<std::size_t>(o)])]++;
      const int kn = pfill[static_cast<std::size_t>(rank[static_cast<std::size_t>(nb)])]++;
      other[static_cast<std::size_t>(ko)] = nb;
      other[static_cast<std::size_t>(kn)] = o;
      face_at[f] = make_int2(ko, kn);
    }
    roff_.upload(poff, stream_);
    rother_.upload(other, stream_);
    rat_.upload(face_at, stream_);
    rank_.upload(rank, stream_);
    ccells_.upload(cells, stream_);
    plan_relax();
    const std::size_t m = 5 * n;
    jd_ = Buffer<double>(25 * n);
    jf_ = Buffer<double>(50 * std::max<std::size_t>(ni, 1));
    jb_ = Buffer<double>(25 * std::max<std::size_t>(static_cast<std::size_t>(nb_), 1));
    diag_ = Buffer<double>(25 * n);
    inv_ = Buffer<double>(25 * n);
    offd_ = Buffer<double>(50 * std::max<std::size_t>(ni, 1));
    kr_ = Buffer<double>(m);
    du_ = Buffer<double>(m);
    zs_ = Buffer<double>(m * kImplDirections);
    ws_ = Buffer<double>(m * kImplDirections);
    words_ = Buffer<long long>(std::size_t(kImplDots) * kWords);
    bits_ = Buffer<unsigned long long>(2);
    flag_ = Buffer<int>(1);
    scalars_ = Buffer<double>(kImplScalars);
    if (nbo_ > 0) tan_ = Buffer<C>(m);
    check(hipHostMalloc(reinterpret_cast<void**>(&host_scalars_), kImplScalars * sizeof(double), hipHostMallocDefault),
          "hipHostMalloc");
    check(hipStreamSynchronize(stream_), "implicit_setup");
    implicit_ready_ = true;
  }

  // kImplSweeps symmetric sweeps on A z = b from z = 0: the colours in increasing, then decreasing
  // order (BlockMatrix::gauss_seidel), as the launches of relax_plan_.
  void gauss_seidel(const double* b, double* z, long long m, int gm) {
    launch("impl_zero", [&] { k_impl_zero<<<gm, kBlock, 0, stream_>>>(z, m); });
    const ImplRows rows = impl_rows();
    for (const RelaxLaunch& l : relax_plan_) {
      if (l.run.phases == 0) {
        launch("impl_relax", [&] {
          k_impl_relax<<<impl_relax_blocks(l.count), kBlock, 0, stream_>>>(rows, l.first, l.count, offd_.get(),
                                                                           inv_.get(), b, z);
        });
      } else {
        launch("impl_relax_run", [&] {
          k_impl_relax_run<<<1, kRelaxRunThreads, 0, stream_>>>(rows, l.run, offd_.get(), inv_.get(), b, z);
        });
      }
    }
  }

  // The sweeps' colour sequence as launches: a colour of more than kRelaxRunCells cells is one
  // launch over the device; consecutive smaller ones are one workgroup's run of phases.
  struct RelaxLaunch {
    int first = 0, count = 0;
    ImplRun run{};
  };
  std::vector<RelaxLaunch> relax_plan_;

  void plan_relax() {
    const int colours = static_cast<int>(colour_off_.size()) - 1;
    std::vector<int> order;
    for (int s = 0; s < kImplSweeps; ++s) {
      for (int k = 0; k < colours; ++k) order.push_back(k);
      for (int k = colours - 1; k >= 0; --k) order.push_back(k);
    }
    relax_plan_.clear();
    for (const int k : order) {
      const int first = colour_off_[static_cast<std::size_t>(k)];
      const int count = colour_off_[static_cast<std::size_t>(k) + 1] - first;
      if (count == 0) continue;
      const bool small = count <= kRelaxRunCells;
      RelaxLaunch* last = relax_plan_.empty() ? nullptr : &relax_plan_.back();
      if (small && last != nullptr && last->run.phases > 0 && last->run.phases < kRelaxRunMax) {
        last->run.first[last->run.phases] = first;
        last->run.count[last->run.phases] = count;
        ++last->run.phases;
        continue;
      }
      RelaxLaunch l;
      l.first = first;
      l.count = count;
      if (small) {
        l.run.first[0] = first;
        l.run.count[0] = count;
        l.run.phases = 1;
      }
      relax_plan_.push_back(l);
    }
    // A run of one phase is an ordinary launch.
    for (RelaxLaunch& l : relax_plan_) {
      if (l.run.phases == 1) l.run.phases = 0;
    }
  }

  // Gcr::solve (solvers/src/gcr.cpp) on kr_ = -R into du_ = 0: returns the residual drop. The
  // scalars (inner products rounded by exact_value, the Gram-Schmidt and step coefficients) are
  // computed and used on the device, each inner product fused into the vector update before it
  // (exact sums: the same bits in any order); the host reads (w, w) and (r, r) once per direction
  // for the stopping rules, and takes every decision the host loop takes, on the same values.
  double gcr(long long m, int gm, int& cycles) {
    double* sc = scalars_.get();  // [0, D) (w_k, w_k), then alpha, (r, r), (r_0, r_0)
    constexpr int kAlpha = kImplDirections, kRr = kAlpha + 1, kR0 = kRr + 1;
    static_assert(kR0 + 1 == kImplScalars);
    // Grid of the kernels with exact sums: about 8 values per thread (each block resolves its
    // limbs' carries serially before adding them), at most 1024 blocks.
    const int gd = static_cast<int>(std::max<long long>(1, std::min<long long>(1024, m / (8LL * kBlock))));
    unsigned long long mbits = 0;
    launch("impl_maxabs", [&] { k_impl_maxabs<<<gd, kBlock, 0, stream_>>>(kr_.get(), m, bits_.get() + 1); });
    check(hipMemcpyAsync(&mbits, bits_.get() + 1, sizeof mbits, hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
    check(hipStreamSynchronize(stream_), "implicit_solve");
    double scale;
    std::memcpy(&scale, &mbits, sizeof scale);
    if (std::isnan(scale)) scale = std::numeric_limits<double>::infinity();
    if (!std::isfinite(scale)) return std::numeric_limits<double>::quiet_NaN();
    if (scale == 0.0) return 0.0;
    launch("impl_scale", [&] { k_impl_scale<<<gm, kBlock, 0, stream_>>>(kr_.get(), m, scale, 0); });
    {
      long long* w0 = next_words();
      launch("impl_dot", [&] { k_impl_dot<<<gd, kBlock, 0, stream_>>>(kr_.get(), kr_.get(), m, w0); });
      launch("impl_scalar", [&] { k_impl_scalar<<<1, kWave, 0, stream_>>>(w0, nullptr, sc + kR0); });
    }
    double r0 = 0.0, rr = 0.0;
    for (int k = 0; k < kImplDirections; ++k) {
      double* z = zs_.get() + std::size_t(k) * std::size_t(m);
      double* w = ws_.get() + std::size_t(k) * std::size_t(m);
      gauss_seidel(kr_.get(), z, m, gm);
      const ImplRows rows = impl_rows();
      launch("impl_multiply", [&] {
        k_impl_multiply<<<impl_relax_blocks(nc_), kBlock, 0, stream_>>>(rows, nc_, diag_.get(), offd_.get(), z, w);
      });
      // At second order the operator adds the boundary faces' gradient term.
      if (jac_second_ && nbo_ > 0) {
        const Dev<C> d = dev();
        launch("impl_tangents", [&] {
          k_impl_tangents<C><<<blocks(nc_), kBlock, 0, stream_>>>(jd_.get(), z, nc_, tan_.get());
        });
        launch("impl_bgrad", [&] {
          k_impl_bgrad<C><<<(nbo_ + kBlock - 1) / kBlock, kBlock, 0, stream_>>>(d, bocells_.get(), nbo_, tan_.get(),
                                                                              jb_.get(), w);
        });
      }
      // Modified Gram-Schmidt: beta_j = (w, w_j) / (w_j, w_j) on the w left by step j - 1, each
      // step's update fused with the next inner product, the last one's with (w, w) and (r, w);
      // the scalars are rounded by the kernels that use them.
      long long* prev = nullptr;
      long long* ww_words = nullptr;
      long long* rw_words = nullptr;
      if (k == 0) {
        ww_words = next_words();
        rw_words = next_words();
        launch("impl_orth_dots", [&] {
          k_impl_orth_dots<<<gd, kBlock, 0, stream_>>>(z, w, nullptr, nullptr, m, nullptr, nullptr, w, ww_words,
                                                      kr_.get(), rw_words);
        });
      } else {
        prev = next_words();
        launch("impl_orth_dots", [&] {
          k_impl_orth_dots<<<gd, kBlock, 0, stream_>>>(z, w, nullptr, nullptr, m, nullptr, nullptr, ws_.get(), prev,
                                                      nullptr, nullptr);
        });
      }
      for (int j = 0; j < k; ++j) {
        const double* zj = zs_.get() + std::size_t(j) * std::size_t(m);
        const double* wj = ws_.get() + std::size_t(j) * std::size_t(m);
        if (j + 1 < k) {
          long long* nxt = next_words();
          const double* wn = ws_.get() + std::size_t(j + 1) * std::size_t(m);
          launch("impl_orth_dots", [&] {
            k_impl_orth_dots<<<gd, kBlock, 0, stream_>>>(z, w, zj, wj, m, prev, sc + j, wn, nxt, nullptr, nullptr);
          });
          prev = nxt;
        } else {
          ww_words = next_words();
          rw_words = next_words();
          launch("impl_orth_dots", [&] {
            k_impl_orth_dots<<<gd, kBlock, 0, stream_>>>(z, w, zj, wj, m, prev, sc + j, w, ww_words, kr_.get(),
                                                        rw_words);
          });
        }
      }
      long long* rr_words = next_words();
      launch("impl_advance_dot", [&] {
        k_impl_advance_dot<<<gd, kBlock, 0, stream_>>>(du_.get(), kr_.get(), z, w, m, bits_.get() + 1, ww_words,
                                                      rw_words, sc + k, sc + kAlpha, rr_words);
      });
      launch("impl_scalar", [&] { k_impl_scalar<<<1, kWave, 0, stream_>>>(rr_words, nullptr, sc + kRr); });
      check(hipMemcpyAsync(host_scalars_, sc, (kR0 + 1) * sizeof(double), hipMemcpyDeviceToHost, stream_),
            "hipMemcpyAsync");
      check(hipStreamSynchronize(stream_), "implicit_solve");
      if (k == 0) {
        r0 = std::sqrt(host_scalars_[kR0]);
        rr = r0 * r0;
      }
      const double ww = host_scalars_[k];
      cycles = k + 1;
      if (!(ww > 0.0)) {
        if (!std::isfinite(ww)) rr = ww;
        break;
      }
      rr = host_scalars_[kRr];
      if (std::sqrt(rr) <= kImplTolerance * r0) break;
    }
    return std::isfinite(rr) ? std::sqrt(rr) / r0 : std::numeric_limits<double>::quiet_NaN();
  }

  ImplRows impl_rows() const { return ImplRows{roff_.get(), rother_.get(), ccells_.get()}; }

  // A fresh exact-sum word set of the solve (cleared at its start).
  long long* next_words() {
    if (dot_next_ >= kImplDots) throw std::logic_error("implicit_solve: more inner products than planned");
    return words_.get() + std::size_t(dot_next_++) * kWords;
  }

  bool implicit_ready_ = false, have_jacobian_ = false, jac_second_ = false;
  int nbo_ = 0;  // cells owning a boundary face
  Buffer<int> bocells_;
  int dot_next_ = 0;
  std::vector<int> colour_host_, colour_off_;
  Buffer<int> roff_, rother_, rank_, ccells_, flag_;
  Buffer<int2> rat_;  // each interior face's two entries (owner's row, neighbour's row)
  Buffer<double> jd_, jf_, jb_, diag_, inv_, offd_, kr_, du_, zs_, ws_;
  Buffer<long long> words_;
  Buffer<unsigned long long> bits_;
  // The GCR's scalars on the device and the pinned host copy of the ones the host reads.
  static constexpr int kImplScalars = kImplDirections + 3;
  Buffer<double> scalars_;
  double* host_scalars_ = nullptr;
  Buffer<C> tan_;  // the boundary term's primitive tangents

 public:
  template <int FLUX>
  void launch_face_jacobian(const Dev<C>& d, double* fj, double* bj) {
    if (second_) {
      k_face_jacobian<C, FLUX, true><<<blocks(nf_), kBlock, 0, stream_>>>(d, fj, bj);
      if (!planar_ && ni_ > 0 && d.moment != nullptr) k_face_jacobian_warp<C><<<blocks(ni_), kBlock, 0, stream_>>>(d, fj);
    } else {
      k_face_jacobian<C, FLUX, false><<<blocks(nf_), kBlock, 0, stream_>>>(d, fj, bj);
    }
  }

  // Device-resident variants: the arrays are device memory, the work is ordered after what the
  // caller's stream holds and before what it is given next (events, no host wait).
  void upload_device(const double* u, hipStream_t caller) override {
    prim_fresh_ = false;
    join(caller);
    check(hipMemcpyAsync(u_.get(), u, 5 * std::size_t(nc_) * sizeof(double), hipMemcpyDeviceToDevice, stream_),
          "hipMemcpyAsync D2D");
    release(caller);
  }

  void download_device(double* u, hipStream_t caller) override {
    join(caller);
    check(hipMemcpyAsync(u, u_.get(), 5 * std::size_t(nc_) * sizeof(double), hipMemcpyDeviceToDevice, stream_),
          "hipMemcpyAsync D2D");
    release(caller);
  }

  void residual_device(const double* u, double* res, hipStream_t caller) override {
    prim_fresh_ = false;
    join(caller);
    check(hipMemcpyAsync(stage_.get(), u, 5 * std::size_t(nc_) * sizeof(double), hipMemcpyDeviceToDevice, stream_),
          "hipMemcpyAsync D2D");
    launch_prim(stage_.get());
    residual_from_prim();
    check(hipMemcpyAsync(res, res_.get(), 5 * std::size_t(nc_) * sizeof(double), hipMemcpyDeviceToDevice, stream_),
          "hipMemcpyAsync D2D");
    release(caller);
  }

  void step(double cfl, bool has_dt, double dt, bool local, double* residuals, double* time_step,
            int* invalid) override {
    if (use_graph_ && !has_dt) {
      graph_step(cfl, local);
    } else {
      enqueue_step(cfl, has_dt, dt, local);
    }
    finish(residuals, invalid);
    *time_step = has_dt ? dt : key_value(host_small_[0].hkey);
  }

  void evaluate(double* residuals, int* invalid) override {
    clear();
    launch_prim(u_.get());
    residual_from_prim(true, 1.0, false);
    launch("check", [&] { k_check<C><<<cell_grid(), kBlock, 0, stream_>>>(dev(), u_.get()); });
    finish(residuals, invalid);
  }

  void set_second_order(bool on) override {
    drop_graphs();
    second_ = on;
    check(hipMemsetAsync(rec_.get(), 0, rec_.size() * sizeof(C), stream_), "hipMemsetAsync");
    check(hipStreamSynchronize(stream_), "set_second_order");
  }

  void set_boundary_data(const double* bc) override {
    drop_graphs();
    std::vector<C> b(8 * std::size_t(std::max(nb_, 1)), C(0));
    for (int i = 0; i < nb_; ++i) {
      for (int k = 0; k < 5; ++k) b[8 * std::size_t(i) + k] = static_cast<C>(bc[std::size_t(k) * nb_ + i]);
    }
    bc_.upload(b, stream_);
    check(hipStreamSynchronize(stream_), "set_boundary_data");
  }

  void set_source(const double* source) override {
    drop_graphs();
    if (source == nullptr) {
      source_ = Buffer<C>();
      return;
    }
    std::vector<C> s(5 * std::size_t(nc_));
    for (std::size_t i = 0; i < s.size(); ++i) s[i] = static_cast<C>(source[i]);
    source_.upload(s, stream_);
    check(hipStreamSynchronize(stream_), "set_source");
  }

  void spectral_radius(double* lam) override {
    clear();
    launch_spectral(1.0, false);
    std::vector<C> l(static_cast<std::size_t>(nc_));
    lam_.download(l.data(), l.size(), stream_);
    check(hipStreamSynchronize(stream_), "spectral_radius");
    for (int c = 0; c < nc_; ++c) lam[c] = static_cast<double>(l[std::size_t(c)]);
  }

  void arrays(double* prim, double* grad, double* phi, double* prim_b) override {
    // The last residual's primitives: after a step that left the next step's in pcur_, the other
    // buffer's.
    const int last = prim_fresh_ ? pcur_ ^ 1 : pcur_;
    const Buffer<C>& pc = last ? prim2_ : prim_;
    std::vector<C> p(pc.size()), r(rec_.size()), pb(prim_b_.size());
    pc.download(p.data(), p.size(), stream_);
    rec_.download(r.data(), r.size(), stream_);
    prim_b_.download(pb.data(), pb.size(), stream_);
    check(hipStreamSynchronize(stream_), "arrays");
    const std::size_t n = std::size_t(nc_);
    for (std::size_t c = 0; c < n; ++c) {
      for (std::size_t k = 0; k < 5; ++k) {
        if (prim != nullptr) prim[k * n + c] = p[kPrimStride * c + k];
        if (phi != nullptr) phi[k * n + c] = r[20 * c + 15 + k];
      }
      for (std::size_t q = 0; q < 15 && grad != nullptr; ++q) grad[q * n + c] = r[20 * c + q];
    }
    for (std::size_t b = 0; b < std::size_t(nb_) && prim_b != nullptr; ++b) {
      for (std::size_t k = 0; k < 5; ++k) prim_b[k * std::size_t(nb_) + b] = pb[kPrimStride * b + k];
    }
  }

  // The pressure force on the listed boundary faces for the current state: the residual of u
  // first (prim, grad, phi), as EulerPhysics::face_pressure evaluates it before reading them.
  void face_force(const int* faces, long long n, double* force) override {
    for (long long i = 0; i < n; ++i) {
      if (faces[i] < ni_ || faces[i] >= nf_) throw std::invalid_argument("face_force: not a boundary face");
    }
    if (n == 0) return;
    if (n > INT32_MAX / 3) throw std::invalid_argument("face_force: too many faces");
    Buffer<int> list(static_cast<std::size_t>(n));
    Buffer<double> out(3 * static_cast<std::size_t>(n));
    face_force_into(faces, n, force, list.get(), out.get());
  }

  // face_force with its staging in the device's static array: no allocation; at most
  // kStaticForceFaces faces.
  void face_force_noalloc(const int* faces, long long n, double* force) override {
    if (n > kStaticForceFaces) {
      throw std::invalid_argument("face_force_noalloc: " + std::to_string(n) + " faces, more than its static staging holds (" +
                                  std::to_string(kStaticForceFaces) + ")");
    }
    for (long long i = 0; i < n; ++i) {
      if (faces[i] < ni_ || faces[i] >= nf_) throw std::invalid_argument("face_force: not a boundary face");
    }
    if (n == 0) return;
    auto* base = static_cast<unsigned char*>(scratch_);
    face_force_into(faces, n, force, reinterpret_cast<int*>(base), reinterpret_cast<double*>(base + kStaticForceOut));
  }

  // The face list staged in `list` (n int32), the forces in `out` (3 n FP64), then copied back.
  void face_force_into(const int* faces, long long n, double* force, int* list, double* out) {
    check(hipMemcpyAsync(list, faces, static_cast<std::size_t>(n) * sizeof(int), hipMemcpyHostToDevice, stream_),
          "hipMemcpyAsync H2D");
    launch_prim(u_.get());
    residual_from_prim();
    const int ni = static_cast<int>(n);
    launch("face_force", [&] {
      k_face_force<C><<<(ni + kBlock - 1) / kBlock, kBlock, 0, stream_>>>(dev(), list, ni, rb_.get(), out);
    });
    check(hipMemcpyAsync(force, out, 3 * static_cast<std::size_t>(n) * sizeof(double), hipMemcpyDeviceToHost, stream_),
          "hipMemcpyAsync D2H");
    check(hipStreamSynchronize(stream_), "face_force");
  }

  // EulerPhysics::check_conservation on the current state: R^F (the residual without its
  // source), every face's flux from the store loop, and their exact sums.
  void conservation(double* out) override {
    Buffer<C> saved = std::move(source_);
    try {
      launch_prim(u_.get());
      residual_from_prim();
    } catch (...) {
      source_ = std::move(saved);
      throw;
    }
    source_ = std::move(saved);
    const Dev<C> d = dev();
    launch("flux", [&] { launch_flux(d); });
    Buffer<long long> words(std::size_t(kConsSums) * kWords);
    std::vector<long long> zero(words.size(), 0), host(words.size());
    words.upload(zero.data(), zero.size(), stream_);
    launch("conservation", [&] { k_conservation<C><<<1024, kBlock, 0, stream_>>>(d, words.get()); });
    words.download(host.data(), host.size(), stream_);
    check(hipStreamSynchronize(stream_), "conservation");
    for (int i = 0; i < kConsSums; ++i) out[i] = exact_value(host.data() + std::size_t(i) * kWords);
  }

  // Kernel variants the autotuner chooses between; every variant gives the same bits.
  void tune(const std::string& key, int value) override {
    if (key == "face_stage") {
      if (value != 0 && !(std::is_same_v<C, float> && fused_ && faces_lds(true, true) <= 65536)) {
        throw std::invalid_argument("face_stage: needs fp32+acc64 and the fused face kernel");
      }
      drop_graphs();
      stage_faces_ = value != 0;
      return;
    }
    if (key == "face_block") {
      if (value != 256 && value != 512 && value != 1024) throw std::invalid_argument("face_block: 256, 512 or 1024");
      if (!fused_) throw std::invalid_argument("face_block: needs the fused face kernel");
      if (value == block_faces_) return;
      drop_graphs();
      const int old = block_faces_;
      block_faces_ = value;
      rebuild_blocks();
      if (!fused_ || (stage_faces_ && faces_lds(true, true) > 65536)) {
        block_faces_ = old;
        rebuild_blocks();
        throw std::invalid_argument("face_block: blocks of " + std::to_string(value) +
                                    " faces do not fit the fused face kernel here");
      }
      return;
    }
    if (key == "graph") {
      drop_graphs();
      use_graph_ = value != 0;
      return;
    }
    if (key == "grad_occupancy") {
      if (value != 0 && value != 7 && value != 8) throw std::invalid_argument("grad_occupancy: 0, 7 or 8");
      drop_graphs();
      grad_occ_ = value;
      return;
    }
    if (key == "grad_nontemporal") {
      if (value != 0 && value != 1) throw std::invalid_argument("grad_nontemporal: 0 or 1");
      drop_graphs();
      grad_nt_ = value != 0;
      return;
    }
    if (key == "face_occupancy") {
      if (value != 0 && value != 5 && value != 8) throw std::invalid_argument("face_occupancy: 0, 5 or 8");
      if (value != 0 && !std::is_same_v<C, float>) throw std::invalid_argument("face_occupancy: needs fp32+acc64");
      if (value == 5 && planar_) throw std::invalid_argument("face_occupancy: 5 is built for non-planar meshes only");
      drop_graphs();
      face_occ_ = value;
      return;
    }
    throw std::invalid_argument("unknown tuning key '" + key + "'");
  }

  double area_scale() const override { return area_scale_; }
  int ncase() const override { return ncase_; }

  // Times `steps` steps of every variant (face kernel staging where it runs, graph replay) from
  // the same state, keeps the fastest and restores the state. Every variant gives the same bits.
  std::string autotune(int steps, double cfl, bool local) override {
    if (steps < 1) throw std::invalid_argument("autotune: steps >= 1");
    const std::size_t n = 5 * std::size_t(nc_);
    Buffer<double> saved(n);
    check(hipMemcpyAsync(saved.get(), u_.get(), n * sizeof(double), hipMemcpyDeviceToDevice, stream_), "hipMemcpyAsync");
    const std::size_t nk = static_cast<std::size_t>(ncase_);
    std::vector<double> r(kSums * nk), ts(nk);
    std::vector<int> inv(nk);
    auto one = [&] {
      if (ncase_ > 1) {
        batch_step(cfl, local, r.data(), ts.data(), inv.data());
      } else {
        double h = 0.0;
        step(cfl, false, 0.0, local, r.data(), &h, inv.data());
      }
    };
    auto restore = [&] {
      prim_fresh_ = false;
      check(hipMemcpyAsync(u_.get(), saved.get(), n * sizeof(double), hipMemcpyDeviceToDevice, stream_),
            "hipMemcpyAsync");
      check(hipStreamSynchronize(stream_), "autotune");
    };
    std::string report;
    double best = std::numeric_limits<double>::infinity();
    int best_stage = 0, best_graph = 0;
    int best_go = 0, best_fo = 0, best_fb = block_faces_;
    constexpr int kGradOcc[3] = {0, 7, 8}, kFaceOcc[3] = {0, 5, 8};
    std::vector<int> sizes{block_faces_};
    if (std::is_same_v<C, float> && fused_) sizes = {1024, 512};
    for (const int fb : sizes) {
      if (fb != block_faces_) {
        tune("face_stage", 0);
        try {
          tune("face_block", fb);
        } catch (const std::invalid_argument&) {
          continue;  // blocks of this size do not fit here
        }
      }
      const bool can_stage = std::is_same_v<C, float> && fused_ && faces_lds(true, true) <= 65536;
      for (int vi = 0; vi < 36; ++vi) {
        const int st = vi & 1, gr = (vi >> 1) & 1, go = kGradOcc[(vi >> 2) % 3], fo = kFaceOcc[(vi >> 2) / 3];
        if (st && !can_stage) continue;
        if (fo && (st || !std::is_same_v<C, float>)) continue;
        if (fo == 5 && planar_) continue;
        tune("face_stage", st);
        tune("graph", gr);
        tune("grad_occupancy", go);
        tune("face_occupancy", fo);
        restore();
        one();  // warm-up (records the graph)
        restore();
        const auto t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < steps; ++k) one();
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
        const std::string name = "face_block=" + std::to_string(block_faces_) + ",face_stage=" + std::to_string(st) +
                                 ",graph=" + std::to_string(gr) + ",grad_occupancy=" + std::to_string(go) +
                                 ",face_occupancy=" + std::to_string(fo);
        report += name + ":" + std::to_string(ms) + "ms;";
        if (ms < best) {
          best = ms;
          best_fb = block_faces_;
          best_stage = st;
          best_graph = gr;
          best_go = go;
          best_fo = fo;
        }
      }
    }
    tune("face_stage", 0);
    if (best_fb != block_faces_) tune("face_block", best_fb);
    tune("face_stage", best_stage);
    tune("graph", best_graph);
    tune("grad_occupancy", best_go);
    tune("face_occupancy", best_fo);
    restore();
    return "best face_block=" + std::to_string(best_fb) + ",face_stage=" + std::to_string(best_stage) +
           ",graph=" + std::to_string(best_graph) + ",grad_occupancy=" + std::to_string(best_go) +
           ",face_occupancy=" + std::to_string(best_fo) + ";" + report;
  }

  // One SSP-RK3 step of every active case: norms, time step and invalid cell per case.
  void batch_step(double cfl, bool local, double* residuals, double* time_steps, int* invalid) override {
    if (use_graph_) {
      graph_step(cfl, local);
    } else {
      enqueue_step(cfl, false, 0.0, local);
    }
    finish(residuals, invalid);
    for (int k = 0; k < ncase_; ++k) time_steps[k] = key_value(host_small_[k].hkey);
  }

  // Per case: bit 0 active (stepped; a case made inactive keeps its state), bit 1 second order
  // (a case turned to first order has its gradients cleared).
  void set_flags(const int* flags) override {
    if (ncase_ == 1) throw std::invalid_argument("case flags need a batch");
    for (int k = 0; k < ncase_; ++k) {
      if (flags[k] < 0 || flags[k] > 3) throw std::invalid_argument("case flags: bit 0 active, bit 1 second order");
      if ((flags[k] & 2) && !second_) throw std::invalid_argument("case flags: the batch was built at first order");
    }
    for (int k = 0; k < ncase_; ++k) {
      const int old = cflags_host_[std::size_t(k)];
      // A case stepped again: its primitives were left behind while it was frozen.
      if (!(old & 1) && (flags[k] & 1)) prim_fresh_ = false;
      const std::size_t first = std::size_t(case_first_[std::size_t(k)]);
      const std::size_t range = std::size_t(case_first_[std::size_t(k) + 1]) - first;
      if ((old & 1) && !(flags[k] & 1)) {
        // Both state buffers hold the frozen state, whichever the next step reads.
        for (std::size_t q = 0; q < 5; ++q) {
          const std::size_t at = q * std::size_t(nc_) + first;
          check(hipMemcpyAsync(stage_.get() + at, u_.get() + at, range * sizeof(double), hipMemcpyDeviceToDevice,
                               stream_),
                "hipMemcpyAsync");
        }
      }
      if ((old & 2) && !(flags[k] & 2)) {
        check(hipMemsetAsync(rec_.get() + 20 * first, 0, 20 * range * sizeof(C), stream_), "hipMemsetAsync");
      }
      cflags_host_[std::size_t(k)] = flags[k];
    }
    cflags_.upload(cflags_host_.data(), cflags_host_.size(), stream_);
    if (build_lists()) drop_graphs();
    check(hipStreamSynchronize(stream_), "set_flags");
  }

  // Norms of R(u) at second order for the cases marked in `which` (the others skipped), then
  // those cases' gradients cleared: the residual a first-order start reports.
  void batch_evaluate_second(const std::vector<int>& which, double* reported) {
    std::vector<int> ef(static_cast<std::size_t>(ncase_), 0);
    for (int k = 0; k < ncase_; ++k) ef[std::size_t(k)] = which[std::size_t(k)] ? 3 : 0;
    eflags_.upload(ef.data(), ef.size(), stream_);
    eval_mode_ = true;
    try {
      clear();
      launch_prim(u_.get());
      residual_from_prim(true, 1.0, false);
      launch("zero_rec", [&] { k_zero_rec<C><<<cell_grid(), kBlock, 0, stream_>>>(dev()); });
    } catch (...) {
      eval_mode_ = false;
      throw;
    }
    eval_mode_ = false;
    check(hipMemcpyAsync(host_small_, small_eval_.get(), sizeof(Small) * std::size_t(ncase_), hipMemcpyDeviceToHost,
                         stream_),
          "hipMemcpyAsync");
    check(hipStreamSynchronize(stream_), "evaluate");
    for (int cs = 0; cs < ncase_; ++cs) {
      for (int k = 0; k < kSums; ++k) {
        reported[kSums * cs + k] =
            std::sqrt(exact_value(host_small_[cs].words[k]) / double(case_cells_[std::size_t(cs)]));
      }
    }
  }

  // Every case's steady run together: each case's first-order start, local time steps and
  // convergence decided on its own norms (mxe::steady::SteadyCase); a case that ends is frozen
  // and its blocks skipped. history: 5 per iteration per case, case k at history + 5 k
  // max_iterations (or null).
  void batch_steady(const mxsteady_options& o, const double* case_scales, const double* case_references,
                    mxsteady_result* results, double* history) override {
    if (ncase_ == 1) throw std::invalid_argument("batch_steady needs a batch");
    steady::check_options(o);
    if (o.start_order == 1 && !second_) throw std::invalid_argument("a first-order start needs a second-order batch");
    std::vector<steady::SteadyCase> runs;
    runs.reserve(std::size_t(ncase_));
    const std::size_t nk = static_cast<std::size_t>(ncase_);
    std::vector<int> flags(nk);
    for (int k = 0; k < ncase_; ++k) {
      mxsteady_options ok = o;
      ok.scales = case_scales != nullptr ? case_scales + kSums * std::size_t(k) : nullptr;
      ok.reference = case_references != nullptr ? case_references + kSums * std::size_t(k) : nullptr;
      runs.emplace_back(ok);
      flags[std::size_t(k)] = 1 | ((runs.back().second() && second_) ? 2 : 0);
    }
    set_flags(flags.data());
    const std::size_t n5 = std::size_t(kSums) * std::size_t(ncase_);
    std::vector<double> r(n5), rep(n5), ts(nk);
    std::vector<int> inv(nk), which(nk);
    for (;;) {
      bool any = false, any_prov = false;
      for (int k = 0; k < ncase_; ++k) {
        const auto& run = runs[std::size_t(k)];
        which[std::size_t(k)] = !run.done() && !run.second();
        any = any || !run.done();
        any_prov = any_prov || which[std::size_t(k)];
      }
      if (!any) break;
      if (any_prov) batch_evaluate_second(which, rep.data());
      batch_step(o.cfl, true, r.data(), ts.data(), inv.data());
      bool changed = false;
      for (int k = 0; k < ncase_; ++k) {
        auto& run = runs[std::size_t(k)];
        if (run.done()) continue;
        const bool prov = !run.second();
        double* h = history != nullptr ? history + std::size_t(k) * kSums * std::size_t(o.max_iterations) : nullptr;
        run.record(&r[kSums * std::size_t(k)], prov ? &rep[kSums * std::size_t(k)] : nullptr, inv[std::size_t(k)], h,
                   nullptr);
        int f = flags[std::size_t(k)];
        if (run.switched()) f |= 2;
        if (run.done()) f &= ~1;
        changed = changed || f != flags[std::size_t(k)];
        flags[std::size_t(k)] = f;
      }
      if (changed) set_flags(flags.data());
    }
    for (int k = 0; k < ncase_; ++k) results[k] = runs[std::size_t(k)].result();
  }
  double gamma() const override { return sc_.gamma; }

  int profile(int steps, double cfl, bool local, double* ms_step, double* ms_kernel, int* launches, int max_kernels,
              std::string* names) override {
    // Warm the path once, then time `steps` steps with an event around every launch.
    enqueue_step(cfl, false, 0.0, local);
    check(hipStreamSynchronize(stream_), "profile warm-up");
    prof_ = true;
    prof_names_.clear();
    prof_events_.clear();
    hipEvent_t start, stop;
    check(hipEventCreate(&start), "hipEventCreate");
    check(hipEventCreate(&stop), "hipEventCreate");
    check(hipEventRecord(start, stream_), "hipEventRecord");
    for (int s = 0; s < steps; ++s) enqueue_step(cfl, false, 0.0, local);
    check(hipEventRecord(stop, stream_), "hipEventRecord");
    check(hipEventSynchronize(stop), "hipEventSynchronize");
    prof_ = false;
    float total = 0.0f;
    check(hipEventElapsedTime(&total, start, stop), "hipEventElapsedTime");
    *ms_step = total / steps;
    std::vector<std::string> order;
    std::vector<double> sum;
    std::vector<int> count;
    for (std::size_t i = 0; i < prof_events_.size(); ++i) {
      float ms = 0.0f;
      check(hipEventElapsedTime(&ms, prof_events_[i].first, prof_events_[i].second), "hipEventElapsedTime");
      auto it = std::find(order.begin(), order.end(), prof_names_[i]);
      if (it == order.end()) {
        order.push_back(prof_names_[i]);
        sum.push_back(0.0);
        count.push_back(0);
        it = order.end() - 1;
      }
      const auto j = static_cast<std::size_t>(it - order.begin());
      sum[j] += ms;
      count[j] += 1;
      (void)hipEventDestroy(prof_events_[i].first);
      (void)hipEventDestroy(prof_events_[i].second);
    }
    prof_events_.clear();
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    const int n = std::min<int>(max_kernels, static_cast<int>(order.size()));
    std::string joined;
    for (int i = 0; i < n; ++i) {
      ms_kernel[i] = sum[std::size_t(i)] / count[std::size_t(i)];
      launches[i] = count[std::size_t(i)] / steps;
      joined += (i ? ";" : "") + order[std::size_t(i)];
    }
    *names = joined;
    return n;
  }

 private:
  Dev<C> dev() {
    Dev<C> d{};
    d.nc = nc_;
    d.nf = nf_;
    d.ni = ni_;
    d.nb = nb_;
    d.lr = lr_.get();
    d.bowner = bowner_.get();
    d.cfo = cfo_.get();
    d.inc = inc_.get();
    d.soff = soff_.get();
    d.slen = slen_.get();
    d.sidx = sidx_.get();
    d.sw0 = sw0_.get();
    d.sw1 = sw1_.get();
    d.sw2 = sw2_.get();
    d.fgeo = fgeo_.get();
    d.frr = frr_.get();
    d.moment = moment_.get();
    d.vol = vol_.get();
    d.vol64 = vol64_.get();
    d.bkind = bkind_.get();
    d.bc = bc_.get();
    d.source = source_.get();
    d.prim = pcur_ ? prim2_.get() : prim_.get();
    d.prim_b = prim_b_.get();
    d.rec = rec_.get();
    d.fflux = fflux_.get();
    d.res = res_.get();
    d.lam = lam_.get();
    d.dt = dt_.get();
    d.small = small_.get();
    d.ptab = ptab_.get();
    d.bcase = ncase_ > 1 ? bcase_.get() : nullptr;
    d.fcase = ncase_ > 1 ? fcase_.get() : nullptr;
    d.fcbeg = fcbeg_.get();
    d.cflags = ncase_ > 1 ? (eval_mode_ ? eflags_.get() : cflags_.get()) : nullptr;
    d.cblk = compact_ && !eval_mode_ ? cblk_.get() : nullptr;
    d.fblk = compact_ && !eval_mode_ ? fblk_.get() : nullptr;
    if (eval_mode_) d.small = small_eval_.get();
    d.ncase = ncase_;
    d.nslot = nslot_;
    d.cbeg = cbeg_.get();
    d.sbeg = sbeg_.get();
    d.srec = srec_.get();
    d.smom = smom_.get();
    d.islot = islot_.get();
    d.max_slots = max_slots_;
    d.nslots = nslots_.get();
    d.tbeg = tbeg_.get();
    d.tent = tent_.get();
    d.sloc = sloc_.get();
    d.lens = lens_.get();
    d.ioff = ioff_.get();
    d.ir0 = ir0_.get();
    d.ir1 = ir1_.get();
    d.ir2 = ir2_.get();
    return d;
  }

  static void validate(const mxeuler_scheme& sc, const mxeuler_topology& t, const mxeuler_geometry& geo,
                       const mxeuler_stencil& st) {
    if (t.n_cells <= 0 || t.n_faces <= 0 || t.n_interior < 0 || t.n_interior > t.n_faces) {
      throw std::invalid_argument("topology sizes");
    }
    if (!t.owner || (t.n_interior > 0 && !t.neighbour) || !t.cell_face_offsets || !t.cell_faces) {
      throw std::invalid_argument("topology arrays missing");
    }
    if (!geo.face_normal || !geo.face_area || !geo.r_left || !geo.r_right || !geo.cell_volume) {
      throw std::invalid_argument("geometry arrays missing");
    }
    if (sc.planar && !geo.all_planar) {
      throw std::invalid_argument("planar loops on a mesh with a face not proven planar");
    }
    if (!sc.planar && !geo.face_moment) throw std::invalid_argument("face moments missing for the full loops");
    if (sc.second_order && (!st.offsets || !st.entries || !st.weights)) {
      throw std::invalid_argument("gradient stencil missing");
    }
    if (sc.flux < 0 || sc.flux > 2) throw std::invalid_argument("unknown flux");
    if (!(sc.mach_ref > 0.0)) throw std::invalid_argument("AUSM+-up needs a positive cut-off Mach number");
    if (!(sc.reference[0] > 0.0)) throw std::invalid_argument("reference density must be positive");
    const long long nc = t.n_cells, nf = t.n_faces, ni = t.n_interior;
    if (t.cell_face_offsets[0] != 0 || t.cell_face_offsets[nc] > std::numeric_limits<int>::max()) {
      throw std::invalid_argument("cell_face_offsets");
    }
    for (long long c = 0; c < nc; ++c) {
      if (t.cell_face_offsets[c + 1] < t.cell_face_offsets[c]) throw std::invalid_argument("cell_face_offsets");
    }
    for (long long f = 0; f < nf; ++f) {
      if (t.owner[f] < 0 || t.owner[f] >= nc) throw std::invalid_argument("owner out of range");
      if (f < ni && (t.neighbour[f] < 0 || t.neighbour[f] >= nc || t.neighbour[f] == t.owner[f])) {
        throw std::invalid_argument("neighbour out of range");
      }
    }
    for (long long k = 0; k < t.cell_face_offsets[nc]; ++k) {
      if (t.cell_faces[k] < 0 || t.cell_faces[k] >= nf) throw std::invalid_argument("cell_faces out of range");
    }
    for (long long c = 0; c < nc; ++c) {
      for (long long k = t.cell_face_offsets[c]; k < t.cell_face_offsets[c + 1]; ++k) {
        const long long f = t.cell_faces[k];
        if (t.owner[f] != c && (f >= ni || t.neighbour[f] != c)) {
          throw std::invalid_argument("cell_faces lists a face of another cell");
        }
      }
    }
    if (sc.second_order) {
      if (st.offsets[0] != 0) throw std::invalid_argument("stencil offsets");
      for (long long c = 0; c < nc; ++c) {
        if (st.offsets[c + 1] < st.offsets[c]) throw std::invalid_argument("stencil offsets");
      }
      for (long long k = 0; k < st.offsets[nc]; ++k) {
        const long long e = st.entries[k];
        if (e >= nc || (e < 0 && (-e - 1 < ni || -e - 1 >= nf))) throw std::invalid_argument("stencil entry out of range");
      }
    }
  }

  static void validate_cases(const mxeuler_scheme* sc, int ncase, const long long* first, const long long* cells,
                             long long nc) {
    if (ncase < 1 || ncase > 4096) throw std::invalid_argument("cases: 1 to 4096");
    if (first[0] != 0 || first[ncase] != nc) throw std::invalid_argument("cases: ranges must cover the cells");
    for (int k = 0; k < ncase; ++k) {
      if (first[k + 1] <= first[k] || (k > 0 && first[k] % kBlock != 0)) {
        throw std::invalid_argument("cases: each case starts on a 256-cell block");
      }
      if (cells[k] < 1 || cells[k] > first[k + 1] - first[k]) throw std::invalid_argument("cases: cell counts");
      const mxeuler_scheme& a = sc[0];
      const mxeuler_scheme& b = sc[k];
      if (a.flux != b.flux || a.second_order != b.second_order || a.limiter != b.limiter || a.planar != b.planar ||
          a.gamma != b.gamma || a.gas_constant != b.gas_constant || a.venkat_k != b.venkat_k || a.length != b.length) {
        throw std::invalid_argument("cases: flux, order, limiter, planarity and gas must agree");
      }
      if (!(b.mach_ref > 0.0) || !(b.reference[0] > 0.0)) throw std::invalid_argument("cases: reference state");
    }
  }

  // One case's parameters (physics/src/euler.cpp:238-245, in host code built with
  // -ffp-contract=off).
  static Params<C> params(const mxeuler_scheme& sc, long long cell0, long long cell_end) {
    Params<C> pa{};
    const double* r = sc.reference;
    const double p_ref = (sc.gamma - 1.0) * (r[4] - 0.5 * (r[1] * r[1] + r[2] * r[2] + r[3] * r[3]) / r[0]);
    const double k3 = std::pow(sc.venkat_k / sc.length, 3);
    pa.g = static_cast<C>(sc.gamma);
    pa.rgas = static_cast<C>(sc.gas_constant);
    pa.pref = static_cast<C>(p_ref);
    pa.minf = static_cast<C>(sc.mach_ref);
    for (int k = 0; k < 5; ++k) {
      pa.ur[k] = static_cast<C>(r[k]);
      pa.scale[k] = static_cast<C>(k3 * sc.scale[k] * sc.scale[k]);
      pa.ref[k] = r[k];
    }
    pa.p_ref = p_ref;
    pa.gamma = sc.gamma;
    pa.cell0 = cell0;
    pa.cell_end = cell_end;
    return pa;
  }

  void build_topology(const mxeuler_topology& t) {
    std::vector<int2> lr(std::size_t(std::max(ni_, 1)));
    for (int f = 0; f < ni_; ++f) lr[std::size_t(f)] = make_int2(t.owner[f], t.neighbour[f]);
    std::vector<int> bowner(std::size_t(std::max(nb_, 1)), 0);
    for (int b = 0; b < nb_; ++b) bowner[std::size_t(b)] = t.owner[ni_ + b];
    std::vector<int> cfo(std::size_t(nc_) + 1);
    for (int c = 0; c <= nc_; ++c) cfo[std::size_t(c)] = static_cast<int>(t.cell_face_offsets[c]);
    std::vector<int> inc(std::size_t(std::max(cfo.back(), 1)));
    for (int c = 0; c < nc_; ++c) {
      for (int k = cfo[std::size_t(c)]; k < cfo[std::size_t(c) + 1]; ++k) {
        const int f = t.cell_faces[k];
        const bool right = f < ni_ && t.owner[f] != c;
        inc[std::size_t(k)] = static_cast<int>(static_cast<unsigned>(f) | (right ? 0x80000000u : 0u));
      }
    }
    lr_.upload(lr, stream_);
    bowner_.upload(bowner, stream_);
    cfo_.upload(cfo, stream_);
    inc_.upload(inc, stream_);
    // Kept for rebuilding the face blocks at another size (the face_block tuning key).
    cell_faces_host_.assign(t.cell_faces, t.cell_faces + cfo.back());
    owner_host_.assign(t.owner, t.owner + nf_);
    neighbour_host_.assign(t.neighbour, t.neighbour + ni_);
    build_slots(cell_faces_host_.data(), owner_host_.data(), cfo);
    cfo_host_ = std::move(cfo);
    check(hipStreamSynchronize(stream_), "topology");
  }

  // The face blocks of the fused face kernel: consecutive cells, at most kBlock of them and at
  // most max_block_faces() distinct faces. For each
  // block its faces ascending, and for every incidence the LDS slot of its face with the cell's
  // side and whether the face is a boundary face.
  void build_slots(const int* cell_faces, const int* owner, const std::vector<int>& cfo) {
    std::vector<int> cbeg{0}, sbeg{0}, sface;
    sface.reserve(std::size_t(cfo.back()) / 2 + std::size_t(nc_));
    std::vector<unsigned short> islot(std::size_t(std::max(cfo.back(), 1)), 0);
    std::vector<int> mark(std::size_t(nf_), -1), faces;
    int max_slots = 0;
    bool fits = true;
    int c = 0;
    std::vector<int> fcase, fcbeg(std::size_t(ncase_) + 1, 0);
    int cs = 0;
    while (c < nc_) {
      const int b = static_cast<int>(cbeg.size()) - 1;
      faces.clear();
      const int c0 = c;
      while (c0 >= case_first_[std::size_t(cs) + 1]) fcbeg[std::size_t(++cs)] = b;
      const int cend = static_cast<int>(case_first_[std::size_t(cs) + 1]);
      fcase.push_back(cs);
      while (c < cend && c - c0 < kBlock) {
        int fresh = 0;
        for (int k = cfo[std::size_t(c)]; k < cfo[std::size_t(c) + 1]; ++k) fresh += mark[std::size_t(cell_faces[k])] != b;
        if (static_cast<int>(faces.size()) + fresh > max_block_faces() && c > c0) break;
        for (int k = cfo[std::size_t(c)]; k < cfo[std::size_t(c) + 1]; ++k) {
          const int f = cell_faces[k];
          if (mark[std::size_t(f)] != b) {
            mark[std::size_t(f)] = b;
            faces.push_back(f);
          }
        }
        ++c;
      }
      std::sort(faces.begin(), faces.end());
      max_slots = std::max(max_slots, static_cast<int>(faces.size()));
      if (faces.size() > 0x3fff) fits = false;
      for (int cc = c0; cc < c && fits; ++cc) {
        for (int k = cfo[std::size_t(cc)]; k < cfo[std::size_t(cc) + 1]; ++k) {
          const int f = cell_faces[k];
          const auto slot = static_cast<unsigned>(std::lower_bound(faces.begin(), faces.end(), f) - faces.begin());
          const bool right = f < ni_ && owner[f] != cc;
          const bool boundary = f >= ni_;
          islot[std::size_t(k)] = static_cast<unsigned short>(slot | (right ? 0x4000u : 0u) | (boundary ? 0x8000u : 0u));
        }
      }
      sface.insert(sface.end(), faces.begin(), faces.end());
      sbeg.push_back(static_cast<int>(sface.size()));
      cbeg.push_back(c);
    }
    face_blocks_ = static_cast<int>(cbeg.size()) - 1;
    while (cs < ncase_) fcbeg[std::size_t(++cs)] = face_blocks_;
    // Pad to the XCD-aware grid: the extra blocks hold no cell and no face.
    while (static_cast<int>(cbeg.size()) < xcd_grid(face_blocks_) + 1) {
      cbeg.push_back(nc_);
      sbeg.push_back(sbeg.back());
      fcase.push_back(ncase_ - 1);
    }
    if (ncase_ == 1) fcbeg[1] = xcd_grid(face_blocks_);
    fcbeg_.upload(fcbeg, stream_);
    fcbeg_host_ = fcbeg;
    if (ncase_ > 1) fcase_.upload(fcase, stream_);
    max_slots_ = max_slots;
    fused_ = fits && faces_lds(true) <= 65536;
    if (!fused_) return;
    cbeg_.upload(cbeg, stream_);
    sbeg_.upload(sbeg, stream_);
    islot_.upload(islot, stream_);
    sface_host_ = std::move(sface);
  }

  int max_block_faces() const { return block_faces_; }

  // Dynamic LDS of the fused face kernel: the face fluxes (and lambdas), the exact-sum words of
  // the first residual, the staged cells.
  std::size_t faces_lds(bool first, bool stage = false) const {
    const std::size_t comp = first ? 7 : 5;
    const std::size_t fl = (comp * std::size_t(max_slots_) * sizeof(C) + 15) & ~std::size_t(15);
    return fl + (first ? std::size_t(kSumsLds) : 0) +
           (stage ? std::size_t(kPrimStride + 20) * kBlock * sizeof(C) : 0);
  }

  void build_stencil(const mxeuler_topology& t, const mxeuler_stencil& st) {
    const int slices = (nc_ + kWave - 1) / kWave;
    std::vector<long long> soff(std::size_t(slices) + 1, 0);
    std::vector<int> slen(std::size_t(nc_), 0);
    const bool have = st.offsets != nullptr;
    for (int s = 0; s < slices; ++s) {
      int longest = 0;
      for (int l = 0; l < kWave && s * kWave + l < nc_; ++l) {
        const int c = s * kWave + l;
        const int len = have ? static_cast<int>(st.offsets[c + 1] - st.offsets[c]) : 0;
        slen[std::size_t(c)] = len;
        longest = std::max(longest, len);
      }
      soff[std::size_t(s) + 1] = soff[std::size_t(s)] + static_cast<long long>(longest) * kWave;
    }
    const std::size_t total = std::max<std::size_t>(std::size_t(soff.back()), 1);
    std::vector<int> idx(total, 0);
    std::vector<C> w0(total, C(0)), w1(total, C(0)), w2(total, C(0));
    const long long ns = have ? st.offsets[nc_] : 0;
    for (int c = 0; have && c < nc_; ++c) {
      const long long base = soff[std::size_t(c / kWave)] + c % kWave;
      for (long long j = 0; j < st.offsets[c + 1] - st.offsets[c]; ++j) {
        const long long k = st.offsets[c] + j;
        const std::size_t at = std::size_t(base + j * kWave);
        const int e = st.entries[k];
        idx[at] = e >= 0 ? e : -((-e - 1) - t.n_interior) - 1;
        w0[at] = static_cast<C>(st.weights[k]);
        w1[at] = static_cast<C>(st.weights[ns + k]);
        w2[at] = static_cast<C>(st.weights[2 * ns + k]);
      }
    }
    build_tables(soff, slen, idx);
    slen_host_ = slen;
    soff_.upload(soff, stream_);
    slen_.upload(slen, stream_);
    sidx_.upload(idx, stream_);
    sw0_.upload(w0, stream_);
    sw1_.upload(w1, stream_);
    sw2_.upload(w2, stream_);
    check(hipStreamSynchronize(stream_), "stencil");
  }

  // The staged gradient's tables: logical block b's cells read, through their stencils, the
  // cells tent[tbeg[b] .. ) (ascending) then the boundary faces -(b + 1) (ascending); every sliced
  // stencil entry holds its position in that list (16 bits), so the kernel loads the list once
  // into LDS, in order, and gathers from LDS.
  void build_tables(const std::vector<long long>& soff, const std::vector<int>& slen, const std::vector<int>& idx) {
    const int nblk = grid(nc_);
    std::vector<int> tbeg(std::size_t(nblk) + 1, 0), tent;
    tent.reserve(std::size_t(nc_) * 3);
    std::vector<unsigned short> sloc(idx.size(), 0);
    std::vector<int> cells, bnd;
    int max_table = 0;
    for (int b = 0; b < nblk; ++b) {
      const int c0 = std::min(nc_, b * kBlock), c1 = std::min(nc_, (b + 1) * kBlock);
      cells.clear();
      bnd.clear();
      for (int c = c0; c < c1; ++c) {
        const long long base = soff[std::size_t(c / kWave)] + c % kWave;
        for (int j = 0; j < slen[std::size_t(c)]; ++j) {
          const int e = idx[std::size_t(base + static_cast<long long>(j) * kWave)];
          (e >= 0 ? cells : bnd).push_back(e >= 0 ? e : -e - 1);
        }
      }
      std::sort(cells.begin(), cells.end());
      cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
      std::sort(bnd.begin(), bnd.end());
      bnd.erase(std::unique(bnd.begin(), bnd.end()), bnd.end());
      const int n = static_cast<int>(cells.size() + bnd.size());
      max_table = std::max(max_table, n);
      for (int c = c0; c < c1 && n < 65536; ++c) {
        const long long base = soff[std::size_t(c / kWave)] + c % kWave;
        for (int j = 0; j < slen[std::size_t(c)]; ++j) {
          const std::size_t at = std::size_t(base + static_cast<long long>(j) * kWave);
          const int e = idx[at];
          const long long pos = e >= 0 ? std::lower_bound(cells.begin(), cells.end(), e) - cells.begin()
                                       : static_cast<long long>(cells.size()) +
                                             (std::lower_bound(bnd.begin(), bnd.end(), -e - 1) - bnd.begin());
          sloc[at] = static_cast<unsigned short>(pos);
        }
      }
      tent.insert(tent.end(), cells.begin(), cells.end());
      for (const int bi : bnd) tent.push_back(-bi - 1);
      tbeg[std::size_t(b) + 1] = static_cast<int>(tent.size());
    }
    max_table_ = max_table;
    staged_ = max_table < 65536 && std::size_t(max_table) * 5 * sizeof(C) <= 65536;
    if (!staged_) return;
    if (tent.empty()) tent.push_back(0);
    tbeg_.upload(tbeg, stream_);
    tent_.upload(tent, stream_);
    sloc_.upload(sloc, stream_);
  }

  // The limiter's per-incidence offsets r = rel(x_f, x_c) of the cell's own side, sliced like the
  // stencil (incidence j of cell 64 s + l at ioff[s] + 64 j + l), and the per-cell counts
  // lens = stencil length | incidences << 16.
  void build_incidence_slices(const mxeuler_topology& t, const mxeuler_geometry& g, const std::vector<int>& slen) {
    const int slices = (nc_ + kWave - 1) / kWave;
    const std::size_t nf = std::size_t(nf_);
    std::vector<long long> ioff(std::size_t(slices) + 1, 0);
    std::vector<int> lens(std::size_t(nc_), 0);
    for (int s = 0; s < slices; ++s) {
      int longest = 0;
      for (int l = 0; l < kWave && s * kWave + l < nc_; ++l) {
        const int c = s * kWave + l;
        const int n = static_cast<int>(t.cell_face_offsets[c + 1] - t.cell_face_offsets[c]);
        lens[std::size_t(c)] = slen[std::size_t(c)] | (n << 16);
        longest = std::max(longest, n);
      }
      ioff[std::size_t(s) + 1] = ioff[std::size_t(s)] + static_cast<long long>(longest) * kWave;
    }
    const std::size_t total = std::max<std::size_t>(std::size_t(ioff.back()), 1);
    std::vector<C> r0(total, C(0)), r1(total, C(0)), r2(total, C(0));
    for (int c = 0; c < nc_; ++c) {
      const long long base = ioff[std::size_t(c / kWave)] + c % kWave;
      for (long long j = 0; j < t.cell_face_offsets[c + 1] - t.cell_face_offsets[c]; ++j) {
        const std::size_t f = std::size_t(t.cell_faces[t.cell_face_offsets[c] + j]);
        const double* r = t.owner[f] == c ? g.r_left : g.r_right;
        const std::size_t at = std::size_t(base + j * kWave);
        r0[at] = static_cast<C>(r[f]);
        r1[at] = static_cast<C>(r[nf + f]);
        r2[at] = static_cast<C>(r[2 * nf + f]);
      }
    }
    ioff_.upload(ioff, stream_);
    lens_.upload(lens, stream_);
    ir0_.upload(r0, stream_);
    ir1_.upload(r1, stream_);
    ir2_.upload(r2, stream_);
  }

  void build_geometry(const mxeuler_topology& t, const mxeuler_geometry& g) {
    const std::size_t nf = std::size_t(nf_), ni = std::size_t(ni_), nc = std::size_t(nc_);
    std::vector<C> fgeo(8 * nf), frr(4 * std::max<std::size_t>(ni, 1), C(0));
    for (std::size_t f = 0; f < nf; ++f) {
      for (std::size_t u = 0; u < 3; ++u) {
        fgeo[8 * f + u] = static_cast<C>(g.face_normal[u * nf + f]);
        fgeo[8 * f + 4 + u] = static_cast<C>(g.r_left[u * nf + f]);
        if (f < ni) frr[4 * f + u] = static_cast<C>(g.r_right[u * nf + f]);
      }
      fgeo[8 * f + 3] = static_cast<C>(g.face_area[f]);
      fgeo[8 * f + 7] = C(0);
    }
    fgeo_.upload(fgeo, stream_);
    frr_.upload(frr, stream_);
    if (!planar_) {
      std::vector<C> m(9 * nf);
      for (std::size_t i = 0; i < 9 * nf; ++i) m[i] = static_cast<C>(g.face_moment[i]);
      moment_.upload(m, stream_);
    }
    std::vector<C> vol(nc);
    std::vector<double> vol64(g.cell_volume, g.cell_volume + nc);
    for (std::size_t c = 0; c < nc; ++c) vol[c] = static_cast<C>(g.cell_volume[c]);
    vol_.upload(vol, stream_);
    vol64_.upload(vol64, stream_);
    if (fused_) build_slot_records();
    // FP64 x_f - x_c of the boundary faces, for the wall loads.
    std::vector<double> rb(3 * std::max<std::size_t>(nf - ni, 1), 0.0);
    for (std::size_t f = ni; f < nf; ++f) {
      for (std::size_t u = 0; u < 3; ++u) rb[3 * (f - ni) + u] = g.r_left[u * nf + f];
    }
    rb_.upload(rb, stream_);
    // sqrt(sum_i A_i^2 / N), A_i the surface of cell i summed in face order, as
    // EulerPhysics::residual_scales does on one rank (physics/src/euler_physics.cpp:440-466).
    {
      std::vector<double> surface(nc, 0.0);
      for (std::size_t f = 0; f < nf; ++f) {
        surface[std::size_t(t.owner[f])] += g.face_area[f];
        if (f < ni) surface[std::size_t(t.neighbour[f])] += g.face_area[f];
      }
      ExactSum squares, count;
      for (std::size_t c = 0; c < nc; ++c) {
        squares.add(surface[c] * surface[c]);
        count.add(1.0);
      }
      area_scale_ = std::sqrt(squares.value() / count.value());
    }
    build_incidence_slices(t, g, slen_host_);
    slen_host_.clear();
    slen_host_.shrink_to_fit();
    check(hipStreamSynchronize(stream_), "geometry");
  }

  // The fused face kernel's per-slot face records (face_core's 12 values) and moments, in slot
  // order: each block reads its faces' geometry as one contiguous run.
  // The slot records of the faces build_slots listed, built on the device from the per-face
  // arrays (fgeo, frr, lr, bowner, moment: uploaded before).
  void build_slot_records() {
    const std::size_t ns = std::max<std::size_t>(sface_host_.size(), 1);
    Buffer<int> sface(ns);
    sface.upload(sface_host_.data(), sface_host_.size(), stream_);
    srec_ = Buffer<C>(12 * ns);
    smom_ = planar_ ? Buffer<C>() : Buffer<C>(9 * ns);
    const long long n = static_cast<long long>(sface_host_.size());
    if (n > 0) {
      k_slot_records<C><<<static_cast<int>((n + kBlock - 1) / kBlock), kBlock, 0, stream_>>>(
          dev(), sface.get(), n, srec_.get(), planar_ ? nullptr : smom_.get());
      check(hipGetLastError(), "k_slot_records");
    }
    check(hipStreamSynchronize(stream_), "slot records");
    sface_host_.clear();
    sface_host_.shrink_to_fit();
  }

  // The face blocks rebuilt at the current block_faces_ (the topology kept on the host), with their
  // records, the batch's block lists and the per-block partials sized for them.
  void rebuild_blocks() {
    check(hipStreamSynchronize(stream_), "rebuild");
    build_slots(cell_faces_host_.data(), owner_host_.data(), cfo_host_);
    if (!fused_) return;
    build_slot_records();
    const std::size_t need = std::size_t(std::max(grid(nc_), xcd_grid(face_blocks_)));
    if (partial_.size() < need) partial_ = Buffer<unsigned long long>(need);
    if (compact_) {
      fblk_ = Buffer<int>(std::size_t(xcd_grid(face_blocks_)));
      cblk_host_.clear();
      fblk_host_.clear();
      build_lists();
      check(hipStreamSynchronize(stream_), "rebuild");
    }
  }

  void set_boundary(const double* bc_kind, const double* bc) {
    std::vector<int> kind(std::size_t(std::max(nb_, 1)), 0);
    for (int b = 0; b < nb_; ++b) kind[std::size_t(b)] = code_of(bc_kind[b]);
    bkind_.upload(kind, stream_);
    if (nb_ > 0 && bc == nullptr) throw std::invalid_argument("boundary data missing");
    set_boundary_data(bc);
  }

  template <class F>
  void launch(const char* name, F&& f) {
    hipEvent_t a = nullptr, b = nullptr;
    if (prof_) {
      check(hipEventCreate(&a), "hipEventCreate");
      check(hipEventCreate(&b), "hipEventCreate");
      check(hipEventRecord(a, stream_), "hipEventRecord");
    }
    f();
    check(hipGetLastError(), name);
    if (prof_) {
      check(hipEventRecord(b, stream_), "hipEventRecord");
      prof_names_.emplace_back(name);
      prof_events_.emplace_back(a, b);
    }
  }

  int norm_blocks() const { return std::max(1, std::min(blocks(nc_), 512)); }

  // Spectral radius, dt and its least value: per-block minima, then one block reduces them.
  void launch_spectral(double cfl, bool local) {
    const int g = grid(nc_);
    launch("spectral", [&] {
      k_spectral<C><<<g, kBlock, 0, stream_>>>(dev(), cfl, local ? 1 : 0, partial_.get());
    });
    launch("min_dt", [&] { k_min_key<<<1, 1024, 0, stream_>>>(partial_.get(), g, small_.get()); });
  }

  void clear() {
    launch("clear", [&] {
      k_clear<<<ncase_ * (1 + nslot_), kBlock, 0, stream_>>>(eval_mode_ ? small_eval_.get() : small_.get(), ncase_,
                                                            nslots_.get());
    });
  }

  // The active cases' blocks (a batch): cblk_ and fblk_ list them, padded with -1 to the
  // XCD-aware grid, and the cell and face launches cover only them. True when the lists changed.
  bool build_lists() {
    if (!compact_) return false;
    std::vector<int> cl, fl;
    for (int k = 0; k < ncase_; ++k) {
      if (!(cflags_host_[std::size_t(k)] & 1)) continue;
      for (int b = static_cast<int>(case_first_[std::size_t(k)] / kBlock); b < blocks(case_first_[std::size_t(k) + 1]);
           ++b) {
        cl.push_back(b);
      }
      for (int b = fcbeg_host_[std::size_t(k)]; b < fcbeg_host_[std::size_t(k) + 1]; ++b) fl.push_back(b);
    }
    cl.resize(std::size_t(std::max(kXcds, xcd_grid(static_cast<int>(cl.size())))), -1);
    fl.resize(std::size_t(std::max(kXcds, xcd_grid(static_cast<int>(fl.size())))), -1);
    if (cl.size() > cblk_.size() || fl.size() > fblk_.size()) throw std::logic_error("block lists");
    if (cl == cblk_host_ && fl == fblk_host_) return false;
    cblk_host_ = std::move(cl);
    fblk_host_ = std::move(fl);
    cblk_.upload(cblk_host_.data(), cblk_host_.size(), stream_);
    fblk_.upload(fblk_host_.data(), fblk_host_.size(), stream_);
    return true;
  }

  // Launch grids of the cell and face kernels.
  int cell_grid() const { return compact_ && !eval_mode_ ? static_cast<int>(cblk_host_.size()) : grid(nc_); }
  int face_grid() const { return compact_ && !eval_mode_ ? static_cast<int>(fblk_host_.size()) : xcd_grid(face_blocks_); }

  void launch_prim(const double* u) {
    launch("prim", [&] { k_prim<C><<<cell_grid(), kBlock, 0, stream_>>>(dev(), u); });
  }

  // Impl::cell_loops and Impl::faces from the primitives already stored. `first` (the step's
  // R(u)) also leaves the spectral radius, the time steps and the exact norms' words.
  void residual_from_prim(bool first = false, double cfl = 0.0, bool local = false) {
    const Dev<C> d = dev();
    if (nb_ > 0) launch("bstate", [&] { k_bstate<C><<<blocks(nb_), kBlock, 0, stream_>>>(d); });
    if (second_) {
      const std::size_t lds = std::size_t(max_table_) * 5 * sizeof(C);
      if (staged_ && sc_.limiter && grad_nt_) {
        launch("grad_limiter", [&] {
          if (grad_occ_ == 8) {
            k_grad_lds<C, true, 8, true><<<cell_grid(), kBlock, lds, stream_>>>(d);
          } else if (grad_occ_ == 7) {
            k_grad_lds<C, true, 7, true><<<cell_grid(), kBlock, lds, stream_>>>(d);
          } else {
            k_grad_lds<C, true, 0, true><<<cell_grid(), kBlock, lds, stream_>>>(d);
          }
        });
      } else if (staged_ && sc_.limiter && grad_occ_ == 8) {
        launch("grad_limiter", [&] { k_grad_lds<C, true, 8><<<cell_grid(), kBlock, lds, stream_>>>(d); });
      } else if (staged_ && sc_.limiter && grad_occ_ == 7) {
        launch("grad_limiter", [&] { k_grad_lds<C, true, 7><<<cell_grid(), kBlock, lds, stream_>>>(d); });
      } else if (staged_ && sc_.limiter) {
        launch("grad_limiter", [&] { k_grad_lds<C, true><<<cell_grid(), kBlock, lds, stream_>>>(d); });
      } else if (staged_) {
        launch("grad", [&] { k_grad_lds<C, false><<<cell_grid(), kBlock, lds, stream_>>>(d); });
      } else if (sc_.limiter) {
        launch("grad_limiter", [&] { k_grad<C, true><<<cell_grid(), kBlock, 0, stream_>>>(d); });
      } else {
        launch("grad", [&] { k_grad<C, false><<<cell_grid(), kBlock, 0, stream_>>>(d); });
      }
    }
    if (fused_) {
      static constexpr const char* kNames[2][3] = {{"faces", "faces_rk", "faces_rk_check"},
                                                   {"faces_first", "faces_first_rk", "faces_first_rk_check"}};
      launch(kNames[first ? 1 : 0][rk_.mode], [&] { launch_faces(d, first, cfl, local); });
      if (first) {
        launch("step_reduce", [&] {
          const Dev<C> dd = dev();
          k_step_reduce<<<ncase_, 1024, 0, stream_>>>(nslots_.get(), nslot_, partial_.get(), fcbeg_.get(), dd.cflags,
                                                      dd.small);
        });
      }
      return;
    }
    launch("flux", [&] { launch_flux(d); });
    launch("gather", [&] { k_gather<C><<<grid(nc_), kBlock, 0, stream_>>>(d); });
    if (first) {
      launch("norms", [&] { k_norms<<<norm_blocks(), kBlock, 0, stream_>>>(res_.get(), nc_, small_.get()); });
      launch_spectral(cfl, local);
    }
  }

  template <int FLUX, bool PLANAR, bool SECOND, bool FIRST, bool STAGE, int OCC>
  void launch_faces_k(const Dev<C>& d, double cfl, bool local) {
    k_faces<C, FLUX, PLANAR, SECOND, FIRST, STAGE, OCC><<<face_grid(), kBlock, faces_lds(FIRST, STAGE), stream_>>>(
        d, cfl, FIRST && local ? 1 : 0, FIRST ? partial_.get() : nullptr, rk_);
  }

  template <int FLUX, bool PLANAR, bool SECOND, bool FIRST>
  void launch_faces_v(const Dev<C>& d, double cfl, bool local) {
    if constexpr (std::is_same_v<C, float>) {
      if (stage_faces_) return launch_faces_k<FLUX, PLANAR, SECOND, FIRST, true, 0>(d, cfl, local);
      if (face_occ_ == 8) return launch_faces_k<FLUX, PLANAR, SECOND, FIRST, false, 8>(d, cfl, local);
      if constexpr (!PLANAR) {
        if (face_occ_ == 5) return launch_faces_k<FLUX, PLANAR, SECOND, FIRST, false, 5>(d, cfl, local);
      }
    }
    launch_faces_k<FLUX, PLANAR, SECOND, FIRST, false, 0>(d, cfl, local);
  }

  template <int FLUX, bool PLANAR, bool SECOND>
  void launch_faces_t(const Dev<C>& d, bool first, double cfl, bool local) {
    if (first) {
      launch_faces_v<FLUX, PLANAR, SECOND, true>(d, cfl, local);
    } else {
      launch_faces_v<FLUX, PLANAR, SECOND, false>(d, cfl, local);
    }
  }

  template <int FLUX>
  void launch_faces_f(const Dev<C>& d, bool first, double cfl, bool local) {
    if (planar_) {
      if (second_) {
        launch_faces_t<FLUX, true, true>(d, first, cfl, local);
      } else {
        launch_faces_t<FLUX, true, false>(d, first, cfl, local);
      }
    } else {
      if (second_) {
        launch_faces_t<FLUX, false, true>(d, first, cfl, local);
      } else {
        launch_faces_t<FLUX, false, false>(d, first, cfl, local);
      }
    }
  }

  void launch_faces(const Dev<C>& d, bool first, double cfl, bool local) {
    switch (sc_.flux) {
      case MXEULER_ROE:
        launch_faces_f<0>(d, first, cfl, local);
        break;
      case MXEULER_HLLC:
        launch_faces_f<1>(d, first, cfl, local);
        break;
      default:
        launch_faces_f<2>(d, first, cfl, local);
        break;
    }
  }

  template <int FLUX>
  void launch_flux_t(const Dev<C>& d) {
    const int g = grid(nf_);
    if (planar_) {
      if (second_) {
        k_flux<C, FLUX, true, true><<<g, kBlock, 0, stream_>>>(d);
      } else {
        k_flux<C, FLUX, true, false><<<g, kBlock, 0, stream_>>>(d);
      }
    } else {
      if (second_) {
        k_flux<C, FLUX, false, true><<<g, kBlock, 0, stream_>>>(d);
      } else {
        k_flux<C, FLUX, false, false><<<g, kBlock, 0, stream_>>>(d);
      }
    }
  }

  void launch_flux(const Dev<C>& d) {
    switch (sc_.flux) {
      case MXEULER_ROE:
        launch_flux_t<0>(d);
        break;
      case MXEULER_HLLC:
        launch_flux_t<1>(d);
        break;
      default:
        launch_flux_t<2>(d);
        break;
    }
  }

  // CudaStepper::step (physics/src/device_stepper.cuh:313-345) without a host round trip.
  void enqueue_step(double cfl, bool has_dt, double dt, bool local) {
    const bool fresh = prim_fresh_;
    prim_fresh_ = false;
    clear();
    if (!fresh) {
      pcur_ = 0;
      launch_prim(u_.get());
    }
    const bool use_local = !has_dt && local;
    const int use_hkey = has_dt ? 0 : 1;
    // Shu and Osher (1988), eq. 2.18.
    static constexpr double w[3][2] = {{0.0, 1.0}, {0.75, 0.25}, {1.0 / 3.0, 2.0 / 3.0}};
    for (int s = 0; s < 3; ++s) {
      const double* in = s == 0 ? u_.get() : stage_.get();
      // The fused face kernel updates the stage as it gathers each cell's residual, unless the
      // step is the least over all cells and not yet known (the first stage of a global step).
      const bool fuse = fused_ && (s > 0 || has_dt || use_local);
      if (fuse) {
        // The last stage writes the new state's primitives too: the next step starts from them.
        rk_ = RkStage<C>{u_.get(), in, stage_.get(), pcur_ ? prim_.get() : prim2_.get(), w[s][0], w[s][1], dt,
                         use_hkey, use_local ? 1 : 0, s < 2 ? 1 : 2};
      }
      try {
        if (s == 0) {
          residual_from_prim(true, has_dt ? 1.0 : cfl, use_local);
        } else {
          residual_from_prim();
        }
      } catch (...) {
        rk_ = RkStage<C>{};
        throw;
      }
      rk_ = RkStage<C>{};
      if (fuse) {
        pcur_ ^= 1;
        if (s == 2) prim_fresh_ = true;
        continue;
      }
      const Dev<C> d = dev();
      if (s < 2) {
        launch("rk_prim", [&] {
          k_rk<C, true, false><<<cell_grid(), kBlock, 0, stream_>>>(d, w[s][0], w[s][1], u_.get(), in,
                                                                     stage_.get(), use_hkey, dt, use_local ? 1 : 0);
        });
      } else {
        launch("rk_check", [&] {
          k_rk<C, false, true><<<cell_grid(), kBlock, 0, stream_>>>(d, w[s][0], w[s][1], u_.get(), in,
                                                                     stage_.get(), use_hkey, dt, use_local ? 1 : 0);
        });
      }
    }
    std::swap(u_, stage_);
  }

  // A step's launches recorded once as a HIP graph and replayed. A graph is keyed by where the step
  // starts: the state buffer, the primitive buffer it reads and whether the last step left them
  // (the first step after an upload computes them), so steady stepping alternates between two
  // graphs. Recorded again when the CFL number, the time stepping or a kernel variant changes, or
  // a buffer the kernels read is replaced.
  void graph_step(double cfl, bool local) {
    if (!graphs_valid_ || graph_cfl_ != cfl || graph_local_ != local) {
      drop_graphs();
      graphs_valid_ = true;
      graph_cfl_ = cfl;
      graph_local_ = local;
    }
    const GraphEntry* g = nullptr;
    for (const auto& e : graphs_) {
      if (e.u == u_.get() && e.pcur == pcur_ && e.fresh == prim_fresh_) g = &e;
    }
    if (g == nullptr) {
      if (graphs_.size() >= 4) {
        drop_graphs();
        graphs_valid_ = true;
      }
      GraphEntry e{u_.get(), pcur_, prim_fresh_, nullptr, 0, false};
      // Recorded from the current state; the host's view of the buffers is put back afterwards,
      // since the recording ran nothing.
      check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      hipGraph_t gr = nullptr;
      try {
        enqueue_step(cfl, false, 0.0, local);
      } catch (...) {
        (void)hipStreamEndCapture(stream_, &gr);
        if (gr != nullptr) (void)hipGraphDestroy(gr);
        prim_fresh_ = false;
        throw;
      }
      std::swap(u_, stage_);
      e.end_pcur = pcur_;
      e.end_fresh = prim_fresh_;
      pcur_ = e.pcur;
      prim_fresh_ = e.fresh;
      check(hipStreamEndCapture(stream_, &gr), "hipStreamEndCapture");
      const hipError_t err = hipGraphInstantiate(&e.exec, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      check(err, "hipGraphInstantiate");
      graphs_.push_back(e);
      g = &graphs_.back();
    }
    check(hipGraphLaunch(g->exec, stream_), "hipGraphLaunch");
    pcur_ = g->end_pcur;
    prim_fresh_ = g->end_fresh;
    std::swap(u_, stage_);
  }

  void drop_graphs() {
    for (auto& e : graphs_) {
      if (e.exec != nullptr) (void)hipGraphExecDestroy(e.exec);
    }
    graphs_.clear();
    graphs_valid_ = false;
  }

  // The step's single copy back: norms, time step and validity.
  void finish(double* residuals, int* invalid) {
    check(hipMemcpyAsync(host_small_, small_.get(), sizeof(Small) * std::size_t(ncase_), hipMemcpyDeviceToHost, stream_),
          "hipMemcpyAsync");
    check(hipStreamSynchronize(stream_), "step");
    for (int cs = 0; cs < ncase_; ++cs) {
      const Small& sm = host_small_[cs];
      if (ncase_ > 1 && !(cflags_host_[std::size_t(cs)] & 1)) {
        // A frozen case: nothing was added to its cleared words (its norms are 0, no invalid cell).
        for (int k = 0; k < kSums; ++k) residuals[kSums * cs + k] = 0.0;
        invalid[cs] = -1;
        continue;
      }
      for (int k = 0; k < kSums; ++k) {
        residuals[kSums * cs + k] = std::sqrt(exact_value(sm.words[k]) / double(case_cells_[std::size_t(cs)]));
      }
      const int a = sm.invalid_state, b = sm.invalid_residual;
      invalid[cs] = a != INT32_MAX ? a : (b != INT32_MAX ? b : -1);
    }
  }

  // The stepper's stream waits for the caller's work so far.
  void join(hipStream_t caller) {
    if (caller == stream_) return;
    if (ev_in_ == nullptr) check(hipEventCreateWithFlags(&ev_in_, hipEventDisableTiming), "hipEventCreate");
    check(hipEventRecord(ev_in_, caller), "hipEventRecord");
    check(hipStreamWaitEvent(stream_, ev_in_, 0), "hipStreamWaitEvent");
  }

  // The caller's stream waits for the stepper's work so far.
  void release(hipStream_t caller) {
    if (caller == stream_) return;
    if (ev_out_ == nullptr) check(hipEventCreateWithFlags(&ev_out_, hipEventDisableTiming), "hipEventCreate");
    check(hipEventRecord(ev_out_, stream_), "hipEventRecord");
    check(hipStreamWaitEvent(caller, ev_out_, 0), "hipStreamWaitEvent");
  }

  mxeuler_scheme sc_;
  Buffer<double> rb_;
  hipEvent_t ev_in_ = nullptr, ev_out_ = nullptr;
  double area_scale_ = 0.0;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  void* scratch_ = nullptr;  // this device's `scratch`
  int nc_ = 0, nf_ = 0, ni_ = 0, nb_ = 0;
  bool second_ = true, planar_ = true;
  int ncase_ = 1;
  int nslot_ = kNormSlots;
  std::vector<long long> case_first_;  // ncase + 1: each case's first cell (a multiple of kBlock)
  std::vector<long long> case_cells_;  // each case's cells (the rest of its range is filler)
  std::vector<Params<C>> ptab_host_;
  std::vector<int> cflags_host_;
  Buffer<int> eflags_;       // the flags of a second-order evaluation of some cases
  Buffer<Small> small_eval_;  // its norms
  bool eval_mode_ = false;
  bool use_graph_ = false, graphs_valid_ = false, graph_local_ = false;
  double graph_cfl_ = 0.0;
  struct GraphEntry {
    const double* u;  // the state buffer the step starts from
    int pcur;         // the primitive buffer it reads
    bool fresh;       // whether those primitives were left by the step before
    hipGraphExec_t exec;
    int end_pcur;
    bool end_fresh;
  };
  std::vector<GraphEntry> graphs_;
  // The primitive buffer pcur_ holds the primitives of the state u_ (the last step's last stage
  // wrote them), so the next step needs no k_prim.
  bool prim_fresh_ = false;
  Buffer<Params<C>> ptab_;
  Buffer<int> bcase_, fcase_, fcbeg_, cflags_;
  std::vector<int> fcbeg_host_;
  std::vector<int> cfo_host_, cell_faces_host_, owner_host_, neighbour_host_;
  // Faces per face block: four rounds of the workgroup in FP32 (a full 256 cells of hexahedra per
  // block), two in FP64 (whose fluxes take twice the LDS); the face_block tuning key changes it.
  int block_faces_ = std::is_same_v<C, float> ? 2 * kMaxSlots : kMaxSlots;
  bool compact_ = false;
  Buffer<int> cblk_, fblk_;
  std::vector<int> cblk_host_, fblk_host_;
  Buffer<int2> lr_;
  Buffer<int> bowner_, cfo_, inc_, slen_, sidx_, bkind_;
  Buffer<long long> soff_;
  Buffer<C> sw0_, sw1_, sw2_, fgeo_, frr_, moment_, vol_, bc_, source_;
  Buffer<double> vol64_;
  Buffer<double> u_, stage_, res_, dt_;
  Buffer<C> prim_, prim_b_, rec_, fflux_, lam_;
  // A second primitive buffer: a stage updated inside the face kernel writes its primitives there
  // while other blocks still read the current ones; pcur_ names the current one.
  Buffer<C> prim2_;
  int pcur_ = 0;
  RkStage<C> rk_{};  // the stage the next face launch updates (mode 0: none)
  Buffer<Small> small_;
  Buffer<unsigned long long> partial_;
  Buffer<int> cbeg_, sbeg_;
  std::vector<int> sface_host_;  // the faces of every slot, until the slot records are built
  bool stage_faces_ = false;     // fp32: the face kernel stages its block's cells in LDS
  // The gradient / face kernel built for at least this many waves per SIMD (0: the compiler's
  // choice); gradient 7 or 8, face 5 (non-planar meshes) or 8.
  int grad_occ_ = 0, face_occ_ = 0;
  // The gradient kernel's stencil stream loaded non-temporally (data each launch reads once; the
  // same values; 1.5 % off the kernel on an MI355X).
  bool grad_nt_ = true;
  Buffer<C> srec_, smom_;
  int face_blocks_ = 0;
  Buffer<int> tbeg_, tent_, lens_;
  Buffer<unsigned short> sloc_;
  Buffer<long long> ioff_;
  Buffer<C> ir0_, ir1_, ir2_;
  std::vector<int> slen_host_;
  int max_table_ = 0;
  bool staged_ = false;
  Buffer<unsigned short> islot_;
  Buffer<long long> nslots_;
  int max_slots_ = 0;
  bool fused_ = false;
  Small* host_small_ = nullptr;
  bool prof_ = false;
  std::vector<std::string> prof_names_;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_events_;
};

}  // namespace mxe

struct mxeuler_stepper {
  std::unique_ptr<mxe::StepperBase> impl;
  int nc = 0, nb = 0;
};

namespace {

// 1: the input was refused (std::invalid_argument, as the reference's stepper throws for it);
// 2: a device or runtime failure (std::runtime_error and the rest).
int fail(const std::exception& e, char* err, size_t err_len) {
  if (err != nullptr && err_len > 0) {
    std::strncpy(err, e.what(), err_len - 1);
    err[err_len - 1] = '\0';
  }
  return dynamic_cast<const std::invalid_argument*>(&e) != nullptr ? 1 : 2;
}

// A callback's status for the steady drivers: 0, or 1 when the call threw.
template <class F>
int device_call(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

template <class F>
int guarded(char* err, size_t err_len, F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, err_len);
  }
}

}  // namespace

extern "C" {

int mxeuler_create(int32_t policy, const mxeuler_scheme* scheme, const mxeuler_topology* topology,
                   const mxeuler_geometry* geometry, const mxeuler_stencil* stencil, const double* bc_kind,
                   const double* bc, const double* source, void* stream, mxeuler_stepper** out, char* err,
                   size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!scheme || !topology || !geometry || !stencil || !out) throw std::invalid_argument("null argument");
    auto h = std::make_unique<mxeuler_stepper>();
    const auto s = static_cast<hipStream_t>(stream);
    if (policy == MXEULER_FP64) {
      h->impl = std::make_unique<mxe::Stepper<double>>(*scheme, *topology, *geometry, *stencil, bc_kind, bc, source, s);
    } else if (policy == MXEULER_FP32_ACC64) {
      h->impl = std::make_unique<mxe::Stepper<float>>(*scheme, *topology, *geometry, *stencil, bc_kind, bc, source, s);
    } else {
      throw std::invalid_argument("unknown precision policy");
    }
    h->nc = topology->n_cells;
    h->nb = topology->n_faces - topology->n_interior;
    *out = h.release();
  });
}

void mxeuler_destroy(mxeuler_stepper* s) { delete s; }

int mxeuler_upload(mxeuler_stepper* s, const double* u, char* err, size_t err_len) {
  return guarded(err, err_len, [&] { s->impl->upload(u); });
}

int mxeuler_download(mxeuler_stepper* s, double* u, char* err, size_t err_len) {
  return guarded(err, err_len, [&] { s->impl->download(u); });
}

int mxeuler_residual(mxeuler_stepper* s, const double* u, double* res, char* err, size_t err_len) {
  return guarded(err, err_len, [&] { s->impl->residual(u, res); });
}

int mxeuler_residual_and_jacobian(mxeuler_stepper* s, const double* u, double* res, double* dwdu, double* face_jac,
                                  double* bnd_jac, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (u == nullptr || res == nullptr || dwdu == nullptr || face_jac == nullptr || bnd_jac == nullptr) {
      throw std::invalid_argument("residual_and_jacobian: null array");
    }
    s->impl->residual_and_jacobian(u, res, dwdu, face_jac, bnd_jac);
  });
}

int mxeuler_upload_device(mxeuler_stepper* s, const double* u, void* stream, char* err, size_t err_len) {
  return guarded(err, err_len, [&] { s->impl->upload_device(u, static_cast<hipStream_t>(stream)); });
}

int mxeuler_download_device(mxeuler_stepper* s, double* u, void* stream, char* err, size_t err_len) {
  return guarded(err, err_len, [&] { s->impl->download_device(u, static_cast<hipStream_t>(stream)); });
}

int mxeuler_residual_device(mxeuler_stepper* s, const double* u, double* res, void* stream, char* err,
                            size_t err_len) {
  return guarded(err, err_len, [&] { s->impl->residual_device(u, res, static_cast<hipStream_t>(stream)); });
}

int mxeuler_step(mxeuler_stepper* s, double cfl, int32_t has_dt, double dt, int32_t local, double residuals[5],
                 double* time_step, int32_t* invalid_cell, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (s->impl->ncase() != 1) throw std::invalid_argument("a batch steps with mxeuler_batch_step");
    if (has_dt && !(dt >= 0.0)) throw std::invalid_argument("time step must be non-negative");
    int inv = -1;
    s->impl->step(cfl, has_dt != 0, dt, local != 0, residuals, time_step, &inv);
    *invalid_cell = inv;
  });
}

int mxeuler_evaluate(mxeuler_stepper* s, double residuals[5], int32_t* invalid_cell, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (s->impl->ncase() != 1) throw std::invalid_argument("evaluate takes one case");
    int inv = -1;
    s->impl->evaluate(residuals, &inv);
    *invalid_cell = inv;
  });
}

int mxeuler_set_second_order(mxeuler_stepper* s, int32_t on, char* err, size_t err_len) {
  return guarded(err, err_len, [&] { s->impl->set_second_order(on != 0); });
}

int mxeuler_set_boundary_data(mxeuler_stepper* s, const double* bc, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (s->nb > 0 && bc == nullptr) throw std::invalid_argument("boundary data missing");
    s->impl->set_boundary_data(bc);
  });
}

int mxeuler_set_source(mxeuler_stepper* s, const double* source, char* err, size_t err_len) {
  return guarded(err, err_len, [&] { s->impl->set_source(source); });
}

int mxeuler_spectral_radius(mxeuler_stepper* s, double* lambda, char* err, size_t err_len) {
  return guarded(err, err_len, [&] { s->impl->spectral_radius(lambda); });
}

int mxeuler_arrays(mxeuler_stepper* s, double* prim, double* grad, double* phi, double* prim_b, char* err,
                   size_t err_len) {
  return guarded(err, err_len, [&] { s->impl->arrays(prim, grad, phi, prim_b); });
}

int mxeuler_profile(mxeuler_stepper* s, int32_t steps, double cfl, int32_t local, double* ms_per_step,
                    double* ms_per_kernel, int32_t* launches_per_step, int32_t max_kernels, int32_t* n_kernels,
                    char* names, size_t names_len, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (steps <= 0) throw std::invalid_argument("steps must be positive");
    std::string joined;
    std::vector<int> launches(static_cast<std::size_t>(std::max(max_kernels, 0)));
    const int n = s->impl->profile(steps, cfl, local != 0, ms_per_step, ms_per_kernel, launches.data(), max_kernels,
                                   &joined);
    for (int i = 0; i < n; ++i) launches_per_step[i] = launches[static_cast<std::size_t>(i)];
    *n_kernels = n;
    if (names != nullptr && names_len > 0) {
      std::strncpy(names, joined.c_str(), names_len - 1);
      names[names_len - 1] = '\0';
    }
  });
}

int mxeuler_probe_math(int32_t policy, const double* x, const double* y, int64_t n, double* out_sqrt,
                       double* out_div, double* out_exp, double* out_log, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    using mxe::Buffer;
    Buffer<double> dx(static_cast<std::size_t>(n)), dy(static_cast<std::size_t>(n)), s(static_cast<std::size_t>(n)),
        q(static_cast<std::size_t>(n)), e(static_cast<std::size_t>(n)), l(static_cast<std::size_t>(n));
    dx.upload(x, static_cast<std::size_t>(n), nullptr);
    dy.upload(y, static_cast<std::size_t>(n), nullptr);
    const int g = static_cast<int>((n + 255) / 256);
    if (policy == 0) {
      mxe::k_probe<double><<<g, 256>>>(dx.get(), dy.get(), n, s.get(), q.get(), e.get(), l.get());
    } else {
      mxe::k_probe<float><<<g, 256>>>(dx.get(), dy.get(), n, s.get(), q.get(), e.get(), l.get());
    }
    mxe::check(hipGetLastError(), "k_probe");
    s.download(out_sqrt, static_cast<std::size_t>(n), nullptr);
    q.download(out_div, static_cast<std::size_t>(n), nullptr);
    e.download(out_exp, static_cast<std::size_t>(n), nullptr);
    l.download(out_log, static_cast<std::size_t>(n), nullptr);
    mxe::check(hipDeviceSynchronize(), "probe");
  });
}

int mxeuler_stream(int64_t n, int32_t repeats, double* gbps, double* ms, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    using mxe::Buffer;
    if (n < 2 || n % 2 != 0 || repeats < 1) throw std::invalid_argument("stream: n even and >= 2, repeats >= 1");
    const long long m = n / 2;  // 16 B elements
    Buffer<double2> a(static_cast<std::size_t>(m)), b(static_cast<std::size_t>(m)), c(static_cast<std::size_t>(m));
    const int g = static_cast<int>((m + 255) / 256);
    hipStream_t s;
    mxe::check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    mxe::k_stream_init<<<g, 256, 0, s>>>(a.get(), b.get(), c.get(), m);
    hipEvent_t e0, e1;
    mxe::check(hipEventCreate(&e0), "hipEventCreate");
    mxe::check(hipEventCreate(&e1), "hipEventCreate");
    const double bytes[4] = {32.0 * m, 32.0 * m, 48.0 * m, 48.0 * m};  // arrays read + written
    for (int k = 0; k < 4; ++k) {
      gbps[k] = 0.0;
      ms[k] = 1e30;
    }
    for (int r = 0; r < repeats; ++r) {
      for (int k = 0; k < 4; ++k) {
        mxe::check(hipEventRecord(e0, s), "hipEventRecord");
        if (k == 0) mxe::k_stream_copy<<<g, 256, 0, s>>>(a.get(), c.get(), m);
        if (k == 1) mxe::k_stream_scale<<<g, 256, 0, s>>>(c.get(), b.get(), 3.0, m);
        if (k == 2) mxe::k_stream_add<<<g, 256, 0, s>>>(a.get(), b.get(), c.get(), m);
        if (k == 3) mxe::k_stream_triad<<<g, 256, 0, s>>>(b.get(), c.get(), a.get(), 3.0, m);
        mxe::check(hipGetLastError(), "stream kernel");
        mxe::check(hipEventRecord(e1, s), "hipEventRecord");
        mxe::check(hipEventSynchronize(e1), "hipEventSynchronize");
        float t = 0.0f;
        mxe::check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
        if (t < ms[k]) {
          ms[k] = t;
          gbps[k] = bytes[k] / (t * 1e-3) / 1e9;
        }
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
  });
}

int mxeuler_face_force(mxeuler_stepper* s, const int32_t* faces, int64_t n, double* force, char* err,
                       size_t err_len) {
  return guarded(err, err_len, [&] {
    if (n < 0 || (n > 0 && (faces == nullptr || force == nullptr))) throw std::invalid_argument("null argument");
    s->impl->face_force(faces, n, force);
  });
}

int mxeuler_face_force_noalloc(mxeuler_stepper* s, const int32_t* faces, int64_t n, double* force, char* err,
                               size_t err_len) {
  return guarded(err, err_len, [&] {
    if (n < 0 || (n > 0 && (faces == nullptr || force == nullptr))) throw std::invalid_argument("null argument");
    s->impl->face_force_noalloc(faces, n, force);
  });
}

int mxeuler_loads(mxeuler_stepper* s, const int32_t* faces, int64_t n, const double* area_vector,
                  const double* face_centroid, const double center[3], double p_inf, double force[3],
                  double moment[3], char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (n < 0 || !center || !force || !moment || (n > 0 && (!faces || !area_vector || !face_centroid))) {
      throw std::invalid_argument("null argument");
    }
    const auto m = static_cast<std::size_t>(n);
    std::vector<double> ff(3 * m);
    s->impl->face_force(faces, n, ff.data());
    // EulerPhysics::loads and wall_force (physics/src/euler_physics.cpp:527-562, 930-948): the
    // pressure force minus p_inf S, and its moment about `center`, as exact sums.
    mxe::ExactSum sums[6];
    for (std::size_t i = 0; i < m; ++i) {
      const double f[3] = {ff[i] - p_inf * area_vector[i], ff[m + i] - p_inf * area_vector[m + i],
                           ff[2 * m + i] - p_inf * area_vector[2 * m + i]};
      const double arm[3] = {face_centroid[i] - center[0], face_centroid[m + i] - center[1],
                             face_centroid[2 * m + i] - center[2]};
      for (int k = 0; k < 3; ++k) sums[k].add(f[k]);
      sums[3].add(arm[1] * f[2] - arm[2] * f[1]);
      sums[4].add(arm[2] * f[0] - arm[0] * f[2]);
      sums[5].add(arm[0] * f[1] - arm[1] * f[0]);
    }
    for (int k = 0; k < 3; ++k) {
      force[k] = sums[k].value();
      moment[k] = sums[3 + k].value();
    }
  });
}

int mxeuler_tune(mxeuler_stepper* s, const char* key, int32_t value, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!s || !key) throw std::invalid_argument("null argument");
    s->impl->tune(key, value);
  });
}

int mxeuler_conservation(mxeuler_stepper* s, double sums[20], char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!sums) throw std::invalid_argument("null argument");
    s->impl->conservation(sums);
  });
}

int mxeuler_residual_scales(mxeuler_stepper* s, double density, double speed, double pressure, double scales[5],
                            char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!scales) throw std::invalid_argument("null argument");
    if (!(density > 0.0) || !(pressure > 0.0) || !(speed >= 0.0)) {
      throw std::invalid_argument("residual_scales: density and pressure must be positive, speed non-negative");
    }
    // EulerPhysics::residual_scales (physics/src/euler_physics.cpp:467-476).
    const double g = s->impl->gamma();
    const double area = s->impl->area_scale();
    const double c = std::sqrt(g * pressure / density);
    const double q = speed + c;
    const double h = c * c / (g - 1.0) + 0.5 * speed * speed;
    const double mass = density * q * area;
    const double momentum = mass * q;
    scales[0] = mass;
    scales[1] = scales[2] = scales[3] = momentum;
    scales[4] = mass * h;
  });
}

int mxeuler_create_batch(int32_t policy, int32_t n_cases, const mxeuler_scheme* schemes, const int64_t* case_first,
                         const int64_t* case_cells, const mxeuler_topology* topology, const mxeuler_geometry* geometry,
                         const mxeuler_stencil* stencil, const double* bc_kind, const double* bc, const double* source,
                         void* stream, mxeuler_stepper** out, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!schemes || !case_first || !case_cells || !topology || !geometry || !stencil || !out || n_cases < 1) {
      throw std::invalid_argument("null argument");
    }
    auto h = std::make_unique<mxeuler_stepper>();
    const auto s = static_cast<hipStream_t>(stream);
    const auto* first = reinterpret_cast<const long long*>(case_first);
    const auto* cells = reinterpret_cast<const long long*>(case_cells);
    if (policy == MXEULER_FP64) {
      h->impl = std::make_unique<mxe::Stepper<double>>(schemes, n_cases, first, cells, *topology, *geometry, *stencil,
                                                       bc_kind, bc, source, s);
    } else if (policy == MXEULER_FP32_ACC64) {
      h->impl = std::make_unique<mxe::Stepper<float>>(schemes, n_cases, first, cells, *topology, *geometry, *stencil,
                                                      bc_kind, bc, source, s);
    } else {
      throw std::invalid_argument("unknown precision policy");
    }
    h->nc = topology->n_cells;
    h->nb = topology->n_faces - topology->n_interior;
    *out = h.release();
  });
}

int mxeuler_cases(const mxeuler_stepper* s) { return s != nullptr ? s->impl->ncase() : 0; }

int mxeuler_batch_step(mxeuler_stepper* s, double cfl, int32_t local, double* residuals, double* time_steps,
                       int32_t* invalid_cells, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!residuals || !time_steps || !invalid_cells) throw std::invalid_argument("null argument");
    const int n = s->impl->ncase();
    std::vector<int> inv(static_cast<std::size_t>(n));
    s->impl->batch_step(cfl, local != 0, residuals, time_steps, inv.data());
    for (int k = 0; k < n; ++k) invalid_cells[k] = inv[std::size_t(k)];
  });
}

int mxeuler_batch_flags(mxeuler_stepper* s, const int32_t* flags, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!flags) throw std::invalid_argument("null argument");
    std::vector<int> f(flags, flags + s->impl->ncase());
    s->impl->set_flags(f.data());
  });
}

int mxeuler_batch_steady(mxeuler_stepper* s, const mxsteady_options* options, const double* case_scales,
                         const double* case_references, mxsteady_result* results, double* history, char* err,
                         size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!options || !results) throw std::invalid_argument("null argument");
    s->impl->batch_steady(*options, case_scales, case_references, results, history);
  });
}

int mxeuler_autotune(mxeuler_stepper* s, int32_t steps, double cfl, int32_t local, char* report, size_t report_len,
                     char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    const std::string r = s->impl->autotune(steps, cfl, local != 0);
    if (report != nullptr && report_len > 0) {
      std::strncpy(report, r.c_str(), report_len - 1);
      report[report_len - 1] = '\0';
    }
  });
}

int mxsteady_run(const mxsteady_ops* ops, const mxsteady_options* options, mxsteady_result* result,
                 double* history, int32_t* provisional, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!ops || !options || !result) throw std::invalid_argument("null argument");
    mxe::steady::run(*ops, *options, *result, history, provisional);
  });
}

namespace {

int steady_step(void* ctx, double cfl, double residuals[5], int32_t* invalid) {
  try {
    double h = 0.0;
    int inv = -1;
    static_cast<mxeuler_stepper*>(ctx)->impl->step(cfl, false, 0.0, true, residuals, &h, &inv);
    *invalid = inv;
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

int steady_evaluate(void* ctx, double residuals[5], int32_t* invalid) {
  try {
    int inv = -1;
    static_cast<mxeuler_stepper*>(ctx)->impl->evaluate(residuals, &inv);
    *invalid = inv;
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

int steady_order(void* ctx, int32_t on) {
  try {
    static_cast<mxeuler_stepper*>(ctx)->impl->set_second_order(on != 0);
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

}  // namespace

int mxeuler_steady(mxeuler_stepper* s, const mxsteady_options* options, mxsteady_result* result, double* history,
                   int32_t* provisional, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!s || !options || !result) throw std::invalid_argument("null argument");
    if (s->impl->ncase() != 1) throw std::invalid_argument("a batch runs with mxeuler_batch_steady");
    const mxsteady_ops ops{s, steady_step, steady_evaluate, steady_order};
    mxe::steady::run(ops, *options, *result, history, provisional);
  });
}

int mxeuler_implicit_evaluate(mxeuler_stepper* s, double norms[5], int32_t* invalid_cell, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!s || !norms || !invalid_cell) throw std::invalid_argument("null argument");
    int inv = -1;
    s->impl->implicit_evaluate(norms, &inv);
    *invalid_cell = inv;
  });
}

int mxeuler_implicit_solve(mxeuler_stepper* s, double cfl, double report[5], char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!s || !report) throw std::invalid_argument("null argument");
    s->impl->implicit_solve(cfl, report);
  });
}

int mxeuler_implicit_colours(mxeuler_stepper* s, int32_t* colours, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!s || !colours) throw std::invalid_argument("null argument");
    s->impl->implicit_colours(colours);
  });
}

int mximplicit_run(const mximplicit_ops* ops, const mxsteady_options* options, mxsteady_result* result,
                   double* history, int32_t* provisional, double* stats, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!ops || !options || !result) throw std::invalid_argument("null argument");
    mxe::steady::run_implicit(*ops, *options, *result, history, provisional, stats);
  });
}

namespace {

int implicit_jacobian(void* ctx, double residuals[5], int32_t* invalid) {
  return device_call([&] {
    int inv = -1;
    static_cast<mxeuler_stepper*>(ctx)->impl->implicit_evaluate(residuals, &inv);
    *invalid = inv;
  });
}

int implicit_solve(void* ctx, double cfl, double report[5]) {
  return device_call([&] { static_cast<mxeuler_stepper*>(ctx)->impl->implicit_solve(cfl, report); });
}

int implicit_worst(void* ctx, int32_t* cell) {
  return device_call([&] { *cell = static_cast<mxeuler_stepper*>(ctx)->impl->implicit_worst_cell(); });
}

}  // namespace

int mxeuler_implicit_steady(mxeuler_stepper* s, const mxsteady_options* options, mxsteady_result* result,
                            double* history, int32_t* provisional, double* stats, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!s || !options || !result) throw std::invalid_argument("null argument");
    if (s->impl->ncase() != 1) throw std::invalid_argument("the implicit path runs one case per stepper");
    const mximplicit_ops ops{s, steady_evaluate, steady_order, implicit_jacobian, implicit_solve, implicit_worst};
    mxe::steady::run_implicit(ops, *options, *result, history, provisional, stats);
  });
}

const char* mxeuler_version(void) { return "mxeuler 0.1 gfx950"; }

}  // extern "C"
