// The steady run around a stepper: first-order start, local time step, convergence rule.
//
// Restates, on one rank:
//   runtime::Plateau, runtime::ResidualPlateau, runtime::RoundOffExemption
//     (runtime/include/autan/runtime/round_off.hpp:14-160);
//   physics::FirstOrderStart (physics/include/autan/physics/first_order_start.hpp:22-137);
//   EulerPhysics::device_step, relative_drop, switch_order (physics/src/euler_physics.cpp:1495-1571);
//   the runtime's steady loop and convergence decision (runtime/src/run.cpp:2262-2389), with
//   round_off_floor = 0.5 * 2^-23 and round_off_band = 10 (runtime/include/autan/runtime/physics.hpp:
//   164-165).
// Host code only: the stepper is reached through the callbacks of mxsteady_ops.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mxeuler.h"

namespace mxe::steady {

constexpr int kEq = 5;
constexpr std::size_t kPlateauSteps = 100;
constexpr double kPlateauRatio = 0.9;
constexpr std::size_t kPlateauWindow = 100;
constexpr double kRoundOffFloor = 0.5 * 0x1p-23;
constexpr double kRoundOffBand = 10.0;
constexpr double kStartDrop = 1e-3;
constexpr int kStartMaxSteps = 200;
constexpr double kStartBand = 10.0;

// Equations 1..3 are the momentum components (residual_names: density, momentum-x, -y, -z, energy).
inline bool momentum(int e) { return e >= 1 && e <= 3; }

// A sequence is on a plateau once kPlateauSteps values pass without one below kPlateauRatio times
// the lowest reference so far.
class Plateau {
 public:
  bool add(double v) {
    ++count_;
    if (v < kPlateauRatio * reference_) {
      reference_ = v;
      since_ = count_;
    }
    return count_ - since_ >= kPlateauSteps;
  }

 private:
  double reference_ = std::numeric_limits<double>::infinity();
  std::size_t count_ = 0, since_ = 0;
};

// The median of a window of n values as the mean of its two middle values.
inline double middle(std::vector<double> v) {
  const auto upper = v.begin() + static_cast<std::ptrdiff_t>(v.size() / 2);
  std::nth_element(v.begin(), upper, v.end());
  return (*std::max_element(v.begin(), upper) + *upper) / 2;
}

// The plateau rule over the level of a residual (the middle of its last kPlateauWindow values).
class ResidualPlateau {
 public:
  bool add(double v) {
    window_[count_ % kPlateauWindow] = v;
    ++count_;
    if (count_ >= kPlateauWindow) {
      const double level = middle(window_);
      if (level < kPlateauRatio * reference_) {
        reference_ = level;
        since_ = count_;
      }
    }
    return since_ > 0 && count_ - since_ >= kPlateauSteps;
  }
  void reset() { *this = ResidualPlateau{}; }

 private:
  std::vector<double> window_ = std::vector<double>(kPlateauWindow, 0.0);
  double reference_ = std::numeric_limits<double>::infinity();
  std::size_t count_ = 0, since_ = 0;
};

// An equation whose reference residual is below 1e-12 of the largest is exempt while it stays at
// or below that floor; once above, it converges from a reference of its kind.
class Exemption {
 public:
  void reference(const double* first) {
    if (set_) return;
    set_ = true;
    floor_ = 1e-12 * *std::max_element(first, first + kEq);
    for (int e = 0; e < kEq; ++e) {
      first_[e] = first[e];
      zero_[e] = first[e] < floor_;
      exempt_[e] = zero_[e];
    }
  }
  void update(const double* r) {
    reference(r);
    for (int e = 0; e < kEq; ++e) {
      if (exempt_[e] && r[e] > floor_) exempt_[e] = false;
    }
  }
  bool exempt(int e) const { return set_ && exempt_[e]; }
  bool woken(int e) const { return zero_[e] && !exempt_[e]; }
  // A momentum component drops from the largest momentum reference, any other from its own.
  double drop_reference(int e, const double* ref) const {
    if (!momentum(e)) return ref[e];
    return std::max({0.0, ref[1], ref[2], ref[3]});
  }
  double woken_reference(int e) const {
    if (momentum(e)) return drop_reference(e, first_.data());
    return *std::max_element(first_.begin(), first_.end());
  }

 private:
  bool set_ = false;
  double floor_ = 0.0;
  std::array<double, kEq> first_{};
  std::array<bool, kEq> zero_{}, exempt_{};
};

// The first-order start's end: a drop of kStartDrop, a stalled drop (fp64) or every watched level
// within kStartBand floors (fp32+acc64), else after kStartMaxSteps + 1 steps.
class FirstOrderStart {
 public:
  explicit FirstOrderStart(bool levels) : levels_(levels), samples_(levels ? kEq * kPlateauWindow : 0, 0.0) {}

  bool ends_after(double drop, const double* residuals, const double* floors) {
    if (!levels_) {
      const bool flat = plateau_.add(drop);
      if (drop < kStartDrop) {
        end_ = MXSTEADY_END_DROP;
      } else if (flat) {
        end_ = MXSTEADY_END_STALL;
      } else if (++steps_ > kStartMaxSteps) {
        end_ = MXSTEADY_END_CAP;
      } else {
        return false;
      }
      return true;
    }
    for (int k = 0; k < kEq; ++k) samples_[std::size_t(k) * kPlateauWindow + count_ % kPlateauWindow] = residuals[k];
    ++count_;
    if (drop < kStartDrop) {
      end_ = MXSTEADY_END_DROP;
    } else if (crossed(floors)) {
      end_ = MXSTEADY_END_LEVEL;
    } else if (++steps_ > kStartMaxSteps) {
      end_ = MXSTEADY_END_CAP;
    } else {
      return false;
    }
    return true;
  }
  int end() const { return end_; }

 private:
  bool crossed(const double* floors) const {
    if (count_ < kPlateauWindow) return false;
    bool watched = false;
    for (int k = 0; k < kEq; ++k) {
      if (!(floors[k] > 0.0)) continue;
      watched = true;
      const auto b = samples_.begin() + static_cast<std::ptrdiff_t>(std::size_t(k) * kPlateauWindow);
      const double level = middle(std::vector<double>(b, b + static_cast<std::ptrdiff_t>(kPlateauWindow)));
      if (!(level <= kStartBand * floors[k])) return false;
    }
    return watched;
  }

  bool levels_;
  Plateau plateau_;
  int steps_ = 0;
  int end_ = MXSTEADY_END_NONE;
  std::size_t count_ = 0;
  std::vector<double> samples_;
};

inline void call(int rc, const char* what) {
  if (rc != 0) throw std::runtime_error(std::string("steady run: ") + what + " failed");
}

// One steady run. history (5 per iteration) and provisional (1 per iteration) may be null.
inline void run(const mxsteady_ops& ops, const mxsteady_options& o, mxsteady_result& out, double* history,
                int32_t* provisional) {
  if (!ops.step || !ops.evaluate || !ops.set_second_order) throw std::invalid_argument("steady run: missing callback");
  if (o.max_iterations < 1) throw std::invalid_argument("steady run: max_iterations must be >= 1");
  if (!(o.residual_orders > 0.0)) throw std::invalid_argument("steady run: residual_orders must be > 0");
  if (!(o.cfl > 0.0)) throw std::invalid_argument("steady run: cfl must be > 0");
  if (o.start_order != 1 && o.start_order != 2) throw std::invalid_argument("steady run: start_order is 1 or 2");
  out = mxsteady_result{};
  out.diverged_cell = -1;
  const bool levels = o.level_rule != 0;
  std::array<double, kEq> floors{};  // the runtime's convergence floors (fp32+acc64), none: zeros
  const bool has_floor = levels && o.scales != nullptr;
  if (has_floor) {
    for (int e = 0; e < kEq; ++e) {
      if (!(std::isfinite(o.scales[e]) && o.scales[e] >= 0.0)) throw std::invalid_argument("steady run: bad scale");
      floors[e] = kRoundOffFloor * o.scales[e];
    }
  }
  // The physics' side: order, exemption, drops and the first-order start.
  bool second = o.start_order == 2;
  call(ops.set_second_order(ops.ctx, second ? 1 : 0), "set_second_order");
  Exemption phys;
  std::vector<double> first_norms;
  FirstOrderStart start(levels);
  bool checked = false;
  // The runtime's side.
  Exemption runtime;
  std::array<double, kEq> first{};
  bool have_first = false;
  if (o.reference != nullptr) {
    std::copy(o.reference, o.reference + kEq, first.begin());
    have_first = true;
  }
  std::vector<ResidualPlateau> plateaus(kEq);
  const double drop_to = std::pow(10.0, -o.residual_orders);

  for (int64_t it = 1; it <= o.max_iterations; ++it) {
    const bool prov = !second;
    double reported[kEq];
    bool have_reported = false;
    int32_t invalid = -1;
    bool stepped = false;
    if (!checked) {
      checked = true;
      int32_t inv = -1;
      call(ops.evaluate(ops.ctx, reported, &inv), "evaluate");
      if (inv >= 0) {
        invalid = inv;
        have_reported = true;
      }
    }
    if (invalid < 0) {
      have_reported = false;
      if (!second) {
        int32_t ignored = -1;
        call(ops.set_second_order(ops.ctx, 1), "set_second_order");
        call(ops.evaluate(ops.ctx, reported, &ignored), "evaluate");
        call(ops.set_second_order(ops.ctx, 0), "set_second_order");
        have_reported = true;
      }
      double r[kEq];
      call(ops.step(ops.ctx, o.cfl, r, &invalid), "step");
      stepped = true;
      // relative_drop, on the residuals the runtime receives.
      const double* rep = have_reported ? reported : r;
      phys.update(rep);
      if (first_norms.empty()) first_norms.assign(r, r + kEq);
      double drop = 0.0;
      for (int k = 0; k < kEq; ++k) {
        if (phys.exempt(k)) continue;
        const double from = phys.woken(k) ? phys.woken_reference(k) : first_norms[std::size_t(k)];
        if (from > 0.0) drop = std::max(drop, r[k] / from);
      }
      // switch_order
      if (!second) {
        double lf[kEq] = {};
        if (levels && o.scales != nullptr) {
          for (int k = 0; k < kEq; ++k) lf[k] = phys.exempt(k) ? 0.0 : kRoundOffFloor * o.scales[k];
        }
        if (start.ends_after(drop, r, lf)) {
          second = true;
          call(ops.set_second_order(ops.ctx, 1), "set_second_order");
          first_norms.clear();
          out.first_order_end = start.end();
        }
      }
      if (!have_reported) std::copy(r, r + kEq, reported);
    }
    (void)stepped;
    bool finite = invalid < 0;
    for (int e = 0; e < kEq; ++e) finite = finite && std::isfinite(reported[e]);
    std::copy(reported, reported + kEq, out.residuals);
    if (!finite) {
      out.status = MXSTEADY_DIVERGED;
      out.diverged_cell = invalid;
      std::copy(first.begin(), first.end(), out.reference);
      return;
    }
    if (prov) ++out.first_order_steps;
    if (history != nullptr) std::copy(reported, reported + kEq, history + kEq * (it - 1));
    if (provisional != nullptr) provisional[it - 1] = prov ? 1 : 0;
    out.iterations = it;
    if (!have_first) {
      std::copy(reported, reported + kEq, first.begin());
      have_first = true;
    }
    runtime.reference(first.data());
    runtime.update(reported);
    bool converged = true;
    int32_t at_plateau = 0;
    for (int e = 0; e < kEq; ++e) {
      const double r = reported[e];
      if (prov) plateaus[std::size_t(e)].reset();
      const bool flat = !prov && plateaus[std::size_t(e)].add(r);
      const bool met = runtime.exempt(e) ||
                       r <= (runtime.woken(e) ? runtime.woken_reference(e) : runtime.drop_reference(e, first.data())) *
                                drop_to;
      const bool plateau = has_floor && flat && r <= kRoundOffBand * floors[std::size_t(e)];
      if (!met && plateau) at_plateau |= 1 << e;
      converged = converged && (met || plateau);
    }
    if (converged) {
      out.status = MXSTEADY_CONVERGED;
      out.plateau_mask = at_plateau;
      std::copy(first.begin(), first.end(), out.reference);
      return;
    }
  }
  out.status = MXSTEADY_BUDGET;
  std::copy(first.begin(), first.end(), out.reference);
}

}  // namespace mxe::steady
