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

inline void check_options(const mxsteady_options& o) {
  if (o.max_iterations < 1) throw std::invalid_argument("steady run: max_iterations must be >= 1");
  if (!(o.residual_orders > 0.0)) throw std::invalid_argument("steady run: residual_orders must be > 0");
  if (!(o.cfl > 0.0)) throw std::invalid_argument("steady run: cfl must be > 0");
  if (o.start_order != 1 && o.start_order != 2) throw std::invalid_argument("steady run: start_order is 1 or 2");
  if (o.level_rule && o.scales != nullptr) {
    for (int e = 0; e < kEq; ++e) {
      if (!(std::isfinite(o.scales[e]) && o.scales[e] >= 0.0)) throw std::invalid_argument("steady run: bad scale");
    }
  }
}

// One steady run's decisions, iteration by iteration: the physics' side (exemption, relative
// drop, first-order start) and the runtime's (history, divergence, convergence).
class SteadyCase {
 public:
  explicit SteadyCase(const mxsteady_options& o) : o_(o), start_(o.level_rule != 0) {
    check_options(o);
    out_.diverged_cell = -1;
    second_ = o.start_order == 2;
    has_floor_ = o.level_rule != 0 && o.scales != nullptr;
    for (int e = 0; e < kEq && has_floor_; ++e) floors_[std::size_t(e)] = kRoundOffFloor * o.scales[e];
    if (o.reference != nullptr) {
      std::copy(o.reference, o.reference + kEq, first_.begin());
      have_first_ = true;
    }
    plateaus_.resize(kEq);
    drop_to_ = std::pow(10.0, -o.residual_orders);
  }

  bool second() const { return second_; }
  bool done() const { return done_; }
  // Whether the last record() ended the first-order start.
  bool switched() const { return switched_; }
  const mxsteady_result& result() const { return out_; }

  // The initial state is invalid (the first evaluate's check): the run diverges before stepping.
  void diverged_before(int32_t invalid, const double* reported) {
    std::copy(reported, reported + kEq, out_.residuals);
    finish(MXSTEADY_DIVERGED);
    out_.diverged_cell = invalid;
  }

  // One iteration: r the step's norms (of R(u) at the order solved), reported what the history
  // records (the second-order residual of the same state during the first-order start; null: r),
  // invalid the step's invalid cell (-1: none). history (5 per iteration) and provisional may be
  // null. Returns done().
  bool record(const double* r, const double* reported, int32_t invalid, double* history, int32_t* provisional) {
    physics(r, reported);
    return runtime(r, reported, invalid, history, provisional);
  }

  // The physics' side of one iteration (EulerPhysics::relative_drop, then switch_order): the largest
  // relative residual over the equations of the order being solved, and whether the first-order
  // start ends (switched()). r: the step's norms; reported: what the history records (null: r).
  double physics(const double* r, const double* reported) {
    switched_ = false;
    prov_ = !second_;
    const double* rep = reported != nullptr ? reported : r;
    // EulerPhysics::relative_drop on the residuals the runtime receives, then switch_order.
    phys_.update(rep);
    if (first_norms_.empty()) first_norms_.assign(r, r + kEq);
    double drop = 0.0;
    for (int k = 0; k < kEq; ++k) {
      if (phys_.exempt(k)) continue;
      const double from = phys_.woken(k) ? phys_.woken_reference(k) : first_norms_[std::size_t(k)];
      if (from > 0.0) drop = std::max(drop, r[k] / from);
    }
    if (!second_) {
      double lf[kEq] = {};
      if (o_.level_rule && o_.scales != nullptr) {
        for (int k = 0; k < kEq; ++k) lf[k] = phys_.exempt(k) ? 0.0 : kRoundOffFloor * o_.scales[k];
      }
      if (start_.ends_after(drop, r, lf)) {
        second_ = true;
        switched_ = true;
        first_norms_.clear();
        out_.first_order_end = start_.end();
      }
    }
    return drop;
  }

  // The runtime's side of the same iteration: divergence, history, convergence. Returns done().
  bool runtime(const double* r, const double* reported, int32_t invalid, double* history, int32_t* provisional) {
    const bool prov = prov_;
    const double* rep = reported != nullptr ? reported : r;
    bool finite = invalid < 0;
    for (int e = 0; e < kEq; ++e) finite = finite && std::isfinite(rep[e]);
    std::copy(rep, rep + kEq, out_.residuals);
    if (!finite) {
      finish(MXSTEADY_DIVERGED);
      out_.diverged_cell = invalid;
      return true;
    }
    const int64_t it = out_.iterations;
    if (prov) ++out_.first_order_steps;
    if (history != nullptr) std::copy(rep, rep + kEq, history + kEq * it);
    if (provisional != nullptr) provisional[it] = prov ? 1 : 0;
    out_.iterations = it + 1;
    if (!have_first_) {
      std::copy(rep, rep + kEq, first_.begin());
      have_first_ = true;
    }
    runtime_.reference(first_.data());
    runtime_.update(rep);
    bool converged = true;
    int32_t at_plateau = 0;
    for (int e = 0; e < kEq; ++e) {
      const double v = rep[e];
      if (prov) plateaus_[std::size_t(e)].reset();
      const bool flat = !prov && plateaus_[std::size_t(e)].add(v);
      const bool met = runtime_.exempt(e) ||
                       v <= (runtime_.woken(e) ? runtime_.woken_reference(e) : runtime_.drop_reference(e, first_.data())) *
                                drop_to_;
      const bool plateau = has_floor_ && flat && v <= kRoundOffBand * floors_[std::size_t(e)];
      if (!met && plateau) at_plateau |= 1 << e;
      converged = converged && (met || plateau);
    }
    if (converged) {
      out_.plateau_mask = at_plateau;
      finish(MXSTEADY_CONVERGED);
      return true;
    }
    if (out_.iterations >= o_.max_iterations) {
      finish(MXSTEADY_BUDGET);
      return true;
    }
    return false;
  }

 private:
  void finish(int32_t status) {
    out_.status = status;
    std::copy(first_.begin(), first_.end(), out_.reference);
    done_ = true;
  }

  mxsteady_options o_;
  mxsteady_result out_{};
  bool second_ = false, switched_ = false, done_ = false, has_floor_ = false, have_first_ = false, prov_ = false;
  std::array<double, kEq> floors_{}, first_{};
  double drop_to_ = 0.0;
  Exemption phys_, runtime_;
  std::vector<double> first_norms_;
  FirstOrderStart start_;
  std::vector<ResidualPlateau> plateaus_;
};

// One steady run through the callbacks (EulerPhysics::device_step and the runtime's loop).
// history (5 per iteration) and provisional (1 per iteration) may be null.
inline void run(const mxsteady_ops& ops, const mxsteady_options& o, mxsteady_result& out, double* history,
                int32_t* provisional) {
  if (!ops.step || !ops.evaluate || !ops.set_second_order) throw std::invalid_argument("steady run: missing callback");
  SteadyCase sc(o);
  call(ops.set_second_order(ops.ctx, sc.second() ? 1 : 0), "set_second_order");
  {
    // As the host's step: a stored state the kernels cannot evaluate diverges at once.
    double reported[kEq];
    int32_t inv = -1;
    call(ops.evaluate(ops.ctx, reported, &inv), "evaluate");
    if (inv >= 0) {
      sc.diverged_before(inv, reported);
      out = sc.result();
      return;
    }
  }
  while (!sc.done()) {
    double reported[kEq];
    const bool prov = !sc.second();
    if (prov) {
      int32_t ignored = -1;
      call(ops.set_second_order(ops.ctx, 1), "set_second_order");
      call(ops.evaluate(ops.ctx, reported, &ignored), "evaluate");
      call(ops.set_second_order(ops.ctx, 0), "set_second_order");
    }
    double r[kEq];
    int32_t invalid = -1;
    call(ops.step(ops.ctx, o.cfl, r, &invalid), "step");
    sc.record(r, prov ? reported : nullptr, invalid, history, provisional);
    if (sc.switched()) call(ops.set_second_order(ops.ctx, 1), "set_second_order");
  }
  out = sc.result();
}

// ---- the steady implicit run (EulerPhysics::steady_step with its switched evolution relaxation,
// physics/src/euler_physics.cpp:59-86, 1056-1250), over callbacks.

constexpr double kCflStart = 1000.0;
constexpr double kCflMin = 1.0;
constexpr double kCflMax = 1e6;
constexpr double kCflRise = 1.1;

// The CFL number of each step: CFL_0 over the largest relative residual, times a penalty halved
// after a step that was shortened or rejected or whose relative residual rose by more than
// kCflRise, recovering by half otherwise; after a shortened or rejected step at most half the CFL
// number used.
class CflControl {
 public:
  double next(double drop) {
    const bool rose = previous_ > 0.0 && drop > kCflRise * previous_;
    penalty_ = (shortened_ || rose) ? 0.5 * penalty_ : std::min(1.0, 1.5 * penalty_);
    previous_ = drop;
    if (drop > 0.0) cfl_ = std::clamp(std::min(penalty_ * kCflStart / drop, cap_), kCflMin, kCflMax);
    cap_ = kCflMax;
    return cfl_;
  }
  // After the solve: report = {cycles, residual drop, alpha, failed, singular}. Returns whether the
  // run has diverged (no usable step at the smallest CFL number).
  bool after(const double* report, double drop) {
    if (report[3] != 0.0) {
      if (drop > 0.0) penalty_ = std::min(penalty_, cfl_ * drop / kCflStart);
      cap_ = 0.5 * cfl_;
      shortened_ = true;
      return cfl_ <= kCflMin;
    }
    const double alpha = report[2];
    shortened_ = alpha < 1.0;
    if (shortened_) {
      if (drop > 0.0) penalty_ = std::min(penalty_, cfl_ * drop / kCflStart);
      cap_ = 0.5 * cfl_;
    }
    return alpha < 1e-6 && cfl_ <= kCflMin;
  }
  double cfl() const { return cfl_; }

 private:
  double cfl_ = kCflStart, penalty_ = 1.0, previous_ = 0.0, cap_ = kCflMax;
  bool shortened_ = false;
};

// One steady implicit run: per iteration (the second-order residual for the history during the
// first-order start), the residual and Jacobian at the order being solved, the physics' decisions
// (relative drop, end of the first-order start), the CFL number, the solve and update, then the
// runtime's (history, divergence, convergence). stats (4 per iteration: CFL, linear cycles, linear
// residual drop, alpha) may be null.
inline void run_implicit(const mximplicit_ops& ops, const mxsteady_options& o, mxsteady_result& out, double* history,
                         int32_t* provisional, double* stats) {
  if (!ops.evaluate || !ops.set_second_order || !ops.jacobian || !ops.solve) {
    throw std::invalid_argument("implicit run: missing callback");
  }
  SteadyCase sc(o);
  CflControl ctl;
  call(ops.set_second_order(ops.ctx, sc.second() ? 1 : 0), "set_second_order");
  {
    double reported[kEq];
    int32_t inv = -1;
    call(ops.evaluate(ops.ctx, reported, &inv), "evaluate");
    if (inv >= 0) {
      sc.diverged_before(inv, reported);
      out = sc.result();
      return;
    }
  }
  while (!sc.done()) {
    double reported[kEq];
    const bool prov = !sc.second();
    if (prov) {
      int32_t ignored = -1;
      call(ops.set_second_order(ops.ctx, 1), "set_second_order");
      call(ops.evaluate(ops.ctx, reported, &ignored), "evaluate");
      call(ops.set_second_order(ops.ctx, 0), "set_second_order");
    }
    double r[kEq];
    int32_t invalid = -1;
    call(ops.jacobian(ops.ctx, r, &invalid), "jacobian");
    const double drop = sc.physics(r, prov ? reported : nullptr);
    const double cfl = ctl.next(drop);
    double report[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (invalid < 0) {
      call(ops.solve(ops.ctx, cfl, report), "solve");
      if (ctl.after(report, drop)) {
        int32_t worst = 0;
        if (ops.worst_cell != nullptr) call(ops.worst_cell(ops.ctx, &worst), "worst_cell");
        invalid = worst;
      }
    }
    if (stats != nullptr) {
      const int64_t it = sc.result().iterations;
      stats[4 * it] = cfl;
      stats[4 * it + 1] = report[0];
      stats[4 * it + 2] = report[1];
      stats[4 * it + 3] = report[2];
    }
    sc.runtime(r, prov ? reported : nullptr, invalid, history, provisional);
    if (sc.switched()) call(ops.set_second_order(ops.ctx, 1), "set_second_order");
  }
  out = sc.result();
}

}  // namespace mxe::steady
