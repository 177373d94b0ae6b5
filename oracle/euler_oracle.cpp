// CPU oracle of the explicit Euler residual loop: test infrastructure, not the product.
//
// A serial C++ restatement of the reference's arithmetic for this path, used only by the tests,
// the smoke check and the CPU baseline of the benchmark to check and time the HIP library. It
// follows, body for body and operation for operation:
//   kernels/euler.py (every @pointwise body and loop of the residual),
//   codegen/src/autan_codegen/emit_cpp.py:117-120,145-161,207-297 (arguments converted to the
//     compute type on use, `rel` as an FP64 difference then rounded, SoA indexing),
//   codegen/src/autan_codegen/frontend.py (abs/min/max as selects: abs(x) = x >= 0 ? x : -x,
//     min(a, b) = a <= b ? a : b, max(a, b) = a >= b ? a : b; left-to-right association; matmul
//     and dot starting from the first product; constants folded in FP64),
//   codegen/src/autan_codegen/emit_cpu.py:126-182,346-464,506-564 (Sum in the accumulation type,
//     Min/Max in the compute type with ties keeping the first entry, scatter and gather orders),
//   physics/src/euler.cpp:113-136,229-245,649-685,925-941 (loop order, p_ref, limiter scales),
//   physics/src/euler_physics.cpp:1040-1054,1362-1433 (norms, SSP-RK3 step, first invalid cell),
//   physics/src/device_stepper.cuh:83-103 (local time step, RK stage),
//   runtime/src/exact_sum.cpp:22-103 (exact sums).
// Built with -O3 -ffp-contract=off, as the reference builds its host code (CMakeLists.txt:45-55).
//
// Parity status, pinned by the tests in tests/test_oracle_*.py: every flux by consistency, the
// stationary contact, rotation invariance and conservation antisymmetry, and HLLC/Roe/AUSM against
// exact Riemann wave structure; the boundary states by their closed forms; the gradient by
// exactness on linear fields; the limiter by its closed form; the spectral radius by its closed
// form; the RK3 stage by exactness on polynomials in time; the exact sums against math.fsum;
// uniform flow steady under every condition; discrete conservation.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace oracle {

using index_t = std::int32_t;
using offset_t = std::int64_t;

// Precision policies (codegen/include/autan/kernels/runtime.hpp:73-99).
struct fp64 {
  using state = double;
  using grad = double;
  using geom_value = double;
  using compute = double;
  static constexpr int id = 0;
};
struct fp32_acc64 {
  using state = float;
  using grad = float;
  using geom_value = float;
  using compute = float;
  static constexpr int id = 1;
};
using A = double;  // accumulation and residual type of both policies

// ------------------------------------------------------------------------------------------ math
// The generator's select semantics (frontend.py:438-442, 486-489).
// The bodies' square root, exponential and logarithm, called unqualified so that the dual numbers
// of the Jacobian (below) take their own.
inline float osqrt(float x) { return std::sqrt(x); }
inline double osqrt(double x) { return std::sqrt(x); }
inline float oexp(float x) { return std::exp(x); }
inline double oexp(double x) { return std::exp(x); }
inline float olog(float x) { return std::log(x); }
inline double olog(double x) { return std::log(x); }

template <class C> inline C sabs(C x) { return x >= C(0) ? x : -x; }
template <class C> inline C smin(C a, C b) { return a <= b ? a : b; }
template <class C> inline C smax(C a, C b) { return a >= b ? a : b; }
template <class C> inline C dot3(const C* a, const C* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// ------------------------------------------------------------------------------ thermodynamics
// kernels/euler.py:94-105
template <class C>
void to_primitive(const C* dq, const C* ur, C g, C* w) {
  const C rho = ur[0] + dq[0];
  const C mr[3] = {ur[1], ur[2], ur[3]};
  const C dm[3] = {dq[1], dq[2], dq[3]};
  const C m[3] = {mr[0] + dm[0], mr[1] + dm[1], mr[2] + dm[2]};
  const C t[3] = {C(2.0) * mr[0] + dm[0], C(2.0) * mr[1] + dm[1], C(2.0) * mr[2] + dm[2]};
  const C dk = (dot3(dm, t) * ur[0] - dot3(mr, mr) * dq[0]) / (rho * ur[0]);
  w[0] = rho;
  w[1] = m[0] / rho;
  w[2] = m[1] / rho;
  w[3] = m[2] / rho;
  w[4] = (g - C(1.0)) * (dq[4] - C(0.5) * dk);
}

// kernels/euler.py:113-115
template <class C>
C total_enthalpy(const C* w, C g, C pr) {
  const C v[3] = {w[1], w[2], w[3]};
  return g / (g - C(1.0)) * (w[4] + pr) / w[0] + C(0.5) * dot3(v, v);
}

// kernels/euler.py:119-130
template <class C>
void physical_flux(const C* w, const C* n, C g, C pr, C* f) {
  const C v[3] = {w[1], w[2], w[3]};
  const C vn = dot3(v, n);
  const C m = w[0] * vn;
  f[0] = m;
  f[1] = m * v[0] + w[4] * n[0];
  f[2] = m * v[1] + w[4] * n[1];
  f[3] = m * v[2] + w[4] * n[2];
  f[4] = m * total_enthalpy(w, g, pr);
}

// ------------------------------------------------------------------------------ fluxes
// kernels/euler.py:136-141
template <class C>
C harten(C lam, C delta) {
  C a = sabs(lam);
  if (a < delta) a = (lam * lam + delta * delta) / (C(2.0) * delta);
  return a;
}

// kernels/euler.py:145-175
template <class C>
void roe(const C* wl, const C* wr, const C* n, C g, C pr, C* out) {
  const C rl = wl[0], rr = wr[0];
  const C vl[3] = {wl[1], wl[2], wl[3]}, vr[3] = {wr[1], wr[2], wr[3]};
  const C hl = total_enthalpy(wl, g, pr), hr = total_enthalpy(wr, g, pr);
  const C s = osqrt(rr / rl);
  const C rho = s * rl;
  C v[3];
  for (int i = 0; i < 3; ++i) v[i] = (vl[i] + s * vr[i]) / (C(1.0) + s);
  const C h = (hl + s * hr) / (C(1.0) + s);
  const C vn = dot3(v, n);
  const C q2 = dot3(v, v);
  const C c = osqrt((g - C(1.0)) * (h - C(0.5) * q2));
  const C dp = wr[4] - wl[4];
  const C drho = rr - rl;
  const C dv[3] = {vr[0] - vl[0], vr[1] - vl[1], vr[2] - vl[2]};
  const C dvn = dot3(dv, n);
  const C delta = C(0.1) * c;
  const C l1 = harten(vn - c, delta);
  const C l2 = sabs(vn);
  const C l5 = harten(vn + c, delta);
  const C a1 = l1 * (dp - rho * c * dvn) / (C(2.0) * c * c);
  const C a5 = l5 * (dp + rho * c * dvn) / (C(2.0) * c * c);
  const C a2 = l2 * (drho - dp / (c * c));
  C shear[3];
  for (int i = 0; i < 3; ++i) shear[i] = dv[i] - dvn * n[i];
  const C d0 = a1 + a2 + a5;
  C dm[3];
  for (int i = 0; i < 3; ++i) {
    dm[i] = a1 * (v[i] - c * n[i]) + a2 * v[i] + l2 * rho * shear[i] + a5 * (v[i] + c * n[i]);
  }
  const C de = a1 * (h - c * vn) + a2 * C(0.5) * q2 + l2 * rho * (dot3(v, dv) - vn * dvn) + a5 * (h + c * vn);
  C fl[5], fr[5];
  physical_flux(wl, n, g, pr, fl);
  physical_flux(wr, n, g, pr, fr);
  const C d[5] = {d0, dm[0], dm[1], dm[2], de};
  for (int k = 0; k < 5; ++k) out[k] = C(0.5) * (fl[k] + fr[k]) - C(0.5) * d[k];
}

// kernels/euler.py:178-191
template <class C>
void hllc_star(const C* w, const C* n, C sk, C sm, C g, C pr, C* out) {
  const C v[3] = {w[1], w[2], w[3]};
  const C vn = dot3(v, n);
  const C rs = w[0] * (sk - vn) / (sk - sm);
  const C p = w[4] + pr;
  const C e = p / ((g - C(1.0)) * w[0]) + C(0.5) * dot3(v, v);
  out[0] = rs;
  for (int i = 0; i < 3; ++i) out[1 + i] = rs * (v[i] + (sm - vn) * n[i]);
  out[4] = rs * (e + (sm - vn) * (sm + p / (w[0] * (sk - vn))));
}

// kernels/euler.py:194-203
template <class C>
void conserved(const C* w, C g, C pr, C* out) {
  const C v[3] = {w[1], w[2], w[3]};
  out[0] = w[0];
  out[1] = w[0] * v[0];
  out[2] = w[0] * v[1];
  out[3] = w[0] * v[2];
  out[4] = (w[4] + pr) / (g - C(1.0)) + C(0.5) * w[0] * dot3(v, v);
}

// kernels/euler.py:206-233
template <class C>
void hllc(const C* wl, const C* wr, const C* n, C g, C pr, C* out) {
  const C rl = wl[0], rr = wr[0];
  const C vl[3] = {wl[1], wl[2], wl[3]}, vr[3] = {wr[1], wr[2], wr[3]};
  const C vnl = dot3(vl, n), vnr = dot3(vr, n);
  const C cl = osqrt(g * (wl[4] + pr) / rl);
  const C cr = osqrt(g * (wr[4] + pr) / rr);
  const C s = osqrt(rr / rl);
  C v[3];
  for (int i = 0; i < 3; ++i) v[i] = (vl[i] + s * vr[i]) / (C(1.0) + s);
  const C h = (total_enthalpy(wl, g, pr) + s * total_enthalpy(wr, g, pr)) / (C(1.0) + s);
  const C c = osqrt((g - C(1.0)) * (h - C(0.5) * dot3(v, v)));
  const C vn = dot3(v, n);
  const C sl = smin(vnl - cl, vn - c);
  const C sr = smax(vnr + cr, vn + c);
  const C sm = (wr[4] - wl[4] + rl * vnl * (sl - vnl) - rr * vnr * (sr - vnr)) / (rl * (sl - vnl) - rr * (sr - vnr));
  C fl[5], fr[5];
  physical_flux(wl, n, g, pr, fl);
  physical_flux(wr, n, g, pr, fr);
  if (sl >= C(0.0)) {
    for (int k = 0; k < 5; ++k) out[k] = fl[k];
  } else if (sm >= C(0.0)) {
    C st[5], u[5];
    hllc_star(wl, n, sl, sm, g, pr, st);
    conserved(wl, g, pr, u);
    for (int k = 0; k < 5; ++k) out[k] = fl[k] + sl * (st[k] - u[k]);
  } else if (sr > C(0.0)) {
    C st[5], u[5];
    hllc_star(wr, n, sr, sm, g, pr, st);
    conserved(wr, g, pr, u);
    for (int k = 0; k < 5; ++k) out[k] = fr[k] + sr * (st[k] - u[k]);
  } else {
    for (int k = 0; k < 5; ++k) out[k] = fr[k];
  }
}

// kernels/euler.py:236-273; 16 beta = 2 folded as the generator folds constant products.
template <class C>
C m4_plus(C m) {
  const C mp = C(0.25) * (m + C(1.0)) * (m + C(1.0));
  const C mm = C(-0.25) * (m - C(1.0)) * (m - C(1.0));
  return sabs(m) >= C(1.0) ? C(0.5) * (m + sabs(m)) : mp * (C(1.0) - C(2.0) * mm);
}
template <class C>
C m4_minus(C m) {
  const C mp = C(0.25) * (m + C(1.0)) * (m + C(1.0));
  const C mm = C(-0.25) * (m - C(1.0)) * (m - C(1.0));
  return sabs(m) >= C(1.0) ? C(0.5) * (m - sabs(m)) : mm * (C(1.0) + C(2.0) * mp);
}
template <class C>
C p5_plus(C m, C alpha) {
  const C mp = C(0.25) * (m + C(1.0)) * (m + C(1.0));
  const C mm = C(-0.25) * (m - C(1.0)) * (m - C(1.0));
  const C supersonic = m > C(0.0) ? C(1.0) : C(0.0);
  return sabs(m) >= C(1.0) ? supersonic : mp * ((C(2.0) - m) - C(16.0) * alpha * m * mm);
}
template <class C>
C p5_minus(C m, C alpha) {
  const C mp = C(0.25) * (m + C(1.0)) * (m + C(1.0));
  const C mm = C(-0.25) * (m - C(1.0)) * (m - C(1.0));
  const C supersonic = m < C(0.0) ? C(1.0) : C(0.0);
  return sabs(m) >= C(1.0) ? supersonic : mm * ((C(-2.0) - m) + C(16.0) * alpha * m * mp);
}

// kernels/euler.py:276-303; the signed constants -K_p, -K_u folded, sigma = 1 multiplied away.
template <class C>
void ausm_up(const C* wl, const C* wr, const C* n, C g, C minf, C pr, C* out) {
  const C rl = wl[0], rr = wr[0];
  const C vl[3] = {wl[1], wl[2], wl[3]}, vr[3] = {wr[1], wr[2], wr[3]};
  const C vnl = dot3(vl, n), vnr = dot3(vr, n);
  const C hl = total_enthalpy(wl, g, pr), hr = total_enthalpy(wr, g, pr);
  const C csl = osqrt(C(2.0) * (g - C(1.0)) / (g + C(1.0)) * hl);
  const C csr = osqrt(C(2.0) * (g - C(1.0)) / (g + C(1.0)) * hr);
  const C al = csl * csl / smax(csl, vnl);
  const C ar = csr * csr / smax(csr, -vnr);
  const C a = smin(al, ar);
  const C ml = vnl / a;
  const C mr = vnr / a;
  const C mbar2 = C(0.5) * (vnl * vnl + vnr * vnr) / (a * a);
  const C mo2 = smin(C(1.0), smax(mbar2, minf * minf));
  const C mo = osqrt(mo2);
  const C fa = mo * (C(2.0) - mo);
  const C alpha = C(0.1875) * (C(-4.0) + C(5.0) * fa * fa);
  const C rho = C(0.5) * (rl + rr);
  const C mp = C(-0.25) / fa * smax(C(1.0) - mbar2, C(0.0)) * (wr[4] - wl[4]) / (rho * a * a);
  const C m = m4_plus(ml) + m4_minus(mr) + mp;
  const C pu = C(-0.75) * p5_plus(ml, alpha) * p5_minus(mr, alpha) * (rl + rr) * fa * a * (vnr - vnl);
  const C pp = p5_plus(ml, alpha);
  const C pm = p5_minus(mr, alpha);
  const C p = pp * wl[4] + pm * wr[4] + pu + (pp + pm - C(1.0)) * pr;
  const bool up = m > C(0.0);
  const C mdot = up ? a * m * rl : a * m * rr;
  const C* vu = up ? vl : vr;
  const C hu = up ? hl : hr;
  out[0] = mdot;
  for (int i = 0; i < 3; ++i) out[1 + i] = mdot * vu[i] + p * n[i];
  out[4] = mdot * hu;
}

// ------------------------------------------------------------------------------ boundary states
// kernels/euler.py:309-337
template <class C>
void farfield_state(const C* wi, const C* wf, const C* n, C g, C* w) {
  const C vi[3] = {wi[1], wi[2], wi[3]}, vf[3] = {wf[1], wf[2], wf[3]};
  const C vni = dot3(vi, n), vnf = dot3(vf, n);
  const C ci = osqrt(g * wi[4] / wi[0]);
  const C cf = osqrt(g * wf[4] / wf[0]);
  const C rp = vni + C(2.0) * ci / (g - C(1.0));
  const C rm = vnf - C(2.0) * cf / (g - C(1.0));
  const C vnb = C(0.5) * (rp + rm);
  const C cb = C(0.25) * (g - C(1.0)) * (rp - rm);
  C ent, vt[3];
  if (vnb < C(0.0)) {
    ent = wf[4] / oexp(g * olog(wf[0]));
    for (int i = 0; i < 3; ++i) vt[i] = vf[i] - vnf * n[i];
  } else {
    ent = wi[4] / oexp(g * olog(wi[0]));
    for (int i = 0; i < 3; ++i) vt[i] = vi[i] - vni * n[i];
  }
  const C rb = oexp(olog(cb * cb / (g * ent)) / (g - C(1.0)));
  if (vni >= ci) {
    for (int k = 0; k < 5; ++k) w[k] = wi[k];
  } else if (vnf <= -cf) {
    for (int k = 0; k < 5; ++k) w[k] = wf[k];
  } else {
    w[0] = rb;
    for (int i = 0; i < 3; ++i) w[1 + i] = vt[i] + vnb * n[i];
    w[4] = rb * cb * cb / g;
  }
}

// kernels/euler.py:340-358
template <class C>
void inlet_state(const C* wi, const C* data, const C* n, C g, C rgas, C* w) {
  const C p0 = data[0], t0 = data[1];
  const C gm = g - C(1.0);
  const C vi[3] = {wi[1], wi[2], wi[3]};
  const C ci = osqrt(g * wi[4] / wi[0]);
  const C r = dot3(vi, n) + C(2.0) * ci / gm;
  const C h0 = g * rgas * t0 / gm;
  const C disc = smax((gm + C(2.0)) * h0 - C(0.5) * gm * r * r, C(0.0));
  const C cb = gm * (r + osqrt(disc)) / (gm + C(2.0));
  const C vnb = smin(r - C(2.0) * cb / gm, C(0.0));
  const C tb = t0 - C(0.5) * gm * vnb * vnb / (g * rgas);
  const C pb = p0 * oexp(g / gm * olog(tb / t0));
  const C rb = pb / (rgas * tb);
  w[0] = rb;
  w[1] = vnb * n[0];
  w[2] = vnb * n[1];
  w[3] = vnb * n[2];
  w[4] = pb;
}

// kernels/euler.py:361-372
template <class C>
void outlet_state(const C* wi, const C* data, const C* n, C g, C* w) {
  const C vi[3] = {wi[1], wi[2], wi[3]};
  const C ci = osqrt(g * wi[4] / wi[0]);
  const C pb = data[4];
  const C dp = wi[4] - pb;
  const C rb = wi[0] - dp / (ci * ci);
  if (dot3(vi, n) >= ci) {
    for (int k = 0; k < 5; ++k) w[k] = wi[k];
    return;
  }
  w[0] = rb;
  for (int i = 0; i < 3; ++i) w[1 + i] = vi[i] + dp / (wi[0] * ci) * n[i];
  w[4] = pb;
}

// kernels/euler.py:375-379
template <class C>
void mirror_state(const C* wi, const C* n, C* w) {
  const C vi[3] = {wi[1], wi[2], wi[3]};
  const C vn2 = C(2.0) * dot3(vi, n);
  w[0] = wi[0];
  for (int i = 0; i < 3; ++i) w[1 + i] = vi[i] - vn2 * n[i];
  w[4] = wi[4];
}

// kernels/euler.py:382-388
template <class C>
void wall_state(const C* wg, const C* data, C rgas, C pr, C* w) {
  w[0] = data[0] > C(0.0) ? (wg[4] + pr) / (rgas * data[0]) : wg[0];
  w[1] = C(0.0);
  w[2] = C(0.0);
  w[3] = C(0.0);
  w[4] = wg[4];
}

// kernels/euler.py:391-422: codes compared at the midpoints in the compute type.
template <class C>
void boundary_value(const C* wg, C kind, const C* data, const C* n, C g, C rgas, C pr, C* out) {
  const C wi[5] = {wg[0], wg[1], wg[2], wg[3], wg[4] + pr};
  C shift = pr;
  C w[5];
  if (kind < C(0.5)) {
    for (int k = 0; k < 5; ++k) w[k] = data[k];
  } else if (kind < C(1.5)) {
    farfield_state(wi, data, n, g, w);
  } else if (kind < C(2.5)) {
    mirror_state(wg, n, w);
    shift = C(0.0) * pr;
  } else if (kind < C(3.5)) {
    inlet_state(wi, data, n, g, rgas, w);
  } else if (kind < C(4.5)) {
    outlet_state(wi, data, n, g, w);
  } else {
    wall_state(wg, data, rgas, pr, w);
    shift = C(0.0) * pr;
  }
  for (int k = 0; k < 4; ++k) out[k] = w[k];
  out[4] = w[4] - shift;
}

// ------------------------------------------------------------------------------ reconstruction
// kernels/euler.py:619-627; gw is the 5 x 3 gradient (row-major), r = rel(x_f, x_c).
template <class C>
void reconstruct(const C* w, const C* gw, const C* lim, const C* r, C pref, C* wf) {
  C out[5];
  for (int i = 0; i < 5; ++i) {
    const C d = gw[3 * i] * r[0] + gw[3 * i + 1] * r[1] + gw[3 * i + 2] * r[2];
    out[i] = w[i] + lim[i] * d;
  }
  const bool ok = out[0] > C(0.0) && out[4] + pref > C(0.0);
  for (int i = 0; i < 5; ++i) wf[i] = ok ? out[i] : w[i];
}

// kernels/euler.py:635-650
template <class C>
void flux_increment(const C* w, const C* dw, const C* e, C g, C pref, C* out) {
  const C v[3] = {w[1], w[2], w[3]}, dv[3] = {dw[1], dw[2], dw[3]};
  const C ve = dot3(v, e);
  const C dve = dot3(dv, e);
  const C rh = g / (g - C(1.0)) * (w[4] + pref) + C(0.5) * w[0] * dot3(v, v);
  const C drh = g / (g - C(1.0)) * dw[4] + C(0.5) * dw[0] * dot3(v, v) + w[0] * dot3(v, dv);
  const C dm = dw[0] * ve + w[0] * dve;
  out[0] = dm;
  for (int i = 0; i < 3; ++i) out[1 + i] = dm * v[i] + (w[0] * ve) * dv[i] + dw[4] * e[i];
  out[4] = drh * ve + rh * dve;
}

// kernels/euler.py:653-677; m is the 3 x 3 face moment (row-major, m[3 u + v]).
template <class C>
void warp(const C* w, const C* gl, const C* pl, const C* gr, const C* pr, const C* m, C g, C pref, C* out) {
  C d[3][5];
  for (int j = 0; j < 3; ++j) {
    const C col[3] = {m[j], m[3 + j], m[6 + j]};
    for (int i = 0; i < 5; ++i) {
      const C a = gl[3 * i] * col[0] + gl[3 * i + 1] * col[1] + gl[3 * i + 2] * col[2];
      const C b = gr[3 * i] * col[0] + gr[3 * i + 1] * col[1] + gr[3 * i + 2] * col[2];
      d[j][i] = C(0.5) * (pl[i] * a + pr[i] * b);
    }
  }
  static const C ex[3] = {C(1.0), C(0.0), C(0.0)}, ey[3] = {C(0.0), C(1.0), C(0.0)}, ez[3] = {C(0.0), C(0.0), C(1.0)};
  C f0[5], f1[5], f2[5];
  flux_increment(w, d[0], ex, g, pref, f0);
  flux_increment(w, d[1], ey, g, pref, f1);
  flux_increment(w, d[2], ez, g, pref, f2);
  for (int k = 0; k < 5; ++k) out[k] = f0[k] + f1[k] + f2[k];
}

// kernels/euler.py:680-682
template <class C>
void wall_warp(const C* gl, const C* pl, const C* m, C* out) {
  const C g4[3] = {gl[12], gl[13], gl[14]};
  out[0] = C(0.0);
  for (int i = 0; i < 3; ++i) out[1 + i] = pl[4] * (m[i] * g4[0] + m[3 + i] * g4[1] + m[6 + i] * g4[2]);
  out[4] = C(0.0);
}

template <class C>
void numerical_flux(int flux, const C* wl, const C* wr, const C* n, C g, C minf, C pref, C* out) {
  if (flux == 0) {
    roe(wl, wr, n, g, pref, out);
  } else if (flux == 1) {
    hllc(wl, wr, n, g, pref, out);
  } else {
    ausm_up(wl, wr, n, g, minf, pref, out);
  }
}

// ------------------------------------------------------------------------------ Jacobians
// The approximate Jacobian's blocks of the reference's implicit solvers (kernels/euler.py:434-440,
// 854-860, 1051-1057): the same bodies evaluated on dual numbers, a value and its derivatives
// along N directions, each operation applying its chain rule; a select takes the derivative of the
// branch its value takes.
template <class C, int N>
struct Dual {
  C v;
  C d[N];
  Dual() = default;
  template <class T, std::enable_if_t<std::is_arithmetic_v<T>, int> = 0>
  Dual(T x) : v(static_cast<C>(x)) {
    for (int i = 0; i < N; ++i) d[i] = C(0);
  }
  friend Dual operator+(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v + b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
  }
  friend Dual operator-(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v - b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
  }
  friend Dual operator-(const Dual& a) {
    Dual r;
    r.v = -a.v;
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
  }
  friend Dual operator*(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v * b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
  }
  friend Dual operator/(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v / b.v;
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
  }
  friend bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
  friend bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
  friend bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
  friend bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
};

template <class C, int N>
Dual<C, N> osqrt(const Dual<C, N>& a) {
  Dual<C, N> r;
  r.v = osqrt(a.v);
  const C h = C(0.5) / r.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * h;
  return r;
}

template <class C, int N>
Dual<C, N> oexp(const Dual<C, N>& a) {
  Dual<C, N> r;
  r.v = oexp(a.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * r.v;
  return r;
}

template <class C, int N>
Dual<C, N> olog(const Dual<C, N>& a) {
  Dual<C, N> r;
  r.v = olog(a.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / a.v;
  return r;
}

// dF/dW_L then dF/dW_R of one face (50 values, entry (k, i, j) at 25 k + 5 i + j), not area-weighted.
template <class C>
void face_jacobian(int flux, const C* wl, const C* wr, const C* n, C g, C minf, C pref, C* out) {
  using D = Dual<C, 5>;
  const D nd[3] = {D(n[0]), D(n[1]), D(n[2])};
  for (int side = 0; side < 2; ++side) {
    D a[5], b[5], o[5];
    for (int k = 0; k < 5; ++k) {
      a[k] = D(wl[k]);
      b[k] = D(wr[k]);
      (side == 0 ? a : b)[k].d[k] = C(1);
    }
    numerical_flux<D>(flux, a, b, nd, D(g), D(minf), D(pref), o);
    for (int i = 0; i < 5; ++i) {
      for (int j = 0; j < 5; ++j) out[25 * side + 5 * i + j] = o[i].d[j];
    }
  }
}

// dF/dW_f of one boundary face (25 values): slip and wall faces the pressure flux's, the other
// kinds the flux against boundary_value of W_f, through it (kernels/euler.py:868-940).
template <class C>
void boundary_jacobian(int flux, const C* wf, C kind, const C* data, const C* n, C g, C rgas, C minf, C pref, C* out) {
  for (int q = 0; q < 25; ++q) out[q] = C(0);
  if ((kind > C(1.5) && kind < C(2.5)) || kind > C(4.5)) {
    for (int k = 0; k < 3; ++k) out[5 * (1 + k) + 4] = n[k];
    return;
  }
  using D = Dual<C, 5>;
  const D nd[3] = {D(n[0]), D(n[1]), D(n[2])};
  D a[5], dd[5], wb[5], o[5];
  for (int k = 0; k < 5; ++k) {
    a[k] = D(wf[k]);
    a[k].d[k] = C(1);
    dd[k] = D(data[k]);
  }
  boundary_value<D>(a, D(kind), dd, nd, D(g), D(rgas), D(pref), wb);
  numerical_flux<D>(flux, a, wb, nd, D(g), D(minf), D(pref), o);
  for (int i = 0; i < 5; ++i) {
    for (int j = 0; j < 5; ++j) out[5 * i + j] = o[i].d[j];
  }
}

// dW/dq of one cell (25 values, entry (i, j) at 5 i + j), q the stored state.
template <class C>
void dwdu(const C* dq, const C* ur, C g, C* out) {
  using D = Dual<C, 5>;
  D q[5], r[5], w[5];
  for (int k = 0; k < 5; ++k) {
    q[k] = D(dq[k]);
    q[k].d[k] = C(1);
    r[k] = D(ur[k]);
  }
  to_primitive<D>(q, r, D(g), w);
  for (int i = 0; i < 5; ++i) {
    for (int j = 0; j < 5; ++j) out[5 * i + j] = w[i].d[j];
  }
}

// The warped-face term's derivative along the five components of each side's reconstructed state
// (kernels/euler.py:858-860 differentiates the whole face body, whose warp term reads
// W = (W_L + W_R) / 2 with the gradients and limiter held): 50 values, entry (k, i, j) at
// 25 k + 5 i + j, to be added to the area-weighted flux derivatives.
template <class C>
void warp_jacobian(const C* fl, const C* fr, const C* gl, const C* pl, const C* gr, const C* pr, const C* m, C g,
                   C pref, C* out) {
  using D = Dual<C, 5>;
  D gld[15], grd[15], pld[5], prd[5], md[9];
  for (int q = 0; q < 15; ++q) {
    gld[q] = D(gl[q]);
    grd[q] = D(gr[q]);
  }
  for (int q = 0; q < 5; ++q) {
    pld[q] = D(pl[q]);
    prd[q] = D(pr[q]);
  }
  for (int q = 0; q < 9; ++q) md[q] = D(m[q]);
  for (int side = 0; side < 2; ++side) {
    D a[5], b[5], w[5], o[5];
    for (int k = 0; k < 5; ++k) {
      a[k] = D(fl[k]);
      b[k] = D(fr[k]);
      (side == 0 ? a : b)[k].d[k] = C(1);
      w[k] = D(0.5) * (a[k] + b[k]);
    }
    warp<D>(w, gld, pld, grd, prd, md, D(g), D(pref), o);
    for (int i = 0; i < 5; ++i) {
      for (int j = 0; j < 5; ++j) out[25 * side + 5 * i + j] = o[i].d[j];
    }
  }
}

// Threads of the gather loops: 1 is the reference's serial CPU path; more runs every per-cell
// (and per-boundary-face) loop of the gather variant in parallel. Each cell's arithmetic and order
// stay those of the serial loop, so the results do not depend on the count.
inline int g_threads = 1;

// ------------------------------------------------------------------------------ exact sums
// runtime/src/exact_sum.cpp:22-103: 68 limbs of 32 bits, limb i weighs 2^(32 i - 1088).
class ExactSum {
 public:
  static constexpr int limbs = 68;
  void add(double term) {
    std::uint64_t bits;
    std::memcpy(&bits, &term, 8);
    const int biased = static_cast<int>((bits >> 52) & 0x7ff);
    const std::uint64_t fraction = bits & ((std::uint64_t{1} << 52) - 1);
    const bool negative = (bits >> 63) != 0;
    if (biased == 0x7ff) {
      ++w_[fraction != 0 ? limbs + 2 : (negative ? limbs + 1 : limbs)];
      return;
    }
    if (biased == 0 && fraction == 0) return;
    const std::uint64_t m = biased == 0 ? fraction : (fraction | (std::uint64_t{1} << 52));
    const int e = biased == 0 ? -1074 : biased - 1075;
    const int p = e + 1088;
    const int i = p >> 5, s = p & 31;
    const std::uint64_t lo = (m & 0xffffffffULL) << s, hi = (m >> 32) << s;
    const std::int64_t sign = negative ? -1 : 1;
    w_[i] += sign * static_cast<std::int64_t>(lo & 0xffffffffULL);
    w_[i + 1] += sign * static_cast<std::int64_t>((lo >> 32) + (hi & 0xffffffffULL));
    w_[i + 2] += sign * static_cast<std::int64_t>(hi >> 32);
    if (++pending_ == (std::int64_t{1} << 29)) normalize();
  }
  void normalize() {
    for (int i = 0; i + 1 < limbs; ++i) {
      const std::int64_t carry = w_[i] >> 32;
      w_[i] -= carry * (std::int64_t{1} << 32);
      w_[i + 1] += carry;
    }
    pending_ = 0;
  }
  [[nodiscard]] double value() const {
    constexpr double inf = std::numeric_limits<double>::infinity();
    if (w_[limbs + 2] > 0 || (w_[limbs] > 0 && w_[limbs + 1] > 0)) return std::numeric_limits<double>::quiet_NaN();
    if (w_[limbs] > 0) return inf;
    if (w_[limbs + 1] > 0) return -inf;
    ExactSum a = *this;
    a.normalize();
    auto& w = a.w_;
    const bool negative = w[limbs - 1] < 0;
    if (negative) {
      for (int i = 0; i < limbs; ++i) w[i] = -w[i];
      a.normalize();
    }
    int t = limbs - 1;
    while (t >= 0 && w[t] == 0) --t;
    if (t < 0) return 0.0;
    auto limb = [&](int k) { return k < 0 ? std::uint64_t{0} : static_cast<std::uint64_t>(w[k]); };
    const std::uint64_t top = limb(t);
    if (top > 0xffffffffULL) return negative ? -inf : inf;
    int b = 0;
    while ((top >> b) != 0) ++b;
    const std::uint64_t below = (limb(t - 1) << 32) | limb(t - 2);
    const std::uint64_t lead = (top << (64 - b)) | (below >> b);
    bool sticky = (below & ((std::uint64_t{1} << b) - 1)) != 0;
    for (int k = t - 3; k >= 0 && !sticky; --k) sticky = limb(k) != 0;
    std::uint64_t mantissa = lead >> 11;
    const std::uint64_t rest = lead & 0x7ff;
    if (rest > 0x400 || (rest == 0x400 && (sticky || (mantissa & 1) != 0))) ++mantissa;
    int exponent = 32 * (t - 2) - 1088 + b + 11;
    if (mantissa == (std::uint64_t{1} << 53)) {
      mantissa >>= 1;
      ++exponent;
    }
    const double magnitude = std::ldexp(static_cast<double>(mantissa), exponent);
    return negative ? -magnitude : magnitude;
  }

 private:
  std::int64_t w_[limbs + 3] = {};
  std::int64_t pending_ = 0;
};

// ------------------------------------------------------------------------------ the solver
struct Scheme {
  int flux;          // 0 roe, 1 hllc, 2 ausm+up
  int second_order;  // linear reconstruction
  int limiter;       // Venkatakrishnan
  int planar;        // the *_planar loops
  double gamma, gas_constant, mach_ref;
  double venkat_k, length;
  double scale[5];
  double reference[5];
};

struct Inputs {
  index_t n_cells, n_faces, n_interior;
  const index_t* owner;
  const index_t* neighbour;
  const offset_t* cell_face_offsets;
  const index_t* cell_faces;
  const double* face_normal;  // 3 x nf
  const double* face_area;    // nf
  const double* r_left;       // 3 x nf
  const double* r_right;      // 3 x nf
  const double* face_moment;  // 9 x nf (may be null when planar)
  const double* cell_volume;  // nc
  const offset_t* stencil_offsets;
  const index_t* stencil;
  const double* stencil_weights;  // 3 x entries
  const double* bc_kind;          // nb
  const double* bc;               // 5 x nb
  const double* source;           // 5 x nc (may be null)
};

class Base {
 public:
  virtual ~Base() = default;
  virtual void residual(const double* u, double* res, bool gather) = 0;
  virtual void spectral_radius(double* lam) = 0;
  virtual void face_fluxes(double* flux) = 0;
  virtual void arrays(double* prim, double* grad, double* phi, double* wmin, double* wmax, double* prim_b) = 0;
  virtual int step(double* u, double cfl, int has_dt, double dt, int local, int gather, double* residuals,
                   double* time_step, index_t* invalid) = 0;
  virtual void face_force(const double* u, const index_t* faces, std::size_t n, double* out) = 0;
  virtual void conservation(const double* u, double* out) = 0;
  virtual int first_invalid(const double* u, const double* res) = 0;
  virtual void set_second_order(int on) = 0;
  virtual void residual_and_jacobian(const double* u, double* res, double* dwdu, double* fj, double* bj) = 0;
  virtual void boundary_gradient_term(const double* v, const double* dwdu, const double* bj, double* y) const = 0;
};

template <class P>
class Solver final : public Base {
  using S = typename P::state;
  using G = typename P::grad;
  using V = typename P::geom_value;
  using C = typename P::compute;

 public:
  Solver(const Scheme& s, const Inputs& in) : sc_(s) {
    nc_ = static_cast<std::size_t>(in.n_cells);
    nf_ = static_cast<std::size_t>(in.n_faces);
    ni_ = static_cast<std::size_t>(in.n_interior);
    owner_.assign(in.owner, in.owner + nf_);
    neighbour_.assign(in.neighbour, in.neighbour + ni_);
    cfo_.assign(in.cell_face_offsets, in.cell_face_offsets + nc_ + 1);
    cf_.assign(in.cell_faces, in.cell_faces + cfo_[nc_]);
    normal_ = conv<V>(in.face_normal, 3 * nf_);
    area_ = conv<V>(in.face_area, nf_);
    rl_.assign(in.r_left, in.r_left + 3 * nf_);
    rr_.assign(in.r_right, in.r_right + 3 * nf_);
    if (in.face_moment != nullptr) moment_ = conv<V>(in.face_moment, 9 * nf_);
    volume_ = conv<V>(in.cell_volume, nc_);
    volume64_.assign(in.cell_volume, in.cell_volume + nc_);
    so_.assign(in.stencil_offsets, in.stencil_offsets + nc_ + 1);
    const auto ns = static_cast<std::size_t>(so_[nc_]);
    ns_ = ns;
    st_.assign(in.stencil, in.stencil + ns);
    weight_ = conv<V>(in.stencil_weights, 3 * ns);
    const std::size_t nb = nf_ - ni_;
    bc_kind_.assign(in.bc_kind, in.bc_kind + nb);
    bc_ = conv<S>(in.bc, 5 * nb);
    source_ = in.source != nullptr ? conv<S>(in.source, 5 * nc_) : std::vector<S>(5 * nc_, S(0));
    // physics/src/euler.cpp:229-245
    const double* r = s.reference;
    p_ref_ = (s.gamma - 1.0) * (r[4] - 0.5 * (r[1] * r[1] + r[2] * r[2] + r[3] * r[3]) / r[0]);
    const double k3 = std::pow(s.venkat_k / s.length, 3);
    for (int k = 0; k < 5; ++k) limiter_scale_[k] = k3 * s.scale[k] * s.scale[k];
    state_.assign(5 * nc_, S(0));
    prim_.assign(5 * nc_, S(0));
    prim_b_.assign(5 * nf_, S(0));
    grad_.assign(15 * nc_, G(0));
    wmin_.assign(5 * nc_, S(0));
    wmax_.assign(5 * nc_, S(0));
    phi_.assign(5 * nc_, S(sc_.second_order && !sc_.limiter ? 1 : 0));
    res_.assign(5 * nc_, 0.0);
  }

  void set_second_order(int on) override {
    sc_.second_order = on;
    if (!on) {
      std::fill(grad_.begin(), grad_.end(), G(0));
      std::fill(phi_.begin(), phi_.end(), S(0));
    } else if (!sc_.limiter) {
      std::fill(phi_.begin(), phi_.end(), S(1));
    }
  }

  // Euler::state then Euler::residual (physics/src/euler.cpp:925-941).
  void residual(const double* u, double* res, bool gather) override {
    for (std::size_t k = 0; k < 5; ++k) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
      for (std::size_t c = 0; c < nc_; ++c) state_[k * nc_ + c] = S(u[k * nc_ + c] - sc_.reference[k]);
    }
    cell_loops();
    faces(gather);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (std::size_t i = 0; i < 5 * nc_; ++i) res[i] = res_[i];
  }

  // kernels/euler.py:1091-1112 over the incidences of each cell, summed in FP64, stored as State.
  void spectral_radius(double* lam) override {
    const C g = C(sc_.gamma), pref = C(p_ref_);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (std::size_t c = 0; c < nc_; ++c) {
      A acc = 0;
      for (auto k = cfo_[c]; k < cfo_[c + 1]; ++k) {
        const auto f = static_cast<std::size_t>(cf_[static_cast<std::size_t>(k)]);
        const bool across = f < ni_;
        const std::size_t nb = across ? (static_cast<std::size_t>(owner_[f]) == c ? neighbour_[f] : owner_[f]) : 0;
        C wc[5], wn[5], n[3];
        for (int q = 0; q < 5; ++q) {
          wc[q] = C(prim_[q * nc_ + c]);
          wn[q] = across ? C(prim_[q * nc_ + nb]) : C(prim_b_[q * nf_ + f]);
        }
        for (int q = 0; q < 3; ++q) n[q] = C(normal_[q * nf_ + f]);
        const C vc[3] = {wc[1], wc[2], wc[3]}, vn[3] = {wn[1], wn[2], wn[3]};
        const C lc = sabs(dot3(vc, n)) + std::sqrt(g * (wc[4] + pref) / wc[0]);
        const C ln = sabs(dot3(vn, n)) + std::sqrt(g * (wn[4] + pref) / wn[0]);
        acc += A(smax(lc, ln) * C(area_[f]));
      }
      lam[c] = static_cast<double>(S(acc));
    }
  }

  // The flux of every interior face in its orientation and of every boundary face, from the
  // bodies of the residual's face loops (kernels/euler.py:1062-1085), after a residual.
  void face_fluxes(double* flux) override {
    for (std::size_t f = 0; f < nf_; ++f) {
      C out[5];
      if (f < ni_) {
        face_body(f, out);
      } else {
        boundary_body(f, out);
      }
      for (int k = 0; k < 5; ++k) flux[k * nf_ + f] = A(out[k]);
    }
  }

  // EulerPhysics::check_conservation (physics/src/euler_physics.cpp:1750-1784) on one rank: R^F is
  // the residual of u without its source, the fluxes those of the store loops; per equation k
  // out[k] E, out[5 + k] D, out[10 + k] boundary, out[15 + k] faces, each an exact sum of its terms.
  void conservation(const double* u, double* out) override {
    std::vector<S> saved(5 * nc_, S(0));
    std::swap(saved, source_);
    std::vector<double> res(5 * nc_), flux(5 * nf_);
    residual(u, res.data(), true);
    std::swap(saved, source_);
    face_fluxes(flux.data());
    for (std::size_t k = 0; k < 5; ++k) {
      ExactSum e, d, b, faces;
      for (std::size_t c = 0; c < nc_; ++c) {
        e.add(res[k * nc_ + c]);
        faces.add(res[k * nc_ + c]);
      }
      for (std::size_t f = ni_; f < nf_; ++f) {
        const double v = flux[k * nf_ + f];
        e.add(-v);
        d.add(std::abs(v));
        b.add(v);
      }
      for (std::size_t f = 0; f < ni_; ++f) {
        d.add(std::abs(flux[k * nf_ + f]));
        d.add(std::abs(flux[k * nf_ + f]));
      }
      out[k] = e.value();
      out[5 + k] = d.value();
      out[10 + k] = b.value();
      out[15 + k] = faces.value();
    }
  }

  void arrays(double* prim, double* grad, double* phi, double* wmin, double* wmax, double* prim_b) override {
    for (std::size_t i = 0; i < 5 * nc_; ++i) {
      prim[i] = prim_[i];
      phi[i] = phi_[i];
      wmin[i] = wmin_[i];
      wmax[i] = wmax_[i];
    }
    for (std::size_t i = 0; i < 15 * nc_; ++i) grad[i] = grad_[i];
    for (std::size_t i = 0; i < 5 * nf_; ++i) prim_b[i] = prim_b_[i];
  }

  // Euler::boundary_pressure (physics/src/euler.cpp:1187-1223) after the residual of u: the face
  // pressure of the boundary flux's reconstruction and the force p_f S + phi_p M^T grad p, in FP64
  // from the stored arrays, on the listed boundary faces; out 3 x n.
  void face_force(const double* u, const index_t* faces, std::size_t n, double* out) override {
    std::vector<double> res(5 * nc_);
    residual(u, res.data(), true);
    for (std::size_t i = 0; i < n; ++i) {
      const auto f = static_cast<std::size_t>(faces[i]);
      const auto c = static_cast<std::size_t>(owner_[f]);
      auto g = [&](std::size_t a, std::size_t j) { return static_cast<double>(grad_[(3 * a + j) * nc_ + c]); };
      double drho = 0.0, dp = 0.0;
      for (std::size_t j = 0; j < 3; ++j) {
        const double r = rl_[j * nf_ + f];
        drho += g(0, j) * r;
        dp += g(4, j) * r;
      }
      const double rho_c = static_cast<double>(prim_[c]);
      const double p_c = p_ref_ + static_cast<double>(prim_[4 * nc_ + c]);
      const double rho_f = rho_c + static_cast<double>(phi_[c]) * drho;
      double p_f = p_c + static_cast<double>(phi_[4 * nc_ + c]) * dp;
      if (!(rho_f > 0.0 && p_f > 0.0)) p_f = p_c;
      const double phi_p = static_cast<double>(phi_[4 * nc_ + c]);
      for (std::size_t k = 0; k < 3; ++k) {
        double warp = 0.0;
        for (std::size_t j = 0; j < 3; ++j) {
          const double mjk = moment_.empty() ? 0.0 : static_cast<double>(moment_[(3 * j + k) * nf_ + f]);
          warp += mjk * g(4, j);
        }
        const double area_k = static_cast<double>(normal_[k * nf_ + f]) * static_cast<double>(area_[f]);
        out[k * n + i] = p_f * area_k + phi_p * warp;
      }
    }
  }

  // EulerPhysics::first_invalid_cell (physics/src/euler_physics.cpp:1399-1433): the first cell
  // whose stored state, read as the kernels read it, has a non-finite or non-positive density or
  // pressure; else the first with a non-finite residual; -1 for none.
  int first_invalid(const double* u, const double* res) override {
    const double* r = sc_.reference;
    const C g = static_cast<C>(sc_.gamma);
    const C r0 = static_cast<C>(r[0]);
    const C mr[3] = {static_cast<C>(r[1]), static_cast<C>(r[2]), static_cast<C>(r[3])};
    const C p_ref = static_cast<C>(p_ref_);
    for (std::size_t c = 0; c < nc_; ++c) {
      C dq[5];
      for (std::size_t k = 0; k < 5; ++k) dq[k] = static_cast<C>(S(u[k * nc_ + c] - r[k]));
      const C rho = r0 + dq[0];
      C dm_m = 0, mr_mr = 0;
      for (std::size_t k = 0; k < 3; ++k) {
        dm_m += dq[k + 1] * (C(2) * mr[k] + dq[k + 1]);
        mr_mr += mr[k] * mr[k];
      }
      const C dk = (dm_m * r0 - mr_mr * dq[0]) / (rho * r0);
      const C p = (g - C(1)) * (dq[4] - C(0.5) * dk) + p_ref;
      if (!std::isfinite(static_cast<double>(rho)) || !std::isfinite(static_cast<double>(p)) || !(rho > C(0)) ||
          !(p > C(0))) {
        return static_cast<int>(c);
      }
    }
    if (res != nullptr) {
      for (std::size_t c = 0; c < nc_; ++c) {
        for (std::size_t k = 0; k < 5; ++k) {
          if (!std::isfinite(res[k * nc_ + c])) return static_cast<int>(c);
        }
      }
    }
    return -1;
  }

  // EulerPhysics::unsteady_step with CudaStepper::step's options: residual norms of R(u), the
  // global (min CFL V / lambda) or given time step, or the local one, then three SSP-RK3 stages
  // (physics/src/euler_physics.cpp:1362-1393; physics/src/device_stepper.cuh:313-345).
  int step(double* u, double cfl, int has_dt, double dt, int local, int gather, double* residuals,
           double* time_step, index_t* invalid) override {
    std::vector<double> res(5 * nc_), stage(u, u + 5 * nc_), lam(nc_), dtc;
    residual(u, res.data(), gather != 0);
    norms(res.data(), residuals);
    double h = has_dt ? dt : 0.0;
    if (!has_dt) {
      spectral_radius(lam.data());
      h = std::numeric_limits<double>::infinity();
      dtc.resize(nc_);
#pragma omp parallel for schedule(static) num_threads(g_threads)
      for (std::size_t c = 0; c < nc_; ++c) dtc[c] = cfl * volume64_[c] / static_cast<double>(S(lam[c]));
      for (std::size_t c = 0; c < nc_; ++c) h = std::min(h, dtc[c]);
    }
    const double weights[3][2] = {{0.0, 1.0}, {0.75, 0.25}, {1.0 / 3.0, 2.0 / 3.0}};
    for (int s = 0; s < 3; ++s) {
      if (s > 0) residual(stage.data(), res.data(), gather != 0);
      for (std::size_t k = 0; k < 5; ++k) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
        for (std::size_t c = 0; c < nc_; ++c) {
          const std::size_t i = k * nc_ + c;
          const double hc = (!has_dt && local) ? dtc[c] : h;
          const double next = stage[i] - hc / volume64_[c] * res[i];
          stage[i] = weights[s][0] * u[i] + weights[s][1] * next;
        }
      }
    }
    for (std::size_t i = 0; i < 5 * nc_; ++i) u[i] = stage[i];
    *time_step = h;
    *invalid = first_invalid(u, res.data());
    return 0;
  }

  // EulerPhysics::record_norms: sqrt(ExactSum(R_k^2) / N).
  void norms(const double* res, double* out) const {
    for (std::size_t k = 0; k < 5; ++k) {
      ExactSum sum, count;
      for (std::size_t c = 0; c < nc_; ++c) sum.add(res[k * nc_ + c] * res[k * nc_ + c]);
      count.add(static_cast<double>(nc_));
      out[k] = std::sqrt(sum.value() / count.value());
    }
  }

 private:
  template <class T>
  static std::vector<T> conv(const double* p, std::size_t n) {
    std::vector<T> out(n);
    for (std::size_t i = 0; i < n; ++i) out[i] = static_cast<T>(p[i]);
    return out;
  }

  void load_prim(std::size_t c, C* w) const {
    for (int q = 0; q < 5; ++q) w[q] = C(prim_[q * nc_ + c]);
  }
  void load_grad(std::size_t c, C* g) const {
    for (int q = 0; q < 15; ++q) g[q] = C(grad_[q * nc_ + c]);
  }
  void load_phi(std::size_t c, C* p) const {
    for (int q = 0; q < 5; ++q) p[q] = C(phi_[q * nc_ + c]);
  }
  void load_normal(std::size_t f, C* n) const {
    for (int q = 0; q < 3; ++q) n[q] = C(normal_[q * nf_ + f]);
  }
  void load_moment(std::size_t f, C* m) const {
    for (int q = 0; q < 9; ++q) m[q] = moment_.empty() ? C(0) : C(moment_[q * nf_ + f]);
  }
  // rel(x_f, x_c) of the owner (left) or neighbour (right) of f: the FP64 difference, rounded.
  void rel(std::size_t f, bool left, C* r) const {
    const std::vector<double>& v = left ? rl_ : rr_;
    for (int q = 0; q < 3; ++q) r[q] = C(v[q * nf_ + f]);
  }

  // Euler::residual_and_jacobian (physics/src/euler.cpp:983-991): the residual (gather loops),
  // then the approximate Jacobian's parts in the device's layout, widened to FP64: dW/dq of every
  // cell at (5 i + j) nc + c; per interior face f dF/dW_L at (5 i + j) ni + f and dF/dW_R at
  // (25 + 5 i + j) ni + f (the reconstructed states of the face body, gradients and limiter held,
  // area-weighted, plus the warped-face term's derivative on non-planar meshes); per boundary face
  // b the derivative of its flux (the pressure flux on slip and wall faces) at (5 i + j) nb + b.
  void residual_and_jacobian(const double* u, double* res, double* dwdu_out, double* fj, double* bj) override {
    residual(u, res, true);
    const C g = C(sc_.gamma), pref = C(p_ref_), minf = C(sc_.mach_ref), rgas = C(sc_.gas_constant);
    C ur[5];
    for (int k = 0; k < 5; ++k) ur[k] = C(sc_.reference[k]);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (std::size_t c = 0; c < nc_; ++c) {
      C dq[5], o[25];
      for (int k = 0; k < 5; ++k) dq[k] = C(state_[k * nc_ + c]);
      dwdu<C>(dq, ur, g, o);
      for (int q = 0; q < 25; ++q) dwdu_out[q * nc_ + c] = static_cast<double>(o[q]);
    }
    const bool second = sc_.second_order != 0;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (std::size_t f = 0; f < ni_; ++f) {
      const auto l = static_cast<std::size_t>(owner_[f]), r = static_cast<std::size_t>(neighbour_[f]);
      C wl[5], gl[15], pl[5], wr[5], gr[15], pr[5], n[3], xl[3], xr[3], fl[5], fr[5], o[50];
      load_prim(l, wl);
      load_prim(r, wr);
      load_grad(l, gl);
      load_grad(r, gr);
      load_phi(l, pl);
      load_phi(r, pr);
      load_normal(f, n);
      if (second) {
        rel(f, true, xl);
        rel(f, false, xr);
        reconstruct(wl, gl, pl, xl, pref, fl);
        reconstruct(wr, gr, pr, xr, pref, fr);
      } else {
        std::copy(wl, wl + 5, fl);
        std::copy(wr, wr + 5, fr);
      }
      face_jacobian<C>(sc_.flux, fl, fr, n, g, minf, pref, o);
      const C area = C(area_[f]);
      C val[50];
      for (int q = 0; q < 50; ++q) val[q] = area * o[q];
      if (!sc_.planar) {
        C m[9], wj[50];
        load_moment(f, m);
        warp_jacobian<C>(fl, fr, gl, pl, gr, pr, m, g, pref, wj);
        for (int q = 0; q < 50; ++q) val[q] = val[q] + wj[q];
      }
      for (int q = 0; q < 50; ++q) fj[q * ni_ + f] = static_cast<double>(val[q]);
    }
    const std::size_t nb = nf_ - ni_;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (std::size_t b = 0; b < nb; ++b) {
      const std::size_t f = ni_ + b;
      const auto l = static_cast<std::size_t>(owner_[f]);
      C wl[5], gl[15], pl[5], n[3], xl[3], wf[5], data[5], o[25];
      load_prim(l, wl);
      load_normal(f, n);
      if (second) {
        load_grad(l, gl);
        load_phi(l, pl);
        rel(f, true, xl);
        reconstruct(wl, gl, pl, xl, pref, wf);
      } else {
        std::copy(wl, wl + 5, wf);
      }
      for (int k = 0; k < 5; ++k) data[k] = C(bc_[k * nb + b]);
      boundary_jacobian<C>(sc_.flux, wf, C(bc_kind_[b]), data, n, g, rgas, minf, pref, o);
      const C area = C(area_[f]);
      for (int q = 0; q < 25; ++q) bj[q * nb + b] = static_cast<double>(area * o[q]);
    }
  }

  // The steady implicit step's Krylov operator adds the boundary faces' derivative through the
  // reconstruction's gradients, which the assembled blocks drop (EulerPhysics::
  // add_boundary_gradient_term with Euler::gradient_tangent, physics/src/euler_physics.cpp:1300-1320,
  // physics/src/euler.cpp:1081-1119), for the state of the last residual_and_jacobian: y += B v.
  // For each cell owning a boundary face: its stencil cells' primitive tangents dW = dW/dq v (the
  // direction rounded to the state type, the products summed in the compute type), the gradient's
  // tangent sum_k (dW_k - dW_c) w_k (FP64 sums, boundary values held), then over its boundary faces
  // in cell_faces order, where the reconstruction did not fall back to the cell value, the face
  // state's tangent phi (grad_tan . r) through the boundary block dF/dW_f, and on warped slip and
  // wall faces the wall-warp term's tangent phi_p M^T grad_tan p.
  void boundary_gradient_term(const double* v, const double* dwdu, const double* bj, double* y) const override {
    if (!sc_.second_order) return;
    const C pref = C(p_ref_);
    const std::size_t nb = nf_ - ni_;
    auto tangent = [&](std::size_t c, C* dw) {
      for (int i = 0; i < 5; ++i) {
        C t = C(0);
        for (int j = 0; j < 5; ++j) t = t + C(dwdu[static_cast<std::size_t>(5 * i + j) * nc_ + c]) * C(S(v[static_cast<std::size_t>(j) * nc_ + c]));
        dw[i] = t;
      }
    };
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (std::size_t c = 0; c < nc_; ++c) {
      bool owns = false;
      for (auto k = cfo_[c]; k < cfo_[c + 1] && !owns; ++k) owns = static_cast<std::size_t>(cf_[static_cast<std::size_t>(k)]) >= ni_;
      if (!owns) continue;
      C dwc[5];
      tangent(c, dwc);
      double acc[15] = {};
      for (auto k = so_[c]; k < so_[c + 1]; ++k) {
        const index_t e = st_[static_cast<std::size_t>(k)];
        C dwn[5] = {C(0), C(0), C(0), C(0), C(0)};
        if (e >= 0) tangent(static_cast<std::size_t>(e), dwn);
        C w[3];
        for (int j = 0; j < 3; ++j) w[j] = C(weight_[static_cast<std::size_t>(j) * ns_ + static_cast<std::size_t>(k)]);
        for (int i = 0; i < 5; ++i) {
          const C d = dwn[i] - dwc[i];
          for (int j = 0; j < 3; ++j) acc[3 * i + j] += static_cast<double>(d * w[j]);
        }
      }
      C gt[15];
      for (int q = 0; q < 15; ++q) gt[q] = C(acc[q]);
      C wc[5], gc[15], pc[5];
      load_prim(c, wc);
      load_grad(c, gc);
      load_phi(c, pc);
      double dres[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      for (auto k = cfo_[c]; k < cfo_[c + 1]; ++k) {
        const auto f = static_cast<std::size_t>(cf_[static_cast<std::size_t>(k)]);
        if (f < ni_) continue;
        const std::size_t b = f - ni_;
        C r[3];
        rel(f, true, r);
        // reconstruct's test: the cell value where the reconstructed density or pressure is not positive.
        const C d0[5] = {gc[0] * r[0] + gc[1] * r[1] + gc[2] * r[2], gc[3] * r[0] + gc[4] * r[1] + gc[5] * r[2],
                         gc[6] * r[0] + gc[7] * r[1] + gc[8] * r[2], gc[9] * r[0] + gc[10] * r[1] + gc[11] * r[2],
                         gc[12] * r[0] + gc[13] * r[1] + gc[14] * r[2]};
        const bool fell_back = !((wc[0] + pc[0] * d0[0]) > C(0.0) && (wc[4] + pc[4] * d0[4]) + pref > C(0.0));
        if (fell_back) continue;
        C dwf[5];
        for (int i = 0; i < 5; ++i) dwf[i] = pc[i] * (gt[3 * i] * r[0] + gt[3 * i + 1] * r[1] + gt[3 * i + 2] * r[2]);
        for (int i = 0; i < 5; ++i) {
          double t = 0.0;
          for (int j = 0; j < 5; ++j) t += bj[static_cast<std::size_t>(5 * i + j) * nb + b] * static_cast<double>(dwf[j]);
          dres[i] += t;
        }
        const C kind = C(bc_kind_[b]);
        if (!sc_.planar && ((kind > C(1.5) && kind < C(2.5)) || kind > C(4.5))) {
          C m[9];
          load_moment(f, m);
          for (int i = 0; i < 3; ++i) {
            dres[1 + i] += static_cast<double>(pc[4] * (m[i] * gt[12] + m[3 + i] * gt[13] + m[6 + i] * gt[14]));
          }
        }
      }
      for (int i = 0; i < 5; ++i) y[static_cast<std::size_t>(i) * nc_ + c] += dres[i];
    }
  }

  // Impl::cell_loops (physics/src/euler.cpp:113-136), on one rank.
  void cell_loops() {
    const C g = C(sc_.gamma), rgas = C(sc_.gas_constant), pref = C(p_ref_);
    // primitives (kernels/euler.py:428-433)
    C ur[5];
    for (int k = 0; k < 5; ++k) ur[k] = C(sc_.reference[k]);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (std::size_t c = 0; c < nc_; ++c) {
      C dq[5], w[5];
      for (int k = 0; k < 5; ++k) dq[k] = C(state_[k * nc_ + c]);
      to_primitive(dq, ur, g, w);
      for (int k = 0; k < 5; ++k) prim_[k * nc_ + c] = S(w[k]);
    }
    // boundary_states (kernels/euler.py:442-466)
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (std::size_t f = ni_; f < nf_; ++f) {
      C wi[5], data[5], n[3], out[5];
      load_prim(static_cast<std::size_t>(owner_[f]), wi);
      for (int k = 0; k < 5; ++k) data[k] = C(bc_[k * (nf_ - ni_) + (f - ni_)]);
      load_normal(f, n);
      boundary_value(wi, C(bc_kind_[f - ni_]), data, n, g, rgas, pref, out);
      for (int k = 0; k < 5; ++k) prim_b_[k * nf_ + f] = S(out[k]);
    }
    if (sc_.second_order) {
      gradients();
      if (sc_.limiter) limiter();
    }
    // sources (kernels/euler.py:611-616): res = -V s
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (std::size_t c = 0; c < nc_; ++c) {
      const C nv = -C(volume_[c]);
      for (int k = 0; k < 5; ++k) res_[k * nc_ + c] = A(nv * C(source_[k * nc_ + c]));
    }
  }

  // gradients (kernels/euler.py:530-549; reduction body emit_cpu.py:126-182).
  void gradients() {
    const C inf = std::numeric_limits<C>::infinity();
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (std::size_t c = 0; c < nc_; ++c) {
      A acc[15];
      C lo[5], hi[5], wc[5];
      for (int q = 0; q < 15; ++q) acc[q] = 0;
      for (int q = 0; q < 5; ++q) {
        lo[q] = inf;
        hi[q] = -inf;
      }
      load_prim(c, wc);
      for (auto k = so_[c]; k < so_[c + 1]; ++k) {
        const index_t e = st_[static_cast<std::size_t>(k)];
        C wn[5], w[3];
        if (e >= 0) {
          load_prim(static_cast<std::size_t>(e), wn);
        } else {
          const auto f = static_cast<std::size_t>(-e - 1);
          for (int q = 0; q < 5; ++q) wn[q] = C(prim_b_[q * nf_ + f]);
        }
        for (int j = 0; j < 3; ++j) w[j] = C(weight_[static_cast<std::size_t>(j) * ns_ + static_cast<std::size_t>(k)]);
        for (int i = 0; i < 5; ++i) {
          const C d = wn[i] - wc[i];
          for (int j = 0; j < 3; ++j) acc[3 * i + j] += A(d * w[j]);
        }
        for (int i = 0; i < 5; ++i) {
          if (wn[i] < lo[i]) lo[i] = wn[i];
          if (wn[i] > hi[i]) hi[i] = wn[i];
        }
      }
      // include_own_value (kernels/euler.py:514-518)
      for (int q = 0; q < 15; ++q) grad_[q * nc_ + c] = G(C(acc[q]));
      for (int i = 0; i < 5; ++i) {
        wmin_[i * nc_ + c] = S(smin(lo[i], wc[i]));
        wmax_[i * nc_ + c] = S(smax(hi[i], wc[i]));
      }
    }
  }

  // limiter (kernels/euler.py:563-608): min over the cell's faces, then min(phi, 1).
  void limiter() {
    const C inf = std::numeric_limits<C>::infinity();
    C scale[5];
    for (int k = 0; k < 5; ++k) scale[k] = C(limiter_scale_[k]);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (std::size_t c = 0; c < nc_; ++c) {
      C wc[5], gc[15], lo[5], hi[5], acc[5];
      load_prim(c, wc);
      load_grad(c, gc);
      for (int i = 0; i < 5; ++i) {
        lo[i] = C(wmin_[i * nc_ + c]);
        hi[i] = C(wmax_[i * nc_ + c]);
        acc[i] = inf;
      }
      const C vol = C(volume_[c]);
      for (auto k = cfo_[c]; k < cfo_[c + 1]; ++k) {
        const auto f = static_cast<std::size_t>(cf_[static_cast<std::size_t>(k)]);
        C r[3];
        rel(f, static_cast<std::size_t>(owner_[f]) == c, r);
        for (int i = 0; i < 5; ++i) {
          const C d = gc[3 * i] * r[0] + gc[3 * i + 1] * r[1] + gc[3 * i + 2] * r[2];
          const C dmax = hi[i] - wc[i], dmin = lo[i] - wc[i], eps2 = scale[i] * vol;
          const C dp = d > C(0.0) ? dmax : dmin;
          const C v = (dp * dp + eps2 + C(2.0) * d * dp) / (dp * dp + C(2.0) * d * d + d * dp + eps2);
          if (v < acc[i]) acc[i] = v;
        }
      }
      for (int i = 0; i < 5; ++i) phi_[i * nc_ + c] = S(smin(acc[i], C(1.0)));
    }
  }

  // {flux}_face[_planar] (kernels/euler.py:687-852) for interior face f.
  void face_body(std::size_t f, C* out) const {
    const C g = C(sc_.gamma), pref = C(p_ref_), minf = C(sc_.mach_ref);
    const auto l = static_cast<std::size_t>(owner_[f]), r = static_cast<std::size_t>(neighbour_[f]);
    C wl[5], gl[15], pl[5], wr[5], gr[15], pr[5], n[3], xl[3], xr[3], fl[5], fr[5];
    load_prim(l, wl);
    load_grad(l, gl);
    load_phi(l, pl);
    load_prim(r, wr);
    load_grad(r, gr);
    load_phi(r, pr);
    load_normal(f, n);
    rel(f, true, xl);
    rel(f, false, xr);
    reconstruct(wl, gl, pl, xl, pref, fl);
    reconstruct(wr, gr, pr, xr, pref, fr);
    C fx[5];
    numerical_flux(sc_.flux, fl, fr, n, g, minf, pref, fx);
    const C area = C(area_[f]);
    for (int k = 0; k < 5; ++k) out[k] = area * fx[k];
    if (!sc_.planar) {
      C m[9], w[5], wp[5];
      load_moment(f, m);
      for (int k = 0; k < 5; ++k) w[k] = C(0.5) * (fl[k] + fr[k]);
      warp(w, gl, pl, gr, pr, m, g, pref, wp);
      for (int k = 0; k < 5; ++k) out[k] = out[k] + wp[k];
    }
  }

  // {flux}_boundary[_planar] (kernels/euler.py:863-1049) for boundary face f.
  void boundary_body(std::size_t f, C* out) const {
    const C g = C(sc_.gamma), pref = C(p_ref_), minf = C(sc_.mach_ref), rgas = C(sc_.gas_constant);
    const auto l = static_cast<std::size_t>(owner_[f]);
    C wl[5], gl[15], pl[5], n[3], xl[3], wf[5], wb[5], data[5];
    load_prim(l, wl);
    load_grad(l, gl);
    load_phi(l, pl);
    load_normal(f, n);
    rel(f, true, xl);
    const C kind = C(bc_kind_[f - ni_]);
    for (int k = 0; k < 5; ++k) data[k] = C(bc_[k * (nf_ - ni_) + (f - ni_)]);
    reconstruct(wl, gl, pl, xl, pref, wf);
    boundary_value(wf, kind, data, n, g, rgas, pref, wb);
    const C area = C(area_[f]);
    const bool slip = kind > C(1.5) && kind < C(2.5);
    const bool wall = !sc_.planar && kind > C(4.5);
    if (slip || wall) {
      // area * wall_flux(wf, n) (+ wall_warp)
      const C wfx[5] = {C(0.0), wf[4] * n[0], wf[4] * n[1], wf[4] * n[2], C(0.0)};
      for (int k = 0; k < 5; ++k) out[k] = area * wfx[k];
      if (!sc_.planar) {
        C m[9], ww[5];
        load_moment(f, m);
        wall_warp(gl, pl, m, ww);
        for (int k = 0; k < 5; ++k) out[k] = out[k] + ww[k];
      }
    } else {
      C fx[5];
      numerical_flux(sc_.flux, wf, wb, n, g, minf, pref, fx);
      for (int k = 0; k < 5; ++k) out[k] = area * fx[k];
    }
  }

  // Impl::faces (physics/src/euler.cpp:649-668): the face loop, then the boundary loop; scatter
  // in face order (emit_cpu.py:435-464) or gather per cell in cell_faces order (emit_cpu.py:506-564).
  void faces(bool gather) {
    if (!gather) {
      for (std::size_t f = 0; f < ni_; ++f) {
        C out[5];
        face_body(f, out);
        const auto l = static_cast<std::size_t>(owner_[f]), r = static_cast<std::size_t>(neighbour_[f]);
        for (int k = 0; k < 5; ++k) res_[k * nc_ + l] += A(out[k]);
        for (int k = 0; k < 5; ++k) res_[k * nc_ + r] -= A(out[k]);
      }
      for (std::size_t f = ni_; f < nf_; ++f) {
        C out[5];
        boundary_body(f, out);
        const auto l = static_cast<std::size_t>(owner_[f]);
        for (int k = 0; k < 5; ++k) res_[k * nc_ + l] += A(out[k]);
      }
      return;
    }
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (std::size_t c = 0; c < nc_; ++c) {
      A acc[5] = {};
      for (auto k = cfo_[c]; k < cfo_[c + 1]; ++k) {
        const auto f = static_cast<std::size_t>(cf_[static_cast<std::size_t>(k)]);
        if (f >= ni_) continue;
        C out[5];
        face_body(f, out);
        const bool left = static_cast<std::size_t>(owner_[f]) == c;
        for (int q = 0; q < 5; ++q) acc[q] += left ? A(out[q]) : -A(out[q]);
      }
      for (int q = 0; q < 5; ++q) res_[q * nc_ + c] += acc[q];
    }
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (std::size_t c = 0; c < nc_; ++c) {
      A acc[5] = {};
      for (auto k = cfo_[c]; k < cfo_[c + 1]; ++k) {
        const auto f = static_cast<std::size_t>(cf_[static_cast<std::size_t>(k)]);
        if (f < ni_) continue;
        C out[5];
        boundary_body(f, out);
        for (int q = 0; q < 5; ++q) acc[q] += A(out[q]);
      }
      for (int q = 0; q < 5; ++q) res_[q * nc_ + c] += acc[q];
    }
  }

  Scheme sc_;
  std::size_t nc_ = 0, nf_ = 0, ni_ = 0, ns_ = 0;
  std::vector<index_t> owner_, neighbour_, cf_, st_;
  std::vector<offset_t> cfo_, so_;
  std::vector<V> normal_, area_, moment_, volume_, weight_;
  std::vector<double> rl_, rr_, volume64_, bc_kind_;
  std::vector<S> bc_, source_;
  double p_ref_ = 0.0, limiter_scale_[5] = {};
  std::vector<S> state_, prim_, prim_b_, wmin_, wmax_, phi_;
  std::vector<G> grad_;
  std::vector<A> res_;
};

// ------------------------------------------------------------------------------ implicit step
// The steady implicit step's linear algebra on one rank, FP64 throughout:
//   Euler::assemble (physics/src/euler.cpp:993-1053): per cell shift_c I + sum over its faces
//     (cell_faces order) of +-J D_c, per interior face A(owner, neighbour) = J_R D_neighbour,
//     A(neighbour, owner) = -J_L D_owner; each product J D summed over k from 0, added as sign * t;
//   BlockMatrix (physics/src/block_matrix.cpp): the 5 x 5 diagonal blocks inverted by Gauss-Jordan
//     elimination with partial pivoting, the greedy distance-1 colouring in cell order, relax,
//     multiply, the symmetric multicolour sweeps;
//   Gcr::solve (solvers/src/gcr.cpp) on the residual scaled by its largest entry, its inner products
//     exact sums over the cells (EulerPhysics::owned_dot, owned_max_abs);
//   EulerPhysics::steady_step's solve and update (physics/src/euler_physics.cpp:1104-1250):
//     shift = lambda / CFL, GCR to 0.1 in at most 10 directions preconditioned by 4 symmetric sweeps
//     from zero, a solve leaving more than 0.9 of its residual rejected, then the admissible fraction
//     of the update (density and pressure kept above half their values, halving down to 1e-6).
namespace implicit {

constexpr int kSweeps = 4;
constexpr int kDirections = 10;
constexpr double kTolerance = 0.1;
constexpr double kFailure = 0.9;
constexpr double kMaxDecrease = 0.5;
constexpr double kMinStep = 1e-6;

struct Topo {
  std::size_t n, ni, nb;
  const index_t* owner;
  const index_t* neighbour;
  const offset_t* cfo;
  const index_t* cf;
};

// out += sign * (j d), each entry's product summed over k from zero.
inline void add_product(double* out, const double* j, const double* d, double sign) {
  for (int r = 0; r < 5; ++r) {
    for (int c = 0; c < 5; ++c) {
      double t = 0.0;
      for (int k = 0; k < 5; ++k) t += j[5 * r + k] * d[5 * k + c];
      out[5 * r + c] += sign * t;
    }
  }
}

// Block k of 25 SoA components over n entities.
inline void block(const double* v, std::size_t n, std::size_t e, int k, double* b) {
  for (int q = 0; q < 25; ++q) b[q] = v[(25 * static_cast<std::size_t>(k) + q) * n + e];
}

inline void assemble(const Topo& t, const double* dwdu, const double* fj, const double* bj, const double* shift,
                     double* diag, double* upper, double* lower) {
  std::fill(diag, diag + 25 * t.n, 0.0);
  std::fill(upper, upper + 25 * t.ni, 0.0);
  std::fill(lower, lower + 25 * t.ni, 0.0);
  for (std::size_t c = 0; c < t.n; ++c) {
    double* dg = diag + 25 * c;
    for (int i = 0; i < 5; ++i) dg[6 * i] += shift[c];
    double dc[25], j[25];
    block(dwdu, t.n, c, 0, dc);
    for (auto k = t.cfo[c]; k < t.cfo[c + 1]; ++k) {
      const auto f = static_cast<std::size_t>(t.cf[k]);
      if (f >= t.ni) {
        block(bj, t.nb, f - t.ni, 0, j);
        add_product(dg, j, dc, 1.0);
      } else if (static_cast<std::size_t>(t.owner[f]) == c) {
        block(fj, t.ni, f, 0, j);
        add_product(dg, j, dc, 1.0);
      } else {
        block(fj, t.ni, f, 1, j);
        add_product(dg, j, dc, -1.0);
      }
    }
  }
  for (std::size_t f = 0; f < t.ni; ++f) {
    double j[25], dm[25];
    block(fj, t.ni, f, 1, j);
    block(dwdu, t.n, static_cast<std::size_t>(t.neighbour[f]), 0, dm);
    add_product(upper + 25 * f, j, dm, 1.0);
    block(fj, t.ni, f, 0, j);
    block(dwdu, t.n, static_cast<std::size_t>(t.owner[f]), 0, dm);
    add_product(lower + 25 * f, j, dm, -1.0);
  }
}

// The inverse of a 5 x 5 row-major block; false when a pivot is zero or not finite.
inline bool invert(const double* a, double* inv) {
  double m[5][10];
  for (int i = 0; i < 5; ++i) {
    for (int j = 0; j < 5; ++j) {
      m[i][j] = a[5 * i + j];
      m[i][5 + j] = i == j ? 1.0 : 0.0;
    }
  }
  for (int k = 0; k < 5; ++k) {
    int p = k;
    for (int i = k + 1; i < 5; ++i) {
      if (std::fabs(m[i][k]) > std::fabs(m[p][k])) p = i;
    }
    if (m[p][k] == 0.0 || !std::isfinite(m[p][k])) return false;
    if (p != k) {
      for (int j = 0; j < 10; ++j) std::swap(m[k][j], m[p][j]);
    }
    const double d = 1.0 / m[k][k];
    for (int j = 0; j < 10; ++j) m[k][j] *= d;
    for (int i = 0; i < 5; ++i) {
      if (i == k) continue;
      const double f = m[i][k];
      for (int j = 0; j < 10; ++j) m[i][j] -= f * m[k][j];
    }
  }
  for (int i = 0; i < 5; ++i) {
    for (int j = 0; j < 5; ++j) inv[5 * i + j] = m[i][5 + j];
  }
  return true;
}

// Each cell's interior faces in ascending face order (the lists BlockMatrix builds), and the
// greedy colour of every cell in cell order (the least colour none of its face neighbours took).
struct Rows {
  std::vector<offset_t> off;
  std::vector<index_t> face;
  std::vector<int> colour;
  int colours = 0;
};

inline Rows rows_of(const Topo& t) {
  Rows r;
  r.off.assign(t.n + 1, 0);
  for (std::size_t f = 0; f < t.ni; ++f) {
    ++r.off[static_cast<std::size_t>(t.owner[f]) + 1];
    ++r.off[static_cast<std::size_t>(t.neighbour[f]) + 1];
  }
  for (std::size_t c = 0; c < t.n; ++c) r.off[c + 1] += r.off[c];
  r.face.resize(static_cast<std::size_t>(r.off[t.n]));
  std::vector<offset_t> fill(r.off.begin(), r.off.end() - 1);
  for (std::size_t f = 0; f < t.ni; ++f) {
    r.face[static_cast<std::size_t>(fill[static_cast<std::size_t>(t.owner[f])]++)] = static_cast<index_t>(f);
    r.face[static_cast<std::size_t>(fill[static_cast<std::size_t>(t.neighbour[f])]++)] = static_cast<index_t>(f);
  }
  r.colour.assign(t.n, -1);
  std::vector<char> seen;
  for (std::size_t c = 0; c < t.n; ++c) {
    seen.assign(static_cast<std::size_t>(r.colours) + 1, 0);
    for (auto k = r.off[c]; k < r.off[c + 1]; ++k) {
      const auto f = static_cast<std::size_t>(r.face[static_cast<std::size_t>(k)]);
      const auto o = static_cast<std::size_t>(static_cast<std::size_t>(t.owner[f]) == c ? t.neighbour[f] : t.owner[f]);
      if (r.colour[o] >= 0) seen[static_cast<std::size_t>(r.colour[o])] = 1;
    }
    int k = 0;
    while (seen[static_cast<std::size_t>(k)] != 0) ++k;
    r.colour[c] = k;
    r.colours = std::max(r.colours, k + 1);
  }
  return r;
}

struct Matrix {
  const Topo* t;
  const Rows* rows;
  const double *diag, *upper, *lower, *inverse;

  // The row of c: the other cell of face f, and its block.
  const double* off(std::size_t c, std::size_t f, std::size_t& other) const {
    const bool owned = static_cast<std::size_t>(t->owner[f]) == c;
    other = static_cast<std::size_t>(owned ? t->neighbour[f] : t->owner[f]);
    return owned ? upper + 25 * f : lower + 25 * f;
  }
  void relax(std::size_t c, const double* b, double* x) const {
    const std::size_t n = t->n;
    double r[5];
    for (int i = 0; i < 5; ++i) r[i] = b[i * n + c];
    for (auto k = rows->off[c]; k < rows->off[c + 1]; ++k) {
      std::size_t o;
      const double* a = off(c, static_cast<std::size_t>(rows->face[static_cast<std::size_t>(k)]), o);
      for (int i = 0; i < 5; ++i) {
        for (int j = 0; j < 5; ++j) r[i] -= a[5 * i + j] * x[j * n + o];
      }
    }
    const double* inv = inverse + 25 * c;
    for (int i = 0; i < 5; ++i) {
      double s = 0.0;
      for (int j = 0; j < 5; ++j) s += inv[5 * i + j] * r[j];
      x[i * n + c] = s;
    }
  }
  void multiply(const double* x, double* y) const {
    const std::size_t n = t->n;
    for (std::size_t c = 0; c < n; ++c) {
      double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      auto add = [&](const double* a, std::size_t col) {
        for (int i = 0; i < 5; ++i) {
          for (int j = 0; j < 5; ++j) r[i] += a[5 * i + j] * x[j * n + col];
        }
      };
      add(diag + 25 * c, c);
      for (auto k = rows->off[c]; k < rows->off[c + 1]; ++k) {
        std::size_t o;
        const double* a = off(c, static_cast<std::size_t>(rows->face[static_cast<std::size_t>(k)]), o);
        add(a, o);
      }
      for (int i = 0; i < 5; ++i) y[i * n + c] = r[i];
    }
  }
  // Symmetric sweeps: the colours in increasing, then decreasing order; within a colour no two
  // cells share a face, so the order of its cells does not matter.
  void gauss_seidel(const double* b, double* x, int sweeps) const {
    for (int s = 0; s < sweeps; ++s) {
      for (int k = 0; k < rows->colours; ++k) {
        for (std::size_t c = 0; c < t->n; ++c)
          if (rows->colour[c] == k) relax(c, b, x);
      }
      for (int k = rows->colours - 1; k >= 0; --k) {
        for (std::size_t c = 0; c < t->n; ++c)
          if (rows->colour[c] == k) relax(c, b, x);
      }
    }
  }
};

inline double exact_dot(const double* u, const double* v, std::size_t m) {
  ExactSum s;
  for (std::size_t i = 0; i < m; ++i) s.add(u[i] * v[i]);
  return s.value();
}

inline double max_abs(const double* u, std::size_t m) {
  double r = 0.0;
  for (std::size_t i = 0; i < m; ++i) {
    const double a = std::fabs(u[i]);
    r = std::isnan(a) ? std::numeric_limits<double>::infinity() : std::max(r, a);
  }
  return r;
}

struct Report {
  int cycles = 0;
  double residual_drop = std::numeric_limits<double>::quiet_NaN();
};

// Gcr::solve with apply = the matrix's product (plus `extra`, when given) and precondition = the
// sweeps from zero.
template <class Extra>
inline Report gcr(const Matrix& a, double* r, double* x, std::size_t m, const Extra& extra) {
  Report rep;
  const double scale = max_abs(r, m);
  if (!std::isfinite(scale)) return rep;
  if (scale == 0.0) {
    rep.residual_drop = 0.0;
    return rep;
  }
  for (std::size_t i = 0; i < m; ++i) r[i] /= scale;
  const double r0 = std::sqrt(exact_dot(r, r, m));
  std::vector<std::vector<double>> z(kDirections), w(kDirections);
  double rr = r0 * r0;
  for (int k = 0; k < kDirections; ++k) {
    z[k].assign(m, 0.0);
    w[k].assign(m, 0.0);
    a.gauss_seidel(r, z[k].data(), kSweeps);
    a.multiply(z[k].data(), w[k].data());
    extra(z[k].data(), w[k].data());
    for (int j = 0; j < k; ++j) {
      const double beta = exact_dot(w[k].data(), w[j].data(), m) / exact_dot(w[j].data(), w[j].data(), m);
      for (std::size_t i = 0; i < m; ++i) {
        z[k][i] -= beta * z[j][i];
        w[k][i] -= beta * w[j][i];
      }
    }
    const double ww = exact_dot(w[k].data(), w[k].data(), m);
    rep.cycles = k + 1;
    if (!(ww > 0.0)) {
      if (!std::isfinite(ww)) rr = ww;
      break;
    }
    const double alpha = exact_dot(r, w[k].data(), m) / ww;
    for (std::size_t i = 0; i < m; ++i) {
      x[i] += scale * alpha * z[k][i];
      r[i] -= alpha * w[k][i];
    }
    rr = exact_dot(r, r, m);
    if (std::sqrt(rr) <= kTolerance * r0) break;
  }
  for (std::size_t i = 0; i < m; ++i) r[i] *= scale;
  rep.residual_drop = std::isfinite(rr) ? std::sqrt(rr) / r0 : std::numeric_limits<double>::quiet_NaN();
  return rep;
}

// EulerPhysics::pressure(c, a) of u + a du.
inline double pressure(const double* u, const double* du, std::size_t n, std::size_t c, double a, double gamma) {
  double v[5];
  for (int k = 0; k < 5; ++k) v[k] = u[k * n + c] + a * du[k * n + c];
  return (gamma - 1.0) * (v[4] - 0.5 * (v[1] * v[1] + v[2] * v[2] + v[3] * v[3]) / v[0]);
}

inline double admissible_step(const double* u, const double* du, std::size_t n, double gamma) {
  double alpha = 1.0;
  for (std::size_t c = 0; c < n; ++c) {
    double a = 1.0;
    const double rho = u[c], drho = du[c];
    if (drho < 0.0) a = std::min(a, kMaxDecrease * rho / -drho);
    const double p = pressure(u, du, n, c, 0.0, gamma);
    while (a >= kMinStep && !(pressure(u, du, n, c, a, gamma) >= (1.0 - kMaxDecrease) * p)) a *= 0.5;
    alpha = std::min(alpha, a);
  }
  return alpha;
}

}  // namespace implicit

}  // namespace oracle

using oracle::Inputs;
using oracle::Scheme;

extern "C" {

// Handle over one policy's solver: 0 = fp64, 1 = fp32+acc64.
int oracle_create(int policy, const Scheme* scheme, const Inputs* in, void** out, char* err, std::size_t err_len) {
  try {
    if (policy == 0) {
      *out = new oracle::Solver<oracle::fp64>(*scheme, *in);
    } else if (policy == 1) {
      *out = new oracle::Solver<oracle::fp32_acc64>(*scheme, *in);
    } else {
      throw std::invalid_argument("unknown policy");
    }
    return 0;
  } catch (const std::exception& e) {
    if (err != nullptr && err_len > 0) {
      std::strncpy(err, e.what(), err_len - 1);
      err[err_len - 1] = '\0';
    }
    return 1;
  }
}

void oracle_destroy(void* h) { delete static_cast<oracle::Base*>(h); }

void oracle_residual(void* h, const double* u, double* res, int gather) {
  static_cast<oracle::Base*>(h)->residual(u, res, gather != 0);
}

void oracle_spectral_radius(void* h, double* lam) { static_cast<oracle::Base*>(h)->spectral_radius(lam); }

void oracle_face_fluxes(void* h, double* flux) { static_cast<oracle::Base*>(h)->face_fluxes(flux); }

void oracle_arrays(void* h, double* prim, double* grad, double* phi, double* wmin, double* wmax, double* prim_b) {
  static_cast<oracle::Base*>(h)->arrays(prim, grad, phi, wmin, wmax, prim_b);
}

void oracle_set_second_order(void* h, int on) { static_cast<oracle::Base*>(h)->set_second_order(on); }

int oracle_step(void* h, double* u, double cfl, int has_dt, double dt, int local, int gather, double* residuals,
                double* time_step, int* invalid) {
  return static_cast<oracle::Base*>(h)->step(u, cfl, has_dt, dt, local, gather, residuals, time_step, invalid);
}

void oracle_face_force(void* h, const double* u, const std::int32_t* faces, std::int64_t n, double* out) {
  static_cast<oracle::Base*>(h)->face_force(u, faces, static_cast<std::size_t>(n), out);
}

void oracle_conservation(void* h, const double* u, double* out) {
  static_cast<oracle::Base*>(h)->conservation(u, out);
}

void oracle_set_threads(int n) { oracle::g_threads = n < 1 ? 1 : n; }

int oracle_threads() { return oracle::g_threads; }

int oracle_first_invalid(void* h, const double* u, const double* res) {
  return static_cast<oracle::Base*>(h)->first_invalid(u, res);
}

double oracle_exact_sum(const double* x, std::int64_t n) {
  oracle::ExactSum s;
  for (std::int64_t i = 0; i < n; ++i) s.add(x[i]);
  return s.value();
}

// Single bodies, for the closed-form and identity tests: policy 0 = double, 1 = float compute.
void oracle_flux(int policy, int flux, const double* wl, const double* wr, const double* n, double g, double minf,
                 double pref, double* out) {
  if (policy == 0) {
    oracle::numerical_flux<double>(flux, wl, wr, n, g, minf, pref, out);
    return;
  }
  float a[5], b[5], m[3], o[5];
  for (int k = 0; k < 5; ++k) {
    a[k] = float(wl[k]);
    b[k] = float(wr[k]);
  }
  for (int k = 0; k < 3; ++k) m[k] = float(n[k]);
  oracle::numerical_flux<float>(flux, a, b, m, float(g), float(minf), float(pref), o);
  for (int k = 0; k < 5; ++k) out[k] = o[k];
}

// The Jacobian blocks of one face, one boundary face and one cell, in the policy's compute type.
extern "C++" {
template <class C>
static void copy_in(const double* x, C* y, int n) {
  for (int k = 0; k < n; ++k) y[k] = static_cast<C>(x[k]);
}

template <class C>
static void jacobian_of(int what, int flux, const double* a, const double* b, double kind, const double* n, double g,
                        double rgas, double minf, double pref, double* out) {
  C x[5], y[5], m[3], o[50];
  copy_in(a, x, 5);
  if (b != nullptr) copy_in(b, y, 5);
  if (n != nullptr) copy_in(n, m, 3);
  int count = 25;
  if (what == 0) {
    oracle::face_jacobian<C>(flux, x, y, m, C(g), C(minf), C(pref), o);
    count = 50;
  } else if (what == 1) {
    oracle::boundary_jacobian<C>(flux, x, C(kind), y, m, C(g), C(rgas), C(minf), C(pref), o);
  } else {
    oracle::dwdu<C>(x, y, C(g), o);
  }
  for (int q = 0; q < count; ++q) out[q] = o[q];
}
}  // extern "C++"

void oracle_face_jacobian(int policy, int flux, const double* wl, const double* wr, const double* n, double g,
                          double minf, double pref, double* out) {
  if (policy == 0) {
    jacobian_of<double>(0, flux, wl, wr, 0.0, n, g, 0.0, minf, pref, out);
  } else {
    jacobian_of<float>(0, flux, wl, wr, 0.0, n, g, 0.0, minf, pref, out);
  }
}

void oracle_boundary_jacobian(int policy, int flux, const double* wf, double kind, const double* data,
                              const double* n, double g, double rgas, double minf, double pref, double* out) {
  if (policy == 0) {
    jacobian_of<double>(1, flux, wf, data, kind, n, g, rgas, minf, pref, out);
  } else {
    jacobian_of<float>(1, flux, wf, data, kind, n, g, rgas, minf, pref, out);
  }
}

void oracle_dwdu(int policy, const double* dq, const double* ur, double g, double* out) {
  if (policy == 0) {
    jacobian_of<double>(2, 0, dq, ur, 0.0, nullptr, g, 0.0, 0.0, 0.0, out);
  } else {
    jacobian_of<float>(2, 0, dq, ur, 0.0, nullptr, g, 0.0, 0.0, 0.0, out);
  }
}

void oracle_physical_flux(const double* w, const double* n, double g, double pref, double* out) {
  oracle::physical_flux<double>(w, n, g, pref, out);
}

void oracle_boundary_value(const double* wg, double kind, const double* data, const double* n, double g,
                           double rgas, double pref, double* out) {
  oracle::boundary_value<double>(wg, kind, data, n, g, rgas, pref, out);
}

void oracle_to_primitive(const double* dq, const double* ur, double g, double* w) {
  oracle::to_primitive<double>(dq, ur, g, w);
}

void oracle_residual_and_jacobian(void* h, const double* u, double* res, double* dwdu, double* fj, double* bj) {
  static_cast<oracle::Base*>(h)->residual_and_jacobian(u, res, dwdu, fj, bj);
}

// One implicit solve and update from the Jacobian's parts (the device's layout), the spectral
// radius lam (per cell, widened) and the residual res at state u (h the oracle that evaluated them,
// second the order of that evaluation: at second order the Krylov operator adds the boundary
// faces' gradient term): assembles lam / cfl I + dR/dU,
// solves it by GCR and the symmetric sweeps, and unless the solve failed (or a diagonal block is
// singular: out[4] = 1) adds the admissible fraction of du to u. out: cycles, residual drop, alpha
// (0 when no update), failed, singular. du receives the solve's result; colours the cells' colours.
void oracle_implicit_solve(void* h, int second, std::int64_t n, std::int64_t ni, const std::int32_t* owner,
                           const std::int32_t* neighbour, const std::int64_t* cfo, const std::int32_t* cf,
                           const double* dwdu, const double* fj, const double* bj, const double* lam, double cfl,
                           const double* res, double gamma, double* u, double* du, double* out, std::int32_t* colours) {
  namespace im = oracle::implicit;
  const std::size_t nc = static_cast<std::size_t>(n), nif = static_cast<std::size_t>(ni);
  const std::size_t nb = static_cast<std::size_t>(cfo[n]) - 2 * nif;
  const im::Topo t{nc, nif, nb, owner, neighbour, cfo, cf};
  const im::Rows rows = im::rows_of(t);
  for (std::size_t c = 0; c < nc; ++c) colours[c] = rows.colour[c];
  std::vector<double> shift(nc), diag(25 * nc), upper(25 * nif), lower(25 * nif), inverse(25 * nc);
  for (std::size_t c = 0; c < nc; ++c) shift[c] = lam[c] / cfl;
  im::assemble(t, dwdu, fj, bj, shift.data(), diag.data(), upper.data(), lower.data());
  for (int q = 0; q < 5; ++q) out[q] = 0.0;
  bool singular = false;
  for (std::size_t c = 0; c < nc && !singular; ++c) singular = !im::invert(&diag[25 * c], &inverse[25 * c]);
  std::fill(du, du + 5 * nc, 0.0);
  if (singular) {
    out[1] = std::numeric_limits<double>::quiet_NaN();
    out[3] = 1.0;
    out[4] = 1.0;
    return;
  }
  const im::Matrix a{&t, &rows, diag.data(), upper.data(), lower.data(), inverse.data()};
  std::vector<double> r(5 * nc);
  for (std::size_t i = 0; i < 5 * nc; ++i) r[i] = -res[i];
  // The boundary faces' gradient term, when the Jacobian was assembled at second order.
  const auto* solver = static_cast<const oracle::Base*>(h);
  auto extra = [&](const double* v, double* y) {
    if (second != 0 && solver != nullptr) solver->boundary_gradient_term(v, dwdu, bj, y);
  };
  const im::Report rep = im::gcr(a, r.data(), du, 5 * nc, extra);
  out[0] = rep.cycles;
  out[1] = rep.residual_drop;
  if (!(rep.residual_drop <= im::kFailure)) {
    out[3] = 1.0;
    return;
  }
  const double alpha = im::admissible_step(u, du, nc, gamma);
  out[2] = alpha;
  if (alpha < im::kMinStep) return;
  for (std::size_t i = 0; i < 5 * nc; ++i) u[i] += alpha * du[i];
}

// EulerPhysics::admissible_step of u + alpha du over n cells (5 x n SoA).
// y += the implicit step's boundary gradient term of direction v for the state of the last
// oracle_residual_and_jacobian (its dwdu and bnd_jac).
void oracle_boundary_term(void* h, const double* v, const double* dwdu, const double* bj, double* y) {
  static_cast<oracle::Base*>(h)->boundary_gradient_term(v, dwdu, bj, y);
}

double oracle_admissible_step(std::int64_t n, const double* u, const double* du, double gamma) {
  return oracle::implicit::admissible_step(u, du, static_cast<std::size_t>(n), gamma);
}

double oracle_venkat(double d, double dmax, double dmin, double eps2) {
  const double dp = d > 0.0 ? dmax : dmin;
  return (dp * dp + eps2 + 2.0 * d * dp) / (dp * dp + 2.0 * d * d + d * dp + eps2);
}

void oracle_warp(const double* w, const double* gl, const double* pl, const double* gr, const double* pr,
                 const double* m, double g, double pref, double* out) {
  oracle::warp<double>(w, gl, pl, gr, pr, m, g, pref, out);
}

}  // extern "C"
