// Pointwise bodies of the explicit Euler residual on the device, in the compute type C.
//
// Each function follows the operation order of its @pointwise body in the reference's
// kernels/euler.py, with the semantics its generator gives them (codegen/src/autan_codegen/
// frontend.py): abs(x) = x >= 0 ? x : -x, min(a, b) = a <= b ? a : b, max(a, b) = a >= b ? a : b,
// operators associate left to right, dot and matrix products start from their first product,
// products of constants are folded in FP64 and x * 1 is dropped. The library is built with
// -ffp-contract=off, so no product is fused into an FMA and every result is rounded as the
// host's. Branches of the DSL (which evaluates both sides and selects) are ordinary branches here:
// the selected value is the same.
#pragma once

#include <hip/hip_runtime.h>

namespace mxe {

template <class C> __device__ __forceinline__ C sabs(C x) { return x >= C(0) ? x : -x; }
template <class C> __device__ __forceinline__ C smin(C a, C b) { return a <= b ? a : b; }
template <class C> __device__ __forceinline__ C smax(C a, C b) { return a >= b ? a : b; }

__device__ __forceinline__ float dsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float dexp(float x) { return expf(x); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ float dlog(float x) { return logf(x); }
__device__ __forceinline__ double dlog(double x) { return log(x); }

template <class C>
struct V3 {
  C x, y, z;
};
template <class C> __device__ __forceinline__ C dot(const V3<C>& a, const V3<C>& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}

// W = (rho, u, v, w, p - p_ref) of a face side.
template <class C>
struct W5 {
  C r, u, v, w, p;
  __device__ __forceinline__ V3<C> vel() const { return {u, v, w}; }
};

template <class C>
struct F5 {
  C f[5];
};

// kernels/euler.py:94-105 (to_primitive): dq = u - u_ref as stored, ur = u_ref.
template <class C>
__device__ __forceinline__ W5<C> to_primitive(const C* dq, const C* ur, C g) {
  const C rho = ur[0] + dq[0];
  const V3<C> mr{ur[1], ur[2], ur[3]};
  const V3<C> dm{dq[1], dq[2], dq[3]};
  const V3<C> t{C(2.0) * mr.x + dm.x, C(2.0) * mr.y + dm.y, C(2.0) * mr.z + dm.z};
  const C dk = (dot(dm, t) * ur[0] - dot(mr, mr) * dq[0]) / (rho * ur[0]);
  return {rho, (mr.x + dm.x) / rho, (mr.y + dm.y) / rho, (mr.z + dm.z) / rho, (g - C(1.0)) * (dq[4] - C(0.5) * dk)};
}

// kernels/euler.py:113-115
template <class C>
__device__ __forceinline__ C enthalpy(const W5<C>& w, C g, C pr) {
  const V3<C> v = w.vel();
  return g / (g - C(1.0)) * (w.p + pr) / w.r + C(0.5) * dot(v, v);
}

// kernels/euler.py:119-130
template <class C>
__device__ __forceinline__ F5<C> physical(const W5<C>& w, const V3<C>& n, C g, C pr) {
  const V3<C> v = w.vel();
  const C vn = dot(v, n);
  const C m = w.r * vn;
  return {{m, m * v.x + w.p * n.x, m * v.y + w.p * n.y, m * v.z + w.p * n.z, m * enthalpy(w, g, pr)}};
}

// kernels/euler.py:136-141
template <class C>
__device__ __forceinline__ C harten(C lam, C delta) {
  C a = sabs(lam);
  if (a < delta) a = (lam * lam + delta * delta) / (C(2.0) * delta);
  return a;
}

// kernels/euler.py:145-175: Roe, wave strengths of Blazek eqs. 4.89-4.91, Harten's fix on the
// acoustic waves.
template <class C>
__device__ __forceinline__ F5<C> roe(const W5<C>& L, const W5<C>& R, const V3<C>& n, C g, C pr) {
  const C rl = L.r, rr = R.r;
  const V3<C> vl = L.vel(), vr = R.vel();
  const C hl = enthalpy(L, g, pr), hr = enthalpy(R, g, pr);
  const C s = dsqrt(rr / rl);
  const C rho = s * rl;
  const C den = C(1.0) + s;
  const V3<C> v{(vl.x + s * vr.x) / den, (vl.y + s * vr.y) / den, (vl.z + s * vr.z) / den};
  const C h = (hl + s * hr) / den;
  const C vn = dot(v, n);
  const C q2 = dot(v, v);
  const C c = dsqrt((g - C(1.0)) * (h - C(0.5) * q2));
  const C dp = R.p - L.p;
  const C drho = rr - rl;
  const V3<C> dv{vr.x - vl.x, vr.y - vl.y, vr.z - vl.z};
  const C dvn = dot(dv, n);
  const C delta = C(0.1) * c;
  const C l1 = harten(vn - c, delta);
  const C l2 = sabs(vn);
  const C l5 = harten(vn + c, delta);
  const C rcd = rho * c * dvn;
  const C cc2 = C(2.0) * c * c;
  const C a1 = l1 * (dp - rcd) / cc2;
  const C a5 = l5 * (dp + rcd) / cc2;
  const C a2 = l2 * (drho - dp / (c * c));
  const V3<C> shear{dv.x - dvn * n.x, dv.y - dvn * n.y, dv.z - dvn * n.z};
  const C d0 = a1 + a2 + a5;
  const C l2r = l2 * rho;
  const C dmx = a1 * (v.x - c * n.x) + a2 * v.x + l2r * shear.x + a5 * (v.x + c * n.x);
  const C dmy = a1 * (v.y - c * n.y) + a2 * v.y + l2r * shear.y + a5 * (v.y + c * n.y);
  const C dmz = a1 * (v.z - c * n.z) + a2 * v.z + l2r * shear.z + a5 * (v.z + c * n.z);
  const C de = a1 * (h - c * vn) + a2 * C(0.5) * q2 + l2r * (dot(v, dv) - vn * dvn) + a5 * (h + c * vn);
  const F5<C> fl = physical(L, n, g, pr), fr = physical(R, n, g, pr);
  const C d[5] = {d0, dmx, dmy, dmz, de};
  F5<C> out;
#pragma unroll
  for (int k = 0; k < 5; ++k) out.f[k] = C(0.5) * (fl.f[k] + fr.f[k]) - C(0.5) * d[k];
  return out;
}

// kernels/euler.py:178-191 (hllc_star) minus kernels/euler.py:194-203 (conserved), the bracket
// of the HLLC star flux F_K + S_K (U*_K - U_K).
template <class C>
__device__ __forceinline__ F5<C> hllc_side(const W5<C>& w, const F5<C>& fk, const V3<C>& n, C sk, C sm, C g,
                                           C pr) {
  const V3<C> v = w.vel();
  const C vn = dot(v, n);
  const C rs = w.r * (sk - vn) / (sk - sm);
  const C p = w.p + pr;
  const C vv = dot(v, v);
  const C e = p / ((g - C(1.0)) * w.r) + C(0.5) * vv;
  const C smv = sm - vn;
  const C st[5] = {rs, rs * (v.x + smv * n.x), rs * (v.y + smv * n.y), rs * (v.z + smv * n.z),
                   rs * (e + smv * (sm + p / (w.r * (sk - vn))))};
  const C u[5] = {w.r, w.r * v.x, w.r * v.y, w.r * v.z, (w.p + pr) / (g - C(1.0)) + C(0.5) * w.r * vv};
  F5<C> out;
#pragma unroll
  for (int k = 0; k < 5; ++k) out.f[k] = fk.f[k] + sk * (st[k] - u[k]);
  return out;
}

// kernels/euler.py:206-233: HLLC of Toro, Spruce and Speares with Einfeldt/Batten wave speeds.
template <class C>
__device__ __forceinline__ F5<C> hllc(const W5<C>& L, const W5<C>& R, const V3<C>& n, C g, C pr) {
  const C rl = L.r, rr = R.r;
  const V3<C> vl = L.vel(), vr = R.vel();
  const C vnl = dot(vl, n), vnr = dot(vr, n);
  const C cl = dsqrt(g * (L.p + pr) / rl);
  const C cr = dsqrt(g * (R.p + pr) / rr);
  const C s = dsqrt(rr / rl);
  const C den = C(1.0) + s;
  const V3<C> v{(vl.x + s * vr.x) / den, (vl.y + s * vr.y) / den, (vl.z + s * vr.z) / den};
  const C h = (enthalpy(L, g, pr) + s * enthalpy(R, g, pr)) / den;
  const C c = dsqrt((g - C(1.0)) * (h - C(0.5) * dot(v, v)));
  const C vn = dot(v, n);
  const C sl = smin(vnl - cl, vn - c);
  const C sr = smax(vnr + cr, vn + c);
  const C sm = (R.p - L.p + rl * vnl * (sl - vnl) - rr * vnr * (sr - vnr)) / (rl * (sl - vnl) - rr * (sr - vnr));
  if (sl >= C(0.0)) return physical(L, n, g, pr);
  if (sm >= C(0.0)) return hllc_side(L, physical(L, n, g, pr), n, sl, sm, g, pr);
  if (sr > C(0.0)) return hllc_side(R, physical(R, n, g, pr), n, sr, sm, g, pr);
  return physical(R, n, g, pr);
}

// kernels/euler.py:236-273 (Liou's split polynomials; 16 beta = 2 folded).
template <class C>
__device__ __forceinline__ C m4p(C m) {
  const C mp = C(0.25) * (m + C(1.0)) * (m + C(1.0));
  const C mm = C(-0.25) * (m - C(1.0)) * (m - C(1.0));
  return sabs(m) >= C(1.0) ? C(0.5) * (m + sabs(m)) : mp * (C(1.0) - C(2.0) * mm);
}
template <class C>
__device__ __forceinline__ C m4m(C m) {
  const C mp = C(0.25) * (m + C(1.0)) * (m + C(1.0));
  const C mm = C(-0.25) * (m - C(1.0)) * (m - C(1.0));
  return sabs(m) >= C(1.0) ? C(0.5) * (m - sabs(m)) : mm * (C(1.0) + C(2.0) * mp);
}
template <class C>
__device__ __forceinline__ C p5p(C m, C alpha) {
  const C mp = C(0.25) * (m + C(1.0)) * (m + C(1.0));
  const C mm = C(-0.25) * (m - C(1.0)) * (m - C(1.0));
  return sabs(m) >= C(1.0) ? (m > C(0.0) ? C(1.0) : C(0.0)) : mp * ((C(2.0) - m) - C(16.0) * alpha * m * mm);
}
template <class C>
__device__ __forceinline__ C p5m(C m, C alpha) {
  const C mp = C(0.25) * (m + C(1.0)) * (m + C(1.0));
  const C mm = C(-0.25) * (m - C(1.0)) * (m - C(1.0));
  return sabs(m) >= C(1.0) ? (m < C(0.0) ? C(1.0) : C(0.0)) : mm * ((C(-2.0) - m) + C(16.0) * alpha * m * mp);
}

// kernels/euler.py:276-303: AUSM+-up (Liou 2006) with the gauge interface pressure.
template <class C>
__device__ __forceinline__ F5<C> ausm(const W5<C>& L, const W5<C>& R, const V3<C>& n, C g, C minf, C pr) {
  const C rl = L.r, rr = R.r;
  const V3<C> vl = L.vel(), vr = R.vel();
  const C vnl = dot(vl, n), vnr = dot(vr, n);
  const C hl = enthalpy(L, g, pr), hr = enthalpy(R, g, pr);
  const C k = C(2.0) * (g - C(1.0)) / (g + C(1.0));
  const C csl = dsqrt(k * hl);
  const C csr = dsqrt(k * hr);
  const C al = csl * csl / smax(csl, vnl);
  const C ar = csr * csr / smax(csr, -vnr);
  const C a = smin(al, ar);
  const C ml = vnl / a;
  const C mr = vnr / a;
  const C aa = a * a;
  const C mbar2 = C(0.5) * (vnl * vnl + vnr * vnr) / aa;
  const C mo2 = smin(C(1.0), smax(mbar2, minf * minf));
  const C mo = dsqrt(mo2);
  const C fa = mo * (C(2.0) - mo);
  const C alpha = C(0.1875) * (C(-4.0) + C(5.0) * fa * fa);
  const C rho = C(0.5) * (rl + rr);
  const C mp = C(-0.25) / fa * smax(C(1.0) - mbar2, C(0.0)) * (R.p - L.p) / (rho * a * a);
  const C m = m4p(ml) + m4m(mr) + mp;
  const C pp = p5p(ml, alpha);
  const C pm = p5m(mr, alpha);
  const C pu = C(-0.75) * pp * pm * (rl + rr) * fa * a * (vnr - vnl);
  const C p = pp * L.p + pm * R.p + pu + (pp + pm - C(1.0)) * pr;
  const bool up = m > C(0.0);
  const C mdot = up ? a * m * rl : a * m * rr;
  const V3<C> vu = up ? vl : vr;
  const C hu = up ? hl : hr;
  return {{mdot, mdot * vu.x + p * n.x, mdot * vu.y + p * n.y, mdot * vu.z + p * n.z, mdot * hu}};
}

template <int FLUX, class C>
__device__ __forceinline__ F5<C> flux(const W5<C>& L, const W5<C>& R, const V3<C>& n, C g, C minf, C pr) {
  if constexpr (FLUX == 0) {
    return roe(L, R, n, g, pr);
  } else if constexpr (FLUX == 1) {
    return hllc(L, R, n, g, pr);
  } else {
    return ausm(L, R, n, g, minf, pr);
  }
}

// kernels/euler.py:309-337: Riemann invariants normal to the face.
template <class C>
__device__ __forceinline__ W5<C> farfield(const W5<C>& wi, const W5<C>& wf, const V3<C>& n, C g) {
  const V3<C> vi = wi.vel(), vf = wf.vel();
  const C vni = dot(vi, n), vnf = dot(vf, n);
  const C ci = dsqrt(g * wi.p / wi.r);
  const C cf = dsqrt(g * wf.p / wf.r);
  if (vni >= ci) return wi;
  if (vnf <= -cf) return wf;
  const C rp = vni + C(2.0) * ci / (g - C(1.0));
  const C rm = vnf - C(2.0) * cf / (g - C(1.0));
  const C vnb = C(0.5) * (rp + rm);
  const C cb = C(0.25) * (g - C(1.0)) * (rp - rm);
  C ent;
  V3<C> vt;
  if (vnb < C(0.0)) {
    ent = wf.p / dexp(g * dlog(wf.r));
    vt = {vf.x - vnf * n.x, vf.y - vnf * n.y, vf.z - vnf * n.z};
  } else {
    ent = wi.p / dexp(g * dlog(wi.r));
    vt = {vi.x - vni * n.x, vi.y - vni * n.y, vi.z - vni * n.z};
  }
  const C rb = dexp(dlog(cb * cb / (g * ent)) / (g - C(1.0)));
  return {rb, vt.x + vnb * n.x, vt.y + vnb * n.y, vt.z + vnb * n.z, rb * cb * cb / g};
}

// kernels/euler.py:340-358: subsonic inflow from total pressure and temperature (Carlson 2011).
template <class C>
__device__ __forceinline__ W5<C> inlet(const W5<C>& wi, const C* data, const V3<C>& n, C g, C rgas) {
  const C p0 = data[0], t0 = data[1];
  const C gm = g - C(1.0);
  const C ci = dsqrt(g * wi.p / wi.r);
  const C r = dot(wi.vel(), n) + C(2.0) * ci / gm;
  const C h0 = g * rgas * t0 / gm;
  const C disc = smax((gm + C(2.0)) * h0 - C(0.5) * gm * r * r, C(0.0));
  const C cb = gm * (r + dsqrt(disc)) / (gm + C(2.0));
  const C vnb = smin(r - C(2.0) * cb / gm, C(0.0));
  const C tb = t0 - C(0.5) * gm * vnb * vnb / (g * rgas);
  const C pb = p0 * dexp(g / gm * dlog(tb / t0));
  const C rb = pb / (rgas * tb);
  return {rb, vnb * n.x, vnb * n.y, vnb * n.z, pb};
}

// kernels/euler.py:361-372: subsonic outflow at a static pressure.
template <class C>
__device__ __forceinline__ W5<C> outlet(const W5<C>& wi, const C* data, const V3<C>& n, C g) {
  const V3<C> vi = wi.vel();
  const C ci = dsqrt(g * wi.p / wi.r);
  if (dot(vi, n) >= ci) return wi;
  const C pb = data[4];
  const C dp = wi.p - pb;
  const C rb = wi.r - dp / (ci * ci);
  const C s = dp / (wi.r * ci);
  return {rb, vi.x + s * n.x, vi.y + s * n.y, vi.z + s * n.z, pb};
}

// kernels/euler.py:391-422 with mirror_state (:375-379) and wall_state (:382-388) inline. Codes
// are whole numbers: kind is the code, as the midpoint comparisons read it.
template <class C>
__device__ __forceinline__ W5<C> boundary_value(const W5<C>& wg, int kind, const C* data, const V3<C>& n, C g,
                                                C rgas, C pr) {
  const W5<C> wi{wg.r, wg.u, wg.v, wg.w, wg.p + pr};
  W5<C> w;
  C shift = pr;
  switch (kind) {
    case 0:
      w = {data[0], data[1], data[2], data[3], data[4]};
      break;
    case 1:
      w = farfield(wi, W5<C>{data[0], data[1], data[2], data[3], data[4]}, n, g);
      break;
    case 2: {
      const V3<C> vi = wg.vel();
      const C s = C(2.0) * dot(vi, n);
      w = {wg.r, vi.x - s * n.x, vi.y - s * n.y, vi.z - s * n.z, wg.p};
      shift = C(0.0) * pr;
      break;
    }
    case 3:
      w = inlet(wi, data, n, g, rgas);
      break;
    case 4:
      w = outlet(wi, data, n, g);
      break;
    default:
      w = {data[0] > C(0.0) ? (wg.p + pr) / (rgas * data[0]) : wg.r, C(0.0), C(0.0), C(0.0), wg.p};
      shift = C(0.0) * pr;
      break;
  }
  w.p = w.p - shift;
  return w;
}

// kernels/euler.py:619-627: W_f = W + phi (grad W . r), the cell value where rho or p would not
// stay positive. g is the 5 x 3 gradient, row-major.
template <class C>
__device__ __forceinline__ W5<C> reconstruct(const W5<C>& w, const C* g, const C* lim, const V3<C>& r, C pref) {
  const C d0 = g[0] * r.x + g[1] * r.y + g[2] * r.z;
  const C d1 = g[3] * r.x + g[4] * r.y + g[5] * r.z;
  const C d2 = g[6] * r.x + g[7] * r.y + g[8] * r.z;
  const C d3 = g[9] * r.x + g[10] * r.y + g[11] * r.z;
  const C d4 = g[12] * r.x + g[13] * r.y + g[14] * r.z;
  const W5<C> f{w.r + lim[0] * d0, w.u + lim[1] * d1, w.v + lim[2] * d2, w.w + lim[3] * d3, w.p + lim[4] * d4};
  return (f.r > C(0.0) && f.p + pref > C(0.0)) ? f : w;
}

// kernels/euler.py:635-650 for the unit direction e_j (j = 0, 1, 2): dot(v, e) keeps the products
// with 0 that the generator keeps (only x * 1 is dropped).
template <int J, class C>
__device__ __forceinline__ F5<C> flux_increment(const W5<C>& w, const W5<C>& dw, C g, C pref) {
  const V3<C> v = w.vel(), dv = dw.vel();
  const C e[3] = {C(J == 0 ? 1.0 : 0.0), C(J == 1 ? 1.0 : 0.0), C(J == 2 ? 1.0 : 0.0)};
  const C ve = (J == 0 ? v.x : v.x * e[0]) + (J == 1 ? v.y : v.y * e[1]) + (J == 2 ? v.z : v.z * e[2]);
  const C dve = (J == 0 ? dv.x : dv.x * e[0]) + (J == 1 ? dv.y : dv.y * e[1]) + (J == 2 ? dv.z : dv.z * e[2]);
  const C vv = dot(v, v);
  const C gg = g / (g - C(1.0));
  const C rh = gg * (w.p + pref) + C(0.5) * w.r * vv;
  const C drh = gg * dw.p + C(0.5) * dw.r * vv + w.r * dot(v, dv);
  const C dm = dw.r * ve + w.r * dve;
  const C wve = w.r * ve;
  const C px = J == 0 ? dw.p : dw.p * e[0];
  const C py = J == 1 ? dw.p : dw.p * e[1];
  const C pz = J == 2 ? dw.p : dw.p * e[2];
  return {{dm, dm * v.x + wve * dv.x + px, dm * v.y + wve * dv.y + py, dm * v.z + wve * dv.z + pz,
           drh * ve + rh * dve}};
}

// kernels/euler.py:653-677: the warped-face term's three directions
// d_j = (phi_L grad W_L + phi_R grad W_R) M[:, j] / 2; m is the 3 x 3 moment, row-major.
template <class C>
__device__ __forceinline__ void warp_directions(const C* gl, const C* pl, const C* gr, const C* pr, const C* m,
                                                W5<C>* d) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const C c0 = m[j], c1 = m[3 + j], c2 = m[6 + j];
    C t[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      const C a = gl[3 * i] * c0 + gl[3 * i + 1] * c1 + gl[3 * i + 2] * c2;
      const C b = gr[3 * i] * c0 + gr[3 * i + 1] * c1 + gr[3 * i + 2] * c2;
      t[i] = C(0.5) * (pl[i] * a + pr[i] * b);
    }
    d[j] = {t[0], t[1], t[2], t[3], t[4]};
  }
}

// The term itself, sum_j dF_j[W; d_j], for the directions of warp_directions. T may be a dual
// number (the Jacobian's), the directions then held fixed.
template <class T, class C>
__device__ __forceinline__ F5<T> warp_of_directions(const W5<T>& w, const W5<C>* d, T g, T pref) {
  const W5<T> d0{T(d[0].r), T(d[0].u), T(d[0].v), T(d[0].w), T(d[0].p)};
  const W5<T> d1{T(d[1].r), T(d[1].u), T(d[1].v), T(d[1].w), T(d[1].p)};
  const W5<T> d2{T(d[2].r), T(d[2].u), T(d[2].v), T(d[2].w), T(d[2].p)};
  const F5<T> f0 = flux_increment<0>(w, d0, g, pref);
  const F5<T> f1 = flux_increment<1>(w, d1, g, pref);
  const F5<T> f2 = flux_increment<2>(w, d2, g, pref);
  F5<T> out;
#pragma unroll
  for (int k = 0; k < 5; ++k) out.f[k] = f0.f[k] + f1.f[k] + f2.f[k];
  return out;
}

template <class C>
__device__ __forceinline__ F5<C> warp(const W5<C>& w, const C* gl, const C* pl, const C* gr, const C* pr,
                                      const C* m, C g, C pref) {
  W5<C> d[3];
  warp_directions(gl, pl, gr, pr, m, d);
  return warp_of_directions<C, C>(w, d, g, pref);
}

// kernels/euler.py:1091-1112 (wave_speed), before the area.
template <class C>
__device__ __forceinline__ C wave_speed(const W5<C>& w, const V3<C>& n, C g, C pref) {
  return sabs(dot(w.vel(), n)) + dsqrt(g * (w.p + pref) / w.r);
}

}  // namespace mxe
