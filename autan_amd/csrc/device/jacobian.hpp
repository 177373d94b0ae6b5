// The implicit solvers' approximate Jacobian (physics/src/euler.cpp:983-991; kernels/euler.py:434-440,
// 854-860, 1051-1057): dW/dU of every cell, and dF/dW_L, dF/dW_R of every face with the
// reconstruction's gradients and limiter held fixed. The derivatives come
// from the same flux bodies evaluated on dual numbers (forward-mode automatic differentiation, as
// the reference's generator differentiates its kernels, codegen/src/autan_codegen/ad.py).
#pragma once

#include <type_traits>

#include "kernels.hpp"

namespace mxe {

// A value and its derivatives along N directions; each operation applies its chain rule.
// Comparisons read the value, so a select takes the derivative of the branch it takes.
template <class C, int N>
struct Dual {
  C v;
  C d[N];
  Dual() = default;
  template <class T, std::enable_if_t<std::is_arithmetic_v<T>, int> = 0>
  __device__ __forceinline__ Dual(T x) : v(static_cast<C>(x)) {
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = C(0);
  }
  friend __device__ __forceinline__ Dual operator+(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
  }
  friend __device__ __forceinline__ Dual operator-(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
  }
  friend __device__ __forceinline__ Dual operator-(const Dual& a) {
    Dual r;
    r.v = -a.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
  }
  friend __device__ __forceinline__ Dual operator*(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
  }
  friend __device__ __forceinline__ Dual operator/(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v / b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
  }
  friend __device__ __forceinline__ bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
  friend __device__ __forceinline__ bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
  friend __device__ __forceinline__ bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
  friend __device__ __forceinline__ bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
};

template <class C, int N>
__device__ __forceinline__ Dual<C, N> dsqrt(const Dual<C, N>& a) {
  Dual<C, N> r;
  r.v = dsqrt(a.v);
  const C h = C(0.5) / r.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * h;
  return r;
}

template <class C, int N>
__device__ __forceinline__ Dual<C, N> dexp(const Dual<C, N>& a) {
  Dual<C, N> r;
  r.v = dexp(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * r.v;
  return r;
}

template <class C, int N>
__device__ __forceinline__ Dual<C, N> dlog(const Dual<C, N>& a) {
  Dual<C, N> r;
  r.v = dlog(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / a.v;
  return r;
}

// w as dual numbers; `seeded`: its five components are the five directions.
template <class C, int N>
__device__ __forceinline__ W5<Dual<C, N>> lift(const W5<C>& w, bool seeded) {
  using D = Dual<C, N>;
  W5<D> r{D(w.r), D(w.u), D(w.v), D(w.w), D(w.p)};
  if (seeded) {
    r.r.d[0] = C(1);
    r.u.d[1] = C(1);
    r.v.d[2] = C(1);
    r.w.d[3] = C(1);
    r.p.d[4] = C(1);
  }
  return r;
}

template <class C, int N>
__device__ __forceinline__ const Dual<C, N>& component(const W5<Dual<C, N>>& w, int k) {
  return k == 0 ? w.r : k == 1 ? w.u : k == 2 ? w.v : k == 3 ? w.w : w.p;
}

// primitives_jacobian (kernels/euler.py:434-440): dW/dq of every cell, q the stored state
// State(u - U_ref); entry (i, j) at (5 i + j) n_cells + c, widened to FP64, or with CELL_MAJOR at
// 25 c + 5 i + j (the implicit solve's layout: a cell's block one contiguous run), the block's
// 256 cells staged in LDS so that the stores stay contiguous across lanes.
template <class C, bool CELL_MAJOR = false>
__global__ __launch_bounds__(kBlock) void k_dwdu(Dev<C> d, const double* __restrict__ u, double* __restrict__ out) {
  const int c0 = blockIdx.x * kBlock;
  const int c = c0 + threadIdx.x;
  double v[25];
  if (c < d.nc) {
    using D = Dual<C, 5>;
    const Params<C>& pa = d.ptab[cell_case(d, c)];
    D dq[5], ur[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      dq[k] = D(static_cast<C>(u[static_cast<long long>(k) * d.nc + c] - pa.ref[k]));
      dq[k].d[k] = C(1);
      ur[k] = D(pa.ur[k]);
    }
    const W5<D> w = to_primitive(dq, ur, D(pa.g));
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      const D& wi = component(w, i);
#pragma unroll
      for (int j = 0; j < 5; ++j) v[5 * i + j] = static_cast<double>(wi.d[j]);
    }
  }
  if constexpr (CELL_MAJOR) {
    __shared__ double tile[25 * kBlock];
    if (c < d.nc) {
#pragma unroll
      for (int q = 0; q < 25; ++q) tile[25 * threadIdx.x + q] = v[q];
    }
    __syncthreads();
    const int count = 25 * (d.nc - c0 < kBlock ? d.nc - c0 : kBlock);
    double* o = out + 25LL * c0;
    for (int t = threadIdx.x; t < count; t += kBlock) o[t] = tile[t];
  } else {
    if (c >= d.nc) return;
#pragma unroll
    for (int q = 0; q < 25; ++q) out[static_cast<long long>(q) * d.nc + c] = v[q];
  }
}

// The face Jacobians of the residual's face and boundary loops
// (kernels/euler.py:854-860, 1051-1057), area-weighted as the fluxes: interior face f (owner to
// neighbour) dF/dW_L at (5 i + j) n_interior + f and dF/dW_R at (25 + 5 i + j) n_interior + f;
// boundary face b dF/dW_L at (5 i + j) n_boundary + b, W_L the cells' primitives: with the
// gradients and limiter fixed, the reconstructed state moves with them (dW_face/dW_cell = I). As
// in the reference's Jacobian loops (the bodies with warped-face terms, roe_boundary and its kin,
// kernels/euler.py:868-940), slip and wall faces take the pressure flux alone, and the other
// boundary kinds the flux against boundary_value of the reconstructed state, differentiated
// through it. On warped meshes k_face_jacobian_warp then adds the face body's warp term's part.
template <class C, int FLUX, bool SECOND>
__global__ __launch_bounds__(kBlock) void k_face_jacobian(Dev<C> d, double* __restrict__ fj, double* __restrict__ bj) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= d.nf) return;
  using D = Dual<C, 5>;
  C g[8];
  load_n<C, 8>(d.fgeo + 8LL * f, g);
  const V3<C> n{g[0], g[1], g[2]};
  const V3<D> nd{D(n.x), D(n.y), D(n.z)};
  const C area = g[3];
  if (f < d.ni) {
    const int2 lr = d.lr[f];
    const Params<C>& pa = d.ptab[cell_case(d, lr.x)];
    W5<C> fl = load_w(d.prim, lr.x), fr = load_w(d.prim, lr.y);
    if constexpr (SECOND) {
      C rr[4];
      load_n<C, 4>(d.frr + 4LL * f, rr);
      C gl[20], gr[20];
      load_rec(d.rec, lr.x, gl);
      load_rec(d.rec, lr.y, gr);
      fl = reconstruct(fl, gl, gl + 15, V3<C>{g[4], g[5], g[6]}, pa.pref);
      fr = reconstruct(fr, gr, gr + 15, V3<C>{rr[0], rr[1], rr[2]}, pa.pref);
    }
    for (int side = 0; side < 2; ++side) {
      const F5<D> fx = flux<FLUX>(lift<C, 5>(fl, side == 0), lift<C, 5>(fr, side == 1), nd, D(pa.g), D(pa.minf),
                                  D(pa.pref));
#pragma unroll
      for (int i = 0; i < 5; ++i) {
#pragma unroll
        for (int j = 0; j < 5; ++j) {
          fj[static_cast<long long>(25 * side + 5 * i + j) * d.ni + f] = static_cast<double>(area * fx.f[i].d[j]);
        }
      }
    }
  } else {
    const int b = f - d.ni;
    const int owner = d.bowner[b];
    const Params<C>& pa = d.ptab[cell_case(d, owner)];
    W5<C> wf = load_w(d.prim, owner);
    if constexpr (SECOND) {
      C gl[20];
      load_rec(d.rec, owner, gl);
      wf = reconstruct(wf, gl, gl + 15, V3<C>{g[4], g[5], g[6]}, pa.pref);
    }
    const int kind = d.bkind[b];
    const long long nb = d.nb;
    if (kind == 2 || kind >= 5) {
      const C an[3] = {area * n.x, area * n.y, area * n.z};
#pragma unroll
      for (int i = 0; i < 5; ++i) {
#pragma unroll
        for (int j = 0; j < 5; ++j) {
          const bool pressure = i >= 1 && i <= 3 && j == 4;
          bj[static_cast<long long>(5 * i + j) * nb + b] = pressure ? static_cast<double>(an[i - 1]) : 0.0;
        }
      }
    } else {
      C data[8];
      load_n<C, 8>(d.bc + 8LL * b, data);
      D dd[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) dd[k] = D(data[k]);
      const W5<D> wfd = lift<C, 5>(wf, true);
      const W5<D> wb = boundary_value(wfd, kind, dd, nd, D(pa.g), D(pa.rgas), D(pa.pref));
      const F5<D> fx = flux<FLUX>(wfd, wb, nd, D(pa.g), D(pa.minf), D(pa.pref));
#pragma unroll
      for (int i = 0; i < 5; ++i) {
#pragma unroll
        for (int j = 0; j < 5; ++j) bj[static_cast<long long>(5 * i + j) * nb + b] = static_cast<double>(area * fx.f[i].d[j]);
      }
    }
  }
}

// The warped-face term's part of the interior faces' blocks (second order, meshes with a face not
// proven planar), added after k_face_jacobian in the compute type: the stored area dF/dW (an FP64
// widening of a compute-type value, so its conversion back is exact) plus the warp term's
// derivative through its state (W_L + W_R) / 2, its directions
// dW_j = (phi_L grad W_L + phi_R grad W_R) M[:, j] / 2 held (kernels/euler.py:704-705 within
// 858-860). A kernel of its own keeps k_face_jacobian's registers.
template <class C>
__global__ __launch_bounds__(kBlock) void k_face_jacobian_warp(Dev<C> d, double* __restrict__ fj) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= d.ni) return;
  using D = Dual<C, 5>;
  C g[8], rr[4];
  load_n<C, 8>(d.fgeo + 8LL * f, g);
  load_n<C, 4>(d.frr + 4LL * f, rr);
  const int2 lr = d.lr[f];
  const Params<C>& pa = d.ptab[cell_case(d, lr.x)];
  C gl[20], gr[20], m9[9];
  load_rec(d.rec, lr.x, gl);
  load_rec(d.rec, lr.y, gr);
  const W5<C> fl = reconstruct(load_w(d.prim, lr.x), gl, gl + 15, V3<C>{g[4], g[5], g[6]}, pa.pref);
  const W5<C> fr = reconstruct(load_w(d.prim, lr.y), gr, gr + 15, V3<C>{rr[0], rr[1], rr[2]}, pa.pref);
#pragma unroll
  for (int q = 0; q < 9; ++q) m9[q] = d.moment[static_cast<long long>(q) * d.nf + f];
  W5<C> wd[3];
  warp_directions(gl, gl + 15, gr, gr + 15, m9, wd);
  for (int side = 0; side < 2; ++side) {
    const W5<D> dl = lift<C, 5>(fl, side == 0), dr = lift<C, 5>(fr, side == 1);
    const D half(0.5);
    const W5<D> w{half * (dl.r + dr.r), half * (dl.u + dr.u), half * (dl.v + dr.v), half * (dl.w + dr.w),
                  half * (dl.p + dr.p)};
    const F5<D> wx = warp_of_directions(w, wd, D(pa.g), D(pa.pref));
#pragma unroll
    for (int i = 0; i < 5; ++i) {
#pragma unroll
      for (int j = 0; j < 5; ++j) {
        double& out = fj[static_cast<long long>(25 * side + 5 * i + j) * d.ni + f];
        out = static_cast<double>(static_cast<C>(out) + wx.f[i].d[j]);
      }
    }
  }
}

}  // namespace mxe
