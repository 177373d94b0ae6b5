// The steady implicit step on the device: pseudo-transient continuation
// (V / dt + dR/dU) du = -R with the approximate Jacobian, as EulerPhysics::steady_step takes it
// with linear_solver = "gauss-seidel" (physics/src/euler_physics.cpp:1056-1250):
//   k_impl_diag, k_impl_off   Euler::assemble (physics/src/euler.cpp:993-1053): per cell
//                             lambda_c / CFL I + the sum over its faces, in cell_faces order, of
//                             +-J D_c; per interior face A(owner, neighbour) = J_R D_neighbour and
//                             A(neighbour, owner) = -J_L D_owner; each J D entry summed over k from 0
//                             and added as sign * t;
//   k_impl_invert             BlockMatrix::factor: each diagonal block's inverse by Gauss-Jordan
//                             elimination with partial pivoting (physics/src/block_matrix.cpp:14-45);
//   k_impl_relax(_run)        BlockMatrix::relax over the cells of one colour (a colour's cells share
//                             no face, so they update together), or of a run of small colours;
//   k_impl_multiply           BlockMatrix::multiply;
//   k_impl_dot, k_impl_maxabs the GCR's inner products as exact sums and its largest entry
//                             (EulerPhysics::owned_dot, owned_max_abs), the sums also fused into
//                             the GCR's updates (k_impl_orth_dots, k_impl_advance_dot);
//   k_impl_admissible         EulerPhysics::admissible_step: each cell's largest fraction of du that
//                             keeps its density and pressure above half their values, the least over
//                             the cells (an exact minimum of the FP64 values);
// and the GCR's vector updates (solvers/src/gcr.cpp). Everything in FP64, no contraction, so every
// value equals the host restatement's. Layouts (chosen for the relaxation, which streams a colour's
// rows): the rows are numbered by position in the colour-major cell order (the cells of colour 0
// ascending, then colour 1's, ...); the diagonal and inverse blocks are 25 contiguous values per
// position; a row's off-diagonal blocks are its entries, 25 contiguous values each, so a colour's
// rows and blocks are one contiguous run; the solve's vectors (the right-hand side, du, the GCR's
// directions) hold a cell's five values together, value q of cell c at 5 c + q.
#pragma once

#include "exact.hpp"
#include "jacobian.hpp"

namespace mxe {

static_assert(kWords == kExactWords, "the device's exact-sum words are exact_value's");

// Row r of out += sign * (j d) for 5 x 5 row-major blocks: out_r[c] += sign * sum_k j_r[k] dm[5 k + c],
// each entry's product summed over k from zero.
__device__ __forceinline__ void impl_add_product_row(double* out_r, const double* j_r, const double* dm, double sign) {
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) t += j_r[k] * dm[5 * k + c];
    out_r[c] += sign * t;
  }
}

// Row r of a block stored as impl_block reads it.
__device__ __forceinline__ void impl_block_row(const double* __restrict__ v, long long n, long long e, int k, int r,
                                               double* b_r) {
#pragma unroll
  for (int c = 0; c < 5; ++c) b_r[c] = v[(25LL * k + 5 * r + c) * n + e];
}

__device__ __forceinline__ void impl_block(const double* __restrict__ v, long long n, long long e, int k, double* b) {
#pragma unroll
  for (int q = 0; q < 25; ++q) b[q] = v[(25LL * k + q) * n + e];
}

// Cell c's 5 x 5 dW/dq block from the cell-major copy (25 contiguous values per cell).
__device__ __forceinline__ void impl_cell_block(const double* __restrict__ dwdu, long long c, double* b) {
  const double* p = dwdu + 25 * c;
#pragma unroll
  for (int q = 0; q < 25; ++q) b[q] = p[q];
}

// Five lanes per cell or face (12 per wave, lanes 60-63 idle), lane r computing row r of its blocks.
constexpr int kRowGroups = kWave / 5;

__device__ __forceinline__ int impl_group_item(int& r) {
  const int lane = threadIdx.x % kWave;
  const int g = lane / 5;
  r = lane % 5;
  if (g >= kRowGroups) return -1;
  return (blockIdx.x * (kBlock / kWave) + static_cast<int>(threadIdx.x / kWave)) * kRowGroups + g;
}

inline int impl_group_blocks(long long items) {
  constexpr int per_block = (kBlock / kWave) * kRowGroups;
  return static_cast<int>((items + per_block - 1) / per_block);
}

// The diagonal blocks: lam_c / cfl on the diagonal, then the cell's faces in cell_faces order.
// An interior face's term is the negated off-diagonal block of the transposed position, which
// k_impl_off has stored: the owner's +J_L D_owner is -A(neighbour, owner) (entry at[f].y), the
// neighbour's -J_R D_neighbour is -A(owner, neighbour) (entry at[f].x). Each such block is the
// same sum from zero, negated exactly, and a running sum from +0 is never -0, so adding the
// negated entry gives the bits of adding the product; it reads each entry's 25 contiguous values
// instead of recomputing the product from the face blocks. A boundary face's term J_b D_c from
// bj (25 x nb) and the cell's dW/dq block (cell-major). Five lanes per cell; stored at the cell's
// row position.
template <class C>
__global__ __launch_bounds__(kBlock) void k_impl_diag(Dev<C> d, const double* __restrict__ dwdu,
                                                      const double* __restrict__ bj, const double* __restrict__ offd,
                                                      const int2* __restrict__ at, double cfl,
                                                      const int* __restrict__ rank, double* __restrict__ diag) {
  int r;
  const int c = impl_group_item(r);
  if (c < 0 || c >= d.nc) return;
  double dg[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const double shift = static_cast<double>(d.lam[c]) / cfl;
  dg[r] += shift;
  for (int k = d.cfo[c]; k < d.cfo[c + 1]; ++k) {
    const int e = d.inc[k];
    const int f = e & 0x7fffffff;
    if (f >= d.ni) {
      double dc[25], j[5];
      impl_cell_block(dwdu, c, dc);
      impl_block_row(bj, d.nb, f - d.ni, 0, r, j);
      impl_add_product_row(dg, j, dc, 1.0);
    } else {
      const int2 a = at[f];
      const double* o = offd + 25LL * (e >= 0 ? a.y : a.x) + 5 * r;
#pragma unroll
      for (int q = 0; q < 5; ++q) dg[q] += -o[q];
    }
  }
  double* out = diag + 25LL * rank[c] + 5 * r;
#pragma unroll
  for (int q = 0; q < 5; ++q) out[q] = dg[q];
}

// The off-diagonal blocks of every interior face, at their entries (at[f].x in the owner's row,
// at[f].y in the neighbour's): J_R D_neighbour (row owner), -J_L D_owner (row neighbour). One
// thread per face: five lanes per face would each load the cells' whole dW/dq blocks, and a block
// computed a row at a time stores in pieces too small (measured twice the time).
template <class C>
__global__ __launch_bounds__(kBlock) void k_impl_off(Dev<C> d, const double* __restrict__ dwdu,
                                                     const double* __restrict__ fj, const int2* __restrict__ at,
                                                     double* __restrict__ offd) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= d.ni) return;
  const int2 lr = d.lr[f];
  const int2 k = at[f];
  double j[25], dm[25], out[25];
  impl_block(fj, d.ni, f, 1, j);
  impl_cell_block(dwdu, lr.y, dm);
#pragma unroll
  for (int q = 0; q < 25; ++q) out[q] = 0.0;
#pragma unroll
  for (int r = 0; r < 5; ++r) impl_add_product_row(out + 5 * r, j + 5 * r, dm, 1.0);
  double* up = offd + 25LL * k.x;
#pragma unroll
  for (int q = 0; q < 25; ++q) up[q] = out[q];
  impl_block(fj, d.ni, f, 0, j);
  impl_cell_block(dwdu, lr.x, dm);
#pragma unroll
  for (int q = 0; q < 25; ++q) out[q] = 0.0;
#pragma unroll
  for (int r = 0; r < 5; ++r) impl_add_product_row(out + 5 * r, j + 5 * r, dm, -1.0);
  double* lo = offd + 25LL * k.y;
#pragma unroll
  for (int q = 0; q < 25; ++q) lo[q] = out[q];
}

// Each diagonal block's inverse: Gauss-Jordan on [A | I] with partial pivoting (the first row
// holding the largest |pivot| below the diagonal, rows swapped whole, the pivot row scaled by the
// pivot's reciprocal, every other row reduced). A zero or non-finite pivot marks the cell in
// *singular (the least such cell). One thread per position p, cell[p] its cell.
__global__ __launch_bounds__(kBlock) void k_impl_invert(const double* __restrict__ diag, const int* __restrict__ cell,
                                                        int n, double* __restrict__ inverse,
                                                        int* __restrict__ singular) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const double* dp = diag + 25LL * p;
  double m[5][10];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      m[i][j] = dp[5 * i + j];
      m[i][5 + j] = i == j ? 1.0 : 0.0;
    }
  }
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    // The pivot row by its value's magnitude carried along, so that no register is indexed at
    // run time (the first strictly larger one wins, as when comparing with m[piv][k]).
    int piv = k;
    double best = fabs(m[k][k]);
#pragma unroll
    for (int i = k + 1; i < 5; ++i) {
      const double a = fabs(m[i][k]);
      if (a > best) {
        best = a;
        piv = i;
      }
    }
    // Swapped by selects, row by row.
#pragma unroll
    for (int i = k + 1; i < 5; ++i) {
      if (i == piv) {
#pragma unroll
        for (int j = 0; j < 10; ++j) {
          const double t = m[k][j];
          m[k][j] = m[i][j];
          m[i][j] = t;
        }
      }
    }
    if (m[k][k] == 0.0 || !__builtin_isfinite(m[k][k])) ok = false;
    const double dv = 1.0 / m[k][k];
#pragma unroll
    for (int j = 0; j < 10; ++j) m[k][j] *= dv;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      if (i == k) continue;
      const double f = m[i][k];
#pragma unroll
      for (int j = 0; j < 10; ++j) m[i][j] -= f * m[k][j];
    }
  }
  if (!ok) atomicMin(singular, cell[p]);
  double* ip = inverse + 25LL * p;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
#pragma unroll
    for (int j = 0; j < 5; ++j) ip[5 * i + j] = m[i][5 + j];
  }
}

// The rows by position p (colour-major): cell[p] the row's cell, its entries off[p] .. off[p + 1]
// (the cell's interior faces in ascending face order), entry k's block at offd + 25 k and its other
// cell other[k].
struct ImplRows {
  const int* off;    // n + 1
  const int* other;  // 2 ni
  const int* cell;   // n
};

// BlockMatrix::relax of one cell by a group of five lanes, lane q computing row q: x_c = D_c^-1
// (b_c - sum_f A_cf x_other). Row q's sum runs over the faces, then j, in the one-thread order;
// the group's five rows reach every lane of the group by shuffles for the inverse's product. Each
// wave holds 12 groups (lanes 60-63 idle); every lane takes part in the shuffles.
constexpr int kRelaxGroups = kWave / 5;

__device__ __forceinline__ void impl_relax_rows(const ImplRows& rows, int p, bool active,
                                                const double* __restrict__ offd, const double* __restrict__ inverse,
                                                const double* __restrict__ b, double* __restrict__ x) {
  const int lane = threadIdx.x % kWave;
  const int q = lane % 5, base = lane - q;
  const long long c = active ? rows.cell[p] : 0;
  double r = 0.0;
  if (active) {
    r = b[5 * c + q];
    for (int k = rows.off[p]; k < rows.off[p + 1]; ++k) {
      const double* a = offd + 25LL * k + 5 * q;
      const long long o = rows.other[k];
      double xo[5];
#pragma unroll
      for (int j = 0; j < 5; ++j) xo[j] = x[5 * o + j];
#pragma unroll
      for (int j = 0; j < 5; ++j) r -= a[j] * xo[j];
    }
  }
  double rj[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) rj[j] = __shfl(r, (base + j) % kWave);
  if (active) {
    const double* ip = inverse + 25LL * p + 5 * q;
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 5; ++j) s += ip[j] * rj[j];
    x[5 * c + q] = s;
  }
}

// BlockMatrix::relax of the rows at positions first .. first + count (one colour): 12 cells per
// wave.
__global__ __launch_bounds__(kBlock) void k_impl_relax(ImplRows rows, int first, int count,
                                                       const double* __restrict__ offd,
                                                       const double* __restrict__ inverse,
                                                       const double* __restrict__ b, double* __restrict__ x) {
  const int lane = threadIdx.x % kWave;
  const int g = lane / 5;
  const int t = (blockIdx.x * (kBlock / kWave) + static_cast<int>(threadIdx.x / kWave)) * kRelaxGroups + g;
  const bool active = g < kRelaxGroups && t < count;
  impl_relax_rows(rows, first + t, active, offd, inverse, b, x);
}

// Launch width of k_impl_relax for `count` cells.
inline int impl_relax_blocks(int count) {
  constexpr int per_block = (kBlock / kWave) * kRelaxGroups;
  return (count + per_block - 1) / per_block;
}

// A run of small colours relaxed in order by one workgroup: phase p relaxes the rows at positions
// first[p] .. first[p] + count[p], the barrier between phases making each phase's x visible to the next (the
// workgroup shares one CU's cache), as separate launches would. Saves a launch per colour where a
// colour is too small to fill the device.
constexpr int kRelaxRunThreads = 1024;
constexpr int kRelaxRunMax = 32;
// Colours of at most this many cells (two passes of the run's workgroup) join a run.
constexpr int kRelaxRunCells = 2 * (kRelaxRunThreads / kWave) * (kWave / 5);
struct ImplRun {
  int first[kRelaxRunMax];
  int count[kRelaxRunMax];
  int phases;
};

__global__ __launch_bounds__(kRelaxRunThreads) void k_impl_relax_run(ImplRows rows, ImplRun run,
                                                                     const double* __restrict__ offd,
                                                                     const double* __restrict__ inverse,
                                                                     const double* __restrict__ b,
                                                                     double* __restrict__ x) {
  constexpr int per_pass = (kRelaxRunThreads / kWave) * kRelaxGroups;
  const int lane = threadIdx.x % kWave;
  const int g = lane / 5;
  const int slot = static_cast<int>(threadIdx.x / kWave) * kRelaxGroups + g;
  for (int p = 0; p < run.phases; ++p) {
    // wave-uniform trip count: every lane reaches the shuffles
    for (int t0 = 0; t0 < run.count[p]; t0 += per_pass) {
      const int t = t0 + slot;
      const bool active = g < kRelaxGroups && t < run.count[p];
      impl_relax_rows(rows, run.first[p] + t, active, offd, inverse, b, x);
    }
    __syncthreads();
  }
}

// BlockMatrix::multiply: y_c = D_c x_c + sum_f A_cf x_other, the diagonal block first; five lanes
// per row position (12 per wave), lane q summing row q in the one-thread order.
__global__ __launch_bounds__(kBlock) void k_impl_multiply(ImplRows rows, int n, const double* __restrict__ diag,
                                                          const double* __restrict__ offd,
                                                          const double* __restrict__ x, double* __restrict__ y) {
  const int lane = threadIdx.x % kWave;
  const int g = lane / 5, q = lane % 5;
  const int p = (blockIdx.x * (kBlock / kWave) + static_cast<int>(threadIdx.x / kWave)) * kRelaxGroups + g;
  if (g >= kRelaxGroups || p >= n) return;
  const long long c = rows.cell[p];
  const double* dp = diag + 25LL * p + 5 * q;
  double r = 0.0;
#pragma unroll
  for (int j = 0; j < 5; ++j) r += dp[j] * x[5 * c + j];
  for (int k = rows.off[p]; k < rows.off[p + 1]; ++k) {
    const double* a = offd + 25LL * k + 5 * q;
    const long long o = rows.other[k];
#pragma unroll
    for (int j = 0; j < 5; ++j) r += a[j] * x[5 * o + j];
  }
  y[5 * c + q] = r;
}

// A block's exact-sum accumulator (carries resolved) added into the global words.
__device__ __forceinline__ void impl_flush(long long* acc, long long* __restrict__ words) {
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int l = 0; l + 1 < kLimbs; ++l) {
      const long long carry = acc[l] >> 32;
      acc[l] -= carry * (1LL << 32);
      acc[l + 1] += carry;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kWords; t += kBlock) {
    const long long w = acc[t];
    if (w != 0) atomicAdd(reinterpret_cast<unsigned long long*>(words + t), static_cast<unsigned long long>(w));
  }
}

// sum_i u_i v_i over m values as exact words (added into `words`, cleared beforehand): each product
// rounded to FP64, then added exactly. Grid-stride with wave-uniform trip counts.
__global__ __launch_bounds__(kBlock) void k_impl_dot(const double* __restrict__ u, const double* __restrict__ v,
                                                     long long m, long long* __restrict__ words) {
  __shared__ long long acc[kWords];
  for (int t = threadIdx.x; t < kWords; t += kBlock) acc[t] = 0;
  __syncthreads();
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const long long stride = static_cast<long long>(gridDim.x) * kBlock;
  for (long long base = static_cast<long long>(blockIdx.x) * kBlock + wave * kWave; base < m; base += stride) {
    const long long i = base + lane;
    const bool active = i < m;
    const double t = active ? u[i] * v[i] : 0.0;
    exact_add_wave(acc, t, active);
  }
  impl_flush(acc, words);
}

// The GCR's modified Gram-Schmidt step fused with the inner products that follow it: when zj is
// given, beta = (w, w_j) / (w_j, w_j) from the words of the last inner product and the device's
// (w_j, w_j) (each block rounds them with exact_value), z -= beta zj and w -= beta wj; then sum
// w_i p_i into wp and, when q is given, sum w_i q_i into wq, each over the updated w. The sums are
// exact, so fusing them changes no bit.
__global__ __launch_bounds__(kBlock) void k_impl_orth_dots(double* __restrict__ z, double* __restrict__ w,
                                                           const double* __restrict__ zj,
                                                           const double* __restrict__ wj, long long m,
                                                           const long long* __restrict__ beta_words,
                                                           const double* __restrict__ beta_den,
                                                           const double* __restrict__ p, long long* __restrict__ wp,
                                                           const double* __restrict__ q, long long* __restrict__ wq) {
  __shared__ long long acc[2][kWords];
  __shared__ double beta_sh;
  for (int t = threadIdx.x; t < 2 * kWords; t += kBlock) acc[t / kWords][t % kWords] = 0;
  if (zj != nullptr && threadIdx.x == 0) beta_sh = exact_value(beta_words) / *beta_den;
  __syncthreads();
  const double beta = zj != nullptr ? beta_sh : 0.0;
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const long long stride = static_cast<long long>(gridDim.x) * kBlock;
  for (long long base = static_cast<long long>(blockIdx.x) * kBlock + wave * kWave; base < m; base += stride) {
    const long long i = base + lane;
    const bool active = i < m;
    double wi = 0.0;
    if (active) {
      wi = w[i];
      if (zj != nullptr) {
        z[i] -= beta * zj[i];
        wi -= beta * wj[i];
        w[i] = wi;
      }
    }
    exact_add_wave(acc[0], active ? wi * p[i] : 0.0, active);
    if (q != nullptr) exact_add_wave(acc[1], active ? wi * q[i] : 0.0, active);
  }
  impl_flush(acc[0], wp);
  if (q != nullptr) impl_flush(acc[1], wq);
}

// The GCR's update fused with the scalars before it and the inner product after it: each block
// rounds (w, w) and alpha = (r, w) / (w, w) from their words (block 0 stores them for the host and
// the next directions), then, when (w, w) > 0 (else the GCR stops before its update), x += sa z
// (sa = scale * alpha, scale the residual's largest magnitude from its bits) and r -= alpha w;
// then sum r_i r_i into words.
__global__ __launch_bounds__(kBlock) void k_impl_advance_dot(double* __restrict__ x, double* __restrict__ r,
                                                             const double* __restrict__ z,
                                                             const double* __restrict__ w, long long m,
                                                             const unsigned long long* __restrict__ scale_bits,
                                                             const long long* __restrict__ ww_words,
                                                             const long long* __restrict__ rw_words,
                                                             double* __restrict__ ww_out, double* __restrict__ alpha_out,
                                                             long long* __restrict__ words) {
  __shared__ long long acc[kWords];
  __shared__ double ww_sh, alpha_sh;
  for (int t = threadIdx.x; t < kWords; t += kBlock) acc[t] = 0;
  if (threadIdx.x == 0) {
    const double ww = exact_value(ww_words);
    const double alpha = exact_value(rw_words) / ww;
    ww_sh = ww;
    alpha_sh = alpha;
    if (blockIdx.x == 0) {
      *ww_out = ww;
      *alpha_out = alpha;
    }
  }
  __syncthreads();
  const bool update = ww_sh > 0.0;
  const double alpha = alpha_sh;
  const double sa = __longlong_as_double(static_cast<long long>(*scale_bits)) * alpha;
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const long long stride = static_cast<long long>(gridDim.x) * kBlock;
  for (long long base = static_cast<long long>(blockIdx.x) * kBlock + wave * kWave; base < m; base += stride) {
    const long long i = base + lane;
    const bool active = i < m;
    double ri = 0.0;
    if (active) {
      ri = r[i];
      if (update) {
        x[i] += sa * z[i];
        ri -= alpha * w[i];
        r[i] = ri;
      }
    }
    exact_add_wave(acc, active ? ri * ri : 0.0, active);
  }
  impl_flush(acc, words);
}

// The largest |u_i| as the bits of a non-negative double (their integer order is their order);
// a NaN entry leaves a NaN's bits, above every number's.
__global__ __launch_bounds__(kBlock) void k_impl_maxabs(const double* __restrict__ u, long long m,
                                                        unsigned long long* __restrict__ out) {
  unsigned long long best = 0;
  const long long stride = static_cast<long long>(gridDim.x) * kBlock;
  for (long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; i < m; i += stride) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(fabs(u[i])));
    best = b > best ? b : best;
  }
  if (best != 0) atomicMax(out, best);
}

// Clears n exact-sum word sets and sets *bits (the admissible fraction's minimum) to 1.0's bits,
// *singular to INT32_MAX and the largest-entry bits to 0.
__global__ void k_impl_clear(long long* __restrict__ words, int n, unsigned long long* __restrict__ bits,
                             int* __restrict__ singular) {
  for (int t = threadIdx.x; t < n * kWords; t += blockDim.x) words[t] = 0;
  if (threadIdx.x == 0) {
    bits[0] = static_cast<unsigned long long>(__double_as_longlong(1.0));
    bits[1] = 0;
    *singular = INT32_MAX;
  }
}

// EulerPhysics::pressure(c, a): the pressure of u + a du in FP64.
__device__ __forceinline__ double impl_pressure(const double* v0, const double* dv, double a, double gamma) {
  double v[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) v[k] = v0[k] + a * dv[k];
  return (gamma - 1.0) * (v[4] - 0.5 * (v[1] * v[1] + v[2] * v[2] + v[3] * v[3]) / v[0]);
}

constexpr double kImplMaxDecrease = 0.5;
constexpr double kImplMinStep = 1e-6;

// Each cell's admissible fraction (its density bound, then halved until its pressure bound holds),
// and their least as the bits of a non-negative double (atomicMin of the bits).
__global__ __launch_bounds__(kBlock) void k_impl_admissible(const double* __restrict__ u,
                                                            const double* __restrict__ du, int n, double gamma,
                                                            unsigned long long* __restrict__ bits) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  unsigned long long key = ~0ULL;
  if (c < n) {
    double v0[5], dv[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      v0[k] = u[static_cast<long long>(k) * n + c];
      dv[k] = du[5LL * c + k];
    }
    double a = 1.0;
    if (dv[0] < 0.0) {
      const double b = kImplMaxDecrease * v0[0] / -dv[0];
      a = b < a ? b : a;  // std::min(a, b): b when b < a
    }
    const double p = impl_pressure(v0, dv, 0.0, gamma);
    while (a >= kImplMinStep && !(impl_pressure(v0, dv, a, gamma) >= (1.0 - kImplMaxDecrease) * p)) a *= 0.5;
    key = static_cast<unsigned long long>(__double_as_longlong(a));
  }
  key = wave_min(key);
  if ((threadIdx.x % kWave) == 0 && key != ~0ULL) atomicMin(bits, key);
}

// The Krylov operator's boundary term (EulerPhysics::add_boundary_gradient_term with
// Euler::gradient_tangent, physics/src/euler_physics.cpp:1300-1320, physics/src/euler.cpp:1081-1119):
// y += B v, the boundary faces' fluxes' derivative through the reconstruction's gradients, which
// the assembled blocks drop. One thread per cell owning a boundary face (`cells`, ascending): the
// primitive tangents dW = dW/dq v of the cell and its stencil cells (the direction rounded to the
// compute type, the products summed in it; k_impl_tangents, for every cell), the gradient's tangent sum_k (dW_k - dW_c) w_k (FP64
// sums, boundary values held), then over its boundary faces in cell_faces order, where the
// reconstruction did not fall back to the cell value, bnd_jac applied to phi (grad_tan . r), and on
// warped slip and wall faces the wall-warp term's tangent phi_p M^T grad_tan p.
// dwdu cell-major (25 c + 5 i + j), v cell-major (5 c + j).
template <class C>
__device__ __forceinline__ void impl_tangent(const double* __restrict__ dwdu, const double* __restrict__ v, long long c,
                                             C* dw) {
  C vc[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) vc[j] = static_cast<C>(v[5 * c + j]);
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    C t = C(0);
#pragma unroll
    for (int j = 0; j < 5; ++j) t = t + static_cast<C>(dwdu[25 * c + 5 * i + j]) * vc[j];
    dw[i] = t;
  }
}

// Every cell's primitive tangent dW/dq v (impl_tangent), 5 x n in the compute type, for the
// boundary term's stencils.
template <class C>
__global__ __launch_bounds__(kBlock) void k_impl_tangents(const double* __restrict__ dwdu, const double* __restrict__ v,
                                                          int n, C* __restrict__ out) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= n) return;
  C dw[5];
  impl_tangent<C>(dwdu, v, c, dw);
#pragma unroll
  for (int i = 0; i < 5; ++i) out[static_cast<long long>(i) * n + c] = dw[i];
}

template <class C>
__global__ __launch_bounds__(kBlock) void k_impl_bgrad(Dev<C> d, const int* __restrict__ cells, int count,
                                                       const C* __restrict__ tangents, const double* __restrict__ bj,
                                                       double* __restrict__ y) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  const int c = cells[t];
  const long long n = d.nc;
  const Params<C>& pa = d.ptab[cell_case(d, c)];
  C dwc[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) dwc[i] = tangents[i * n + c];
  double acc[15];
#pragma unroll
  for (int q = 0; q < 15; ++q) acc[q] = 0.0;
  const long long base = d.soff[c / kWave] + (c % kWave);
  const int len = d.slen[c];
  for (int j = 0; j < len; ++j) {
    const long long k = base + static_cast<long long>(j) * kWave;
    const int e = d.sidx[k];
    const C w0 = d.sw0[k], w1 = d.sw1[k], w2 = d.sw2[k];
    C dwn[5] = {C(0), C(0), C(0), C(0), C(0)};
    if (e >= 0) {
#pragma unroll
      for (int i = 0; i < 5; ++i) dwn[i] = tangents[i * n + e];
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      const C dd = dwn[i] - dwc[i];
      acc[3 * i] += static_cast<double>(dd * w0);
      acc[3 * i + 1] += static_cast<double>(dd * w1);
      acc[3 * i + 2] += static_cast<double>(dd * w2);
    }
  }
  C gt[15];
#pragma unroll
  for (int q = 0; q < 15; ++q) gt[q] = static_cast<C>(acc[q]);
  const W5<C> w = load_w(d.prim, c);
  const C wc[5] = {w.r, w.u, w.v, w.w, w.p};
  C rec[20];
  load_rec(d.rec, c, rec);
  const C* gc = rec;
  const C* pc = rec + 15;
  double dres[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int k = d.cfo[c]; k < d.cfo[c + 1]; ++k) {
    const int f = d.inc[k] & 0x7fffffff;
    if (f < d.ni) continue;
    const int b = f - d.ni;
    C g8[8];
    load_n<C, 8>(d.fgeo + 8LL * f, g8);
    const C r[3] = {g8[4], g8[5], g8[6]};
    const C d00 = gc[0] * r[0] + gc[1] * r[1] + gc[2] * r[2];
    const C d04 = gc[12] * r[0] + gc[13] * r[1] + gc[14] * r[2];
    const bool fell_back = !((wc[0] + pc[0] * d00) > C(0.0) && (wc[4] + pc[4] * d04) + pa.pref > C(0.0));
    if (fell_back) continue;
    C dwf[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) dwf[i] = pc[i] * (gt[3 * i] * r[0] + gt[3 * i + 1] * r[1] + gt[3 * i + 2] * r[2]);
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < 5; ++j) s += bj[static_cast<long long>(5 * i + j) * d.nb + b] * static_cast<double>(dwf[j]);
      dres[i] += s;
    }
    const int kind = d.bkind[b];
    if (d.moment != nullptr && (kind == 2 || kind >= 5)) {
      C m9[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) m9[q] = d.moment[static_cast<long long>(q) * d.nf + f];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        dres[1 + i] += static_cast<double>(pc[4] * (m9[i] * gt[12] + m9[3 + i] * gt[13] + m9[6 + i] * gt[14]));
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 5; ++i) y[5 * static_cast<long long>(c) + i] += dres[i];
}

// ---- the GCR's vector operations (solvers/src/gcr.cpp), m values each

// y = -x, x as the residual holds it (value q of cell c at q n + c), y cell-major (at 5 c + q).
__global__ __launch_bounds__(kBlock) void k_impl_negate(const double* __restrict__ x, double* __restrict__ y,
                                                        long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < 5 * n) y[i] = -x[(i % 5) * n + i / 5];
}

__global__ __launch_bounds__(kBlock) void k_impl_zero(double* __restrict__ x, long long m) {
  const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < m) x[i] = 0.0;
}

// r /= s (mode 0) or r *= s (mode 1)
__global__ __launch_bounds__(kBlock) void k_impl_scale(double* __restrict__ r, long long m, double s, int mode) {
  const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < m) r[i] = mode == 0 ? r[i] / s : r[i] * s;
}

// *out = the exact sum's value (exact_value, the host's rounding), divided by *den when given: the
// GCR's scalars stay on the device, so its vector updates need no host round trip.
__global__ void k_impl_scalar(const long long* __restrict__ words, const double* __restrict__ den,
                              double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double v = exact_value(words);
  *out = den != nullptr ? v / *den : v;
}

// u += alpha du, u as the state holds it (value q of cell c at q n + c), du cell-major.
__global__ __launch_bounds__(kBlock) void k_impl_update(double* __restrict__ u, const double* __restrict__ du,
                                                        long long n, double alpha) {
  const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < 5 * n) u[i] += alpha * du[5 * (i % n) + i / n];
}

}  // namespace mxe
