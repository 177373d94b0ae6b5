// HIP kernels of the explicit Euler residual loop for gfx950 (wave64).
//
// Device layout (the boundary of include/mxeuler.h converts to it once, at creation):
//   cell records   prim  8 x C per cell (rho, u, v, w, p - p_ref, 3 pad): one 32 B record per
//                        gather in fp32, so a stencil or face neighbour costs one cache line;
//                  rec   20 x C per cell: the gradient (15, row 3 i + j) then the limiter (5);
//   face records   fgeo  8 x C per face (n_x, n_y, n_z, |A|, r_left, 0) and frr 4 x C per
//                        interior face (r_right, 0), r = rel(x_f, x_c) already rounded to C;
//   stencil        sliced by wave: entry j of cell 64 s + l at soff[s] + 64 j + l, so the 64
//                  lanes of a wave read 64 consecutive indices and weights per entry;
//   incidences     CSR per cell, face index with bit 31 set where the cell is the neighbour;
//   conserved u, stage and res stay FP64 structure of arrays, k * n + c, as at the boundary.
// Every loop keeps the reference's summation order, so FP64 results can match the host bit for
// bit: the store-gather order of its device path (codegen/src/autan_codegen/emit_gpu.py:189-238).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "bodies.hpp"

namespace mxe {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kLimbs = 68;        // exact-sum limbs of 32 bits, limb i weighs 2^(32 i - 1088)
constexpr int kWords = kLimbs + 3;  // then the counts of +inf, -inf and NaN terms
constexpr int kSums = 5;
// LDS bytes of a block's exact-sum words, rounded up to 16 so that what follows them stays 16 B
// aligned.
constexpr int kSumsLds = (kSums * kWords * 8 + 15) & ~15;

// Results a step leaves for the host: one copy back per step.
struct Small {
  long long words[kSums][kWords];
  unsigned long long hkey;  // ordered key of the least time step
  int invalid_state;        // least cell with an invalid state, INT32_MAX for none
  int invalid_residual;     // least cell with a non-finite residual
};

template <class C>
struct Params {
  C g, rgas, pref, minf;
  C ur[5];
  C scale[5];
  double ref[5];
  double p_ref;  // FP64 p_ref for the validity check, rounded there as the reference rounds it
  double gamma;
  long long cell0;     // the case's first cell: invalid cells are reported relative to it
  long long cell_end;  // one past its last real cell; the filler cells after it keep their state
};

template <class C>
struct Dev {
  int nc, nf, ni, nb;
  const int2* lr;      // ni: owner, neighbour
  const int* bowner;   // nb
  const int* cfo;      // nc + 1
  const int* inc;      // incidences
  const long long* soff;  // per slice of 64 cells
  const int* slen;     // per cell
  const int* sidx;     // sliced entries: cell, or -(b + 1) for boundary face ni + b
  const C* sw0;
  const C* sw1;
  const C* sw2;
  const C* fgeo;    // 8 per face
  const C* frr;     // 4 per interior face
  const C* moment;  // 9 x nf SoA, or null (planar)
  const C* vol;     // nc
  const double* vol64;
  const int* bkind;  // nb codes 0..5
  const C* bc;       // 8 per boundary face
  const C* source;   // 5 x nc SoA or null
  C* prim;           // 8 per cell
  C* prim_b;         // 8 per boundary face
  C* rec;            // 20 per cell
  C* fflux;          // 8 per face
  double* res;       // 5 x nc
  C* lam;            // nc
  double* dt;        // nc
  Small* small;  // one per case
  // Cases: meshes stepped together in one index space, each case's cells starting on a 256-cell
  // block (one case: ncase = 1 and no tables). Case k's parameters are ptab[k]; bcase gives the
  // case of each 256-cell block, fcase of each face block, fcbeg each case's first face block;
  // cflags bit 0 marks a case active (stepped), bit 1 at second order (null: all active).
  const Params<C>* ptab;
  const int* bcase;
  const int* fcase;
  const int* fcbeg;
  const int* cflags;
  // Launches over the active cases' blocks only: grid position p (after the XCD mapping) takes
  // logical cell block cblk[p] and face block fblk[p], -1 padding the grid to a multiple of the
  // XCD count; null: position p is block p.
  const int* cblk;
  const int* fblk;
  int ncase;
  int nslot;  // exact-sum slots per case
  // Block-local faces: the faces of the cells of logical block b are the slots sbeg[b] ..
  // sbeg[b + 1] (ascending face order; a face on two blocks has a slot in each); incidence k of a
  // cell reads its face's flux from LDS slot islot[k] & 0x3fff, bit 14 set where the cell is the
  // face's neighbour, bit 15 on boundary faces.
  const int* cbeg;  // cells of face block b: [cbeg[b], cbeg[b + 1])
  const int* sbeg;
  const unsigned short* islot;
  const C* srec;  // per slot: the face's 12 values (face_core)
  const C* smom;  // per slot: the face's nine moments (non-planar meshes)
  int max_slots;
  long long* nslots;  // kNormSlots x kSums x kWords exact-sum words, spread to avoid contention
  // Staged gradient: block b's table of cells then boundary faces tent[tbeg[b] .. tbeg[b + 1]),
  // sloc the table position of every sliced stencil entry, lens = stencil length | incidences
  // << 16, and the limiter's offsets of each incidence (ir0, ir1, ir2) sliced from ioff.
  const int* tbeg;
  const int* tent;
  const unsigned short* sloc;
  const int* lens;
  const long long* ioff;
  const C* ir0;
  const C* ir1;
  const C* ir2;
};

constexpr int kNormSlots = 64;  // exact-sum slots of one case
constexpr int kCaseSlots = 8;   // per case when several cases run together
constexpr int kMaxSlots = 2 * kBlock;  // faces per face block in FP64 (twice that in FP32)

// Workgroups are handed to the eight XCDs in turn (workgroup b to XCD b mod 8), each XCD with its
// own L2. Remapped so that XCD k takes a contiguous run of logical blocks: the cells an XCD works
// on at once are neighbours, and their stencil and face neighbours share its L2. The grid is
// launched with xcd_grid(blocks) workgroups; the logical blocks past `blocks` do nothing.
constexpr int kXcds = 8;
__host__ __device__ inline int xcd_grid(int blocks) { return kXcds * ((blocks + kXcds - 1) / kXcds); }
__device__ __forceinline__ int xcd_block() {
  const int per = gridDim.x / kXcds;
  return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}
__device__ __forceinline__ int xcd_index() { return xcd_block() * kBlock + threadIdx.x; }

// The logical cell block and face block of this workgroup (-1: none).
template <class C>
__device__ __forceinline__ int cell_block(const Dev<C>& d) {
  const int p = xcd_block();
  return d.cblk != nullptr ? d.cblk[p] : p;
}
template <class C>
__device__ __forceinline__ int face_block(const Dev<C>& d) {
  const int p = xcd_block();
  return d.fblk != nullptr ? d.fblk[p] : p;
}

template <class C>
__device__ __forceinline__ int block_case(const Dev<C>& d, int blk) {
  return d.bcase != nullptr ? d.bcase[blk] : 0;
}
template <class C>
__device__ __forceinline__ int cell_case(const Dev<C>& d, int c) {
  return d.bcase != nullptr ? d.bcase[c / kBlock] : 0;
}
template <class C>
__device__ __forceinline__ bool case_active(const Dev<C>& d, int cs) {
  return d.cflags == nullptr || (d.cflags[cs] & 1) != 0;
}
template <class C>
__device__ __forceinline__ bool case_second(const Dev<C>& d, int cs) {
  return d.cflags == nullptr || (d.cflags[cs] & 2) != 0;
}

// A cell's (or boundary face's) primitive record: rho, u, v, w, p - p_ref, kPrimStride values per
// record and no padding, so that the gathers of neighbouring cells share cache lines.
constexpr int kPrimStride = 5;

template <class C>
__device__ __forceinline__ W5<C> load_w(const C* __restrict__ p, long long i) {
  const C* q = p + kPrimStride * i;
  return {q[0], q[1], q[2], q[3], q[4]};
}

template <class C>
__device__ __forceinline__ void store_w(C* __restrict__ p, long long i, const W5<C>& w) {
  C* q = p + kPrimStride * i;
  q[0] = w.r;
  q[1] = w.u;
  q[2] = w.v;
  q[3] = w.w;
  q[4] = w.p;
}

// N consecutive values at a 16 B aligned address.
template <class C, int N>
__device__ __forceinline__ void load_n(const C* __restrict__ p, C* out) {
  if constexpr (std::is_same_v<C, float>) {
    static_assert(N % 4 == 0);
#pragma unroll
    for (int i = 0; i < N / 4; ++i) {
      const float4 v = reinterpret_cast<const float4*>(p)[i];
      out[4 * i] = v.x;
      out[4 * i + 1] = v.y;
      out[4 * i + 2] = v.z;
      out[4 * i + 3] = v.w;
    }
  } else {
    static_assert(N % 2 == 0);
#pragma unroll
    for (int i = 0; i < N / 2; ++i) {
      const double2 v = reinterpret_cast<const double2*>(p)[i];
      out[2 * i] = v.x;
      out[2 * i + 1] = v.y;
    }
  }
}

template <class C, int N>
__device__ __forceinline__ void store_n(C* __restrict__ p, const C* in) {
  if constexpr (std::is_same_v<C, float>) {
    static_assert(N % 4 == 0);
#pragma unroll
    for (int i = 0; i < N / 4; ++i) {
      reinterpret_cast<float4*>(p)[i] = make_float4(in[4 * i], in[4 * i + 1], in[4 * i + 2], in[4 * i + 3]);
    }
  } else {
    static_assert(N % 2 == 0);
#pragma unroll
    for (int i = 0; i < N / 2; ++i) reinterpret_cast<double2*>(p)[i] = make_double2(in[2 * i], in[2 * i + 1]);
  }
}

template <class C>
__device__ __forceinline__ C infinity() {
  if constexpr (std::is_same_v<C, float>) {
    return __builtin_huge_valf();
  } else {
    return __builtin_huge_val();
  }
}

// ------------------------------------------------------------------------ state and primitives

// Euler::state then the primitives loop (physics/src/euler.cpp:925-932; kernels/euler.py:428-433):
// s = State(u - U_ref), W = to_primitive(s).
template <class C>
__device__ __forceinline__ W5<C> primitive_of(const double* __restrict__ u, int nc, int c, const Params<C>& pa) {
  C dq[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) dq[k] = static_cast<C>(u[static_cast<long long>(k) * nc + c] - pa.ref[k]);
  return to_primitive(dq, pa.ur, pa.g);
}

template <class C>
__global__ __launch_bounds__(kBlock) void k_prim(Dev<C> d, const double* __restrict__ u) {
  const int blk = cell_block(d);
  if (blk < 0) return;
  const int c = blk * kBlock + threadIdx.x;
  if (c >= d.nc) return;
  const int cs = block_case(d, blk);
  if (!case_active(d, cs)) return;
  store_w(d.prim, c, primitive_of(u, d.nc, c, d.ptab[cs]));
}

// boundary_states (kernels/euler.py:442-466): the boundary value of every boundary face from its
// owner's primitives.
template <class C>
__global__ __launch_bounds__(kBlock) void k_bstate(Dev<C> d) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= d.nb) return;
  const int owner = d.bowner[b];
  const int cs = cell_case(d, owner);
  if (!case_active(d, cs)) return;
  const Params<C>& pa = d.ptab[cs];
  const W5<C> w = load_w(d.prim, owner);
  C geo[8], data[8];
  load_n<C, 8>(d.fgeo + 8LL * (d.ni + b), geo);
  load_n<C, 8>(d.bc + 8LL * b, data);
  const V3<C> n{geo[0], geo[1], geo[2]};
  store_w(d.prim_b, b, boundary_value(w, d.bkind[b], data, n, pa.g, pa.rgas, pa.pref));
}

// ------------------------------------------------------------------------ gradient and limiter

// gradients then limiter (kernels/euler.py:530-549, 563-608; reduction body
// codegen/src/autan_codegen/emit_cpu.py:126-182), fused per cell: the bounds stay in registers.
// Sums accumulate in FP64 and are stored in C; bounds and the limiter's minimum are taken in C,
// ties keeping the first entry. Without the limiter phi = 1 (second order unlimited).
template <class C, bool LIMITER>
__global__ __launch_bounds__(kBlock) void k_grad(Dev<C> d) {
  const int blk = cell_block(d);
  if (blk < 0) return;
  const int c = blk * kBlock + threadIdx.x;
  if (c >= d.nc) return;
  const int cs = block_case(d, blk);
  if (!case_active(d, cs) || !case_second(d, cs)) return;
  const Params<C>& pa = d.ptab[cs];
  const W5<C> wc = load_w(d.prim, c);
  const C wcv[5] = {wc.r, wc.u, wc.v, wc.w, wc.p};
  double acc[15];
#pragma unroll
  for (int q = 0; q < 15; ++q) acc[q] = 0.0;
  C lo[5], hi[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    lo[q] = infinity<C>();
    hi[q] = -infinity<C>();
  }
  const long long base = d.soff[c / kWave] + (c % kWave);
  const int len = d.slen[c];
  for (int j = 0; j < len; ++j) {
    const long long k = base + static_cast<long long>(j) * kWave;
    const int e = d.sidx[k];
    const C w0 = d.sw0[k], w1 = d.sw1[k], w2 = d.sw2[k];
    const W5<C> wn = e >= 0 ? load_w(d.prim, e) : load_w(d.prim_b, -e - 1);
    const C wnv[5] = {wn.r, wn.u, wn.v, wn.w, wn.p};
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      const C dd = wnv[i] - wcv[i];
      acc[3 * i] += static_cast<double>(dd * w0);
      acc[3 * i + 1] += static_cast<double>(dd * w1);
      acc[3 * i + 2] += static_cast<double>(dd * w2);
      if (wnv[i] < lo[i]) lo[i] = wnv[i];
      if (wnv[i] > hi[i]) hi[i] = wnv[i];
    }
  }
  C out[20];
#pragma unroll
  for (int q = 0; q < 15; ++q) out[q] = static_cast<C>(acc[q]);
  if constexpr (LIMITER) {
    C phi[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      lo[i] = smin(lo[i], wcv[i]);
      hi[i] = smax(hi[i], wcv[i]);
      phi[i] = infinity<C>();
    }
    const C vol = d.vol[c];
    C eps2[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) eps2[i] = pa.scale[i] * vol;
    for (int k = d.cfo[c]; k < d.cfo[c + 1]; ++k) {
      const int e = d.inc[k];
      const int f = e & 0x7fffffff;
      C r[4];
      if (e < 0) {
        load_n<C, 4>(d.frr + 4LL * f, r);
      } else {
        load_n<C, 4>(d.fgeo + 8LL * f + 4, r);
      }
#pragma unroll
      for (int i = 0; i < 5; ++i) {
        const C dd = out[3 * i] * r[0] + out[3 * i + 1] * r[1] + out[3 * i + 2] * r[2];
        const C dmax = hi[i] - wcv[i], dmin = lo[i] - wcv[i];
        const C dp = dd > C(0.0) ? dmax : dmin;
        const C v = (dp * dp + eps2[i] + C(2.0) * dd * dp) / (dp * dp + C(2.0) * dd * dd + dd * dp + eps2[i]);
        if (v < phi[i]) phi[i] = v;
      }
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) out[15 + i] = smin(phi[i], C(1.0));
  } else {
#pragma unroll
    for (int i = 0; i < 5; ++i) out[15 + i] = C(1.0);
  }
  store_n<C, 20>(d.rec + 20LL * c, out);
}

// The same gradient and limiter with the block's stencil values staged in LDS: the block's
// table of cells and boundary faces is read once, in ascending order (runs of consecutive
// records), then every stencil entry gathers its primitives from LDS through a 16-bit position.
// Weights, positions and the limiter's face offsets stream in the wave-sliced layout. The sums,
// bounds and minima run in the same order as k_grad.
// OCC > 0 asks the compiler for OCC waves per SIMD (registers capped accordingly): a variant for
// the autotuner, same bits.
// A load of data read once per launch: with NT, marked non-temporal (streamed past the caches'
// replacement, leaving them to the data other blocks read again); the same value either way.
template <bool NT, class T>
__device__ __forceinline__ T stream_load(const T* p) {
  if constexpr (NT) {
    return __builtin_nontemporal_load(p);
  } else {
    return *p;
  }
}

template <class C, bool LIMITER, int OCC = 0, bool NT = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1))) void k_grad_lds(Dev<C> d) {
  extern __shared__ __align__(16) unsigned char smem[];
  C* tab = reinterpret_cast<C*>(smem);
  const int blk = cell_block(d);
  if (blk < 0) return;
  const int cs = block_case(d, blk);
  if (!case_active(d, cs) || !case_second(d, cs)) return;
  const Params<C>& pa = d.ptab[cs];
  const int t0 = d.tbeg[blk], tn = d.tbeg[blk + 1] - t0;
  for (int j = threadIdx.x; j < tn; j += kBlock) {
    const int e = d.tent[t0 + j];
    const W5<C> w = e >= 0 ? load_w(d.prim, e) : load_w(d.prim_b, -e - 1);
    tab[j] = w.r;
    tab[tn + j] = w.u;
    tab[2 * tn + j] = w.v;
    tab[3 * tn + j] = w.w;
    tab[4 * tn + j] = w.p;
  }
  __syncthreads();
  const int c = blk * kBlock + threadIdx.x;
  if (c >= d.nc) return;
  const W5<C> wc = load_w(d.prim, c);
  const C wcv[5] = {wc.r, wc.u, wc.v, wc.w, wc.p};
  double acc[15];
#pragma unroll
  for (int q = 0; q < 15; ++q) acc[q] = 0.0;
  C lo[5], hi[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    lo[q] = infinity<C>();
    hi[q] = -infinity<C>();
  }
  const int lens = d.lens[c];
  const int len = lens & 0xffff;
  const long long base = d.soff[c / kWave] + (c % kWave);
  // Entry j + 1's position and weights are loaded while entry j is summed (its loads' latency
  // hidden behind the arithmetic); the sums run in entry order as before.
  int l_next = 0;
  C w0_next = C(0), w1_next = C(0), w2_next = C(0);
  if (len > 0) {
    l_next = stream_load<NT>(d.sloc + base);
    w0_next = stream_load<NT>(d.sw0 + base);
    w1_next = stream_load<NT>(d.sw1 + base);
    w2_next = stream_load<NT>(d.sw2 + base);
  }
  for (int j = 0; j < len; ++j) {
    const int l = l_next;
    const C w0 = w0_next, w1 = w1_next, w2 = w2_next;
    if (j + 1 < len) {
      const long long k = base + static_cast<long long>(j + 1) * kWave;
      l_next = stream_load<NT>(d.sloc + k);
      w0_next = stream_load<NT>(d.sw0 + k);
      w1_next = stream_load<NT>(d.sw1 + k);
      w2_next = stream_load<NT>(d.sw2 + k);
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      const C wn = tab[i * tn + l];
      const C dd = wn - wcv[i];
      acc[3 * i] += static_cast<double>(dd * w0);
      acc[3 * i + 1] += static_cast<double>(dd * w1);
      acc[3 * i + 2] += static_cast<double>(dd * w2);
      if (wn < lo[i]) lo[i] = wn;
      if (wn > hi[i]) hi[i] = wn;
    }
  }
  C out[20];
#pragma unroll
  for (int q = 0; q < 15; ++q) out[q] = static_cast<C>(acc[q]);
  if constexpr (LIMITER) {
    C phi[5], eps2[5];
    const C vol = d.vol[c];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      lo[i] = smin(lo[i], wcv[i]);
      hi[i] = smax(hi[i], wcv[i]);
      phi[i] = infinity<C>();
      eps2[i] = pa.scale[i] * vol;
    }
    const int ilen = lens >> 16;
    const long long ib = d.ioff[c / kWave] + (c % kWave);
    C r0_next = C(0), r1_next = C(0), r2_next = C(0);
    if (ilen > 0) {
      r0_next = stream_load<NT>(d.ir0 + ib);
      r1_next = stream_load<NT>(d.ir1 + ib);
      r2_next = stream_load<NT>(d.ir2 + ib);
    }
    for (int j = 0; j < ilen; ++j) {
      const C r0 = r0_next, r1 = r1_next, r2 = r2_next;
      if (j + 1 < ilen) {
        const long long k = ib + static_cast<long long>(j + 1) * kWave;
        r0_next = stream_load<NT>(d.ir0 + k);
        r1_next = stream_load<NT>(d.ir1 + k);
        r2_next = stream_load<NT>(d.ir2 + k);
      }
#pragma unroll
      for (int i = 0; i < 5; ++i) {
        const C dd = out[3 * i] * r0 + out[3 * i + 1] * r1 + out[3 * i + 2] * r2;
        const C dmax = hi[i] - wcv[i], dmin = lo[i] - wcv[i];
        const C dp = dd > C(0.0) ? dmax : dmin;
        const C v = (dp * dp + eps2[i] + C(2.0) * dd * dp) / (dp * dp + C(2.0) * dd * dd + dd * dp + eps2[i]);
        if (v < phi[i]) phi[i] = v;
      }
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) out[15 + i] = smin(phi[i], C(1.0));
  } else {
#pragma unroll
    for (int i = 0; i < 5; ++i) out[15 + i] = C(1.0);
  }
  store_n<C, 20>(d.rec + 20LL * c, out);
}

// ------------------------------------------------------------------------ face fluxes

template <class C>
__device__ __forceinline__ void load_rec(const C* __restrict__ rec, int c, C* g) {
  load_n<C, 20>(rec + 20LL * c, g);
}

// A face as the flux bodies read it, 12 values: unit normal and area, x_f - x_c of the left cell
// and the left cell's index, x_f - x_c of the right cell and its index (the indices' bits held in
// the compute type; a right index -(b + 1) marks boundary face ni + b, whose right offsets are 0).
template <class C>
__device__ __forceinline__ C index_bits(int i) {
  if constexpr (std::is_same_v<C, float>) {
    return __int_as_float(i);
  } else {
    return __longlong_as_double(static_cast<long long>(i));
  }
}

template <class C>
__device__ __forceinline__ int bits_index(C v) {
  if constexpr (std::is_same_v<C, float>) {
    return __float_as_int(v);
  } else {
    return static_cast<int>(__double_as_longlong(v));
  }
}

// A slot record's left index with its sign bit set: the face's nine moments are all +0 (bit for
// bit), so the warp terms take zeros in registers instead of reading them; the same values, so
// the same bits. Left indices are cells, never negative otherwise.
constexpr unsigned kFlatFace = 0x80000000u;

template <class C>
__device__ __forceinline__ bool plus_zero(C v) {
  if constexpr (std::is_same_v<C, float>) {
    return __float_as_uint(v) == 0u;
  } else {
    return __double_as_longlong(v) == 0LL;
  }
}

// Where face_core reads the cells' primitives and gradient records: global memory, or LDS for the
// cells [cb, cb + ncb) of a face block staged there (others from global memory).
template <class C>
struct GlobalCells {
  const C* prim;
  const C* rec;
  __device__ __forceinline__ W5<C> w(int c) const { return load_w(prim, c); }
  __device__ __forceinline__ void r(int c, C* g) const { load_rec(rec, c, g); }
};

template <class C>
struct StagedCells {
  const C* prim;
  const C* rec;
  const C* sp;  // LDS: 8 per staged cell
  const C* sr;  // LDS: 20 per staged cell
  int cb, ncb;
  __device__ __forceinline__ W5<C> w(int c) const {
    const unsigned l = static_cast<unsigned>(c - cb);
    return l < static_cast<unsigned>(ncb) ? load_w(sp, static_cast<long long>(l)) : load_w(prim, c);
  }
  __device__ __forceinline__ void r(int c, C* g) const {
    const unsigned l = static_cast<unsigned>(c - cb);
    if (l < static_cast<unsigned>(ncb)) {
      load_rec(sr, static_cast<int>(l), g);
    } else {
      load_rec(rec, c, g);
    }
  }
};

// {flux}_face[_planar] and {flux}_boundary[_planar] (kernels/euler.py:687-1049) for the face `g`
// (12 values, above): the area-weighted flux in C. The warp terms read the face's nine moments at
// m[q * ms], or take zeros when the left index carries kFlatFace. With LAMBDA it also leaves the face's convective spectral radius
// (kernels/euler.py:1091-1112) as each side takes it: lam[0] = max(s_owner, s_other) A for the
// owner, lam[1] = max(s_other, s_owner) A for the neighbour, s = |v . n| + c of the cells' (not
// reconstructed) primitives and of the boundary state.
template <class C, int FLUX, bool PLANAR, bool SECOND, bool LAMBDA, class Src>
__device__ __forceinline__ F5<C> face_core(const Dev<C>& d, const Params<C>& pa, const Src& src, const C* g,
                                           const C* m9, long long ms, C* lam) {
  const V3<C> n{g[0], g[1], g[2]};
  const C area = g[3];
  const int lbits = bits_index(g[7]), right = bits_index(g[11]);
  const bool flat = lbits < 0;  // kFlatFace: the moments are all +0
  const int left = lbits & 0x7fffffff;
  F5<C> out;
  if (right >= 0) {
    W5<C> wl = src.w(left), wr = src.w(right);
    if constexpr (LAMBDA) {
      const C sl = wave_speed(wl, n, pa.g, pa.pref), sr = wave_speed(wr, n, pa.g, pa.pref);
      lam[0] = smax(sl, sr) * area;
      lam[1] = smax(sr, sl) * area;
    }
    C gl[20], gr[20];
    W5<C> fl = wl, fr = wr;
    if constexpr (SECOND) {
      src.r(left, gl);
      src.r(right, gr);
      fl = reconstruct(wl, gl, gl + 15, V3<C>{g[4], g[5], g[6]}, pa.pref);
      fr = reconstruct(wr, gr, gr + 15, V3<C>{g[8], g[9], g[10]}, pa.pref);
    }
    const F5<C> fx = flux<FLUX>(fl, fr, n, pa.g, pa.minf, pa.pref);
#pragma unroll
    for (int k = 0; k < 5; ++k) out.f[k] = area * fx.f[k];
    if constexpr (!PLANAR) {
      C m[9];
      if (flat) {
#pragma unroll
        for (int q = 0; q < 9; ++q) m[q] = C(0);
      } else {
#pragma unroll
        for (int q = 0; q < 9; ++q) m[q] = m9[q * ms];
      }
      if constexpr (!SECOND) {
#pragma unroll
        for (int q = 0; q < 20; ++q) gl[q] = gr[q] = C(0);
      }
      const W5<C> wm{C(0.5) * (fl.r + fr.r), C(0.5) * (fl.u + fr.u), C(0.5) * (fl.v + fr.v), C(0.5) * (fl.w + fr.w),
                     C(0.5) * (fl.p + fr.p)};
      const F5<C> wp = warp(wm, gl, gl + 15, gr, gr + 15, m, pa.g, pa.pref);
#pragma unroll
      for (int k = 0; k < 5; ++k) out.f[k] = out.f[k] + wp.f[k];
    }
  } else {
    const int b = -right - 1;
    const W5<C> wl = src.w(left);
    if constexpr (LAMBDA) {
      const C sl = wave_speed(wl, n, pa.g, pa.pref), sb = wave_speed(load_w(d.prim_b, b), n, pa.g, pa.pref);
      lam[0] = smax(sl, sb) * area;
      lam[1] = lam[0];
    }
    C gl[20];
    W5<C> wf = wl;
    if constexpr (SECOND || !PLANAR) {
      if constexpr (SECOND) {
        src.r(left, gl);
      } else {
#pragma unroll
        for (int q = 0; q < 20; ++q) gl[q] = C(0);
      }
    }
    if constexpr (SECOND) wf = reconstruct(wl, gl, gl + 15, V3<C>{g[4], g[5], g[6]}, pa.pref);
    C data[8];
    load_n<C, 8>(d.bc + 8LL * b, data);
    const int kind = d.bkind[b];
    const bool slip = kind == 2;
    const bool wall = !PLANAR && kind >= 5;
    if (slip || wall) {
      const C wfx[5] = {C(0.0), wf.p * n.x, wf.p * n.y, wf.p * n.z, C(0.0)};
#pragma unroll
      for (int k = 0; k < 5; ++k) out.f[k] = area * wfx[k];
      if constexpr (!PLANAR) {
        C m[9];
        if (flat) {
#pragma unroll
          for (int q = 0; q < 9; ++q) m[q] = C(0);
        } else {
#pragma unroll
          for (int q = 0; q < 9; ++q) m[q] = m9[q * ms];
        }
        const C p4 = gl[15 + 4];
        const C g4[3] = {gl[12], gl[13], gl[14]};
        const C ww[5] = {C(0.0), p4 * (m[0] * g4[0] + m[3] * g4[1] + m[6] * g4[2]),
                         p4 * (m[1] * g4[0] + m[4] * g4[1] + m[7] * g4[2]),
                         p4 * (m[2] * g4[0] + m[5] * g4[1] + m[8] * g4[2]), C(0.0)};
#pragma unroll
        for (int k = 0; k < 5; ++k) out.f[k] = out.f[k] + ww[k];
      }
    } else {
      const W5<C> wb = boundary_value(wf, kind, data, n, pa.g, pa.rgas, pa.pref);
      const F5<C> fx = flux<FLUX>(wf, wb, n, pa.g, pa.minf, pa.pref);
#pragma unroll
      for (int k = 0; k < 5; ++k) out.f[k] = area * fx.f[k];
    }
  }
  return out;
}

// face_core for face f from the per-face arrays (fgeo, frr, lr, bowner, the SoA moments).
template <class C, int FLUX, bool PLANAR, bool SECOND, bool LAMBDA = false>
__device__ __forceinline__ F5<C> face_flux(const Dev<C>& d, int f, C* lam = nullptr) {
  C g[12];
  load_n<C, 8>(d.fgeo + 8LL * f, g);
  if (f < d.ni) {
    const int2 lr = d.lr[f];
    C rr[4];
    load_n<C, 4>(d.frr + 4LL * f, rr);
    g[7] = index_bits<C>(lr.x);
    g[8] = rr[0];
    g[9] = rr[1];
    g[10] = rr[2];
    g[11] = index_bits<C>(lr.y);
  } else {
    const int b = f - d.ni;
    g[7] = index_bits<C>(d.bowner[b]);
    g[8] = g[9] = g[10] = C(0);
    g[11] = index_bits<C>(-b - 1);
  }
  const Params<C>& pa = d.ptab[cell_case(d, bits_index(g[7]))];
  return face_core<C, FLUX, PLANAR, SECOND, LAMBDA>(d, pa, GlobalCells<C>{d.prim, d.rec}, g,
                                                   PLANAR ? nullptr : d.moment + f, d.nf, lam);
}

// The fused face kernel's slot records (face_core's 12 values) of the faces sface[0 .. ns), built
// from the per-face arrays as face_flux assembles them, and their moments (non-planar meshes); a
// face whose nine moments are all +0 has kFlatFace in its left index, and face_core then reads
// none of them.
template <class C>
__global__ __launch_bounds__(kBlock) void k_slot_records(Dev<C> d, const int* __restrict__ sface, long long ns,
                                                         C* __restrict__ srec, C* __restrict__ smom) {
  const long long sl = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (sl >= ns) return;
  const int f = sface[sl];
  bool flat = false;
  if (smom != nullptr) {
    flat = true;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      const C m = d.moment[static_cast<long long>(q) * d.nf + f];
      smom[9 * sl + q] = m;
      flat = flat && plus_zero(m);
    }
  }
  const unsigned fl = flat ? kFlatFace : 0u;
  C g[12];
  load_n<C, 8>(d.fgeo + 8LL * f, g);
  if (f < d.ni) {
    const int2 lr = d.lr[f];
    C rr[4];
    load_n<C, 4>(d.frr + 4LL * f, rr);
    g[7] = index_bits<C>(static_cast<int>(static_cast<unsigned>(lr.x) | fl));
    g[8] = rr[0];
    g[9] = rr[1];
    g[10] = rr[2];
    g[11] = index_bits<C>(lr.y);
  } else {
    const int b = f - d.ni;
    g[7] = index_bits<C>(static_cast<int>(static_cast<unsigned>(d.bowner[b]) | fl));
    g[8] = g[9] = g[10] = C(0);
    g[11] = index_bits<C>(-b - 1);
  }
  store_n<C, 12>(srec + 12 * sl, g);
}

// The face loops' fluxes stored per face for the per-cell gather (the store half of the
// reference's store_gather), for meshes whose blocks have too many faces for the fused kernel.
template <class C, int FLUX, bool PLANAR, bool SECOND>
__global__ __launch_bounds__(kBlock) void k_flux(Dev<C> d) {
  const int f = xcd_index();
  if (f >= d.nf) return;
  const F5<C> out = face_flux<C, FLUX, PLANAR, SECOND>(d, f);
  C o8[8] = {out.f[0], out.f[1], out.f[2], out.f[3], out.f[4], C(0), C(0), C(0)};
  store_n<C, 8>(d.fflux + 8LL * f, o8);
}

// sources (kernels/euler.py:611-616), then each cell's faces in its cell_faces order: the
// interior faces' fluxes added with the cell's sign, then the boundary faces', each set summed
// from zero in FP64 and added to res (codegen/src/autan_codegen/emit_cpu.py:506-564).
template <class C>
__global__ __launch_bounds__(kBlock) void k_gather(Dev<C> d) {
  const int c = xcd_index();
  if (c >= d.nc) return;
  const C nv = -d.vol[c];
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, accb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int k = d.cfo[c]; k < d.cfo[c + 1]; ++k) {
    const int e = d.inc[k];
    const int f = e & 0x7fffffff;
    C fx[8];
    load_n<C, 8>(d.fflux + 8LL * f, fx);
    if (f < d.ni) {
#pragma unroll
      for (int q = 0; q < 5; ++q) acc[q] += e < 0 ? -static_cast<double>(fx[q]) : static_cast<double>(fx[q]);
    } else {
#pragma unroll
      for (int q = 0; q < 5; ++q) accb[q] += static_cast<double>(fx[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const C s = d.source != nullptr ? d.source[static_cast<long long>(q) * d.nc + c] : C(0);
    const double base = static_cast<double>(nv * s);
    d.res[static_cast<long long>(q) * d.nc + c] = (base + acc[q]) + accb[q];
  }
}

// ------------------------------------------------------------------------ time step

// Order-preserving key of a double: a < b (not NaN) iff key(a) < key(b) as unsigned.
__device__ __forceinline__ unsigned long long order_key(double x) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
  return (b >> 63) != 0 ? ~b : (b | 0x8000000000000000ULL);
}
__host__ __device__ inline double key_value(unsigned long long k) {
  const unsigned long long b = (k >> 63) != 0 ? (k & 0x7fffffffffffffffULL) : ~k;
  double x;
  __builtin_memcpy(&x, &b, 8);
  return x;
}

template <class T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o /= 2) {
    const T w = __shfl_xor(v, o, kWave);
    v = w < v ? w : v;
  }
  return v;
}
template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o /= 2) {
    const T w = __shfl_xor(v, o, kWave);
    v = w > v ? w : v;
  }
  return v;
}
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o /= 2) v += __shfl_xor(v, o, kWave);
  return v;
}

// A 64-bit integer moved across lanes by a DPP control (zero where the source lane is out of the
// row, or the destination row is masked off), as two 32-bit halves.
template <int CTRL, int ROWS>
__device__ __forceinline__ long long dpp_move(long long v) {
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, ROWS, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(v >> 32), CTRL, ROWS, 0xf, true);
  return static_cast<long long>((static_cast<unsigned long long>(static_cast<unsigned int>(hi)) << 32) |
                                static_cast<unsigned int>(lo));
}

// The wave's least or greatest 32-bit integer, in registers, by the same row shifts and
// broadcasts (lanes without a source keep `identity`).
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_move32(int v, int identity) {
  return __builtin_amdgcn_update_dpp(identity, v, CTRL, ROWS, 0xf, false);
}
template <bool MIN>
__device__ __forceinline__ int wave_extreme_rows(int v) {
  constexpr int id = MIN ? INT32_MAX : INT32_MIN;
  auto pick = [](int a, int b) { return MIN ? (b < a ? b : a) : (b > a ? b : a); };
  v = pick(v, dpp_move32<0x111, 0xf>(v, id));
  v = pick(v, dpp_move32<0x112, 0xf>(v, id));
  v = pick(v, dpp_move32<0x114, 0xf>(v, id));
  v = pick(v, dpp_move32<0x118, 0xf>(v, id));
  v = pick(v, dpp_move32<0x142, 0xa>(v, id));
  v = pick(v, dpp_move32<0x143, 0xc>(v, id));
  return __builtin_amdgcn_readlane(v, kWave - 1);
}

// The wave's sum of a 64-bit integer, in registers: prefix sums within each 16-lane row (row_shr
// 1, 2, 4, 8), then each row's total carried into the rows after it (row_bcast 15 into rows 1 and
// 3, row_bcast 31 into rows 2 and 3); lane 63 then holds the sum, read by every lane. Integer
// sums: the same words as any other order.
__device__ __forceinline__ long long wave_sum_rows(long long v) {
  v += dpp_move<0x111, 0xf>(v);
  v += dpp_move<0x112, 0xf>(v);
  v += dpp_move<0x114, 0xf>(v);
  v += dpp_move<0x118, 0xf>(v);
  v += dpp_move<0x142, 0xa>(v);
  v += dpp_move<0x143, 0xc>(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(v), kWave - 1);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(v >> 32), kWave - 1);
  return static_cast<long long>((static_cast<unsigned long long>(static_cast<unsigned int>(hi)) << 32) |
                                static_cast<unsigned int>(lo));
}

// spectral_radius (kernels/euler.py:1091-1112), then local_time_step (dt_c = CFL V / lambda,
// physics/src/device_stepper.cuh:83-89) and each block's least dt_c as an ordered key in
// partial[block]; k_min_key reduces those exactly (no same-address atomics from every wave).
template <class C>
__global__ __launch_bounds__(kBlock) void k_spectral(Dev<C> d, double cfl, int store_dt,
                                                     unsigned long long* partial) {
  __shared__ unsigned long long wave_key[kBlock / kWave];
  const int blk = xcd_block();
  const int c = blk * kBlock + threadIdx.x;
  const Params<C>& pa = d.ptab[block_case(d, blk)];
  unsigned long long key = ~0ULL;
  if (c < d.nc) {
    const W5<C> wc = load_w(d.prim, c);
    double acc = 0.0;
    for (int k = d.cfo[c]; k < d.cfo[c + 1]; ++k) {
      const int e = d.inc[k];
      const int f = e & 0x7fffffff;
      C geo[4];
      load_n<C, 4>(d.fgeo + 8LL * f, geo);
      const V3<C> n{geo[0], geo[1], geo[2]};
      W5<C> wn;
      if (f < d.ni) {
        const int2 lr = d.lr[f];
        wn = load_w(d.prim, e < 0 ? lr.x : lr.y);
      } else {
        wn = load_w(d.prim_b, f - d.ni);
      }
      const C lc = wave_speed(wc, n, pa.g, pa.pref);
      const C ln = wave_speed(wn, n, pa.g, pa.pref);
      acc += static_cast<double>(smax(lc, ln) * geo[3]);
    }
    const C lam = static_cast<C>(acc);
    d.lam[c] = lam;
    const double dtc = (cfl * d.vol64[c]) / static_cast<double>(lam);
    if (store_dt) d.dt[c] = dtc;
    if (!__builtin_isnan(dtc)) key = order_key(dtc);
  }
  key = wave_min(key);
  if ((threadIdx.x % kWave) == 0) wave_key[threadIdx.x / kWave] = key;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = wave_key[0];
#pragma unroll
    for (int w = 1; w < kBlock / kWave; ++w) m = wave_key[w] < m ? wave_key[w] : m;
    partial[blk] = m;
  }
}

// The least of n keys (one block), into small->hkey if it is less.
__global__ __launch_bounds__(1024) void k_min_key(const unsigned long long* __restrict__ partial, int n, Small* small) {
  __shared__ unsigned long long wave_key[1024 / kWave];
  unsigned long long m = ~0ULL;
  for (int i = threadIdx.x; i < n; i += 1024) m = partial[i] < m ? partial[i] : m;
  m = wave_min(m);
  if ((threadIdx.x % kWave) == 0) wave_key[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 1024 / kWave; ++w) m = wave_key[w] < m ? wave_key[w] : m;
    m = wave_key[0] < m ? wave_key[0] : m;
    if (m < small->hkey) small->hkey = m;
  }
}

// The validity checks of physics/src/device_stepper.cuh:107-145 for cell c of state u and
// residual res: the stored state read in C with the reference's gauge pressure.
template <class C>
__device__ __forceinline__ bool invalid_state(const double* __restrict__ u, int nc, int c, const Params<C>& pa) {
  const C g = static_cast<C>(pa.gamma);
  const C r0 = static_cast<C>(pa.ref[0]);
  C dq[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) dq[k] = static_cast<C>(u[static_cast<long long>(k) * nc + c] - pa.ref[k]);
  const C rho = r0 + dq[0];
  C dm_m = 0, mr_mr = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const C mr = static_cast<C>(pa.ref[k + 1]);
    dm_m = dm_m + dq[k + 1] * (C(2) * mr + dq[k + 1]);
    mr_mr = mr_mr + mr * mr;
  }
  const C dk = (dm_m * r0 - mr_mr * dq[0]) / (rho * r0);
  const C p = (g - C(1)) * (dq[4] - C(0.5) * dk) + static_cast<C>(pa.p_ref);
  return !__builtin_isfinite(static_cast<double>(rho)) || !__builtin_isfinite(static_cast<double>(p)) ||
         !(rho > C(0)) || !(p > C(0));
}

__device__ __forceinline__ void flag_least(int* slot, bool flagged, int c) {
  int v = flagged ? c : INT32_MAX;
  v = wave_min(v);
  if ((threadIdx.x % kWave) == 0 && v != INT32_MAX) atomicMin(slot, v);
}

// An SSP-RK3 stage done by the fused face kernel as each cell's residual is gathered (mode 1: then
// the new stage's primitives into prim_out, a second buffer since other blocks still read this
// block's cells; mode 2, the last stage: the validity of the new state and of the residual, and
// the new state's primitives into prim_out unless it is null (the next step's first primitives);
// 0: no update, the residual is stored). The arithmetic is k_rk's.
template <class C>
struct RkStage {
  const double* u;   // the step's state
  const double* in;  // the stage's input (u on the first stage)
  double* out;       // the new stage (may be `in`: each cell reads its own values first)
  C* prim_out;
  double a, b, hvalue;
  int use_hkey, local, mode;
};

// One SSP-RK3 stage (physics/src/device_stepper.cuh:92-103): next = stage - (h / V) R,
// stage = a u + b next; on the first stage `stage` is u itself. Then either the primitives of the
// new stage for the next residual, or (last stage) the validity of the new state and residual.
// The time step is the cell's own (local), the case's least (use_hkey) or hvalue.
template <class C, bool PRIM, bool CHECK>
__global__ __launch_bounds__(kBlock) void k_rk(Dev<C> d, double a, double b, const double* __restrict__ u,
                                               const double* stage_in, double* stage_out, int use_hkey, double hvalue,
                                               int local) {
  const int blk = cell_block(d);
  if (blk < 0) return;
  const int cs = block_case(d, blk);
  if (!case_active(d, cs)) return;
  const Params<C>& pa = d.ptab[cs];
  const int c = blk * kBlock + threadIdx.x;
  bool bad_state = false, bad_res = false;
  if (c >= pa.cell_end && c < d.nc) {
    // A filler cell (no faces, so no time step of its own): its state stays.
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const long long i = static_cast<long long>(k) * d.nc + c;
      stage_out[i] = u[i];
    }
    if constexpr (PRIM) store_w(d.prim, c, primitive_of(stage_out, d.nc, c, pa));
  } else if (c < d.nc) {
    const double h = local ? d.dt[c] : (use_hkey ? key_value(d.small[cs].hkey) : hvalue);
    const double hv = h / d.vol64[c];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const long long i = static_cast<long long>(k) * d.nc + c;
      const double r = d.res[i];
      const double next = stage_in[i] - hv * r;
      stage_out[i] = a * u[i] + b * next;
      if constexpr (CHECK) bad_res = bad_res || !__builtin_isfinite(r);
    }
    if constexpr (PRIM) store_w(d.prim, c, primitive_of(stage_out, d.nc, c, pa));
    if constexpr (CHECK) bad_state = invalid_state(stage_out, d.nc, c, pa);
  }
  if constexpr (CHECK) {
    const int cl = c - static_cast<int>(pa.cell0);
    flag_least(&d.small[cs].invalid_state, bad_state, cl);
    flag_least(&d.small[cs].invalid_residual, bad_res, cl);
  }
}

// The gradient records of the active cases' cells cleared (their first-order state restored after
// a second-order evaluation).
template <class C>
__global__ __launch_bounds__(kBlock) void k_zero_rec(Dev<C> d) {
  const int blk = cell_block(d);
  if (blk < 0) return;
  const int cs = block_case(d, blk);
  if (!case_active(d, cs)) return;
  const int c = blk * kBlock + threadIdx.x;
  if (c >= d.nc) return;
  C z[20];
#pragma unroll
  for (int q = 0; q < 20; ++q) z[q] = C(0);
  store_n<C, 20>(d.rec + 20LL * c, z);
}

// DeviceStepper::evaluate's validity checks alone.
template <class C>
__global__ __launch_bounds__(kBlock) void k_check(Dev<C> d, const double* __restrict__ u) {
  const int blk = cell_block(d);
  if (blk < 0) return;
  const int cs = block_case(d, blk);
  if (!case_active(d, cs)) return;
  const Params<C>& pa = d.ptab[cs];
  const int c = blk * kBlock + threadIdx.x;
  bool bad_state = false, bad_res = false;
  if (c < d.nc) {
    bad_state = invalid_state(u, d.nc, c, pa);
#pragma unroll
    for (int k = 0; k < 5; ++k) bad_res = bad_res || !__builtin_isfinite(d.res[static_cast<long long>(k) * d.nc + c]);
  }
  const int cl = c - static_cast<int>(pa.cell0);
  flag_least(&d.small[cs].invalid_state, bad_state, cl);
  flag_least(&d.small[cs].invalid_residual, bad_res, cl);
}

// ------------------------------------------------------------------------ exact norms

// Adds each lane's term to the block's fixed-point limbs `acc` (LDS), exactly: the term split
// into three 32-bit limbs from limb i as an exact sum splits it (runtime/src/exact_sum.cpp:25-48),
// the wave's limbs summed lane-wide per limb index, one LDS atomic per limb per wave. Integer
// sums commute, so the words do not depend on the launch or the order of the terms.
__device__ __forceinline__ void exact_add_wave(long long* acc, double t, bool active) {
  const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(t));
  const int biased = static_cast<int>((bits >> 52) & 0x7ff);
  const unsigned long long fraction = bits & ((1ULL << 52) - 1);
  const bool negative = (bits >> 63) != 0;
  if (active && biased == 0x7ff) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&acc[fraction != 0 ? kLimbs + 2 : (negative ? kLimbs + 1 : kLimbs)]),
              1ULL);
  }
  const bool finite = active && biased != 0x7ff && !(biased == 0 && fraction == 0);
  int i = 1 << 30;
  long long part[3] = {0, 0, 0};
  if (finite) {
    const unsigned long long m = biased == 0 ? fraction : (fraction | (1ULL << 52));
    const int p = (biased == 0 ? -1074 : biased - 1075) + 1088;
    i = p >> 5;
    const int s = p & 31;
    const unsigned long long lo = (m & 0xffffffffULL) << s, hi = (m >> 32) << s;
    part[0] = static_cast<long long>(lo & 0xffffffffULL);
    part[1] = static_cast<long long>((lo >> 32) + (hi & 0xffffffffULL));
    part[2] = static_cast<long long>(hi >> 32);
    if (negative) {
      part[0] = -part[0];
      part[1] = -part[1];
      part[2] = -part[2];
    }
  }
  const int first = wave_extreme_rows<true>(i);
  const int last = wave_extreme_rows<false>(finite ? i : -1);
  for (int j = first; j <= last + 2; ++j) {
    const int dj = j - i;
    const long long v = wave_sum_rows(finite && dj >= 0 && dj < 3 ? part[dj] : 0LL);
    if ((threadIdx.x % kWave) == 0 && v != 0) atomicAdd(reinterpret_cast<unsigned long long*>(&acc[j]), static_cast<unsigned long long>(v));
  }
}

// The five sums of R_k^2 over the cells (EulerPhysics::record_norms), as exact fixed-point words
// added into small->words (cleared beforehand). Grid-stride; each block normalises its limbs
// into [0, 2^32) before adding them, so any number of blocks fits the int64 words.
__global__ __launch_bounds__(kBlock) void k_norms(const double* __restrict__ res, int nc, Small* small) {
  __shared__ long long acc[kSums][kWords];
  for (int t = threadIdx.x; t < kSums * kWords; t += kBlock) (&acc[0][0])[t] = 0;
  __syncthreads();
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const long long stride = static_cast<long long>(gridDim.x) * kBlock;
  for (long long base = static_cast<long long>(blockIdx.x) * kBlock + wave * kWave; base < nc; base += stride) {
    const long long c = base + lane;
    const bool active = c < nc;
#pragma unroll
    for (int k = 0; k < kSums; ++k) {
      const double r = active ? res[k * static_cast<long long>(nc) + c] : 0.0;
      exact_add_wave(acc[k], r * r, active);
    }
  }
  __syncthreads();
  if (threadIdx.x < kSums) {
    long long* a = acc[threadIdx.x];
    for (int l = 0; l + 1 < kLimbs; ++l) {
      const long long carry = a[l] >> 32;
      a[l] -= carry * (1LL << 32);
      a[l + 1] += carry;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kSums * kWords; t += kBlock) {
    const long long v = (&acc[0][0])[t];
    if (v != 0) atomicAdd(reinterpret_cast<unsigned long long*>(&small->words[0][0] + t), static_cast<unsigned long long>(v));
  }
}

// Clears the step's results: block 0 the Small record, block b > 0 exact-sum slot b - 1.
// Block k < ncase clears case k's record, block ncase + j the exact-sum slot j.
__global__ void k_clear(Small* small, int ncase, long long* nslots) {
  if (static_cast<int>(blockIdx.x) >= ncase) {
    long long* w = nslots + static_cast<long long>(blockIdx.x - ncase) * kSums * kWords;
    for (int t = threadIdx.x; t < kSums * kWords; t += blockDim.x) w[t] = 0;
    return;
  }
  Small* sm = small + blockIdx.x;
  for (int t = threadIdx.x; t < kSums * kWords; t += blockDim.x) (&sm->words[0][0])[t] = 0;
  if (threadIdx.x == 0) {
    sm->hkey = 0xfff0000000000000ULL;  // order_key(+inf): no cell, an infinite step
    sm->invalid_state = INT32_MAX;
    sm->invalid_residual = INT32_MAX;
  }
}

// The faces and boundaries loops fused with their gather, one block of at most 256 consecutive
// cells and at most 2 kMaxSlots faces (FP32; kMaxSlots in FP64) per workgroup. Phase 1:
// every face of the block's cells (its interior faces once, the faces it
// shares with another block here and there too) evaluates its flux into LDS. Phase 2: each cell
// adds its faces' fluxes in its cell_faces order, interior faces then boundary faces, each set
// from zero in FP64, onto -V s (kernels/euler.py:611-616): the store_gather order of the
// reference's device path, with no face array in HBM. FIRST (the step's R(u)) also gathers the
// spectral radius, the cell's time step and the block's least one, and the exact sums of R^2.
template <class C, int FLUX, bool PLANAR, bool SECOND, bool FIRST, bool STAGE, int OCC = 0>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1))) void k_faces(
    Dev<C> d, double cfl, int store_dt, unsigned long long* partial, RkStage<C> rk) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int kComp = FIRST ? 7 : 5;
  const int ms = d.max_slots;
  C* lf = reinterpret_cast<C*>(smem);
  const int lf_bytes = (kComp * ms * static_cast<int>(sizeof(C)) + 15) & ~15;
  long long* acc = reinterpret_cast<long long*>(smem + lf_bytes);
  __shared__ unsigned long long wave_key[kBlock / kWave];
  const int blk = face_block(d);
  if (blk < 0) return;
  const int cs = d.fcase != nullptr ? d.fcase[blk] : 0;
  if (!case_active(d, cs)) return;
  const Params<C>& pa = d.ptab[cs];
  const int s0 = d.sbeg[blk], ns = d.sbeg[blk + 1] - s0;
  if constexpr (FIRST) {
    for (int t = threadIdx.x; t < kSums * kWords; t += kBlock) acc[t] = 0;
  }
  const int cb = d.cbeg[blk], ncb = d.cbeg[blk + 1] - cb;
  // STAGE: the block's own cells' primitives (kPrimStride per cell) and gradient records (20)
  // copied to LDS first, as contiguous runs; faces read them there and only their other blocks'
  // cells from global memory.
  C* sp = reinterpret_cast<C*>(smem + lf_bytes + (FIRST ? kSumsLds : 0));
  C* sr = sp + kPrimStride * kBlock;  // 16 B aligned: kPrimStride * kBlock values
  if constexpr (STAGE) {
    constexpr int kVec = 16 / sizeof(C);
    using V = std::conditional_t<std::is_same_v<C, float>, float4, double2>;
    const C* gp = d.prim + static_cast<long long>(kPrimStride) * cb;
    const V* gr = reinterpret_cast<const V*>(d.rec + 20LL * cb);
    V* lr = reinterpret_cast<V*>(sr);
    for (int i = threadIdx.x; i < ncb * kPrimStride; i += kBlock) sp[i] = gp[i];
    for (int i = threadIdx.x; i < ncb * (20 / kVec); i += kBlock) lr[i] = gr[i];
    __syncthreads();
  }
  for (int i = threadIdx.x; i < ns; i += kBlock) {
    const long long sl = s0 + i;
    C g[12];
    load_n<C, 12>(d.srec + 12 * sl, g);
    C lam[2];
    const C* m9 = PLANAR ? nullptr : d.smom + 9 * sl;
    F5<C> fx;
    if constexpr (STAGE) {
      fx = face_core<C, FLUX, PLANAR, SECOND, FIRST>(d, pa, StagedCells<C>{d.prim, d.rec, sp, sr, cb, ncb}, g, m9, 1, lam);
    } else {
      fx = face_core<C, FLUX, PLANAR, SECOND, FIRST>(d, pa, GlobalCells<C>{d.prim, d.rec}, g, m9, 1, lam);
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) lf[q * ms + i] = fx.f[q];
    if constexpr (FIRST) {
      lf[5 * ms + i] = lam[0];
      lf[6 * ms + i] = lam[1];
    }
  }
  __syncthreads();
  const int c = cb + threadIdx.x;
  const bool active = threadIdx.x < ncb;
  double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double lam_acc = 0.0;
  if (active) {
    double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, b[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = d.cfo[c]; k < d.cfo[c + 1]; ++k) {
      const unsigned v = d.islot[k];
      const int slot = v & 0x3fff;
      if (v & 0x8000u) {
#pragma unroll
        for (int q = 0; q < 5; ++q) b[q] += static_cast<double>(lf[q * ms + slot]);
      } else if (v & 0x4000u) {
#pragma unroll
        for (int q = 0; q < 5; ++q) a[q] += -static_cast<double>(lf[q * ms + slot]);
      } else {
#pragma unroll
        for (int q = 0; q < 5; ++q) a[q] += static_cast<double>(lf[q * ms + slot]);
      }
      if constexpr (FIRST) lam_acc += static_cast<double>(lf[((v & 0x4000u) ? 6 : 5) * ms + slot]);
    }
    const C nv = -d.vol[c];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const C sq = d.source != nullptr ? d.source[static_cast<long long>(q) * d.nc + c] : C(0);
      r[q] = (static_cast<double>(nv * sq) + a[q]) + b[q];
      if (rk.mode == 0) d.res[static_cast<long long>(q) * d.nc + c] = r[q];
    }
  }
  double dtc = 0.0;
  if constexpr (FIRST) {
    unsigned long long key = ~0ULL;
    if (active) {
      const C lam = static_cast<C>(lam_acc);
      d.lam[c] = lam;
      dtc = (cfl * d.vol64[c]) / static_cast<double>(lam);
      if (store_dt) d.dt[c] = dtc;
      if (!__builtin_isnan(dtc)) key = order_key(dtc);
    }
    key = wave_min(key);
    if ((threadIdx.x % kWave) == 0) wave_key[threadIdx.x / kWave] = key;
#pragma unroll
    for (int q = 0; q < kSums; ++q) exact_add_wave(acc + q * kWords, r[q] * r[q], active);
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long m = wave_key[0];
#pragma unroll
      for (int w = 1; w < kBlock / kWave; ++w) m = wave_key[w] < m ? wave_key[w] : m;
      partial[blk] = m;
    }
    if (threadIdx.x < kSums) {
      long long* w = acc + threadIdx.x * kWords;
      for (int l = 0; l + 1 < kLimbs; ++l) {
        const long long carry = w[l] >> 32;
        w[l] -= carry * (1LL << 32);
        w[l + 1] += carry;
      }
    }
    __syncthreads();
    long long* out = d.nslots + (static_cast<long long>(cs) * d.nslot + blk % d.nslot) * kSums * kWords;
    for (int t = threadIdx.x; t < kSums * kWords; t += kBlock) {
      const long long v = acc[t];
      if (v != 0) atomicAdd(reinterpret_cast<unsigned long long*>(out + t), static_cast<unsigned long long>(v));
    }
  }
  if (rk.mode != 0) {
    bool bad_state = false, bad_res = false;
    if (active) {
      double un[5];
      if (c >= pa.cell_end) {
        // A filler cell of a batch: its state stays.
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          const long long i = static_cast<long long>(k) * d.nc + c;
          un[k] = rk.u[i];
          rk.out[i] = un[k];
        }
      } else {
        const double h = rk.local ? (FIRST ? dtc : d.dt[c]) : (rk.use_hkey ? key_value(d.small[cs].hkey) : rk.hvalue);
        const double hv = h / d.vol64[c];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          const long long i = static_cast<long long>(k) * d.nc + c;
          const double next = rk.in[i] - hv * r[k];
          un[k] = rk.a * rk.u[i] + rk.b * next;
          rk.out[i] = un[k];
          bad_res = bad_res || !__builtin_isfinite(r[k]);
        }
      }
      if (rk.mode == 1) {
        store_w(rk.prim_out, c, primitive_of(un, 1, 0, pa));
      } else {
        bad_state = invalid_state(un, 1, 0, pa);
        if (rk.prim_out != nullptr) store_w(rk.prim_out, c, primitive_of(un, 1, 0, pa));
      }
    }
    if (rk.mode == 2) {
      const int cl = c - static_cast<int>(pa.cell0);
      flag_least(&d.small[cs].invalid_state, bad_state, cl);
      flag_least(&d.small[cs].invalid_residual, bad_res, cl);
    }
  }
}

// The step's reductions, one block per case: its exact-sum slots added into its words (integer
// sums: any order gives the same words), and the least of its face blocks' time-step keys.
__global__ __launch_bounds__(1024) void k_step_reduce(const long long* __restrict__ nslots, int nslot,
                                                      const unsigned long long* __restrict__ partial,
                                                      const int* __restrict__ fcbeg, const int* __restrict__ cflags,
                                                      Small* small) {
  __shared__ unsigned long long wave_key[1024 / kWave];
  const int cs = blockIdx.x;
  if (cflags != nullptr && (cflags[cs] & 1) == 0) return;
  Small* sm = small + cs;
  const long long* slots = nslots + static_cast<long long>(cs) * nslot * kSums * kWords;
  for (int t = threadIdx.x; t < kSums * kWords; t += 1024) {
    long long v = 0;
    for (int q = 0; q < nslot; ++q) v += slots[static_cast<long long>(q) * kSums * kWords + t];
    (&sm->words[0][0])[t] += v;
  }
  unsigned long long m = ~0ULL;
  for (int i = fcbeg[cs] + threadIdx.x; i < fcbeg[cs + 1]; i += 1024) m = partial[i] < m ? partial[i] : m;
  m = wave_min(m);
  if ((threadIdx.x % kWave) == 0) wave_key[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < 1024 / kWave; ++w) m = wave_key[w] < m ? wave_key[w] : m;
    if (m < sm->hkey) sm->hkey = m;
  }
}

// ------------------------------------------------------------------------ conservation check

// EulerPhysics::check_conservation's sums on one rank (physics/src/euler_physics.cpp:1750-1784), per
// equation k as exact sums, every term on its own: words[k] E = sum_c R^F_c - sum_b F_b,
// words[5 + k] D = sum |F| over the incidences (an interior face twice), words[10 + k] the boundary
// fluxes, words[15 + k] sum_c R^F_c. res holds R^F (the residual without its source), fflux every
// face's flux. Grid-stride with block-uniform trip counts (the wave-wide exact adds need every lane).
constexpr int kConsSums = 20;

template <class C>
__global__ __launch_bounds__(kBlock) void k_conservation(Dev<C> d, long long* __restrict__ words) {
  __shared__ long long acc[kConsSums][kWords];
  for (int i = threadIdx.x; i < kConsSums * kWords; i += kBlock) (&acc[0][0])[i] = 0;
  __syncthreads();
  const long long stride = static_cast<long long>(gridDim.x) * kBlock;
  for (long long base = static_cast<long long>(blockIdx.x) * kBlock; base < d.nc; base += stride) {
    const long long c = base + threadIdx.x;
    const bool active = c < d.nc;
    for (int k = 0; k < kSums; ++k) {
      const double r = active ? d.res[k * static_cast<long long>(d.nc) + c] : 0.0;
      exact_add_wave(acc[k], r, active);
      exact_add_wave(acc[15 + k], r, active);
    }
  }
  for (long long base = static_cast<long long>(blockIdx.x) * kBlock; base < d.nf; base += stride) {
    const long long f = base + threadIdx.x;
    const bool active = f < d.nf;
    const bool boundary = active && f >= d.ni;
    for (int k = 0; k < kSums; ++k) {
      const double v = active ? static_cast<double>(d.fflux[8 * f + k]) : 0.0;
      exact_add_wave(acc[5 + k], boundary ? fabs(v) : 2.0 * fabs(v), active);
      exact_add_wave(acc[k], -v, boundary);
      exact_add_wave(acc[10 + k], v, boundary);
    }
  }
  __syncthreads();
  if (threadIdx.x < kConsSums) {
    long long* a = acc[threadIdx.x];
    for (int l = 0; l + 1 < kLimbs; ++l) {
      const long long carry = a[l] >> 32;
      a[l] -= carry * (1LL << 32);
      a[l + 1] += carry;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kConsSums * kWords; i += kBlock) {
    const long long v = (&acc[0][0])[i];
    if (v != 0) atomicAdd(reinterpret_cast<unsigned long long*>(words + i), static_cast<unsigned long long>(v));
  }
}

// ------------------------------------------------------------------------ wall loads

// Euler::boundary_pressure (physics/src/euler.cpp:1187-1223) on the listed boundary faces, in FP64
// from the stored values: the face pressure of the boundary flux's reconstruction (the cell value
// where the reconstructed density or pressure is not positive) and the pressure force
// p_f S + phi_p M^T grad p. rb: FP64 x_f - x_c of every boundary face (3 per face); out: 3 x n SoA.
template <class C>
__global__ __launch_bounds__(kBlock) void k_face_force(Dev<C> d, const int* __restrict__ faces, int n,
                                                      const double* __restrict__ rb, double* __restrict__ out) {
  const int i = static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x);
  if (i >= n) return;
  const int f = faces[i];
  const int b = f - d.ni;
  const long long c = d.bowner[b];
  const Params<C>& pa = d.ptab[cell_case(d, static_cast<int>(c))];
  const C* g = d.rec + 20 * c;
  double drho = 0.0, dp = 0.0;
  for (int j = 0; j < 3; ++j) {
    const double r = rb[3 * static_cast<long long>(b) + j];
    drho += static_cast<double>(g[j]) * r;
    dp += static_cast<double>(g[12 + j]) * r;
  }
  const double rho_c = static_cast<double>(d.prim[kPrimStride * c]);
  const double p_c = pa.p_ref + static_cast<double>(d.prim[kPrimStride * c + 4]);
  const double rho_f = rho_c + static_cast<double>(g[15]) * drho;
  double p_f = p_c + static_cast<double>(g[19]) * dp;
  if (!(rho_f > 0.0 && p_f > 0.0)) p_f = p_c;
  const double phi_p = static_cast<double>(g[19]);
  const C* fg = d.fgeo + 8 * static_cast<long long>(f);
  for (int k = 0; k < 3; ++k) {
    double warp = 0.0;
    if (d.moment != nullptr) {
      for (int j = 0; j < 3; ++j) {
        warp += static_cast<double>(d.moment[(3 * j + k) * static_cast<long long>(d.nf) + f]) *
                static_cast<double>(g[12 + j]);
      }
    }
    const double area_k = static_cast<double>(fg[k]) * static_cast<double>(fg[3]);
    out[k * static_cast<long long>(n) + i] = p_f * area_k + phi_p * warp;
  }
}

// ------------------------------------------------------------------------ math probe

template <class C>
__global__ void k_probe(const double* x, const double* y, long long n, double* s, double* q, double* e, double* l) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const C a = static_cast<C>(x[i]), b = static_cast<C>(y[i]);
  s[i] = dsqrt(a);
  q[i] = a / b;
  e[i] = dexp(a);
  l[i] = dlog(a);
}

}  // namespace mxe
