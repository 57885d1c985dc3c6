// CDNA4 (gfx950) kernels of the preconditioner-application path.
//
// Every sparse product is "one row -> one value" followed by an epilogue
// that consumes the row value in registers (residual, Jacobi update,
// prolongation add, Krylov dot).  Three storage formats:
//   ELL  : column-major (entry k of row i at k*ld+i), one lane per row; each
//          load instruction of a wavefront reads 64 consecutive values.
//   SELL : sliced ELL, slices of 64 rows (one wavefront) stored column-major,
//          rows length-sorted inside windows so that a slice is as wide as
//          its longest row; one lane per row.
//   CSR  : LANES lanes per row (1..64) reading the row in coalesced strides,
//          reduced with cross-lane shuffles.
// ELL, SELL and one-lane CSR sum each row in its stored order, exactly as
// the sequential CPU loop does.  Epilogues may also return up to two dot
// contributions; the kernel writes per-block partial sums and a one-block
// kernel adds them in block order (deterministic, no atomics).
// Built with -ffp-contract=off: products and sums are rounded separately.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mld {
namespace k {
// Internal linkage: every translation unit that includes this header gets
// its own copy of the kernels.
namespace {

constexpr int BS = 256;  // threads per block (4 wavefronts)
constexpr int SELL_C = 64;

// Matrix values and column indices are read exactly once per product:
// non-temporal loads keep them from evicting the gathered vectors from the
// caches.
template <typename T>
__device__ __forceinline__ T ldnt(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Outputs of products over large levels are streamed out with non-temporal
// stores (nt): they are not re-read by the same kernel and would otherwise
// take L2 lines from the gathered vector (fine-level residual at 256^3:
// 240 -> 216 us).  Small levels keep ordinary stores, their vectors then stay
// in L2 for the next kernel.  NT is a type parameter (a runtime switch lets
// the compiler merge both stores into one plain store); spmv_launch picks
// the NT form of an epilogue for the large DIA levels.
template <bool NT, typename T>
__device__ __forceinline__ void stnt(T* p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <typename S>
using vec2_t = S __attribute__((ext_vector_type(2)));
__host__ __device__ inline bool al2(const void* p, int bytes) { return ((uintptr_t)p % (2 * bytes)) == 0; }

// ---------------------------------------------------------------- gathers
template <typename V, typename XT>
struct GatherX {
  const XT* __restrict__ x;
  __device__ __forceinline__ V operator()(int j) const { return (V)x[j]; }
};
// first Jacobi sweep from a zero guess: y_j = d_j * x_j
template <typename V, typename XT>
struct GatherDX {
  const V* __restrict__ d;
  const XT* __restrict__ x;
  __device__ __forceinline__ V operator()(int j) const { return d[j] * (V)x[j]; }
};

// -------------------------------------------------------------- epilogues
// operator()(i, acc, dots): consume the row value acc of row i.
// pair(i, a0, a1): rows i and i+1 (i even) with 2-element loads and stores,
// the same arithmetic per row; usable when aligned2() holds.
template <typename V, typename YT, bool NT = false>
struct EpiSet {  // y = A x
  YT* __restrict__ y;
  __device__ __forceinline__ void operator()(int i, V acc, double*) const { stnt<NT>(y + i, (YT)acc); }
  __device__ __forceinline__ void pair(int i, V a0, V a1, double*) const {
    stnt<NT>((vec2_t<YT>*)(y + i), vec2_t<YT>{(YT)a0, (YT)a1});
  }
  bool aligned2() const { return al2(y, sizeof(YT)); }
};
template <typename V, bool NT = false>
struct EpiAdd {  // y = y + A x      (prolongation, psb_map_Y2X with beta = 1)
  V* __restrict__ y;
  __device__ __forceinline__ void operator()(int i, V acc, double*) const { stnt<NT>(y + i, y[i] + acc); }
  __device__ __forceinline__ void pair(int i, V a0, V a1, double*) const {
    const vec2_t<V> yy = *(const vec2_t<V>*)(y + i);
    stnt<NT>((vec2_t<V>*)(y + i), vec2_t<V>{yy.x + a0, yy.y + a1});
  }
  bool aligned2() const { return al2(y, sizeof(V)); }
};
template <typename V, typename BT, bool NT = false>
struct EpiRes {  // r = b - A x     (psb_geaxpby + psb_spmm(-1,A,x,1,r))
  const BT* __restrict__ b;
  V* __restrict__ r;
  __device__ __forceinline__ void operator()(int i, V acc, double*) const { stnt<NT>(r + i, (V)b[i] - acc); }
  __device__ __forceinline__ void pair(int i, V a0, V a1, double*) const {
    const vec2_t<BT> bb = *(const vec2_t<BT>*)(b + i);
    stnt<NT>((vec2_t<V>*)(r + i), vec2_t<V>{(V)bb.x - a0, (V)bb.y - a1});
  }
  bool aligned2() const { return al2(b, sizeof(BT)) && al2(r, sizeof(V)); }
};
template <typename V, typename BT, typename YT, bool ADD, bool NT = false>
struct EpiJac {  // t = d*(b - ND y);  y = t  or  y = y + t
  const V* __restrict__ d;
  const BT* __restrict__ b;
  YT* __restrict__ y;
  __device__ __forceinline__ void operator()(int i, V acc, double*) const {
    V t = d[i] * ((V)b[i] - acc);
    if (ADD) stnt<NT>(y + i, (YT)((V)y[i] + t));
    else stnt<NT>(y + i, (YT)t);
  }
  __device__ __forceinline__ void pair(int i, V a0, V a1, double*) const {
    const vec2_t<V> dd = *(const vec2_t<V>*)(d + i);
    const vec2_t<BT> bb = *(const vec2_t<BT>*)(b + i);
    const V t0 = dd.x * ((V)bb.x - a0), t1 = dd.y * ((V)bb.y - a1);
    if (ADD) {
      const vec2_t<YT> yy = *(const vec2_t<YT>*)(y + i);
      stnt<NT>((vec2_t<YT>*)(y + i), vec2_t<YT>{(YT)((V)yy.x + t0), (YT)((V)yy.y + t1)});
    } else {
      stnt<NT>((vec2_t<YT>*)(y + i), vec2_t<YT>{(YT)t0, (YT)t1});
    }
  }
  bool aligned2() const { return al2(d, sizeof(V)) && al2(b, sizeof(BT)) && al2(y, sizeof(YT)); }
};
// Last post-smoothing sweep of the fine level, writing the cycle's output:
// yn = yin + d*(b - ND y);  yout = yn (FP64, or in the cycle's own precision
// when the caller keeps the FP32 result);  dot u.yn when u != nullptr.
// yin and yout may be the same array.
template <typename V, typename BT, bool NT = false, typename YO = double>
struct EpiJacOut {
  const V* d;
  const BT* b;
  const V* yin;
  YO* yout;  // FP64 output, or (YO = V) yin itself
  const double* u;
  __device__ __forceinline__ void operator()(int i, V acc, double* dd) const {
    const V t = d[i] * ((V)b[i] - acc);
    const V yn = yin[i] + t;
    stnt<NT>(yout + i, (YO)yn);
    if (u) dd[0] = u[i] * (double)yn;
  }
  // two rows; their dot terms are added to dd
  __device__ __forceinline__ void pair(int i, V a0, V a1, double* dd) const {
    const vec2_t<V> dv = *(const vec2_t<V>*)(d + i), yv = *(const vec2_t<V>*)(yin + i);
    const vec2_t<BT> bb = *(const vec2_t<BT>*)(b + i);
    const V y0 = yv.x + dv.x * ((V)bb.x - a0), y1 = yv.y + dv.y * ((V)bb.y - a1);
    stnt<NT>((vec2_t<YO>*)(yout + i), vec2_t<YO>{(YO)y0, (YO)y1});
    if (u) {
      const vec2_t<double> uu = *(const vec2_t<double>*)(u + i);
      dd[0] += uu.x * (double)y0;
      dd[0] += uu.y * (double)y1;
    }
  }
  bool aligned2() const {
    return al2(d, sizeof(V)) && al2(b, sizeof(BT)) && al2(yin, sizeof(V)) && al2(yout, sizeof(YO)) &&
           (!u || al2(u, sizeof(double)));
  }
};
// Krylov SpMV: y = A x and dots u.y (u == nullptr: y.y) and w.y
template <typename V, bool NT = false>
struct EpiSetDot {
  V* y;
  const V* u;
  const V* w;
  __device__ __forceinline__ void operator()(int i, V acc, double* d) const {
    stnt<NT>(y + i, acc);
    const double a = (double)acc;
    d[0] = (u ? (double)u[i] : a) * a;
    if (w) d[1] = (double)w[i] * a;
  }
  __device__ __forceinline__ void pair(int i, V a0, V a1, double* d) const {
    stnt<NT>((vec2_t<V>*)(y + i), vec2_t<V>{a0, a1});
    const double b0 = (double)a0, b1 = (double)a1;
    if (u) {
      const vec2_t<V> uu = *(const vec2_t<V>*)(u + i);
      d[0] += (double)uu.x * b0;
      d[0] += (double)uu.y * b1;
    } else {
      d[0] += b0 * b0;
      d[0] += b1 * b1;
    }
    if (w) {
      const vec2_t<V> ww = *(const vec2_t<V>*)(w + i);
      d[1] += (double)ww.x * b0;
      d[1] += (double)ww.y * b1;
    }
  }
  bool aligned2() const { return al2(y, sizeof(V)) && (!u || al2(u, sizeof(V))) && (!w || al2(w, sizeof(V))); }
};

// FCG's Krylov SpMV: y = A x with x = u (the direction p); dots u.y (p.Ap)
// and u.w (p.r with w = r, the step's numerator, taken here where p is read
// anyway instead of in a pass of its own)
template <typename V, bool NT = false>
struct EpiSetDotUW {
  V* y;
  const V* u;
  const V* w;
  __device__ __forceinline__ void operator()(int i, V acc, double* d) const {
    stnt<NT>(y + i, acc);
    const double ui = (double)u[i];
    d[0] = ui * (double)acc;
    d[1] = ui * (double)w[i];
  }
  __device__ __forceinline__ void pair(int i, V a0, V a1, double* d) const {
    stnt<NT>((vec2_t<V>*)(y + i), vec2_t<V>{a0, a1});
    const vec2_t<V> uu = *(const vec2_t<V>*)(u + i), ww = *(const vec2_t<V>*)(w + i);
    d[0] += (double)uu.x * (double)a0;
    d[0] += (double)uu.y * (double)a1;
    d[1] += (double)uu.x * (double)ww.x;
    d[1] += (double)uu.y * (double)ww.y;
  }
  bool aligned2() const { return al2(y, sizeof(V)) && al2(u, sizeof(V)) && al2(w, sizeof(V)); }
};

// The NT form of an epilogue (same data, non-temporal stores).
template <class E>
struct NtForm { using type = void; };
template <typename V, typename YT>
struct NtForm<EpiSet<V, YT, false>> {
  using type = EpiSet<V, YT, true>;
  static type make(const EpiSet<V, YT, false>& e) { return type{e.y}; }
};
template <typename V>
struct NtForm<EpiAdd<V, false>> {
  using type = EpiAdd<V, true>;
  static type make(const EpiAdd<V, false>& e) { return type{e.y}; }
};
template <typename V, typename BT>
struct NtForm<EpiRes<V, BT, false>> {
  using type = EpiRes<V, BT, true>;
  static type make(const EpiRes<V, BT, false>& e) { return type{e.b, e.r}; }
};
template <typename V, typename BT, typename YT, bool ADD>
struct NtForm<EpiJac<V, BT, YT, ADD, false>> {
  using type = EpiJac<V, BT, YT, ADD, true>;
  static type make(const EpiJac<V, BT, YT, ADD, false>& e) { return type{e.d, e.b, e.y}; }
};
template <typename V, typename BT, typename YO>
struct NtForm<EpiJacOut<V, BT, false, YO>> {
  using type = EpiJacOut<V, BT, true, YO>;
  static type make(const EpiJacOut<V, BT, false, YO>& e) { return type{e.d, e.b, e.yin, e.yout, e.u}; }
};
template <typename V>
struct NtForm<EpiSetDot<V, false>> {
  using type = EpiSetDot<V, true>;
  static type make(const EpiSetDot<V, false>& e) { return type{e.y, e.u, e.w}; }
};
template <typename V>
struct NtForm<EpiSetDotUW<V, false>> {
  using type = EpiSetDotUW<V, true>;
  static type make(const EpiSetDotUW<V, false>& e) { return type{e.y, e.u, e.w}; }
};

// ------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum of v over the block, valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) s += lds[w];
  }
  return s;
}

// K partial sums per block -> partial[q*nblocks + block].  Every thread of
// the block must call it.  The partials are added in block order by
// finalize_kernel, launched right after on the same stream (no atomics).
template <int K>
__device__ __forceinline__ void block_partials(const double* v, double* partial, int slot) {
  __shared__ double lds[BS / 64];
  const int nb = gridDim.x;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const double s = block_sum(v[q], lds);
    if (threadIdx.x == 0) partial[(size_t)q * nb + slot] = s;
  }
}
template <int K>
__device__ __forceinline__ void block_partials(const double* v, double* partial) {
  block_partials<K>(v, partial, blockIdx.x);
}

// One workgroup of FIN_T threads adds the nb partials of each of K sums:
// thread t sums entries t, t+FIN_T, t+2*FIN_T, ... in index order, loading 8
// at a time (independent loads in flight; a plain strided loop was latency
// bound at ~60 us for 64k partials), then a fixed-order wave / block tree.
constexpr int FIN_T = 1024;
template <int K>
__global__ __launch_bounds__(FIN_T) void finalize_kernel(int nb, const double* __restrict__ partial,
                                                          double* __restrict__ result) {
  __shared__ double lds[K][FIN_T / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const double* pp = partial + (size_t)q * nb;
    double acc = 0.0;
    int b = threadIdx.x;
    for (; b + 7 * FIN_T < nb; b += 8 * FIN_T) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = pp[b + u * FIN_T];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += v[u];
    }
    for (; b < nb; b += FIN_T) acc += pp[b];
    acc = wave_sum(acc);
    if (lane == 0) lds[q][wid] = acc;
  }
  __syncthreads();
  if ((int)threadIdx.x < K) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < FIN_T / 64; ++w) s += lds[threadIdx.x][w];
    result[threadIdx.x] = s;
  }
}

struct Red {  // reduction buffers of a kernel with dot outputs
  double* partial;
  double* result;
};

// Block -> row-block order of a sparse product.  Workgroups reach the 8
// XCDs round-robin in blockIdx order, and each XCD has its own L2.  With
// chunk = C > 0, row blocks are handed out in chunks of C consecutive blocks,
// XCD x taking chunks x, x+8, x+16, ...; blocks past the last full period of
// 8C keep their index.  chunk = 0 is the identity.  The launcher picks C per
// operator (autotuned among the identity, chunks of 8192, 32768 and 131072
// rows and one contiguous range per XCD): with 8192-row chunks an XCD's rows
// recur every 65536 rows,
// so for a 256^2-plane grid the planes a block gathers sit in that XCD's L2
// (fine-level traffic 1.20x -> 1.01x of the algorithmic bytes); operators
// whose rows gather from far apart (restriction) prefer a contiguous range.
__device__ __forceinline__ int xcd_block(int chunk) {
  const int p = blockIdx.x;
  if (chunk <= 0) return p;
  const int full = (gridDim.x / (8 * chunk)) * (8 * chunk);
  if (p >= full) return p;
  const int x = p & 7, k = p >> 3;
  return ((k / chunk) * 8 + x) * chunk + (k % chunk);
}

// ----------------------------------------------------------- SpMV kernels
// ITERS > 0: the first ITERS strided entries of each lane are loaded up front
// (predicated), so that all index/value loads and then all gathers of a lane
// are in flight together; rows longer than ITERS*LANES finish in a loop.
// The lane's summation order is the same as the plain strided loop.
template <typename V, int LANES, int ITERS, int K, class G, class E, typename S>
__global__ __launch_bounds__(BS) void csr_kernel(int nr, const int* __restrict__ irp,
                                                  const int* __restrict__ ja, const S* __restrict__ val,
                                                  G g, E e, Red red, int xchunk) {
  const int lb = xcd_block(xchunk);
  const int gid = lb * BS + threadIdx.x;
  const int row = gid / LANES;
  const int lane = gid & (LANES - 1);
  V acc = 0;
  if (row < nr) {
    const int b = irp[row], en = irp[row + 1];
    int k0 = b + lane;
    if constexpr (ITERS > 0) {
      int jj[ITERS];
      V vv[ITERS];
#pragma unroll
      for (int it = 0; it < ITERS; ++it) {
        const int k = k0 + it * LANES;
        const bool in = k < en;
        jj[it] = in ? ldnt(ja + k) : -1;
        vv[it] = in ? (V)ldnt(val + k) : (V)0;
      }
      V xx[ITERS];
#pragma unroll
      for (int it = 0; it < ITERS; ++it) xx[it] = (jj[it] >= 0) ? g(jj[it]) : (V)0;
#pragma unroll
      for (int it = 0; it < ITERS; ++it)
        if (jj[it] >= 0) acc = acc + vv[it] * xx[it];
      k0 += ITERS * LANES;
    }
    for (int k = k0; k < en; k += LANES) acc = acc + (V)ldnt(val + k) * g(ldnt(ja + k));
  }
  if (LANES > 1) {
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) acc = acc + __shfl_down(acc, off, LANES);
  }
  double d[2] = {0.0, 0.0};
  if (lane == 0 && row < nr) e(row, acc, d);
  if constexpr (K > 0) block_partials<K>(d, red.partial, lb);
}

template <typename V, int W, int K, class G, class E, typename S>
__global__ __launch_bounds__(BS) void ell_kernel(int nr, int w, int ld, const int* __restrict__ ja,
                                                  const S* __restrict__ val, G g, E e, Red red, int xchunk) {
  const int lb = xcd_block(xchunk);
  const int row = lb * BS + threadIdx.x;
  double d[2] = {0.0, 0.0};
  if (row < nr) {
    V acc = 0;
    if constexpr (W > 0) {
#pragma unroll
      for (int k = 0; k < W; ++k) {
        const int idx = k * ld + row;
        acc = acc + (V)ldnt(val + idx) * g(ldnt(ja + idx));
      }
    } else {
      for (int k = 0; k < w; ++k) {
        const int idx = k * ld + row;
        acc = acc + (V)ldnt(val + idx) * g(ldnt(ja + idx));
      }
    }
    e(row, acc, d);
  }
  if constexpr (K > 0) block_partials<K>(d, red.partial, lb);
}

// DIA: nd diagonals with offsets off.o[0..nd) ascending; entry (i, i+o_d) at
// val[d*nr + i], zero where the row has no such entry (its column is then
// clamped into range; 0 * x adds nothing).  No column indices are stored or
// loaded, and the gathers of a wavefront are contiguous.  Summed over the
// diagonals in ascending offset order = the row's stored column order.
struct DiaOff {
  int o[8];
};
// PAIR (plain products, nr even): each thread takes two adjacent rows and
// reads their values of a diagonal with one 2-element load (fine-level
// residual at 256^3: 216 -> 198 us).  Otherwise RPT rows BS apart.
template <typename V, int ND, int K, int RPT, class G, class E, typename S, bool PAIR = false>
__global__ __launch_bounds__(BS) void dia_kernel(int nr, int nc, int nd, DiaOff off, const S* __restrict__ val, G g,
                                                  E e, Red red, int xchunk) {
  const int lb = xcd_block(xchunk);
  if constexpr (PAIR) {
    const int i0 = 2 * (lb * BS + threadIdx.x);
    double pd[2] = {0.0, 0.0};
    if (i0 < nr) {
      V a0 = 0, a1 = 0;
      const int ndg = (ND > 0) ? ND : nd;
#pragma unroll
      for (int k = 0; k < ((ND > 0) ? ND : 8); ++k) {
        if (ND == 0 && k >= ndg) break;
        const vec2_t<S> vv = ldnt((const vec2_t<S>*)(val + (int64_t)k * nr + i0));
        const int j0 = min(max(i0 + off.o[k], 0), nc - 1), j1 = min(max(i0 + 1 + off.o[k], 0), nc - 1);
        a0 = a0 + (V)vv.x * g(j0);
        a1 = a1 + (V)vv.y * g(j1);
      }
      e.pair(i0, a0, a1, pd);
    }
    if constexpr (K > 0) block_partials<K>(pd, red.partial, lb);
    return;
  }
  double dd[2] = {0.0, 0.0};
#pragma unroll
  for (int rr = 0; rr < RPT; ++rr) {
    const int row = (lb * RPT + rr) * BS + threadIdx.x;
    if (row < nr) {
      V acc = 0;
      const int ndg = (ND > 0) ? ND : nd;
#pragma unroll
      for (int k = 0; k < ((ND > 0) ? ND : 8); ++k) {
        if (ND == 0 && k >= ndg) break;
        const int j = min(max(row + off.o[k], 0), nc - 1);
        acc = acc + (V)ldnt(val + (int64_t)k * nr + row) * g(j);
      }
      double d1[2] = {0.0, 0.0};
      e(row, acc, d1);
      dd[0] += d1[0];
      dd[1] += d1[1];
    }
  }
  if constexpr (K > 0) block_partials<K>(dd, red.partial, lb);
}

// SELL: position p = 64*slice + lane holds row perm[p]; slice s occupies
// [sptr[s], sptr[s+1]) with entry k of its lane at sptr[s] + 64*k + lane.
template <typename V, int K, class G, class E, typename S>
__global__ __launch_bounds__(BS) void sell_kernel(int nr, const int* __restrict__ sptr,
                                                   const int* __restrict__ perm, const int* __restrict__ ja,
                                                   const S* __restrict__ val, G g, E e, Red red, int xchunk) {
  const int lb = xcd_block(xchunk);
  const int p = lb * BS + threadIdx.x;
  double d[2] = {0.0, 0.0};
  if (p < nr) {
    const int s = p >> 6, l = p & 63;
    const int base = sptr[s] + l;
    const int w = (sptr[s + 1] - sptr[s]) >> 6;
    V acc = 0;
    int k = 0;
    // eight entries per step: all index and value loads, then all gathers,
    // in flight together
    for (; k + 8 <= w; k += 8) {
      const int i0 = base + 64 * k;
      int j[8];
      V v[8], x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) j[u] = ldnt(ja + i0 + 64 * u);
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = (V)ldnt(val + i0 + 64 * u);
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = g(j[u]);
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = acc + v[u] * x[u];
    }
    for (; k + 4 <= w; k += 4) {
      const int i0 = base + 64 * k;
      const int j0 = ldnt(ja + i0), j1 = ldnt(ja + i0 + 64), j2 = ldnt(ja + i0 + 128), j3 = ldnt(ja + i0 + 192);
      const V v0 = (V)ldnt(val + i0), v1 = (V)ldnt(val + i0 + 64), v2 = (V)ldnt(val + i0 + 128),
              v3 = (V)ldnt(val + i0 + 192);
      const V x0 = g(j0), x1 = g(j1), x2 = g(j2), x3 = g(j3);
      acc = acc + v0 * x0;
      acc = acc + v1 * x1;
      acc = acc + v2 * x2;
      acc = acc + v3 * x3;
    }
    for (; k < w; ++k) {
      const int i0 = base + 64 * k;
      acc = acc + (V)ldnt(val + i0) * g(ldnt(ja + i0));
    }
    e(perm[p], acc, d);
  }
  if constexpr (K > 0) block_partials<K>(d, red.partial, lb);
}

// SELD: the SELL product with 16-bit column steps (entry k's column is the
// previous one + dja[k] + 1, starting from col0 - 1; padding entries, value
// zero, keep stepping by one and gather the clamped column).  Same stored
// order as SELL, so the same sums.
template <typename V, int K, class G, class E, typename S>
__global__ __launch_bounds__(BS) void seld_kernel(int nr, int nc, const int* __restrict__ sptr,
                                                  const int* __restrict__ perm, const int* __restrict__ col0,
                                                  const uint16_t* __restrict__ dja, const S* __restrict__ val, G g,
                                                  E e, Red red, int xchunk) {
  const int lb = xcd_block(xchunk);
  const int p = lb * BS + threadIdx.x;
  double d[2] = {0.0, 0.0};
  if (p < nr) {
    const int s = p >> 6, l = p & 63;
    const int base = sptr[s] + l;
    const int w = (sptr[s + 1] - sptr[s]) >> 6;
    const int cmax = nc - 1;
    int col = col0[p] - 1;
    V acc = 0;
    int k = 0;
    for (; k + 8 <= w; k += 8) {
      const int i0 = base + 64 * k;
      int j[8];
      V v[8], x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) j[u] = (int)ldnt(dja + i0 + 64 * u);
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = (V)ldnt(val + i0 + 64 * u);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        col += j[u] + 1;
        j[u] = min(col, cmax);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = g(j[u]);
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = acc + v[u] * x[u];
    }
    for (; k < w; ++k) {
      const int i0 = base + 64 * k;
      col += (int)ldnt(dja + i0) + 1;
      acc = acc + (V)ldnt(val + i0) * g(min(col, cmax));
    }
    e(perm[p], acc, d);
  }
  if constexpr (K > 0) block_partials<K>(d, red.partial, lb);
}

// SELL column indices (ja, stored order per position) -> SELD steps and
// first columns; positions past nr and entries past a row's length are
// padding (step 0).  Rows are strictly increasing with steps <= 65536
// (checked on the host before this layout is built).
__global__ __launch_bounds__(BS) void seld_convert_kernel(int nr, int npos, const int* __restrict__ perm,
                                                          const int* __restrict__ sptr, const int* __restrict__ irp,
                                                          const int* __restrict__ ja, int* __restrict__ col0,
                                                          uint16_t* __restrict__ dja) {
  const int p = blockIdx.x * BS + threadIdx.x;
  if (p >= npos) return;
  const int s = p >> 6, l = p & 63;
  const int base = sptr[s] + l;
  const int w = (sptr[s + 1] - sptr[s]) >> 6;
  const int len = (p < nr) ? irp[perm[p] + 1] - irp[perm[p]] : 0;
  int prev = (len > 0 ? ja[base] : 0) - 1;
  col0[p] = prev + 1;
  for (int k = 0; k < w; ++k) {
    const int i = base + 64 * k;
    if (k < len) {
      const int c = ja[i];
      dja[i] = (uint16_t)(c - prev - 1);
      prev = c;
    } else {
      dja[i] = 0;
    }
  }
}

// --------------------------------------------------------- vector kernels
template <typename V, typename XT, bool ADD>
__global__ __launch_bounds__(BS) void vmul_kernel(int n, const V* __restrict__ d, const XT* __restrict__ x,
                                                   V* __restrict__ y) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  V t = d[i] * (V)x[i];
  if (ADD) y[i] = y[i] + t;
  else y[i] = t;
}

// One-sweep post-smoothing of the fine level writing the cycle's output:
// yout = yin + d*x (FP64), with the dot u.yout when u != nullptr.
template <typename V, int K, typename YO = double>
__global__ __launch_bounds__(BS) void vmul_out_kernel(int n, const V* __restrict__ d, const V* __restrict__ x,
                                                       const V* yin, YO* yout, const double* __restrict__ u,
                                                       Red red) {
  const int i = blockIdx.x * BS + threadIdx.x;
  double dd[2] = {0.0, 0.0};
  if (i < n) {
    const V yn = yin[i] + d[i] * x[i];
    yout[i] = (YO)yn;
    if (K > 0) dd[0] = u[i] * (double)yn;
  }
  if constexpr (K > 0) block_partials<K>(dd, red.partial);
}

template <typename TO, typename TI>
__global__ __launch_bounds__(BS) void convert_kernel(int n, const TI* __restrict__ x, TO* __restrict__ y) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) y[i] = (TO)x[i];
}

// y = x + y (axpby with alpha = beta = 1) or y = x (beta = 0)
template <typename V, typename XT, bool ADD>
__global__ __launch_bounds__(BS) void add_kernel(int n, const XT* __restrict__ x, V* __restrict__ y) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  if (ADD) y[i] = (V)x[i] + y[i];
  else y[i] = (V)x[i];
}

template <typename V>
__global__ __launch_bounds__(BS) void fill_kernel(int n, V* __restrict__ y, V v) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) y[i] = v;
}

// y = a*x + b*y, general coefficients (PSBLAS geaxpby)
template <typename V>
__global__ __launch_bounds__(BS) void axpby_kernel(int n, V a, const V* __restrict__ x, V b, V* __restrict__ y) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  if (b == (V)0) y[i] = a * x[i];
  else y[i] = a * x[i] + b * y[i];
}

// Elements per thread of the streaming reduction kernels (each is followed
// by finalize_kernel over its red_blocks(n) partials).
constexpr int RED_EPT = 2;
__host__ __device__ constexpr int red_blocks(int n) { return (n + BS * RED_EPT - 1) / (BS * RED_EPT); }

// dot products: K pairs (x_q, y_q) in one pass, the y operands of type W
template <typename V, int K, typename W = V>
__global__ __launch_bounds__(BS) void dot_kernel(int n, const V* __restrict__ x0, const W* __restrict__ y0,
                                                  const V* __restrict__ x1, const W* __restrict__ y1, Red red) {
  double s[2] = {0.0, 0.0};
  const int b = blockIdx.x * BS * RED_EPT + threadIdx.x;
#pragma unroll
  for (int e = 0; e < RED_EPT; ++e) {
    const int i = b + e * BS;
    if (i < n) {
      s[0] += (double)x0[i] * (double)y0[i];
      if (K > 1) s[1] += (double)x1[i] * (double)y1[i];
    }
  }
  block_partials<K>(s, red.partial);
}

// ------------------------------------------------------------ Krylov steps
// Scalar slots in device memory.
enum { S_RHO0 = 0, S_RHO1 = 1, S_SIGMA = 2, S_RR = 3, S_PR = 4, S_GAMMA = 5, S_DELTA = 6, S_DPR = 7,
       S_TMP0 = 8, S_TMP1 = 9, S_OMEGA = 10, S_ALPHA = 11, S_K0 = 12, S_K1 = 13, S_ABORT = 15, S_NSCAL = 16 };
// (two-dot reductions write slot and slot + 1: S_DELTA/S_DPR (FCG's p.Ap and
// p.r), S_TMP0/S_TMP1, S_K1/14; S_ABORT stays clear of them)

// Step length num/den from device slots; 0 when den == 0 (breakdown, detected
// on the host), which leaves x and r unchanged.
__device__ __forceinline__ double step_len(const double* __restrict__ scal, int num_slot, int den_slot) {
  const double den = scal[den_slot];
  return (den != 0.0) ? scal[num_slot] / den : 0.0;
}

// The solution update x += alpha p of an iteration is deferred to the next
// direction update, which reads p anyway (one pass over p fewer per
// iteration); x_update applies the last one when the iteration ends.  The
// arithmetic is that of the update in place.

// CG: p = z (first) or p = z + (rho_new/rho_old) p; z in FP64 or the FP32
// output of an FP32 preconditioner.  From the second iteration also the
// previous iteration's x += alpha p (alpha = scal[xnum] / scal[xden]).
template <typename ZT>
__global__ __launch_bounds__(BS) void cg_pupdate(int n, const ZT* __restrict__ z, double* __restrict__ p,
                                                  const double* __restrict__ scal, int it,
                                                  double* __restrict__ x, int xnum, int xden) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  if (it == 1) {
    p[i] = (double)z[i];
  } else {
    const double rn = scal[(it & 1) ? S_RHO1 : S_RHO0], ro = scal[(it & 1) ? S_RHO0 : S_RHO1];
    const double beta = rn / ro;
    const double pi = p[i];
    x[i] = step_len(scal, xnum, xden) * pi + x[i];
    p[i] = (double)z[i] + beta * pi;
  }
}

// r -= alpha q; sum r_i^2.  r32 (may be null): also writes r in FP32, the
// input of the next FP32 preconditioner application (saves its conversion
// pass).
__global__ __launch_bounds__(BS) void r_update(int n, double* __restrict__ r, const double* __restrict__ q,
                                                const double* __restrict__ scal, int num_slot, int den_slot,
                                                Red red, float* __restrict__ r32) {
  const double alpha = step_len(scal, num_slot, den_slot);
  double s[2] = {0.0, 0.0};
  const int b = blockIdx.x * BS * RED_EPT + threadIdx.x;
#pragma unroll
  for (int e = 0; e < RED_EPT; ++e) {
    const int i = b + e * BS;
    if (i < n) {
      const double ri = (-alpha) * q[i] + r[i];
      r[i] = ri;
      if (r32) r32[i] = (float)ri;
      s[0] += ri * ri;
    }
  }
  block_partials<1>(s, red.partial);
}

// x += alpha p (the last deferred update)
__global__ __launch_bounds__(BS) void x_update(int n, double* __restrict__ x, const double* __restrict__ p,
                                                const double* __restrict__ scal, int num_slot, int den_slot) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) x[i] = step_len(scal, num_slot, den_slot) * p[i] + x[i];
}

// FCG(1): p = z - (gamma/delta) p (z as in cg_pupdate; from the second
// iteration also the previous iteration's x += alpha p).  The step's
// numerator p.r is reduced by the following product (EpiSetDotUW).
template <typename ZT>
__global__ __launch_bounds__(BS) void fcg_pupdate(int n, const ZT* __restrict__ z, double* __restrict__ p,
                                                   const double* __restrict__ scal, int it,
                                                   double* __restrict__ x, int xnum, int xden) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  if (it == 1) {
    p[i] = (double)z[i];
  } else {
    const double c = -(scal[S_GAMMA] / scal[S_DELTA]);
    const double po = p[i];
    x[i] = step_len(scal, xnum, xden) * po + x[i];
    p[i] = (double)z[i] + c * po;
  }
}

// ---------------------------------------------------- triangular solves
// One row of a triangular solve by a group of LANES lanes, off-diagonal
// entries f_ij of the row in CSR.  SC selects the diagonal treatment:
//   SC = 0: unit diagonal,          t_i = b_i - acc;
//   SC = 1: scaled right-hand side, t_i = d_i * b_i - acc (ILU's second solve,
//           the D scaling of psb_spsm);
//   SC = 2: stored diagonal d_i,    t_i = (b_i - acc) / d_i (Gauss-Seidel).
// The lanes form the products f_ij * t_j of LANES stored entries at once
// (independent loads in flight together); every lane then adds them up in the
// sequential order with cross-lane reads, so the result is the sequential
// loop's:
//   TR = false: acc = f_ij1 t_j1 + f_ij2 t_j2 + ... (stored order);
//   TR = true (rows of U^T / L^T, the transposed ILU apply): products
//     subtracted one at a time in the order of the column-oriented sweep over
//     the untransposed factor (U^T, SC = 0: ascending; L^T, SC = 1:
//     descending column index), t_i = (...((b'_i - f_ij1 t_j1) - f_ij2 t_j2)...).
template <typename V, int SC, bool TR, int LANES>
__device__ __forceinline__ V tri_row(int i, int lane, const int* __restrict__ irp, const int* __restrict__ ja,
                                     const V* __restrict__ val, const V* __restrict__ b,
                                     const V* __restrict__ d, const V* t) {
  const int k0 = irp[i], k1 = irp[i + 1];
  const V bi = (SC == 1) ? d[i] * b[i] : b[i];
  V acc = TR ? bi : (V)0;
  if constexpr (!(TR && SC == 1)) {
    for (int kb = k0; kb < k1; kb += LANES) {
      const int k = kb + lane;
      V pk = (V)0;
      if (k < k1) pk = val[k] * t[ja[k]];
      const int cnt = min(LANES, k1 - kb);
      if constexpr (LANES == 1) {
        acc = TR ? acc - pk : acc + pk;
      } else {
        for (int q = 0; q < cnt; ++q) {
          const V pq = __shfl(pk, q, LANES);
          acc = TR ? acc - pq : acc + pq;
        }
      }
    }
  } else {
    for (int ke = k1; ke > k0; ke -= LANES) {
      const int kb = max(k0, ke - LANES);
      const int k = kb + lane;
      V pk = (V)0;
      if (k < ke) pk = val[k] * t[ja[k]];
      if constexpr (LANES == 1) {
        acc = acc - pk;
      } else {
        for (int q = ke - kb - 1; q >= 0; --q) acc = acc - __shfl(pk, q, LANES);
      }
    }
  }
  if constexpr (TR) return acc;
  if constexpr (SC == 2) return (bi - acc) / d[i];
  return bi - acc;
}

// The rows of one level set, LANES lanes per row.  BETA: 0 also writes
// y = t, 1 writes y = y + t, 2 writes t only.
template <typename V, int SC, int BETA, bool TR, int LANES>
__global__ __launch_bounds__(BS) void trsv_level_kernel(int cnt, const int* __restrict__ rows,
                                                         const int* __restrict__ irp, const int* __restrict__ ja,
                                                         const V* __restrict__ val, const V* __restrict__ b,
                                                         const V* __restrict__ d, V* t, V* __restrict__ y) {
  const int gid = blockIdx.x * BS + threadIdx.x;
  const int idx = gid / LANES, lane = gid % LANES;
  if (idx >= cnt) return;  // whole lane groups leave together
  const int i = rows[idx];
  const V ti = tri_row<V, SC, TR, LANES>(i, lane, irp, ja, val, b, d, t);
  if (lane == 0) {
    t[i] = ti;
    if (BETA == 0) y[i] = ti;
    else if (BETA == 1) y[i] = y[i] + ti;
  }
}

// A run of consecutive level sets [lv0, lv1) of one triangular solve in one
// workgroup, with a barrier between level sets; per row the same arithmetic
// as trsv_level_kernel.  Replaces one launch per small level set.
template <typename V, int SC, int BETA, bool TR, int LANES>
__global__ __launch_bounds__(1024) void trsv_run_kernel(int lv0, int lv1, const int* __restrict__ ptr,
                                                         const int* __restrict__ rows, const int* __restrict__ irp,
                                                         const int* __restrict__ ja, const V* __restrict__ val,
                                                         const V* __restrict__ b, const V* __restrict__ d, V* t,
                                                         V* __restrict__ y) {
  const int lane = threadIdx.x % LANES, grp = threadIdx.x / LANES, ngrp = blockDim.x / LANES;
  for (int lv = lv0; lv < lv1; ++lv) {
    for (int q = ptr[lv] + grp; q < ptr[lv + 1]; q += ngrp) {
      const int i = rows[q];
      const V ti = tri_row<V, SC, TR, LANES>(i, lane, irp, ja, val, b, d, t);
      if (lane == 0) {
        t[i] = ti;
        if (BETA == 0) y[i] = ti;
        else if (BETA == 1) y[i] = y[i] + ti;
      }
    }
    __syncthreads();
  }
}

// Sync-free substitution (no launch or barrier per level set): the rows in
// level-schedule order are cut into chunks of at most BS/LANES rows that
// never span two level sets (the rows of a chunk are independent).
// Workgroups take chunks in order through an atomic ticket, so every chunk a
// row depends on was taken earlier by a running workgroup.  The solution
// vector itself signals completion: it is filled with a NaN sentinel first,
// and a lane reading t_j polls it (device-scope atomic load) until it is not
// the sentinel, then uses that value in its product; t_i is written with a
// device-scope atomic store.  Value and signal are one word, so no fence is
// needed between them.  The per-row arithmetic and summation order are
// tri_row's.  A poll that exceeds spin_max attempts (or sees another's abort)
// stops waiting and sets *abort, so the kernel always ends; the host reports
// the failure.  The workgroup that draws the last ticket resets the counter
// for the next launch.
template <typename V>
struct SfSentinel;
template <>
struct SfSentinel<double> {
  static constexpr uint64_t bits = 0x7ff4deadbeefcafeull;
  __device__ static bool is(double v) { return __double_as_longlong(v) == (long long)bits; }
  __device__ static double value() { return __longlong_as_double((long long)bits); }
};
template <>
struct SfSentinel<float> {
  static constexpr uint32_t bits = 0x7fbadbadu;
  __device__ static bool is(float v) { return __float_as_uint(v) == bits; }
  __device__ static float value() { return __uint_as_float(bits); }
};
template <typename V>
__global__ __launch_bounds__(BS) void sf_fill_kernel(int n, V* __restrict__ t) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) t[i] = SfSentinel<V>::value();
}
template <typename V>
__device__ __forceinline__ V sf_wait(const V* tj, double* abort, int spin_max) {
  V v = __hip_atomic_load(tj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int n = 0;
  while (SfSentinel<V>::is(v)) {
    if (++n > spin_max || __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0.0) {
      __hip_atomic_store(abort, 1.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return (V)0;
    }
    __builtin_amdgcn_s_sleep(2);
    v = __hip_atomic_load(tj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return v;
}
// tri_row with every t_j taken from sf_wait
template <typename V, int SC, bool TR, int LANES>
__device__ __forceinline__ V tri_row_sf(int i, int lane, const int* __restrict__ irp, const int* __restrict__ ja,
                                        const V* __restrict__ val, const V* __restrict__ b,
                                        const V* __restrict__ d, const V* t, double* abort, int spin_max) {
  const int k0 = irp[i], k1 = irp[i + 1];
  const V bi = (SC == 1) ? d[i] * b[i] : b[i];
  V acc = TR ? bi : (V)0;
  if constexpr (!(TR && SC == 1)) {
    for (int kb = k0; kb < k1; kb += LANES) {
      const int k = kb + lane;
      V pk = (V)0;
      if (k < k1) pk = val[k] * sf_wait(t + ja[k], abort, spin_max);
      const int cnt = min(LANES, k1 - kb);
      if constexpr (LANES == 1) {
        acc = TR ? acc - pk : acc + pk;
      } else {
        for (int q = 0; q < cnt; ++q) {
          const V pq = __shfl(pk, q, LANES);
          acc = TR ? acc - pq : acc + pq;
        }
      }
    }
  } else {
    for (int ke = k1; ke > k0; ke -= LANES) {
      const int kb = max(k0, ke - LANES);
      const int k = kb + lane;
      V pk = (V)0;
      if (k < ke) pk = val[k] * sf_wait(t + ja[k], abort, spin_max);
      if constexpr (LANES == 1) {
        acc = acc - pk;
      } else {
        for (int q = ke - kb - 1; q >= 0; --q) acc = acc - __shfl(pk, q, LANES);
      }
    }
  }
  if constexpr (TR) return acc;
  if constexpr (SC == 2) return (bi - acc) / d[i];
  return bi - acc;
}
template <typename V, int SC, int BETA, bool TR, int LANES>
__global__ __launch_bounds__(BS) void trsv_sf_kernel(int nchunks, const int* __restrict__ cptr,
                                                      const int* __restrict__ rows, const int* __restrict__ irp,
                                                      const int* __restrict__ ja, const V* __restrict__ val,
                                                      const V* __restrict__ b, const V* __restrict__ d, V* t,
                                                      V* y, int* ticket, double* abort, int spin_max) {
  __shared__ int chunk;
  const int lane = threadIdx.x % LANES, grp = threadIdx.x / LANES;
  for (;;) {
    if (threadIdx.x == 0) {
      const int c = atomicAdd(ticket, 1);
      if (c == nchunks + (int)gridDim.x - 1) atomicExch(ticket, 0);  // every draw of this launch is done
      chunk = c;
    }
    __syncthreads();
    const int c = chunk;
    __syncthreads();
    if (c >= nchunks) return;
    const int q = cptr[c] + grp;
    if (q < cptr[c + 1]) {
      const int i = rows[q];
      const V ti = tri_row_sf<V, SC, TR, LANES>(i, lane, irp, ja, val, b, d, t, abort, spin_max);
      if (lane == 0) {
        if (BETA == 0) y[i] = ti;
        else if (BETA == 1) y[i] = y[i] + ti;
        __hip_atomic_store(t + i, ti, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

// Triangular solve with a factor of the 7-point stencil on an nx x ny x nz
// grid in natural order (lower: columns r - nx ny, r - nx, r - 1 where they
// exist; BACK, upper: r + 1, r + nx, r + nx ny): the ILU(0) factors (SC 0,
// unit lower: t_i = b_i - acc; SC 1, upper with the D scaling:
// t_i = d_i b_i - acc) or the triangular parts of Gauss-Seidel (SC 2:
// t_i = (b_i - acc) / d_i).  One wavefront per tile of 8 x 8 grid lines
// along x: lane (jj, kk) owns the line (j, k) = (8 J + jj, 8 K + kk) and the
// tile steps through its diagonal planes i + jj + kk = s (mirrored for
// BACK).  A row's neighbours in j and k inside the tile were computed by the
// neighbouring lanes in the previous step and come by cross-lane reads; its
// neighbour in i is the lane's own previous value; only the tile's two faces
// wait for values in memory (sentinel and bounded wait as in trsv_sf_kernel).
// Tiles are drawn through a ticket in an order in which both tiles a tile
// waits for were drawn before it.  Per row the arithmetic of tri_row_sf
// (products added in ascending column order), so the values are those of
// the other substitutions bit for bit.
// Every operand is in the tiles' step order: position ((T (nx + 14) + s) 64
// + lane) for the point that lane (jj, kk) of tile T reaches at step
// s = i + jj + kk of the forward solve (g: the factor's values by stencil
// position, 3 arrays of sz; d, b and t), so that at every step the 64 lanes
// read and write 64 consecutive values instead of one value in each of 64
// grid lines; grid_to_steps_kernel / grid_from_steps_kernel convert the
// input and the output.
template <typename V, int SC, bool BACK>
__global__ __launch_bounds__(64) void trsv_grid_kernel(int nx, int ny, int nz, int64_t sz, const V* __restrict__ g,
                                                       const V* __restrict__ d, const V* __restrict__ b, V* t,
                                                       int* ticket, double* abort, int spin_max) {
  const int lane = threadIdx.x;
  const int jj = lane & 7, kk = lane >> 3;
  const int ntj = (ny + 7) >> 3, ntk = (nz + 7) >> 3, ntiles = ntj * ntk;
  const int64_t ns = nx + 14;
  const int mj = BACK ? 7 - jj : jj, mk = BACK ? 7 - kk : kk;  // position in the solve's order
  const int ljn = BACK ? (lane + 1) & 63 : (lane + 63) & 63;  // lane of the neighbour in j
  const int lkn = BACK ? (lane + 8) & 63 : (lane + 56) & 63;  // lane of the neighbour in k
  for (;;) {
    int c = 0;
    if (lane == 0) {
      c = atomicAdd(ticket, 1);
      if (c == ntiles + (int)gridDim.x - 1) atomicExch(ticket, 0);  // every draw of this launch is done
    }
    c = __shfl(c, 0);
    if (c >= ntiles) return;
    const int tix = BACK ? ntiles - 1 - c : c;
    const int j = (tix % ntj) * 8 + jj, k = (tix / ntj) * 8 + kk;
    const bool valid = j < ny && k < nz;
    const bool hasj = BACK ? j + 1 < ny : j > 0;
    const bool hask = BACK ? k + 1 < nz : k > 0;
    const bool facej = hasj && mj == 0, facek = hask && mk == 0;
    const int64_t base = (int64_t)tix * ns * 64;  // this tile's first position
    // the face neighbours' positions at i = 0: (i, j -/+ 1, k) is lane (7 or 0,
    // kk) of the tile before / after in j, (i, j, k -/+ 1) lane (jj, 7 or 0)
    // of the tile before / after in k
    const int64_t fj0 = BACK ? base + ns * 64 + (int64_t)kk * 64 + 8 * kk
                             : base - ns * 64 + (int64_t)(7 + kk) * 64 + 7 + 8 * kk;
    const int64_t fk0 = BACK ? base + (int64_t)ntj * ns * 64 + (int64_t)jj * 64 + jj
                             : base - (int64_t)ntj * ns * 64 + (int64_t)(jj + 7) * 64 + jj + 56;
    V prev = (V)0;
    for (int st = 0; st < nx + 14; ++st) {
      const V vj = __shfl(prev, ljn), vk = __shfl(prev, lkn);
      const int mi = st - mj - mk;
      if (valid && mi >= 0 && mi < nx) {
        const int i = BACK ? nx - 1 - mi : mi;
        const int64_t q = base + (int64_t)(i + jj + kk) * 64 + lane;  // this point's position
        const bool hasi = BACK ? i + 1 < nx : i > 0;
        const V g0 = g[q], g1 = g[sz + q], g2 = g[2 * sz + q];
        const V bi = (SC == 1) ? d[q] * b[q] : b[q];
        const V tj = facej ? sf_wait(t + fj0 + (int64_t)i * 64, abort, spin_max) : vj;
        const V tk = facek ? sf_wait(t + fk0 + (int64_t)i * 64, abort, spin_max) : vk;
        V acc = (V)0;
        if (!BACK) {
          if (hask) acc = acc + g0 * tk;
          if (hasj) acc = acc + g1 * tj;
          if (hasi) acc = acc + g2 * prev;
        } else {
          if (hasi) acc = acc + g0 * prev;
          if (hasj) acc = acc + g1 * tj;
          if (hask) acc = acc + g2 * tk;
        }
        const V ti = (SC == 2) ? (bi - acc) / d[q] : bi - acc;
        __hip_atomic_store(t + q, ti, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        prev = ti;
      }
    }
  }
}
// xs = x in the tiles' step order (0 at the positions no grid point takes).
// One workgroup per tile and 64 steps: the 78 values of each of the tile's 64
// grid lines that those steps take are read line by line into LDS, then the
// 64 steps' positions are written row by row (both sides consecutive).
template <typename V>
__global__ __launch_bounds__(BS) void grid_to_steps_kernel(int nx, int ny, int nz, const V* __restrict__ x,
                                                           V* __restrict__ xs) {
  constexpr int SC = 64, W = SC + 14;
  __shared__ V sh[64][W + 1];
  const int ntj = (ny + 7) >> 3, nsc = (nx + 14 + SC - 1) / SC;
  const int tile = blockIdx.x / nsc, s0 = (blockIdx.x % nsc) * SC;
  const int64_t ns = nx + 14;
  const int J = tile % ntj, K = tile / ntj;
  for (int e = threadIdx.x; e < 64 * W; e += BS) {  // line l, i = s0 - 14 + w
    const int l = e / W, w = e % W;
    const int j = J * 8 + (l & 7), k = K * 8 + (l >> 3), i = s0 - 14 + w;
    V v = (V)0;
    if (j < ny && k < nz && i >= 0 && i < nx) v = x[i + (int64_t)nx * (j + (int64_t)ny * k)];
    sh[l][w] = v;
  }
  __syncthreads();
  const int64_t base = (int64_t)tile * ns * 64;
  for (int e = threadIdx.x; e < SC * 64; e += BS) {  // step s0 + e / 64, lane e % 64
    const int st = s0 + e / 64, l = e % 64;
    if (st >= ns) break;
    const int w = st - (l & 7) - (l >> 3) - (s0 - 14);  // i - (s0 - 14), in [0, W)
    xs[base + (int64_t)st * 64 + l] = sh[l][w];
  }
}
// ts back into natural order: t = ts (when t is not null) and y = ts
// (BETA 0), y = y + ts (BETA 1) or y untouched (BETA 2).  One workgroup per
// tile and 64 points of its lines: the 78 steps holding them are read row by
// row into LDS, then each line's 64 points are written (both sides
// consecutive).
template <typename V, int BETA>
__global__ __launch_bounds__(BS) void grid_from_steps_kernel(int nx, int ny, int nz, const V* __restrict__ ts,
                                                             V* __restrict__ t, V* __restrict__ y) {
  constexpr int IC = 64, W = IC + 14;
  __shared__ V sh[W][64 + 1];
  const int ntj = (ny + 7) >> 3, nic = (nx + IC - 1) / IC;
  const int tile = blockIdx.x / nic, i0 = (blockIdx.x % nic) * IC;
  const int64_t ns = nx + 14;
  const int J = tile % ntj, K = tile / ntj;
  const int64_t base = (int64_t)tile * ns * 64;
  for (int e = threadIdx.x; e < W * 64; e += BS) {  // step i0 + e / 64, lane e % 64
    const int w = e / 64, l = e % 64;
    const int st = i0 + w;
    sh[w][l] = st < ns ? ts[base + (int64_t)st * 64 + l] : (V)0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * IC; e += BS) {  // line e / IC, i = i0 + e % IC
    const int l = e / IC, i = i0 + e % IC;
    const int j = J * 8 + (l & 7), k = K * 8 + (l >> 3);
    if (j >= ny || k >= nz || i >= nx) continue;
    const V v = sh[i - i0 + (l & 7) + (l >> 3)][l];
    const int64_t r = i + (int64_t)nx * (j + (int64_t)ny * k);
    if (t) t[r] = v;
    if (BETA == 1) y[r] = y[r] + v;
    else if (BETA == 0) y[r] = v;
  }
}

// Whole ILU(0) apply for a small matrix in one workgroup: level sets
// processed in order with a barrier in between.
template <typename V, int BETA, bool TR, int LANES, bool LDS = false>
__global__ __launch_bounds__(1024) void ilu_small_kernel(
    int n, const int* __restrict__ lirp, const int* __restrict__ lja, const V* __restrict__ lval,
    const int* __restrict__ uirp, const int* __restrict__ uja, const V* __restrict__ uval,
    const V* __restrict__ d, const int* __restrict__ lptr, const int* __restrict__ lrows, int nlevl,
    const int* __restrict__ uptr, const int* __restrict__ urows, int nlevu, const V* __restrict__ x,
    V* t1, V* t2, V* __restrict__ y) {
  // LDS: both intermediate vectors live in the workgroup's LDS (2 n values,
  // dynamic shared memory), so a level set's reads of the previous ones
  // never leave the CU; the result goes to y only
  extern __shared__ double ilu_lds[];
  V* s1 = LDS ? (V*)ilu_lds : t1;
  V* s2 = LDS ? (V*)ilu_lds + n : t2;
  const int lane = threadIdx.x % LANES, grp = threadIdx.x / LANES, ngrp = blockDim.x / LANES;
  for (int lv = 0; lv < nlevl; ++lv) {
    for (int q = lptr[lv] + grp; q < lptr[lv + 1]; q += ngrp) {
      const int i = lrows[q];
      const V ti = tri_row<V, 0, TR, LANES>(i, lane, lirp, lja, lval, x, nullptr, s1);
      if (lane == 0) s1[i] = ti;
    }
    __syncthreads();
  }
  for (int lv = 0; lv < nlevu; ++lv) {
    for (int q = uptr[lv] + grp; q < uptr[lv + 1]; q += ngrp) {
      const int i = urows[q];
      const V ti = tri_row<V, 1, TR, LANES>(i, lane, uirp, uja, uval, s1, d, s2);
      if (lane == 0) {
        s2[i] = ti;
        if (BETA == 0) y[i] = ti;
        else y[i] = y[i] + ti;
      }
    }
    __syncthreads();
  }
}

// Dense LU solve (partial pivoting, row-major LU) in one workgroup;
// coarsest-level DENSE_LU solver for small coarse matrices.
template <typename V, int BETA>
__global__ __launch_bounds__(1024) void dense_lu_small_kernel(int n, const double* __restrict__ lu,
                                                              const int* __restrict__ piv,
                                                              const V* __restrict__ x, double* __restrict__ w,
                                                              V* __restrict__ y) {
  __shared__ double red[1024 / 64];
  for (int i = threadIdx.x; i < n; i += blockDim.x) w[i] = (double)x[i];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < n; ++k) {
      const int p = piv[k];
      if (p != k) { double tmp = w[k]; w[k] = w[p]; w[p] = tmp; }
    }
  __syncthreads();
  for (int i = 0; i < n; ++i) {  // forward (unit lower)
    double s = 0.0;
    for (int j = threadIdx.x; j < i; j += blockDim.x) s += lu[(size_t)i * n + j] * w[j];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += red[q];
      w[i] = w[i] - t;
    }
    __syncthreads();
  }
  for (int i = n - 1; i >= 0; --i) {  // backward
    double s = 0.0;
    for (int j = i + 1 + threadIdx.x; j < n; j += blockDim.x) s += lu[(size_t)i * n + j] * w[j];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += red[q];
      w[i] = (w[i] - t) / lu[(size_t)i * n + i];
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (BETA == 0) y[i] = (V)w[i];
    else y[i] = y[i] + (V)w[i];
  }
}

// ------------------------------------------------- layout construction
// Device-side conversion of an uploaded CSR operator (int32 row pointers,
// int32 columns, FP64 values) into the layouts of the products; each writes
// exactly the arrays the sequential host conversion would.
__global__ __launch_bounds__(BS) void irp_to32_kernel(int n1, const int64_t* __restrict__ a, int* __restrict__ b) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n1) b[i] = (int)a[i];
}
template <typename V>
__global__ __launch_bounds__(BS) void cvt_val_kernel(int64_t n, const double* __restrict__ a, V* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) b[i] = (V)a[i];
}
// ELL, column-major with leading dimension ld = nr: padding is a zero value
// on column min(i, nc-1)
template <typename V>
__global__ __launch_bounds__(BS) void ell_fill_kernel(int nr, int nc, int w, const int* __restrict__ irp,
                                                      const int* __restrict__ ja, const double* __restrict__ val,
                                                      int* __restrict__ oja, V* __restrict__ oval) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= nr) return;
  const int k0 = irp[i], len = irp[i + 1] - k0;
  for (int k = 0; k < w; ++k) {
    const int64_t q = (int64_t)k * nr + i;
    if (k < len) {
      oja[q] = ja[k0 + k];
      oval[q] = (V)val[k0 + k];
    } else {
      oja[q] = min(i, nc - 1);
      oval[q] = (V)0;
    }
  }
}
// SELL-64: position p of slice p/64 holds row perm[p] (or padding past nr)
template <typename V>
__global__ __launch_bounds__(BS) void sell_fill_kernel(int nr, int nc, int npos, const int* __restrict__ perm,
                                                       const int* __restrict__ sptr, const int* __restrict__ irp,
                                                       const int* __restrict__ ja, const double* __restrict__ val,
                                                       int* __restrict__ oja, V* __restrict__ oval) {
  const int p = blockIdx.x * BS + threadIdx.x;
  if (p >= npos) return;
  const int sl = p / 64, l = p % 64;
  const int row = (p < nr) ? perm[p] : -1;
  const int base = sptr[sl], w = (sptr[sl + 1] - base) / 64;
  const int k0 = (row >= 0) ? irp[row] : 0;
  const int len = (row >= 0) ? irp[row + 1] - k0 : 0;
  for (int k = 0; k < w; ++k) {
    const int64_t q = (int64_t)base + (int64_t)k * 64 + l;
    if (k < len) {
      oja[q] = ja[k0 + k];
      oval[q] = (V)val[k0 + k];
    } else {
      oja[q] = (row >= 0) ? min(row, nc - 1) : 0;
      oval[q] = (V)0;
    }
  }
}
// DIA: val[d * nr + i] = a(i, i + off[d]), zero where row i has no such entry
template <typename V>
__global__ __launch_bounds__(BS) void dia_fill_kernel(int nr, int w, DiaOff off, const int* __restrict__ irp,
                                                      const int* __restrict__ ja, const double* __restrict__ val,
                                                      V* __restrict__ oval) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= nr) return;
  for (int d = 0; d < w; ++d) oval[(int64_t)d * nr + i] = (V)0;
  for (int k = irp[i]; k < irp[i + 1]; ++k) {
    const int o = ja[k] - i;
    int d = 0;
    while (d < w - 1 && off.o[d] != o) ++d;
    oval[(int64_t)d * nr + i] = (V)val[k];
  }
}

// DIA operator with its columns scaled: oval[d*nr + i] = val[d*nr + i] * c[j],
// j = i + off[d] clamped as in the products (padding stays zero)
template <typename V>
__global__ __launch_bounds__(BS) void dia_colscale_kernel(int nr, int nc, int w, DiaOff off,
                                                          const V* __restrict__ val, const V* __restrict__ c,
                                                          V* __restrict__ oval) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= nr) return;
  for (int d = 0; d < w; ++d) {
    const int j = min(max(i + off.o[d], 0), nc - 1);
    oval[(int64_t)d * nr + i] = val[(int64_t)d * nr + i] * c[j];
  }
}

// ------------------------------------------------- dense tail operator
// y = B x for a dense row-major n x n operator: one wavefront per row, the
// lanes over the columns, then a fixed-order reduction across the lanes.
template <typename V>
__global__ __launch_bounds__(BS) void tail_gemv_kernel(int n, const V* __restrict__ B, const V* __restrict__ x,
                                                       V* __restrict__ y) {
  const int row = blockIdx.x * (BS / 64) + (int)(threadIdx.x / 64);
  const int lane = threadIdx.x % 64;
  if (row >= n) return;  // whole wavefronts
  const V* b = B + (size_t)row * n;
  V acc = (V)0;
  for (int j = lane; j < n; j += 64) acc += b[j] * x[j];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) y[row] = acc;
}
// x = e_j
template <typename V>
__global__ __launch_bounds__(BS) void unit_vec_kernel(int n, V* __restrict__ x, int j) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) x[i] = (i == j) ? (V)1 : (V)0;
}
// column j of the row-major n x n B = y
template <typename V>
__global__ __launch_bounds__(BS) void col_store_kernel(int n, const V* __restrict__ y, V* __restrict__ B, int j) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) B[(size_t)i * n + j] = y[i];
}
// ND D of a SELD operator: val times c at each entry's column (decoded as
// in seld_kernel), one thread per position
template <typename V>
__global__ __launch_bounds__(BS) void seld_colscale_kernel(int npos, int nc, const int* __restrict__ sptr,
                                                           const int* __restrict__ col0,
                                                           const uint16_t* __restrict__ dja,
                                                           const V* __restrict__ val, const V* __restrict__ c,
                                                           V* __restrict__ oval) {
  const int p = blockIdx.x * BS + threadIdx.x;
  if (p >= npos) return;
  const int s = p >> 6, l = p & 63;
  const int base = sptr[s] + l;
  const int w = (sptr[s + 1] - sptr[s]) >> 6;
  int col = col0[p] - 1;
  for (int k = 0; k < w; ++k) {
    const int i = base + 64 * k;
    col += (int)dja[i] + 1;
    oval[i] = val[i] * c[min(col, nc - 1)];
  }
}

// oval[k] = val[k] * c[ja[k]] over every stored entry of an ELL, SELL or CSR
// operator (padding entries hold a zero value on a valid column)
template <typename V>
__global__ __launch_bounds__(BS) void colscale_kernel(int64_t tot, const int* __restrict__ ja,
                                                      const V* __restrict__ val, const V* __restrict__ c,
                                                      V* __restrict__ oval) {
  const int64_t k = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (k < tot) oval[k] = val[k] * c[ja[k]];
}

}  // namespace
}  // namespace k
}  // namespace mld
