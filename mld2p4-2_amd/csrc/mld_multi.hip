// Batched preconditioner application and Krylov solves on one GPU:
//   * block V-cycle: NB = 32 right-hand sides per system, every sparse
//     operator applied as a sparse-times-dense product (SpMM) over
//     row-interleaved multi-vectors;
//   * grouped systems: G independent systems of different sizes, each with
//     its own hierarchy, run as one V-cycle over block-diagonal unions of
//     their levels, and their CG/FCG iterations captured once into a
//     hipGraph that is replayed until every system has converged.
// The cycle is mld_d_inner_mult (mlprec/impl/mld_dmlprec_aply.f90:590-809)
// for each system; the smoothers are the Jacobi-type sweeps of
// mld_d_jac_smoother_apply_vect.f90:38-215 (point or l1 diagonal); the
// coarsest solve of each system is applied as its explicit operator
// (formed on the host from that level's smoother data: ILU(0) solve, dense
// LU solve or Jacobi sweeps), a dense block of the union.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "mld_devcommon.hpp"
#include "mld_kernels.hpp"
#include "mld_multi.hpp"

namespace mld {
namespace mk {

using namespace mld::k;

constexpr int MB = 32;  // columns of a block multi-vector

// ----------------------------------------------------- block epilogues
// operator()(row, idx = row*MB + column, acc) for one column; vec<R>(row,
// idx, acc) for R adjacent columns with R-wide loads and stores (idx is a
// multiple of R, and every multi-vector starts on a 256-byte boundary)
template <typename V, int R>
using bv_t = V __attribute__((ext_vector_type(R)));

template <typename V, int R>
__device__ __forceinline__ bv_t<V, R>& bvat(V* p, int64_t i) {
  return *reinterpret_cast<bv_t<V, R>*>(p + i);
}
template <typename V, int R>
__device__ __forceinline__ bv_t<V, R> bvld(const V* p, int64_t i) {
  return *reinterpret_cast<const bv_t<V, R>*>(p + i);
}

// Stores of the epilogues: NT = non-temporal (outputs of the large DIA
// levels, not re-read by the same kernel; see stnt)
template <bool NT, typename V>
__device__ __forceinline__ void bst(V* p, int64_t i, V v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p + i);
  else p[i] = v;
}
template <bool NT, typename V, int R>
__device__ __forceinline__ void bvst(V* p, int64_t i, bv_t<V, R> v) {
  if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<bv_t<V, R>*>(p + i));
  else bvat<V, R>(p, i) = v;
}

template <typename V, bool NT = false>
struct BSet {
  V* __restrict__ y;
  __device__ __forceinline__ void operator()(int, int64_t i, V acc) const { bst<NT>(y, i, acc); }
  template <int R>
  __device__ __forceinline__ void vec(int, int64_t i, bv_t<V, R> a) const { bvst<NT, V, R>(y, i, a); }
};
template <typename V, bool NT = false>
struct BAdd {
  V* __restrict__ y;
  __device__ __forceinline__ void operator()(int, int64_t i, V acc) const { bst<NT>(y, i, y[i] + acc); }
  template <int R>
  __device__ __forceinline__ void vec(int, int64_t i, bv_t<V, R> a) const {
    bvst<NT, V, R>(y, i, bvld<V, R>(y, i) + a);
  }
};
template <typename V, bool NT = false>
struct BRes {
  const V* __restrict__ b;
  V* __restrict__ r;
  __device__ __forceinline__ void operator()(int, int64_t i, V acc) const { bst<NT>(r, i, b[i] - acc); }
  template <int R>
  __device__ __forceinline__ void vec(int, int64_t i, bv_t<V, R> a) const {
    bvst<NT, V, R>(r, i, bvld<V, R>(b, i) - a);
  }
};
template <typename V, bool ADD, bool NT = false>
struct BJac {  // t = d_row*(b - ND y);  y = t  or  y = y + t
  const V* __restrict__ d;
  const V* __restrict__ b;
  V* __restrict__ y;
  __device__ __forceinline__ void operator()(int row, int64_t i, V acc) const {
    const V t = d[row] * (b[i] - acc);
    if (ADD) bst<NT>(y, i, y[i] + t);
    else bst<NT>(y, i, t);
  }
  template <int R>
  __device__ __forceinline__ void vec(int row, int64_t i, bv_t<V, R> a) const {
    const bv_t<V, R> t = d[row] * (bvld<V, R>(b, i) - a);
    if (ADD) bvst<NT, V, R>(y, i, bvld<V, R>(y, i) + t);
    else bvst<NT, V, R>(y, i, t);
  }
};

// Last post-smoothing sweep of the fine level writing the cycle's FP64
// output directly: yn = yin + d_row*(b - ND y); yout = (double)yn.
template <typename V, bool NT = false>
struct BJacOut {
  const V* __restrict__ d;
  const V* __restrict__ b;
  const V* __restrict__ yin;
  double* __restrict__ yout;
  __device__ __forceinline__ void operator()(int row, int64_t i, V acc) const {
    const V t = d[row] * (b[i] - acc);
    bst<NT>(yout, i, (double)(yin[i] + t));
  }
  template <int R>
  __device__ __forceinline__ void vec(int row, int64_t i, bv_t<V, R> a) const {
    const bv_t<V, R> t = d[row] * (bvld<V, R>(b, i) - a);
    bvst<NT, double, R>(yout, i, __builtin_convertvector(bvld<V, R>(yin, i) + t, bv_t<double, R>));
  }
};

// The NT form of a block epilogue (same data, non-temporal stores)
template <class E>
struct BNtForm { using type = void; };
template <typename V>
struct BNtForm<BSet<V, false>> {
  using type = BSet<V, true>;
  static type make(const BSet<V, false>& e) { return type{e.y}; }
};
template <typename V>
struct BNtForm<BAdd<V, false>> {
  using type = BAdd<V, true>;
  static type make(const BAdd<V, false>& e) { return type{e.y}; }
};
template <typename V>
struct BNtForm<BRes<V, false>> {
  using type = BRes<V, true>;
  static type make(const BRes<V, false>& e) { return type{e.b, e.r}; }
};
template <typename V, bool ADD>
struct BNtForm<BJac<V, ADD, false>> {
  using type = BJac<V, ADD, true>;
  static type make(const BJac<V, ADD, false>& e) { return type{e.d, e.b, e.y}; }
};
template <typename V>
struct BNtForm<BJacOut<V, false>> {
  using type = BJacOut<V, true>;
  static type make(const BJacOut<V, false>& e) { return type{e.d, e.b, e.yin, e.yout}; }
};
// block DIA levels from this many rows store their outputs non-temporally
// (MLD_BNT_MIN_ROWS overrides it, for measurements)
constexpr int BNT_MIN_ROWS = 1 << 20;
static int bnt_min_rows() {
  const char* e = std::getenv("MLD_BNT_MIN_ROWS");
  return e ? std::atoi(e) : BNT_MIN_ROWS;
}

// Y = A X for row-interleaved X (nc x 32) and Y (nr x 32), CSR A.  A group
// of 32/RPL lanes owns one row; each lane owns RPL adjacent columns.  The
// group reads 32/RPL entries of the row at once (one coalesced load) and
// broadcasts them by cross-lane shuffles; for every entry the group reads
// the 32 values of row ja[k] of X (RPL*sizeof(V) bytes per lane).  Each
// column's sum runs over the row in stored order, as the sequential loop.
template <typename V, int RPL, class E>
__global__ __launch_bounds__(BS) void bspmm_kernel(int nr, const int* __restrict__ irp, const int* __restrict__ ja,
                                                   const V* __restrict__ val, const V* __restrict__ x, E e,
                                                   int xchunk) {
  constexpr int LPR = MB / RPL;
  constexpr int RPB = BS / LPR;
  const int row = xcd_block(xchunk) * RPB + threadIdx.x / LPR;
  const int lane = threadIdx.x % LPR;
  if (row >= nr) return;  // whole lane groups leave together
  const int k0 = irp[row], k1 = irp[row + 1];
  V acc[RPL];
#pragma unroll
  for (int r = 0; r < RPL; ++r) acc[r] = (V)0;
  typedef V vec_t __attribute__((ext_vector_type(RPL)));
  for (int kb = k0; kb < k1; kb += LPR) {
    const int kk = kb + lane;
    V v = (V)0;
    int c = 0;
    if (kk < k1) {
      v = ldnt(val + kk);
      c = ldnt(ja + kk);
    }
    const int cnt = min(LPR, k1 - kb);
    // eight entries at a time: broadcast, then all eight gathers in flight,
    // then the products added in stored order
#pragma unroll
    for (int t0 = 0; t0 < LPR; t0 += 8) {
      if (t0 >= cnt) break;
      V vt[8];
      vec_t xv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        vt[u] = __shfl(v, t0 + u, LPR);
        const int ct = __shfl(c, t0 + u, LPR);
        if (t0 + u < cnt) {
          if constexpr (RPL == 1) xv[u][0] = x[(int64_t)ct * MB + lane];
          else xv[u] = *reinterpret_cast<const vec_t*>(x + (int64_t)ct * MB + lane * RPL);
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (t0 + u < cnt) {
#pragma unroll
          for (int r = 0; r < RPL; ++r) acc[r] = acc[r] + vt[u] * xv[u][r];
        }
      }
    }
  }
  const int64_t base = (int64_t)row * MB + lane * RPL;
  if constexpr (RPL == 1) {
    e(row, base, acc[0]);
  } else {
    bv_t<V, RPL> a;
#pragma unroll
    for (int r = 0; r < RPL; ++r) a[r] = acc[r];
    e.template vec<RPL>(row, base, a);
  }
}

// Y = A X for a DIA operator (see dia_kernel): a lane group per row reads
// the row's diagonal values (one address per group) and the 32-column rows
// i + o_d of X, diagonals in ascending offset order (= stored column order).
// DX: the gathered values are d_j * x_j (the first Jacobi sweep from zero,
// formed here instead of in a separate pass; same products).  With a fixed
// diagonal count all index, value, X and d loads of a lane are issued before
// the first product.
template <typename V, int RPL, int ND, bool DX, class E>
__global__ __launch_bounds__(BS) void bdia_kernel(int nr, int nc, int nd, DiaOff off, const V* __restrict__ val,
                                                  const V* __restrict__ x, E e, const V* __restrict__ dx,
                                                  int xchunk) {
  constexpr int LPR = MB / RPL;
  constexpr int RPB = BS / LPR;
  const int row = xcd_block(xchunk) * RPB + threadIdx.x / LPR;
  const int lane = threadIdx.x % LPR;
  if (row >= nr) return;
  typedef V vec_t __attribute__((ext_vector_type(RPL)));
  V acc[RPL];
#pragma unroll
  for (int r = 0; r < RPL; ++r) acc[r] = (V)0;
  if constexpr (ND > 0) {
    int j[ND];
    V v[ND], dj[ND];
    vec_t w[ND];
#pragma unroll
    for (int k = 0; k < ND; ++k) j[k] = min(max(row + off.o[k], 0), nc - 1);
#pragma unroll
    for (int k = 0; k < ND; ++k) v[k] = val[(int64_t)k * nr + row];
#pragma unroll
    for (int k = 0; k < ND; ++k) w[k] = *reinterpret_cast<const vec_t*>(x + (int64_t)j[k] * MB + lane * RPL);
    if constexpr (DX) {
#pragma unroll
      for (int k = 0; k < ND; ++k) dj[k] = dx[j[k]];
    }
#pragma unroll
    for (int k = 0; k < ND; ++k) {
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        V t = w[k][r];
        if constexpr (DX) t = dj[k] * t;
        acc[r] = acc[r] + v[k] * t;
      }
    }
  } else {
    for (int k = 0; k < nd; ++k) {
      const int j = min(max(row + off.o[k], 0), nc - 1);
      const V v = val[(int64_t)k * nr + row];
      vec_t w = *reinterpret_cast<const vec_t*>(x + (int64_t)j * MB + lane * RPL);
      if constexpr (DX) {
        const V d = dx[j];
#pragma unroll
        for (int r = 0; r < RPL; ++r) w[r] = d * w[r];
      }
#pragma unroll
      for (int r = 0; r < RPL; ++r) acc[r] = acc[r] + v * w[r];
    }
  }
  const int64_t base = (int64_t)row * MB + lane * RPL;
  if constexpr (RPL == 1) {
    e(row, base, acc[0]);
  } else {
    bv_t<V, RPL> a;
#pragma unroll
    for (int r = 0; r < RPL; ++r) a[r] = acc[r];
    e.template vec<RPL>(row, base, a);
  }
}

// y = d_row .* x  (ADD: y += d_row .* x) over n rows of 32 columns
// Chunk (in blocks) of the row-block order of a block product (see
// xcd_block).  A row of a 32-column multi-vector is 128 (FP32) or 256 (FP64)
// bytes, so an XCD's 4 MB L2 holds only 16-32K rows of X.  xmap 1 on a DIA
// operator: chunks of (largest |offset|)/8 rows, so that the eight XCDs
// together advance one offset span per period and the rows a block gathers
// at -o and +o were, or will be, fetched by the same XCD (for a 192^3 grid:
// 4608-row chunks, three of them per XCD in flight = 1.8 MB of FP32 X);
// xmap 1 on CSR: 4096-row chunks; xmap 2: one contiguous range per XCD.
template <typename T>
static int bxchunk(const DMat<T>& m, int rpb, int nb) {
  if (m.xmap == 2) return std::max(1, nb / 8);
  if (m.xmap != 1) return 0;
  int rows = 4096;
  if (m.fmt == FMT_DIA) {
    int mo = 0;
    for (int k = 0; k < m.w; ++k) mo = std::max(mo, std::abs(m.off.o[k]));
    rows = mo / 8;
  }
  return std::max(1, rows / rpb);
}

template <typename V, bool ADD>
__global__ __launch_bounds__(BS) void bvmul_kernel(int n, const V* __restrict__ d, const V* __restrict__ x,
                                                   V* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= (int64_t)n * MB) return;
  const V t = d[i / MB] * x[i];
  if (ADD) y[i] = y[i] + t;
  else y[i] = t;
}

template <typename V>
__global__ __launch_bounds__(BS) void fill64_kernel(int64_t n, V* __restrict__ y, V v) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) y[i] = v;
}

template <typename TO, typename TI>
__global__ __launch_bounds__(BS) void convert64_kernel(int64_t n, const TI* __restrict__ x, TO* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) y[i] = (TO)x[i];
}

// Column-major (ld) <-> row-interleaved (32 columns) through a 64 x 32 LDS
// tile, both sides coalesced.  Columns >= nrhs are zero in the interleaved
// copy and not written back.  The interleaved side is FP64 or the FP32 copy
// of an FP32 preconditioner (rounded on the way in, widened on the way out).
template <typename T>
__global__ __launch_bounds__(BS) void col2il_kernel(int n, int nrhs, const double* __restrict__ src, int64_t ld,
                                                    T* __restrict__ dst) {
  __shared__ double tile[64][MB + 1];
  const int row0 = blockIdx.x * 64;
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < MB / 4; ++k) {
    const int j = t / 64 + 4 * k, i = row0 + t % 64;
    tile[t % 64][j] = (j < nrhs && i < n) ? src[(int64_t)j * ld + i] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int e = t + BS * k, i = e / MB, j = e % MB;
    if (row0 + i < n) dst[(int64_t)(row0 + i) * MB + j] = (T)tile[i][j];
  }
}
template <typename T>
__global__ __launch_bounds__(BS) void il2col_kernel(int n, int nrhs, const T* __restrict__ src,
                                                    double* __restrict__ dst, int64_t ld) {
  __shared__ double tile[64][MB + 1];
  const int row0 = blockIdx.x * 64;
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int e = t + BS * k, i = e / MB, j = e % MB;
    tile[i][j] = (row0 + i < n) ? (double)src[(int64_t)(row0 + i) * MB + j] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < MB / 4; ++k) {
    const int j = t / 64 + 4 * k, i = row0 + t % 64;
    if (j < nrhs && i < n) dst[(int64_t)j * ld + i] = tile[t % 64][j];
  }
}

// ----------------------------------------------- per-column reductions
// A reduction block covers rows [bt.x, bt.y) of system bt.z (never two
// systems).  Column c = system*NB + j.  Partial sums land in
// partial[block*NB + j]; mfin_kernel adds a system's blocks in order.
template <int NB>
__device__ __forceinline__ void col_partial(double acc, double* partial) {
  constexpr int RPS = BS / NB;
  __shared__ double sh[BS];
  sh[threadIdx.x] = acc;
  __syncthreads();
#pragma unroll
  for (int w = RPS / 2; w >= 1; w >>= 1) {
    if ((int)threadIdx.x / NB < w) sh[threadIdx.x] += sh[threadIdx.x + w * NB];
    __syncthreads();
  }
  if ((int)threadIdx.x < NB) partial[(int64_t)blockIdx.x * NB + threadIdx.x] = sh[threadIdx.x];
}

template <int NB>
__global__ __launch_bounds__(BS) void mdot_kernel(const int4* __restrict__ bt, const double* __restrict__ x,
                                                  const double* __restrict__ y, double* __restrict__ partial) {
  constexpr int RPS = BS / NB;
  const int4 b = bt[blockIdx.x];
  const int j = threadIdx.x % NB;
  double acc = 0.0;
  for (int row = b.x + (int)threadIdx.x / NB; row < b.y; row += RPS) {
    const int64_t i = (int64_t)row * NB + j;
    acc += x[i] * y[i];
  }
  col_partial<NB>(acc, partial);
}

// One block per column: sum of the column's partials in block order.
template <int NB>
__global__ __launch_bounds__(BS) void mfin_kernel(const int* __restrict__ bstart, const double* __restrict__ partial,
                                                  double* __restrict__ out) {
  __shared__ double sh[BS];
  const int c = blockIdx.x, s = c / NB, j = c % NB;
  const int b1 = bstart[s + 1];
  double acc = 0.0;
  int b = bstart[s] + threadIdx.x;
  for (; b + 7 * BS < b1; b += 8 * BS) {  // 8 independent loads in flight, added in order
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = partial[(int64_t)(b + u * BS) * NB + j];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += v[u];
  }
  for (; b < b1; b += BS) acc += partial[(int64_t)b * NB + j];
  sh[threadIdx.x] = acc;
  __syncthreads();
#pragma unroll
  for (int w = BS / 2; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = sh[0];
}

// CG direction: p = z (first iteration of the column) or z + (rho/rho_old) p
template <int NB>
__global__ __launch_bounds__(BS) void mp_cg_kernel(const int4* __restrict__ bt, const double* __restrict__ z,
                                                   double* __restrict__ p, const double* __restrict__ rho,
                                                   const double* __restrict__ rho_old, const int* __restrict__ kit,
                                                   const int* __restrict__ done) {
  constexpr int RPS = BS / NB;
  const int4 b = bt[blockIdx.x];
  const int j = threadIdx.x % NB, c = b.z * NB + j;
  if (done[c]) return;
  const bool first = (kit[c] == 0);
  const double beta = first ? 0.0 : rho[c] / rho_old[c];
  for (int row = b.x + (int)threadIdx.x / NB; row < b.y; row += RPS) {
    const int64_t i = (int64_t)row * NB + j;
    p[i] = first ? z[i] : z[i] + beta * p[i];
  }
}

// FCG(1) direction p = z - (gamma/delta_old) p, and sum p.r per column
template <int NB>
__global__ __launch_bounds__(BS) void mp_fcg_kernel(const int4* __restrict__ bt, const double* __restrict__ z,
                                                    double* __restrict__ p, const double* __restrict__ r,
                                                    const double* __restrict__ gamma,
                                                    const double* __restrict__ delta, const int* __restrict__ kit,
                                                    const int* __restrict__ done, double* __restrict__ partial) {
  constexpr int RPS = BS / NB;
  const int4 b = bt[blockIdx.x];
  const int j = threadIdx.x % NB, c = b.z * NB + j;
  const bool skip = done[c] != 0;
  const bool first = (kit[c] == 0);
  const double cf = first ? 0.0 : -(gamma[c] / delta[c]);
  double acc = 0.0;
  if (!skip)
    for (int row = b.x + (int)threadIdx.x / NB; row < b.y; row += RPS) {
      const int64_t i = (int64_t)row * NB + j;
      const double pi = first ? z[i] : z[i] + cf * p[i];
      p[i] = pi;
      acc += pi * r[i];
    }
  col_partial<NB>(acc, partial);
}

// x += a p; r -= a q (a = num/den of the column, 0 when converged or den = 0);
// sum r.r per column
template <int NB>
__global__ __launch_bounds__(BS) void mxr_kernel(const int4* __restrict__ bt, double* __restrict__ x,
                                                 double* __restrict__ r, const double* __restrict__ p,
                                                 const double* __restrict__ q, const double* __restrict__ num,
                                                 const double* __restrict__ den, const int* __restrict__ done,
                                                 double* __restrict__ partial, float* __restrict__ r32) {
  constexpr int RPS = BS / NB;
  const int4 b = bt[blockIdx.x];
  const int j = threadIdx.x % NB, c = b.z * NB + j;
  const double dn = den[c];
  const double a = (done[c] || dn == 0.0) ? 0.0 : num[c] / dn;
  double acc = 0.0;
  for (int row = b.x + (int)threadIdx.x / NB; row < b.y; row += RPS) {
    const int64_t i = (int64_t)row * NB + j;
    if (a != 0.0) {
      x[i] = a * p[i] + x[i];
      r[i] = (-a) * q[i] + r[i];
    }
    const double ri = r[i];
    if (r32) r32[i] = (float)ri;  // FP32 input of the next preconditioner application
    acc += ri * ri;
  }
  col_partial<NB>(acc, partial);
}

// Start of a solve: column norms, converged-at-start columns, active count.
__global__ __launch_bounds__(BS) void mstart_kernel(int ncol, double eps, const double* __restrict__ bn2,
                                                    const double* __restrict__ rr, double* __restrict__ bn,
                                                    int* __restrict__ kit, int* __restrict__ done,
                                                    int* __restrict__ iters, double* __restrict__ errs,
                                                    int* __restrict__ active) {
  __shared__ double lds[BS / 64];
  double cnt = 0.0;
  for (int c = threadIdx.x; c < ncol; c += BS) {
    double b = std::sqrt(bn2[c]);
    if (b == 0.0) b = 1.0;
    bn[c] = b;
    const double err = std::sqrt(rr[c]) / b;
    errs[c] = err;
    iters[c] = 0;
    kit[c] = 0;
    done[c] = (err <= eps) ? 1 : 0;
    if (!done[c]) cnt += 1.0;
  }
  const double tot = block_sum(cnt, lds);
  if (threadIdx.x == 0) active[0] = (int)tot;
}

// End of an iteration: per-column residual, convergence, breakdown.
// brk0/brk1: a zero value means breakdown (CG: rho, sigma; FCG: delta).
__global__ __launch_bounds__(BS) void mcheck_kernel(int ncol, double eps, int itmax, const double* __restrict__ rr,
                                                    const double* __restrict__ bn, const double* __restrict__ brk0,
                                                    const double* __restrict__ brk1, const double* __restrict__ rho,
                                                    double* __restrict__ rho_old, int* __restrict__ kit,
                                                    int* __restrict__ done, int* __restrict__ iters,
                                                    double* __restrict__ errs, int* __restrict__ active) {
  __shared__ double lds[BS / 64];
  double cnt = 0.0;
  for (int c = threadIdx.x; c < ncol; c += BS) {
    if (done[c]) continue;
    const int k = kit[c] + 1;
    kit[c] = k;
    const double err = std::sqrt(rr[c]) / bn[c];
    errs[c] = err;
    iters[c] = k;
    if (rho) rho_old[c] = rho[c];
    const bool brk = (brk0 && brk0[c] == 0.0) || (brk1 && brk1[c] == 0.0);
    if (err <= eps || brk || k >= itmax) done[c] = 1;
    else cnt += 1.0;
  }
  const double tot = block_sum(cnt, lds);
  if (threadIdx.x == 0) active[0] = (int)tot;
}

}  // namespace mk

using namespace mk;

// ---------------------------------------------------------------------------
// Host side: union of the systems' hierarchies.
// ---------------------------------------------------------------------------
struct HULevel {
  int n = 0, ni = 0;
  HCsr a, nd, nd2, p, r, mc;
  std::vector<double> dv, dv2;
  bool has2 = false;
  int sweeps_pre = 0, sweeps_post = 0, ml_cycle = VCYCLE;
};

static void csr_start(HCsr& u) {
  u.clear();
  u.irp.assign(1, 0);
}
// Host threads for the union build (OMP_NUM_THREADS, default 8).
static int host_threads() {
  int nt = 8;
  if (const char* s = std::getenv("OMP_NUM_THREADS")) nt = std::atoi(s);
  return std::max(1, std::min(nt, 64));
}

// u = the rows of parts[0], parts[1], ... (each with its column shift): sizes
// summed first, then the parts' rows copied into their own ranges by host
// threads, in tasks of at most 2^18 rows.
static void concat_rows(HCsr& u, const std::vector<std::pair<const HCsr*, int64_t>>& parts, int nc) {
  const size_t np = parts.size();
  std::vector<int64_t> r0(np + 1, 0), e0(np + 1, 0);
  for (size_t q = 0; q < np; ++q) {
    const HCsr& m = *parts[q].first;
    r0[q + 1] = r0[q] + m.nr;
    e0[q + 1] = e0[q] + (m.nr ? m.irp[m.nr] - m.irp[0] : 0);
  }
  u.clear();
  u.nr = (int)r0[np];
  u.nc = nc;
  u.irp.resize((size_t)u.nr + 1);
  u.ja.resize((size_t)e0[np]);
  u.val.resize((size_t)e0[np]);
  u.irp[0] = 0;
  constexpr int TASK_ROWS = 1 << 18;
  std::vector<std::pair<size_t, int>> tasks;  // (part, first row)
  for (size_t q = 0; q < np; ++q)
    for (int i = 0; i < parts[q].first->nr; i += TASK_ROWS) tasks.push_back({q, i});
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (size_t t = next++; t < tasks.size(); t = next++) {
      const size_t q = tasks[t].first;
      const HCsr& m = *parts[q].first;
      const int64_t co = parts[q].second;
      const int i0 = tasks[t].second, i1 = std::min(m.nr, i0 + TASK_ROWS);
      const int64_t base = m.irp[0];
      for (int64_t k = m.irp[i0]; k < m.irp[i1]; ++k) {
        u.ja[(size_t)(e0[q] + k - base)] = (int)(m.ja[k] + co);
        u.val[(size_t)(e0[q] + k - base)] = m.val[k];
      }
      for (int i = i0; i < i1; ++i) u.irp[(size_t)(r0[q] + i + 1)] = e0[q] + (m.irp[i + 1] - base);
    }
  };
  const int nt = (int)std::min<size_t>((size_t)host_threads(), std::max<size_t>(1, tasks.size()));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
}


// y = S x for the coarsest-level smoother data of one system
// (mld_d_jac_smoother_apply_vect.f90:107-109 shortcut or Jacobi sweeps,
// init 'Z'); the explicit operator of the coarse solve is built from it.
static int coarse_apply_host(const SmootherData& sd, int sweeps, const std::vector<double>& lu,
                             const std::vector<int>& piv, const std::vector<double>& x, std::vector<double>& y) {
  const int n = (int)x.size();
  y.assign(n, 0.0);
  if (sd.smoother == NOPREC) {
    y = x;
    return 0;
  }
  const bool shortcut = (sweeps <= 1 || sd.nnz_nd == 0);
  if (!shortcut && !(sd.solver == DIAG_SCALE || sd.solver == L1_DIAG)) return E_UNSUPPORTED;
  switch (sd.solver) {
    case DIAG_SCALE:
    case L1_DIAG: {
      for (int i = 0; i < n; ++i) y[i] = sd.dv[i] * x[i];
      if (!shortcut) {
        std::vector<double> t(n);
        for (int sw = 2; sw <= sweeps; ++sw) {
          for (int i = 0; i < n; ++i) {
            double acc = 0.0;
            for (int64_t k = sd.nd.irp[i]; k < sd.nd.irp[i + 1]; ++k) acc = acc + sd.nd.val[k] * y[sd.nd.ja[k]];
            t[i] = sd.dv[i] * (x[i] - acc);
          }
          y.swap(t);
        }
      }
      return 0;
    }
    case ILU_N: {  // y = U^-1 (d .* (L^-1 x)), unit triangular factors
      std::vector<double> t(n);
      for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int64_t k = sd.l.irp[i]; k < sd.l.irp[i + 1]; ++k) acc = acc + sd.l.val[k] * t[sd.l.ja[k]];
        t[i] = x[i] - acc;
      }
      for (int i = n - 1; i >= 0; --i) {
        double acc = 0.0;
        for (int64_t k = sd.u.irp[i]; k < sd.u.irp[i + 1]; ++k) acc = acc + sd.u.val[k] * y[sd.u.ja[k]];
        y[i] = sd.dv[i] * t[i] - acc;
      }
      return 0;
    }
    case DENSE_LU: {  // P A = L U (row-major, unit L), solve
      std::vector<double> t(x);
      for (int i = 0; i < n; ++i) std::swap(t[i], t[piv[i]]);
      for (int i = 0; i < n; ++i) {
        double acc = t[i];
        for (int k = 0; k < i; ++k) acc -= lu[(size_t)i * n + k] * t[k];
        t[i] = acc;
      }
      for (int i = n - 1; i >= 0; --i) {
        double acc = t[i];
        for (int k = i + 1; k < n; ++k) acc -= lu[(size_t)i * n + k] * t[k];
        t[i] = acc / lu[(size_t)i * n + i];
      }
      y = t;
      return 0;
    }
  }
  return E_UNSUPPORTED;
}

static int host_lu(int n, std::vector<double>& a, std::vector<int>& piv) {
  piv.assign(n, 0);
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(a[(size_t)i * n + k]) > std::fabs(a[(size_t)p * n + k])) p = i;
    piv[k] = p;
    if (a[(size_t)p * n + k] == 0.0) return E_PIVOT;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(a[(size_t)k * n + j], a[(size_t)p * n + j]);
    for (int i = k + 1; i < n; ++i) {
      const double f = a[(size_t)i * n + k] / a[(size_t)k * n + k];
      a[(size_t)i * n + k] = f;
      for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= f * a[(size_t)k * n + j];
    }
  }
  return 0;
}

// Explicit n x n operator of the coarse solve, appended to `mc` as dense rows
// with column offset co.
static int append_coarse_operator(HCsr& mc, const Level& L, int64_t co, std::string& err) {
  const SmootherData& sd = L.smd;
  const int n = L.a.nr;
  if (n > 4096) {
    err = "coarsest level of a batched system larger than 4096 rows";
    return E_UNSUPPORTED;
  }
  std::vector<double> lu;
  std::vector<int> piv;
  if (sd.solver == DENSE_LU && sd.smoother != NOPREC) {
    lu = sd.lu;
    int rc = host_lu(n, lu, piv);
    if (rc) {
      err = "dense LU of a coarsest level: zero pivot";
      return rc;
    }
  }
  // column j = S e_j; the columns are independent, so they are formed by
  // host threads (OMP_NUM_THREADS of them, default 8), each into its own
  // contiguous column of mcol (column-major)
  std::vector<double> mcol((size_t)n * n);
  int nt = 8;
  if (const char* s = std::getenv("OMP_NUM_THREADS")) nt = std::atoi(s);
  nt = std::max(1, std::min({nt, 64, (n + 15) / 16}));
  std::vector<int> rcs(nt, 0);
  // an exception (bad_alloc) must not leave a worker thread: it is turned
  // into the thread's return code and reported after the join
  auto work = [&](int t) {
    try {
      std::vector<double> e(n, 0.0), y;
      for (int j = t; j < n; j += nt) {
        e[j] = 1.0;
        const int rc = coarse_apply_host(sd, L.parms.sweeps_pre, lu, piv, e, y);
        e[j] = 0.0;
        if (rc) {
          rcs[t] = rc;
          return;
        }
        std::copy(y.begin(), y.end(), mcol.begin() + (size_t)j * n);
      }
    } catch (const std::exception&) {
      rcs[t] = E_INTERNAL;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  for (int t = 0; t < nt; ++t)
    if (rcs[t]) {
      err = rcs[t] == E_INTERNAL ? "host memory exhausted while forming a coarsest-level operator"
                                 : "coarsest-level solver not supported by the batched engines";
      return rcs[t];
    }
  const size_t o = mc.ja.size(), ro = mc.irp.size();
  mc.ja.resize(o + (size_t)n * n);
  mc.val.resize(o + (size_t)n * n);
  mc.irp.resize(ro + (size_t)n);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      mc.ja[o + (size_t)i * n + j] = (int)(co + j);
      mc.val[o + (size_t)i * n + j] = mcol[(size_t)j * n + i];
    }
    mc.irp[ro + (size_t)i] = (int64_t)(o + (size_t)(i + 1) * n);
  }
  mc.nr += n;
  return 0;
}

static int build_union(const std::vector<const Prec*>& S, std::vector<HULevel>& U, std::vector<int64_t>& off0,
                       std::string& err) {
  const int G = (int)S.size();
  if (G == 0) {
    err = "no systems";
    return E_BAD_ARG;
  }
  std::vector<int> nl(G);
  int maxlev = 0, minlev = 1 << 30;
  for (int s = 0; s < G; ++s) {
    if (!S[s]->hier_built || !S[s]->smoothers_built) {
      err = "system " + std::to_string(s) + " is not built";
      return E_NOT_BUILT;
    }
    nl[s] = (int)S[s]->precv.size();
    maxlev = std::max(maxlev, nl[s]);
    minlev = std::min(minlev, nl[s]);
    if (S[s]->outer_sweeps != 1) {
      err = "outer sweeps > 1 are not supported by the batched engines";
      return E_UNSUPPORTED;
    }
  }
  if (minlev == 1 && maxlev > 1) {
    err = "a group must have either one level in every system or at least two in every system";
    return E_UNSUPPORTED;
  }
  std::vector<int> ord(G);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return nl[a] > nl[b]; });
  std::vector<std::vector<int>> present(maxlev);
  std::vector<std::vector<int64_t>> off(maxlev, std::vector<int64_t>(G, -1));
  U.assign(maxlev, HULevel());
  for (int l = 0; l < maxlev; ++l) {
    for (int q = 0; q < G; ++q) {
      const int s = (l == 0) ? q : ord[q];
      if (nl[s] > l) present[l].push_back(s);
    }
    int64_t n = 0, ni = 0;
    bool seen_coarse = false;
    for (int s : present[l]) {
      off[l][s] = n;
      const int ns = S[s]->precv[l].a.nr;
      n += ns;
      if (nl[s] > l + 1) {
        if (seen_coarse) {
          err = "internal: level order";
          return E_INTERNAL;
        }
        ni += ns;
      } else {
        seen_coarse = true;
      }
    }
    if (n >= ((int64_t)1 << 31) / MB) {
      err = "batched union too large";
      return E_UNSUPPORTED;
    }
    U[l].n = (int)n;
    U[l].ni = (int)ni;
  }
  for (int l = 0; l < maxlev; ++l) {
    HULevel& H = U[l];
    csr_start(H.a);
    csr_start(H.nd);
    csr_start(H.nd2);
    csr_start(H.mc);
    std::vector<std::pair<const HCsr*, int64_t>> pa, pnd, pnd2;
    bool first = true;
    for (int s : present[l]) {
      const Level& L = S[s]->precv[l];
      if (nl[s] > l + 1) {  // continuing system: cycle data
        const SmootherData& a1 = L.smd;
        const bool has2 = L.sm2a.allocated;
        const SmootherData& a2 = has2 ? L.smd2 : L.smd;
        for (const SmootherData* sd : {&a1, &a2})
          if (!(sd->solver == DIAG_SCALE || sd->solver == L1_DIAG) || sd->smoother == NOPREC) {
            err = "the batched engines provide point/l1-Jacobi smoothers above the coarsest level";
            return E_UNSUPPORTED;
          }
        const int cyc = (L.parms.ml_cycle == MULT_ML) ? VCYCLE : L.parms.ml_cycle;  // MULT is the V-cycle
        if (first) {
          H.sweeps_pre = L.parms.sweeps_pre;
          H.sweeps_post = L.parms.sweeps_post;
          H.ml_cycle = cyc;
          H.has2 = has2;
          first = false;
        } else if (H.sweeps_pre != L.parms.sweeps_pre || H.sweeps_post != L.parms.sweeps_post ||
                   H.ml_cycle != cyc || H.has2 != has2) {
          err = "systems of a group must share sweeps, cycle and smoother layout per level";
          return E_UNSUPPORTED;
        }
        if (H.ml_cycle != VCYCLE && H.ml_cycle != WCYCLE) {
          err = "only V- and W-cycles in the batched engines";
          return E_UNSUPPORTED;
        }
        pa.push_back({&L.a, off[l][s]});
        pnd.push_back({&a1.nd, off[l][s]});
        H.dv.insert(H.dv.end(), a1.dv.begin(), a1.dv.end());
        if (has2) {
          pnd2.push_back({&a2.nd, off[l][s]});
          H.dv2.insert(H.dv2.end(), a2.dv.begin(), a2.dv.end());
        }
      } else {  // coarsest level of this system: explicit coarse operator
        int rc = append_coarse_operator(H.mc, L, off[l][s] - H.ni, err);
        if (rc) return rc;
      }
    }
    concat_rows(H.a, pa, H.ni);
    concat_rows(H.nd, pnd, H.ni);
    concat_rows(H.nd2, pnd2, H.ni);
    H.mc.nc = H.n - H.ni;
    if (l + 1 < maxlev) {
      HULevel& C = U[l + 1];
      std::vector<std::pair<const HCsr*, int64_t>> pp, pr;
      for (int s : present[l])
        if (nl[s] > l + 1) pp.push_back({&S[s]->precv[l + 1].p, off[l + 1][s]});
      for (int s : present[l + 1]) pr.push_back({&S[s]->precv[l + 1].r, off[l][s]});
      concat_rows(C.p, pp, C.n);
      concat_rows(C.r, pr, H.ni);
      if (C.p.nr != H.ni || C.r.nr != C.n) {
        err = "internal: transfer operator sizes";
        return E_INTERNAL;
      }
    }
  }
  off0.assign(G + 1, 0);
  for (int s = 0; s < G; ++s) off0[s] = off[0][s];
  off0[G] = U[0].n;
  return 0;
}

// ---------------------------------------------------------------------------
// Device engine.
// ---------------------------------------------------------------------------
template <typename V, int NB>
class MultiT : public MultiEngine {
  struct MLevel {
    int n = 0, ni = 0;
    DMat<V> a, nd, nd2, p, r, mc;
    DMat<V> ndd, ndd2;  // DIA ND (ND2) with columns scaled by dv (dv2): first sweep from zero
    V* dv = nullptr;
    V* dv2 = nullptr;
    int64_t nnz_nd = 0, nnz_nd2 = 0;
    bool has2 = false;
    int sweeps_pre = 0, sweeps_post = 0, ml_cycle = VCYCLE;
    V *x2l = nullptr, *y2l = nullptr, *tx = nullptr, *ty = nullptr, *t1 = nullptr, *t2 = nullptr;
    int64_t nnz[5] = {0, 0, 0, 0, 0};
  };

 public:
  MultiT(int device) : device_(device) {
    HIPCHECK(hipSetDevice(device_));
    HIPCHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    HIPCHECK(hipEventCreateWithFlags(&ev_, hipEventDisableTiming));
  }
  ~MultiT() override {
    (void)hipSetDevice(device_);
    for (auto& c : graphs_) {
      (void)hipGraphExecDestroy(c.exec);
      (void)hipGraphDestroy(c.graph);
    }
    if (hstat_) (void)hipHostFree(hstat_);
    (void)hipStreamSynchronize(s_);
    (void)hipEventDestroy(ev_);
    (void)hipStreamDestroy(s_);
  }

  int build(const std::vector<const Prec*>& systems, std::string& err) {
    std::vector<HULevel> U;
    double t0 = wall_seconds();
    int rc = build_union(systems, U, off0_, err);
    if (rc) return rc;
    phase_log("group: union on host", -1, wall_seconds() - t0);
    t0 = wall_seconds();
    G_ = (int)systems.size();
    const int nlev = (int)U.size();
    L_.resize(nlev);
    for (int l = 0; l < nlev; ++l) {
      HULevel& H = U[l];
      MLevel& D = L_[l];
      D.n = H.n;
      D.ni = H.ni;
      D.sweeps_pre = H.sweeps_pre;
      D.sweeps_post = H.sweeps_post;
      D.ml_cycle = H.ml_cycle;
      D.has2 = H.has2;
      if (D.ni > 0) {
        D.a = up(H.a);
        D.nnz_nd = H.nd.nnz();
        if (D.nnz_nd > 0) D.nd = up(H.nd);
        D.dv = upv(H.dv);
        if (D.nnz_nd > 0) D.ndd = scaled_cols(D.nd, D.dv);
        if (D.has2) {
          D.nnz_nd2 = H.nd2.nnz();
          if (D.nnz_nd2 > 0) D.nd2 = up(H.nd2);
          D.dv2 = upv(H.dv2);
          if (D.nnz_nd2 > 0) D.ndd2 = scaled_cols(D.nd2, D.dv2);
        }
      }
      if (l > 0) {
        D.p = up(H.p);
        D.r = up(H.r);
      }
      if (D.n > D.ni) D.mc = up(H.mc);
      D.nnz[0] = H.a.nnz(); D.nnz[1] = l > 0 ? H.p.nnz() : 0; D.nnz[2] = l > 0 ? H.r.nnz() : 0;
      D.nnz[3] = H.nd.nnz() + H.nd2.nnz(); D.nnz[4] = H.mc.nnz();
      const size_t nv = (size_t)D.n * NB, ni = (size_t)std::max(1, D.ni) * NB;
      D.x2l = ar_.alloc<V>(nv);
      D.y2l = ar_.alloc<V>(nv);
      D.tx = ar_.alloc<V>(ni);
      D.ty = ar_.alloc<V>(ni);
      D.t1 = ar_.alloc<V>(ni);
      D.t2 = ar_.alloc<V>(ni);
    }
    phase_log("group: level uploads", -1, wall_seconds() - t0);
    t0 = wall_seconds();
    n0_ = L_[0].n;
    // Krylov: FP64 union of the fine matrices (every system's level 0)
    {
      HCsr A0;
      std::vector<std::pair<const HCsr*, int64_t>> pa;
      for (int s = 0; s < G_; ++s) pa.push_back({&systems[s]->precv[0].a, off0_[s]});
      concat_rows(A0, pa, n0_);
      akry_ = up64(A0);
    }
    phase_log("group: Krylov operator", -1, wall_seconds() - t0);
    // reduction blocks: rows of one system each
    const int rb = std::max(1, 8192 / NB);
    std::vector<int4> bt;
    std::vector<int> bstart(G_ + 1, 0);
    for (int s = 0; s < G_; ++s) {
      bstart[s] = (int)bt.size();
      for (int64_t r0 = off0_[s]; r0 < off0_[s + 1]; r0 += rb)
        bt.push_back(make_int4((int)r0, (int)std::min<int64_t>(off0_[s + 1], r0 + rb), s, 0));
    }
    bstart[G_] = (int)bt.size();
    nrb_ = (int)bt.size();
    bt_ = ar_.upload(bt.data(), bt.size(), s_);
    bstart_ = ar_.upload(bstart.data(), bstart.size(), s_);
    partial_ = ar_.alloc<double>((size_t)nrb_ * NB);
    ncol_ = G_ * NB;
    for (int i = 0; i < NSC; ++i) sc_[i] = ar_.alloc<double>(ncol_);
    for (int i = 0; i < NIC; ++i) ic_[i] = ar_.alloc<int>(ncol_ + 1);
    for (int i = 0; i < 6; ++i) kv_[i] = ar_.alloc<double>((size_t)n0_ * NB);
    HIPCHECK(hipHostMalloc((void**)&hstat_, sizeof(int) * 4, hipHostMallocDefault));
    HIPCHECK(hipStreamSynchronize(s_));
    return 0;
  }

  int64_t rows() const override { return n0_; }
  int nsys() const override { return G_; }
  int nb() const override { return NB; }
  const std::vector<int64_t>& offsets() const override { return off0_; }
  size_t device_bytes() const override { return ar_.bytes; }
  int nlev() const override { return (int)L_.size(); }
  int level_info(int lev, int64_t* info) const override {
    if (lev < 0 || lev >= (int)L_.size()) return E_BAD_LEVEL;
    const MLevel& D = L_[lev];
    info[0] = D.n; info[1] = D.ni;
    for (int i = 0; i < 5; ++i) info[2 + i] = D.nnz[i];
    info[7] = D.sweeps_pre; info[8] = D.sweeps_post;
    info[9] = (NB == MB) ? D.a.rpl : D.a.fmt;
    return 0;
  }

  int apply_cm(int nrhs, const double* x, int64_t ldx, double* y, int64_t ldy, hipStream_t caller) override {
    HIPCHECK(hipSetDevice(device_));
    if (nrhs < 1 || nrhs > NB || ldx < n0_ || ldy < n0_) return E_BAD_ARG;
    wait_caller(caller);
    if constexpr (std::is_same<V, float>::value && NB == MB) {
      // FP32 preconditioner: the layout change rounds the input into the
      // fine level's FP32 buffer and widens the FP32 result on the way out
      // (the same values as rounding, and widening, in separate passes)
      MLevel& L0 = L_[0];
      to_il(nrhs, x, ldx, L0.x2l);
      apply_il(nullptr, nullptr, true);
      from_il(nrhs, L0.y2l, y, ldy);
    } else {
      double* xi = kv_[0];
      double* yi = kv_[1];
      to_il(nrhs, x, ldx, xi);
      apply_il(xi, yi);
      from_il(nrhs, yi, y, ldy);
    }
    HIPCHECK(hipEventRecord(ev_, s_));
    HIPCHECK(hipStreamWaitEvent(caller, ev_, 0));
    return 0;
  }

  int krylov_cm(int method, int nrhs, const double* b, int64_t ldb, double* x, int64_t ldx, double eps, int itmax,
                int* iters, double* errs, int* launches, hipStream_t caller) override {
    HIPCHECK(hipSetDevice(device_));
    if (nrhs < 1 || nrhs > NB || ldb < n0_ || ldx < n0_ || itmax < 0 || (method != 0 && method != 1))
      return E_BAD_ARG;
    wait_caller(caller);
    double* kb = kv_[0];
    double* kx = kv_[1];
    double* r = kv_[2];
    to_il(nrhs, b, ldb, kb);
    to_il(nrhs, x, ldx, kx);
    // r = b - A x, norms, converged-at-start columns
    mv64_res(kx, kb, r);
    if (float* r32p = r32()) {  // FP32 input of the first application (later ones: the residual update)
      const int64_t cnt = (int64_t)n0_ * NB;
      convert64_kernel<float, double><<<nblk64(cnt), BS, 0, s_>>>(cnt, r, r32p);
    }
    mdot(kb, kb, sc_[C_BN2]);
    mdot(r, r, sc_[C_RR]);
    mstart_kernel<<<1, BS, 0, s_>>>(ncol_, eps, sc_[C_BN2], sc_[C_RR], sc_[C_BN], ic_[I_KIT], ic_[I_DONE],
                                    ic_[I_ITERS], sc_[C_ERR], ic_[I_ACTIVE]);
    HIPCHECK(hipGetLastError());
    int active = read_active();
    int nl = 0, done_it = 0;
    if (active > 0 && itmax > 0) {
      // first launch: as many iterations as the last solve with the same
      // method and tolerances took (one launch when the systems converge in
      // as many steps again); then launches of chunk() iterations until every
      // system has converged.  Iterations after a system's convergence leave
      // its vectors unchanged.
      const bool same = (pred_method_ == method && pred_eps_ == eps && pred_itmax_ == itmax);
      const int first = (same && pred_iters_ > 0) ? std::min(pred_iters_, itmax) : chunk();
      while (done_it < itmax) {
        const int ch = (nl == 0) ? first : chunk();
        HIPCHECK(hipGraphLaunch(graph_for(method, ch, eps, itmax), s_));
        ++nl;
        done_it += ch;
        if (read_active() == 0) break;
      }
    }
    from_il(nrhs, kx, x, ldx);
    std::vector<int> it(ncol_);
    std::vector<double> er(ncol_);
    HIPCHECK(hipMemcpyAsync(it.data(), ic_[I_ITERS], sizeof(int) * ncol_, hipMemcpyDeviceToHost, s_));
    HIPCHECK(hipMemcpyAsync(er.data(), sc_[C_ERR], sizeof(double) * ncol_, hipMemcpyDeviceToHost, s_));
    HIPCHECK(hipStreamSynchronize(s_));
    int itmx = 0;
    for (int j = 0; j < nrhs; ++j)
      for (int s = 0; s < G_; ++s) {
        iters[j * G_ + s] = it[s * NB + j];
        errs[j * G_ + s] = er[s * NB + j];
        itmx = std::max(itmx, it[s * NB + j]);
      }
    pred_method_ = method;
    pred_eps_ = eps;
    pred_itmax_ = itmax;
    pred_iters_ = itmx;
    if (launches) *launches = nl;
    HIPCHECK(hipEventRecord(ev_, s_));
    HIPCHECK(hipStreamWaitEvent(caller, ev_, 0));
    return 0;
  }

  int time_apply(int reps, double* ms) override {
    HIPCHECK(hipSetDevice(device_));
    double* xi = kv_[0];
    double* yi = kv_[1];
    fill64_kernel<double><<<nblk64((int64_t)n0_ * NB), BS, 0, s_>>>((int64_t)n0_ * NB, xi, 1.0);
    apply_il(xi, yi);
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, s_));
    for (int i = 0; i < reps; ++i) apply_il(xi, yi);
    HIPCHECK(hipEventRecord(e1, s_));
    HIPCHECK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHECK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / std::max(1, reps);
    return 0;
  }

 private:
  enum { C_BN2, C_BN, C_RR, C_RHO, C_RHO_OLD, C_SIGMA, C_GAMMA, C_DELTA, C_PR, C_ERR, NSC };
  enum { I_KIT, I_DONE, I_ITERS, I_ACTIVE, NIC };
  int device_;
  hipStream_t s_ = nullptr;
  hipEvent_t ev_ = nullptr;
  DevArena ar_;
  std::vector<MLevel> L_;
  std::vector<int64_t> off0_;
  int G_ = 0, n0_ = 0, nrb_ = 0, ncol_ = 0;
  DMat<double> akry_;
  int4* bt_ = nullptr;
  int* bstart_ = nullptr;
  double* partial_ = nullptr;
  double* sc_[NSC] = {};
  int* ic_[NIC] = {};
  double* kv_[6] = {};
  int* hstat_ = nullptr;
  // captured iteration sequences, one per (method, iterations, eps, itmax)
  struct Captured {
    int method, iters, itmax;
    double eps;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<Captured> graphs_;
  // iteration count of the last solve (the length of the next first launch)
  int pred_method_ = -1, pred_iters_ = 0, pred_itmax_ = 0;
  double pred_eps_ = 0.0;

  static int nblk64(int64_t n) { return (int)((n + BS - 1) / BS); }

  void wait_caller(hipStream_t caller) {
    HIPCHECK(hipEventRecord(ev_, caller));
    HIPCHECK(hipStreamWaitEvent(s_, ev_, 0));
  }

  // ------------------------------------------------------------- upload
  // Grouped systems: storage autotuned as for one system.  Block products:
  // CSR with the columns-per-lane split picked by timing.
  DMat<V> up(const HCsr& h) {
    if constexpr (NB == 1) {
      return upload_tuned<V>(h, ar_, s_, 0);
    } else {
      DMat<V> m = upload_mat<V>(h, ar_, s_, dia_eligible(h, dia_offsets(h)) ? 6 : 1, 0);
      if (m.fmt != FMT_DIA && m.fmt != FMT_CSR) m = upload_mat<V>(h, ar_, s_, 1, 1);
      tune_rpl<V>(m);
      return m;
    }
  }
  DMat<double> up64(const HCsr& h) {
    if constexpr (NB == 1) {
      return upload_tuned<double>(h, ar_, s_, 0);
    } else {
      DMat<double> m = upload_mat<double>(h, ar_, s_, dia_eligible(h, dia_offsets(h)) ? 6 : 1, 0);
      if (m.fmt != FMT_DIA && m.fmt != FMT_CSR) m = upload_mat<double>(h, ar_, s_, 1, 1);
      tune_rpl<double>(m);
      return m;
    }
  }
  // ND D (ND with its columns scaled by d, sharing ND's indices): the first
  // Jacobi sweep from zero, ND (D x), becomes the plain product (ND D) x (one
  // gather per entry, no separate D x pass; the rounding of each product is
  // re-associated, (a_ij d_j) x_j for a_ij (d_j x_j))
  DMat<V> scaled_cols(const DMat<V>& m, const V* d) {
    DMat<V> r = m;
    r.val = ar_.alloc<V>((size_t)std::max<int64_t>(1, m.stored));
    if (m.fmt == FMT_DIA) {
      dia_colscale_kernel<V><<<nblk(m.nr), BS, 0, s_>>>(m.nr, m.nc, m.w, m.off, m.val, d, r.val);
    } else if (m.fmt == FMT_SELD) {
      const int npos = (m.nr + 63) / 64 * 64;
      seld_colscale_kernel<V><<<nblk(npos), BS, 0, s_>>>(npos, m.nc, m.sptr, m.col0, m.dja, m.val, d, r.val);
    } else if (m.stored > 0) {
      colscale_kernel<V><<<(int)((m.stored + BS - 1) / BS), BS, 0, s_>>>(m.stored, m.ja, m.val, d, r.val);
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s_));
    return r;
  }
  V* upv(const std::vector<double>& v) {
    std::vector<V> t(v.begin(), v.end());
    V* d = ar_.upload(t.data(), t.size(), s_);
    HIPCHECK(hipStreamSynchronize(s_));
    return d;
  }

  template <typename T, bool DX, class E>
  void bdia_launch(const DMat<T>& m, const T* x, E e, const T* dx, int rpl, int nb, int xc) {
    switch (rpl) {
      case 4:
        if (m.w == 7) bdia_kernel<T, 4, 7, DX, E><<<nb, BS, 0, s_>>>(m.nr, m.nc, m.w, m.off, m.val, x, e, dx, xc);
        else bdia_kernel<T, 4, 0, DX, E><<<nb, BS, 0, s_>>>(m.nr, m.nc, m.w, m.off, m.val, x, e, dx, xc);
        break;
      case 2:
        if (m.w == 7) bdia_kernel<T, 2, 7, DX, E><<<nb, BS, 0, s_>>>(m.nr, m.nc, m.w, m.off, m.val, x, e, dx, xc);
        else bdia_kernel<T, 2, 0, DX, E><<<nb, BS, 0, s_>>>(m.nr, m.nc, m.w, m.off, m.val, x, e, dx, xc);
        break;
      default:
        if (m.w == 7) bdia_kernel<T, 1, 7, DX, E><<<nb, BS, 0, s_>>>(m.nr, m.nc, m.w, m.off, m.val, x, e, dx, xc);
        else bdia_kernel<T, 1, 0, DX, E><<<nb, BS, 0, s_>>>(m.nr, m.nc, m.w, m.off, m.val, x, e, dx, xc);
        break;
    }
    HIPCHECK(hipGetLastError());
  }

  template <typename T, class E>
  void bspmm(const DMat<T>& m, const T* x, E e, int rpl, const T* dx = nullptr) {
    if (m.nr == 0) return;
    const int lpr = MB / rpl, rpb = BS / lpr, nb = (m.nr + rpb - 1) / rpb;
    const int xc = bxchunk(m, rpb, nb);
    // the first Jacobi sweep from zero uses ND D (no d .* x gathers here)
    if (dx) throw DevError("bspmm: d .* x gathers are not used by block products");
    if (m.fmt == FMT_DIA) {
      if constexpr (!std::is_void<typename BNtForm<E>::type>::value) {
        if (m.nr >= bnt_min_rows()) {
          bdia_launch<T, false>(m, x, BNtForm<E>::make(e), dx, rpl, nb, xc);
          return;
        }
      }
      bdia_launch<T, false>(m, x, e, dx, rpl, nb, xc);
      return;
    }
    switch (rpl) {
      case 4:
        bspmm_kernel<T, 4, E><<<nb, BS, 0, s_>>>(m.nr, m.irp, m.ja, m.val, x, e, xc);
        break;
      case 2:
        bspmm_kernel<T, 2, E><<<nb, BS, 0, s_>>>(m.nr, m.irp, m.ja, m.val, x, e, xc);
        break;
      default:
        bspmm_kernel<T, 1, E><<<nb, BS, 0, s_>>>(m.nr, m.irp, m.ja, m.val, x, e, xc);
        break;
    }
    HIPCHECK(hipGetLastError());
  }

  // columns per lane and row-block order of a block product, by timing
  // Y = A X (large operators); ties within 2 % go to the earlier candidate.
  // MLD_BLOCK_SPLIT="R,X" fixes them for every operator (R in 1,2,4 columns
  // per lane, X in 0,1,2 as xmap), for tests of each variant.
  template <typename T>
  void tune_rpl(DMat<T>& m) {
    m.rpl = 1;
    m.xmap = 0;
    if (const char* f = std::getenv("MLD_BLOCK_SPLIT")) {
      int r = 0, xm = -1;
      if (std::sscanf(f, "%d,%d", &r, &xm) != 2 || !(r == 1 || r == 2 || r == 4) || xm < 0 || xm > 2)
        throw DevError(std::string("MLD_BLOCK_SPLIT must be R,X with R in 1,2,4 and X in 0,1,2: ") + f);
      m.rpl = r;
      m.xmap = xm;
      return;
    }
    if (m.nr < 4096 || !autotune_enabled()) return;
    // the temporaries are released also when a timing launch throws
    DevArena tmp;
    T* x = tmp.alloc<T>((size_t)std::max(1, m.nc) * MB);
    T* y = tmp.alloc<T>((size_t)m.nr * MB);
    fill64_kernel<T><<<nblk64((int64_t)std::max(1, m.nc) * MB), BS, 0, s_>>>((int64_t)std::max(1, m.nc) * MB, x,
                                                                              (T)1);
    struct Events {
      hipEvent_t e[2] = {nullptr, nullptr};
      ~Events() {
        for (hipEvent_t v : e)
          if (v) (void)hipEventDestroy(v);
      }
    } ev;
    HIPCHECK(hipEventCreate(&ev.e[0]));
    HIPCHECK(hipEventCreate(&ev.e[1]));
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    int best_r = 1, best_x = 0;
    float tbest = 1e30f;
    for (int xm : {1, 0, 2}) {
      for (int rpl : {4, 2, 1}) {
        m.xmap = xm;
        bspmm(m, x, BSet<T>{y}, rpl);
        float tmin = 1e30f;
        for (int trial = 0; trial < 3; ++trial) {
          HIPCHECK(hipEventRecord(e0, s_));
          for (int r = 0; r < 3; ++r) bspmm(m, x, BSet<T>{y}, rpl);
          HIPCHECK(hipEventRecord(e1, s_));
          HIPCHECK(hipEventSynchronize(e1));
          float t = 0.f;
          HIPCHECK(hipEventElapsedTime(&t, e0, e1));
          tmin = std::min(tmin, t);
        }
        if (std::getenv("MLD_VERBOSE"))
          std::fprintf(stderr, "[mld] block product %d x %d (%s): %d columns per lane, row-block order %d, %.1f us\n",
                       m.nr, m.nc, m.fmt == FMT_DIA ? "DIA" : "CSR", rpl, xm, tmin / 3 * 1e3);
        if (tmin < 0.98f * tbest) {
          tbest = tmin;
          best_r = rpl;
          best_x = xm;
        }
      }
    }
    HIPCHECK(hipStreamSynchronize(s_));  // the timing launches are done before tmp is freed
    m.rpl = best_r;
    m.xmap = best_x;
  }

  // ------------------------------------------------- operators on levels
  void op_set(const DMat<V>& m, const V* x, V* y) {
    if constexpr (NB == 1) spmv_launch(m, GatherX<V, V>{x}, EpiSet<V, V>{y}, s_);
    else bspmm(m, x, BSet<V>{y}, m.rpl);
  }
  void op_add(const DMat<V>& m, const V* x, V* y) {
    if constexpr (NB == 1) spmv_launch(m, GatherX<V, V>{x}, EpiAdd<V>{y}, s_);
    else bspmm(m, x, BAdd<V>{y}, m.rpl);
  }
  void op_res(const DMat<V>& m, const V* x, const V* b, V* r) {
    if constexpr (NB == 1) spmv_launch(m, GatherX<V, V>{x}, EpiRes<V, V>{b, r}, s_);
    else bspmm(m, x, BRes<V>{b, r}, m.rpl);
  }
  // y (+)= d .* (b - ND x); dx != null: x is gathered as dx_j * x_j (first
  // sweep from zero, DIA or ELL operators)
  template <bool ADD>
  void op_jac(const DMat<V>& m, const V* x, const V* d, const V* b, V* y, const V* dx = nullptr) {
    if constexpr (NB == 1) {
      if (dx) spmv_launch(m, GatherDX<V, V>{dx, x}, EpiJac<V, V, V, ADD>{d, b, y}, s_);
      else spmv_launch(m, GatherX<V, V>{x}, EpiJac<V, V, V, ADD>{d, b, y}, s_);
    } else {
      bspmm(m, x, BJac<V, ADD>{d, b, y}, m.rpl, dx);
    }
  }
  // last post-smoothing sweep of the fine level: FP64 output written directly
  void op_jac_out(const DMat<V>& m, const V* x, const V* d, const V* b, const V* yin, double* yout,
                  const V* dx) {
    if constexpr (NB == 1) {
      if (dx) spmv_launch(m, GatherDX<V, V>{dx, x}, EpiJacOut<V, V>{d, b, yin, yout, nullptr}, s_);
      else spmv_launch(m, GatherX<V, V>{x}, EpiJacOut<V, V>{d, b, yin, yout, nullptr}, s_);
    } else {
      bspmm(m, x, BJacOut<V>{d, b, yin, yout}, m.rpl, dx);
    }
  }
  template <bool ADD>
  void vmul(int n, const V* d, const V* x, V* y) {
    if (n == 0) return;
    if constexpr (NB == 1) vmul_kernel<V, V, ADD><<<nblk(n), BS, 0, s_>>>(n, d, x, y);
    else bvmul_kernel<V, ADD><<<nblk64((int64_t)n * NB), BS, 0, s_>>>(n, d, x, y);
    HIPCHECK(hipGetLastError());
  }
  void fill0(V* y, int64_t cnt) {
    if (cnt == 0) return;
    fill64_kernel<V><<<nblk64(cnt), BS, 0, s_>>>(cnt, y, (V)0);
    HIPCHECK(hipGetLastError());
  }

  // Jacobi-type smoother of the continuing rows (mld_d_jac_smoother_apply_vect,
  // init 'Z', alpha = 1, beta in {0, 1}); the first sweep from zero is folded
  // into the gather of the second (d .* x formed first).
  void smooth(MLevel& D, bool post, const V* x, int beta, V* y, int sweeps) {
    const bool two = post && D.has2;
    const DMat<V>& nd = two ? D.nd2 : D.nd;
    const DMat<V>& ndd = two ? D.ndd2 : D.ndd;
    const V* dv = two ? D.dv2 : D.dv;
    const int64_t nnz = two ? D.nnz_nd2 : D.nnz_nd;
    const int n = D.ni;
    if (sweeps <= 1 || nnz == 0) {
      if (beta) vmul<true>(n, dv, x, y);
      else vmul<false>(n, dv, x, y);
      return;
    }
    // the cycle's FP64 output straight from the fine level's last sweep
    const bool fo = out64_ != nullptr && beta == 1 && &D == &L_[0] && y == D.y2l;
    // first sweep from zero: the product with ND D on DIA, a gather of
    // d_j x_j where it is cheap (ELL for one column), else d .* x formed first
    const DMat<V>& nd1 = ndd.val ? ndd : nd;
    const bool dxg = !ndd.val && NB == 1 && nd.fmt == FMT_ELL;
    const V* xin = x;
    const V* dxv = dxg ? dv : nullptr;
    if (!ndd.val && !dxg) {
      vmul<false>(n, dv, x, D.t2);
      xin = D.t2;
    }
    if (sweeps == 2) {
      if (fo) {
        op_jac_out(nd1, xin, dv, x, y, out64_, dxv);
        out_done_ = true;
      } else if (beta) {
        op_jac<true>(nd1, xin, dv, x, y, dxv);
      } else {
        op_jac<false>(nd1, xin, dv, x, y, dxv);
      }
      return;
    }
    V* ta = D.tx;
    V* tb = D.t1;
    op_jac<false>(nd1, xin, dv, x, ta, dxv);
    for (int i = 2; i <= sweeps - 2; ++i) {
      op_jac<false>(nd, ta, dv, x, tb);
      std::swap(ta, tb);
    }
    if (fo) {
      op_jac_out(nd, ta, dv, x, y, out64_, nullptr);
      out_done_ = true;
    } else if (beta) {
      op_jac<true>(nd, ta, dv, x, y);
    } else {
      op_jac<false>(nd, ta, dv, x, y);
    }
  }

  // mld_d_inner_mult for every system of the union at once
  void inner(int l) {
    MLevel& D = L_[l];
    if (D.ni > 0) {
      MLevel& C = L_[l + 1];
      const bool pre = D.sweeps_pre > 0, post = D.sweeps_post > 0;
      if (pre) {
        smooth(D, false, D.x2l, 0, D.y2l, D.sweeps_pre);
        op_res(D.a, D.y2l, D.x2l, D.ty);
        op_set(C.r, D.ty, C.x2l);
      } else {
        fill0(D.y2l, (int64_t)D.ni * NB);
        op_set(C.r, D.x2l, C.x2l);
      }
      inner(l + 1);
      op_add(C.p, C.y2l, D.y2l);
      if (D.ml_cycle == WCYCLE) {
        op_res(D.a, D.y2l, D.x2l, D.ty);
        op_set(C.r, D.ty, C.x2l);
        inner(l + 1);
        op_add(C.p, C.y2l, D.y2l);
      }
      if (post) {
        op_res(D.a, D.y2l, D.x2l, D.ty);
        smooth(D, true, D.ty, 1, D.y2l, D.sweeps_post);
      }
    }
    if (D.n > D.ni) op_set(D.mc, D.x2l + (size_t)D.ni * NB, D.y2l + (size_t)D.ni * NB);
  }

  // y = M^-1 x, row-interleaved FP64 device vectors of n0 x NB.  FP32 form:
  // preconverted = the fine level's input buffer already holds x in FP32;
  // the last post-smoothing sweep of the fine level writes y in FP64, or,
  // with y = null, leaves the result in the fine level's FP32 output buffer.
  void apply_il(const double* x, double* y, bool preconverted = false) {
    MLevel& L0 = L_[0];
    const int64_t cnt = (int64_t)n0_ * NB;
    if constexpr (std::is_same<V, double>::value) {
      V* xs = L0.x2l;
      V* ys = L0.y2l;
      L0.x2l = const_cast<double*>(x);
      L0.y2l = y;
      inner(0);
      L0.x2l = xs;
      L0.y2l = ys;
    } else {
      if (!preconverted) convert64_kernel<V, double><<<nblk64(cnt), BS, 0, s_>>>(cnt, x, L0.x2l);
      out64_ = y;
      out_done_ = false;
      inner(0);
      out64_ = nullptr;
      if (y && !out_done_) convert64_kernel<double, V><<<nblk64(cnt), BS, 0, s_>>>(cnt, L0.y2l, y);
    }
    HIPCHECK(hipGetLastError());
  }
  double* out64_ = nullptr;
  bool out_done_ = false;
  // FP32 form: where the residual update leaves r in FP32 (the fine level's
  // input buffer); FP64 form: nowhere (the cycle reads r itself)
  float* r32() {
    if constexpr (std::is_same<V, double>::value) return nullptr;
    else return L_[0].x2l;
  }

  // --------------------------------------------------------- layout
  template <typename T>
  void to_il(int nrhs, const double* src, int64_t ld, T* dst) {
    if constexpr (NB == 1) {
      static_assert(std::is_same<T, double>::value, "one-column copies are FP64");
      HIPCHECK(hipMemcpyAsync(dst, src, sizeof(double) * n0_, hipMemcpyDeviceToDevice, s_));
    } else {
      col2il_kernel<T><<<(n0_ + 63) / 64, BS, 0, s_>>>(n0_, nrhs, src, ld, dst);
      HIPCHECK(hipGetLastError());
    }
  }
  template <typename T>
  void from_il(int nrhs, const T* src, double* dst, int64_t ld) {
    if constexpr (NB == 1) {
      static_assert(std::is_same<T, double>::value, "one-column copies are FP64");
      HIPCHECK(hipMemcpyAsync(dst, src, sizeof(double) * n0_, hipMemcpyDeviceToDevice, s_));
    } else {
      il2col_kernel<T><<<(n0_ + 63) / 64, BS, 0, s_>>>(n0_, nrhs, src, dst, ld);
      HIPCHECK(hipGetLastError());
    }
  }

  // --------------------------------------------------------- Krylov
  void mv64_res(const double* x, const double* b, double* r) {
    if constexpr (NB == 1) spmv_launch(akry_, GatherX<double, double>{x}, EpiRes<double, double>{b, r}, s_);
    else bspmm(akry_, x, BRes<double>{b, r}, akry_.rpl);
  }
  void mv64_set(const double* x, double* y) {
    if constexpr (NB == 1) spmv_launch(akry_, GatherX<double, double>{x}, EpiSet<double, double>{y}, s_);
    else bspmm(akry_, x, BSet<double>{y}, akry_.rpl);
  }
  void mfin(double* out) {
    mfin_kernel<NB><<<ncol_, BS, 0, s_>>>(bstart_, partial_, out);
    HIPCHECK(hipGetLastError());
  }
  void mdot(const double* x, const double* y, double* out) {
    mdot_kernel<NB><<<nrb_, BS, 0, s_>>>(bt_, x, y, partial_);
    mfin(out);
  }
  int read_active() {
    HIPCHECK(hipMemcpyAsync(hstat_, ic_[I_ACTIVE], sizeof(int), hipMemcpyDeviceToHost, s_));
    HIPCHECK(hipStreamSynchronize(s_));
    return hstat_[0];
  }
  static int chunk() {
    const char* e = std::getenv("MLD_GRAPH_CHUNK");
    const int c = e ? std::atoi(e) : 4;
    return std::max(1, std::min(64, c));
  }

  // One iteration of every column (no host synchronisation).
  void iteration(int method, double eps, int itmax) {
    double* r = kv_[2];
    double* z = kv_[3];
    double* p = kv_[4];
    double* q = kv_[5];
    double* x = kv_[1];
    apply_il(r, z, true);  // r's FP32 copy was written by the previous residual update
    if (method == 0) {  // CG
      mdot(r, z, sc_[C_RHO]);
      mp_cg_kernel<NB><<<nrb_, BS, 0, s_>>>(bt_, z, p, sc_[C_RHO], sc_[C_RHO_OLD], ic_[I_KIT], ic_[I_DONE]);
      mv64_set(p, q);
      mdot(p, q, sc_[C_SIGMA]);
      mxr_kernel<NB><<<nrb_, BS, 0, s_>>>(bt_, x, r, p, q, sc_[C_RHO], sc_[C_SIGMA], ic_[I_DONE], partial_, r32());
      mfin(sc_[C_RR]);
      mcheck_kernel<<<1, BS, 0, s_>>>(ncol_, eps, itmax, sc_[C_RR], sc_[C_BN], sc_[C_RHO], sc_[C_SIGMA], sc_[C_RHO],
                                      sc_[C_RHO_OLD], ic_[I_KIT], ic_[I_DONE], ic_[I_ITERS], sc_[C_ERR],
                                      ic_[I_ACTIVE]);
    } else {  // FCG(1)
      mdot(q, z, sc_[C_GAMMA]);  // q of the previous iteration (unused in the first)
      mp_fcg_kernel<NB><<<nrb_, BS, 0, s_>>>(bt_, z, p, r, sc_[C_GAMMA], sc_[C_DELTA], ic_[I_KIT], ic_[I_DONE],
                                             partial_);
      mfin(sc_[C_PR]);
      mv64_set(p, q);
      mdot(p, q, sc_[C_DELTA]);
      mxr_kernel<NB><<<nrb_, BS, 0, s_>>>(bt_, x, r, p, q, sc_[C_PR], sc_[C_DELTA], ic_[I_DONE], partial_, r32());
      mfin(sc_[C_RR]);
      mcheck_kernel<<<1, BS, 0, s_>>>(ncol_, eps, itmax, sc_[C_RR], sc_[C_BN], sc_[C_DELTA], nullptr, nullptr,
                                      nullptr, ic_[I_KIT], ic_[I_DONE], ic_[I_ITERS], sc_[C_ERR], ic_[I_ACTIVE]);
    }
    HIPCHECK(hipGetLastError());
  }

  // The sequence of `ch` iterations captured once as a graph; its vectors
  // are the engine's own buffers, so the graph is reused by later solves with
  // the same method and tolerances (a few lengths are kept).
  hipGraphExec_t graph_for(int method, int ch, double eps, int itmax) {
    for (auto& c : graphs_)
      if (c.method == method && c.iters == ch && c.eps == eps && c.itmax == itmax) return c.exec;
    if (graphs_.size() >= 4) {  // drop the oldest
      (void)hipGraphExecDestroy(graphs_.front().exec);
      (void)hipGraphDestroy(graphs_.front().graph);
      graphs_.erase(graphs_.begin());
    }
    hipGraph_t g = nullptr;
    HIPCHECK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    try {
      for (int i = 0; i < ch; ++i) iteration(method, eps, itmax);
    } catch (...) {
      hipGraph_t gg = nullptr;
      (void)hipStreamEndCapture(s_, &gg);
      if (gg) (void)hipGraphDestroy(gg);
      throw;
    }
    HIPCHECK(hipStreamEndCapture(s_, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(g);
      HIPCHECK(e);
    }
    graphs_.push_back(Captured{method, ch, itmax, eps, g, ex});
    return ex;
  }
};

template <typename V, int NB>
static MultiEngine* make_one(const std::vector<const Prec*>& systems, int device, int& rc, std::string& err) {
  std::unique_ptr<MultiT<V, NB>> e(new MultiT<V, NB>(device));
  rc = e->build(systems, err);
  return rc ? nullptr : e.release();
}

MultiEngine* make_multi(const std::vector<const Prec*>& systems, int value_kind, int nb, int device, int& rc,
                        std::string& err) {
  rc = 0;
  try {
    if (nb == 1) return value_kind == 1 ? make_one<float, 1>(systems, device, rc, err)
                                        : make_one<double, 1>(systems, device, rc, err);
    if (nb == MB) return value_kind == 1 ? make_one<float, MB>(systems, device, rc, err)
                                         : make_one<double, MB>(systems, device, rc, err);
    rc = E_BAD_ARG;
    err = "block width must be 1 or 32";
    return nullptr;
  } catch (const std::exception& ex) {
    err = ex.what();
    rc = E_DEVICE;
    return nullptr;
  }
}

}  // namespace mld
