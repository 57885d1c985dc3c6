// Device-side building blocks shared by the single-system engine and the
// batched engines: device matrix storage (CSR / ELL / SELL-C-sigma), per-level
// smoother data, the allocation arena, the upload of host CSR matrices, the
// generic sparse launch and the storage autotuner.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "mld_device.hpp"
#include "mld_host.hpp"
#include "mld_kernels.hpp"

namespace mld {

using namespace mld::k;

struct DevError : std::runtime_error {
  explicit DevError(const std::string& s) : std::runtime_error(s) {}
};

#define HIPCHECK(x)                                                                                 \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess)                                                                           \
      throw DevError(std::string(#x) + ": " + hipGetErrorString(e_) + " at " + __FILE__ + ":" +     \
                     std::to_string(__LINE__));                                                     \
  } while (0)

enum { FMT_CSR = 0, FMT_ELL = 1, FMT_SELL = 2, FMT_DIA = 3, FMT_SELD = 4 };

template <typename V>
struct DMat {
  int nr = 0, nc = 0;
  int64_t nnz = 0;
  int64_t stored = 0;   // entries in val (padding included; DIA: nr * w)
  int fmt = FMT_CSR, lanes = 1, w = 0, ld = 0;
  int iters = 0;        // CSR: entries per lane loaded up front (0, 2, 4, 8)
  int rpl = 1;          // block products (32 columns): columns per lane (1, 2, 4)
  DiaOff off{};         // DIA: diagonal offsets (w diagonals)
  int xmap = 1;         // row-block order: 0 identity, 1 / 3 / 4 chunks of 8192 / 32768 / 131072 rows,
                        // 2 one range per XCD
  int* irp = nullptr;   // CSR row pointers
  int* sptr = nullptr;  // SELL slice offsets (nslices + 1)
  int* perm = nullptr;  // SELL position -> row
  // SELD: SELL whose column indices are 16-bit steps: entry k of a position
  // is column col0 + k + (dja_0 + ... + dja_k) (padding: steps of 0)
  int* col0 = nullptr;
  uint16_t* dja = nullptr;
  int* ja = nullptr;
  V* val = nullptr;
};

template <typename V>
struct DSmoother {
  int smoother = JAC, solver = DIAG_SCALE;
  int64_t nnz_nd = 0;
  DMat<V> nd, ndt;     // ND and its transpose (uploaded for 'T' applies)
  DMat<V> ndd;         // ND with its columns scaled by d (ND D): first sweep from zero
  V* dv = nullptr;
  // ILU(0): factors and level schedules; for 'T': transposed factors
  DMat<V> l, u, lt, ut;
  int* lrows = nullptr; int* urows = nullptr; int* lptr_d = nullptr; int* uptr_d = nullptr;
  std::vector<int> lptr, uptr;
  int* ltrows = nullptr; int* utrows = nullptr; int* ltptr_d = nullptr; int* utptr_d = nullptr;
  std::vector<int> ltptr, utptr;
  bool has_t = false;
  int tri_lanes = 1;  // lanes per row of the ILU triangular solves (1, 4, 16, 32)
  // sync-free substitutions: chunks of each schedule (positions in its row
  // list, at most BS/tri_lanes rows, never across level sets) and the chunk
  // ticket
  int *lch = nullptr, *uch = nullptr, *ltch = nullptr, *utch = nullptr;
  int nlch = 0, nuch = 0, nltch = 0, nutch = 0;
  int* sf_ticket = nullptr;
  // Triangular factors (ILU(0), or the parts of Gauss-Seidel) of a 7-point
  // stencil on a gnx x gny x gnz grid in natural order: their values by
  // stencil position, d and two work vectors in the tiles' step order for
  // trsv_grid_kernel, gsz entries each (null when the factors have another
  // pattern)
  int gnx = 0, gny = 0, gnz = 0;
  V *lgs = nullptr, *ugs = nullptr, *dgs = nullptr, *tgs = nullptr, *xgs = nullptr;
  int64_t gsz = 0;
  // dense LU
  double* lu = nullptr; int* piv = nullptr;
  double* dinv = nullptr;  // DENSE_LU: explicit inverse (row-major)
  int ndense = 0;
};

template <typename V>
struct DLevel {
  int n = 0;
  DMat<V> a, at, p, r;  // at: A^T for 'T' applies
  DSmoother<V> sm, sm2;
  bool has_sm2 = false;
  int sweeps_pre = 1, sweeps_post = 1, ml_cycle = VCYCLE;
  V *x2l = nullptr, *y2l = nullptr, *tx = nullptr, *ty = nullptr, *t1 = nullptr, *t2 = nullptr;
  // K-cycle inner iteration at this level: w, v, v1, d0, d1 (the iterate
  // accumulates in y2l)
  V *kw = nullptr, *kv = nullptr, *kv1 = nullptr, *kd0 = nullptr, *kd1 = nullptr;
  double* wlu = nullptr;
};

// ---------------------------------------------------------------------------
// Allocation bookkeeping: every device allocation of a handle is freed with it.
// ---------------------------------------------------------------------------
static inline void h2d_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
  host_to_device(dst, src, bytes, s);
}

struct DevArena {
  std::vector<void*> ptrs;
  size_t bytes = 0;
  template <typename T>
  T* alloc(size_t n) {
    if (n == 0) n = 1;
    void* p = nullptr;
    HIPCHECK(hipMalloc(&p, n * sizeof(T)));
    ptrs.push_back(p);
    bytes += n * sizeof(T);
    return (T*)p;
  }
  template <typename T>
  T* upload(const T* h, size_t n, hipStream_t s) {
    T* d = alloc<T>(n);
    if (n) h2d_copy(d, h, n * sizeof(T), s);
    return d;
  }
  ~DevArena() {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

static inline int nblk(int64_t n) { return (int)((n + BS - 1) / BS); }

// Rows per thread of a DIA product with dot outputs: two rows per thread
// halve the blocks and their reduction epilogues (Krylov product at 256^3:
// 265 -> 239 us FP64); plain products keep one row per thread (faster for
// them: residual 239 vs 247 us).
template <int K>
constexpr int dia_rpt() { return K > 0 ? 2 : 1; }

// ---------------------------------------------------------------------------
// Storage choice (force: 0 automatic, 1 CSR, 2 ELL, 3 SELL).  Automatic:
// ELL when the widest row is short and padding small (the 7-point fine
// level: 0.3 % padding); SELL for other matrices with enough rows to fill
// slices; CSR with a power-of-two number of lanes per row (matched to the
// mean row length) for small matrices.
// ---------------------------------------------------------------------------
struct FmtChoice {
  int fmt, lanes, w;
  std::vector<int> offs;  // DIA: the diagonal offsets (ascending)
};

// Diagonal offsets of h when it has at most 8 distinct ones (ascending), else
// empty.
// Columns ascend within a row, so the offsets of a row ascend too: one merge
// walk per row against the sorted set found so far.
static std::vector<int> dia_offsets(const HCsr& h) { return host_dia_offsets(h); }

// DIA pays when the padded diagonals hold at most 10 % more entries than A.
static bool dia_eligible(const HCsr& h, const std::vector<int>& offs) {
  return !offs.empty() && h.nr == h.nc && (double)h.nr * offs.size() <= 1.10 * (double)h.nnz();
}

// SELD (force 8: sigma 256, force 9: sigma 4096, force 10: sigma 128): SELL with each row's
// column indices stored as 16-bit steps between consecutive stored columns
// (step - 1, so 1 .. 65536), and the first column per position.  Entries
// become 10 bytes instead of 12 (FP64), 6 instead of 8 (FP32); the products
// keep the stored order, so the sums are the SELL sums.  Eligible when every
// row's columns increase by at most 65536 from one entry to the next.
static bool host_seld_ok(const HCsr& h) {
  bool ok = true;
#pragma omp parallel for schedule(static) reduction(&& : ok)
  for (int i = 0; i < h.nr; ++i)
    for (int64_t k = h.irp[i] + 1; k < h.irp[i + 1]; ++k) {
      const int64_t d = (int64_t)h.ja[k] - h.ja[k - 1];
      if (d < 1 || d > 65536) ok = false;
    }
  return ok;
}

static FmtChoice choose_format(const HCsr& h, int force) {
  const int maxrow = host_max_row(h);
  const double nnz = (double)h.nnz();
  const double avg = h.nr ? nnz / h.nr : 0.0;
  FmtChoice c{FMT_CSR, 1, maxrow};
  int lanes = 1;
  while (lanes < 64 && lanes * 4 <= avg) lanes *= 2;
  c.lanes = lanes;
  if (force == 1 || maxrow == 0) return c;
  if (force == 0 || force == 6) {
    std::vector<int> offs = dia_offsets(h);
    if (dia_eligible(h, offs)) {
      c.fmt = FMT_DIA; c.w = (int)offs.size(); c.offs = std::move(offs);
      return c;
    }
  }
  if (force == 2 && maxrow <= 64) { c.fmt = FMT_ELL; return c; }
  if ((force >= 3 && force <= 5) || force == 7) { c.fmt = FMT_SELL; return c; }
  if (force >= 8 && force <= 10) {  // SELL where the steps do not fit 16 bits
    c.fmt = host_seld_ok(h) ? FMT_SELD : FMT_SELL;
    return c;
  }
  if (maxrow <= 16 && (double)h.nr * maxrow <= 1.10 * nnz) { c.fmt = FMT_ELL; return c; }
  if (h.nr >= 4096) { c.fmt = FMT_SELL; return c; }
  return c;
}

// Rows are sorted by decreasing length inside windows of sigma rows (stable,
// so equal-length rows keep their order); slices of 64 positions.  Large
// windows cut padding; small windows keep consecutive rows (and so the
// columns they gather) together in a wavefront.  force 3: sigma 4096,
// force 4: sigma 256, force 5: no sorting, force 7: sigma 128 (the level-1
// Galerkin operator at 256^3: padding 1.16x unsorted, 1.09x at 128 with the
// same gathered cache lines per entry, 1.05x at 256 with 16 % more).
static int sell_sigma(int force) {
  return (force == 4 || force == 8) ? 256 : (force == 5 ? 1 : ((force == 7 || force == 10) ? 128 : 4096));
}

// An operator uploaded once as CSR (int32 row pointers, FP64 values); every
// candidate layout is then built from it on the device.
struct DevSrc {
  int nr = 0, nc = 0;
  int64_t nnz = 0;
  int maxrow = 0;
  int* irp = nullptr;
  int* ja = nullptr;
  double* val = nullptr;
};

static DevSrc make_src(const HCsr& h, DevArena& tmp, hipStream_t s) {
  DevSrc d;
  d.nr = h.nr; d.nc = h.nc; d.nnz = h.nnz();
  if (d.nnz >= ((int64_t)1 << 31)) throw DevError("operator exceeds 2^31 entries");
  d.maxrow = host_max_row(h);
  int64_t* irp64 = tmp.alloc<int64_t>((size_t)h.nr + 1);
  if (!h.irp.empty()) h2d_copy(irp64, h.irp.data(), ((size_t)h.nr + 1) * sizeof(int64_t), s);
  else HIPCHECK(hipMemsetAsync(irp64, 0, ((size_t)h.nr + 1) * sizeof(int64_t), s));
  d.irp = tmp.alloc<int>((size_t)h.nr + 1);
  irp_to32_kernel<<<nblk(h.nr + 1), BS, 0, s>>>(h.nr + 1, irp64, d.irp);
  d.ja = tmp.upload(h.ja.data(), h.ja.size(), s);
  d.val = tmp.upload(h.val.data(), h.val.size(), s);
  HIPCHECK(hipGetLastError());
  return d;
}

// Host SELL layouts (position -> row, slice offsets) by sort window, kept
// while the autotuner builds its candidates (SELL and SELD of one window
// share a layout).
struct SellLayouts {
  std::vector<int> sigma;
  std::vector<std::vector<int>> perm, sptr;
};

// Layout c of the operator h (uploaded as src) in arena ar.
template <typename V>
static DMat<V> build_mat(const HCsr& h, const DevSrc& src, DevArena& ar, hipStream_t s, FmtChoice c, int force,
                         SellLayouts* cache = nullptr) {
  DMat<V> m;
  m.nr = h.nr; m.nc = h.nc; m.nnz = h.nnz();
  m.fmt = c.fmt; m.lanes = c.lanes; m.w = c.w;
  {
    const int per_lane = (c.w + c.lanes - 1) / std::max(1, c.lanes);
    m.iters = per_lane <= 2 ? 2 : (per_lane <= 4 ? 4 : 8);
  }
  if (c.fmt == FMT_DIA) {
    const std::vector<int> offs = c.offs.empty() ? dia_offsets(h) : c.offs;
    m.w = (int)offs.size();
    for (int d = 0; d < 8; ++d) m.off.o[d] = d < m.w ? offs[d] : 0;
    m.stored = (int64_t)h.nr * m.w;
    m.val = ar.alloc<V>((size_t)m.stored);
    dia_fill_kernel<V><<<nblk(h.nr), BS, 0, s>>>(h.nr, m.w, m.off, src.irp, src.ja, src.val, m.val);
  } else if (c.fmt == FMT_ELL) {
    m.ld = h.nr;
    m.stored = (int64_t)m.ld * m.w;
    m.ja = ar.alloc<int>((size_t)m.ld * m.w);
    m.val = ar.alloc<V>((size_t)m.ld * m.w);
    ell_fill_kernel<V><<<nblk(h.nr), BS, 0, s>>>(h.nr, h.nc, m.w, src.irp, src.ja, src.val, m.ja, m.val);
  } else if (c.fmt == FMT_SELL || c.fmt == FMT_SELD) {
    const int nr = h.nr;
    const int nsl = (nr + SELL_C - 1) / SELL_C;
    std::vector<int> perm_own, sptr_own;
    const std::vector<int>* pp = &perm_own;
    const std::vector<int>* sp = &sptr_own;
    const int sigma = sell_sigma(force);
    size_t ci = cache ? cache->sigma.size() : 0;
    if (cache)
      for (size_t k = 0; k < cache->sigma.size(); ++k)
        if (cache->sigma[k] == sigma) ci = k;
    if (cache && ci < cache->sigma.size()) {
      pp = &cache->perm[ci];
      sp = &cache->sptr[ci];
    } else {
      if (!host_sell_layout(h, sigma, SELL_C, src.maxrow, perm_own, sptr_own))
        throw DevError("SELL storage exceeds 2^31 entries");
      if (cache) {
        cache->sigma.push_back(sigma);
        cache->perm.push_back(perm_own);
        cache->sptr.push_back(sptr_own);
      }
    }
    const std::vector<int>& perm = *pp;
    const std::vector<int>& sptr = *sp;
    const size_t tot = (size_t)sptr[nsl];
    m.stored = (int64_t)tot;
    m.sptr = ar.upload(sptr.data(), sptr.size(), s);
    m.perm = ar.upload(perm.data(), perm.size(), s);
    DevArena jtmp;  // SELD: the 32-bit indices only until they are converted
    int* ja = (c.fmt == FMT_SELL) ? ar.alloc<int>(tot) : jtmp.alloc<int>(std::max<size_t>(1, tot));
    m.val = ar.alloc<V>(tot);
    sell_fill_kernel<V><<<nblk((int64_t)nsl * SELL_C), BS, 0, s>>>(nr, h.nc, nsl * SELL_C, m.perm, m.sptr, src.irp,
                                                                  src.ja, src.val, ja, m.val);
    if (c.fmt == FMT_SELL) {
      m.ja = ja;
    } else {
      m.col0 = ar.alloc<int>((size_t)nsl * SELL_C);
      m.dja = ar.alloc<uint16_t>(std::max<size_t>(1, tot));
      seld_convert_kernel<<<nblk((int64_t)nsl * SELL_C), BS, 0, s>>>(nr, nsl * SELL_C, m.perm, m.sptr, src.irp, ja,
                                                                   m.col0, m.dja);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipStreamSynchronize(s));  // before jtmp is released
    }
  } else {
    m.stored = (int64_t)h.val.size();
    m.irp = ar.alloc<int>((size_t)h.nr + 1);
    m.ja = ar.alloc<int>(h.ja.size());
    m.val = ar.alloc<V>(h.val.size());
    HIPCHECK(hipMemcpyAsync(m.irp, src.irp, ((size_t)h.nr + 1) * sizeof(int), hipMemcpyDeviceToDevice, s));
    if (!h.ja.empty())
      HIPCHECK(hipMemcpyAsync(m.ja, src.ja, h.ja.size() * sizeof(int), hipMemcpyDeviceToDevice, s));
    if (!h.val.empty())
      cvt_val_kernel<V><<<(int)((h.val.size() + BS - 1) / BS), BS, 0, s>>>((int64_t)h.val.size(), src.val, m.val);
  }
  HIPCHECK(hipGetLastError());
  // the host's arrays (sptr, perm) must outlive the asynchronous copies
  HIPCHECK(hipStreamSynchronize(s));
  return m;
}

template <typename V>
static DMat<V> upload_mat(const HCsr& h, DevArena& ar, hipStream_t s, int force, int lanes_override = 0) {
  FmtChoice c = choose_format(h, force);
  if (lanes_override > 0) { c.fmt = FMT_CSR; c.lanes = lanes_override; }
  DevArena tmp;
  const DevSrc src = make_src(h, tmp, s);
  DMat<V> m = build_mat<V>(h, src, ar, s, c, force);
  return m;  // tmp is released after build_mat's synchronisation
}

static HCsr host_transpose(const HCsr& a) {
  HCsr t;
  t.nr = a.nc; t.nc = a.nr;
  t.irp.assign((size_t)t.nr + 1, 0);
  for (int64_t k = 0; k < a.nnz(); ++k) t.irp[a.ja[k] + 1]++;
  for (int i = 0; i < t.nr; ++i) t.irp[i + 1] += t.irp[i];
  t.ja.resize(a.nnz()); t.val.resize(a.nnz());
  std::vector<int64_t> pos(t.irp.begin(), t.irp.end() - 1);
  for (int i = 0; i < a.nr; ++i)
    for (int64_t k = a.irp[i]; k < a.irp[i + 1]; ++k) {
      int64_t q = pos[a.ja[k]]++;
      t.ja[q] = i; t.val[q] = a.val[k];
    }
  return t;
}

// Chunks of a level schedule (ptr: level -> first position) for the
// sync-free substitution: at most `per` consecutive positions of one level
// set each; returns the chunk start positions plus the end.
static std::vector<int> sched_chunks(const std::vector<int>& ptr, int per) {
  std::vector<int> c;
  for (size_t lv = 0; lv + 1 < ptr.size(); ++lv)
    for (int q = ptr[lv]; q < ptr[lv + 1]; q += per) c.push_back(q);
  c.push_back(ptr.empty() ? 0 : ptr.back());
  return c;
}

// Blocks of the grid of a sparse launch (row- or lane-per-thread).
template <typename V>
static int spmv_blocks(const DMat<V>& m) {
  if (m.fmt == FMT_CSR) return nblk((int64_t)m.nr * m.lanes);
  return nblk(m.nr);  // DIA: an upper bound (products with dots use fewer blocks)
}

// Chunk (in blocks) of the row-block order for a launch of nb blocks of
// rows_per_block rows (see xcd_block).
static inline int xcd_chunk(int xmap, int nb, int rows_per_block) {
  if (xmap == 1) return std::max(1, 8192 / rows_per_block);
  if (xmap == 2) return std::max(1, nb / 8);
  if (xmap == 3) return std::max(1, 32768 / rows_per_block);
  if (xmap == 4) return std::max(1, 131072 / rows_per_block);
  return 0;
}

template <int LANES, int K, typename C, typename V, class G, class E>
static void csr_launch_l(const DMat<V>& m, int nb, G g, E e, hipStream_t s, Red red) {
  const int xc = xcd_chunk(m.xmap, nb, BS / LANES);
  switch (m.iters) {
    case 2: csr_kernel<C, LANES, 2, K><<<nb, BS, 0, s>>>(m.nr, m.irp, m.ja, m.val, g, e, red, xc); break;
    case 4: csr_kernel<C, LANES, 4, K><<<nb, BS, 0, s>>>(m.nr, m.irp, m.ja, m.val, g, e, red, xc); break;
    case 8: csr_kernel<C, LANES, 8, K><<<nb, BS, 0, s>>>(m.nr, m.irp, m.ja, m.val, g, e, red, xc); break;
    default: csr_kernel<C, LANES, 0, K><<<nb, BS, 0, s>>>(m.nr, m.irp, m.ja, m.val, g, e, red, xc); break;
  }
}

// ---------------------------------------------------------------------------
// Generic sparse launch: dispatch on format / width / lanes.  K > 0: the
// epilogue's dot contributions are reduced into red.result[0..K).  CT: the
// arithmetic type when it differs from the stored value type (FP32-stored
// values that are exact in FP32, summed in FP64).
// ---------------------------------------------------------------------------
// DIA levels from this many rows store their outputs non-temporally
constexpr int NT_MIN_ROWS = 1 << 22;
// (MLD_NT_MIN_ROWS overrides it, for measurements)
static int nt_min_rows() {
  static const int v = [] {
    const char* e = std::getenv("MLD_NT_MIN_ROWS");
    return e ? std::atoi(e) : NT_MIN_ROWS;
  }();
  return v;
}

// epilogues with a two-row form (pair / aligned2)
template <class E, class = void>
struct has_pair : std::false_type {};
template <class E>
struct has_pair<E, std::void_t<decltype(std::declval<const E&>().aligned2())>> : std::true_type {};

// DIA launch (with its final reduction when K > 0): two adjacent rows per
// thread when the epilogue has the two-row form and the arrays allow it.
template <int K, typename C, typename V, class G, class E>
static void dia_launch(const DMat<V>& m, G g, E e, hipStream_t s, Red red) {
  // (FP64 values only: with FP32 values the two-row form was slower, the
  // fine-level FP32 residual 94 -> 127 us at 256^3, the FP32-stored Krylov
  // product summed in FP64 144 -> 150 us)
  if constexpr (has_pair<E>::value && std::is_same<V, double>::value) {
    if ((m.nr & 1) == 0 && al2(m.val, sizeof(V)) && e.aligned2()) {
      const int nbp = (int)(((int64_t)m.nr / 2 + BS - 1) / BS);
      const int xc = xcd_chunk(m.xmap, nbp, 2 * BS);
      if (m.w == 7)
        dia_kernel<C, 7, K, 1, G, E, V, true><<<nbp, BS, 0, s>>>(m.nr, m.nc, m.w, m.off, m.val, g, e, red, xc);
      else
        dia_kernel<C, 0, K, 1, G, E, V, true><<<nbp, BS, 0, s>>>(m.nr, m.nc, m.w, m.off, m.val, g, e, red, xc);
      if constexpr (K > 0) finalize_kernel<K><<<1, FIN_T, 0, s>>>(nbp, red.partial, red.result);
      HIPCHECK(hipGetLastError());
      return;
    }
  }
  constexpr int RPT = dia_rpt<K>();
  const int nb = (int)(((int64_t)m.nr + BS * RPT - 1) / (BS * RPT));
  const int xc = xcd_chunk(m.xmap, nb, BS * RPT);
  if (m.w == 7) dia_kernel<C, 7, K, RPT><<<nb, BS, 0, s>>>(m.nr, m.nc, m.w, m.off, m.val, g, e, red, xc);
  else dia_kernel<C, 0, K, RPT><<<nb, BS, 0, s>>>(m.nr, m.nc, m.w, m.off, m.val, g, e, red, xc);
  if constexpr (K > 0) finalize_kernel<K><<<1, FIN_T, 0, s>>>(nb, red.partial, red.result);
  HIPCHECK(hipGetLastError());
}

template <int K = 0, typename CT = void, typename V, class G, class E>
static void spmv_launch(const DMat<V>& m, G g, E e, hipStream_t s, Red red = Red{nullptr, nullptr}) {
  using C = typename std::conditional<std::is_void<CT>::value, V, CT>::type;
  if (m.nr == 0) return;
  if (m.fmt == FMT_DIA) {
    if constexpr (!std::is_void<typename NtForm<E>::type>::value) {
      if (m.nr >= nt_min_rows()) {
        dia_launch<K, C>(m, g, NtForm<E>::make(e), s, red);
        return;
      }
    }
    dia_launch<K, C>(m, g, e, s, red);
    return;
  }
  const int nb = spmv_blocks(m);
  if (m.fmt == FMT_ELL) {
    const int xc = xcd_chunk(m.xmap, nb, BS);
    switch (m.w) {
      case 4: ell_kernel<C, 4, K><<<nb, BS, 0, s>>>(m.nr, m.w, m.ld, m.ja, m.val, g, e, red, xc); break;
      case 5: ell_kernel<C, 5, K><<<nb, BS, 0, s>>>(m.nr, m.w, m.ld, m.ja, m.val, g, e, red, xc); break;
      case 6: ell_kernel<C, 6, K><<<nb, BS, 0, s>>>(m.nr, m.w, m.ld, m.ja, m.val, g, e, red, xc); break;
      case 7: ell_kernel<C, 7, K><<<nb, BS, 0, s>>>(m.nr, m.w, m.ld, m.ja, m.val, g, e, red, xc); break;
      default: ell_kernel<C, 0, K><<<nb, BS, 0, s>>>(m.nr, m.w, m.ld, m.ja, m.val, g, e, red, xc); break;
    }
  } else if (m.fmt == FMT_SELL) {
    sell_kernel<C, K><<<nb, BS, 0, s>>>(m.nr, m.sptr, m.perm, m.ja, m.val, g, e, red, xcd_chunk(m.xmap, nb, BS));
  } else if (m.fmt == FMT_SELD) {
    seld_kernel<C, K><<<nb, BS, 0, s>>>(m.nr, m.nc, m.sptr, m.perm, m.col0, m.dja, m.val, g, e, red,
                                         xcd_chunk(m.xmap, nb, BS));
  } else {
    switch (m.lanes) {
      case 1: csr_launch_l<1, K, C>(m, nb, g, e, s, red); break;
      case 2: csr_launch_l<2, K, C>(m, nb, g, e, s, red); break;
      case 4: csr_launch_l<4, K, C>(m, nb, g, e, s, red); break;
      case 8: csr_launch_l<8, K, C>(m, nb, g, e, s, red); break;
      case 16: csr_launch_l<16, K, C>(m, nb, g, e, s, red); break;
      case 32: csr_launch_l<32, K, C>(m, nb, g, e, s, red); break;
      default: csr_launch_l<64, K, C>(m, nb, g, e, s, red); break;
    }
  }
  if constexpr (K > 0) finalize_kernel<K><<<1, FIN_T, 0, s>>>(nb, red.partial, red.result);
  HIPCHECK(hipGetLastError());
}

// True when every value of h is exactly representable in FP32 (then an
// FP32-stored copy, summed in FP64, gives bit-identical FP64 products).
static bool fp32_exact(const HCsr& h) {
  for (double v : h.val) {
    const float f = (float)v;
    if (!std::isfinite(v) || (double)f != v) return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Autotuner: storage of each large operator picked by timing y = M x in every
// candidate layout (ELL when eligible, SELL, CSR with 1..16 lanes per row up
// to the mean row length) on the device.  Layouts that keep each row's stored
// order (ELL, SELL, one-lane CSR) win ties within 3 %.  Disabled by
// MLD_AUTOTUNE=0 (then the static rule of choose_format applies).
// ---------------------------------------------------------------------------
constexpr int TUNE_MIN_ROWS = 32768;

static bool autotune_enabled() {
  const char* e = std::getenv("MLD_AUTOTUNE");
  return !(e && e[0] == '0');
}

// Average time (ms) of y = M x over reps launches, best of three batches.
template <typename V>
static float time_spmv(const DMat<V>& m, const V* x, V* y, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  spmv_launch(m, GatherX<V, V>{x}, EpiSet<V, V>{y}, s);  // warm-up
  float msmin = 1e30f;
  const int reps = 5;
  for (int trial = 0; trial < 3; ++trial) {
    HIPCHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) spmv_launch(m, GatherX<V, V>{x}, EpiSet<V, V>{y}, s);
    HIPCHECK(hipEventRecord(e1, s));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    msmin = std::min(msmin, ms);
  }
  return msmin / reps;
}

template <typename V>
static DMat<V> upload_tuned(const HCsr& h, DevArena& ar, hipStream_t s, int force) {
  if (force != 0 || h.nr < TUNE_MIN_ROWS || !autotune_enabled()) return upload_mat<V>(h, ar, s, force);
  // every temporary is released also when a candidate layout or a timing
  // launch throws
  DevArena tmp;
  V* x = tmp.alloc<V>((size_t)std::max(1, h.nc));
  V* y = tmp.alloc<V>((size_t)std::max(1, h.nr));
  fill_kernel<V><<<nblk(h.nc), BS, 0, s>>>(h.nc, x, (V)1);
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
  const DevSrc src = make_src(h, tmp, s);
  int bforce = 6, blanes = 0;
  SellLayouts layouts;
  std::vector<int> offs = dia_offsets(h);
  const bool dia = dia_eligible(h, offs);
  if (!dia) {  // DIA reads strictly fewer bytes than any other layout
    const int maxrow = host_max_row(h);
    const double avg = (double)h.nnz() / std::max(1, h.nr);
    struct Cand { int force, lanes; bool ordered; };
    std::vector<Cand> cands;
    if (maxrow <= 16 && (double)h.nr * maxrow <= 1.02 * (double)h.nnz()) {
      cands.push_back({2, 0, true});  // an almost padding-free ELL is not worth a search
    } else {
      if (maxrow <= 64 && (double)h.nr * maxrow <= 1.25 * (double)h.nnz()) cands.push_back({2, 0, true});
      cands.push_back({3, 0, true});
      cands.push_back({4, 0, true});
      cands.push_back({5, 0, true});
      cands.push_back({7, 0, true});
      // (rows of a few entries: the per-row first column eats the saving)
      if (avg >= 8.0 && host_seld_ok(h)) {
        cands.push_back({8, 0, true});
        cands.push_back({9, 0, true});
        cands.push_back({10, 0, true});
      }
      for (int l = 1; l <= 64; l *= 2)
        if (l == 1 || l <= avg) cands.push_back({1, l, l == 1});
    }
    double best = 1e300;
    int bi = 0;
    for (size_t c = 0; c < cands.size() && cands.size() > 1; ++c) {
      DevArena ctmp;
      FmtChoice fc = choose_format(h, cands[c].force);
      if (cands[c].lanes > 0) { fc.fmt = FMT_CSR; fc.lanes = cands[c].lanes; }
      DMat<V> m = build_mat<V>(h, src, ctmp, s, fc, cands[c].force, &layouts);
      const double t = time_spmv(m, x, y, s, e0, e1) * (cands[c].ordered ? 1.0 : 1.03);
      if (t < best) { best = t; bi = (int)c; }
    }
    bforce = cands[bi].force;
    blanes = cands[bi].lanes;
    if (std::getenv("MLD_VERBOSE"))
      std::fprintf(stderr, "[mld] %d x %d operator: layout %s (%s %d), %.1f us\n", h.nr, h.nc,
                   blanes > 0 ? "CSR" : (bforce == 2 ? "ELL" : (bforce >= 8 ? "SELD" : "SELL")),
                   blanes > 0 ? "lanes" : "sigma",
                   blanes > 0 ? blanes : (bforce == 2 ? 0 : sell_sigma(bforce)), best * 1e3);
  }
  FmtChoice fb = dia ? FmtChoice{FMT_DIA, 1, (int)offs.size(), offs} : choose_format(h, bforce);
  if (blanes > 0) { fb.fmt = FMT_CSR; fb.lanes = blanes; }
  DMat<V> m = build_mat<V>(h, src, ar, s, fb, bforce, &layouts);
  // row-block order of the chosen layout: chunks of 8192 rows per XCD
  // (default), identity, one contiguous range per XCD, chunks of 32768 or
  // 131072 rows; the default wins ties within 2 % (MLD_XMAP=0..4 fixes it)
  if (const char* e = std::getenv("MLD_XMAP")) {
    m.xmap = std::max(0, std::min(4, std::atoi(e)));
  } else {
    float tb = 1e30f;
    int xb = 1;
    for (int xm : {1, 0, 2, 3, 4}) {
      m.xmap = xm;
      const float t = time_spmv(m, x, y, s, e0, e1) * (xm == 1 ? 1.0f : 1.02f);
      if (t < tb) { tb = t; xb = xm; }
    }
    m.xmap = xb;
    if (std::getenv("MLD_VERBOSE"))
      std::fprintf(stderr, "[mld] %d x %d operator (%s): row-block order %d, %.1f us\n", h.nr, h.nc,
                   m.fmt == FMT_DIA ? "DIA" : "tuned layout", xb, tb * 1e3);
  }
  HIPCHECK(hipStreamSynchronize(s));  // the timing launches are done before tmp is freed
  return m;
}

}  // namespace mld
