// Device part of the hierarchy setup: the smoothed-aggregation products of
// one level (mlprec/impl/mld_daggrmat_smth_asb.f90:291,316,345,399),
//   P  = (I - omega D^-1 A) Ptilde,   AP = A P,   R = P^T,   Ac = R (A P),
// on the GPU.  Aggregation (the greedy, order-dependent phases of
// mld_d_dec_map_bld) and the omega estimate stay on the host.  Every product
// is the host's row-wise Gustavson product term for term: each C_ij is the
// sum over the left factor's stored row i, in that order, of x_ik y_kj,
// accumulated from 0.0, columns ascending; the transpose lists each row's
// columns ascending.  The results are therefore bit-identical to the host
// build, and they are copied back into the host hierarchy (the smoother
// build, the next level's aggregation and the inspection entry points read
// them there).
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "mld_devcommon.hpp"

namespace mld {
namespace {

// distinct columns one output row may have (beyond: the host products)
constexpr int SG_MAXC = 128;

// Row-wise Gustavson C = X Y, one thread per row of X.  PASS 0 counts the
// distinct columns of row i (cnt[i]; more than SG_MAXC sets *over); PASS 1
// accumulates the row in X's stored order and writes it with ascending
// columns at cptr[i].
template <int PASS>
__global__ __launch_bounds__(256) void sg_rowprod_kernel(int nr, const int64_t* __restrict__ xirp,
                                                         const int* __restrict__ xja, const double* __restrict__ xval,
                                                         const int64_t* __restrict__ yirp,
                                                         const int* __restrict__ yja, const double* __restrict__ yval,
                                                         int64_t* __restrict__ cnt, const int64_t* __restrict__ cptr,
                                                         int* __restrict__ cja, double* __restrict__ cval,
                                                         int* __restrict__ over) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nr) return;
  int col[SG_MAXC];
  double acc[SG_MAXC];
  int m = 0;
  bool ovf = false;
  for (int64_t ka = xirp[i]; ka < xirp[i + 1]; ++ka) {
    const int k = xja[ka];
    const double xk = xval[ka];
    for (int64_t kb = yirp[k]; kb < yirp[k + 1]; ++kb) {
      const int j = yja[kb];
      int q = 0;
      while (q < m && col[q] != j) ++q;
      if (q == m) {
        if (m == SG_MAXC) {
          ovf = true;
          continue;
        }
        col[m] = j;
        acc[m] = 0.0;
        ++m;
      }
      if (PASS == 1) acc[q] = acc[q] + xk * yval[kb];
    }
  }
  if (PASS == 0) {
    cnt[i] = m;
    if (ovf) atomicOr(over, 1);
    return;
  }
  for (int a = 1; a < m; ++a) {  // ascending columns (values follow)
    const int c = col[a];
    const double v = acc[a];
    int b = a - 1;
    while (b >= 0 && col[b] > c) {
      col[b + 1] = col[b];
      acc[b + 1] = acc[b];
      --b;
    }
    col[b + 1] = c;
    acc[b + 1] = v;
  }
  const int64_t o = cptr[i];
  for (int a = 0; a < m; ++a) {
    cja[o + a] = col[a];
    cval[o + a] = acc[a];
  }
}

// S = I - omega D^-1 A with A's structure: s_ik = 1 - omega (a_ik dinv_i)
// on the diagonal, -(omega (a_ik dinv_i)) elsewhere (the host's expressions).
__global__ __launch_bounds__(256) void sg_smooth_kernel(int n, const int64_t* __restrict__ irp,
                                                        const int* __restrict__ ja, const double* __restrict__ aval,
                                                        const double* __restrict__ dinv, double omega,
                                                        double* __restrict__ sval) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double di = dinv[i];
  for (int64_t k = irp[i]; k < irp[i + 1]; ++k) {
    const double v = aval[k] * di;
    sval[k] = (ja[k] == i) ? 1.0 - omega * v : -(omega * v);
  }
}

__global__ __launch_bounds__(256) void sg_iota_kernel(int64_t n, int64_t* __restrict__ a) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = i;
}
__global__ __launch_bounds__(256) void sg_fill_kernel(int64_t n, double* __restrict__ a, double v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = v;
}
// transpose keys: (column << 32) | row for every entry of P
__global__ __launch_bounds__(256) void sg_tkeys_kernel(int nr, const int64_t* __restrict__ irp,
                                                       const int* __restrict__ ja, uint64_t* __restrict__ key,
                                                       int64_t* __restrict__ ccnt) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nr) return;
  for (int64_t k = irp[i]; k < irp[i + 1]; ++k) {
    key[k] = ((uint64_t)(uint32_t)ja[k] << 32) | (uint32_t)i;
    atomicAdd((unsigned long long*)(ccnt + ja[k]), 1ull);
  }
}
__global__ __launch_bounds__(256) void sg_tsplit_kernel(int64_t nz, const uint64_t* __restrict__ key,
                                                        int* __restrict__ ja) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k < nz) ja[k] = (int)(uint32_t)(key[k] & 0xffffffffull);
}

struct DCsr {
  int nr = 0, nc = 0;
  int64_t nnz = 0;
  int64_t* irp = nullptr;
  int* ja = nullptr;
  double* val = nullptr;
};

inline int nblk64(int64_t n) { return (int)((n + 255) / 256); }

// exclusive scan of cnt[0..n) into ptr[0..n] (ptr[n] = total), returned
int64_t scan_counts(const int64_t* cnt, int64_t* ptr, int n, DevArena& tmp, hipStream_t s) {
  size_t bytes = 0;
  HIPCHECK(rocprim::exclusive_scan(nullptr, bytes, cnt, ptr, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), s));
  void* ws = tmp.alloc<char>(bytes);
  HIPCHECK(rocprim::exclusive_scan(ws, bytes, cnt, ptr, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), s));
  int64_t last_ptr = 0, last_cnt = 0;
  if (n > 0) {
    HIPCHECK(hipMemcpyAsync(&last_ptr, ptr + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(&last_cnt, cnt + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  }
  HIPCHECK(hipStreamSynchronize(s));
  const int64_t tot = last_ptr + last_cnt;
  HIPCHECK(hipMemcpyAsync(ptr + n, &tot, sizeof(int64_t), hipMemcpyHostToDevice, s));
  HIPCHECK(hipStreamSynchronize(s));
  return tot;
}

// C = X Y (X: nr rows; Y: nc columns); false when a row has more than
// SG_MAXC distinct columns
bool rowprod(const DCsr& X, const DCsr& Y, DCsr& C, DevArena& ar, DevArena& tmp, int* over, hipStream_t s) {
  C.nr = X.nr;
  C.nc = Y.nc;
  int64_t* cnt = tmp.alloc<int64_t>((size_t)X.nr + 1);
  C.irp = ar.alloc<int64_t>((size_t)X.nr + 1);
  HIPCHECK(hipMemsetAsync(over, 0, sizeof(int), s));
  sg_rowprod_kernel<0><<<nblk64(X.nr), 256, 0, s>>>(X.nr, X.irp, X.ja, X.val, Y.irp, Y.ja, Y.val, cnt, nullptr,
                                                     nullptr, nullptr, over);
  HIPCHECK(hipGetLastError());
  int hover = 0;
  HIPCHECK(hipMemcpyAsync(&hover, over, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  if (hover) return false;
  C.nnz = scan_counts(cnt, C.irp, X.nr, tmp, s);
  if (C.nnz >= ((int64_t)1 << 31)) return false;
  C.ja = ar.alloc<int>((size_t)C.nnz);
  C.val = ar.alloc<double>((size_t)C.nnz);
  sg_rowprod_kernel<1><<<nblk64(X.nr), 256, 0, s>>>(X.nr, X.irp, X.ja, X.val, Y.irp, Y.ja, Y.val, nullptr, C.irp,
                                                     C.ja, C.val, over);
  HIPCHECK(hipGetLastError());
  return true;
}

// R = P^T with each row's columns ascending: entries sorted by
// (column, row) with a radix sort of 64-bit keys
void transpose_dev(const DCsr& P, DCsr& R, DevArena& ar, DevArena& tmp, hipStream_t s) {
  R.nr = P.nc;
  R.nc = P.nr;
  R.nnz = P.nnz;
  uint64_t* k0 = tmp.alloc<uint64_t>((size_t)P.nnz);
  uint64_t* k1 = tmp.alloc<uint64_t>((size_t)P.nnz);
  double* v1 = tmp.alloc<double>((size_t)P.nnz);
  int64_t* ccnt = tmp.alloc<int64_t>((size_t)P.nc + 1);
  HIPCHECK(hipMemsetAsync(ccnt, 0, sizeof(int64_t) * ((size_t)P.nc + 1), s));
  sg_tkeys_kernel<<<nblk64(P.nr), 256, 0, s>>>(P.nr, P.irp, P.ja, k0, ccnt);
  HIPCHECK(hipGetLastError());
  int hibits = 1;
  while (hibits < 32 && ((int64_t)1 << hibits) < (int64_t)P.nc) ++hibits;
  size_t bytes = 0;
  HIPCHECK(rocprim::radix_sort_pairs(nullptr, bytes, k0, k1, P.val, v1, (size_t)P.nnz, 0, 32 + hibits, s));
  void* ws = tmp.alloc<char>(bytes);
  HIPCHECK(rocprim::radix_sort_pairs(ws, bytes, k0, k1, P.val, v1, (size_t)P.nnz, 0, 32 + hibits, s));
  R.irp = ar.alloc<int64_t>((size_t)R.nr + 1);
  scan_counts(ccnt, R.irp, R.nr, tmp, s);
  R.ja = ar.alloc<int>((size_t)R.nnz);
  R.val = v1;  // tmp-owned: copied out before tmp is released
  sg_tsplit_kernel<<<nblk64(R.nnz), 256, 0, s>>>(R.nnz, k1, R.ja);
  HIPCHECK(hipGetLastError());
}

void upload_csr(const HCsr& h, DCsr& d, DevArena& ar, hipStream_t s, bool vals = true) {
  d.nr = h.nr;
  d.nc = h.nc;
  d.nnz = h.nnz();
  d.irp = ar.alloc<int64_t>((size_t)h.nr + 1);
  d.ja = ar.alloc<int>((size_t)std::max<int64_t>(1, d.nnz));
  host_to_device(d.irp, h.irp.data(), sizeof(int64_t) * ((size_t)h.nr + 1), s);
  if (d.nnz) host_to_device(d.ja, h.ja.data(), sizeof(int) * (size_t)d.nnz, s);
  if (vals) {
    d.val = ar.alloc<double>((size_t)std::max<int64_t>(1, d.nnz));
    if (d.nnz) host_to_device(d.val, h.val.data(), sizeof(double) * (size_t)d.nnz, s);
  }
}

void download_csr(const DCsr& d, HCsr& h, hipStream_t s) {
  h.clear();
  h.nr = d.nr;
  h.nc = d.nc;
  h.irp.resize((size_t)d.nr + 1);
  h.ja.resize((size_t)d.nnz);
  h.val.resize((size_t)d.nnz);
  device_to_host(h.irp.data(), d.irp, sizeof(int64_t) * ((size_t)d.nr + 1), s);
  if (d.nnz) {
    device_to_host(h.ja.data(), d.ja, sizeof(int) * (size_t)d.nnz, s);
    device_to_host(h.val.data(), d.val, sizeof(double) * (size_t)d.nnz, s);
  }
  HIPCHECK(hipStreamSynchronize(s));
}

}  // namespace

int dev_smoothed_products(int device, const HCsr& a, const HCsr* s_filtered, const std::vector<double>& dinv,
                          double omega, const std::vector<int>& ilaggr, int naggr, HCsr& p, HCsr& r, HCsr& ac,
                          std::string& err) {
  try {
    HIPCHECK(hipSetDevice(device));
    hipStream_t s = nullptr;
    HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
      }
    } sg{s};
    DevArena ar, tmp;
    int* over = tmp.alloc<int>(1);
    const int n = a.nr;
    // A and S (S shares A's structure unless filtered)
    DCsr A, S;
    upload_csr(a, A, ar, s);
    if (s_filtered) {
      upload_csr(*s_filtered, S, ar, s);
    } else {
      S = A;
      S.val = ar.alloc<double>((size_t)std::max<int64_t>(1, A.nnz));
      double* dd = tmp.alloc<double>((size_t)std::max(1, n));
      host_to_device(dd, dinv.data(), sizeof(double) * (size_t)n, s);
      sg_smooth_kernel<<<nblk64(n), 256, 0, s>>>(n, A.irp, A.ja, A.val, dd, omega, S.val);
      HIPCHECK(hipGetLastError());
    }
    // Ptilde: row i holds 1.0 in column ilaggr[i]
    DCsr T;
    T.nr = n;
    T.nc = naggr;
    T.nnz = n;
    T.irp = tmp.alloc<int64_t>((size_t)n + 1);
    T.ja = tmp.alloc<int>((size_t)std::max(1, n));
    T.val = tmp.alloc<double>((size_t)std::max(1, n));
    sg_iota_kernel<<<nblk64((int64_t)n + 1), 256, 0, s>>>((int64_t)n + 1, T.irp);
    sg_fill_kernel<<<nblk64(n), 256, 0, s>>>(n, T.val, 1.0);
    host_to_device(T.ja, ilaggr.data(), sizeof(int) * (size_t)n, s);
    HIPCHECK(hipGetLastError());
    DCsr P, AP, R, AC;
    if (!rowprod(S, T, P, ar, tmp, over, s) || !rowprod(A, P, AP, ar, tmp, over, s)) {
      err = "device setup: a product row exceeds the accumulator";
      return E_UNSUPPORTED;
    }
    transpose_dev(P, R, ar, tmp, s);
    if (!rowprod(R, AP, AC, ar, tmp, over, s)) {
      err = "device setup: a product row exceeds the accumulator";
      return E_UNSUPPORTED;
    }
    download_csr(P, p, s);
    download_csr(R, r, s);
    download_csr(AC, ac, s);
    return 0;
  } catch (const std::exception& e) {
    err = e.what();
    return E_DEVICE;
  }
}

}  // namespace mld
