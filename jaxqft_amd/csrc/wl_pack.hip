// Conversion between the reference array layouts and the internal
// structure-of-arrays layout.
//
// Reference spinor: (cols, L0, L1, L2, L3, Ns=4, Nc) complex, C-order
// (su3_wilson_nf2.py:815-817).  Reference links: BMXYIJ (B, 4, L0..L3, Nc, Nc)
// or BXYMIJ (B, L0..L3, 4, Nc, Nc); U(1) drops the (Nc, Nc) axes
// (su3_ym.py:246-249, u1_ym.py:189-192).
#include "wl_device.h"
#include "wl_kernels.h"

namespace wl {

template <int P, int NC>
struct LinkStore;

template <int NC>
struct LinkStore<F64, NC> {
  template <int RECON>
  __device__ static void store(void* base, size_t V, size_t site, const cx<double> (&u)[NC][NC]) {
    double2* p = reinterpret_cast<double2*>(base);
    constexpr int ROWS = (NC == 3 && RECON == 12) ? 2 : NC;
    for (int a = 0; a < ROWS; ++a)
      for (int b = 0; b < NC; ++b) p[(a * NC + b) * V + site] = make_double2(u[a][b].re, u[a][b].im);
  }
};

template <int NC>
struct LinkStore<F32, NC> {
  template <int RECON>
  __device__ static void store(void* base, size_t V, size_t site, const cx<double> (&u)[NC][NC]) {
    if (NC == 3 && RECON == 12) {
      float4* p = reinterpret_cast<float4*>(base);
      for (int j = 0; j < 3; ++j) {
        int c0 = 2 * j, c1 = 2 * j + 1;
        p[j * V + site] = make_float4((float)u[c0 / NC][c0 % NC].re, (float)u[c0 / NC][c0 % NC].im,
                                      (float)u[c1 / NC][c1 % NC].re, (float)u[c1 / NC][c1 % NC].im);
      }
    } else {
      float2* p = reinterpret_cast<float2*>(base);
      for (int c = 0; c < NC * NC; ++c)
        p[c * V + site] = make_float2((float)u[c / NC][c % NC].re, (float)u[c / NC][c % NC].im);
    }
  }
};

template <int NC>
struct LinkStore<F16, NC> {
  template <int RECON>
  __device__ static void store(void* base, size_t V, size_t site, const cx<double> (&u)[NC][NC]) {
    if (NC == 3 && RECON == 12) {
      uint4* pa = reinterpret_cast<uint4*>(base);
      uint2* pb = reinterpret_cast<uint2*>(pa + V);
      uint32_t h[6];
      for (int c = 0; c < 6; ++c) h[c] = f2h((float)u[c / NC][c % NC].re, (float)u[c / NC][c % NC].im);
      pa[site] = make_uint4(h[0], h[1], h[2], h[3]);
      pb[site] = make_uint2(h[4], h[5]);
    } else {
      uint32_t* p = reinterpret_cast<uint32_t*>(base);
      for (int c = 0; c < NC * NC; ++c) p[c * V + site] = f2h((float)u[c / NC][c % NC].re, (float)u[c / NC][c % NC].im);
    }
  }
};

template <int P, int NC, class S>
__global__ __launch_bounds__(NT) void k_pack_spinor(const S* __restrict__ src, void* __restrict__ dst, long long V,
                                                    long long Vp, int ncol, size_t col_bytes) {
  const size_t site = (size_t)blockIdx.x * NT + threadIdx.x;
  const int col = blockIdx.y;
  if (site >= (size_t)V) return;
  using T = typename PT<P>::acc;
  const S* s = src + ((size_t)col * V + site) * (4 * NC) * 2;
  cx<T> v[4 * NC];
#pragma unroll
  for (int c = 0; c < 4 * NC; ++c) v[c] = {(T)s[2 * c], (T)s[2 * c + 1]};
  SpinorIO<P, NC>::store(reinterpret_cast<char*>(dst) + col * col_bytes, Vp, site, v);
}

template <int P, int NC, class S>
__global__ __launch_bounds__(NT) void k_unpack_spinor(const void* __restrict__ src, S* __restrict__ dst, long long V,
                                                      long long Vp, int ncol, size_t col_bytes) {
  const size_t site = (size_t)blockIdx.x * NT + threadIdx.x;
  const int col = blockIdx.y;
  if (site >= (size_t)V) return;
  using T = typename PT<P>::acc;
  cx<T> v[4 * NC];
  SpinorIO<P, NC>::load(reinterpret_cast<const char*>(src) + col * col_bytes, Vp, site, v);
  S* d = dst + ((size_t)col * V + site) * (4 * NC) * 2;
#pragma unroll
  for (int c = 0; c < 4 * NC; ++c) {
    d[2 * c] = (S)v[c].re;
    d[2 * c + 1] = (S)v[c].im;
  }
}

// layout: 0 = BMXYIJ, 1 = BXYMIJ
template <int P, int NC, int RECON, class S>
__global__ __launch_bounds__(NT) void k_pack_gauge(const S* __restrict__ src, void* __restrict__ dst, long long V,
                                                   long long Vp, int layout, size_t cfg_bytes, size_t dir_bytes) {
  const size_t site = (size_t)blockIdx.x * NT + threadIdx.x;
  const int mu = blockIdx.y;
  const int cfg = blockIdx.z;
  if (site >= (size_t)V) return;
  size_t off = (layout == 0) ? (((size_t)cfg * 4 + mu) * V + site) : (((size_t)cfg * V + site) * 4 + mu);
  const S* s = src + off * NC * NC * 2;
  cx<double> u[NC][NC];
  for (int a = 0; a < NC; ++a)
    for (int b = 0; b < NC; ++b) u[a][b] = {(double)s[2 * (a * NC + b)], (double)s[2 * (a * NC + b) + 1]};
  char* d = reinterpret_cast<char*>(dst) + cfg * cfg_bytes + mu * dir_bytes;
  LinkStore<P, NC>::template store<RECON>(d, Vp, site, u);
}

// Re-encode an internal FP64 full gauge field into another precision/recon.
template <int P, int NC, int RECON>
__global__ __launch_bounds__(NT) void k_convert_gauge(const void* __restrict__ src, void* __restrict__ dst, long long V,
                                                      long long Vp, size_t src_cfg_bytes, size_t src_dir_bytes, size_t cfg_bytes,
                                                      size_t dir_bytes) {
  const size_t site = (size_t)blockIdx.x * NT + threadIdx.x;
  const int mu = blockIdx.y;
  const int cfg = blockIdx.z;
  if (site >= (size_t)V) return;
  const double2* s =
      reinterpret_cast<const double2*>(reinterpret_cast<const char*>(src) + cfg * src_cfg_bytes + mu * src_dir_bytes);
  cx<double> u[NC][NC];
  for (int a = 0; a < NC; ++a)
    for (int b = 0; b < NC; ++b) {
      double2 t = s[(a * NC + b) * Vp + site];
      u[a][b] = {t.x, t.y};
    }
  char* d = reinterpret_cast<char*>(dst) + cfg * cfg_bytes + mu * dir_bytes;
  LinkStore<P, NC>::template store<RECON>(d, Vp, site, u);
}

// Largest |U_2j - conj(U_0 x U_1)_j| over the SU(3) links of the FP64 master:
// how far the stored third row is from the one a 12-real encoding rebuilds.
// A wavefront max (NaN kept), then one integer atomic max per wavefront on the
// bits of the non-negative double (for those, bit order is value order).
__device__ __forceinline__ double max_keep_nan(double a, double b) { return (a > b || a != a) ? a : b; }

__global__ __launch_bounds__(NT) void k_row2_deviation(const void* __restrict__ src, long long V, long long Vp,
                                                       size_t cfg_bytes, size_t dir_bytes,
                                                       unsigned long long* __restrict__ out) {
  const size_t site = (size_t)blockIdx.x * NT + threadIdx.x;
  const int mu = blockIdx.y;
  const int cfg = blockIdx.z;
  double dev = 0.0;
  if (site < (size_t)V) {
    const double2* s =
        reinterpret_cast<const double2*>(reinterpret_cast<const char*>(src) + cfg * cfg_bytes + mu * dir_bytes);
    cx<double> u[3][3], r[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        const double2 t = s[(a * 3 + b) * Vp + site];
        u[a][b] = {t.x, t.y};
        r[a][b] = u[a][b];
      }
    reconstruct_row2(r);
    for (int b = 0; b < 3; ++b) {
      const double dr = u[2][b].re - r[2][b].re, di = u[2][b].im - r[2][b].im;
      dev = max_keep_nan(dev, sqrt(dr * dr + di * di));
    }
  }
  for (int off = 32; off > 0; off >>= 1) dev = max_keep_nan(dev, __shfl_xor(dev, off));
  if ((threadIdx.x & 63u) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(dev));
}

// Even-odd layouts.  Spinors: the sites of parity q of a full-lattice
// reference array, compact (ascending full index).  Unpacking writes the
// whole full-lattice array: the compact values at parity-q sites and zero at
// the others (the parity projection of su3_wilson_nf2.py:884-892).
template <int P, int NC, class S>
__global__ __launch_bounds__(NT) void k_pack_spinor_eo(const S* __restrict__ src, void* __restrict__ dst, Geom g,
                                                       long long Vh, long long Vhp, int q, size_t col_bytes) {
  const size_t i = (size_t)blockIdx.x * NT + threadIdx.x;
  const int col = blockIdx.y;
  if (i >= (size_t)Vh) return;
  using T = typename PT<P>::acc;
  const size_t site = eo_full_index(g, q, i);
  const S* s = src + ((size_t)col * g.V + site) * (4 * NC) * 2;
  cx<T> v[4 * NC];
#pragma unroll
  for (int c = 0; c < 4 * NC; ++c) v[c] = {(T)s[2 * c], (T)s[2 * c + 1]};
  SpinorIO<P, NC>::store(reinterpret_cast<char*>(dst) + col * col_bytes, Vhp, i, v);
}

template <int P, int NC, class S>
__global__ __launch_bounds__(NT) void k_unpack_spinor_eo(const void* __restrict__ src, S* __restrict__ dst, Geom g,
                                                         long long Vhp, int q, size_t col_bytes) {
  const size_t site = (size_t)blockIdx.x * NT + threadIdx.x;
  const int col = blockIdx.y;
  if (site >= (size_t)g.V) return;
  using T = typename PT<P>::acc;
  size_t rem = site;
  const int x3 = (int)(rem % g.L[3]);
  rem /= g.L[3];
  const int x2 = (int)(rem % g.L[2]);
  rem /= g.L[2];
  const int x1 = (int)(rem % g.L[1]);
  const int x0 = (int)(rem / g.L[1]);
  cx<T> v[4 * NC];
  if (((x0 + x1 + x2 + x3) & 1) == q) {
    SpinorIO<P, NC>::load(reinterpret_cast<const char*>(src) + col * col_bytes, Vhp, site >> 1, v);
  } else {
#pragma unroll
    for (int c = 0; c < 4 * NC; ++c) v[c] = {T(0), T(0)};
  }
  S* d = dst + ((size_t)col * g.V + site) * (4 * NC) * 2;
#pragma unroll
  for (int c = 0; c < 4 * NC; ++c) {
    d[2 * c] = (S)v[c].re;
    d[2 * c + 1] = (S)v[c].im;
  }
}

// Links per parity: dst[cfg][q][mu] holds U_mu at the parity-q sites, compact.
template <int P, int NC, int RECON>
__global__ __launch_bounds__(NT) void k_convert_gauge_eo(const void* __restrict__ src, void* __restrict__ dst, Geom g,
                                                         long long Vh, long long Vhp, size_t src_cfg_bytes,
                                                         size_t src_dir_bytes, size_t cfg_bytes, size_t dir_bytes) {
  const size_t i = (size_t)blockIdx.x * NT + threadIdx.x;
  const int mu = blockIdx.y & 3;
  const int q = blockIdx.y >> 2;
  const int cfg = blockIdx.z;
  if (i >= (size_t)Vh) return;
  const size_t site = eo_full_index(g, q, i);
  const double2* s =
      reinterpret_cast<const double2*>(reinterpret_cast<const char*>(src) + cfg * src_cfg_bytes + mu * src_dir_bytes);
  cx<double> u[NC][NC];
  for (int a = 0; a < NC; ++a)
    for (int b = 0; b < NC; ++b) {
      double2 t = s[(a * NC + b) * g.Vp + site];
      u[a][b] = {t.x, t.y};
    }
  char* d = reinterpret_cast<char*>(dst) + cfg * cfg_bytes + (q * 4 + mu) * dir_bytes;
  LinkStore<P, NC>::template store<RECON>(d, Vhp, i, u);
}

// ---------------------------------------------------------------------------
// host dispatch.  dtype: 0 complex128, 1 complex64
// ---------------------------------------------------------------------------
template <int P, int NC>
static hipError_t pack_spinor_t(const void* src, int dtype, void* dst, long long V, long long Vp, int ncol, size_t col_bytes,
                                hipStream_t st) {
  dim3 grid((unsigned)((V + NT - 1) / NT), (unsigned)ncol), block(NT);
  if (dtype == 0)
    hipLaunchKernelGGL((k_pack_spinor<P, NC, double>), grid, block, 0, st, (const double*)src, dst, V, Vp, ncol, col_bytes);
  else
    hipLaunchKernelGGL((k_pack_spinor<P, NC, float>), grid, block, 0, st, (const float*)src, dst, V, Vp, ncol, col_bytes);
  return hipGetLastError();
}

template <int P, int NC>
static hipError_t unpack_spinor_t(const void* src, void* dst, int dtype, long long V, long long Vp, int ncol, size_t col_bytes,
                                  hipStream_t st) {
  dim3 grid((unsigned)((V + NT - 1) / NT), (unsigned)ncol), block(NT);
  if (dtype == 0)
    hipLaunchKernelGGL((k_unpack_spinor<P, NC, double>), grid, block, 0, st, src, (double*)dst, V, Vp, ncol, col_bytes);
  else
    hipLaunchKernelGGL((k_unpack_spinor<P, NC, float>), grid, block, 0, st, src, (float*)dst, V, Vp, ncol, col_bytes);
  return hipGetLastError();
}

hipError_t pack_spinor(int prec, int nc, const void* src, int dtype, void* dst, long long V, long long Vp, int ncol,
                       size_t col_bytes, hipStream_t st) {
  if (nc == 3) {
    if (prec == F64) return pack_spinor_t<F64, 3>(src, dtype, dst, V, Vp, ncol, col_bytes, st);
    if (prec == F32) return pack_spinor_t<F32, 3>(src, dtype, dst, V, Vp, ncol, col_bytes, st);
    return pack_spinor_t<F16, 3>(src, dtype, dst, V, Vp, ncol, col_bytes, st);
  }
  if (prec == F64) return pack_spinor_t<F64, 1>(src, dtype, dst, V, Vp, ncol, col_bytes, st);
  if (prec == F32) return pack_spinor_t<F32, 1>(src, dtype, dst, V, Vp, ncol, col_bytes, st);
  return pack_spinor_t<F16, 1>(src, dtype, dst, V, Vp, ncol, col_bytes, st);
}

hipError_t unpack_spinor(int prec, int nc, const void* src, void* dst, int dtype, long long V, long long Vp, int ncol,
                         size_t col_bytes, hipStream_t st) {
  if (nc == 3) {
    if (prec == F64) return unpack_spinor_t<F64, 3>(src, dst, dtype, V, Vp, ncol, col_bytes, st);
    if (prec == F32) return unpack_spinor_t<F32, 3>(src, dst, dtype, V, Vp, ncol, col_bytes, st);
    return unpack_spinor_t<F16, 3>(src, dst, dtype, V, Vp, ncol, col_bytes, st);
  }
  if (prec == F64) return unpack_spinor_t<F64, 1>(src, dst, dtype, V, Vp, ncol, col_bytes, st);
  if (prec == F32) return unpack_spinor_t<F32, 1>(src, dst, dtype, V, Vp, ncol, col_bytes, st);
  return unpack_spinor_t<F16, 1>(src, dst, dtype, V, Vp, ncol, col_bytes, st);
}

hipError_t pack_gauge_f64(int nc, const void* src, int dtype, int layout, void* dst, long long V, long long Vp, int ncfg,
                          size_t cfg_bytes, size_t dir_bytes, hipStream_t st) {
  dim3 grid((unsigned)((V + NT - 1) / NT), 4, (unsigned)ncfg), block(NT);
  if (nc == 3) {
    if (dtype == 0)
      hipLaunchKernelGGL((k_pack_gauge<F64, 3, 18, double>), grid, block, 0, st, (const double*)src, dst, V, Vp, layout,
                         cfg_bytes, dir_bytes);
    else
      hipLaunchKernelGGL((k_pack_gauge<F64, 3, 18, float>), grid, block, 0, st, (const float*)src, dst, V, Vp, layout,
                         cfg_bytes, dir_bytes);
  } else {
    if (dtype == 0)
      hipLaunchKernelGGL((k_pack_gauge<F64, 1, 18, double>), grid, block, 0, st, (const double*)src, dst, V, Vp, layout,
                         cfg_bytes, dir_bytes);
    else
      hipLaunchKernelGGL((k_pack_gauge<F64, 1, 18, float>), grid, block, 0, st, (const float*)src, dst, V, Vp, layout,
                         cfg_bytes, dir_bytes);
  }
  return hipGetLastError();
}

hipError_t row2_deviation(const void* src, long long V, long long Vp, int ncfg, size_t cfg_bytes, size_t dir_bytes,
                          unsigned long long* out, hipStream_t st) {
  dim3 grid((unsigned)((V + NT - 1) / NT), 4, (unsigned)ncfg), block(NT);
  hipLaunchKernelGGL(k_row2_deviation, grid, block, 0, st, src, V, Vp, cfg_bytes, dir_bytes, out);
  return hipGetLastError();
}

hipError_t convert_gauge(int prec, int nc, int recon, const void* src, void* dst, long long V, long long Vp, int ncfg,
                         size_t src_cfg_bytes, size_t src_dir_bytes, size_t cfg_bytes, size_t dir_bytes,
                         hipStream_t st) {
  dim3 grid((unsigned)((V + NT - 1) / NT), 4, (unsigned)ncfg), block(NT);
#define WL_CG(P, NCV, R)                                                                                        \
  hipLaunchKernelGGL((k_convert_gauge<P, NCV, R>), grid, block, 0, st, src, dst, V, Vp, src_cfg_bytes, src_dir_bytes, \
                     cfg_bytes, dir_bytes)
  if (nc == 3) {
    if (prec == F64 && recon == 12) WL_CG(F64, 3, 12);
    else if (prec == F64) WL_CG(F64, 3, 18);
    else if (prec == F32 && recon == 12) WL_CG(F32, 3, 12);
    else if (prec == F32) WL_CG(F32, 3, 18);
    else if (prec == F16 && recon == 12) WL_CG(F16, 3, 12);
    else WL_CG(F16, 3, 18);
  } else {
    if (prec == F64) WL_CG(F64, 1, 18);
    else if (prec == F32) WL_CG(F32, 1, 18);
    else WL_CG(F16, 1, 18);
  }
#undef WL_CG
  return hipGetLastError();
}

template <int P, int NC>
static hipError_t pack_spinor_eo_t(const void* src, int dtype, void* dst, const Geom& g, long long Vh, long long Vhp,
                                   int q, int ncol, size_t col_bytes, hipStream_t st) {
  dim3 grid((unsigned)((Vh + NT - 1) / NT), (unsigned)ncol), block(NT);
  if (dtype == 0)
    hipLaunchKernelGGL((k_pack_spinor_eo<P, NC, double>), grid, block, 0, st, (const double*)src, dst, g, Vh, Vhp, q,
                       col_bytes);
  else
    hipLaunchKernelGGL((k_pack_spinor_eo<P, NC, float>), grid, block, 0, st, (const float*)src, dst, g, Vh, Vhp, q,
                       col_bytes);
  return hipGetLastError();
}

template <int P, int NC>
static hipError_t unpack_spinor_eo_t(const void* src, void* dst, int dtype, const Geom& g, long long Vhp, int q,
                                     int ncol, size_t col_bytes, hipStream_t st) {
  dim3 grid((unsigned)((g.V + NT - 1) / NT), (unsigned)ncol), block(NT);
  if (dtype == 0)
    hipLaunchKernelGGL((k_unpack_spinor_eo<P, NC, double>), grid, block, 0, st, src, (double*)dst, g, Vhp, q, col_bytes);
  else
    hipLaunchKernelGGL((k_unpack_spinor_eo<P, NC, float>), grid, block, 0, st, src, (float*)dst, g, Vhp, q, col_bytes);
  return hipGetLastError();
}

hipError_t pack_spinor_eo(int prec, int nc, const void* src, int dtype, void* dst, const Geom& g, long long Vh,
                          long long Vhp, int q, int ncol, size_t col_bytes, hipStream_t st) {
  if (nc == 3) {
    if (prec == F64) return pack_spinor_eo_t<F64, 3>(src, dtype, dst, g, Vh, Vhp, q, ncol, col_bytes, st);
    if (prec == F32) return pack_spinor_eo_t<F32, 3>(src, dtype, dst, g, Vh, Vhp, q, ncol, col_bytes, st);
    return pack_spinor_eo_t<F16, 3>(src, dtype, dst, g, Vh, Vhp, q, ncol, col_bytes, st);
  }
  if (prec == F64) return pack_spinor_eo_t<F64, 1>(src, dtype, dst, g, Vh, Vhp, q, ncol, col_bytes, st);
  if (prec == F32) return pack_spinor_eo_t<F32, 1>(src, dtype, dst, g, Vh, Vhp, q, ncol, col_bytes, st);
  return pack_spinor_eo_t<F16, 1>(src, dtype, dst, g, Vh, Vhp, q, ncol, col_bytes, st);
}

hipError_t unpack_spinor_eo(int prec, int nc, const void* src, void* dst, int dtype, const Geom& g, long long Vhp, int q,
                            int ncol, size_t col_bytes, hipStream_t st) {
  if (nc == 3) {
    if (prec == F64) return unpack_spinor_eo_t<F64, 3>(src, dst, dtype, g, Vhp, q, ncol, col_bytes, st);
    if (prec == F32) return unpack_spinor_eo_t<F32, 3>(src, dst, dtype, g, Vhp, q, ncol, col_bytes, st);
    return unpack_spinor_eo_t<F16, 3>(src, dst, dtype, g, Vhp, q, ncol, col_bytes, st);
  }
  if (prec == F64) return unpack_spinor_eo_t<F64, 1>(src, dst, dtype, g, Vhp, q, ncol, col_bytes, st);
  if (prec == F32) return unpack_spinor_eo_t<F32, 1>(src, dst, dtype, g, Vhp, q, ncol, col_bytes, st);
  return unpack_spinor_eo_t<F16, 1>(src, dst, dtype, g, Vhp, q, ncol, col_bytes, st);
}

hipError_t convert_gauge_eo(int prec, int nc, int recon, const void* src, void* dst, const Geom& g, long long Vh,
                            long long Vhp, int ncfg, size_t src_cfg_bytes, size_t src_dir_bytes, size_t cfg_bytes,
                            size_t dir_bytes, hipStream_t st) {
  dim3 grid((unsigned)((Vh + NT - 1) / NT), 8, (unsigned)ncfg), block(NT);
#define WL_CGE(P, NCV, R)                                                                                      \
  hipLaunchKernelGGL((k_convert_gauge_eo<P, NCV, R>), grid, block, 0, st, src, dst, g, Vh, Vhp, src_cfg_bytes, \
                     src_dir_bytes, cfg_bytes, dir_bytes)
  if (nc == 3) {
    if (prec == F64 && recon == 12) WL_CGE(F64, 3, 12);
    else if (prec == F64) WL_CGE(F64, 3, 18);
    else if (prec == F32 && recon == 12) WL_CGE(F32, 3, 12);
    else if (prec == F32) WL_CGE(F32, 3, 18);
    else if (prec == F16 && recon == 12) WL_CGE(F16, 3, 12);
    else WL_CGE(F16, 3, 18);
  } else {
    if (prec == F64) WL_CGE(F64, 1, 18);
    else if (prec == F32) WL_CGE(F32, 1, 18);
    else WL_CGE(F16, 1, 18);
  }
#undef WL_CGE
  return hipGetLastError();
}

}  // namespace wl
