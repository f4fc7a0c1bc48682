// Device-side building blocks for the Wilson-fermion kernels (gfx950, wave64).
//
// Internal field layouts in HBM (all structure-of-arrays, 16-byte "chunks" so
// that a wavefront's load of one plane is one contiguous 1 KiB run):
//
//   spinor column, Nc=3 (24 reals/site), Nc=1 (8 reals/site):
//     F64: 12 (Nc=3) / 4 (Nc=1) planes of double2, plane c = spin*Nc + colour
//     F32:  6 / 2 planes of float4, plane j holds complex 2j, 2j+1
//     F16:  3 / 1 planes of 8 halves (complex 4j..4j+3) followed by one plane
//           of float per-site scales; value = half * scale
//   link field, per direction mu (planes of V sites each):
//     F64: 9 double2 (full) or 6 double2 (rows 0,1: 12-real reconstruction)
//     F32: 9 float2  (full) or 3 float4 (12-real)
//     F16: 9 half2   (full) or 1 uint4 + 1 uint2 (12-real)
//     Nc=1: one complex per link (double2 / float2 / half2)
//   Links are stored without the fermion boundary phase; the kernels apply
//   p_mu on the hop that crosses x_mu = L_mu - 1 (and conj(p_mu) backwards),
//   which is where the reference bakes the phase into the link
//   (su3_wilson_nf2.py:726-759).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace wl {

enum Prec { F64 = 0, F32 = 1, F16 = 2 };

template <class T>
struct cx {
  T re, im;
};

template <class T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// conj(a) * b
template <class T>
__device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {
  return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
// i * a
template <class T>
__device__ __forceinline__ cx<T> ci(cx<T> a) { return {-a.im, a.re}; }
template <class T>
__device__ __forceinline__ cx<T> cscale(T s, cx<T> a) { return {s * a.re, s * a.im}; }
template <class T>
__device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.re, -a.im}; }

// ---------------------------------------------------------------------------
// precision traits
// ---------------------------------------------------------------------------
template <int P>
struct PT;
template <>
struct PT<F64> {
  using acc = double;
};
template <>
struct PT<F32> {
  using acc = float;
};
template <>
struct PT<F16> {
  using acc = float;
};

// bytes of one spinor column (planes + scales) for V sites
__host__ __device__ inline size_t spinor_col_bytes(int prec, int nc, size_t V) {
  size_t reals = (nc == 3) ? 24 : 8;
  if (prec == F64) return reals * 8 * V;
  if (prec == F32) return reals * 4 * V;
  return reals * 2 * V + 4 * V;
}

// bytes of one direction of one gauge configuration
__host__ __device__ inline size_t link_dir_bytes(int prec, int nc, int recon, size_t V) {
  size_t reals = (nc == 3) ? (recon == 12 ? 12 : 18) : 2;
  size_t b = (prec == F64) ? 8 : (prec == F32 ? 4 : 2);
  return reals * b * V;
}

// ---------------------------------------------------------------------------
// Global accesses, plain (S = false) or streaming (S = true: non-temporal
// hint, for vectors read or written once per kernel)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 nt_ld(const double2* p) {
  return make_double2(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
}
__device__ __forceinline__ float4 nt_ld(const float4* p) {
  return make_float4(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y),
                     __builtin_nontemporal_load(&p->z), __builtin_nontemporal_load(&p->w));
}
__device__ __forceinline__ uint4 nt_ld(const uint4* p) {
  return make_uint4(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y),
                    __builtin_nontemporal_load(&p->z), __builtin_nontemporal_load(&p->w));
}
__device__ __forceinline__ float nt_ld(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void nt_st(double2* p, double2 v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
}
__device__ __forceinline__ void nt_st(float4* p, float4 v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
  __builtin_nontemporal_store(v.z, &p->z);
  __builtin_nontemporal_store(v.w, &p->w);
}
__device__ __forceinline__ void nt_st(uint4* p, uint4 v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
  __builtin_nontemporal_store(v.z, &p->z);
  __builtin_nontemporal_store(v.w, &p->w);
}
__device__ __forceinline__ void nt_st(float* p, float v) { __builtin_nontemporal_store(v, p); }

template <bool S, class T>
__device__ __forceinline__ T gld(const T* p) {
  if constexpr (S) return nt_ld(p);
  else return *p;
}

template <bool S, class T>
__device__ __forceinline__ void gst(T* p, T v) {
  if constexpr (S) nt_st(p, v);
  else *p = v;
}

// ---------------------------------------------------------------------------
// spinor load/store, NS=4 spins x NC colours, index c = s*NC + col
// (S: streaming accesses, see gld / gst)
// ---------------------------------------------------------------------------
template <int P, int NC>
struct SpinorIO;

template <int NC>
struct SpinorIO<F64, NC> {
  static constexpr int N = 4 * NC;
  template <bool S = false>
  __device__ static __forceinline__ void load(const void* __restrict__ base, size_t V, size_t site,
                                              cx<double> (&v)[N]) {
    const double2* p = reinterpret_cast<const double2*>(base);
#pragma unroll
    for (int c = 0; c < N; ++c) {
      double2 t = gld<S>(p + c * V + site);
      v[c] = {t.x, t.y};
    }
  }
  template <bool S = false>
  __device__ static __forceinline__ void store(void* __restrict__ base, size_t V, size_t site,
                                               const cx<double> (&v)[N]) {
    double2* p = reinterpret_cast<double2*>(base);
#pragma unroll
    for (int c = 0; c < N; ++c) gst<S>(p + c * V + site, make_double2(v[c].re, v[c].im));
  }
};

template <int NC>
struct SpinorIO<F32, NC> {
  static constexpr int N = 4 * NC;
  template <bool S = false>
  __device__ static __forceinline__ void load(const void* __restrict__ base, size_t V, size_t site,
                                              cx<float> (&v)[N]) {
    const float4* p = reinterpret_cast<const float4*>(base);
#pragma unroll
    for (int j = 0; j < N / 2; ++j) {
      float4 t = gld<S>(p + j * V + site);
      v[2 * j] = {t.x, t.y};
      v[2 * j + 1] = {t.z, t.w};
    }
  }
  template <bool S = false>
  __device__ static __forceinline__ void store(void* __restrict__ base, size_t V, size_t site,
                                               const cx<float> (&v)[N]) {
    float4* p = reinterpret_cast<float4*>(base);
#pragma unroll
    for (int j = 0; j < N / 2; ++j)
      gst<S>(p + j * V + site, make_float4(v[2 * j].re, v[2 * j].im, v[2 * j + 1].re, v[2 * j + 1].im));
  }
};

__device__ __forceinline__ float2 h2f(uint32_t u) {
  __half2 h = *reinterpret_cast<__half2*>(&u);
  return __half22float2(h);
}
__device__ __forceinline__ uint32_t f2h(float a, float b) {
  __half2 h = __floats2half2_rn(a, b);
  return *reinterpret_cast<uint32_t*>(&h);
}

template <int NC>
struct SpinorIO<F16, NC> {
  static constexpr int N = 4 * NC;
  static constexpr int PLANES = N / 4;
  template <bool S = false>
  __device__ static __forceinline__ void load(const void* __restrict__ base, size_t V, size_t site,
                                              cx<float> (&v)[N]) {
    const uint4* p = reinterpret_cast<const uint4*>(base);
    const float* sc = reinterpret_cast<const float*>(p + PLANES * V);
    float s = gld<S>(sc + site);
#pragma unroll
    for (int j = 0; j < PLANES; ++j) {
      uint4 t = gld<S>(p + j * V + site);
      float2 a = h2f(t.x), b = h2f(t.y), c = h2f(t.z), d = h2f(t.w);
      v[4 * j + 0] = {a.x * s, a.y * s};
      v[4 * j + 1] = {b.x * s, b.y * s};
      v[4 * j + 2] = {c.x * s, c.y * s};
      v[4 * j + 3] = {d.x * s, d.y * s};
    }
  }
  template <bool S = false>
  __device__ static __forceinline__ void store(void* __restrict__ base, size_t V, size_t site,
                                               const cx<float> (&v)[N]) {
    uint4* p = reinterpret_cast<uint4*>(base);
    float* sc = reinterpret_cast<float*>(p + PLANES * V);
    float m = 0.f;
#pragma unroll
    for (int c = 0; c < N; ++c) m = fmaxf(m, fmaxf(fabsf(v[c].re), fabsf(v[c].im)));
    float inv = (m > 0.f) ? 1.0f / m : 0.f;
#pragma unroll
    for (int j = 0; j < PLANES; ++j) {
      uint4 t;
      t.x = f2h(v[4 * j + 0].re * inv, v[4 * j + 0].im * inv);
      t.y = f2h(v[4 * j + 1].re * inv, v[4 * j + 1].im * inv);
      t.z = f2h(v[4 * j + 2].re * inv, v[4 * j + 2].im * inv);
      t.w = f2h(v[4 * j + 3].re * inv, v[4 * j + 3].im * inv);
      gst<S>(p + j * V + site, t);
    }
    gst<S>(sc + site, m);
  }
};

// ---------------------------------------------------------------------------
// link load: U[a][b], row a, column b.  RECON = 18 (full) or 12 (rows 0,1).
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void reconstruct_row2(cx<T> (&u)[3][3]) {
  // row2 = conj(row0 x row1): valid for det U = 1
  cx<T> a = csub(cmul(u[0][1], u[1][2]), cmul(u[0][2], u[1][1]));
  cx<T> b = csub(cmul(u[0][2], u[1][0]), cmul(u[0][0], u[1][2]));
  cx<T> c = csub(cmul(u[0][0], u[1][1]), cmul(u[0][1], u[1][0]));
  u[2][0] = cconj(a);
  u[2][1] = cconj(b);
  u[2][2] = cconj(c);
}

template <int P, int NC, int RECON>
struct LinkIO;

template <int RECON>
struct LinkIO<F64, 3, RECON> {
  __device__ static __forceinline__ void load(const void* __restrict__ base, size_t V, size_t site,
                                              cx<double> (&u)[3][3]) {
    const double2* p = reinterpret_cast<const double2*>(base);
    constexpr int ROWS = (RECON == 12) ? 2 : 3;
#pragma unroll
    for (int a = 0; a < ROWS; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        double2 t = p[(a * 3 + b) * V + site];
        u[a][b] = {t.x, t.y};
      }
    if (RECON == 12) reconstruct_row2(u);
  }
};

template <int RECON>
struct LinkIO<F32, 3, RECON> {
  __device__ static __forceinline__ void load(const void* __restrict__ base, size_t V, size_t site,
                                              cx<float> (&u)[3][3]) {
    if (RECON == 12) {
      const float4* p = reinterpret_cast<const float4*>(base);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        float4 t = p[j * V + site];
        int c0 = 2 * j, c1 = 2 * j + 1;
        u[c0 / 3][c0 % 3] = {t.x, t.y};
        u[c1 / 3][c1 % 3] = {t.z, t.w};
      }
      reconstruct_row2(u);
    } else {
      const float2* p = reinterpret_cast<const float2*>(base);
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        float2 t = p[c * V + site];
        u[c / 3][c % 3] = {t.x, t.y};
      }
    }
  }
};

template <int RECON>
struct LinkIO<F16, 3, RECON> {
  __device__ static __forceinline__ void load(const void* __restrict__ base, size_t V, size_t site,
                                              cx<float> (&u)[3][3]) {
    if (RECON == 12) {
      const uint4* pa = reinterpret_cast<const uint4*>(base);
      const uint2* pb = reinterpret_cast<const uint2*>(pa + V);
      uint4 t = pa[site];
      uint2 s = pb[site];
      float2 c0 = h2f(t.x), c1 = h2f(t.y), c2 = h2f(t.z), c3 = h2f(t.w), c4 = h2f(s.x), c5 = h2f(s.y);
      u[0][0] = {c0.x, c0.y};
      u[0][1] = {c1.x, c1.y};
      u[0][2] = {c2.x, c2.y};
      u[1][0] = {c3.x, c3.y};
      u[1][1] = {c4.x, c4.y};
      u[1][2] = {c5.x, c5.y};
      reconstruct_row2(u);
    } else {
      const uint32_t* p = reinterpret_cast<const uint32_t*>(base);
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        float2 t = h2f(p[c * V + site]);
        u[c / 3][c % 3] = {t.x, t.y};
      }
    }
  }
};

template <int RECON>
struct LinkIO<F64, 1, RECON> {
  __device__ static __forceinline__ void load(const void* __restrict__ base, size_t V, size_t site,
                                              cx<double> (&u)[1][1]) {
    double2 t = reinterpret_cast<const double2*>(base)[site];
    u[0][0] = {t.x, t.y};
  }
};
template <int RECON>
struct LinkIO<F32, 1, RECON> {
  __device__ static __forceinline__ void load(const void* __restrict__ base, size_t V, size_t site,
                                              cx<float> (&u)[1][1]) {
    float2 t = reinterpret_cast<const float2*>(base)[site];
    u[0][0] = {t.x, t.y};
  }
};
template <int RECON>
struct LinkIO<F16, 1, RECON> {
  __device__ static __forceinline__ void load(const void* __restrict__ base, size_t V, size_t site,
                                              cx<float> (&u)[1][1]) {
    float2 t = h2f(reinterpret_cast<const uint32_t*>(base)[site]);
    u[0][0] = {t.x, t.y};
  }
};

// ---------------------------------------------------------------------------
// wavefront / block reductions (wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over a block of NT threads; result valid in thread 0.  Fixed order.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) s += lds[i];
  }
  __syncthreads();
  return s;
}

// Map the hardware workgroup id to a logical tile so that each XCD (workgroups
// are dispatched round-robin over the 8 XCDs) walks one contiguous range of the
// lattice; neighbour planes then stay in that XCD's L2.  Bijective for any n.
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned n, int remap = 1) {
  if (!remap) return b;
  const unsigned X = 8;
  unsigned per = n / X, rem = n % X;
  unsigned xcd = b % X, idx = b / X;
  return (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// Split a logical workgroup index over (column unit, tile) for n units and
// ntiles tiles.  cb = 0: units fastest (the units of one tile run back to
// back); cb >= 1 (dividing n): blocks of cb units, each block sweeping every
// tile with its units fastest, blocks slowest (cb = 1: units slowest).
__device__ __forceinline__ void split_units(unsigned logical, int ntiles, int n, int cb, int& unit, int& tile) {
  if (cb <= 0) {
    unit = (int)(logical % (unsigned)n);
    tile = (int)(logical / (unsigned)n);
    return;
  }
  const unsigned per = (unsigned)ntiles * (unsigned)cb;
  const unsigned blk = logical / per;
  const unsigned r = logical - blk * per;
  unit = (int)(blk * (unsigned)cb + r % (unsigned)cb);
  tile = (int)(r / (unsigned)cb);
}

}  // namespace wl
