// Decode of a SciDAC LIME binary field record on the device
// (jaxqft/io/lime.py:247-331 with the site reorder of :236-244 and the
// direction-axis move of decode_scidac_gauge :334-352).
//
// The payload holds, for every site in lexicographic order (x0 fastest, then
// x1, x2, x3), `datacount` items of `ncomplex` complex values, each value two
// big-endian IEEE reals (re, im) of 4 (precision F) or 8 (precision D) bytes.
// A site's record is therefore Wc = datacount * ncomplex contiguous complex
// values.  The output is complex64 or complex128 in one of three orders:
//   LEXI   (V, datacount, ncomplex)        sites lexicographic, as stored
//   SITE   (L0, .., L3, datacount, ncomplex) C order, last lattice axis fastest
//   COUNT  (datacount, L0, .., L3, ncomplex) direction-major links (BM...IJ)
// SITE and COUNT reverse the lattice axes.  A workgroup owns a T0 x T3 tile of
// the (x0, x3) plane at one (x1, x2): it reads T3 contiguous runs of T0 site
// records (the payload's fast axis), byte-swaps them into LDS, and writes T0
// (SITE) or datacount * T0 (COUNT) contiguous runs along x3 (the output's fast
// axis), so both the HBM reads and the HBM writes are whole contiguous runs.
// Lower-dimensional lattices are padded to four axes by the host so that both
// orders stay the same (nd=1: (L0,1,1,1); nd=2: (L0,1,1,L1); nd=3: (L0,L1,1,L2)).
#include <hip/hip_runtime.h>

#include "wl_kernels.h"

namespace wl {

template <int WB>
struct BigEndian;

template <>
struct BigEndian<4> {
  using Raw = uint2;     // one complex value as stored
  using Val = float2;    // decoded
  static __device__ __forceinline__ Val decode(Raw r) {
    return make_float2(__uint_as_float(__builtin_bswap32(r.x)), __uint_as_float(__builtin_bswap32(r.y)));
  }
};

template <>
struct BigEndian<8> {
  using Raw = ulonglong2;
  using Val = double2;
  static __device__ __forceinline__ Val decode(Raw r) {
    return make_double2(__longlong_as_double((long long)__builtin_bswap64(r.x)),
                        __longlong_as_double((long long)__builtin_bswap64(r.y)));
  }
};

template <class O>
struct OutComplex;
template <>
struct OutComplex<float> {
  using T = float2;
  template <class V>
  static __device__ __forceinline__ T make(V v) { return make_float2((float)v.x, (float)v.y); }
};
template <>
struct OutComplex<double> {
  using T = double2;
  template <class V>
  static __device__ __forceinline__ T make(V v) { return make_double2((double)v.x, (double)v.y); }
};

#ifndef WL_LIME_THREADS
#define WL_LIME_THREADS 256
#endif
constexpr int kLimeThreads = WL_LIME_THREADS;

template <int WB, class O>
__global__ __launch_bounds__(kLimeThreads) void k_lime_decode(LimeArgs a) {
  using BE = BigEndian<WB>;
  using Raw = typename BE::Raw;
  using Val = typename BE::Val;
  using Out = typename OutComplex<O>::T;
  extern __shared__ __align__(16) unsigned char lime_lds_raw[];
  Val* lds = reinterpret_cast<Val*>(lime_lds_raw);

  const long long tiles_per_mid = (long long)a.nt0 * a.nt3;
  const long long b = blockIdx.x;
  const int mid = (int)(b / tiles_per_mid);
  const int tile = (int)(b - (long long)mid * tiles_per_mid);
  const int x1 = mid % a.L[1];
  const int x2 = mid / a.L[1];
  const int x0a = (tile % a.nt0) * a.T0;
  const int x3a = (tile / a.nt0) * a.T3;
  const int n0 = min(a.T0, a.L[0] - x0a);
  const int n3 = min(a.T3, a.L[3] - x3a);
  const int Wc = a.Wc;
  const Raw* __restrict__ in = reinterpret_cast<const Raw*>(a.in);
  Out* __restrict__ out = reinterpret_cast<Out*>(a.out);

  // lexicographic index of (x0a, x1, x2, x3a + t3) and C-order index of (x0a + t0, x1, x2, x3a)
  const long long lex0 = x0a + (long long)a.L[0] * (x1 + (long long)a.L[1] * (x2 + (long long)a.L[2] * x3a));
  const long long lex_step = (long long)a.L[0] * a.L[1] * a.L[2];
  const long long c0 = (((long long)x0a * a.L[1] + x1) * a.L[2] + x2) * a.L[3] + x3a;
  const long long c_step = (long long)a.L[1] * a.L[2] * a.L[3];

  // LDS holds the tile as [t3][t0][Wc]: each x3 row is one contiguous payload run
  const int row_in = n0 * Wc;
  for (int i = threadIdx.x; i < n3 * row_in; i += kLimeThreads) {
    const int t3 = i / row_in;
    const int w = i - t3 * row_in;
    lds[t3 * a.T0 * Wc + w] = BE::decode(in[(lex0 + t3 * lex_step) * Wc + w]);
  }
  __syncthreads();

  if (a.order == 0) {  // LEXI: the same runs, decoded
    for (int i = threadIdx.x; i < n3 * row_in; i += kLimeThreads) {
      const int t3 = i / row_in;
      const int w = i - t3 * row_in;
      out[(lex0 + t3 * lex_step) * Wc + w] = OutComplex<O>::make(lds[t3 * a.T0 * Wc + w]);
    }
  } else if (a.order == 1) {  // SITE: one run of n3 site records per x0
    const int row_out = n3 * Wc;
    for (int i = threadIdx.x; i < n0 * row_out; i += kLimeThreads) {
      const int t0 = i / row_out;
      const int r = i - t0 * row_out;
      const int t3 = r / Wc;
      const int w = r - t3 * Wc;
      out[(c0 + t0 * c_step) * Wc + r] = OutComplex<O>::make(lds[(t3 * a.T0 + t0) * Wc + w]);
    }
  } else {  // COUNT: one run of n3 items per (item index d, x0)
    const int it = a.item;
    const int row_out = n3 * it;
    for (int i = threadIdx.x; i < a.dc * n0 * row_out; i += kLimeThreads) {
      const int row = i / row_out;
      const int r = i - row * row_out;
      const int d = row / n0;
      const int t0 = row - d * n0;
      const int t3 = r / it;
      const int k = r - t3 * it;
      out[((long long)d * a.V + c0 + t0 * c_step) * it + r] =
          OutComplex<O>::make(lds[(t3 * a.T0 + t0) * Wc + d * it + k]);
    }
  }
}

hipError_t launch_lime_decode(int word_bytes, int out_double, const LimeArgs& a, size_t lds_bytes, hipStream_t st) {
  const long long nblk = (long long)a.nt0 * a.nt3 * a.L[1] * a.L[2];
  dim3 grid((unsigned)nblk), block(kLimeThreads);
  if (word_bytes == 8) {
    if (out_double)
      hipLaunchKernelGGL((k_lime_decode<8, double>), grid, block, lds_bytes, st, a);
    else
      hipLaunchKernelGGL((k_lime_decode<8, float>), grid, block, lds_bytes, st, a);
  } else {
    if (out_double)
      hipLaunchKernelGGL((k_lime_decode<4, double>), grid, block, lds_bytes, st, a);
    else
      hipLaunchKernelGGL((k_lime_decode<4, float>), grid, block, lds_bytes, st, a);
  }
  return hipGetLastError();
}

}  // namespace wl
