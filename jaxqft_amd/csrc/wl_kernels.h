// Kernel argument blocks and launch wrappers shared by the host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace wl {

constexpr int NT = 256;  // threads per workgroup (4 wavefronts)

struct Geom {
  int L[4];
  int stride[4];  // flat-index strides, stride[3] = 1 (last axis fastest)
  long long V;    // sites
  long long Vp;   // plane stride of the internal layouts (V + padding)
};

// Planes are padded so that the planes of one site (and the same site in
// different fields) do not all fall in one L2 set: a power-of-two plane
// stride maps every plane of a site to the same set index.
constexpr long long PLANE_PAD = 288;

// Per-column solver scalars, FP64, resident in device memory.
struct ColState {
  double rr;        // |r|^2 of the iterated residual
  double rr_old;    // |r|^2 at the start of the current iteration
  double pAp;       // <p, A p> = |D p|^2
  double alpha, beta;
  double alpha_prev;  // alpha of the first iteration of a pair whose alpha p is not yet in y (pend)
  double b2;        // |b|^2
  double stop2;     // inner stopping threshold on rr
  double tol2b2;    // true-residual target tol^2 |b|^2
  double r0norm;    // |r| at the last reliable update
  double delta;     // reliable-update trigger (0 disables)
  double rtrue2;    // last true residual |b - A x|^2
  double m0;        // diagonal m + 4r of this column's operator
  double gap2;      // (true / iterated residual)^2 measured at the last reliable update (0: none yet);
                    // the inner stopping test takes it as at least 1
  int iter;
  int maxiter;
  int active;       // iterate this column
  int done;         // finished (see status)
  int status;       // WL_STATUS_*
  int need_update;  // 1: reliable update, 2: convergence check
  int nupdates;
  int mode;         // 0: stop on true residual, 1: stop on iterated residual (reference rule)
  int last_update;  // iteration of the last reliable update
  int pairwise;     // mixed precision: y takes alpha p of two iterations at a time
  int pend;         // pairwise: alpha_prev p of the previous iteration is still to be added
  int p_par;        // search-direction buffer (0 or 1) holding the column's latest p
};

// A reliable update is also forced after this many inner iterations without
// one (guards against an iterated residual that stalls above the trigger).
constexpr int MAX_INNER_ITERS = 100;

// Which per-column action the last workgroup of a reduction performs.
enum RedAction {
  RA_NONE = 0,
  RA_PAP = 1,     // pAp = sum (= |D p|^2 = <p, D^dag D p>); alpha = rr / pAp
  RA_ITER = 2,    // rr update, beta, iteration count, stopping / update triggers
  RA_B2 = 3,      // b2 = sum
  RA_TRUERES = 4, // rtrue2 = sum: convergence decision after a reliable update
};

struct Reduction {
  double* partials;      // [ncol][nblk_per_col]
  ColState* cols;        // [ncol]
  int action;
  int only_col;          // column of gate 3
  int par;               // RA_ITER: search-direction buffer of the iteration (pairwise y updates)
  int then_finalize;     // RA_TRUERES: close the reliable update in the same kernel (finalize_col)
};

// Stencil epilogues.  EPI_NORM and the CG steps feed the fused CG iteration
// (see wl_dslash_kernels.h); EPI_RESID is the FP64 true residual.
enum Epilogue {
  EPI_STORE = 0,  // out = result
  EPI_NORM = 1,   // out = result; partial |result|^2
  EPI_RESID = 2,  // out = w - result; partial |out|^2
  EPI_CG64 = 3,   // CG step with A p = result (not stored): y += alpha p, r -= alpha A p; partial |r|^2; y in F64
  EPI_CG32 = 4,   //   ... y in F32
  EPI_CG16 = 5,   //   ... y in F16
  EPI_RESID32 = 6,  // FP64 true residual stored in the inner precision: res = w - result in F32
                    // (column stride col_bytes_y), partial |w - result|^2 from the FP64 values
  EPI_RESID16 = 7,  //   ... in F16 with per-site scales
};

struct DslashArgs {
  Geom g;
  const void* U;          // gauge configurations, internal layout
  size_t cfg_bytes;       // bytes of one configuration (4 directions)
  size_t dir_bytes;       // bytes of one direction
  const void* in;
  void* out;
  const void* w;          // epilogue operand: b of EPI_RESID, p of the CG steps (precision of out)
  const void* w2;         // CG steps: the previous iteration's search direction (pairwise y updates)
  void* res;              // CG steps: iterated residual r (precision of out)
  void* y;                // CG steps: solution accumulator (precision given by the epilogue)
  size_t col_bytes;       // bytes of one spinor column (in, out, w and r share precision)
  size_t col_bytes_y;     // bytes of one column of y
  int ncol;
  int cols_per_cfg;
  int ntiles;             // nchunk * nseg
  int march;              // x0 sites per thread (divides L0)
  int nseg;               // L0 / march
  int xcd_remap;          // map workgroups so that each XCD walks a contiguous tile range
  int mrhs;               // right-hand sides per thread (1, 2 or 4; > 1 implies one site per thread)
  int slab;               // one site per thread: 0 flat order, else x1 extent of a slab (x0-fastest slabs)
  int slab_tiles;         // tiles per slab
  int slab2;              // x2 extent of a slab (divides L2; L2 = whole x1 planes)
  int box;                // 1: workgroups cover compact 4D boxes (bsh, bn below); overrides slab
  int bsh[4];             // log2 of the box extent per direction
  int bn[4];              // boxes per direction (L_mu >> bsh_mu)
  int lbox;               // 1: LDS box kernel (F16 SU(3)): a workgroup owns one 256-site box (bsh, bn)
                          //    of one configuration and walks all of its columns, two per pass
  int lregions;           // LDS box kernel: 1 when the box grid is split into 8 XCD regions (lsplit)
  int lsplit[4];          // regions per direction (product 8), each region's boxes run on one XCD
  int cols_outer;         // column order: 0 columns (groups) fastest; n >= 1 blocks of n sweeping the lattice in turn
  int plane;              // 1: LDS-staged plane kernel (F16, one column per thread, marching x1; march = x1 sites)
  int prow;               // plane kernel: x2 rows per workgroup tile (prow * L3 <= NT, prow divides L2)
  int lds_bytes;          // dynamic LDS reserved per workgroup (caps resident workgroups; unused by the code)
  double bc_re[4], bc_im[4];
  double m0;              // diagonal coefficient (used when m0_from_state == 0)
  double r;               // Wilson r
  int m0_from_state;      // take m0 per column from red.cols[col].m0
  int gate;               // 0 all, 1 active columns, 2 columns awaiting a reliable update, 3 column red.only_col
  int search_rec;         // CG epilogues with the search direction by recurrence: EPI_NORM forms
                          // out = result + beta out_prev (in = r), the CG steps form p = r + beta p_prev
                          // in out (in place) beside r -= alpha A p; beta = the column's beta
  Reduction red;
  // even-odd (checkerboard) kernels: compact fields of one parity hold the
  // sites of that parity in ascending full-lattice order (compact index =
  // full index / 2 when L3 is even); links are stored per parity the same way.
  int pt;                 // parity of the target sites
  int normalized;         // Schur combine: 1 - K K / m0^2 (1) or m0 - K K / m0 (0)
  long long Vh;           // sites of one parity
  const void* ctr;        // Schur combine: the field the Schur operator acts on (target parity)
};

// Full-lattice index of compact index i of parity q (L3 even; see the
// compact layout above).
__host__ __device__ inline size_t eo_full_index(const Geom& g, int q, size_t i) {
  const size_t L3h = (size_t)(g.L[3] >> 1);
  const size_t row = i / L3h;
  const size_t j = i - row * L3h;
  const size_t x2 = row % (size_t)g.L[2];
  const size_t r2 = row / (size_t)g.L[2];
  const size_t x1 = r2 % (size_t)g.L[1];
  const size_t x0 = r2 / (size_t)g.L[1];
  return row * (size_t)g.L[3] + 2 * j + ((x0 + x1 + x2 + (size_t)q) & 1);
}

// Analytic pseudofermion force (wl_force.hip): reference-layout X, Y and F,
// FP64 internal links.
struct ForceArgs {
  Geom g;
  const void* U;     // FP64 links, internal layout
  size_t cfg_bytes;  // bytes of one configuration
  size_t dir_bytes;  // bytes of one direction
  const void* X;     // (B, *L, 4, NC) complex
  const void* Y;
  void* F;           // (B, 4, *L, NC, NC) (layout 0) or (B, *L, 4, NC, NC) (layout 1); U(1): no (NC, NC)
  int layout;
  size_t col_bytes;  // X, Y as resident FP64 fields (internal layout, one column per configuration):
                     // bytes of one column; 0: X, Y in the reference layout
  double bc_re[4], bc_im[4];
  double r;
};

// SciDAC LIME payload decode (wl_lime.hip).
struct LimeArgs {
  const void* in;  // big-endian payload, sites lexicographic
  void* out;       // complex output in the chosen order
  int L[4];        // lattice padded to four axes
  long long V;
  int Wc;          // complex values per site record (dc * item)
  int item;        // complex values per item
  int dc;          // items per site
  int order;       // 0 LEXI, 1 SITE, 2 COUNT
  int T0, T3;      // tile extents along x0 and x3
  int nt0, nt3;    // tiles along x0 and x3
};

struct BlasArgs {
  long long V;
  long long Vp;
  int ncol;
  int ntiles;
  int gate;  // 0 none, 1 active columns, 2 columns awaiting a reliable update, 3 column red.only_col
  size_t col_bytes_a, col_bytes_b, col_bytes_c, col_bytes_d;
  void* a;
  void* b;
  void* c;
  void* d;
  Reduction red;
};

}  // namespace wl
