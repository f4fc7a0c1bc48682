/*
 * wilson_amd.h -- C ABI of the MI355X (gfx950) Wilson-fermion operator and
 * normal-equation CG library (libwilson_amd.so).
 *
 * The entry points replace the operator/solver calls of the reference theory
 * object (reference paths relative to the jaxQFT tree):
 *
 *   wl_apply(WL_OP_D)      <- SU3WilsonNf2.apply_D      jaxqft/models/su3_wilson_nf2.py:1036-1041
 *                             U1WilsonNf2.apply_D       jaxqft/models/u1_wilson_nf2.py:526-540
 *   wl_apply(WL_OP_DDAG)   <- apply_Ddag                su3_wilson_nf2.py:1043-1045, u1_wilson_nf2.py:542-553
 *   wl_apply(WL_OP_DDAGD)  <- apply_normal              su3_wilson_nf2.py:1101-1105, u1_wilson_nf2.py:634-635
 *   wl_apply(WL_OP_G5D)    <- gamma5 . apply_D          su3_wilson_nf2.py:1688-1690, :2044-2050
 *   wl_apply(WL_OP_K/KDAG) <- apply_offdiag[_dagger]    su3_wilson_nf2.py:1047-1055
 *   wl_cg_normal           <- solve_normal[_with_guess] su3_wilson_nf2.py:1410-1432 (+ _solve_linear :1243-1269)
 *                             U1WilsonNf2.solve_normal  u1_wilson_nf2.py:873-883
 *   wl_solve_stats.true_resid <- normal_residual_stats  su3_wilson_nf2.py:1456-1466
 *   wl_gauge_upload        <- _links_with_fermion_bc    su3_wilson_nf2.py:751-759 (phases are applied in-kernel)
 *
 * Conventions
 *   - Spinor arrays use the reference layout (cols, L0, L1, L2, L3, 4, Nc),
 *     complex128 or complex64, C-order, last lattice axis fastest.  Column c
 *     uses gauge configuration c / nrhs.
 *   - Links use the reference layouts BMXYIJ (B, 4, L0..L3, Nc, Nc) or BXYMIJ
 *     (B, L0..L3, 4, Nc, Nc); U(1) drops the (Nc, Nc) axes.  Links are given
 *     WITHOUT the fermion boundary phase; bc[mu] is applied on the x_mu = L_mu-1
 *     slice exactly as the reference does.
 *   - D psi = (m + 4r) psi - 1/2 sum_mu [(r - g_mu) U_mu(x) psi(x+mu)
 *                                       + (r + g_mu) U_mu(x-mu)^dag psi(x-mu)].
 *   - All array pointers are device pointers unless stated; buffers are
 *     caller-owned, gauge and field handles are library-owned.
 *   - Calls are ordered on the given HIP stream (void* = hipStream_t, NULL =
 *     default stream).  Calls on one context share its device workspace: the
 *     host side holds a per-context lock for the duration of a call, and a
 *     call issued on a stream other than the previous call's first makes its
 *     stream wait for the end of that call (an event recorded at the end of
 *     every call), so calls on one context run in the order they were issued
 *     whatever streams they use.  Under graph capture nothing is waited for or
 *     recorded, and the caller orders the captured work against other calls
 *     on the context.  Separate contexts may be used concurrently.
 *   - Every function returns 0 on success or a WL_ERR_* code; wl_last_error()
 *     gives the message of the calling thread's last failure.
 */
#ifndef WILSON_AMD_H
#define WILSON_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WL_OK 0
#define WL_ERR_BAD_ARG 1
#define WL_ERR_HIP 2
#define WL_ERR_OOM 3
#define WL_ERR_UNSUPPORTED 4

/* storage / arithmetic precision of internal fields */
#define WL_F64 0
#define WL_F32 1
#define WL_F16 2 /* FP16 storage with a per-site FP32 scale, FP32 arithmetic */

/* caller array element types */
#define WL_C128 0
#define WL_C64 1

/* link layouts (su3_ym.py:195-202) */
#define WL_LAYOUT_BMXYIJ 0
#define WL_LAYOUT_BXYMIJ 1

/* operators */
#define WL_OP_D 0
#define WL_OP_DDAG 1
#define WL_OP_G5D 2
#define WL_OP_K 3
#define WL_OP_KDAG 4
#define WL_OP_DDAGD 5

/* even-odd (checkerboard) operators of wl_eo_apply; K = D - m0 (the hopping
 * term), P_e / P_o the parity projections, S_e = m0 - K_eo K_oe / m0 and
 * S_o = m0 - K_oe K_eo / m0 (normalized: divided by m0) */
#define WL_EO_K_OE 0        /* P_o K P_e          su3_wilson_nf2.py:1057-1061 */
#define WL_EO_K_EO 1        /* P_e K P_o          :1063-1067 */
#define WL_EO_KDAG_OE 2     /* P_o K^dag P_e      :1069-1073 */
#define WL_EO_KDAG_EO 3     /* P_e K^dag P_o      :1075-1079 */
#define WL_EO_SCHUR_E 4     /* S_e                :1112-1121 */
#define WL_EO_SCHUR_E_DAG 5 /* S_e^dag            :1138-1146 */
#define WL_EO_SCHUR_O 6     /* S_o                :1148-1156 */
#define WL_EO_SCHUR_O_DAG 7 /* S_o^dag (no reference counterpart; completes the set) */

/* CG stopping rule */
#define WL_STOP_TRUE_RESIDUAL 0 /* ||b - A x|| <= tol ||b||, recomputed in FP64 */
#define WL_STOP_ITERATED 1      /* reference rule: iterated residual only (jax.scipy.sparse.linalg.cg) */

/* per-column solve status */
#define WL_STATUS_RUNNING (-1)
#define WL_STATUS_CONVERGED 0
#define WL_STATUS_MAXITER 1
#define WL_STATUS_NAN 2
#define WL_STATUS_BREAKDOWN 3

typedef struct wl_ctx wl_ctx;
typedef struct wl_gauge wl_gauge;
typedef struct wl_field wl_field;

typedef struct {
  int nd;    /* must be 4 */
  int L[4];  /* lattice extents, last axis fastest */
  int nc;    /* 3 (SU(3)) or 1 (U(1)) */
  int batch; /* number of gauge configurations B */
} wl_geom;

typedef struct {
  double mass;     /* bare mass m (su3_wilson_nf2.py:595) */
  double wilson_r; /* Wilson r (must be 1: the projector kernels need rank-2 (1 +- g)) */
  double bc_re[4]; /* fermion boundary phases p_mu (su3_wilson_nf2.py:711-724) */
  double bc_im[4];
} wl_params;

typedef struct {
  double tol;       /* relative residual target (cg_tol) */
  int maxiter;      /* cg_maxiter */
  int stop_rule;    /* WL_STOP_* */
  int inner_prec;   /* WL_F64 (plain CG) / WL_F32 / WL_F16 (reliable-update mixed precision) */
  int inner_recon;  /* 18 or 12 (12-real link reconstruction) for the inner operator */
  double delta;     /* reliable-update trigger on the residual drop (mixed precision), e.g. 0.1;
                       at least 3e-3 (FP16 inner) or 1e-5 (FP32 inner) is used, and an update
                       is also forced after 100 inner iterations without one */
  int poll_every;   /* iterations queued between host polls of the convergence flags (rounded up
                       to an even number: iterations alternate two search-direction buffers) */
  int use_graph;    /* capture the poll_every-iteration body as a HIP graph */
  const double* mass; /* host array of per-column masses, or NULL for the context mass */
  int accum_prec;     /* WL_ACCUM_*: storage of the inner solution accumulator between reliable updates */
  int search_update;  /* WL_SEARCH_*: how a mixed-precision solve forms the search direction p and
                         t = D p (an FP64 solve always takes WL_SEARCH_XPAY) */
} wl_solve_params;

#define WL_ACCUM_AUTO 0 /* FP32 for a mixed-precision solve */
#define WL_ACCUM_F64 1
#define WL_ACCUM_F32 2
#define WL_ACCUM_F16 3

/* WL_SEARCH_RECURRENCE: the D launch of an iteration takes r and forms
 * t = D r + beta t_prev (= D (r + beta p_prev)), and the D^dag launch forms
 * p = r + beta p_prev beside its update of r; two launches per iteration.
 * WL_SEARCH_XPAY: p = r + beta p_prev in a pass of its own and t = D p;
 * three launches per iteration.  Both give the same iterates in exact
 * arithmetic; they differ in FP16/FP32 rounding. */
#define WL_SEARCH_RECURRENCE 0
#define WL_SEARCH_XPAY 1

typedef struct {
  int iters;            /* CG iterations performed */
  int status;           /* WL_STATUS_* */
  int reliable_updates; /* FP64 true-residual recomputations */
  int pad;
  double true_resid;    /* ||b - D^dag D x|| / ||b|| in FP64 at the last update */
  double iter_resid;    /* iterated residual / ||b|| */
} wl_solve_stats;

int wl_version(void);
const char* wl_last_error(void);

int wl_ctx_create(const wl_geom* geom, const wl_params* params, wl_ctx** out);
void wl_ctx_destroy(wl_ctx* ctx);
int wl_ctx_set_mass(wl_ctx* ctx, double mass);
/* Stencil launch tuning (no counterpart in the reference), per storage
 * precision and separately for single-RHS (multi_rhs = 0) and multi-RHS
 * (multi_rhs = 1) launches. */
typedef struct wl_tuning {
  int march;          /* x0 sites each GPU thread marches over (0 = automatic; must
                         divide L0; 1 = one site per thread) */
  int rhs_per_thread; /* right-hand sides sharing each link load in a thread (multi-RHS
                         only; 0 = automatic, 1, 2 or 4; used when it divides the
                         right-hand sides per configuration) */
  int xcd_remap;      /* 1: each XCD walks one contiguous range of tiles; 0: hardware
                         workgroup order */
  int slab;           /* one-site-per-thread kernels: 0 = flat site order; s > 0 (must
                         divide L1) = tiles walk slabs of s consecutive x1 values,
                         x0 fastest, so that x0 neighbours are reused from L2 */
  int cols_outer;     /* several spinor columns: 0 = the columns (column groups of a
                         multi-RHS kernel) of one tile run back to back (links shared in
                         L2); n = 1..64 = blocks of n columns (groups) sweep the lattice
                         in turn, the columns of a block back to back per tile, so that
                         a block's spinor neighbours fit in L2 and its links are shared
                         (1 = one column sweeps the lattice before the next); a size
                         that does not divide the columns is lowered until it does */
  int block_shape;    /* one-site-per-thread kernels: 0 = flat / slab order; k = 1..8 =
                         each workgroup covers a compact 4D box of 256 sites (preset k,
                         x3 extent >= 8 so loads stay whole cache lines), which keeps
                         about half the neighbour loads inside the workgroup; boxes are
                         ordered x3-fastest.  Presets (b0,b1,b2,b3): 1 (8,4,1,8),
                         2 (4,4,2,8), 3 (2,4,4,8), 4 (4,2,4,8), 5 (8,2,2,8),
                         6 (16,2,1,8), 7 (4,4,1,16), 8 (8,2,1,16); each b_mu must
                         divide L_mu.  Takes precedence over slab. */
  int blocks_per_cu;  /* 0 = as many resident workgroups per CU as registers allow;
                         n > 0 = at most n (by reserving LDS), which narrows the window
                         of lattice sites in flight on each XCD to what its L2 holds */
  int plane;          /* FP16, one right-hand side per thread: 0 = off; n > 0 (must divide
                         L1) = the LDS-staged plane kernel, each workgroup marching n x1
                         values over a tile of full x3 rows with its x2 and x3 neighbours
                         read from LDS (needs 2 L3 <= 256); takes precedence over march,
                         slab and block_shape */
  int eo_slab;        /* even-odd hop (wl_eo_apply, wl_cg_eo_normal): 0 = flat order of the
                         compact sites; s > 0 (must divide L1) = slabs of s consecutive x1
                         planes walked x0-fastest */
  int lds_box;        /* FP16 SU(3) with 12-real links (18-real launches keep the other
                         settings): 0 = off; k = 1..3 = the LDS box kernel, one workgroup
                         per 256-site box of one configuration walking all of its columns
                         two per pass, with the box's links loaded once and in-box
                         neighbours read from LDS; box extents (x0,x1,x2,x3) 1 (2,2,4,16),
                         2 (2,4,4,8), 3 (4,2,4,8), which must divide the lattice; takes
                         precedence over every other shape setting */
  int slab_x2;        /* slab order (slab > 0): 0 = a slab spans whole x1 planes; t > 0
                         (must divide L2) = a slab spans t consecutive x2 values, so the
                         slabs are (slab x t x L3) tiles walked x0-fastest, then by x2
                         block, then by x1 block: each XCD sweeps x0 over a compact
                         (x1, x2) block, whose x0 neighbours stay in its L2 */
} wl_tuning;
/* wl_ctx_get_tuning reports the values in effect for nrhs right-hand sides. */
int wl_ctx_set_tuning(wl_ctx* ctx, int prec, int multi_rhs, const wl_tuning* t);
int wl_ctx_get_tuning(wl_ctx* ctx, int prec, int nrhs, wl_tuning* t);
/* Sets only the eo_slab field (no reference counterpart): the even-odd hop is
 * tuned apart from the full-lattice stencil, whose settings stay as they are
 * (automatic ones included). */
int wl_ctx_set_eo_slab(wl_ctx* ctx, int prec, int multi_rhs, int eo_slab);

/* where: 0 = device pointer, 1 = host pointer.  The call is synchronous and
 * takes no stream: a device pointer is read after all device work queued so far
 * (on every stream) has finished.  The upload also measures how
 * far each SU(3) link's stored third row is from conj(row0 x row1); when every
 * link agrees to 1e-14 the FP64 operator of the reliable updates and true
 * residuals reads 12-real links (the same operator to rounding). */
int wl_gauge_upload(wl_ctx* ctx, const void* U, int layout, int dtype, int where, wl_gauge** out);
void wl_gauge_release(wl_gauge* g);
/* prebuild the (prec, recon) encoding of a gauge field (otherwise built on
 * first use).  A new encoding is built on `stream` and the call waits for
 * it, so any stream may use it afterwards; it cannot be built while
 * `stream` is capturing a graph (prepare it before the capture). */
int wl_gauge_prepare(wl_gauge* g, int prec, int recon, void* stream);

/* Operator on reference-layout arrays of ncol = batch * nrhs spinors. */
int wl_apply(wl_ctx* ctx, wl_gauge* g, int op, const void* in, void* out, int nrhs, int dtype, void* stream);

/* wl_apply without device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * operator, arguments and results as wl_apply (the same kernels), with the
 * FP64 staging buffers in a statically sized per-device scratch array of
 * WL_NOALLOC_SCRATCH_BYTES (part of the library's code object, held by each GPU
 * that loads it until the process ends) instead of the context workspace:
 * 2 (3 for WL_OP_DDAGD) buffers of batch * nrhs FP64 columns must fit (one
 * FP64 column holds 192 bytes per site plus 288 padding sites, so e.g. one
 * column at 16^4 or twelve at 8^4); larger calls return WL_ERR_BAD_ARG and
 * need wl_apply.  Every call on a GPU shares that GPU's array and it keeps
 * the last call's contents: calls that overlap on one GPU (from any context
 * or stream) race, so the caller orders them. */
#define WL_NOALLOC_SCRATCH_BYTES (64u << 20)
int wl_apply_noalloc(wl_ctx* ctx, wl_gauge* g, int op, const void* in, void* out, int nrhs, int dtype, void* stream);

/* Resident internal-layout fields: per column, structure-of-arrays planes of
 * 16-byte chunks over the padded site count (F64: double2 per complex; F32:
 * float4 per two complex; F16: eight halves per four complex, then one plane
 * of float per-site scales). */
int wl_field_create(wl_ctx* ctx, int prec, int ncol, wl_field** out);
void wl_field_destroy(wl_field* f);
int wl_field_import(wl_field* f, const void* src, int dtype, void* stream);
int wl_field_export(const wl_field* f, void* dst, int dtype, void* stream);
size_t wl_field_bytes(const wl_field* f);
/* out = op(in) with links encoded as (in->prec, recon); in and out share precision */
int wl_dslash(wl_ctx* ctx, wl_gauge* g, int op, int recon, const wl_field* in, wl_field* out, int nrhs, void* stream);

/* Solve D^dag D x = b per column.  b, x: reference-layout arrays (ncol =
 * batch * nrhs columns); x0 (nullable) is the initial guess; stats: ncol
 * entries (host memory).  Returns WL_OK even if some columns did not converge
 * -- check stats[i].status. */
int wl_cg_normal(wl_ctx* ctx, wl_gauge* g, const void* b, void* x, const void* x0, int nrhs, int dtype,
                 const wl_solve_params* sp, wl_solve_stats* stats, void* stream);
/* Same on resident FP64 fields; x is used as the initial guess when use_x0. */
int wl_cg_normal_field(wl_ctx* ctx, wl_gauge* g, const wl_field* b, wl_field* x, int use_x0, int nrhs,
                       const wl_solve_params* sp, wl_solve_stats* stats, void* stream);

/* Even-odd operators (replaces SU3WilsonNf2.apply_K_oe / apply_K_eo /
 * apply_eo_schur_even / apply_eo_schur_even_dagger / apply_eo_schur_odd,
 * su3_wilson_nf2.py:1057-1156, and the compact blocks of :934-988).  in, out:
 * full-lattice reference-layout arrays (batch * nrhs columns); the sites of the
 * source parity of in are read and out is written everywhere (zero on the
 * other parity).  normalized applies to the Schur operators.  Needs every
 * lattice extent even. */
int wl_eo_apply(wl_ctx* ctx, wl_gauge* g, int op, int normalized, const void* in, void* out, int nrhs, int dtype,
                void* stream);
/* Analytic force of the Nf=2 pseudofermion action phi^dag (D^dag D)^-1 phi
 * from X = (D^dag D)^-1 phi and Y = D X (replaces
 * SU3WilsonNf2.pseudofermion_force_from_solution, su3_wilson_nf2.py:1556-1576,
 * and the U(1) one, u1_wilson_nf2.py:1004-1025): F_mu(x) in the su(3) (u(1))
 * algebra for every link.  X, Y: (batch, *L, 4, Nc) reference-layout arrays;
 * F: (batch, 4, *L, Nc, Nc) for WL_LAYOUT_BMXYIJ or (batch, *L, 4, Nc, Nc) for
 * WL_LAYOUT_BXYMIJ (U(1): no (Nc, Nc) axes); all three of element type dtype. */
int wl_pf_force(wl_ctx* ctx, wl_gauge* g, const void* X, const void* Y, void* F, int layout, int dtype, void* stream);
/* The same force from resident FP64 fields X and Y (one column per
 * configuration, e.g. X from wl_cg_normal_field and Y = D X from wl_dslash),
 * so that a monomial force step (su3_wilson_nf2.py:378-392) keeps its
 * solution on the device in the internal layout; F as for wl_pf_force.
 * X and Y are read in FP64 whatever dtype F has: the result equals
 * wl_pf_force on the fields exported as complex128, rounded once to F's dtype. */
int wl_pf_force_field(wl_ctx* ctx, wl_gauge* g, const wl_field* X, const wl_field* Y, void* F, int layout, int dtype,
                      void* stream);
/* Solve S_e^dag S_e x_e = P_e b with the normalized even Schur complement
 * (replaces solve_eop_even_normal / solve_eop_even_normal_with_guess,
 * su3_wilson_nf2.py:1373-1392); same solver parameters and statistics as
 * wl_cg_normal, the true residual being that of the Schur normal equations.
 * b, x, x0: full-lattice arrays; even sites are read, x is zero on odd sites. */
int wl_cg_eo_normal(wl_ctx* ctx, wl_gauge* g, const void* b, void* x, const void* x0, int nrhs, int dtype,
                    const wl_solve_params* sp, wl_solve_stats* stats, void* stream);

/* Output orders of wl_lime_decode. */
#define WL_LIME_LEXI 0  /* (V, datacount, ncomplex), sites lexicographic (x0 fastest): site_data, lime.py:296-297 */
#define WL_LIME_SITE 1  /* (L0, .., L[nd-1], datacount, ncomplex), last axis fastest: lattice_data, :236-244 */
#define WL_LIME_COUNT 2 /* (datacount, L0, .., L[nd-1], ncomplex): BM...IJ links, decode_scidac_gauge :334-352 */
/* Decode the binary payload of a SciDAC field record (replaces the payload
 * decode of decode_scidac_field, jaxqft/io/lime.py:247-331).  payload: device
 * copy of the record's nbytes, big-endian IEEE reals of word_bytes 4 (SciDAC
 * precision F) or 8 (D); for each site in lexicographic order, datacount items
 * of ncomplex (re, im) values.  dims[nd] (nd 1..4): lattice extents, x0 first.
 * out: device array of complex128 (WL_C128) or complex64 (WL_C64) in the given
 * order; narrowing D to complex64 rounds to nearest.  nbytes must equal
 * V * datacount * ncomplex * 2 * word_bytes. */
int wl_lime_decode(const void* payload, size_t nbytes, int word_bytes, const int* dims, int nd, int datacount,
                   int ncomplex, int order, void* out, int dtype, void* stream);

/* Kernel timing of CG iterations (no counterpart in the reference; a
 * measurement aid).  While profiling is on, solves on the context run their
 * iterations host-driven (no graphs) with HIP events on the solve's stream
 * around the two stencil launches and the search-direction update of every
 * iteration; wl_ctx_get_profile reports the averages since profiling was
 * switched on, split by the parity of the iteration within its pair (a
 * mixed-precision solve adds to its solution accumulator every second
 * iteration, in the launches of parity 1).  Switching on (or off) resets the
 * sums.  The operator launch M (D, or the two half-lattice launches of the
 * even Schur complement) is timed with its |M p|^2 reduction, M^dag with its
 * fused CG step. */
typedef struct {
  int iterations[2];   /* iterations timed, by parity */
  double us_op[2];     /* t = M p with |t|^2 partials */
  double us_op_dag[2]; /* M^dag t with r -= alpha M^dag t, |r|^2 partials and the y update */
  double us_xpay;      /* p' = r + beta p */
  int iterations_without_update; /* of those, iterations queued without the fused (gated) update launches */
  int pad;
} wl_profile;
int wl_ctx_set_profiling(wl_ctx* ctx, int on);
int wl_ctx_get_profile(wl_ctx* ctx, wl_profile* out);

/* HBM streaming probes (no counterpart in the reference; the bandwidth the
 * stencils are compared with): 16 bytes per lane, `unroll` (1, 2, 4) loads or
 * stores in flight per lane, non-temporal when `nontemporal`; blocks = 0 covers
 * the buffer once with one workgroup per 4 KiB x unroll, else `blocks`
 * workgroups stride over it.  bytes: a multiple of 16; src and dst 16-byte
 * aligned device pointers (READ: dst receives at most one word, COPY: dst of
 * `bytes`).  Returns WL_ERR_BAD_ARG for bad arguments. */
#define WL_STREAM_READ 0
#define WL_STREAM_WRITE 1
#define WL_STREAM_COPY 2
int wl_stream_probe(int kind, int nontemporal, int unroll, const void* src, void* dst, size_t bytes, int blocks,
                    void* stream);

/* Continuous-batching solve engine (no counterpart in the reference; it serves
 * the point-source propagator loop of core/measurements.py:277-371 and request
 * streams).  A batch holds `capacity` columns on one resident configuration
 * (context batch must be 1).  wl_batch_admit packs the right-hand side b
 * (device, reference layout, one spinor) into a free column with its own mass;
 * wl_batch_step queues poll_every CG iterations for every occupied column
 * (asynchronous, one HIP graph when use_graph); wl_batch_query waits for the
 * stream and reports, per column, state (0 free, 1 running, 2 finished) and
 * stats; wl_batch_retire writes a column's solution x (device, reference
 * layout) and frees the column.  Columns are independent: a right-hand side
 * gives the same solution alone or batched.  Admitting into an occupied slot,
 * retiring a free one or naming a slot outside [0, capacity) returns
 * WL_ERR_BAD_ARG; retiring a running column returns its current iterate.
 * The batch keeps using `g`: release the gauge only after wl_batch_destroy. */
typedef struct wl_batch wl_batch;
size_t wl_batch_bytes(wl_ctx* ctx, int capacity, int inner_prec);
int wl_batch_create(wl_ctx* ctx, wl_gauge* g, int capacity, const wl_solve_params* sp, void* stream, wl_batch** out);
void wl_batch_destroy(wl_batch* b);
int wl_batch_admit(wl_batch* b, int slot, const void* rhs, int dtype, double mass);
int wl_batch_step(wl_batch* b);
int wl_batch_query(wl_batch* b, wl_solve_stats* stats, int* state);
int wl_batch_retire(wl_batch* b, int slot, void* x, int dtype);

#ifdef __cplusplus
}
#endif
#endif /* WILSON_AMD_H */
