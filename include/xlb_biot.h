/*
 * xlb_biot.h -- C ABI of the MI355X (gfx950) lattice-Boltzmann elasticity / Darcy /
 * Biot kernels.
 *
 * Every entry point replaces one jitted method of the reference solvers
 * (s-lunowa/xlb-biot); the replaced method is cited in the comment above it.
 * The Python solver classes in xlb-biot_amd/src bind these symbols with ctypes.
 *
 * Conventions
 *  - Plain device pointers; no allocation inside any call.  Buffers that a call
 *    needs beyond its inputs and outputs (norm partials) are passed in, sized by
 *    the matching *_count() query.
 *  - Populations are structure-of-arrays: f[d][x][y], y fastest, shape (q, nx, ny).
 *    Vector fields are (2, nx, ny), stresses (3, nx, ny), scalars (nx, ny).
 *    The reference's (nx, ny, c) arrays are the permuted views of these.
 *  - dtype: XB_F64 or XB_F32; all operands of one call share it (the reference's
 *    "f64/f64" and "f32/f32" policies).  Norms are always accumulated in FP64.
 *    XB_F64_S32 is the reference's "f64/f32" (elasticity.py:606-629: FP64
 *    arithmetic, FP32 storage): FP64 buffers and arithmetic, and every population
 *    value rounded to FP32 where the reference stores it in its FP32 arrays (the
 *    collision's cast_to_output, each boundary-rule stage, the sweep's output, the
 *    prolongation's sum).  XB_F64_S16 and XB_F32_S16 are "f64/f16" and "f32/f16"
 *    the same way: FP64 / FP32 buffers and arithmetic, every stored population value
 *    rounded to FP16 (one correctly rounded conversion, as the reference's astype).
 *    ("f16/f16", FP16 arithmetic, has no dtype here.)  Only xb_elast_step,
 *    xb_elast_partials_count, xb_darcy_step, xb_elast_prolong_add and
 *    xb_elast_prolong_add_div take these three; every other entry point returns
 *    XB_EUNSUPPORTED for them.
 *  - Every call takes the HIP stream to launch on (hipStream_t as void*) and
 *    returns an int status: XB_OK, or a negative XB_E* code; the message of the
 *    last failure on the calling thread is returned by xb_last_error().
 *  - The domain is periodic (the reference streams with jnp.roll); walls are
 *    the per-cell rule tables (xb_bc_table) compiled on the host from the
 *    boundary-condition objects.
 */
#ifndef XLB_BIOT_H
#define XLB_BIOT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XB_ABI_VERSION 1

enum { XB_F64 = 0, XB_F32 = 1, XB_F64_S32 = 2, XB_F64_S16 = 3, XB_F32_S16 = 4 };

enum {
  XB_OK = 0,
  XB_EINVAL = -1,       /* bad argument (null pointer, shape, box, code) */
  XB_EUNSUPPORTED = -2, /* dtype or mode not supported */
  XB_EHIP = -3,         /* HIP runtime error on launch */
  XB_ENONFINITE = -4    /* non-finite value found by a checked reduction */
};

/* Sweep modes of xb_elast_step. */
enum {
  XB_MODE_STEP = 0,     /* f_out = S(f_in)                                     */
  XB_MODE_SMOOTH = 1,   /* f_out = (1-gamma) f_in + gamma S(f_in - r)          */
  XB_MODE_RESIDUAL = 2  /* coarse_out = restrict(f_in - S(f_in - r)) on the
                           interior box, norm_partials = per-wave sums of squares */
};

/* Rule codes of xb_bc_table (per slot and direction d):
 *   0           keep the default value (collision output / streamed value)
 *   k in 1..q   +src[k-1]      (src = pre-collision populations of the cell for
 *   -k          -src[k-1]       the post-collision stage, post-collision
 *                               populations of the cell for the post-streaming stage)
 *   XB_RULE_ZERO  0
 * and the constant const_*[slot*q + d] is added afterwards. */
#define XB_RULE_ZERO 100

typedef struct xb_bc_table {
  int32_t nslots;           /* boundary cells carrying a rule                       */
  const int32_t *cell_slot; /* (nx*ny) slot of each cell or -1                      */
  const int8_t *code_pc;    /* (nslots*q) post-collision rules (may be null)        */
  const int8_t *code_ps;    /* (nslots*q) post-streaming rules (may be null)        */
  const void *const_pc;     /* (nslots*q) dtype                                     */
  const void *const_ps;     /* (nslots*q) dtype                                     */
} xb_bc_table;

typedef struct xb_elast_args {
  int32_t nx, ny;           /* grid of this level                                   */
  int32_t dtype;
  int32_t mode;             /* XB_MODE_*                                            */
  double lam, mu;           /* Lame parameters (lattice units)                      */
  double gamma;             /* damping of XB_MODE_SMOOTH                            */
  const void *f_in;         /* (8, nx, ny)                                          */
  void *f_out;              /* STEP/SMOOTH: (8, nx, ny); RESIDUAL: unused           */
  void *fpc_out;            /* STEP only, optional: post-collision populations      */
  const void *force_a;      /* optional (2, nx, ny): body force (finest level)      */
  const void *force_b;      /* optional (2, nx, ny): external force                 */
  const void *resid;        /* optional (8, ix1-ix0, iy1-iy0): subtracted on box    */
  int32_t ix0, ix1, iy0, iy1; /* interior box (cells without multigrid BC layers) */
  const xb_bc_table *bc;    /* optional                                             */
  void *coarse_out;         /* RESIDUAL: (8, (ix1-ix0)/2, (iy1-iy0)/2) restriction out;
                               xb_elast_smoothn (SMOOTH), optional: the next level's
                               (8, nx-(ix1-ix0)/2, ny-(iy1-iy0)/2) correction, read   */
  double *norm_partials;    /* RESIDUAL: xb_elast_partials_count() doubles          */
  int32_t rows;             /* launch shape: rows per wavefront (power of 2 in
                               [1, 64]; >= 2 for RESIDUAL) or 0 for the built-in    */
  int32_t pad_;
  /* Fused Biot coupling force (all optional; active when div_in is set):
   * force_b is replaced by force_b + G(div_in), G(q) = sum_d qs_d (-3 nl_scale w_d c_d)
   * over the D2Q9 stencil, qs_d(x) = q(x - c_d) or the Darcy wall rule of the cell
   * (grad_bc); STEP and SMOOTH also write div_out = div(f_out). */
  const void *div_in;       /* (nx, ny) divergence of f_in                          */
  void *div_out;            /* (nx, ny) divergence of f_out (STEP / SMOOTH)         */
  double nl_scale;
  const xb_bc_table *grad_bc;
  /* Optional per-cell Lame parameters (nx, ny) each, both or neither; they replace
   * lam, mu (also in the divergence carried by the fused coupling force). */
  const void *lam_field;
  const void *mu_field;
} xb_elast_args;

typedef struct xb_darcy_args {
  int32_t nx, ny;
  int32_t dtype;
  int32_t pad_;
  double omega;             /* scalar relaxation rate (used if omega_field null)    */
  const void *omega_field;  /* optional (nx, ny) per-cell relaxation rate           */
  const void *f_in;         /* (9, nx, ny)                                          */
  void *f_out;              /* (9, nx, ny)                                          */
  void *fpc_out;            /* optional post-collision populations                  */
  const void *source_a;     /* optional (nx, ny): static source                     */
  const void *source_b;     /* optional (nx, ny): external source                   */
  const xb_bc_table *bc;    /* optional                                             */
  int32_t rows;             /* launch shape, as in xb_elast_args                    */
  int32_t pad2_;
} xb_darcy_args;

/* ---- library ---------------------------------------------------------------- */
int xb_abi_version(void);
const char *xb_last_error(void);

/* ---- elasticity (D2Q8) --------------------------------------------------------- */

/* One fused collide -> post-collision BCs -> stream -> post-streaming BCs pass.
 * STEP replaces LBMElasticity._step (reference src/elasticity.py:1169-1219, with
 * collision :959-1000, apply_bc :1060-1091, streaming :738-803); SMOOTH is one
 * damped multigrid smoothing step (:1258-1267, :1296-1305 with
 * substract_residual :1314-1344); RESIDUAL is the residual, its 2x2 restriction
 * and the squared norm of the V-cycle (:1278-1287, _restrict :850-876, the norm
 * at :1307-1311). */
int xb_elast_step(const xb_elast_args *args, void *stream);

/* Two plain steps in one pass: f_out = S(S(f_in)) with S the XB_MODE_STEP step of these
 * args (populations read and written once per two steps; bitwise the same as two
 * xb_elast_step calls).  STEP mode only: no fpc_out, resid or div_in; grid >= 2 x 2. */
int xb_elast_step2(const xb_elast_args *args, void *stream);

/* `steps` (1..8) consecutive plain steps in one launch on 32 x 32 LDS tiles
 * (temporal blocking; bit-identical to `steps` XB_MODE_STEP calls).  The same
 * restrictions as xb_elast_step2.  Here args->rows selects the tile shape: 256 or
 * 512 threads per 32 x 32 tile, 704 threads per 44 x 48 tile; any other value lets the
 * library choose (32 x 32 tiles, 512 threads for grids of at most four tiles per
 * CU, else 256).  Results do not depend on it. */
int xb_elast_stepn(const xb_elast_args *args, int32_t steps, void *stream);

/* `steps` consecutive XB_MODE_SMOOTH sweeps of one multigrid level in one launch on
 * 32 x 32 LDS tiles (residual, external force, fused coupling force with div_out,
 * per-cell Lame fields as xb_elast_step; bit-identical to `steps` calls).  steps <= 6
 * (<= 3 with the coupling force, which reaches two cells per step).  A gradient wall
 * table of more than 32 slots: XB_EUNSUPPORTED (use the sweeps).
 * With coarse_out set, the sweeps start from f_in plus the injection of that coarse
 * correction over the interior box: the post-smoothing of a V-cycle with its
 * prolong-add (reference src/elasticity.py:1290-1293, _prolong :878-906) fused in,
 * bit-identical to xb_elast_prolong_add followed by this call; f_in is not written
 * and the divergence of the corrected state is formed in the tile (div_in is then
 * only read as the coupling flag).  A null f_in is the zero state (the first sweeps
 * of a coarse level's V-cycle, reference :1288-1289): nothing is read for it, and a
 * non-null div_in (not read) still marks the coupling force. */
int xb_elast_smoothn(const xb_elast_args *args, int32_t steps, void *stream);

/* Number of doubles in norm_partials for a RESIDUAL call with these args. */
int64_t xb_elast_partials_count(const xb_elast_args *args);

/* f[interior] += prolong(coarse) (piecewise-constant 2x2 injection).
 * Replaces _prolong (src/elasticity.py:878-906) and the .at[].add of :1290-1293.
 * coarse is the full coarse array (8, cnx, cny); its interior starts at (ix0, iy0). */
int xb_elast_prolong_add(int32_t nx, int32_t ny, int32_t dtype, void *f,
                         const void *coarse, int32_t cnx, int32_t cny,
                         int32_t ix0, int32_t ix1, int32_t iy0, int32_t iy1,
                         void *stream);

/* As xb_elast_prolong_add, and div[interior] = div(f) of the updated populations
 * (the divergence carried by the fused Biot coupling force), with per-cell Lame
 * fields (N each, both or neither) when given. */
int xb_elast_prolong_add_div(int32_t nx, int32_t ny, int32_t dtype, double lam, double mu, void *f,
                             const void *coarse, int32_t cnx, int32_t cny,
                             int32_t ix0, int32_t ix1, int32_t iy0, int32_t iy1,
                             void *div, const void *lam_field, const void *mu_field, void *stream);

/* eta (2,nx,ny), sigma (3,nx,ny) from f and the optional forces.
 * Replaces update_macroscopic (src/elasticity.py:1002-1038). Either output may be null.
 * lam_field / mu_field (optional, both or neither): per-cell Lame parameters. */
int xb_elast_macro(int32_t nx, int32_t ny, int32_t dtype, double lam, double mu,
                   const void *f, const void *force_a, const void *force_b,
                   void *eta_out, void *sigma_out, const void *lam_field, const void *mu_field,
                   void *stream);

/* div = (-1.5/(3(lam+mu)+1)) * m_s(f - f_sub) (f_sub optional).
 * Replaces update_divergence (src/elasticity.py:1040-1058). */
int xb_elast_divergence(int32_t nx, int32_t ny, int32_t dtype, double lam, double mu,
                        const void *f, const void *f_sub, void *div_out,
                        const void *lam_field, const void *mu_field, void *stream);

/* f = equilibrium(eta, sigma) with eta shifted by -F/2 for each given force.
 * Replaces equilibrium (src/elasticity.py:908-957). */
int xb_elast_equilibrium(int32_t nx, int32_t ny, int32_t dtype, double lam, double mu,
                         const void *eta, const void *sigma, const void *force_a,
                         const void *force_b, void *f_out, const void *lam_field,
                         const void *mu_field, void *stream);

/* ---- coarse tail of a V-cycle in one workgroup ------------------------------------
 *
 * Below a few thousand cells a level's sweep is bound by launch and load latency.
 * xb_vcycle_tail runs the V-cycle of the levels L0 .. L0+nlevels-1 (the recursion of
 * _mg_step below the level L0-1 that calls it, src/elasticity.py:1221-1312) from the
 * zero state in one workgroup: per sweep every cell collides, then every cell pulls
 * and applies its rules, with the per-cell arithmetic of xb_elast_step (bit-identical
 * results).  The first level's resid is the input (the restricted residual of level
 * L0-1); the result is in the first level's f0, for xb_elast_prolong_add(_div). */
typedef struct xb_tail_level {
  int32_t nx, ny, ix0, ix1, iy0, iy1; /* grid and interior box of the level           */
  const xb_bc_table *bc;        /* the level's rules (may be null)                     */
  const xb_bc_table *grad_bc;   /* coupling force: the gradient's wall rules (or null) */
  void *f0, *f1;                /* (8, nx, ny) states; the result is in f0             */
  void *div0, *div1;            /* (nx, ny) divergence carry of the coupling, or null  */
  void *resid;                  /* (8, ix1-ix0, iy1-iy0) the level's residual          */
  void *scratch;                /* (8, nx, ny) post-collision populations              */
} xb_tail_level;

typedef struct xb_tail_plan {   /* filled by xb_vcycle_tail_prepare                    */
  int32_t dtype, nlevels, k_smooth, coupled;
  double lam, mu, gamma, nl_scale;
  void *table;                  /* device copy of the level descriptors                */
  int64_t lds_bytes;            /* LDS the levels' arrays and rule tables take in the
                                   kernel's 128 KB arena, or 0: then they stay in global
                                   memory (the plan's buffers)                          */
} xb_tail_plan;

/* Device bytes of the level descriptors of a tail of nlevels (<= 16) levels. */
int64_t xb_vcycle_tail_bytes(int32_t nlevels);

/* Checks the levels (each coarsens the previous one) and copies their descriptors
 * into table (device memory of xb_vcycle_tail_bytes; synchronous).  coupled: the
 * Biot coupling force G_scale(D f) with the divergence carry (nl_scale as
 * xb_elast_args.nl_scale), else no force (coarse levels carry none). */
int xb_vcycle_tail_prepare(int32_t dtype, double lam, double mu, double gamma, int32_t k_smooth,
                           double nl_scale, int32_t coupled, const xb_tail_level *levels,
                           int32_t nlevels, void *table, int64_t table_bytes, xb_tail_plan *plan);

/* The tail V-cycle: one workgroup, graph-capturable.  With plan->lds_bytes set it
 * copies the rule tables and the first level's residual into LDS, runs there and
 * writes the first level's f0 (the other buffers are then not written). */
int xb_vcycle_tail(const xb_tail_plan *plan, void *stream);

/* ---- adjoint steps ---------------------------------------------------------------
 * No reference counterpart as kernels: they give what jax.grad through the jitted
 * _step (src/elasticity.py:1169-1219) / AdvectionDiffusionReactionBGK.step
 * (src/darcy.py:131-171) would.  With S the step (collide, post-collision rules,
 * stream, post-streaming rules; affine in f), they compute
 *   abar_out = (dS/df)^T abar_in  (+ abar_add)
 * and add abar_in . dS/dtheta, evaluated at the step's input f_in, to the FP64
 * gradient accumulators that are given.  Walls as in the forward step; no multigrid
 * residual, no fused coupling force. */
typedef struct xb_elast_adjoint_args {
  int32_t nx, ny, dtype, pad_;
  double lam, mu;             /* used where the fields are null                       */
  const void *lam_field;      /* optional (nx, ny), with mu_field                     */
  const void *mu_field;
  const void *abar_in;        /* (8, nx, ny) adjoint of the step's output             */
  const void *f_in;           /* (8, nx, ny) populations entering the step            */
  void *abar_out;             /* (8, nx, ny) adjoint of the step's input (!= abar_in) */
  const void *abar_add;       /* optional (8, nx, ny) added to abar_out               */
  double *grad_lam;           /* optional (nx, ny): += dJ/dlambda(x)                  */
  double *grad_mu;            /* optional (nx, ny): += dJ/dmu(x)                      */
  double *grad_force;         /* optional (2, nx, ny): += dJ/dF(x), F the total force */
  const xb_bc_table *bc;      /* the forward step's rule table (optional)             */
} xb_elast_adjoint_args;

typedef struct xb_darcy_adjoint_args {
  int32_t nx, ny, dtype, pad_;
  double omega;               /* used where omega_field is null                      */
  const void *omega_field;    /* optional (nx, ny)                                   */
  const void *source_a;       /* the forward step's sources (optional)               */
  const void *source_b;
  const void *abar_in;        /* (9, nx, ny)                                         */
  const void *f_in;           /* (9, nx, ny)                                         */
  void *abar_out;             /* (9, nx, ny), != abar_in                             */
  const void *abar_add;       /* optional (9, nx, ny)                                */
  double *grad_omega;         /* optional (nx, ny): += dJ/domega(x)                  */
  double *grad_source;        /* optional (nx, ny): += dJ/dS(x), S = source_a + _b   */
  const xb_bc_table *bc;
} xb_darcy_adjoint_args;

int xb_elast_adjoint(const xb_elast_adjoint_args *args, void *stream);
int xb_darcy_adjoint(const xb_darcy_adjoint_args *args, void *stream);

/* ---- adjoint of the multigrid V-cycle and of the Biot coupler -----------------------
 *
 * What jax.grad through _mg_step (src/elasticity.py:1221-1312) and through the coupled
 * step (src/poroelasticity.py:303-339) would give; the reference has no adjoint.
 * Every adjoint output below accumulates (+=) unless it says "written". */

enum { XB_ADJ_SMOOTH = 1, XB_ADJ_RESIDUAL = 2 };

/* Transposed sweep of one V-cycle level.  Forward (xb_elast_step modes SMOOTH and
 * RESIDUAL): o = (1-gamma) f + gamma S(f - r; F) on the whole level, and
 * c = R(f - S(f - r; F)) with R the 2x2 sum over the interior box, r the level's
 * residual (interior box; none on the finest level), F the sweep's total force. */
typedef struct xb_elast_mg_adjoint_args {
  int32_t nx, ny, dtype, mode;  /* mode: XB_ADJ_SMOOTH or XB_ADJ_RESIDUAL              */
  double lam, mu, gamma;        /* lam, mu used where the fields are null              */
  const void *lam_field;        /* optional (nx, ny), with mu_field (finest level)     */
  const void *mu_field;
  int32_t ix0, ix1, iy0, iy1;   /* interior box of the level                           */
  const void *abar_in;          /* SMOOTH: (8, nx, ny) adjoint of o; RESIDUAL: (8, (ix1-ix0)/2, (iy1-iy0)/2) adjoint of c */
  const void *abar_add;         /* optional (8, nx, ny) added to abar_out              */
  void *abar_out;               /* written: (8, nx, ny) adjoint of f (!= abar_in)      */
  void *rbar;                   /* optional (8, ix1-ix0, iy1-iy0): += adjoint of r     */
  void *force_bar;              /* optional, written: (2, nx, ny) adjoint of F         */
  const void *f_in;             /* with grad_lam / grad_mu: the sweep's f              */
  const void *resid;            /* with f_in, where the level subtracts a residual     */
  double *grad_lam;             /* optional (nx, ny), with grad_mu and the fields      */
  double *grad_mu;
  const xb_bc_table *bc;        /* the level's rule table (optional)                   */
} xb_elast_mg_adjoint_args;

int xb_elast_mg_adjoint(const xb_elast_mg_adjoint_args *args, void *stream);

/* `sweeps` consecutive XB_ADJ_SMOOTH transposed sweeps of a grid of at most 65536
 * cells, each followed by xb_elast_coupling_adjoint of its force_bar (scale, grad_bc,
 * the args' lam / mu and fields), in one workgroup: bit-identical to the sequence of
 * those calls, in 1 launch instead of 2 * sweeps.  Grids whose 18 nx ny (+ 8 per
 * interior cell with rbar) values fit 64 KB run with these arrays in LDS (tmp unused).  abar_out receives the result; tmp
 * is (8, nx, ny) scratch; abar_in, abar_out and tmp differ; force_bar is written.  No
 * abar_add and no parameter derivatives (the coarse levels of the transposed
 * V-cycle).  No reference counterpart (the reference's gradients are jax.grad). */
int xb_elast_smooth_adjoint_small(const xb_elast_mg_adjoint_args *args, int32_t sweeps, double scale,
                                  const xb_bc_table *grad_bc, void *tmp, void *stream);

/* `steps` (1..3) consecutive XB_ADJ_SMOOTH transposed sweeps of one level on 32 x 32
 * LDS tiles, each followed by xb_elast_coupling_adjoint of its force adjoint (scale,
 * grad_bc, the args' lam / mu and fields): bit-identical to that sequence of
 * 2 * steps calls (the abar_out of a sweep the abar_in of the next), with abar_in read
 * and abar_out written once.  rbar (optional) accumulates the residual's adjoint;
 * with grad_lam / grad_mu (finest level: the Lame fields, no residual) `states` is a
 * host array of the sweeps' recorded inputs f (device pointers, in the order the
 * sweeps are transposed) and the parameter derivatives of the collision and of the
 * divergence accumulate; base_bar (optional, (2, nx, ny) FP64) += each force adjoint.
 * No abar_add; force_bar must be null (the force adjoint stays in the tile).  Replaces
 * the transposed smoothing of the reference's V-cycle (src/elasticity.py:1258-1267,
 * :1296-1305, differentiated by jax.grad there). */
int xb_elast_smooth_adjoint_tiles(const xb_elast_mg_adjoint_args *args, int32_t steps, double scale,
                                  const xb_bc_table *grad_bc, const void *const *states, double *base_bar,
                                  void *stream);

/* Transposed Biot coupling force F = G_scale(D f) of one sweep (the fused force of
 * xb_elast_args.div_in): abar[d] += k w2_d (G^T force_bar) with w2 = 1,1,1,1,2,2,2,2
 * and k the divergence coefficient; with f_in and the fields, the Lame-parameter
 * derivative of D at f; base_bar (2, nx, ny, FP64) += force_bar when given.
 * grad_bc: the gradient's wall table for this grid (as xb_elast_args.grad_bc). */
int xb_elast_coupling_adjoint(int32_t nx, int32_t ny, int32_t dtype, double lam, double mu,
                              const void *lam_field, const void *mu_field, const void *force_bar,
                              double scale, const xb_bc_table *grad_bc, void *abar,
                              const void *f_in, double *grad_lam, double *grad_mu,
                              double *base_bar, void *stream);

/* qbar (1, nx, ny) += b * G_scale^T gbar: the transpose of xb_darcy_gradient's G. */
int xb_darcy_gradient_adjoint(int32_t nx, int32_t ny, int32_t dtype, const void *gbar, double scale,
                              const xb_bc_table *bc, double b, void *qbar, void *stream);

/* coarse (8, cnx, cny), written: the sum of the four children of every coarse
 * interior cell, zero on the coarse boundary layers (the transpose of the injection
 * of xb_elast_prolong_add; arguments as there). */
int xb_elast_restrict_sum(int32_t nx, int32_t ny, int32_t dtype, const void *f, void *coarse,
                          int32_t cnx, int32_t cny, int32_t ix0, int32_t ix1, int32_t iy0,
                          int32_t iy1, void *stream);

/* fbar += (d eta/d f)^T eta_bar + (d sigma/d f)^T sigma_bar (xb_elast_macro; either
 * bar may be null); with the fields and grad_lam / grad_mu, the stresses'
 * Lame-parameter derivatives.  The forces' adjoint is eta_bar / 2. */
int xb_elast_macro_adjoint(int32_t nx, int32_t ny, int32_t dtype, double lam, double mu,
                           const void *f, const void *eta_bar, const void *sigma_bar, void *fbar,
                           const void *lam_field, const void *mu_field, double *grad_lam,
                           double *grad_mu, void *stream);

/* div = k (m_s(f) - m_s(f_sub)) (xb_elast_divergence): fbar += k w2 div_bar,
 * fsub_bar -= k w2 div_bar (either may be null); with the fields and grad_lam /
 * grad_mu (f_sub may be null), the derivative of k. */
int xb_elast_divergence_adjoint(int32_t nx, int32_t ny, int32_t dtype, double lam, double mu,
                                const void *f, const void *f_sub, const void *div_bar, void *fbar,
                                void *fsub_bar, const void *lam_field, const void *mu_field,
                                double *grad_lam, double *grad_mu, void *stream);

/* xb_darcy_macro transposed: fbar_d += p_bar + (-3 omega) c_d . grad_bar,
 * source_bar += p_bar / 2 (the adjoint of either source), grad_omega += the
 * relaxation-rate derivative of the gradient (needs f).  Bars may be null. */
int xb_darcy_macro_adjoint(int32_t nx, int32_t ny, int32_t dtype, double omega, const void *omega_field,
                           const void *f, const void *p_bar, const void *grad_bar, void *fbar,
                           void *source_bar, double *grad_omega, void *stream);

/* ---- Darcy (D2Q9 BGK advection-diffusion-reaction, zero velocity) ------------------ */

/* Fused BGK collide -> BCs -> stream -> BCs.  Replaces
 * AdvectionDiffusionReactionBGK.step (src/darcy.py:131-171, collision :106-129). */
int xb_darcy_step(const xb_darcy_args *args, void *stream);

/* Two Darcy steps in one pass (as xb_elast_step2; no fpc_out). */
int xb_darcy_step2(const xb_darcy_args *args, void *stream);

/* `steps` (1..8) consecutive Darcy steps in one launch on 32 x 32 LDS tiles
 * (temporal blocking; bit-identical to `steps` xb_darcy_step calls; no fpc_out).
 * args->rows is the tile block size, as in xb_elast_stepn. */
int xb_darcy_stepn(const xb_darcy_args *args, int32_t steps, void *stream);

/* p = sum_d f_d + (s_a + s_b)/2 ; grad = sum_d f_d (-3 omega) c_d.
 * Replaces update_macroscopic (src/darcy.py:36-73). Either output may be null. */
int xb_darcy_macro(int32_t nx, int32_t ny, int32_t dtype, double omega,
                   const void *omega_field, const void *f, const void *source_a,
                   const void *source_b, void *p_out, void *grad_out, void *stream);

/* out = a*base + b*G(q) with G(q) = sum_d qs_d (-3 scale w_d c_d), qs_d(x) = q(x - c_d)
 * or the boundary rule of the cell (bounce-back / anti-bounce-back walls only).
 * Replaces gradient (src/darcy.py:75-104). base may be null (then a is ignored). */
int xb_darcy_gradient(int32_t nx, int32_t ny, int32_t dtype, const void *q, double scale,
                      const xb_bc_table *bc, const void *base, double a, double b,
                      void *out, void *stream);

/* ---- helpers of the Biot coupler (src/poroelasticity.py:303-339) --------------------- */

/* out = a*x + b*y over n elements (y may be null). */
int xb_lincomb(int64_t n, int32_t dtype, double a, const void *x, double b, const void *y,
               void *out, void *stream);

/* out[0..n) = value.  A kernel, not a memset, so that it may sit in a captured graph. */
int xb_fill(int64_t n, int32_t dtype, double value, void *out, void *stream);

/* partials[k] = sum over a slice of (a - b)^2 (b may be null) in FP64; returns the
 * number of partials written through *count.  Used by convergence checks. */
int xb_sqdiff_partials(int64_t n, int32_t dtype, const void *a, const void *b,
                       double *partials, int32_t max_partials, int32_t *count, void *stream);

/* out[0] = sqrt(sum partials[0..n)) in a fixed order; out[1] = 1 if non-finite. */
int xb_norm_finalize(const double *partials, int32_t n, double *out, void *stream);

/* ---- batches: many independent grids per launch -------------------------------------
 *
 * The reference solves one problem per solver object; its parameter sweeps run the
 * solvers one after another (the run scripts of the examples).  A batch puts
 * one operation of many instances (grids of any sizes and parameters) into one
 * launch.  A plan is prepared once (host-side job table copied into a device
 * workspace the caller owns; this call synchronises the stream and must not be
 * captured) and launched any number of times (capturable in a HIP graph).  Every
 * launch takes an optional device array `active` of int32 flags indexed by the
 * instance numbers of the jobs: the jobs of an instance whose flag is 0 do nothing. */
enum {
  XB_BATCH_ELAST_STEP = 1,  /* jobs: xb_elast_args[] (one xb_elast_step each)             */
  XB_BATCH_PROLONG_DIV = 2, /* jobs: xb_prolong_job[] (xb_elast_prolong_add[_div] each)   */
  XB_BATCH_FILL = 3,        /* jobs: xb_fill_job[] (xb_fill each)                         */
  XB_BATCH_NORM = 4,        /* jobs: xb_norm_job[] (for xb_batch_converge), one per instance */
  XB_BATCH_ELAST_SMOOTHN = 5 /* jobs: xb_elast_args[] in XB_MODE_SMOOTH, rows = steps per
                                launch (one xb_elast_smoothn each)                        */
};

typedef struct xb_prolong_job {
  int32_t nx, ny, cnx, cny;
  int32_t ix0, ix1, iy0, iy1;
  double lam, mu;           /* divergence factor of div                              */
  void *f;                  /* (8, nx, ny)                                           */
  const void *coarse;       /* (8, cnx, cny)                                         */
  void *div;                /* optional (nx, ny): div(f) on the interior afterwards  */
} xb_prolong_job;

typedef struct xb_fill_job {
  int64_t n;
  void *out;
  double value;
} xb_fill_job;

typedef struct xb_norm_job {
  const double *partials;   /* residual partials of the instance's finest level      */
  int32_t count;            /* xb_elast_partials_count() of that residual sweep      */
  int32_t pad_;
} xb_norm_job;

typedef struct xb_batch_plan { /* filled by xb_batch_prepare                          */
  int32_t kind, dtype, mode, flags;
  int32_t njobs, nblocks;
  void *table;
  int64_t table_bytes;
  int64_t jobs_offset, blocks_offset, insts_offset;
  /* per job: first block, block count, instance (3 x njobs int32) */
  int64_t jobinfo_offset;
  /* work list written by xb_batch_compact: active jobs (njobs), prefix of their block
   * counts (njobs + 1), count of active jobs (1); 0: the kind has none (NORM) */
  int64_t work_offset;
  /* set by the caller: 1 = launch over the work list (a grid of at most as many
   * workgroups as the GPU holds at once, each taking listed blocks in turn) instead
   * of one workgroup per table block */
  int32_t use_list;
  /* set by the caller: workgroups per resident slot of a list launch (0 or 1: as many
   * as the GPU holds at once; more: hardware scheduling balances the blocks) */
  int32_t list_waves;
} xb_batch_plan;

/* Per-instance convergence state (device arrays of ninst entries, nactive one int).
 * Before a solve: r0 < 0, cycles 0, active 1 (or 0 to leave an instance out). */
typedef struct xb_converge_state {
  double *r0;               /* residual norm of the first cycle                      */
  double *rlast;            /* residual norm of the last cycle                       */
  int32_t *cycles;          /* cycles done                                           */
  int32_t *active;          /* 1 while iterating                                     */
  int32_t *status;          /* 0 running, 1 converged, 2 out of cycles, 3 non-finite */
  int32_t *nactive;         /* number of active instances after the update           */
  const double *rref;       /* optional: residual scale per instance (<= 0: use r0)   */
} xb_converge_state;

/* Bytes of device workspace a plan of these jobs needs (-1 on bad arguments). */
int64_t xb_batch_table_bytes(int32_t kind, int32_t dtype, const void *jobs, int32_t njobs);

/* Validate the jobs (as the single-grid calls do) and upload their table.  inst[k]
 * is job k's instance number (null: k).  Elasticity jobs must share dtype, mode and
 * the presence of bc, resid and div_in. */
int xb_batch_prepare(int32_t kind, int32_t dtype, const void *jobs, const int32_t *inst, int32_t njobs,
                     void *table, int64_t table_bytes, xb_batch_plan *plan, void *stream);

/* One launch of every job of the plan (kinds ELAST_STEP, ELAST_SMOOTHN, PROLONG_DIV,
 * FILL).  With plan->use_list the launch covers the blocks of the jobs listed by the
 * last xb_batch_compact of that plan (whatever `active` says now); otherwise every
 * table block is a workgroup and blocks of inactive instances return at once. */
int xb_batch_launch(const xb_batch_plan *plan, const int32_t *active, void *stream);

/* Work lists of many plans in one launch: for each plan of `plans` (a DEVICE array of
 * nplans xb_batch_plan structs, copies of prepared plans), the jobs whose instance is
 * active (active[inst] != 0; all when active is null), in job order, and the prefix
 * of their block counts, written into that plan's table.  Capture-safe (one kernel).
 * No single reference method: the per-instance convergence masks of the batched
 * loop over solver objects, turned into the work of the next launches. */
int xb_batch_compact(const xb_batch_plan *plans, int32_t nplans, const int32_t *active, void *stream);

/* Convergence update after one V-cycle of every active instance (plan of kind NORM):
 * r = sqrt(sum of its partials) (same order as xb_norm_finalize); cycles += 1;
 * the first cycle sets r0; the instance stops when r <= tol * s (status 1), with
 * s = rref[i] where given and positive, else r0; when
 * cycles reaches max_cycles (2), or when r is not finite (3). Replaces the
 * host-side test of the residual-driven inner loop of the Biot step. */
int xb_batch_converge(const xb_batch_plan *plan, int32_t ninst, const xb_converge_state *state, double tol,
                      int32_t max_cycles, void *stream);

/* out[inst] = sqrt(sum of the job's partials) for every job of a NORM plan (same
 * order as xb_norm_finalize).  With the partials of a RESIDUAL sweep of the zero
 * state this is the residual scale |S(0)| of each instance. */
int xb_batch_norms(const xb_batch_plan *plan, double *out, void *stream);

/* Capacity of the static scratch of xb_elast_residual_noalloc, in FP64 partials
 * (one per wavefront item of the residual sweep: 1 MiB per GPU). */
#define XB_STATIC_PARTIALS (1 << 17)

/* Allocation-free residual sweep: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid it).  The same sweep as xb_elast_step in XB_MODE_RESIDUAL
 * (restricted residual into coarse_out, with the same arguments and checks)
 * followed by xb_norm_finalize, writing (norm, non-finite flag) to norm_out[0..1];
 * the squared-norm partials live in a static device array of XB_STATIC_PARTIALS
 * doubles held by each GPU the library runs on, instead of a caller workspace, so
 * args->norm_partials must be null.  A sweep of more wavefront items than that is
 * rejected with XB_EINVAL.  Calls no allocation routine.  Calls that overlap on one
 * GPU (several streams or host threads) race on the shared array; the stream must
 * belong to the current device. */
int xb_elast_residual_noalloc(const xb_elast_args *args, double *norm_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* XLB_BIOT_H */
