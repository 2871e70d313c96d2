// gfx950 kernels and C ABI of the lattice-Boltzmann elasticity / Darcy / Biot path.
//
// The hot kernel is k_sweep: one wavefront owns a strip of 62 output cells along
// y (plus one halo lane on each side) and walks down a block of rows along x.
// For every row it loads the populations once (coalesced SoA loads), collides
// them in registers, applies the post-collision boundary rules, and keeps three
// consecutive rows of post-collision populations in registers.  Streaming is a
// pull: direction d of the output row x takes row x - cx_d of the register ring
// and lane (lane - cy_d) through a one-lane wavefront shift.  Post-streaming
// boundary rules, damping (multigrid smoothing) or residual + 2x2 restriction +
// squared norm are fused behind it.  Each population is read and written once
// per step; halo rows and lanes are re-read from L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <exception>
#include <mutex>
#include <new>
#include <type_traits>
#include <vector>

#include "xb_device.h"

using namespace xb;

// The library is compiled as one source in several parts (XB_PART = 0, 1, ...: each
// part defines a disjoint group of the C entry points, so the parts compile in
// parallel and link into one library); without XB_PART one compilation holds all.
#ifndef XB_PART
#define XB_IN_PART(n) 1
#else
#define XB_IN_PART(n) (XB_PART == (n))
#endif

// batched launches of the sweep and smoothing-tile kinds, each compiled in a part of
// its own (they instantiate most of the kernels); the dtype is the plan's
int xb_part_elast_batch_step(const xb_batch_plan *p, const int32_t *active, hipStream_t st);
int xb_part_elast_batch_smooth(const xb_batch_plan *p, const int32_t *active, hipStream_t st);
int xb_part_elast_batch_residual(const xb_batch_plan *p, const int32_t *active, hipStream_t st);
int xb_part_smoothn_batch(const xb_batch_plan *p, const int32_t *active, hipStream_t st);

// the calling thread's last error message (one definition, shared by every part)
extern thread_local char xb_err_msg[512];
#if XB_IN_PART(0)
thread_local char xb_err_msg[512] = "";
#endif

namespace {

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(xb_err_msg, sizeof(xb_err_msg), fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(XB_EHIP, "%s: %s", what, hipGetErrorString(e));
  return XB_OK;
}

// Runs `body` (returning an int status) and turns any C++ exception into an error
// status, so none crosses the C ABI (an exception leaving an extern "C" function
// would end the process).  Used by the entry points that build host-side tables
// in std containers; the others call only the HIP C API.
template <class F> int guarded(const char *what, F &&body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    return fail(XB_EHIP, "%s: host memory allocation failed", what);
  } catch (const std::exception &e) {
    return fail(XB_EHIP, "%s: %s", what, e.what());
  } catch (...) {
    return fail(XB_EHIP, "%s: unexpected C++ exception", what);
  }
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

struct SweepGeom {
  int nx, ny;
  int ox, oy, ex, ey;     // output box
  int rows;               // rows per wavefront item
  int nstrips, nitems;
  int ix0, ix1, iy0, iy1; // box carrying the subtracted residual
  int rmode;              // storage rounding of population values where stored:
                          // 0 none, 1 FP32 (XB_F64_S32), 2 FP16 (XB_F64_S16, XB_F32_S16)
};

// Storage in a narrower type than the arithmetic: the value an FP32 (mode 1) or FP16
// (mode 2, one correctly rounded conversion from T, as the reference's astype) array
// would hold
template <typename T, int Q> __device__ __forceinline__ void store_round(T (&v)[Q], int mode) {
  if (mode == 1) {
#pragma unroll
    for (int d = 0; d < Q; ++d) v[d] = T(float(v[d]));
  } else {
#pragma unroll
    for (int d = 0; d < Q; ++d) v[d] = T(_Float16(v[d]));
  }
}

// Rows per wavefront: enough items to give every SIMD several waves, at most 64
// and at least min_rows (2 for the residual sweep, whose restriction pairs rows).
SweepGeom make_geom(int nx, int ny, int ox, int oy, int ex, int ey, int ix0, int ix1, int iy0, int iy1,
                    int rows_hint, int strip = STRIP, int min_rows = 2) {
  SweepGeom g{};
  g.nx = nx; g.ny = ny; g.ox = ox; g.oy = oy; g.ex = ex; g.ey = ey;
  g.ix0 = ix0; g.ix1 = ix1; g.iy0 = iy0; g.iy1 = iy1;
  g.nstrips = cdiv(ey, strip);
  int rows = 64;
  if (rows_hint >= min_rows && rows_hint <= 64 && (rows_hint & (rows_hint - 1)) == 0) {
    rows = rows_hint;
  } else {
    const int target = 8192;
    while (rows > min_rows && (int64_t)g.nstrips * cdiv(ex, rows) < target) rows >>= 1;
  }
  g.rows = rows;
  g.nitems = g.nstrips * cdiv(ex, rows);
  return g;
}

// ---------------------------------------------------------------------------
// One wavefront item of the sweep: a strip of 62 cells along y walked down `g.rows`
// rows.  Shared by the single-grid kernel and the batched kernel, so both do the
// same arithmetic in the same order.
template <typename T, class Coll, int MODE, int BC, bool HAS_RES, bool HAS_NL>
__device__ __forceinline__ void sweep_item(const SweepGeom &g, const Coll &coll, const T *__restrict__ fin,
                                           T *__restrict__ fout, T *__restrict__ fpc, const T *__restrict__ res,
                                           const BCDev<T> &bc, const T omg, const T gamma, T *__restrict__ coarse,
                                           double *__restrict__ partials, const NLDev<T> &nl,
                                           BCShared<T, Coll::Q> *sh, const int item) {
  constexpr int Q = Coll::Q;
  constexpr bool HAS_BC = BC != 0;
  bool staged = false; // BC == 2: the table is in this wavefront's LDS region
  auto stage_if_needed = [&](int slot) {
    if constexpr (BC == 2) {
      if (!staged && __ballot(slot >= 0) != 0) { // uniform per wavefront
        stage_rules_wave<T, Q>(bc, *sh);
        staged = true;
      }
    }
  };
  // rule rows of a slot, from LDS (BC == 2) or global memory (BC == 1)
  auto rules_pc = [&](T(&v)[Q], const T(&src)[Q], int slot) {
    if constexpr (BC == 2) apply_rules<T, Q>(v, src, sh->code_pc + slot * Q, sh->const_pc + slot * Q);
    else apply_rules<T, Q>(v, src, bc.code_pc + (size_t)slot * Q, bc.const_pc + (size_t)slot * Q);
  };
  auto rules_ps = [&](T(&v)[Q], const T(&src)[Q], int slot) {
    if constexpr (BC == 2) apply_rules<T, Q>(v, src, sh->code_ps + slot * Q, sh->const_ps + slot * Q);
    else apply_rules<T, Q>(v, src, bc.code_ps + (size_t)slot * Q, bc.const_ps + (size_t)slot * Q);
  };
  const int lane = threadIdx.x & (WAVE - 1);

  const int strip = item % g.nstrips;
  const int rb = item / g.nstrips;
  const int ys = g.oy + strip * STRIP;
  const int yend = min(ys + STRIP, g.oy + g.ey);
  const int yv = ys - 1 + lane;
  const bool out_lane = lane >= 1 && yv < yend;
  const int y = wrapi(min(yv, yend), g.ny);
  const int xs = g.ox + rb * g.rows;
  const int xe = min(xs + g.rows, g.ox + g.ex);
  const size_t N = (size_t)g.nx * (size_t)g.ny;
  const int nyi = g.iy1 - g.iy0;
  const size_t Ni = (size_t)(g.ix1 - g.ix0) * (size_t)nyi;
  const bool y_in_box = y >= g.iy0 && y < g.iy1;

  using Aux = typename Coll::Aux;
  // Everything one row needs from memory.  Rows are loaded one step ahead of their
  // collision (two rows in flight per wavefront) so that the load latency of row
  // x+2 overlaps the collision of row x+1 and the stores of row x.
  struct Row {
    T f[Q];
    T r[Q];
    T dv[9];   // HAS_NL: divergence on the D2Q9 stencil around the cell
    Aux aux;
    int slot;
    int gslot; // HAS_NL: Darcy wall rule slot of the cell
    bool has_r;
  };
  auto load_row = [&](int x, Row &rw) {
    const size_t cell = (size_t)x * g.ny + y;
#pragma unroll
    for (int d = 0; d < Q; ++d) rw.f[d] = fin[d * N + cell];
    rw.has_r = false;
    if constexpr (HAS_RES) {
      rw.has_r = y_in_box && x >= g.ix0 && x < g.ix1;
      if (rw.has_r) {
        const size_t ci = (size_t)(x - g.ix0) * nyi + (size_t)(y - g.iy0);
#pragma unroll
        for (int d = 0; d < Q; ++d) rw.r[d] = res[d * Ni + ci];
      }
    }
    rw.aux = coll.load_aux(cell);
    rw.slot = -1;
    if constexpr (HAS_BC) rw.slot = bc.slot[cell];
    if constexpr (HAS_NL) {
      using D = DarcyColl<T>;
#pragma unroll
      for (int d = 0; d < 9; ++d)
        rw.dv[d] = nl.div_in[(size_t)wrapi(x - D::cx(d), g.nx) * g.ny + wrapi(y - D::cy(d), g.ny)];
      rw.gslot = nl.gbc.slot != nullptr ? nl.gbc.slot[cell] : -1;
    }
  };
  // Subtract the residual where it lives, collide, post-collision rules.
  auto finish_row = [&](const Row &rw, T(&pc)[Q]) {
    T pre[Q];
#pragma unroll
    for (int d = 0; d < Q; ++d) pre[d] = rw.f[d];
    if constexpr (HAS_RES) {
      if (rw.has_r) {
#pragma unroll
        for (int d = 0; d < Q; ++d) pre[d] = pre[d] - rw.r[d];
      }
    }
    if constexpr (HAS_NL) {
      // external force + grad(div f) (reference poroelasticity.py:324, elasticity.py:1259-1265)
      T qs[9];
#pragma unroll
      for (int d = 0; d < 9; ++d) qs[d] = rw.dv[d];
      if (rw.gslot >= 0) {
        T src[9];
#pragma unroll
        for (int d = 0; d < 9; ++d) src[d] = rw.dv[0];
        apply_rules<T, 9>(qs, src, nl.gbc.code_ps + (size_t)rw.gslot * 9, nl.gbc.const_ps + (size_t)rw.gslot * 9);
      }
      T g0, g1;
      stencil_grad<T>(qs, nl.m3s, g0, g1);
      Aux ax = rw.aux;
      ax.b0 = nl.base != nullptr ? ax.b0 + g0 : g0;
      ax.b1 = nl.base != nullptr ? ax.b1 + g1 : g1;
      coll.collide(pre, ax, pc);
    } else {
      coll.collide(pre, rw.aux, pc);
    }
    if (g.rmode) store_round(pc, g.rmode);
    if constexpr (HAS_BC) {
      stage_if_needed(rw.slot);
      if (rw.slot >= 0 && bc.code_pc != nullptr) {
        rules_pc(pc, pre, rw.slot);
        if (g.rmode) store_round(pc, g.rmode);
      }
    }
  };

  T R0[Q], R1[Q], R2[Q], F1[Q];
  T K1 = nl.kdiv; // HAS_NL: divergence coefficient of the output row (per cell with Lame fields)
  int S1 = -1;
  double nrm = 0.0;
  T acc[Q];
#pragma unroll
  for (int d = 0; d < Q; ++d) acc[d] = T(0);
  const int ncy = g.ey >> 1;
  const size_t Nc = (size_t)(g.ex >> 1) * (size_t)ncy;
  const bool first_of_pair = out_lane && (((yv - g.oy) & 1) == 0);

  // Output row x from the ring (R0, R1, R2) = post-collision rows (x-1, x, x+1).
  auto emit = [&](int x) {
    T v[Q];
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      T s = Coll::cx(d) == 1 ? R0[d] : (Coll::cx(d) == 0 ? R1[d] : R2[d]);
      if (Coll::cy(d) == 1) s = from_prev_lane(s);
      else if (Coll::cy(d) == -1) s = from_next_lane(s);
      v[d] = s;
    }
    if constexpr (HAS_BC) {
      if (S1 >= 0 && bc.code_ps != nullptr) {
        rules_ps(v, R1, S1);
        if (g.rmode) store_round(v, g.rmode);
      }
    }
    const size_t cell = (size_t)x * g.ny + y;
    if constexpr (MODE == XB_MODE_STEP) {
      if (out_lane) {
#pragma unroll
        for (int d = 0; d < Q; ++d) fout[d * N + cell] = v[d];
        if (fpc != nullptr) {
#pragma unroll
          for (int d = 0; d < Q; ++d) fpc[d * N + cell] = R1[d];
        }
        if constexpr (HAS_NL && Q == 8) {
          if (nl.div_out != nullptr) nl.div_out[cell] = elast_div_of<T>(v, K1);
        }
      }
    } else if constexpr (MODE == XB_MODE_SMOOTH) {
      if (out_lane) {
        T o[Q];
#pragma unroll
        for (int d = 0; d < Q; ++d) o[d] = omg * F1[d] + gamma * v[d];
        if (g.rmode) store_round(o, g.rmode);
#pragma unroll
        for (int d = 0; d < Q; ++d) fout[d * N + cell] = o[d];
        if constexpr (HAS_NL && Q == 8) {
          if (nl.div_out != nullptr) nl.div_out[cell] = elast_div_of<T>(o, K1);
        }
      }
    } else {
      // residual r = f - S(f - r_in) on the interior box, 2x2 sum, squared norm
      T pair[Q], r[Q];
#pragma unroll
      for (int d = 0; d < Q; ++d) r[d] = F1[d] - v[d];
      if (g.rmode) store_round(r, g.rmode);
#pragma unroll
      for (int d = 0; d < Q; ++d) {
        if (out_lane) nrm += (double)r[d] * (double)r[d];
        pair[d] = r[d] + from_next_lane(r[d]);
      }
      if (((x - g.ox) & 1) == 0) {
#pragma unroll
        for (int d = 0; d < Q; ++d) acc[d] = pair[d];
      } else if (first_of_pair) {
        const size_t cc = (size_t)((x - g.ox) >> 1) * ncy + (size_t)((yv - g.oy) >> 1);
#pragma unroll
        for (int d = 0; d < Q; ++d) coarse[d * Nc + cc] = acc[d] + pair[d];
      }
    }
  };
  // Finish row x+1 from `cur`, emit row x, rotate the ring.
  auto advance = [&](int x, const Row &cur) {
    finish_row(cur, R2);
    emit(x);
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      R0[d] = R1[d];
      R1[d] = R2[d];
      F1[d] = cur.f[d];
    }
    S1 = cur.slot;
    if constexpr (HAS_NL && Q == 8) K1 = coll.kdiv(cur.aux, nl.kdiv);
  };

  // rows xs-1 .. xe are read (xe: the halo row after the last output row)
  Row A, B;
  load_row(wrapi(xs - 1, g.nx), A);
  load_row(xs, B);
  finish_row(A, R0);
  load_row(wrapi(xs + 1, g.nx), A);
  finish_row(B, R1);
#pragma unroll
  for (int d = 0; d < Q; ++d) F1[d] = B.f[d];
  S1 = B.slot;
  if constexpr (HAS_NL && Q == 8) K1 = coll.kdiv(B.aux, nl.kdiv);
  // A holds row x+1 at even steps, B at odd steps; the other buffer takes row x+2
  for (int x = xs; x < xe; x += 2) {
    if (x + 2 <= xe) load_row(wrapi(x + 2, g.nx), B);
    advance(x, A);
    if (x + 1 >= xe) break;
    if (x + 3 <= xe) load_row(wrapi(x + 3, g.nx), A);
    advance(x + 1, B);
  }

  if constexpr (MODE == XB_MODE_RESIDUAL) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off, WAVE);
    if (lane == 0) partials[item] = nrm;
  }
}

template <typename T, class Coll, int MODE, int BC, bool HAS_RES, bool HAS_NL>
__global__ void __launch_bounds__(BLOCK)
k_sweep(const SweepGeom g, const Coll coll, const T *__restrict__ fin, T *__restrict__ fout,
        T *__restrict__ fpc, const T *__restrict__ res, const BCDev<T> bc, const T omg, const T gamma,
        T *__restrict__ coarse, double *__restrict__ partials, const NLDev<T> nl) {
  const int item = xcd_block_id() * WAVES_PER_BLOCK + (int)(threadIdx.x >> 6);
  if constexpr (BC == 2) {
    __shared__ BCShared<T, Coll::Q> sh[WAVES_PER_BLOCK];
    if (item >= g.nitems) return; // uniform per wavefront
    sweep_item<T, Coll, MODE, BC, HAS_RES, HAS_NL>(g, coll, fin, fout, fpc, res, bc, omg, gamma, coarse, partials,
                                                   nl, &sh[threadIdx.x >> 6], item);
  } else {
    if (item >= g.nitems) return;
    sweep_item<T, Coll, MODE, BC, HAS_RES, HAS_NL>(g, coll, fin, fout, fpc, res, bc, omg, gamma, coarse, partials,
                                                   nl, nullptr, item);
  }
}

// ---------------------------------------------------------------------------
// Two plain steps per launch (temporal blocking): f_out = S(S(f_in)).  A wavefront
// owns 60 output cells along y (two halo lanes on each side) and walks input rows
// xs-2 .. xe+1.  Stage 1 collides each input row into a ring of three rows and
// streams the middle one (post-streaming rules included): the step-1 populations of
// that row, still in registers.  Stage 2 collides those into a second ring and
// streams again into the output.  Each population is read and written once per two
// steps; per cell the arithmetic is exactly that of two k_sweep STEP launches.
constexpr int STRIP2 = WAVE - 4;

template <typename T, class Coll, int BC>
__device__ __forceinline__ void sweep2_item(const SweepGeom &g, const Coll &coll, const T *__restrict__ fin,
                                            T *__restrict__ fout, const BCDev<T> &bc, BCShared<T, Coll::Q> *sh,
                                            const int item) {
  constexpr int Q = Coll::Q;
  constexpr bool HAS_BC = BC != 0;
  bool staged = false;
  using Aux = typename Coll::Aux;
  const int lane = threadIdx.x & (WAVE - 1);
  const int strip = item % g.nstrips;
  const int rb = item / g.nstrips;
  const int ys = g.oy + strip * STRIP2;
  const int yend = min(ys + STRIP2, g.oy + g.ey);
  const int yv = ys - 2 + lane;
  const bool out_lane = lane >= 2 && yv < yend;
  const int y = wrapi(min(yv, yend + 1), g.ny);
  const int xs = g.ox + rb * g.rows;
  const int xe = min(xs + g.rows, g.ox + g.ex);
  const size_t N = (size_t)g.nx * g.ny;

  auto load = [&](int x, T(&f)[Q], Aux &aux, int &slot) {
    const size_t cell = (size_t)x * g.ny + y;
#pragma unroll
    for (int d = 0; d < Q; ++d) f[d] = fin[d * N + cell];
    aux = coll.load_aux(cell);
    slot = -1;
    if constexpr (HAS_BC) slot = bc.slot[cell];
  };
  auto collide = [&](const T(&pre)[Q], const Aux &aux, int slot, T(&pc)[Q]) {
    coll.collide(pre, aux, pc);
    if constexpr (BC == 2) {
      if (!staged && __ballot(slot >= 0) != 0) { // uniform per wavefront
        stage_rules_wave<T, Q>(bc, *sh);
        staged = true;
      }
    }
    if constexpr (HAS_BC) {
      if (slot >= 0 && bc.code_pc != nullptr) {
        if constexpr (BC == 2) apply_rules<T, Q>(pc, pre, sh->code_pc + slot * Q, sh->const_pc + slot * Q);
        else apply_rules<T, Q>(pc, pre, bc.code_pc + (size_t)slot * Q, bc.const_pc + (size_t)slot * Q);
      }
    }
  };
  // streamed populations of the middle row of (r0, r1, r2) = rows (x-1, x, x+1)
  auto stream = [&](const T(&r0)[Q], const T(&r1)[Q], const T(&r2)[Q], int slot, T(&v)[Q]) {
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      T s = Coll::cx(d) == 1 ? r0[d] : (Coll::cx(d) == 0 ? r1[d] : r2[d]);
      if (Coll::cy(d) == 1) s = from_prev_lane(s);
      else if (Coll::cy(d) == -1) s = from_next_lane(s);
      v[d] = s;
    }
    if constexpr (HAS_BC) {
      if (slot >= 0 && bc.code_ps != nullptr) {
        if constexpr (BC == 2) apply_rules<T, Q>(v, r1, sh->code_ps + slot * Q, sh->const_ps + slot * Q);
        else apply_rules<T, Q>(v, r1, bc.code_ps + (size_t)slot * Q, bc.const_ps + (size_t)slot * Q);
      }
    }
  };

  T R0[Q], R1[Q], R2[Q], Q0[Q], Q1[Q], Q2[Q], cur[Q], nxt[Q];
#pragma unroll
  for (int d = 0; d < Q; ++d) R1[d] = R2[d] = Q1[d] = Q2[d] = T(0);
  Aux acur, anxt, aP;
  int scur, snxt = -1, sP = -1, sPP = -1;
  load(wrapi(xs - 2, g.nx), cur, acur, scur);
  for (int i = xs - 2; i <= xe + 1; ++i) {
    if (i + 1 <= xe + 1) load(wrapi(i + 1, g.nx), nxt, anxt, snxt);
    T C[Q];
    collide(cur, acur, scur, C);
#pragma unroll
    for (int d = 0; d < Q; ++d) {
      R0[d] = R1[d];
      R1[d] = R2[d];
      R2[d] = C[d];
    }
    if (i >= xs) { // (R0, R1, R2) = collided rows (i-2, i-1, i): step 1 of row i-1
      T v1[Q], D[Q];
      stream(R0, R1, R2, sP, v1);
      collide(v1, aP, sP, D);
#pragma unroll
      for (int d = 0; d < Q; ++d) {
        Q0[d] = Q1[d];
        Q1[d] = Q2[d];
        Q2[d] = D[d];
      }
      if (i >= xs + 2) { // (Q0, Q1, Q2) = step-1 rows (i-3, i-2, i-1) collided: output row i-2
        T o[Q];
        stream(Q0, Q1, Q2, sPP, o);
        if (out_lane) {
          const size_t cell = (size_t)(i - 2) * g.ny + y;
#pragma unroll
          for (int d = 0; d < Q; ++d) fout[d * N + cell] = o[d];
        }
      }
    }
    sPP = sP;
    sP = scur;
    aP = acur;
#pragma unroll
    for (int d = 0; d < Q; ++d) cur[d] = nxt[d];
    acur = anxt;
    scur = snxt;
  }
}

template <typename T, class Coll, int BC>
__global__ void __launch_bounds__(BLOCK)
k_sweep2(const SweepGeom g, const Coll coll, const T *__restrict__ fin, T *__restrict__ fout, const BCDev<T> bc) {
  const int item = xcd_block_id() * WAVES_PER_BLOCK + (int)(threadIdx.x >> 6);
  if constexpr (BC == 2) {
    __shared__ BCShared<T, Coll::Q> sh[WAVES_PER_BLOCK];
    if (item >= g.nitems) return;
    sweep2_item<T, Coll, BC>(g, coll, fin, fout, bc, &sh[threadIdx.x >> 6], item);
  } else {
    if (item >= g.nitems) return;
    sweep2_item<T, Coll, BC>(g, coll, fin, fout, bc, nullptr, item);
  }
}

template <typename T> BCDev<T> bcdev(const xb_bc_table *t);

template <typename T, class Coll>
int launch_sweep2(const SweepGeom &g, const Coll &coll, const void *fin, void *fout, const xb_bc_table *bct,
                  hipStream_t st) {
  const BCDev<T> bc = bcdev<T>(bct);
  const dim3 grid(cdiv(g.nitems, WAVES_PER_BLOCK)), block(BLOCK);
  const T *fi = static_cast<const T *>(fin);
  T *fo = static_cast<T *>(fout);
  switch (bc_mode(bc.slot, bc.nslots)) {
  case 2: hipLaunchKernelGGL((k_sweep2<T, Coll, 2>), grid, block, 0, st, g, coll, fi, fo, bc); break;
  case 1: hipLaunchKernelGGL((k_sweep2<T, Coll, 1>), grid, block, 0, st, g, coll, fi, fo, bc); break;
  default: hipLaunchKernelGGL((k_sweep2<T, Coll, 0>), grid, block, 0, st, g, coll, fi, fo, bc); break;
  }
  return check_launch("two-step sweep kernel");
}

// ---------------------------------------------------------------------------
// Batched sweeps: one launch advances many independent grids.  Every grid (a
// "job") owns a contiguous range of blocks; block_job maps a block to its job
// (-1 for the padding that rounds the grid up to a multiple of the 8 XCDs).  A job
// whose instance is inactive (active[inst] == 0) returns at once, so converged
// instances cost one scalar load per block.
template <typename T, class Coll> struct SweepJob {
  SweepGeom g;
  Coll coll;
  BCDev<T> bc;
  NLDev<T> nl;
  const T *fin;
  T *fout;
  T *fpc;
  const T *res;
  T *coarse;
  double *partials;
  T omg, gm;
  int32_t inst;   // index of the instance's activity flag
  int32_t block0; // first block of this job in the launch
};

// Work-list launches of the batched kernels.  xb_batch_compact lists, per plan, the
// jobs of active instances and the prefix of their block counts; a launch of G
// workgroups (G a multiple of 8, at most what the GPU holds at once) then covers
// exactly those blocks.  Workgroups are dealt round-robin to the 8 XCDs, so
// workgroup i serves XCD i % 8: each XCD takes a contiguous run of the listed blocks
// (neighbouring tiles share halo lines in that XCD's L2) and each of its G / 8
// workgroups every (G / 8)-th block of the run, in order.  Idle instances cost
// nothing: no workgroup is dispatched for them.
struct WorkListDev {
  const int32_t *ajob = nullptr; // active jobs; null: one workgroup per table block
  const int32_t *aoff = nullptr; // prefix of their block counts (naj + 1 entries)
  const int32_t *naj = nullptr;  // number of active jobs
};

template <class F> __device__ __forceinline__ void for_listed_blocks(const WorkListDev &wl, F &&body) {
  const int na = *wl.naj;
  const int total = wl.aoff[na];
  const int x = (int)(blockIdx.x & 7), k = (int)(blockIdx.x >> 3), g8 = (int)(gridDim.x >> 3);
  const int lo = (int)(((int64_t)total * x) >> 3), hi = (int)(((int64_t)total * (x + 1)) >> 3);
  int a = 0;
  for (int w = lo + k; w < hi; w += g8) { // w is uniform per workgroup
    while (wl.aoff[a + 1] <= w) ++a;      // monotone in w
    body(wl.ajob[a], w - wl.aoff[a]);
  }
}

// block `local` of one batched sweep job
template <typename T, class Coll, int MODE, int BC, bool HAS_RES, bool HAS_NL>
__device__ __forceinline__ void sweep_batch_block(const SweepJob<T, Coll> *__restrict__ jobs, int j, int local) {
  const SweepJob<T, Coll> J = jobs[j];
  const int item = local * WAVES_PER_BLOCK + (int)(threadIdx.x >> 6);
  if (item >= J.g.nitems) return; // uniform per wavefront
  if constexpr (BC == 2) {
    __shared__ BCShared<T, Coll::Q> sh[WAVES_PER_BLOCK];
    sweep_item<T, Coll, MODE, BC, HAS_RES, HAS_NL>(J.g, J.coll, J.fin, J.fout, J.fpc, J.res, J.bc, J.omg, J.gm,
                                                   J.coarse, J.partials, J.nl, &sh[threadIdx.x >> 6], item);
  } else {
    sweep_item<T, Coll, MODE, BC, HAS_RES, HAS_NL>(J.g, J.coll, J.fin, J.fout, J.fpc, J.res, J.bc, J.omg, J.gm,
                                                   J.coarse, J.partials, J.nl, nullptr, item);
  }
}

template <typename T, class Coll, int MODE, int BC, bool HAS_RES, bool HAS_NL>
__global__ void __launch_bounds__(BLOCK)
k_sweep_batch(const SweepJob<T, Coll> *__restrict__ jobs, const int32_t *__restrict__ block_job,
              const int32_t *__restrict__ block_inst, const int32_t *__restrict__ active, const WorkListDev wl) {
  if (wl.ajob != nullptr) {
    // the waves of a workgroup share nothing but the rule staging of their own LDS
    // regions, so consecutive blocks need no barrier between them
    for_listed_blocks(wl, [&](int j, int local) { sweep_batch_block<T, Coll, MODE, BC, HAS_RES, HAS_NL>(jobs, j, local); });
    return;
  }
  const int b = xcd_block_id();
  const int j = block_job[b];
  const int i = block_inst[b];
  if (j < 0) return;                               // uniform per block
  if (active != nullptr && active[i] == 0) return; // uniform per block
  sweep_batch_block<T, Coll, MODE, BC, HAS_RES, HAS_NL>(jobs, j, b - jobs[j].block0);
}

template <typename T, class Coll, int MODE>
int launch_sweep(const SweepGeom &g, const Coll &coll, const void *fin, void *fout, void *fpc, const void *res,
                 const xb_bc_table *bct, double gamma, void *coarse, double *partials, hipStream_t st,
                 const NLDev<T> *nlp = nullptr) {
  BCDev<T> bc;
  const bool has_bc = bct != nullptr && bct->cell_slot != nullptr;
  if (has_bc) {
    bc.slot = bct->cell_slot;
    bc.code_pc = bct->code_pc;
    bc.code_ps = bct->code_ps;
    bc.const_pc = static_cast<const T *>(bct->const_pc);
    bc.const_ps = static_cast<const T *>(bct->const_ps);
    bc.nslots = bct->nslots;
  }
  const int bcm = bc_mode(bc.slot, bc.nslots);
  const bool has_res = res != nullptr;
  const dim3 grid(cdiv(g.nitems, WAVES_PER_BLOCK)), block(BLOCK);
  const T omg = T(1.0 - gamma), gm = T(gamma);
  const T *fi = static_cast<const T *>(fin);
  T *fo = static_cast<T *>(fout);
  T *fp = static_cast<T *>(fpc);
  const T *rs = static_cast<const T *>(res);
  T *co = static_cast<T *>(coarse);
  const NLDev<T> nl = nlp != nullptr ? *nlp : NLDev<T>();
#define XB_SWEEP(BC, RES, NL) \
  hipLaunchKernelGGL((k_sweep<T, Coll, MODE, BC, RES, NL>), grid, block, 0, st, g, coll, fi, fo, fp, rs, bc, omg, gm, \
                     co, partials, nl)
#define XB_SWEEP_BC(RES, NL)                                                                                        \
  do {                                                                                                             \
    if (bcm == 2) XB_SWEEP(2, RES, NL);                                                                            \
    else if (bcm == 1) XB_SWEEP(1, RES, NL);                                                                       \
    else XB_SWEEP(0, RES, NL);                                                                                     \
  } while (0)
  if constexpr (Coll::Q == 8) {
    if (nlp != nullptr) {
      if (has_res) XB_SWEEP_BC(true, true);
      else XB_SWEEP_BC(false, true);
      return check_launch("sweep kernel");
    }
  }
  if (has_res) XB_SWEEP_BC(true, false);
  else XB_SWEEP_BC(false, false);
#undef XB_SWEEP_BC
#undef XB_SWEEP
  return check_launch("sweep kernel");
}

// ---------------------------------------------------------------------------
// Element-wise kernels (grid-stride over cells).
inline dim3 ew_grid(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65535 * 8) b = 65535 * 8;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

template <typename T>
__global__ void k_prolong_add(int nx, int ny, T *__restrict__ f, const T *__restrict__ c, int cnx, int cny, int ix0,
                              int ix1, int iy0, int iy1, int rmode) {
  const int nxi = ix1 - ix0, nyi = iy1 - iy0;
  const int64_t n = (int64_t)nxi * nyi;
  const size_t N = (size_t)nx * ny, Nc = (size_t)cnx * cny;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int xi = (int)(t / nyi), yi = (int)(t % nyi);
    const size_t cell = (size_t)(ix0 + xi) * ny + (iy0 + yi);
    const size_t cc = (size_t)(ix0 + (xi >> 1)) * cny + (iy0 + (yi >> 1));
    T v[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) v[d] = f[d * N + cell] + c[d * Nc + cc];
    if (rmode) store_round(v, rmode);
#pragma unroll
    for (int d = 0; d < 8; ++d) f[d * N + cell] = v[d];
  }
}

template <typename T>
__global__ void k_prolong_add_div(int nx, int ny, T *__restrict__ f, const T *__restrict__ c, int cnx, int cny,
                                  int ix0, int ix1, int iy0, int iy1, T k, T *__restrict__ div,
                                  const T *__restrict__ lamf, const T *__restrict__ muf, int rmode) {
  const int nxi = ix1 - ix0, nyi = iy1 - iy0;
  const int64_t n = (int64_t)nxi * nyi;
  const size_t N = (size_t)nx * ny, Nc = (size_t)cnx * cny;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int xi = (int)(t / nyi), yi = (int)(t % nyi);
    const size_t cell = (size_t)(ix0 + xi) * ny + (iy0 + yi);
    const size_t cc = (size_t)(ix0 + (xi >> 1)) * cny + (iy0 + (yi >> 1));
    T v[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) v[d] = f[d * N + cell] + c[d * Nc + cc];
    if (rmode) store_round(v, rmode);
#pragma unroll
    for (int d = 0; d < 8; ++d) f[d * N + cell] = v[d];
    div[cell] = elast_div_of<T>(v, lamf != nullptr ? elast_kdiv<T>(lamf[cell], muf[cell]) : k);
  }
}

template <typename T>
__global__ void k_elast_macro(int64_t N, T cs0, T cd0, const T *__restrict__ f, const T *__restrict__ fa,
                              const T *__restrict__ fb, T *__restrict__ eta, T *__restrict__ sig,
                              const T *__restrict__ lamf, const T *__restrict__ muf) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    T cs = cs0, cd = cd0;
    if (lamf != nullptr) {
      const T lm = lamf[i] + muf[i];
      cs = T(-1.5) * lm / (T(3) * lm + T(1));
      cd = T(-3) * muf[i] / (T(6) * muf[i] + T(1));
    }
    T v[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) v[d] = f[d * N + i];
    T e0 = v[0] - v[2] + v[4] - v[5] - v[6] + v[7];
    T e1 = v[1] - v[3] + v[4] + v[5] - v[6] - v[7];
    if (eta != nullptr) {
      if (fa != nullptr) { e0 = e0 + T(0.5) * fa[i]; e1 = e1 + T(0.5) * fa[N + i]; }
      if (fb != nullptr) { e0 = e0 + T(0.5) * fb[i]; e1 = e1 + T(0.5) * fb[N + i]; }
      eta[i] = e0;
      eta[N + i] = e1;
    }
    if (sig != nullptr) {
      const T ms = (v[0] + v[1] + v[2] + v[3] + T(2) * (v[4] + v[5] + v[6] + v[7])) * cs;
      const T md = (v[0] - v[1] + v[2] - v[3]) * cd;
      const T m11 = v[4] - v[5] + v[6] - v[7];
      sig[i] = ms + md;
      sig[N + i] = (T(2) * cd) * m11;
      sig[2 * N + i] = ms - md;
    }
  }
}

template <typename T>
__global__ void k_elast_div(int64_t N, T k0, const T *__restrict__ f, const T *__restrict__ fs, T *__restrict__ out,
                            const T *__restrict__ lamf, const T *__restrict__ muf) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    T v[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) v[d] = fs != nullptr ? f[d * N + i] - fs[d * N + i] : f[d * N + i];
    const T k = lamf != nullptr ? elast_kdiv<T>(lamf[i], muf[i]) : k0;
    out[i] = elast_div_of<T>(v, k);
  }
}

template <typename T>
__global__ void k_elast_equilibrium(int64_t N, T cs0, T cd0, T h0, T n0, const T *__restrict__ eta,
                                    const T *__restrict__ sig, const T *__restrict__ fa, const T *__restrict__ fb,
                                    T *__restrict__ f, const T *__restrict__ lamf, const T *__restrict__ muf) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    T cs = cs0, cd = cd0, h = h0, n = n0;
    if (lamf != nullptr) {
      T r0, r1;
      elast_coef<T>(lamf[i], muf[i], r0, r1, h, n);
      cs = T(-1) - T(1) / (T(3) * (lamf[i] + muf[i]));
      cd = T(-1) - T(1) / (T(6) * muf[i]);
    }
    T e0 = eta[i], e1 = eta[N + i];
    if (fa != nullptr) { e0 = e0 - T(0.5) * fa[i]; e1 = e1 - T(0.5) * fa[N + i]; }
    if (fb != nullptr) { e0 = e0 - T(0.5) * fb[i]; e1 = e1 - T(0.5) * fb[N + i]; }
    const T sxx = sig[i], sxy = sig[N + i], syy = sig[2 * N + i];
    const T M2 = (sxx + syy) * cs, M3 = (sxx - syy) * cd, M4 = sxy * cd;
    const T M5 = e0 / T(3), M6 = e1 / T(3);
    const T hM2 = h * M2, nM2 = n * M2;
    f[0 * N + i] = T(0.5) * e0 + hM2 + T(0.25) * M3 - T(0.5) * M5;
    f[1 * N + i] = T(0.5) * e1 + hM2 - T(0.25) * M3 - T(0.5) * M6;
    f[2 * N + i] = -T(0.5) * e0 + hM2 + T(0.25) * M3 + T(0.5) * M5;
    f[3 * N + i] = -T(0.5) * e1 + hM2 - T(0.25) * M3 + T(0.5) * M6;
    f[4 * N + i] = nM2 + T(0.25) * M4 + T(0.25) * M5 + T(0.25) * M6;
    f[5 * N + i] = nM2 - T(0.25) * M4 - T(0.25) * M5 + T(0.25) * M6;
    f[6 * N + i] = nM2 + T(0.25) * M4 - T(0.25) * M5 - T(0.25) * M6;
    f[7 * N + i] = nM2 - T(0.25) * M4 + T(0.25) * M5 - T(0.25) * M6;
  }
}

template <typename T>
__global__ void k_darcy_macro(int64_t N, T m3om, const T *__restrict__ omf, const T *__restrict__ f,
                              const T *__restrict__ sa, const T *__restrict__ sb, T *__restrict__ p,
                              T *__restrict__ grad) {
  using C = DarcyColl<T>;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    T v[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) v[d] = f[d * N + i];
    if (p != nullptr) {
      T rho = v[0];
#pragma unroll
      for (int d = 1; d < 9; ++d) rho = rho + v[d];
      if (sa != nullptr && sb != nullptr) rho = rho + T(0.5) * (sa[i] + sb[i]);
      else if (sa != nullptr) rho = rho + T(0.5) * sa[i];
      else if (sb != nullptr) rho = rho + T(0.5) * sb[i];
      p[i] = rho;
    }
    if (grad != nullptr) {
      if (omf == nullptr) {
        T g0 = T(0), g1 = T(0);
#pragma unroll
        for (int d = 0; d < 9; ++d) {
          if (C::cx(d) != 0) g0 = g0 + v[d] * (C::cx(d) > 0 ? m3om : -m3om);
          if (C::cy(d) != 0) g1 = g1 + v[d] * (C::cy(d) > 0 ? m3om : -m3om);
        }
        grad[i] = g0;
        grad[N + i] = g1;
      } else {
        T g0 = T(0), g1 = T(0);
#pragma unroll
        for (int d = 0; d < 9; ++d) {
          if (C::cx(d) != 0) g0 = C::cx(d) > 0 ? g0 + v[d] : g0 - v[d];
          if (C::cy(d) != 0) g1 = C::cy(d) > 0 ? g1 + v[d] : g1 - v[d];
        }
        const T s = T(-3) * omf[i];
        grad[i] = g0 * s;
        grad[N + i] = g1 * s;
      }
    }
  }
}

// out = a*base + b * sum_d qs_d * ((-3 scale) w_d) c_d, qs_d(x) = q(x - c_d) or the rule of the cell.
template <typename T>
__global__ void k_darcy_gradient(int nx, int ny, const T *__restrict__ q, T m3s, const BCDev<T> bc,
                                 const T *__restrict__ base, T a, T b, T *__restrict__ out) {
  using C = DarcyColl<T>;
  const int64_t N = (int64_t)nx * ny;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i / ny), y = (int)(i % ny);
    T qs[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) qs[d] = q[(size_t)wrapi(x - C::cx(d), nx) * ny + wrapi(y - C::cy(d), ny)];
    if (bc.slot != nullptr) {
      const int s = bc.slot[i];
      if (s >= 0) {
        T src[9];
#pragma unroll
        for (int d = 0; d < 9; ++d) src[d] = qs[0];
        apply_rules<T, 9>(qs, src, bc.code_ps + (size_t)s * 9, bc.const_ps + (size_t)s * 9);
      }
    }
    T g0, g1;
    stencil_grad<T>(qs, m3s, g0, g1);
    if (base != nullptr) {
      out[i] = a * base[i] + b * g0;
      out[N + i] = a * base[N + i] + b * g1;
    } else {
      out[i] = b * g0;
      out[N + i] = b * g1;
    }
  }
}

template <typename T>
__global__ void k_lincomb(int64_t n, T a, const T *__restrict__ x, T b, const T *__restrict__ y, T *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = y != nullptr ? a * x[i] + b * y[i] : a * x[i];
}

template <typename T> __global__ void k_fill(int64_t n, T v, T *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = v;
}

// One partial per block, fixed order inside the block (deterministic for a fixed grid).
template <typename T>
__global__ void __launch_bounds__(256) k_sqdiff(int64_t n, const T *__restrict__ a, const T *__restrict__ b,
                                                double *__restrict__ partials) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double d = b != nullptr ? (double)a[i] - (double)b[i] : (double)a[i];
    s += d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// Sum of p[0..n) by one block of NORM_THREADS threads in a fixed order: thread t
// adds p[t + k * NORM_THREADS] into four interleaved accumulators (k mod 4), which
// are combined, then the block's values are summed pairwise in LDS.  Every norm of
// residual partials goes through this one function, so single-instance and batched
// convergence tests see bit-identical norms.  Call from all threads of the block.
constexpr int NORM_THREADS = 1024;
__device__ __forceinline__ double block_sum(const double *__restrict__ p, int n, double *red) {
  const int t = threadIdx.x;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int i = t;
  for (; i + 3 * NORM_THREADS < n; i += 4 * NORM_THREADS) {
    a0 += p[i];
    a1 += p[i + NORM_THREADS];
    a2 += p[i + 2 * NORM_THREADS];
    a3 += p[i + 3 * NORM_THREADS];
  }
  for (; i < n; i += NORM_THREADS) a0 += p[i];
  __syncthreads(); // red may still be read by a previous call
  red[t] = (a0 + a1) + (a2 + a3);
  __syncthreads();
#pragma unroll
  for (int off = NORM_THREADS / 2; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  return red[0];
}

__global__ void __launch_bounds__(NORM_THREADS) k_norm_finalize(const double *__restrict__ partials, int n,
                                                                double *__restrict__ out) {
  __shared__ double red[NORM_THREADS];
  const double s = block_sum(partials, n, red);
  if (threadIdx.x == 0) {
    out[0] = sqrt(s);
    out[1] = (isfinite(s) ? 0.0 : 1.0);
  }
}

template <typename T> BCDev<T> bcdev(const xb_bc_table *t) {
  BCDev<T> b;
  if (t != nullptr && t->cell_slot != nullptr) {
    b.slot = t->cell_slot;
    b.code_pc = t->code_pc;
    b.code_ps = t->code_ps;
    b.const_pc = static_cast<const T *>(t->const_pc);
    b.const_ps = static_cast<const T *>(t->const_ps);
    b.nslots = t->nslots;
  }
  return b;
}

// ---------------------------------------------------------------------------
// Several plain steps per launch on LDS tiles (temporal blocking).  A workgroup owns
// a TILE x TILE region of cells (periodic wrap at the grid edges), keeps the
// populations of its cells in registers and runs `steps` steps on them: every thread
// collides its cells and applies their post-collision rules, writes the
// post-collision populations to LDS and, after a barrier, pulls its cells' streamed
// populations from the neighbours' LDS slots and applies the post-streaming rules.
// A cell on the region's edge pulls from itself where its neighbour lies outside the
// region; that error moves inward one cell per step, so after `steps` steps the inner
// (TILE - 2 steps)^2 cells are exact, and those are written.  Each population is read
// and written once per launch; per cell the arithmetic is that of `steps` k_sweep
// STEP launches (bit-identical).  A grid resident in the 256 MB last-level cache
// (the 514^2 Darcy grid is 19 MB) is bound by the latency of its launches, not by
// HBM: fewer, larger launches are the speed-up.
constexpr int TILE = 32;
constexpr int TILE_THREADS = 256;
constexpr int TILE_CPT = TILE * TILE / TILE_THREADS;
// Multi-step tiles also run as blocks of 512 threads (2 cells each): fewer registers
// per thread and twice the waves per tile, for grids of a few tiles per CU, where a
// tile's latency is the launch time.  Blocks of 256 threads (4 cells each) keep two
// blocks per CU (their register budget is capped at 256), whose loads and barriers
// overlap on grids of many tiles per CU.
constexpr int TILE_THREADS_WIDE = 512;
// One-wave tiles: 44 x 48 cells in blocks of 704 threads (3 cells each), the largest
// region whose D2Q9 FP64 populations and rule table fit the 160 KB of LDS of a CU.  A
// grid of up to about 16 x 15 such tiles (the 514^2 C2 grid at 6 steps per launch:
// 255 tiles) runs as a single wave of blocks on the 256 CUs: every launch costs one
// tile's load, steps and store instead of two or more waves of 32 x 32 tiles.
constexpr int TILE_ONE_WAVE_X = 44, TILE_ONE_WAVE_Y = 48, TILE_THREADS_ONE_WAVE = 704;
constexpr int TILE_MAX_STEPS = 8;
// LDS slots on either side of a tile's population planes: the largest offset of an
// upstream cell in LDS order (one row plus one cell).
#define TILE_LDS_PAD(TY) ((TY) + 1)

__device__ __forceinline__ int wrapm(int i, int n) {
  const int r = i % n;
  return r < 0 ? r + n : r;
}

// A tile cell's coordinate x along one axis (the tile's origin plus the cell's offset)
// wrapped onto [0, n).  On a grid at least a tile wide along that axis (`big`) every
// such x lies in [-n, 2n), because the tile's halo is narrower than the tile, and one
// conditional add or subtract wraps it; on smaller grids (coarse levels) the remainder
// does.  `big` is uniform over the block, so the choice is one branch per tile.
__device__ __forceinline__ int wrap_axis(int x, int n, bool big) { return big ? wrapi(x, n) : wrapm(x, n); }

// Rules of one cell without branches (the tile kernel unrolls its cells, and the
// branchy form of apply_rules would not fit the instruction cache): the same values
// as apply_rules, with the rule sources read from the cell's LDS slots (source
// population d at own[d * TT]) instead of selected among registers.
template <typename T, int Q, int TT>
__device__ __forceinline__ void rules_lds(T (&v)[Q], const T *own, const int8_t *code, const T *cst) {
  T out[Q];
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    const int c = code[d];
    const int k = min(max((c > 0 ? c : -c) - 1, 0), Q - 1);
    const T s = own[k * TT];
    T base = c > 0 ? s : -s;
    base = c == XB_RULE_ZERO ? T(0) : base;
    base = c == 0 ? v[d] : base;
    out[d] = base + cst[d];
  }
#pragma unroll
  for (int d = 0; d < Q; ++d) v[d] = out[d];
}

// The rule table of a tile in LDS, with one more row of identity rules (codes 0,
// constants 0) for the cells without rules: a tile holding any wall cell applies the
// rules of all its cells in straight-line code, so that the four cells of a thread
// share their LDS round trips (per-cell branches would serialize them).
template <typename T, int Q> struct TileRules {
  int8_t code_pc[(LDS_BC_SLOTS + 1) * Q], code_ps[(LDS_BC_SLOTS + 1) * Q];
  T const_pc[(LDS_BC_SLOTS + 1) * Q], const_ps[(LDS_BC_SLOTS + 1) * Q];
};

// Pull plan of one tile cell: for every direction the LDS slot its streamed
// population comes from (the neighbour's slot, or the cell's own slot of the source
// population where a post-streaming rule replaces it) and a sign (0: the rule gives
// zero, 1: +, 2: -), 16 bits per direction.  Rules are fixed per cell, so the plan is
// built once per launch and the rules cost nothing in the steps but the constants of
// rule cells.
template <int Q> struct PullPlan {
  uint32_t w[(Q + 1) / 2];
};

template <class Lat, int Q, int TX, int TY>
__device__ __forceinline__ PullPlan<Q> make_pull_plan(int c, const int8_t *code) {
  constexpr int TT = TX * TY;
  static_assert(Q * TT <= (1 << 14), "pull plan offsets are 14 bits");
  PullPlan<Q> p;
  const int rx = c / TY, ry = c % TY;
#pragma unroll
  for (int i = 0; i < (Q + 1) / 2; ++i) p.w[i] = 0u;
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    const int sx = rx - Lat::cx(d), sy = ry - Lat::cy(d);
    const bool in = sx >= 0 && sx < TX && sy >= 0 && sy < TY;
    uint32_t off = (uint32_t)(d * TT + (in ? sx * TY + sy : c)), sg = 1u;
    const int cc = code != nullptr ? code[d] : 0;
    if (cc == XB_RULE_ZERO) {
      sg = 0u;
    } else if (cc != 0) {
      off = (uint32_t)(((cc > 0 ? cc : -cc) - 1) * TT + c);
      sg = cc > 0 ? 1u : 2u;
    }
    p.w[d >> 1] |= (off | (sg << 14)) << ((d & 1) * 16);
  }
  return p;
}

template <typename T, int Q>
__device__ __forceinline__ void pull_with_plan(const T *lds, const PullPlan<Q> &p, T (&v)[Q]) {
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    const uint32_t e = (p.w[d >> 1] >> ((d & 1) * 16)) & 0xFFFFu;
    const T x = lds[e & 0x3FFFu];
    const uint32_t sg = e >> 14;
    v[d] = sg == 1u ? x : (sg == 2u ? -x : T(0));
  }
}

// Streaming inside a tile: population d of cell c comes from the upstream cell's
// slot, or from the cell's own slot at the tile border (a halo cell, whose value is
// not written out).
// The upstream cell of a border cell lies outside the tile; the pull then reads the
// neighbouring slot in LDS order (another row, or the adjacent direction's plane, or
// the padding of TILE_LDS_PAD(TY) slots on either side of the array).  Any value will
// do there: a border cell is in the tile's halo, whose error moves inward one cell per
// step and never reaches the written inner cells, so the pull needs no bounds test
// and every direction is one LDS read at a constant offset from the cell's own slot.
template <typename T, class Lat, int TX, int TY>
__device__ __forceinline__ void tile_pull(const T *lds, int c, T (&v)[Lat::Q]) {
  constexpr int TT = TX * TY;
#pragma unroll
  for (int d = 0; d < Lat::Q; ++d) v[d] = lds[d * TT + c - (Lat::cx(d) * TY + Lat::cy(d))];
}

// The steps of one tile.  Blocks of 512 threads (PLAN): tiles that hold a rule cell of
// a table in LDS run their own copy of the loop (WALLS), where the post-collision
// rules read their sources through the cells' own LDS slots and the post-streaming
// rules are folded into a per-cell pull plan; the other tiles run the plain loop, in
// which no plan registers are live.  Blocks of 256 threads (four cells per thread,
// whose register budget has no room for plans): one loop for all tiles, in which the
// tiles that hold a rule cell (``walls``, uniform per block) apply both kinds of
// rules through the cells' own LDS slots.  The plain pull's LDS addresses are
// immediate offsets from the cell's own.  BC == 1 (a table in global memory): per-cell
// rules through the own slots.
template <typename T, class Coll, int BC, bool WALLS, int TX, int TY, int NT, int CPT = TX * TY / NT>
__device__ __forceinline__ void tile_step_loop(int steps, const Coll &coll, T (&f)[CPT][Coll::Q],
                                               const typename Coll::Aux (&aux)[CPT], const int (&slot)[CPT],
                                               bool walls, const BCDev<T> &bc, T *lds, TileRules<T, Coll::Q> *sh) {
  constexpr int Q = Coll::Q;
  constexpr int TT = TX * TY;
  PullPlan<Q> plan[WALLS ? CPT : 1];
  int cslot[WALLS ? CPT : 1];
  int row[CPT]; // BC == 2: the cell's rule row (cells without rules: the identity row)
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    row[k] = slot[k] >= 0 ? slot[k] : bc.nslots;
    if constexpr (WALLS) {
      const int c = threadIdx.x + k * NT;
      const int8_t *code = (slot[k] >= 0 && bc.code_ps != nullptr) ? sh->code_ps + slot[k] * Q : nullptr;
      plan[k] = make_pull_plan<Coll, Q, TX, TY>(c, code);
      cslot[k] = code != nullptr ? slot[k] : -1;
    }
  }
  for (int s = 0; s < steps; ++s) {
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int c = threadIdx.x + k * NT;
      T pc[Q];
      coll.collide(f[k], aux[k], pc);
      // post-collision rules: their sources, the cell's pre-collision populations, go
      // through its own LDS slots (only this thread touches them before the barrier)
      if constexpr (BC == 2) {
        if ((WALLS || walls) && bc.code_pc != nullptr) {
#pragma unroll
          for (int d = 0; d < Q; ++d) lds[d * TT + c] = f[k][d];
          rules_lds<T, Q, TT>(pc, lds + c, sh->code_pc + row[k] * Q, sh->const_pc + row[k] * Q);
        }
      } else if constexpr (BC == 1) {
        const int sl = slot[k];
        if (sl >= 0 && bc.code_pc != nullptr) {
#pragma unroll
          for (int d = 0; d < Q; ++d) lds[d * TT + c] = f[k][d];
          rules_lds<T, Q, TT>(pc, lds + c, bc.code_pc + (size_t)sl * Q, bc.const_pc + (size_t)sl * Q);
        }
      }
#pragma unroll
      for (int d = 0; d < Q; ++d) lds[d * TT + c] = pc[d];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int c = threadIdx.x + k * NT;
      if constexpr (WALLS) {
        // the post-streaming rules are folded into the pull; rule cells add their constants
        pull_with_plan<T, Q>(lds, plan[k], f[k]);
        if (cslot[k] >= 0) {
          const T *cst = sh->const_ps + cslot[k] * Q;
#pragma unroll
          for (int d = 0; d < Q; ++d) f[k][d] = f[k][d] + cst[d];
        }
      } else {
        T v[Q];
        tile_pull<T, Coll, TX, TY>(lds, c, v);
        if constexpr (BC == 2) {
          if (walls && bc.code_ps != nullptr)
            rules_lds<T, Q, TT>(v, lds + c, sh->code_ps + row[k] * Q, sh->const_ps + row[k] * Q);
        } else if constexpr (BC == 1) {
          const int sl = slot[k];
          if (sl >= 0 && bc.code_ps != nullptr)
            rules_lds<T, Q, TT>(v, lds + c, bc.code_ps + (size_t)sl * Q, bc.const_ps + (size_t)sl * Q);
        }
#pragma unroll
        for (int d = 0; d < Q; ++d) f[k][d] = v[d];
      }
    }
    __syncthreads();
  }
}

template <typename T, class Coll, int BC, int TX, int TY, int NT>
__device__ __forceinline__ void tile_steps_body(int nx, int ny, int steps, int tiles_y, const Coll &coll,
                                                const T *__restrict__ fin, T *__restrict__ fout, const BCDev<T> &bc,
                                                T *lds, TileRules<T, Coll::Q> *sh) {
  constexpr int Q = Coll::Q;
  static_assert(TX * TY % NT == 0, "every thread owns the same number of cells");
  constexpr int CPT = TX * TY / NT;
  constexpr bool HAS_BC = BC != 0;
  using Aux = typename Coll::Aux;
  const int tile = xcd_block_id();
  const int bx = tile / tiles_y, by = tile % tiles_y;
  const int x0 = bx * (TX - 2 * steps) - steps, y0 = by * (TY - 2 * steps) - steps;
  const size_t N = (size_t)nx * ny;
  if constexpr (BC == 2) {
    // the (small, deduplicated) rule table into LDS by the whole block, then the
    // identity row
    const int n = bc.nslots * Q;
    for (int k = threadIdx.x; k < n + Q; k += blockDim.x) {
      const bool id = k >= n;
      sh->code_pc[k] = (!id && bc.code_pc != nullptr) ? bc.code_pc[k] : int8_t(0);
      sh->const_pc[k] = (!id && bc.const_pc != nullptr) ? bc.const_pc[k] : T(0);
      sh->code_ps[k] = (!id && bc.code_ps != nullptr) ? bc.code_ps[k] : int8_t(0);
      sh->const_ps[k] = (!id && bc.const_ps != nullptr) ? bc.const_ps[k] : T(0);
    }
  }
  T f[CPT][Q];
  Aux aux[CPT];
  int slot[CPT];
  const bool big = nx >= TX && ny >= TY; // the halo (steps) is below the tile size
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int c = threadIdx.x + k * NT;
    const int gx = wrap_axis(x0 + c / TY, nx, big), gy = wrap_axis(y0 + c % TY, ny, big);
    const size_t cell = (size_t)gx * ny + gy;
#pragma unroll
    for (int d = 0; d < Q; ++d) f[k][d] = fin[d * N + cell];
    aux[k] = coll.load_aux(cell);
    slot[k] = HAS_BC ? bc.slot[cell] : -1;
  }
  if constexpr (BC == 2) {
    // whether this tile holds a cell with rules (uniform over the block; the barrier
    // also orders the table staging above before its reads)
    bool mine = false;
#pragma unroll
    for (int k = 0; k < CPT; ++k) mine = mine || slot[k] >= 0;
    const bool walls = __syncthreads_or(mine) != 0;
    if constexpr (NT == TILE_THREADS_WIDE) {
      if (walls)
        tile_step_loop<T, Coll, 2, true, TX, TY, NT>(steps, coll, f, aux, slot, true, bc, lds, sh);
      else
        tile_step_loop<T, Coll, 2, false, TX, TY, NT>(steps, coll, f, aux, slot, false, bc, lds, sh);
    } else {
      tile_step_loop<T, Coll, 2, false, TX, TY, NT>(steps, coll, f, aux, slot, walls, bc, lds, sh);
    }
  } else {
    tile_step_loop<T, Coll, BC, false, TX, TY, NT>(steps, coll, f, aux, slot, false, bc, lds, sh);
  }
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int c = threadIdx.x + k * NT;
    const int rx = c / TY, ry = c % TY;
    if (rx < steps || rx >= TX - steps || ry < steps || ry >= TY - steps) continue;
    const int gx = x0 + rx, gy = y0 + ry;
    if (gx >= nx || gy >= ny) continue;
    const size_t cell = (size_t)gx * ny + gy;
#pragma unroll
    for (int d = 0; d < Q; ++d) fout[d * N + cell] = f[k][d];
  }
}

template <typename T, class Coll, int BC, int TX, int TY, int NT>
__global__ void __launch_bounds__(NT, NT == TILE_THREADS ? 2 : 1)
k_tile_steps(int nx, int ny, int steps, int tiles_y, const Coll coll, const T *__restrict__ fin, T *__restrict__ fout,
             const BCDev<T> bc) {
  __shared__ T lds_raw[Coll::Q * TX * TY + 2 * TILE_LDS_PAD(TY)];
  T *lds = lds_raw + TILE_LDS_PAD(TY);
  if constexpr (BC == 2) {
    __shared__ TileRules<T, Coll::Q> sh;
    tile_steps_body<T, Coll, BC, TX, TY, NT>(nx, ny, steps, tiles_y, coll, fin, fout, bc, lds, &sh);
  } else {
    tile_steps_body<T, Coll, BC, TX, TY, NT>(nx, ny, steps, tiles_y, coll, fin, fout, bc, lds, nullptr);
  }
}

// Tile shape of a multi-step launch from the asked threads per block: 256, 512 (32 x 32
// tiles) or 704 (44 x 48); otherwise 512 for grids of at most four 32 x 32 tiles per
// CU (256 CUs), else 256.
inline int tile_block_threads(int asked, int nx, int ny, int steps) {
  if (asked == TILE_THREADS || asked == TILE_THREADS_WIDE || asked == TILE_THREADS_ONE_WAVE) return asked;
  const int tiles = cdiv(nx, TILE - 2 * steps) * cdiv(ny, TILE - 2 * steps);
  return tiles <= 4 * 256 ? TILE_THREADS_WIDE : TILE_THREADS;
}

template <typename T, class Coll, int TX, int TY, int NT>
void launch_tile_steps_shape(int nx, int ny, int steps, const Coll &coll, const T *fi, T *fo, const BCDev<T> &bc,
                             hipStream_t st) {
  const int tx = cdiv(nx, TX - 2 * steps), ty = cdiv(ny, TY - 2 * steps);
  const dim3 grid((unsigned)(tx * ty)), block(NT);
  switch (bc_mode(bc.slot, bc.nslots)) {
  case 2:
    hipLaunchKernelGGL((k_tile_steps<T, Coll, 2, TX, TY, NT>), grid, block, 0, st, nx, ny, steps, ty, coll, fi, fo, bc);
    break;
  case 1:
    hipLaunchKernelGGL((k_tile_steps<T, Coll, 1, TX, TY, NT>), grid, block, 0, st, nx, ny, steps, ty, coll, fi, fo, bc);
    break;
  default:
    hipLaunchKernelGGL((k_tile_steps<T, Coll, 0, TX, TY, NT>), grid, block, 0, st, nx, ny, steps, ty, coll, fi, fo, bc);
    break;
  }
}

template <typename T, class Coll>
int launch_tile_steps(int nx, int ny, int steps, const Coll &coll, const void *fin, void *fout,
                      const xb_bc_table *bct, int threads, hipStream_t st) {
  const BCDev<T> bc = bcdev<T>(bct);
  const T *fi = static_cast<const T *>(fin);
  T *fo = static_cast<T *>(fout);
  switch (tile_block_threads(threads, nx, ny, steps)) {
  case TILE_THREADS_ONE_WAVE:
    launch_tile_steps_shape<T, Coll, TILE_ONE_WAVE_X, TILE_ONE_WAVE_Y, TILE_THREADS_ONE_WAVE>(nx, ny, steps, coll, fi,
                                                                                              fo, bc, st);
    break;
  case TILE_THREADS_WIDE:
    launch_tile_steps_shape<T, Coll, TILE, TILE, TILE_THREADS_WIDE>(nx, ny, steps, coll, fi, fo, bc, st);
    break;
  default:
    launch_tile_steps_shape<T, Coll, TILE, TILE, TILE_THREADS>(nx, ny, steps, coll, fi, fo, bc, st);
    break;
  }
  return check_launch("multi-step tile kernel");
}

int check_dtype(int32_t dtype) {
  if (dtype != XB_F64 && dtype != XB_F32) return fail(XB_EUNSUPPORTED, "unsupported dtype %d", dtype);
  return XB_OK;
}

// A storage-rounding dtype as its arithmetic dtype and rounding mode (0 none, 1 to
// FP32, 2 to FP16); other dtypes pass unchanged with mode 0.
int split_storage(int32_t dtype, int32_t &arith) {
  switch (dtype) {
  case XB_F64_S32: arith = XB_F64; return 1;
  case XB_F64_S16: arith = XB_F64; return 2;
  case XB_F32_S16: arith = XB_F32; return 2;
  default: arith = dtype; return 0;
  }
}

int check_bc(const xb_bc_table *t, bool need_ps) {
  if (t == nullptr) return XB_OK;
  if (t->nslots < 0) return fail(XB_EINVAL, "bc table: negative slot count");
  if (t->nslots > 0 && t->cell_slot == nullptr) return fail(XB_EINVAL, "bc table: null cell_slot");
  if (t->code_pc != nullptr && t->const_pc == nullptr) return fail(XB_EINVAL, "bc table: null const_pc");
  if (t->code_ps != nullptr && t->const_ps == nullptr) return fail(XB_EINVAL, "bc table: null const_ps");
  if (need_ps && t->cell_slot != nullptr && t->code_ps == nullptr) return fail(XB_EINVAL, "bc table: null code_ps");
  return XB_OK;
}

int elast_geom(const xb_elast_args *a, SweepGeom &g) {
  if (a->mode == XB_MODE_RESIDUAL)
    g = make_geom(a->nx, a->ny, a->ix0, a->iy0, a->ix1 - a->ix0, a->iy1 - a->iy0, a->ix0, a->ix1, a->iy0, a->iy1,
                  a->rows);
  else
    g = make_geom(a->nx, a->ny, 0, 0, a->nx, a->ny, a->ix0, a->ix1, a->iy0, a->iy1, a->rows, STRIP, 1);
  return XB_OK;
}

template <typename T> void elast_coll(const xb_elast_args *a, ElastColl<T> &c, NLDev<T> &nl) {
  const double lm = a->lam + a->mu;
  double g = 1.0 / (12.0 * lm - 4.0);
  if (std::fabs(g) > 100.0) g = 0.0;
  c.cs = T((3 * lm - 1) / (3 * lm + 1));
  c.cd = T((6 * a->mu - 1) / (6 * a->mu + 1));
  c.h = T(0.25 + 0.5 * g);
  c.n = T(-0.25 * g);
  c.fa = static_cast<const T *>(a->force_a);
  c.fb = static_cast<const T *>(a->force_b);
  c.N = (size_t)a->nx * a->ny;
  c.has_b = a->force_b != nullptr || a->div_in != nullptr;
  c.lamf = static_cast<const T *>(a->lam_field);
  c.muf = static_cast<const T *>(a->mu_field);
  nl = NLDev<T>();
  if (a->div_in != nullptr) {
    nl.div_in = static_cast<const T *>(a->div_in);
    nl.div_out = static_cast<T *>(a->div_out);
    nl.m3s = T(-3 * a->nl_scale);
    nl.kdiv = T(-1.5 / (3 * (a->lam + a->mu) + 1));
    nl.gbc = bcdev<T>(a->grad_bc);
    nl.base = static_cast<const T *>(a->force_b);
  }
}

template <typename T> int elast_step_t(const xb_elast_args *a, hipStream_t st, int rmode = 0) {
  ElastColl<T> c;
  NLDev<T> nl;
  elast_coll<T>(a, c, nl);
  const NLDev<T> *nlp = a->div_in != nullptr ? &nl : nullptr;
  SweepGeom geo;
  elast_geom(a, geo);
  geo.rmode = rmode;
  switch (a->mode) {
  case XB_MODE_STEP:
    return launch_sweep<T, ElastColl<T>, XB_MODE_STEP>(geo, c, a->f_in, a->f_out, a->fpc_out, a->resid, a->bc, 1.0,
                                                       nullptr, nullptr, st, nlp);
  case XB_MODE_SMOOTH:
    return launch_sweep<T, ElastColl<T>, XB_MODE_SMOOTH>(geo, c, a->f_in, a->f_out, nullptr, a->resid, a->bc,
                                                         a->gamma, nullptr, nullptr, st, nlp);
  default:
    return launch_sweep<T, ElastColl<T>, XB_MODE_RESIDUAL>(geo, c, a->f_in, nullptr, nullptr, a->resid, a->bc, 1.0,
                                                           a->coarse_out, a->norm_partials, st, nlp);
  }
}

// Argument checks of one elasticity sweep (xb_elast_step and every job of a batch).
// smoothn: the tile smoothing of xb_elast_smoothn, which may inject a coarse correction
// (coarse_out in SMOOTH mode) and start from the zero state (null f_in).
int validate_elast(const xb_elast_args *a, const char *who, bool smoothn = false) {
  if (a == nullptr) return fail(XB_EINVAL, "%s: null args", who);
  if (int r = check_dtype(a->dtype)) return r;
  if (a->nx <= 0 || a->ny <= 0) return fail(XB_EINVAL, "%s: bad grid %d x %d", who, a->nx, a->ny);
  if (a->f_in == nullptr && !smoothn) return fail(XB_EINVAL, "%s: null f_in", who);
  if (a->mode < XB_MODE_STEP || a->mode > XB_MODE_RESIDUAL) return fail(XB_EINVAL, "%s: bad mode %d", who, a->mode);
  if (a->ix0 < 0 || a->ix1 > a->nx || a->iy0 < 0 || a->iy1 > a->ny || a->ix0 >= a->ix1 || a->iy0 >= a->iy1)
    return fail(XB_EINVAL, "%s: bad interior box [%d,%d)x[%d,%d)", who, a->ix0, a->ix1, a->iy0, a->iy1);
  if (a->mode != XB_MODE_RESIDUAL && a->f_out == nullptr) return fail(XB_EINVAL, "%s: null f_out", who);
  if (a->mode != XB_MODE_STEP && a->fpc_out != nullptr)
    return fail(XB_EINVAL, "%s: fpc_out is only produced by XB_MODE_STEP", who);
  if (a->mode == XB_MODE_RESIDUAL) {
    if (a->coarse_out == nullptr || a->norm_partials == nullptr)
      return fail(XB_EINVAL, "%s: residual mode needs coarse_out and norm_partials", who);
    if (((a->ix1 - a->ix0) & 1) || ((a->iy1 - a->iy0) & 1))
      return fail(XB_EINVAL, "%s: interior box must have even extents for restriction", who);
  }
  if (a->mode != XB_MODE_RESIDUAL && a->coarse_out != nullptr && !smoothn)
    return fail(XB_EINVAL, "%s: coarse_out is written in residual mode; in smoothing mode only xb_elast_smoothn "
                "reads it (as a coarse correction)", who);
  // (the tile smoothing takes no sweep launch shape: a batch job's rows is its steps)
  if (!smoothn && a->rows != 0 && (a->rows < 1 || a->rows > 64 || (a->rows & (a->rows - 1))))
    return fail(XB_EINVAL, "%s: rows must be 0 or a power of two in [1, 64]", who);
  if (a->mode == XB_MODE_RESIDUAL && a->rows == 1)
    return fail(XB_EINVAL, "%s: the residual sweep restricts row pairs (rows >= 2)", who);
  if (a->mode == XB_MODE_SMOOTH && !(a->gamma >= 0.0 && a->gamma <= 1.0))
    return fail(XB_EINVAL, "%s: gamma must be in [0,1]", who);
  if (int r = check_bc(a->bc, false)) return r;
  if (a->div_in == nullptr && (a->div_out != nullptr || a->grad_bc != nullptr))
    return fail(XB_EINVAL, "%s: div_out / grad_bc without div_in", who);
  if (a->div_in != nullptr) {
    if (int r = check_bc(a->grad_bc, true)) return r;
  }
  if ((a->lam_field == nullptr) != (a->mu_field == nullptr))
    return fail(XB_EINVAL, "%s: lam_field and mu_field go together", who);
  return XB_OK;
}

// ---------------------------------------------------------------------------
// Batched element-wise kernels: as the sweep batch, a job owns `nblocks`
// consecutive blocks and walks its cells with a block-strided loop.
template <typename T> struct ProlongJob {
  int nx, ny, cnx, cny, ix0, ix1, iy0, iy1;
  T k;     // divergence factor (used when div is set)
  T *f;
  const T *c;
  T *div;
  int32_t inst, block0, nblocks;
};

template <typename T> __device__ __forceinline__ void prolong_batch_block(const ProlongJob<T> &J, int local);

template <typename T>
__global__ void __launch_bounds__(256)
k_prolong_batch(const ProlongJob<T> *__restrict__ jobs, const int32_t *__restrict__ block_job,
                const int32_t *__restrict__ block_inst, const int32_t *__restrict__ active, const WorkListDev wl) {
  if (wl.ajob != nullptr) {
    for_listed_blocks(wl, [&](int j, int local) { prolong_batch_block<T>(jobs[j], local); });
    return;
  }
  const int b = blockIdx.x;
  const int j = block_job[b];
  const int i = block_inst[b];
  if (j < 0) return;
  if (active != nullptr && active[i] == 0) return;
  prolong_batch_block<T>(jobs[j], b - jobs[j].block0);
}

// block `local` of one batched prolongation job (a grid-stride loop over the job's
// fine interior with the job's block count as the stride)
template <typename T> __device__ __forceinline__ void prolong_batch_block(const ProlongJob<T> &Jr, int local) {
  const ProlongJob<T> J = Jr;
  const int nxi = J.ix1 - J.ix0, nyi = J.iy1 - J.iy0;
  const int64_t n = (int64_t)nxi * nyi;
  const size_t N = (size_t)J.nx * J.ny, Nc = (size_t)J.cnx * J.cny;
  for (int64_t t = (int64_t)local * 256 + threadIdx.x; t < n; t += (int64_t)J.nblocks * 256) {
    const int xi = (int)(t / nyi), yi = (int)(t % nyi);
    const size_t cell = (size_t)(J.ix0 + xi) * J.ny + (J.iy0 + yi);
    const size_t cc = (size_t)(J.ix0 + (xi >> 1)) * J.cny + (J.iy0 + (yi >> 1));
    T v[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      v[d] = J.f[d * N + cell] + J.c[d * Nc + cc];
      J.f[d * N + cell] = v[d];
    }
    if (J.div != nullptr) J.div[cell] = elast_div_of<T>(v, J.k);
  }
}

template <typename T> struct FillJob {
  T *out;
  int64_t n;
  T value;
  int32_t inst, block0, nblocks;
};

template <typename T> __device__ __forceinline__ void fill_batch_block(const FillJob<T> &Jr, int local) {
  const FillJob<T> J = Jr;
  for (int64_t i = (int64_t)local * 256 + threadIdx.x; i < J.n; i += (int64_t)J.nblocks * 256) J.out[i] = J.value;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_fill_batch(const FillJob<T> *__restrict__ jobs, const int32_t *__restrict__ block_job,
             const int32_t *__restrict__ block_inst, const int32_t *__restrict__ active, const WorkListDev wl) {
  if (wl.ajob != nullptr) {
    for_listed_blocks(wl, [&](int j, int local) { fill_batch_block<T>(jobs[j], local); });
    return;
  }
  const int b = blockIdx.x;
  const int j = block_job[b];
  const int i = block_inst[b];
  if (j < 0) return;
  if (active != nullptr && active[i] == 0) return;
  fill_batch_block<T>(jobs[j], b - jobs[j].block0);
}

struct NormJob {
  const double *partials;
  int32_t count;
  int32_t inst;
};

// Per-instance convergence test after a batched V-cycle, one block per job (one job
// per instance): the block sums the job's residual partials (block_sum, as
// k_norm_finalize does) if the instance is active, then thread 0 updates that
// instance's state and clears its activity flag once converged (r <= tol * s, s the
// instance's residual scale if given, else r0), out of cycles, or non-finite.
// k_count_active then counts the active instances.
__global__ void __launch_bounds__(NORM_THREADS)
k_batch_converge(const NormJob *__restrict__ jobs, double tol, int max_cycles, double *__restrict__ r0,
                 double *__restrict__ rlast, int32_t *__restrict__ cycles, int32_t *__restrict__ active,
                 int32_t *__restrict__ status, const double *__restrict__ rref) {
  __shared__ double red[NORM_THREADS];
  const NormJob J = jobs[blockIdx.x];
  // every thread reads the flag before block_sum's barriers; thread 0 writes it after
  if (active[J.inst] == 0) return; // uniform over the block
  const double s = block_sum(J.partials, J.count, red);
  if (threadIdx.x == 0) {
    const int i = J.inst;
    const double r = sqrt(s);
    const int c = cycles[i] + 1;
    cycles[i] = c;
    rlast[i] = r;
    if (!isfinite(s)) {
      status[i] = 3;
      active[i] = 0;
    } else {
      if (r0[i] < 0.0) r0[i] = r;
      const double scale = (rref != nullptr && rref[i] > 0.0) ? rref[i] : r0[i];
      if (r <= tol * scale) {
        status[i] = 1;
        active[i] = 0;
      } else if (c >= max_cycles) {
        status[i] = 2;
        active[i] = 0;
      }
    }
  }
}

__global__ void __launch_bounds__(NORM_THREADS)
k_count_active(const int32_t *__restrict__ active, int ninst, int32_t *__restrict__ nactive) {
  __shared__ int red[NORM_THREADS];
  int n = 0;
  for (int i = threadIdx.x; i < ninst; i += NORM_THREADS) n += active[i] != 0;
  red[threadIdx.x] = n;
  __syncthreads();
  for (int w = NORM_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *nactive = red[0];
}

// out[inst] = sqrt(sum of the job's partials), one block per job.
__global__ void __launch_bounds__(NORM_THREADS)
k_batch_norms(const NormJob *__restrict__ jobs, double *__restrict__ out) {
  __shared__ double red[NORM_THREADS];
  const NormJob J = jobs[blockIdx.x];
  const double s = block_sum(J.partials, J.count, red);
  if (threadIdx.x == 0) out[J.inst] = sqrt(s);
}

// Host side of the batches: job tables are built here and copied into the
// caller's device workspace: [jobs, padded to 256 B][block_job int32 per block].
inline int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

struct BlockTable {
  std::vector<int32_t> block_job;  // job of each block (-1: padding)
  std::vector<int32_t> block_inst; // instance of each block's job (read beside block_job)
  // per job (in job order): first block, block count, instance (the work lists)
  std::vector<int32_t> job_block0, job_nblocks, job_inst;
  void add(int job, int inst, int nblocks) {
    job_block0.push_back((int32_t)block_job.size());
    job_nblocks.push_back(nblocks);
    job_inst.push_back(inst);
    block_job.insert(block_job.end(), (size_t)nblocks, job);
    block_inst.insert(block_inst.end(), (size_t)nblocks, inst);
  }
  void pad_xcd() {
    while (block_job.size() % 8) {
      block_job.push_back(-1);
      block_inst.push_back(0);
    }
  }
};

template <typename T> int build_sweep_jobs(const xb_elast_args *args, const int32_t *inst, int32_t n,
                                           std::vector<SweepJob<T, ElastColl<T>>> &jobs, BlockTable &bt,
                                           int &flags) {
  for (int32_t k = 0; k < n; ++k) {
    const xb_elast_args *a = &args[k];
    if (int r = validate_elast(a, "xb_batch_prepare(elasticity job)")) return r;
    if (a->dtype != args[0].dtype || a->mode != args[0].mode)
      return fail(XB_EINVAL, "xb_batch_prepare: job %d: dtype and mode must agree across the batch", k);
    const int f = (a->bc != nullptr && a->bc->cell_slot != nullptr ? 1 : 0) | (a->resid != nullptr ? 2 : 0) |
                  (a->div_in != nullptr ? 4 : 0);
    if (k == 0) flags = f | 8;
    else if (f != (flags & 7))
      return fail(XB_EINVAL, "xb_batch_prepare: job %d: boundary table / residual / coupling presence must agree "
                  "across the batch", k);
    if ((f & 1) && a->bc->nslots > LDS_BC_SLOTS) flags &= ~8;
    SweepJob<T, ElastColl<T>> J{};
    elast_geom(a, J.g);
    elast_coll<T>(a, J.coll, J.nl);
    J.bc = bcdev<T>(a->bc);
    J.fin = static_cast<const T *>(a->f_in);
    J.fout = static_cast<T *>(a->f_out);
    J.fpc = static_cast<T *>(a->fpc_out);
    J.res = static_cast<const T *>(a->resid);
    J.coarse = static_cast<T *>(a->coarse_out);
    J.partials = a->norm_partials;
    const double gamma = a->mode == XB_MODE_SMOOTH ? a->gamma : 1.0;
    J.omg = T(1.0 - gamma);
    J.gm = T(gamma);
    J.inst = inst != nullptr ? inst[k] : k;
    J.block0 = (int32_t)bt.block_job.size();
    bt.add(k, J.inst, cdiv(J.g.nitems, WAVES_PER_BLOCK));
    jobs.push_back(J);
  }
  bt.pad_xcd();
  return XB_OK;
}

// ---------------------------------------------------------------------------
// Damped smoothing sweeps of a multigrid level, several per launch on LDS tiles:
// f <- (1 - gamma) f + gamma S(f - r; F(f)), r the level's residual on its interior
// box, F the body force (finest level) + the external force + the fused coupling
// force G_s(D f).  Per step every cell writes its divergence D f to LDS (coupled
// solves); barrier; every cell forms the coupling force from its neighbours'
// divergences (the gradient's wall rules included), collides f - r, applies its
// post-collision rules and writes the post-collision populations to LDS; barrier;
// every cell pulls, applies the post-streaming rules and damps.  The coupling force
// reaches two cells per step (divergence, gradient) and the pull one more, so the
// exact core shrinks by R = 2 cells per side and step with the coupling (1 without):
// after m steps the inner (32 - 2 R m)^2 cells are exact and are written, with the
// divergence of the result.  Per cell the arithmetic is that of m XB_MODE_SMOOTH
// sweeps (bit-identical).  Rule tables of at most LDS_BC_SLOTS slots only (staged in
// LDS with an identity row, as k_tile_steps); larger ones stay on the sweeps.
// Blocks of the smoothing tiles (cells per thread: SMOOTH_CPT).
constexpr int SMOOTH_THREADS = TILE_THREADS;
constexpr int SMOOTH_CPT = TILE * TILE / SMOOTH_THREADS;

template <typename T> struct TileSmoothDev {
  int nx, ny, steps, tiles_y, ix0, ix1, iy0, iy1;
  T gamma, omg;
  ElastColl<T> coll;
  NLDev<T> nl;
  const T *fin, *res;
  T *fout;
  BCDev<T> bc;
  // optional coarse correction of the next level (cnx x cny, Nc cells per direction),
  // injected into f_in over the interior box as it is loaded (k_prolong_add fused)
  const T *coarse;
  int cny;
  size_t Nc;
};

template <typename T> struct TileGradRules {
  int8_t code[(LDS_BC_SLOTS + 1) * 9];
  T cst[(LDS_BC_SLOTS + 1) * 9];
};

template <typename T, int Q>
__device__ __forceinline__ void stage_rule_rows(const int8_t *gc, const T *gk, int nslots, int8_t *code, T *cst) {
  const int n = nslots * Q;
  for (int k = threadIdx.x; k < n + Q; k += blockDim.x) {
    const bool id = k >= n || gc == nullptr;
    code[k] = id ? int8_t(0) : gc[k];
    cst[k] = id ? T(0) : gk[k];
  }
}

template <typename T, int BC, bool HAS_RES, bool HAS_NL, bool GWALLS>
__device__ __forceinline__ void tile_smooth_body(const TileSmoothDev<T> &A, int tile, T *lds, T *ldiv,
                                                 TileRules<T, 8> *sh, TileGradRules<T> *gsh) {
  using C = ElastColl<T>;
  using D = DarcyColl<T>;
  using Aux = typename C::Aux;
  constexpr int Q = 8;
  constexpr int TT = TILE * TILE;
  constexpr int R = HAS_NL ? 2 : 1;
  const int nx = A.nx, ny = A.ny, steps = A.steps;
  const int inner = TILE - 2 * R * steps;
  const int bx = tile / A.tiles_y, by = tile % A.tiles_y;
  const int x0 = bx * inner - R * steps, y0 = by * inner - R * steps;
  const size_t N = (size_t)nx * ny;
  const int nyi = A.iy1 - A.iy0;
  const size_t Ni = (size_t)(A.ix1 - A.ix0) * nyi;
  constexpr bool WALLS = BC != 0;
  if constexpr (BC == 2) {
    stage_rule_rows<T, Q>(A.bc.code_pc, A.bc.const_pc, A.bc.nslots, sh->code_pc, sh->const_pc);
    stage_rule_rows<T, Q>(A.bc.code_ps, A.bc.const_ps, A.bc.nslots, sh->code_ps, sh->const_ps);
  }
  if constexpr (GWALLS) stage_rule_rows<T, 9>(A.nl.gbc.code_ps, A.nl.gbc.const_ps, A.nl.gbc.nslots, gsh->code, gsh->cst);
  T f[SMOOTH_CPT][Q];
  Aux aux[SMOOTH_CPT];
  T kd[SMOOTH_CPT];
  int slot[SMOOTH_CPT], gslot[SMOOTH_CPT], ri[SMOOTH_CPT];
  bool mine = false, gmine = false;
  // grid coordinates of the thread's cells: one column (SMOOTH_THREADS is a multiple
  // of TILE), one row per cell; the halo (R steps) is below the tile size
  static_assert(SMOOTH_THREADS % TILE == 0, "a thread's cells share their column");
  const bool big = nx >= TILE && ny >= TILE;
  const int gy = wrap_axis(y0 + (int)threadIdx.x % TILE, ny, big);
  int gxs[SMOOTH_CPT];
#pragma unroll
  for (int k = 0; k < SMOOTH_CPT; ++k)
    gxs[k] = wrap_axis(x0 + ((int)threadIdx.x + k * SMOOTH_THREADS) / TILE, nx, big);
  if (A.fin == nullptr) {
    // a coarse level's V-cycle starts from the zero state: nothing to load
#pragma unroll
    for (int k = 0; k < SMOOTH_CPT; ++k) {
#pragma unroll
      for (int d = 0; d < Q; ++d) f[k][d] = T(0);
    }
  } else if (A.coarse != nullptr) {
    // the prolongation's f + c (the sum of k_prolong_add) on the interior box: no
    // branch per cell, so the correction's loads go out with the populations'
#pragma unroll
    for (int k = 0; k < SMOOTH_CPT; ++k) {
      const int gx = gxs[k];
      const size_t cell = (size_t)gx * ny + gy;
      const bool in = gx >= A.ix0 && gx < A.ix1 && gy >= A.iy0 && gy < A.iy1;
      const size_t cc = in ? (size_t)(A.ix0 + ((gx - A.ix0) >> 1)) * A.cny + (A.iy0 + ((gy - A.iy0) >> 1)) : 0;
#pragma unroll
      for (int d = 0; d < Q; ++d) {
        const T v = A.fin[d * N + cell];
        const T cv = A.coarse[d * A.Nc + cc];
        f[k][d] = in ? v + cv : v;
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < SMOOTH_CPT; ++k) {
      const size_t cell = (size_t)gxs[k] * ny + gy;
#pragma unroll
      for (int d = 0; d < Q; ++d) f[k][d] = A.fin[d * N + cell];
    }
  }
#pragma unroll
  for (int k = 0; k < SMOOTH_CPT; ++k) {
    const int gx = gxs[k];
    const size_t cell = (size_t)gx * ny + gy;
    aux[k] = A.coll.load_aux(cell);
    kd[k] = A.coll.kdiv(aux[k], A.nl.kdiv);
    slot[k] = (WALLS && A.bc.slot != nullptr) ? A.bc.slot[cell] : -1;
    gslot[k] = (GWALLS && A.nl.gbc.slot != nullptr) ? A.nl.gbc.slot[cell] : -1;
    mine = mine || slot[k] >= 0;
    gmine = gmine || gslot[k] >= 0;
    ri[k] = -1;
    if constexpr (HAS_RES) {
      if (gx >= A.ix0 && gx < A.ix1 && gy >= A.iy0 && gy < A.iy1) ri[k] = (gx - A.ix0) * nyi + (gy - A.iy0);
    }
  }
  const bool walls = BC == 2 && __syncthreads_or(mine) != 0;
  const bool gwalls = GWALLS && __syncthreads_or(gmine) != 0;
#pragma unroll
  for (int k = 0; k < SMOOTH_CPT; ++k) {
    if constexpr (BC == 2) slot[k] = slot[k] >= 0 ? slot[k] : A.bc.nslots;
    if constexpr (GWALLS) gslot[k] = gslot[k] >= 0 ? gslot[k] : A.nl.gbc.nslots;
  }
  for (int s = 0; s < steps; ++s) {
    // The residual is read again in every step (from L2): a pointer that the compiler
    // must reload keeps it from hoisting the 32 values of the thread's cells out of
    // the step loop and holding them in registers (which made this kernel spill).
    const T *res = HAS_RES ? *reinterpret_cast<const T *const volatile *>(&A.res) : nullptr;
    if constexpr (HAS_NL) {
#pragma unroll
      for (int k = 0; k < SMOOTH_CPT; ++k) ldiv[threadIdx.x + k * SMOOTH_THREADS] = elast_div_of<T>(f[k], kd[k]);
      __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < SMOOTH_CPT; ++k) {
      const int c = threadIdx.x + k * SMOOTH_THREADS;
      const int rx = c / TILE, ry = c % TILE;
      T pre[Q];
#pragma unroll
      for (int d = 0; d < Q; ++d) pre[d] = f[k][d];
      if constexpr (HAS_RES) {
        if (ri[k] >= 0) {
#pragma unroll
          for (int d = 0; d < Q; ++d) pre[d] = pre[d] - res[d * Ni + ri[k]];
        }
      }
      Aux ax = aux[k];
      if constexpr (HAS_NL) {
        // as tile_pull: border cells read a neighbouring slot (they are in the halo)
        T qs[9];
#pragma unroll
        for (int d = 0; d < 9; ++d) qs[d] = ldiv[c - (D::cx(d) * TILE + D::cy(d))];
        if (gwalls) {
          // the gradient's rules read the cell's own divergence
          const int8_t *code = gsh->code + gslot[k] * 9;
          const T *cst = gsh->cst + gslot[k] * 9;
          const T s0 = qs[0];
          T o[9];
#pragma unroll
          for (int d = 0; d < 9; ++d) {
            const int cc = code[d];
            T b = cc > 0 ? s0 : -s0;
            b = cc == XB_RULE_ZERO ? T(0) : b;
            b = cc == 0 ? qs[d] : b;
            o[d] = b + cst[d];
          }
#pragma unroll
          for (int d = 0; d < 9; ++d) qs[d] = o[d];
        }
        T g0, g1;
        stencil_grad<T>(qs, A.nl.m3s, g0, g1);
        ax.b0 = A.nl.base != nullptr ? ax.b0 + g0 : g0;
        ax.b1 = A.nl.base != nullptr ? ax.b1 + g1 : g1;
      }
      T pc[Q];
      A.coll.collide(pre, ax, pc);
      if constexpr (BC == 2) {
        if (walls && A.bc.code_pc != nullptr) {
#pragma unroll
          for (int d = 0; d < Q; ++d) lds[d * TT + c] = pre[d];
          rules_lds<T, Q, TT>(pc, lds + c, sh->code_pc + slot[k] * Q, sh->const_pc + slot[k] * Q);
        }
      } else if constexpr (BC == 1) {
        if (slot[k] >= 0 && A.bc.code_pc != nullptr) {
#pragma unroll
          for (int d = 0; d < Q; ++d) lds[d * TT + c] = pre[d];
          rules_lds<T, Q, TT>(pc, lds + c, A.bc.code_pc + (size_t)slot[k] * Q, A.bc.const_pc + (size_t)slot[k] * Q);
        }
      }
#pragma unroll
      for (int d = 0; d < Q; ++d) lds[d * TT + c] = pc[d];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SMOOTH_CPT; ++k) {
      const int c = threadIdx.x + k * SMOOTH_THREADS;
      const int rx = c / TILE, ry = c % TILE;
      T v[Q];
      tile_pull<T, C, TILE, TILE>(lds, c, v);
      if constexpr (BC == 2) {
        if (walls && A.bc.code_ps != nullptr)
          rules_lds<T, Q, TT>(v, lds + c, sh->code_ps + slot[k] * Q, sh->const_ps + slot[k] * Q);
      } else if constexpr (BC == 1) {
        if (slot[k] >= 0 && A.bc.code_ps != nullptr)
          rules_lds<T, Q, TT>(v, lds + c, A.bc.code_ps + (size_t)slot[k] * Q, A.bc.const_ps + (size_t)slot[k] * Q);
      }
#pragma unroll
      for (int d = 0; d < Q; ++d) f[k][d] = A.omg * f[k][d] + A.gamma * v[d];
    }
    // the pull's LDS reads end before the next step's post-collision writes: with the
    // coupling force the next step's divergence barrier (ldiv is another array) already
    // orders them, and after the last step only registers are read
    if constexpr (!HAS_NL) __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < SMOOTH_CPT; ++k) {
    const int c = threadIdx.x + k * SMOOTH_THREADS;
    const int rx = c / TILE, ry = c % TILE;
    const int h = R * steps;
    if (rx < h || rx >= TILE - h || ry < h || ry >= TILE - h) continue;
    const int gx = x0 + rx, gy = y0 + ry;
    if (gx >= nx || gy >= ny) continue;
    const size_t cell = (size_t)gx * ny + gy;
#pragma unroll
    for (int d = 0; d < Q; ++d) A.fout[d * N + cell] = f[k][d];
    if constexpr (HAS_NL) {
      if (A.nl.div_out != nullptr) A.nl.div_out[cell] = elast_div_of<T>(f[k], kd[k]);
    }
  }
}

template <typename T, int BC, bool HAS_RES, bool HAS_NL, bool GWALLS>
__device__ __forceinline__ void tile_smooth_run(const TileSmoothDev<T> &A, int tile) {
  constexpr bool WALLS = BC == 2;
  __shared__ T lds_raw[8 * TILE * TILE + 2 * TILE_LDS_PAD(TILE)];
  T *lds = lds_raw + TILE_LDS_PAD(TILE);
  if constexpr (HAS_NL) {
    __shared__ T ldiv_raw[TILE * TILE + 2 * TILE_LDS_PAD(TILE)];
    T *ldiv = ldiv_raw + TILE_LDS_PAD(TILE);
    if constexpr (WALLS) {
      __shared__ TileRules<T, 8> sh;
      if constexpr (GWALLS) {
        __shared__ TileGradRules<T> gsh;
        tile_smooth_body<T, BC, HAS_RES, HAS_NL, GWALLS>(A, tile, lds, ldiv, &sh, &gsh);
      } else {
        tile_smooth_body<T, BC, HAS_RES, HAS_NL, GWALLS>(A, tile, lds, ldiv, &sh, nullptr);
      }
    } else {
      if constexpr (GWALLS) {
        __shared__ TileGradRules<T> gsh;
        tile_smooth_body<T, BC, HAS_RES, HAS_NL, GWALLS>(A, tile, lds, ldiv, nullptr, &gsh);
      } else {
        tile_smooth_body<T, BC, HAS_RES, HAS_NL, GWALLS>(A, tile, lds, ldiv, nullptr, nullptr);
      }
    }
  } else {
    if constexpr (WALLS) {
      __shared__ TileRules<T, 8> sh;
      tile_smooth_body<T, BC, HAS_RES, false, false>(A, tile, lds, nullptr, &sh, nullptr);
    } else {
      tile_smooth_body<T, BC, HAS_RES, false, false>(A, tile, lds, nullptr, nullptr, nullptr);
    }
  }
}

template <typename T, int BC, bool HAS_RES, bool HAS_NL, bool GWALLS>
__global__ void __launch_bounds__(SMOOTH_THREADS, HAS_RES ? 2 : 1) k_tile_smooth(const TileSmoothDev<T> A) {
  tile_smooth_run<T, BC, HAS_RES, HAS_NL, GWALLS>(A, xcd_block_id());
}

// Batched form: every block looks up its job (one level of one instance) and tile.
template <typename T> struct TileSmoothJob {
  TileSmoothDev<T> A;
  int32_t inst, block0;
};

template <typename T, int BC, bool HAS_RES, bool HAS_NL, bool GWALLS>
__global__ void __launch_bounds__(SMOOTH_THREADS, HAS_RES ? 2 : 1)
k_tile_smooth_batch(const TileSmoothJob<T> *__restrict__ jobs, const int32_t *__restrict__ block_job,
                    const int32_t *__restrict__ block_inst, const int32_t *__restrict__ active, const WorkListDev wl) {
  if (wl.ajob != nullptr) {
    for_listed_blocks(wl, [&](int j, int local) {
      tile_smooth_run<T, BC, HAS_RES, HAS_NL, GWALLS>(jobs[j].A, local);
      __syncthreads(); // the next tile reuses the LDS arrays
    });
    return;
  }
  const int b = xcd_block_id();
  const int j = block_job[b];
  if (j < 0) return;                                         // uniform per block
  if (active != nullptr && active[block_inst[b]] == 0) return; // uniform per block
  const TileSmoothJob<T> &J = jobs[j];
  tile_smooth_run<T, BC, HAS_RES, HAS_NL, GWALLS>(J.A, b - J.block0);
}

template <typename T, int BC, bool HAS_RES>
void launch_tile_smooth_nl(const TileSmoothDev<T> &A, dim3 grid, hipStream_t st, bool nl, bool gwalls) {
  if (!nl) hipLaunchKernelGGL((k_tile_smooth<T, BC, HAS_RES, false, false>), grid, dim3(SMOOTH_THREADS), 0, st, A);
  else if (gwalls) hipLaunchKernelGGL((k_tile_smooth<T, BC, HAS_RES, true, true>), grid, dim3(SMOOTH_THREADS), 0, st, A);
  else hipLaunchKernelGGL((k_tile_smooth<T, BC, HAS_RES, true, false>), grid, dim3(SMOOTH_THREADS), 0, st, A);
}

// the tile-smoothing operands of one xb_elast_args; returns the number of tiles
template <typename T> int make_tile_smooth(const xb_elast_args *a, int steps, TileSmoothDev<T> &A) {
  A = TileSmoothDev<T>{};
  NLDev<T> nl;
  elast_coll<T>(a, A.coll, nl);
  A.nl = nl;
  A.nx = a->nx; A.ny = a->ny; A.steps = steps;
  A.ix0 = a->ix0; A.ix1 = a->ix1; A.iy0 = a->iy0; A.iy1 = a->iy1;
  A.gamma = T(a->gamma);
  A.omg = T(1.0 - a->gamma);
  A.fin = static_cast<const T *>(a->f_in);
  A.res = static_cast<const T *>(a->resid);
  A.fout = static_cast<T *>(a->f_out);
  A.bc = bcdev<T>(a->bc);
  A.coarse = static_cast<const T *>(a->coarse_out);
  A.cny = a->ny - (a->iy1 - a->iy0) / 2;
  A.Nc = (size_t)(a->nx - (a->ix1 - a->ix0) / 2) * A.cny;
  const int R = a->div_in != nullptr ? 2 : 1;
  const int inner = TILE - 2 * R * steps;
  const int tx = cdiv(a->nx, inner), ty = cdiv(a->ny, inner);
  A.tiles_y = ty;
  return tx * ty;
}

template <typename T> int elast_smoothn_t(const xb_elast_args *a, int steps, hipStream_t st) {
  TileSmoothDev<T> A;
  const dim3 grid((unsigned)make_tile_smooth<T>(a, steps, A));
  const NLDev<T> &nl = A.nl;
  const bool has_nl = a->div_in != nullptr;
  const int bcm = bc_mode(A.bc.slot, A.bc.nslots);
  const bool gwalls = has_nl && nl.gbc.slot != nullptr;
  const bool res = a->resid != nullptr;
  if (bcm == 2 && res) launch_tile_smooth_nl<T, 2, true>(A, grid, st, has_nl, gwalls);
  else if (bcm == 2) launch_tile_smooth_nl<T, 2, false>(A, grid, st, has_nl, gwalls);
  else if (bcm == 1 && res) launch_tile_smooth_nl<T, 1, true>(A, grid, st, has_nl, gwalls);
  else if (bcm == 1) launch_tile_smooth_nl<T, 1, false>(A, grid, st, has_nl, gwalls);
  else if (res) launch_tile_smooth_nl<T, 0, true>(A, grid, st, has_nl, gwalls);
  else launch_tile_smooth_nl<T, 0, false>(A, grid, st, has_nl, gwalls);
  return check_launch("multi-step smoothing tile kernel");
}


int check_smoothn(const xb_elast_args *a, int steps, const char *who) {
  if (int r = validate_elast(a, who, true)) return r;
  if (a->mode != XB_MODE_SMOOTH) return fail(XB_EINVAL, "%s: needs XB_MODE_SMOOTH", who);
  if (a->f_in == a->f_out) return fail(XB_EINVAL, "%s: f_in and f_out must differ", who);
  if (a->coarse_out != nullptr && (((a->ix1 - a->ix0) & 1) || ((a->iy1 - a->iy0) & 1)))
    return fail(XB_EINVAL, "%s: a coarse correction needs an interior box of even sides", who);
  if (a->coarse_out != nullptr && (a->coarse_out == a->f_out || a->coarse_out == a->f_in))
    return fail(XB_EINVAL, "%s: the coarse correction must not alias f_in or f_out", who);
  if (a->f_in == nullptr && a->coarse_out != nullptr)
    return fail(XB_EINVAL, "%s: a zero start state (null f_in) takes no coarse correction", who);
  const int R = a->div_in != nullptr ? 2 : 1;
  if (steps < 1 || TILE - 2 * R * steps < 8)
    return fail(XB_EINVAL, "%s: steps must be in [1, %d]", who, (TILE - 8) / (2 * R));
  if (a->grad_bc != nullptr && a->grad_bc->cell_slot != nullptr && a->grad_bc->nslots > LDS_BC_SLOTS)
    return fail(XB_EUNSUPPORTED, "%s: gradient rule tables of more than %d slots stay on the sweeps", who,
                LDS_BC_SLOTS);
  return XB_OK;
}

template <typename T> int build_smoothn_jobs(const xb_elast_args *args, const int32_t *inst, int32_t n,
                                             std::vector<TileSmoothJob<T>> &jobs, BlockTable &bt, int &flags) {
  for (int32_t k = 0; k < n; ++k) {
    const xb_elast_args *a = &args[k];
    if (int r = check_smoothn(a, a->rows, "xb_batch_prepare(smoothing tile job)")) return r;
    if (a->dtype != args[0].dtype) return fail(XB_EINVAL, "xb_batch_prepare: job %d: dtype must agree", k);
    const int f = (a->bc != nullptr && a->bc->cell_slot != nullptr ? 1 : 0) | (a->resid != nullptr ? 2 : 0) |
                  (a->div_in != nullptr ? 4 : 0) |
                  (a->div_in != nullptr && a->grad_bc != nullptr && a->grad_bc->cell_slot != nullptr ? 16 : 0);
    if (k == 0) flags = f | 8;
    else if (f != (flags & 23))
      return fail(XB_EINVAL, "xb_batch_prepare: job %d: boundary tables / residual / coupling presence must agree "
                  "across the batch", k);
    if ((f & 1) && a->bc->nslots > LDS_BC_SLOTS) flags &= ~8;
    TileSmoothJob<T> J{};
    const int tiles = make_tile_smooth<T>(a, a->rows, J.A);
    J.inst = inst != nullptr ? inst[k] : k;
    J.block0 = (int32_t)bt.block_job.size();
    bt.add(k, J.inst, tiles);
    jobs.push_back(J);
  }
  bt.pad_xcd();
  return XB_OK;
}

// Work list of a plan launched over it (empty: one workgroup per table block).  The
// list region holds the active jobs (njobs), the prefix of their block counts
// (njobs + 1) and their number.
WorkListDev plan_worklist(const xb_batch_plan *p) {
  WorkListDev wl;
  if (p->use_list && p->work_offset > 0) {
    const int32_t *w = reinterpret_cast<const int32_t *>(static_cast<const char *>(p->table) + p->work_offset);
    wl.ajob = w;
    wl.aoff = w + p->njobs;
    wl.naj = w + 2 * p->njobs + 1;
  }
  return wl;
}

// Workgroups of a batched launch: one per table block, or over a work list as many
// as the GPU holds at once for this kernel (a multiple of 8, at most the table's
// block count): each takes listed blocks in turn.
int batch_grid(const xb_batch_plan *p, const WorkListDev &wl, const void *kernel, int threads) {
  if (wl.ajob == nullptr) return p->nblocks;
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 256;
  const int waves = std::max(1, std::min(p->list_waves, 64));
  const int g = (int)std::min<int64_t>((int64_t)per_cu * cus * waves, p->nblocks);
  return std::max(8, (g + 7) / 8 * 8);
}

template <typename T, int BC, bool RES, bool NL, bool GW>
void launch_smoothn_batch_k(const xb_batch_plan *p, const TileSmoothJob<T> *jobs, const int32_t *bj,
                            const int32_t *bi, const int32_t *active, const WorkListDev &wl, hipStream_t st) {
  const int grid = batch_grid(p, wl, reinterpret_cast<const void *>(&k_tile_smooth_batch<T, BC, RES, NL, GW>),
                              SMOOTH_THREADS);
  hipLaunchKernelGGL((k_tile_smooth_batch<T, BC, RES, NL, GW>), dim3(grid), dim3(SMOOTH_THREADS), 0, st, jobs, bj, bi,
                     active, wl);
}

template <typename T, int BC, bool RES>
void launch_smoothn_batch_nl(const xb_batch_plan *p, const TileSmoothJob<T> *jobs, const int32_t *bj,
                             const int32_t *bi, const int32_t *active, const WorkListDev &wl, hipStream_t st, bool nl,
                             bool gw) {
  if (!nl) launch_smoothn_batch_k<T, BC, RES, false, false>(p, jobs, bj, bi, active, wl, st);
  else if (gw) launch_smoothn_batch_k<T, BC, RES, true, true>(p, jobs, bj, bi, active, wl, st);
  else launch_smoothn_batch_k<T, BC, RES, true, false>(p, jobs, bj, bi, active, wl, st);
}

template <typename T> int launch_smoothn_batch(const xb_batch_plan *p, const int32_t *active, hipStream_t st) {
  const char *base = static_cast<const char *>(p->table);
  const auto *jobs = reinterpret_cast<const TileSmoothJob<T> *>(base + p->jobs_offset);
  const int32_t *bj = reinterpret_cast<const int32_t *>(base + p->blocks_offset);
  const int32_t *bi = reinterpret_cast<const int32_t *>(base + p->insts_offset);
  const WorkListDev wl = plan_worklist(p);
  const int bcm = (p->flags & 1) ? ((p->flags & 8) ? 2 : 1) : 0;
  const bool res = p->flags & 2, nl = p->flags & 4, gw = p->flags & 16;
  if (bcm == 2 && res) launch_smoothn_batch_nl<T, 2, true>(p, jobs, bj, bi, active, wl, st, nl, gw);
  else if (bcm == 2) launch_smoothn_batch_nl<T, 2, false>(p, jobs, bj, bi, active, wl, st, nl, gw);
  else if (bcm == 1 && res) launch_smoothn_batch_nl<T, 1, true>(p, jobs, bj, bi, active, wl, st, nl, gw);
  else if (bcm == 1) launch_smoothn_batch_nl<T, 1, false>(p, jobs, bj, bi, active, wl, st, nl, gw);
  else if (res) launch_smoothn_batch_nl<T, 0, true>(p, jobs, bj, bi, active, wl, st, nl, gw);
  else launch_smoothn_batch_nl<T, 0, false>(p, jobs, bj, bi, active, wl, st, nl, gw);
  return check_launch("batched smoothing tile kernel");
}

int ew_blocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65535 * 8)); }

template <typename T> int build_prolong_jobs(const xb_prolong_job *args, const int32_t *inst, int32_t n,
                                             std::vector<ProlongJob<T>> &jobs, BlockTable &bt) {
  for (int32_t k = 0; k < n; ++k) {
    const xb_prolong_job *a = &args[k];
    if (a->f == nullptr || a->coarse == nullptr) return fail(XB_EINVAL, "xb_batch_prepare: prolong job %d: null pointer", k);
    if (a->ix0 < 0 || a->ix1 > a->nx || a->iy0 < 0 || a->iy1 > a->ny || a->ix0 >= a->ix1 || a->iy0 >= a->iy1 ||
        ((a->ix1 - a->ix0) & 1) || ((a->iy1 - a->iy0) & 1))
      return fail(XB_EINVAL, "xb_batch_prepare: prolong job %d: bad interior box", k);
    if (a->cnx != a->ix0 + (a->ix1 - a->ix0) / 2 + (a->nx - a->ix1) ||
        a->cny != a->iy0 + (a->iy1 - a->iy0) / 2 + (a->ny - a->iy1))
      return fail(XB_EINVAL, "xb_batch_prepare: prolong job %d: coarse grid does not match the fine box", k);
    ProlongJob<T> J{};
    J.nx = a->nx; J.ny = a->ny; J.cnx = a->cnx; J.cny = a->cny;
    J.ix0 = a->ix0; J.ix1 = a->ix1; J.iy0 = a->iy0; J.iy1 = a->iy1;
    J.k = T(-1.5 / (3 * (a->lam + a->mu) + 1));
    J.f = static_cast<T *>(a->f);
    J.c = static_cast<const T *>(a->coarse);
    J.div = static_cast<T *>(a->div);
    J.inst = inst != nullptr ? inst[k] : k;
    J.block0 = (int32_t)bt.block_job.size();
    J.nblocks = ew_blocks((int64_t)(a->ix1 - a->ix0) * (a->iy1 - a->iy0));
    bt.add(k, J.inst, J.nblocks);
    jobs.push_back(J);
  }
  return XB_OK;
}

template <typename T> int build_fill_jobs(const xb_fill_job *args, const int32_t *inst, int32_t n,
                                          std::vector<FillJob<T>> &jobs, BlockTable &bt) {
  for (int32_t k = 0; k < n; ++k) {
    const xb_fill_job *a = &args[k];
    if (a->n < 0 || (a->n > 0 && a->out == nullptr)) return fail(XB_EINVAL, "xb_batch_prepare: fill job %d: bad arguments", k);
    FillJob<T> J{};
    J.out = static_cast<T *>(a->out);
    J.n = a->n;
    J.value = T(a->value);
    J.inst = inst != nullptr ? inst[k] : k;
    J.block0 = (int32_t)bt.block_job.size();
    J.nblocks = a->n > 0 ? ew_blocks(a->n) : 0;
    bt.add(k, J.inst, J.nblocks);
    jobs.push_back(J);
  }
  return XB_OK;
}

int build_norm_jobs(const xb_norm_job *args, const int32_t *inst, int32_t n, std::vector<NormJob> &jobs) {
  std::vector<int32_t> seen;
  for (int32_t k = 0; k < n; ++k) {
    if (args[k].partials == nullptr || args[k].count <= 0)
      return fail(XB_EINVAL, "xb_batch_prepare: norm job %d: bad arguments", k);
    const int32_t i = inst != nullptr ? inst[k] : k;
    if (i < 0) return fail(XB_EINVAL, "xb_batch_prepare: norm job %d: negative instance", k);
    seen.push_back(i);
    jobs.push_back(NormJob{args[k].partials, args[k].count, i});
  }
  // the convergence and norm kernels give each job a block of its own and write its
  // instance's state: one job per instance
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
    return fail(XB_EINVAL, "xb_batch_prepare: a norm plan takes one job per instance");
  return XB_OK;
}

// Table size and, when `table` is set, the upload of the tables.
template <typename J>
int finish_table(const std::vector<J> &jobs, const BlockTable &bt, int32_t kind, int32_t dtype, int32_t mode,
                 int32_t flags, void *table, int64_t bytes, xb_batch_plan *plan, hipStream_t st, int64_t *need) {
  // layout: jobs | block_job | block_inst | per-job info (3 x njobs) | work list
  // (2 x njobs + 2: active jobs, prefix of their block counts, their number)
  const bool listed = !bt.job_block0.empty(); // every kind with blocks (not NORM)
  const int64_t nj = (int64_t)bt.job_block0.size();
  const int64_t jb = align256((int64_t)(jobs.size() * sizeof(J)));
  const int64_t ib = jb + align256((int64_t)bt.block_job.size() * 4);
  const int64_t fb = ib + align256((int64_t)bt.block_inst.size() * 4);
  const int64_t wb = fb + align256(3 * nj * 4);
  const int64_t total = listed ? wb + (2 * nj + 2) * 4 : ib + (int64_t)bt.block_inst.size() * 4;
  if (need != nullptr) {
    *need = std::max<int64_t>(total, 256);
    return XB_OK;
  }
  if (table == nullptr || bytes < total)
    return fail(XB_EINVAL, "xb_batch_prepare: table of %lld bytes, %lld needed", (long long)bytes, (long long)total);
  if (!jobs.empty()) {
    if (hipMemcpyAsync(table, jobs.data(), jobs.size() * sizeof(J), hipMemcpyHostToDevice, st) != hipSuccess)
      return fail(XB_EHIP, "xb_batch_prepare: job table upload failed");
  }
  if (!bt.block_job.empty()) {
    if (hipMemcpyAsync(static_cast<char *>(table) + jb, bt.block_job.data(), bt.block_job.size() * 4,
                       hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(static_cast<char *>(table) + ib, bt.block_inst.data(), bt.block_inst.size() * 4,
                       hipMemcpyHostToDevice, st) != hipSuccess)
      return fail(XB_EHIP, "xb_batch_prepare: block table upload failed");
  }
  std::vector<int32_t> info, work;
  if (listed) {
    info.reserve((size_t)(3 * nj));
    info.insert(info.end(), bt.job_block0.begin(), bt.job_block0.end());
    info.insert(info.end(), bt.job_nblocks.begin(), bt.job_nblocks.end());
    info.insert(info.end(), bt.job_inst.begin(), bt.job_inst.end());
    work.assign((size_t)(2 * nj + 2), 0); // nothing listed until the first xb_batch_compact
    if (hipMemcpyAsync(static_cast<char *>(table) + fb, info.data(), info.size() * 4, hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(static_cast<char *>(table) + wb, work.data(), work.size() * 4, hipMemcpyHostToDevice, st) !=
            hipSuccess)
      return fail(XB_EHIP, "xb_batch_prepare: work list upload failed");
  }
  if (hipStreamSynchronize(st) != hipSuccess) return fail(XB_EHIP, "xb_batch_prepare: synchronisation failed");
  plan->jobinfo_offset = listed ? fb : 0;
  plan->work_offset = listed ? wb : 0;
  plan->use_list = 0;
  plan->list_waves = 1;
  plan->kind = kind;
  plan->dtype = dtype;
  plan->mode = mode;
  plan->flags = flags;
  plan->njobs = (int32_t)jobs.size();
  plan->nblocks = (int32_t)bt.block_job.size();
  plan->table = table;
  plan->table_bytes = bytes;
  plan->jobs_offset = 0;
  plan->blocks_offset = jb;
  plan->insts_offset = ib;
  return XB_OK;
}

template <typename T>
int batch_prepare_t(int32_t kind, const void *jobs, const int32_t *inst, int32_t n, void *table, int64_t bytes,
                    xb_batch_plan *plan, hipStream_t st, int64_t *need) {
  BlockTable bt;
  const int32_t dt = std::is_same<T, double>::value ? XB_F64 : XB_F32;
  switch (kind) {
  case XB_BATCH_ELAST_STEP: {
    std::vector<SweepJob<T, ElastColl<T>>> v;
    int flags = 0;
    const xb_elast_args *a = static_cast<const xb_elast_args *>(jobs);
    if (int r = build_sweep_jobs<T>(a, inst, n, v, bt, flags)) return r;
    return finish_table(v, bt, kind, dt, a[0].mode, flags, table, bytes, plan, st, need);
  }
  case XB_BATCH_ELAST_SMOOTHN: {
    std::vector<TileSmoothJob<T>> v;
    int flags = 0;
    const xb_elast_args *a = static_cast<const xb_elast_args *>(jobs);
    if (int r = build_smoothn_jobs<T>(a, inst, n, v, bt, flags)) return r;
    return finish_table(v, bt, kind, dt, a[0].mode, flags, table, bytes, plan, st, need);
  }
  case XB_BATCH_PROLONG_DIV: {
    std::vector<ProlongJob<T>> v;
    if (int r = build_prolong_jobs<T>(static_cast<const xb_prolong_job *>(jobs), inst, n, v, bt)) return r;
    return finish_table(v, bt, kind, dt, 0, 0, table, bytes, plan, st, need);
  }
  case XB_BATCH_FILL: {
    std::vector<FillJob<T>> v;
    if (int r = build_fill_jobs<T>(static_cast<const xb_fill_job *>(jobs), inst, n, v, bt)) return r;
    return finish_table(v, bt, kind, dt, 0, 0, table, bytes, plan, st, need);
  }
  default: {
    std::vector<NormJob> v;
    if (int r = build_norm_jobs(static_cast<const xb_norm_job *>(jobs), inst, n, v)) return r;
    return finish_table(v, bt, kind, dt, 0, 0, table, bytes, plan, st, need);
  }
  }
}

template <typename T, int MODE>
int launch_elast_batch(const xb_batch_plan *p, const int32_t *active, hipStream_t st) {
  using J = SweepJob<T, ElastColl<T>>;
  const J *jobs = reinterpret_cast<const J *>(static_cast<const char *>(p->table) + p->jobs_offset);
  const int32_t *bj = reinterpret_cast<const int32_t *>(static_cast<const char *>(p->table) + p->blocks_offset);
  const int32_t *bi = reinterpret_cast<const int32_t *>(static_cast<const char *>(p->table) + p->insts_offset);
  const WorkListDev wl = plan_worklist(p);
  const dim3 block(BLOCK);
#define XB_SB(BC, RES, NL)                                                                                          \
  hipLaunchKernelGGL(                                                                                              \
      (k_sweep_batch<T, ElastColl<T>, MODE, BC, RES, NL>),                                                         \
      dim3(batch_grid(p, wl, reinterpret_cast<const void *>(&k_sweep_batch<T, ElastColl<T>, MODE, BC, RES, NL>),  \
                      BLOCK)),                                                                                     \
      block, 0, st, jobs, bj, bi, active, wl)
  const int bcm = (p->flags & 1) ? ((p->flags & 8) ? 2 : 1) : 0;
#define XB_SB_BC(RES, NL)                                                                                           \
  do {                                                                                                             \
    if (bcm == 2) XB_SB(2, RES, NL);                                                                               \
    else if (bcm == 1) XB_SB(1, RES, NL);                                                                          \
    else XB_SB(0, RES, NL);                                                                                        \
  } while (0)
  switch (p->flags & 6) {
  case 0: XB_SB_BC(false, false); break;
  case 2: XB_SB_BC(true, false); break;
  case 4: XB_SB_BC(false, true); break;
  default: XB_SB_BC(true, true); break;
  }
#undef XB_SB_BC
#undef XB_SB
  return check_launch("batched sweep kernel");
}

template <typename T> int batch_launch_t(const xb_batch_plan *p, const int32_t *active, hipStream_t st) {
  const char *base = static_cast<const char *>(p->table);
  const int32_t *bj = reinterpret_cast<const int32_t *>(base + p->blocks_offset);
  const int32_t *bi = reinterpret_cast<const int32_t *>(base + p->insts_offset);
  const WorkListDev wl = plan_worklist(p);
  switch (p->kind) {
  case XB_BATCH_ELAST_STEP:
    if (p->mode == XB_MODE_STEP) return xb_part_elast_batch_step(p, active, st);
    if (p->mode == XB_MODE_SMOOTH) return xb_part_elast_batch_smooth(p, active, st);
    return xb_part_elast_batch_residual(p, active, st);
  case XB_BATCH_ELAST_SMOOTHN:
    return xb_part_smoothn_batch(p, active, st);
  case XB_BATCH_PROLONG_DIV:
    hipLaunchKernelGGL(k_prolong_batch<T>, dim3(batch_grid(p, wl, reinterpret_cast<const void *>(&k_prolong_batch<T>), 256)),
                       dim3(256), 0, st, reinterpret_cast<const ProlongJob<T> *>(base + p->jobs_offset), bj, bi, active,
                       wl);
    return check_launch("batched prolongation kernel");
  default:
    hipLaunchKernelGGL(k_fill_batch<T>, dim3(batch_grid(p, wl, reinterpret_cast<const void *>(&k_fill_batch<T>), 256)),
                       dim3(256), 0, st, reinterpret_cast<const FillJob<T> *>(base + p->jobs_offset), bj, bi, active, wl);
    return check_launch("batched fill kernel");
  }
}

// Work lists of many plans (one workgroup per plan): the jobs whose instance is active,
// in job order, and the exclusive prefix of their block counts (a block-wide scan of
// 256 jobs at a time: ballot-free shuffle scans per wavefront, wavefront totals in LDS).
__global__ void __launch_bounds__(256) k_batch_compact(const xb_batch_plan *__restrict__ plans,
                                                       const int32_t *__restrict__ active) {
  const xb_batch_plan P = plans[blockIdx.x];
  if (P.work_offset <= 0 || P.jobinfo_offset <= 0) return; // uniform per workgroup
  const int32_t *jnb = reinterpret_cast<const int32_t *>(static_cast<const char *>(P.table) + P.jobinfo_offset) +
                       P.njobs;
  const int32_t *jin = jnb + P.njobs;
  int32_t *ajob = reinterpret_cast<int32_t *>(static_cast<char *>(P.table) + P.work_offset);
  int32_t *aoff = ajob + P.njobs;
  int32_t *naj = aoff + P.njobs + 1;
  __shared__ int32_t s_cnt[4], s_nb[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int carry_cnt = 0, carry_nb = 0;
  for (int j0 = 0; j0 < P.njobs; j0 += 256) {
    const int j = j0 + (int)threadIdx.x;
    const bool on = j < P.njobs && jnb[j] > 0 && (active == nullptr || active[jin[j]] != 0);
    const int c = on ? 1 : 0, nb = on ? jnb[j] : 0;
    int ic = c, inb = nb; // inclusive scans over the wavefront
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int tc = __shfl_up(ic, o, 64), tn = __shfl_up(inb, o, 64);
      if (lane >= o) {
        ic += tc;
        inb += tn;
      }
    }
    if (lane == 63) {
      s_cnt[wv] = ic;
      s_nb[wv] = inb;
    }
    __syncthreads();
    int wc = 0, wn = 0, tc = 0, tn = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      if (w < wv) {
        wc += s_cnt[w];
        wn += s_nb[w];
      }
      tc += s_cnt[w];
      tn += s_nb[w];
    }
    if (on) {
      const int pos = carry_cnt + wc + ic - 1;
      ajob[pos] = j;
      aoff[pos] = carry_nb + wn + inb - nb;
    }
    carry_cnt += tc;
    carry_nb += tn;
    __syncthreads(); // s_cnt / s_nb are rewritten by the next chunk
  }
  if (threadIdx.x == 0) {
    aoff[carry_cnt] = carry_nb;
    *naj = carry_cnt;
  }
}

// ---------------------------------------------------------------------------
// Adjoint (transposed) steps.  A step is S(f) = PS(Stream(PC(C(f)))) with C the
// collision, PC / PS the post-collision / post-streaming rule tables and Stream the
// pull f_d(x) <- p_d(x - c_d); it is affine in f.  Given abar = dJ/d(S(f)) these
// kernels write dJ/df = C^T PC^T Stream^T PS^T abar (Stream^T pulls from x + c_d) and
// add the parameter derivatives of this step, abar . dS/dtheta at the state f, to
// FP64 accumulators.  One thread per cell, grid-stride; the only cross-cell reads are
// the 8 (9) neighbours of abar and, with walls, their rule slots.

// adjoint of the post-collision populations of cell i (Stream^T and PS^T); ab(d, j)
// is the adjoint of post-streaming population d of cell j
template <typename T, int Q, class Lat, bool HAS_BC, class AB>
__device__ __forceinline__ void adj_post_collision(const AB &ab, const BCDev<T> &bc, int nx, int ny, int x, int y,
                                                   size_t i, T (&pb)[Q]) {
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    const size_t jn = (size_t)wrapi(x + Lat::cx(d), nx) * ny + wrapi(y + Lat::cy(d), ny);
    T v = ab(d, jn);
    if constexpr (HAS_BC) {
      if (bc.code_ps != nullptr) {
        const int s = bc.slot[jn];
        if (s >= 0 && bc.code_ps[(size_t)s * Q + d] != 0) v = T(0); // the neighbour's rule replaced it
      }
    }
    pb[d] = v;
  }
  if constexpr (HAS_BC) {
    const int s = bc.slot[i];
    if (s >= 0 && bc.code_ps != nullptr) {
      // post-streaming rules of this cell read its own post-collision populations
#pragma unroll
      for (int d = 0; d < Q; ++d) {
        const int c = bc.code_ps[(size_t)s * Q + d];
        if (c != 0 && c != XB_RULE_ZERO) {
          const T a = ab(d, i);
          const int k = (c > 0 ? c : -c) - 1;
#pragma unroll
          for (int e = 0; e < Q; ++e)
            if (e == k) pb[e] = c > 0 ? pb[e] + a : pb[e] - a;
        }
      }
    }
  }
}

// PC^T: split the post-collision adjoint into the collision output's adjoint (ob) and
// the direct pre-collision contributions of the rules (fpre)
template <typename T, int Q, bool HAS_BC>
__device__ __forceinline__ void adj_pc_rules(const BCDev<T> &bc, size_t i, const T (&pb)[Q], T (&ob)[Q],
                                             T (&fpre)[Q]) {
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    ob[d] = pb[d];
    fpre[d] = T(0);
  }
  if constexpr (HAS_BC) {
    const int s = bc.slot[i];
    if (s >= 0 && bc.code_pc != nullptr) {
#pragma unroll
      for (int d = 0; d < Q; ++d) {
        const int c = bc.code_pc[(size_t)s * Q + d];
        if (c != 0) {
          ob[d] = T(0);
          if (c != XB_RULE_ZERO) {
            const int k = (c > 0 ? c : -c) - 1;
#pragma unroll
            for (int e = 0; e < Q; ++e)
              if (e == k) fpre[e] = c > 0 ? fpre[e] + pb[d] : fpre[e] - pb[d];
          }
        }
      }
    }
  }
}

// C^T of the D2Q8 moment collision at one cell: out = C^T ob + fpre.  The
// adjoints of the post-collision moments (B0..B6) and the sums S03, S47 are kept
// for the parameter and force derivatives.
template <typename T> struct ElastMomBar {
  T S03, S47, B0, B1, B2, B3, B4, B5, B6;
};

template <typename T>
__device__ __forceinline__ ElastMomBar<T> elast_collision_adjoint(const T (&ob)[8], const T (&fpre)[8], T cs, T cd,
                                                                  T h, T n, T (&out)[8]) {
  ElastMomBar<T> M;
  M.S03 = ob[0] + ob[1] + ob[2] + ob[3];
  M.S47 = ob[4] + ob[5] + ob[6] + ob[7];
  M.B0 = T(0.5) * (ob[0] - ob[2]);
  M.B1 = T(0.5) * (ob[1] - ob[3]);
  M.B2 = h * M.S03 + n * M.S47;
  M.B3 = T(0.25) * (ob[0] - ob[1] + ob[2] - ob[3]);
  M.B4 = T(0.25) * (ob[4] - ob[5] + ob[6] - ob[7]);
  M.B5 = T(0.5) * (ob[2] - ob[0]) + T(0.25) * (ob[4] - ob[5] - ob[6] + ob[7]);
  M.B6 = T(0.5) * (ob[3] - ob[1]) + T(0.25) * (ob[4] + ob[5] - ob[6] - ob[7]);
  // moments of the populations: m10, m01 (+F), m_s * cs, m_d * cd, m_11 * cd
  const T b10 = M.B0 + div_const(M.B5, T(3), T(1) / T(3)), b01 = M.B1 + div_const(M.B6, T(3), T(1) / T(3));
  const T bs = cs * M.B2, bd = cd * M.B3, b11 = cd * M.B4;
  out[0] = b10 + bs + bd + fpre[0];
  out[1] = b01 + bs - bd + fpre[1];
  out[2] = -b10 + bs + bd + fpre[2];
  out[3] = -b01 + bs - bd + fpre[3];
  out[4] = b10 + b01 + T(2) * bs + b11 + fpre[4];
  out[5] = -b10 + b01 + T(2) * bs - b11 + fpre[5];
  out[6] = -b10 - b01 + T(2) * bs + b11 + fpre[6];
  out[7] = b10 - b01 + T(2) * bs - b11 + fpre[7];
  return M;
}

// d(abar . C(f)) / d(lambda, mu) at the pre-collision populations f of one cell
template <typename T>
__device__ __forceinline__ void elast_param_bar(const ElastMomBar<T> &M, const T (&f)[8], T lam, T mu, T cs,
                                                double &g_lam, double &g_mu) {
  const T ms = f[0] + f[1] + f[2] + f[3] + T(2) * (f[4] + f[5] + f[6] + f[7]);
  const T md = f[0] - f[1] + f[2] - f[3];
  const T m11 = f[4] - f[5] + f[6] - f[7];
  const double M2 = (double)ms * (double)cs;
  const double bcs = (double)M.B2 * (double)ms;
  const double bcd = (double)M.B3 * (double)md + (double)M.B4 * (double)m11;
  const double bh = (double)M.S03 * M2, bn = (double)M.S47 * M2;
  const double lm = (double)lam + (double)mu;
  double g = 1.0 / (12.0 * lm - 4.0);
  const double dg = std::fabs(g) > 100.0 ? 0.0 : -12.0 * g * g;
  const double dcs = 6.0 / ((3.0 * lm + 1.0) * (3.0 * lm + 1.0));
  const double dcd = 12.0 / ((6.0 * (double)mu + 1.0) * (6.0 * (double)mu + 1.0));
  const double bL = bcs * dcs + (0.5 * bh - 0.25 * bn) * dg;
  g_lam = bL;
  g_mu = bL + bcd * dcd;
}

template <typename T> struct ElastAdjDev {
  int nx, ny;
  T lam, mu;
  const T *lamf, *muf;
  const T *ab_in, *f_in;
  T *ab_out;
  const T *ab_add;
  double *g_lam, *g_mu, *g_force;
  BCDev<T> bc;
};

template <typename T, bool HAS_BC>
__global__ void __launch_bounds__(256) k_elast_adjoint(const ElastAdjDev<T> A) {
  using C = ElastColl<T>;
  const int nx = A.nx, ny = A.ny;
  const size_t N = (size_t)nx * ny;
  const auto ab = [&](int d, size_t j) { return A.ab_in[d * N + j]; };
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(i / ny), y = (int)(i % ny);
    T pb[8], ob[8], fpre[8];
    adj_post_collision<T, 8, C, HAS_BC>(ab, A.bc, nx, ny, x, y, i, pb);
    adj_pc_rules<T, 8, HAS_BC>(A.bc, i, pb, ob, fpre);
    T lam = A.lam, mu = A.mu;
    if (A.lamf != nullptr) {
      lam = A.lamf[i];
      mu = A.muf[i];
    }
    T cs, cd, h, n;
    elast_coef<T>(lam, mu, cs, cd, h, n);
    T out[8];
    const ElastMomBar<T> M = elast_collision_adjoint<T>(ob, fpre, cs, cd, h, n, out);
#pragma unroll
    for (int d = 0; d < 8; ++d) A.ab_out[d * N + i] = A.ab_add != nullptr ? out[d] + A.ab_add[d * N + i] : out[d];
    if (A.g_lam != nullptr || A.g_mu != nullptr) {
      T f[8];
#pragma unroll
      for (int d = 0; d < 8; ++d) f[d] = A.f_in[d * N + i];
      double gl, gm;
      elast_param_bar<T>(M, f, lam, mu, cs, gl, gm);
      if (A.g_lam != nullptr) A.g_lam[i] += gl;
      if (A.g_mu != nullptr) A.g_mu[i] += gm;
    }
    if (A.g_force != nullptr) {
      A.g_force[i] += (double)(M.B0 + div_const(M.B5, T(6), T(1) / T(6)));
      A.g_force[N + i] += (double)(M.B1 + div_const(M.B6, T(6), T(1) / T(6)));
    }
  }
}

// Transposed multigrid sweep (smoothing or residual) of one level.  Forward:
//   SMOOTH    o = (1 - gamma) f + gamma S(f - r; F)                  (all cells)
//   RESIDUAL  c = R(f - S(f - r; F)), R the 2x2 sum over the interior box
// with r the level's residual (interior box; none on the finest level) and F the
// sweep's total force.  Given the adjoint of o (of c), the kernel writes the adjoint
// of f, adds the adjoint of r, writes the adjoint of F, and adds the Lame-parameter
// derivatives at f - r (finest level, per-cell fields).
template <typename T> struct ElastMGAdjDev {
  int nx, ny, ix0, ix1, iy0, iy1;
  T lam, mu, gamma;
  const T *lamf, *muf;
  const T *ab_in, *ab_add;
  T *ab_out, *rbar, *fbar;
  const T *f_in, *resid;
  double *g_lam, *g_mu;
  BCDev<T> bc;
};

// One cell of a transposed smoothing / residual sweep.
template <typename T, int MODE, bool HAS_BC>
__device__ __forceinline__ void elast_mg_adjoint_cell(const ElastMGAdjDev<T> &A, size_t i) {
  using C = ElastColl<T>;
  const int nx = A.nx, ny = A.ny;
  const size_t N = (size_t)nx * ny;
  const int nyi = A.iy1 - A.iy0;
  const size_t Ni = (size_t)(A.ix1 - A.ix0) * nyi;
  const int cnyi = nyi >> 1;
  const size_t Nci = Ni >> 2;
  // adjoint of the residual f - S(f - r) at cell j (zero outside the interior box)
  const auto res_bar = [&](int d, int xj, int yj) {
    if (xj < A.ix0 || xj >= A.ix1 || yj < A.iy0 || yj >= A.iy1) return T(0);
    return A.ab_in[d * Nci + (size_t)((xj - A.ix0) >> 1) * cnyi + ((yj - A.iy0) >> 1)];
  };
  // adjoint of S's output
  const auto vbar = [&](int d, size_t j) {
    if constexpr (MODE == XB_ADJ_SMOOTH) {
      return A.gamma * A.ab_in[d * N + j];
    } else {
      return -res_bar(d, (int)(j / ny), (int)(j % ny));
    }
  };
  {
    const int x = (int)(i / ny), y = (int)(i % ny);
    T pb[8], ob[8], fpre[8];
    adj_post_collision<T, 8, C, HAS_BC>(vbar, A.bc, nx, ny, x, y, i, pb);
    adj_pc_rules<T, 8, HAS_BC>(A.bc, i, pb, ob, fpre);
    T lam = A.lam, mu = A.mu;
    if (A.lamf != nullptr) {
      lam = A.lamf[i];
      mu = A.muf[i];
    }
    T cs, cd, h, n;
    elast_coef<T>(lam, mu, cs, cd, h, n);
    T yv[8];
    const ElastMomBar<T> M = elast_collision_adjoint<T>(ob, fpre, cs, cd, h, n, yv);
    const bool in_box = x >= A.ix0 && x < A.ix1 && y >= A.iy0 && y < A.iy1;
    T o[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      T v = yv[d];
      if constexpr (MODE == XB_ADJ_SMOOTH) v = v + (T(1) - A.gamma) * A.ab_in[d * N + i];
      else v = v + res_bar(d, x, y);
      if (A.ab_add != nullptr) v = v + A.ab_add[d * N + i];
      o[d] = v;
    }
#pragma unroll
    for (int d = 0; d < 8; ++d) A.ab_out[d * N + i] = o[d];
    if (A.rbar != nullptr && in_box) {
      const size_t ci = (size_t)(x - A.ix0) * nyi + (y - A.iy0);
#pragma unroll
      for (int d = 0; d < 8; ++d) A.rbar[d * Ni + ci] = A.rbar[d * Ni + ci] - yv[d];
    }
    if (A.fbar != nullptr) {
      A.fbar[i] = M.B0 + div_const(M.B5, T(6), T(1) / T(6));
      A.fbar[N + i] = M.B1 + div_const(M.B6, T(6), T(1) / T(6));
    }
    if (A.g_lam != nullptr) {
      T f[8];
#pragma unroll
      for (int d = 0; d < 8; ++d) f[d] = A.f_in[d * N + i];
      if (A.resid != nullptr && in_box) {
        const size_t ci = (size_t)(x - A.ix0) * nyi + (y - A.iy0);
#pragma unroll
        for (int d = 0; d < 8; ++d) f[d] = f[d] - A.resid[d * Ni + ci];
      }
      double gl, gm;
      elast_param_bar<T>(M, f, lam, mu, cs, gl, gm);
      A.g_lam[i] += gl;
      A.g_mu[i] += gm;
    }
  }
}

template <typename T, int MODE, bool HAS_BC>
__global__ void __launch_bounds__(256) k_elast_mg_adjoint(const ElastMGAdjDev<T> A) {
  const size_t N = (size_t)A.nx * A.ny;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x)
    elast_mg_adjoint_cell<T, MODE, HAS_BC>(A, i);
}

// Transposed lattice gradient G_s (xb_darcy_gradient with walls): qbar(x) =
// sum over the stencil slots d of the cells that read q(x) of (-3 s w_d) c_d . gbar.
// COUPLE: the transposed Biot coupling force F = G_s(D f) of a sweep instead adds
// k(x) w2_d qbar(x) to the populations' adjoint (D^T, w2 = 1 1 1 1 2 2 2 2), the
// Lame-parameter derivative of D at f, and gbar to a FP64 accumulator of the
// sweep's base force.
template <typename T> struct GradAdjDev {
  int nx, ny;
  T m3s, b, kdiv;
  const T *gbar;
  T *qbar;                // !COUPLE: qbar += b * G^T gbar
  T *fbar;                // COUPLE
  const T *f_in;          // COUPLE with g_lam
  const T *lamf, *muf;
  double *g_lam, *g_mu, *base_bar;
  BCDev<T> bc;
};

// One cell of a transposed gradient / coupling force.
template <typename T, bool HAS_BC, bool COUPLE>
__device__ __forceinline__ void grad_adjoint_cell(const GradAdjDev<T> &A, size_t i) {
  using D = DarcyColl<T>;
  const int nx = A.nx, ny = A.ny;
  const size_t N = (size_t)nx * ny;
  const auto slot_bar = [&](int d, size_t j) {
    const T cw = A.m3s * T(D::wd(d));
    T v = T(0);
    if (D::cx(d) != 0) v = v + (D::cx(d) > 0 ? cw : -cw) * A.gbar[j];
    if (D::cy(d) != 0) v = v + (D::cy(d) > 0 ? cw : -cw) * A.gbar[N + j];
    return v;
  };
  {
    const int x = (int)(i / ny), y = (int)(i % ny);
    T pb[9];
    adj_post_collision<T, 9, D, HAS_BC>(slot_bar, A.bc, nx, ny, x, y, i, pb);
    T q = pb[0];
#pragma unroll
    for (int d = 1; d < 9; ++d) q = q + pb[d];
    if constexpr (!COUPLE) {
      A.qbar[i] = A.qbar[i] + A.b * q;
    } else {
      T k = A.kdiv;
      double dk = 0.0;
      if (A.lamf != nullptr) {
        const T lm = A.lamf[i] + A.muf[i];
        k = elast_kdiv<T>(A.lamf[i], A.muf[i]);
        const double l3 = 3.0 * (double)lm + 1.0;
        dk = 4.5 / (l3 * l3);
      }
      const T kq = k * q;
#pragma unroll
      for (int d = 0; d < 8; ++d) A.fbar[d * N + i] = A.fbar[d * N + i] + (d < 4 ? kq : T(2) * kq);
      if (A.g_lam != nullptr) {
        T ms = A.f_in[i] + A.f_in[N + i] + A.f_in[2 * N + i] + A.f_in[3 * N + i];
        ms = ms + T(2) * (A.f_in[4 * N + i] + A.f_in[5 * N + i] + A.f_in[6 * N + i] + A.f_in[7 * N + i]);
        const double g = (double)q * (double)ms * dk;
        A.g_lam[i] += g;
        A.g_mu[i] += g;
      }
      if (A.base_bar != nullptr) {
        A.base_bar[i] += (double)A.gbar[i];
        A.base_bar[N + i] += (double)A.gbar[N + i];
      }
    }
  }
}

template <typename T, bool HAS_BC, bool COUPLE>
__global__ void __launch_bounds__(256) k_grad_adjoint(const GradAdjDev<T> A) {
  const size_t N = (size_t)A.nx * A.ny;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x)
    grad_adjoint_cell<T, HAS_BC, COUPLE>(A, i);
}

// `sweeps` transposed smoothing sweeps of a small grid, each followed by the
// transposed coupling force of its force adjoint, in one workgroup (global memory,
// barriers between the phases): per cell the arithmetic of the sequence of
// k_elast_mg_adjoint<SMOOTH> and k_grad_adjoint<COUPLE> launches (bit-identical).
// The sweeps alternate between tmp and A.ab_out so that the last one writes ab_out.
constexpr int ADJ_SMALL_THREADS = 1024;
constexpr int ADJ_SMALL_MAX_CELLS = 65536;

template <typename T, bool HAS_BC, bool HAS_GBC>
__global__ void __launch_bounds__(ADJ_SMALL_THREADS)
k_smooth_adjoint_small(const ElastMGAdjDev<T> A, const GradAdjDev<T> G, int sweeps, T *tmp) {
  const size_t N = (size_t)A.nx * A.ny;
  const T *in = A.ab_in;
  for (int s = 0; s < sweeps; ++s) {
    T *out = ((sweeps - 1 - s) & 1) == 0 ? A.ab_out : tmp;
    ElastMGAdjDev<T> a = A;
    a.ab_in = in;
    a.ab_out = out;
    for (size_t i = threadIdx.x; i < N; i += ADJ_SMALL_THREADS) elast_mg_adjoint_cell<T, XB_ADJ_SMOOTH, HAS_BC>(a, i);
    __syncthreads();
    GradAdjDev<T> g = G;
    g.fbar = out;
    for (size_t i = threadIdx.x; i < N; i += ADJ_SMALL_THREADS) grad_adjoint_cell<T, HAS_GBC, true>(g, i);
    __syncthreads();
    in = out;
  }
}

// The same runs with the grid's adjoint state, force adjoint and residual adjoint held
// in LDS (copied in at the start, out at the end): each phase then waits on LDS
// instead of global memory.  Same per-cell functions in the same order, so
// bit-identical to k_smooth_adjoint_small; used when (18 nx ny + 8 nbox) values fit
// ADJ_LDS_BYTES.
constexpr int ADJ_LDS_BYTES = 65536;

template <typename T, bool HAS_BC, bool HAS_GBC>
__global__ void __launch_bounds__(ADJ_SMALL_THREADS)
k_smooth_adjoint_small_lds(const ElastMGAdjDev<T> A, const GradAdjDev<T> G, int sweeps) {
  __shared__ T arena[ADJ_LDS_BYTES / sizeof(T)];
  const int N = A.nx * A.ny;
  const int Ni = (A.ix1 - A.ix0) * (A.iy1 - A.iy0);
  T *buf0 = arena, *buf1 = arena + 8 * N, *fb = arena + 16 * N;
  T *rb = A.rbar != nullptr ? arena + 18 * N : nullptr;
  for (int i = threadIdx.x; i < 8 * N; i += ADJ_SMALL_THREADS) buf0[i] = A.ab_in[i];
  if (rb != nullptr)
    for (int i = threadIdx.x; i < 8 * Ni; i += ADJ_SMALL_THREADS) rb[i] = A.rbar[i];
  __syncthreads();
  T *in = buf0;
  for (int s = 0; s < sweeps; ++s) {
    T *out = in == buf0 ? buf1 : buf0;
    ElastMGAdjDev<T> a = A;
    a.ab_in = in;
    a.ab_out = out;
    a.fbar = fb;
    a.rbar = rb;
    for (int i = threadIdx.x; i < N; i += ADJ_SMALL_THREADS) elast_mg_adjoint_cell<T, XB_ADJ_SMOOTH, HAS_BC>(a, i);
    __syncthreads();
    GradAdjDev<T> g = G;
    g.gbar = fb;
    g.fbar = out;
    for (int i = threadIdx.x; i < N; i += ADJ_SMALL_THREADS) grad_adjoint_cell<T, HAS_GBC, true>(g, i);
    __syncthreads();
    in = out;
  }
  for (int i = threadIdx.x; i < 8 * N; i += ADJ_SMALL_THREADS) A.ab_out[i] = in[i];
  for (int i = threadIdx.x; i < 2 * N; i += ADJ_SMALL_THREADS) A.fbar[i] = fb[i];
  if (rb != nullptr)
    for (int i = threadIdx.x; i < 8 * Ni; i += ADJ_SMALL_THREADS) A.rbar[i] = rb[i];
}

// coarse = P^T fine: the sum of the four children of each coarse interior cell, zero
// on the coarse boundary layers (the transpose of the injection of xb_elast_prolong_add)
template <typename T>
__global__ void k_restrict_sum(int nx, int ny, const T *__restrict__ f, T *__restrict__ c, int cnx, int cny, int ix0,
                               int ix1, int iy0, int iy1) {
  const size_t N = (size_t)nx * ny, Nc = (size_t)cnx * cny;
  const int cx1 = ix0 + ((ix1 - ix0) >> 1), cy1 = iy0 + ((iy1 - iy0) >> 1);
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < Nc; t += (size_t)gridDim.x * blockDim.x) {
    const int X = (int)(t / cny), Y = (int)(t % cny);
    const bool in = X >= ix0 && X < cx1 && Y >= iy0 && Y < cy1;
    const size_t a = in ? (size_t)(ix0 + 2 * (X - ix0)) * ny + (iy0 + 2 * (Y - iy0)) : 0;
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      T v = T(0);
      if (in) {
        const T *p = f + d * N + a;
        v = (p[0] + p[ny]) + (p[1] + p[ny + 1]);
      }
      c[d * Nc + t] = v;
    }
  }
}

// Adjoints of the pointwise maps of the Biot coupler; every output accumulates.
// stresses and displacement: fbar += (d eta / d f)^T eta_bar + (d sigma / d f)^T sigma_bar
template <typename T>
__global__ void k_elast_macro_adjoint(int64_t N, T lam, T mu, const T *__restrict__ f, const T *__restrict__ eb,
                                      const T *__restrict__ sb, T *__restrict__ fbar, const T *__restrict__ lamf,
                                      const T *__restrict__ muf, double *__restrict__ g_lam,
                                      double *__restrict__ g_mu) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    T o[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) o[d] = T(0);
    if (eb != nullptr) {
      const T e0 = eb[i], e1 = eb[N + i];
#pragma unroll
      for (int d = 0; d < 8; ++d) o[d] = T(ElastColl<T>::cx(d)) * e0 + T(ElastColl<T>::cy(d)) * e1;
    }
    if (sb != nullptr) {
      const T l = lamf != nullptr ? lamf[i] : lam, m = lamf != nullptr ? muf[i] : mu;
      const T lm = l + m;
      const T cs = T(-1.5) * lm / (T(3) * lm + T(1));
      const T cd = T(-3) * m / (T(6) * m + T(1));
      const T s0 = sb[i], s1 = sb[N + i], s2 = sb[2 * N + i];
      const T as = cs * (s0 + s2), ad = cd * (s0 - s2), a11 = T(2) * cd * s1;
      o[0] = o[0] + as + ad;
      o[1] = o[1] + as - ad;
      o[2] = o[2] + as + ad;
      o[3] = o[3] + as - ad;
      o[4] = o[4] + T(2) * as + a11;
      o[5] = o[5] + T(2) * as - a11;
      o[6] = o[6] + T(2) * as + a11;
      o[7] = o[7] + T(2) * as - a11;
      if (g_lam != nullptr) {
        T v[8];
#pragma unroll
        for (int d = 0; d < 8; ++d) v[d] = f[d * N + i];
        const double ms = (double)(v[0] + v[1] + v[2] + v[3] + T(2) * (v[4] + v[5] + v[6] + v[7]));
        const double md = (double)(v[0] - v[1] + v[2] - v[3]);
        const double m11 = (double)(v[4] - v[5] + v[6] - v[7]);
        const double l3 = 3.0 * (double)lm + 1.0, m6 = 6.0 * (double)m + 1.0;
        const double gL = ((double)s0 + (double)s2) * ms * (-1.5 / (l3 * l3));
        const double gM = (((double)s0 - (double)s2) * md + 2.0 * (double)s1 * m11) * (-3.0 / (m6 * m6));
        g_lam[i] += gL;
        g_mu[i] += gL + gM;
      }
    }
#pragma unroll
    for (int d = 0; d < 8; ++d) fbar[d * N + i] = fbar[d * N + i] + o[d];
  }
}

// div = k (m_s(f) - m_s(f_sub)): fbar += k w2 dbar, fsub_bar -= k w2 dbar
template <typename T>
__global__ void k_elast_div_adjoint(int64_t N, T k0, const T *__restrict__ f, const T *__restrict__ fs,
                                    const T *__restrict__ db, T *__restrict__ fbar, T *__restrict__ fsbar,
                                    const T *__restrict__ lamf, const T *__restrict__ muf, double *__restrict__ g_lam,
                                    double *__restrict__ g_mu) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const T k = lamf != nullptr ? elast_kdiv<T>(lamf[i], muf[i]) : k0;
    const T a = k * db[i];
    if (fbar != nullptr) {
#pragma unroll
      for (int d = 0; d < 8; ++d) fbar[d * N + i] = fbar[d * N + i] + (d < 4 ? a : T(2) * a);
    }
    if (fsbar != nullptr) {
#pragma unroll
      for (int d = 0; d < 8; ++d) fsbar[d * N + i] = fsbar[d * N + i] - (d < 4 ? a : T(2) * a);
    }
    if (g_lam != nullptr) {
      T v[8];
#pragma unroll
      for (int d = 0; d < 8; ++d) v[d] = fs != nullptr ? f[d * N + i] - fs[d * N + i] : f[d * N + i];
      const double ms = (double)(v[0] + v[1] + v[2] + v[3] + T(2) * (v[4] + v[5] + v[6] + v[7]));
      const double l3 = 3.0 * ((double)lamf[i] + (double)muf[i]) + 1.0;
      const double g = (double)db[i] * ms * (4.5 / (l3 * l3));
      g_lam[i] += g;
      g_mu[i] += g;
    }
  }
}

// p = sum f + S/2, grad = -3 omega sum c f: fbar_d += pbar + (-3 omega) c_d . gbar,
// sbar += pbar / 2, grad_omega += -3 sum_d (c_d . gbar) f_d
template <typename T>
__global__ void k_darcy_macro_adjoint(int64_t N, T omega, const T *__restrict__ omf, const T *__restrict__ f,
                                      const T *__restrict__ pb, const T *__restrict__ gb, T *__restrict__ fbar,
                                      T *__restrict__ sbar, double *__restrict__ g_omega) {
  using D = DarcyColl<T>;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const T p = pb != nullptr ? pb[i] : T(0);
    const T g0 = gb != nullptr ? gb[i] : T(0), g1 = gb != nullptr ? gb[N + i] : T(0);
    const T om = omf != nullptr ? omf[i] : omega;
    double go = 0.0;
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      const T cg = T(D::cx(d)) * g0 + T(D::cy(d)) * g1;
      if (fbar != nullptr) fbar[d * N + i] = fbar[d * N + i] + (p + T(-3) * om * cg);
      if (g_omega != nullptr && gb != nullptr) go += -3.0 * (double)cg * (double)f[d * N + i];
    }
    if (sbar != nullptr && pb != nullptr) sbar[i] = sbar[i] + T(0.5) * p;
    if (g_omega != nullptr && gb != nullptr) g_omega[i] += go;
  }
}

template <typename T> struct DarcyAdjDev {
  int nx, ny;
  T omega;
  const T *omf, *sa, *sb;
  const T *ab_in, *f_in;
  T *ab_out;
  const T *ab_add;
  double *g_omega, *g_source;
  BCDev<T> bc;
};

template <typename T, bool HAS_BC>
__global__ void __launch_bounds__(256) k_darcy_adjoint(const DarcyAdjDev<T> A) {
  using C = DarcyColl<T>;
  const int nx = A.nx, ny = A.ny;
  const size_t N = (size_t)nx * ny;
  const auto ab = [&](int d, size_t j) { return A.ab_in[d * N + j]; };
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(i / ny), y = (int)(i % ny);
    T pb[9], ob[9], fpre[9];
    adj_post_collision<T, 9, C, HAS_BC>(ab, A.bc, nx, ny, x, y, i, pb);
    adj_pc_rules<T, 9, HAS_BC>(A.bc, i, pb, ob, fpre);
    const T om = A.omf != nullptr ? A.omf[i] : A.omega;
    T Sw = T(0);
#pragma unroll
    for (int d = 0; d < 9; ++d) Sw = Sw + T(C::wd(d)) * ob[d];
    // o_d = f_d - om (f_d - rho w_d) + S (1 - om/2) w_d, rho = sum f + S/2
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      const T v = (T(1) - om) * ob[d] + om * Sw + fpre[d];
      A.ab_out[d * N + i] = A.ab_add != nullptr ? v + A.ab_add[d * N + i] : v;
    }
    if (A.g_omega != nullptr) {
      double rho = 0.0, fo = 0.0;
      T f[9];
#pragma unroll
      for (int d = 0; d < 9; ++d) {
        f[d] = A.f_in[d * N + i];
        rho += (double)f[d];
        fo += (double)f[d] * (double)ob[d];
      }
      double S = 0.0;
      if (A.sa != nullptr) S += (double)A.sa[i];
      if (A.sb != nullptr) S += (double)A.sb[i];
      rho += 0.5 * S;
      // d o_d / d om = rho w_d - f_d - S w_d / 2
      A.g_omega[i] += (double)Sw * rho - fo - 0.5 * S * (double)Sw;
    }
    if (A.g_source != nullptr) A.g_source[i] += (double)Sw;
  }
}

template <typename T> ElastMGAdjDev<T> mg_adjoint_dev(const xb_elast_mg_adjoint_args *a) {
  ElastMGAdjDev<T> A{};
  A.nx = a->nx; A.ny = a->ny;
  A.ix0 = a->ix0; A.ix1 = a->ix1; A.iy0 = a->iy0; A.iy1 = a->iy1;
  A.lam = (T)a->lam; A.mu = (T)a->mu; A.gamma = (T)a->gamma;
  A.lamf = static_cast<const T *>(a->lam_field);
  A.muf = static_cast<const T *>(a->mu_field);
  A.ab_in = static_cast<const T *>(a->abar_in);
  A.ab_add = static_cast<const T *>(a->abar_add);
  A.ab_out = static_cast<T *>(a->abar_out);
  A.rbar = static_cast<T *>(a->rbar);
  A.fbar = static_cast<T *>(a->force_bar);
  A.f_in = static_cast<const T *>(a->f_in);
  A.resid = static_cast<const T *>(a->resid);
  A.g_lam = a->grad_lam;
  A.g_mu = a->grad_mu;
  A.bc = bcdev<T>(a->bc);
  return A;
}

template <typename T> int elast_mg_adjoint_t(const xb_elast_mg_adjoint_args *a, hipStream_t st) {
  const ElastMGAdjDev<T> A = mg_adjoint_dev<T>(a);
  const bool has_bc = A.bc.slot != nullptr;
  const dim3 gr = ew_grid((int64_t)a->nx * a->ny);
  if (a->mode == XB_ADJ_SMOOTH) {
    if (has_bc) hipLaunchKernelGGL((k_elast_mg_adjoint<T, XB_ADJ_SMOOTH, true>), gr, dim3(256), 0, st, A);
    else hipLaunchKernelGGL((k_elast_mg_adjoint<T, XB_ADJ_SMOOTH, false>), gr, dim3(256), 0, st, A);
  } else {
    if (has_bc) hipLaunchKernelGGL((k_elast_mg_adjoint<T, XB_ADJ_RESIDUAL, true>), gr, dim3(256), 0, st, A);
    else hipLaunchKernelGGL((k_elast_mg_adjoint<T, XB_ADJ_RESIDUAL, false>), gr, dim3(256), 0, st, A);
  }
  return check_launch("xb_elast_mg_adjoint");
}

template <typename T>
int grad_adjoint_t(int nx, int ny, double scale, double b, double kdiv, const void *gbar, void *qbar, void *fbar,
                   const void *f_in, const void *lamf, const void *muf, double *g_lam, double *g_mu,
                   double *base_bar, const xb_bc_table *bct, bool couple, hipStream_t st) {
  GradAdjDev<T> A{};
  A.nx = nx; A.ny = ny;
  A.m3s = T(-3 * scale);
  A.b = T(b);
  A.kdiv = T(kdiv);
  A.gbar = static_cast<const T *>(gbar);
  A.qbar = static_cast<T *>(qbar);
  A.fbar = static_cast<T *>(fbar);
  A.f_in = static_cast<const T *>(f_in);
  A.lamf = static_cast<const T *>(lamf);
  A.muf = static_cast<const T *>(muf);
  A.g_lam = g_lam;
  A.g_mu = g_mu;
  A.base_bar = base_bar;
  A.bc = bcdev<T>(bct);
  const bool has_bc = A.bc.slot != nullptr;
  const dim3 gr = ew_grid((int64_t)nx * ny);
  if (couple) {
    if (has_bc) hipLaunchKernelGGL((k_grad_adjoint<T, true, true>), gr, dim3(256), 0, st, A);
    else hipLaunchKernelGGL((k_grad_adjoint<T, false, true>), gr, dim3(256), 0, st, A);
  } else {
    if (has_bc) hipLaunchKernelGGL((k_grad_adjoint<T, true, false>), gr, dim3(256), 0, st, A);
    else hipLaunchKernelGGL((k_grad_adjoint<T, false, false>), gr, dim3(256), 0, st, A);
  }
  return check_launch(couple ? "xb_elast_coupling_adjoint" : "xb_darcy_gradient_adjoint");
}

template <typename T>
int smooth_adjoint_small_t(const xb_elast_mg_adjoint_args *a, int sweeps, double scale, const xb_bc_table *gbc,
                           void *tmp, hipStream_t st) {
  const ElastMGAdjDev<T> A = mg_adjoint_dev<T>(a);
  GradAdjDev<T> G{};
  G.nx = a->nx; G.ny = a->ny;
  G.m3s = T(-3 * scale);
  G.b = T(1);
  G.kdiv = T(-1.5 / (3 * (a->lam + a->mu) + 1));
  G.gbar = static_cast<const T *>(a->force_bar);
  G.lamf = static_cast<const T *>(a->lam_field);
  G.muf = static_cast<const T *>(a->mu_field);
  G.bc = bcdev<T>(gbc);
  T *t = static_cast<T *>(tmp);
  const bool hb = A.bc.slot != nullptr, hg = G.bc.slot != nullptr;
  const dim3 gr(1), bl(ADJ_SMALL_THREADS);
  const int64_t N = (int64_t)a->nx * a->ny, Ni = (int64_t)(a->ix1 - a->ix0) * (a->iy1 - a->iy0);
  if ((18 * N + (a->rbar != nullptr ? 8 * Ni : 0)) * (int64_t)sizeof(T) <= ADJ_LDS_BYTES) {
    if (hb && hg) hipLaunchKernelGGL((k_smooth_adjoint_small_lds<T, true, true>), gr, bl, 0, st, A, G, sweeps);
    else if (hb) hipLaunchKernelGGL((k_smooth_adjoint_small_lds<T, true, false>), gr, bl, 0, st, A, G, sweeps);
    else if (hg) hipLaunchKernelGGL((k_smooth_adjoint_small_lds<T, false, true>), gr, bl, 0, st, A, G, sweeps);
    else hipLaunchKernelGGL((k_smooth_adjoint_small_lds<T, false, false>), gr, bl, 0, st, A, G, sweeps);
    return check_launch("xb_elast_smooth_adjoint_small");
  }
  if (hb && hg) hipLaunchKernelGGL((k_smooth_adjoint_small<T, true, true>), gr, bl, 0, st, A, G, sweeps, t);
  else if (hb) hipLaunchKernelGGL((k_smooth_adjoint_small<T, true, false>), gr, bl, 0, st, A, G, sweeps, t);
  else if (hg) hipLaunchKernelGGL((k_smooth_adjoint_small<T, false, true>), gr, bl, 0, st, A, G, sweeps, t);
  else hipLaunchKernelGGL((k_smooth_adjoint_small<T, false, false>), gr, bl, 0, st, A, G, sweeps, t);
  return check_launch("xb_elast_smooth_adjoint_small");
}

// ---------------------------------------------------------------------------
// Transposed smoothing on LDS tiles: `steps` XB_ADJ_SMOOTH transposed sweeps of one
// level in one launch, each followed by the transposed coupling force of its force
// adjoint (the arithmetic of one k_elast_mg_adjoint<SMOOTH> launch and one
// k_grad_adjoint<COUPLE> launch per sweep, in the same order, so bit-identical to
// that sequence).  A 32 x 32 region per workgroup: the transposed pull reads the
// neighbours' vbar at distance one and the transposed gradient their force adjoint,
// so the exact core shrinks by two cells per side and sweep.  Per sweep:
//   A  vbar = gamma abar to LDS, zeroed in the directions the cell's own
//      post-streaming rule replaces (the transposed pull then skips them);
//   B  gather pb (transposed pull + the rule's own-cell terms), transposed
//      post-collision rules and collision; o = yv + (1 - gamma) abar; the residual
//      adjoint (rbar -= yv, output cells, global); the force adjoint Fbar to LDS; the
//      Lame-parameter derivative at the sweep's recorded state (finest level);
//   C  qbar = sum of the transposed gradient slots over the neighbours' Fbar (minus
//      the slots their gradient rules replace, plus the cell's own rule terms);
//      abar = o + k w2 qbar; the divergence's parameter derivative; base_bar += Fbar.
// The FP64 accumulators of the output cells stay in registers over the launch.
constexpr int ADJ_TILE_MAX_STEPS = 3;
// one cell per thread: the transposed sweep holds more per cell than the forward one
constexpr int ADJ_TILE_THREADS = TILE * TILE;
constexpr int ADJ_CPT = 1;

template <typename T> struct TileAdjDev {
  int nx, ny, steps, tiles_y, ix0, ix1, iy0, iy1;
  T lam, mu, gamma, m3s, kdiv;
  const T *lamf, *muf;
  const T *ab_in;
  T *ab_out, *rbar;
  const T *state[ADJ_TILE_MAX_STEPS];
  double *g_lam, *g_mu, *base_bar;
  BCDev<T> bc, gbc;
};

template <typename T, bool HAS_BC, bool HAS_GBC, bool GRADS>
__global__ void __launch_bounds__(ADJ_TILE_THREADS) k_tile_smooth_adjoint(const TileAdjDev<T> A) {
  using C = ElastColl<T>;
  using D = DarcyColl<T>;
  constexpr int Q = 8;
  constexpr int PAD = TILE_LDS_PAD(TILE);
  constexpr int TP = TILE * TILE + 2 * PAD; // one LDS plane with its padding
  constexpr int R = 2;
  __shared__ T lv_raw[Q * TP];
  __shared__ T lf_raw[2 * TP];
  __shared__ int lgs_raw[HAS_GBC ? TP : 1];
  const int tile = xcd_block_id();
  const int nx = A.nx, ny = A.ny, steps = A.steps;
  const int inner = TILE - 2 * R * steps;
  const int bx = tile / A.tiles_y, by = tile % A.tiles_y;
  const int x0 = bx * inner - R * steps, y0 = by * inner - R * steps;
  const size_t N = (size_t)nx * ny;
  const int nyi = A.iy1 - A.iy0;
  const size_t Ni = (size_t)(A.ix1 - A.ix0) * nyi;
  const int h = R * steps;
  T ab[ADJ_CPT][Q];
  T lam[ADJ_CPT], mu[ADJ_CPT];
  int slot[ADJ_CPT], gslot[ADJ_CPT], ri[ADJ_CPT];
  size_t cell[ADJ_CPT];
  bool out[ADJ_CPT];
  double gl[ADJ_CPT], gm[ADJ_CPT], bb0[ADJ_CPT], bb1[ADJ_CPT];
  if constexpr (HAS_GBC) {
    // the padding around the slot plane reads as "no rule": the gathers of the tile's
    // border cells reach into it (their results are discarded, but a slot read there
    // indexes the rule table)
    for (int i = threadIdx.x; i < PAD; i += ADJ_TILE_THREADS) {
      lgs_raw[i] = -1;
      lgs_raw[PAD + TILE * TILE + i] = -1;
    }
  }
#pragma unroll
  for (int k = 0; k < ADJ_CPT; ++k) {
    const int c = threadIdx.x + k * ADJ_TILE_THREADS;
    const int rx = c / TILE, ry = c % TILE;
    const int gx = wrapm(x0 + rx, nx), gy = wrapm(y0 + ry, ny);
    cell[k] = (size_t)gx * ny + gy;
    out[k] = rx >= h && rx < TILE - h && ry >= h && ry < TILE - h && x0 + rx < nx && y0 + ry < ny;
#pragma unroll
    for (int d = 0; d < Q; ++d) ab[k][d] = A.ab_in[d * N + cell[k]];
    lam[k] = A.lamf != nullptr ? A.lamf[cell[k]] : A.lam;
    mu[k] = A.muf != nullptr ? A.muf[cell[k]] : A.mu;
    slot[k] = HAS_BC ? A.bc.slot[cell[k]] : -1;
    gslot[k] = HAS_GBC ? A.gbc.slot[cell[k]] : -1;
    if constexpr (HAS_GBC) lgs_raw[PAD + c] = gslot[k];
    ri[k] = -1;
    if (A.rbar != nullptr && out[k] && gx >= A.ix0 && gx < A.ix1 && gy >= A.iy0 && gy < A.iy1)
      ri[k] = (gx - A.ix0) * nyi + (gy - A.iy0);
    gl[k] = gm[k] = bb0[k] = bb1[k] = 0.0;
    if (out[k]) {
      if (GRADS) {
        gl[k] = A.g_lam[cell[k]];
        gm[k] = A.g_mu[cell[k]];
      }
      if (A.base_bar != nullptr) {
        bb0[k] = A.base_bar[cell[k]];
        bb1[k] = A.base_bar[N + cell[k]];
      }
    }
  }
  T *lv = lv_raw + PAD;
  T *lf = lf_raw + PAD;
  const int *lgs = lgs_raw + PAD;
  for (int s = 0; s < steps; ++s) {
    // A: vbar to LDS, without the directions the cell's own post-streaming rule replaces
#pragma unroll
    for (int k = 0; k < ADJ_CPT; ++k) {
      const int c = threadIdx.x + k * ADJ_TILE_THREADS;
#pragma unroll
      for (int d = 0; d < Q; ++d) {
        T v = A.gamma * ab[k][d];
        if constexpr (HAS_BC) {
          if (slot[k] >= 0 && A.bc.code_ps != nullptr && A.bc.code_ps[(size_t)slot[k] * Q + d] != 0) v = T(0);
        }
        lv[d * TP + c] = v;
      }
    }
    __syncthreads();
    // B: transposed pull, rules and collision; the force adjoint to LDS
    T msv[ADJ_CPT];
#pragma unroll
    for (int k = 0; k < ADJ_CPT; ++k) {
      const int c = threadIdx.x + k * ADJ_TILE_THREADS;
      T pb[Q];
#pragma unroll
      for (int d = 0; d < Q; ++d) pb[d] = lv[d * TP + c + (C::cx(d) * TILE + C::cy(d))];
      T ob[Q], fpre[Q];
      if constexpr (HAS_BC) {
        if (slot[k] >= 0) {
          if (A.bc.code_ps != nullptr) {
#pragma unroll
            for (int d = 0; d < Q; ++d) {
              const int cc = A.bc.code_ps[(size_t)slot[k] * Q + d];
              if (cc != 0 && cc != XB_RULE_ZERO) {
                const T a = A.gamma * ab[k][d];
                const int kk = (cc > 0 ? cc : -cc) - 1;
#pragma unroll
                for (int e = 0; e < Q; ++e)
                  if (e == kk) pb[e] = cc > 0 ? pb[e] + a : pb[e] - a;
              }
            }
          }
        }
      }
#pragma unroll
      for (int d = 0; d < Q; ++d) {
        ob[d] = pb[d];
        fpre[d] = T(0);
      }
      if constexpr (HAS_BC) {
        if (slot[k] >= 0 && A.bc.code_pc != nullptr) {
#pragma unroll
          for (int d = 0; d < Q; ++d) {
            const int cc = A.bc.code_pc[(size_t)slot[k] * Q + d];
            if (cc != 0) {
              ob[d] = T(0);
              if (cc != XB_RULE_ZERO) {
                const int kk = (cc > 0 ? cc : -cc) - 1;
#pragma unroll
                for (int e = 0; e < Q; ++e)
                  if (e == kk) fpre[e] = cc > 0 ? fpre[e] + pb[d] : fpre[e] - pb[d];
              }
            }
          }
        }
      }
      T cs, cd, hh, nn;
      elast_coef<T>(lam[k], mu[k], cs, cd, hh, nn);
      T yv[Q];
      const ElastMomBar<T> M = elast_collision_adjoint<T>(ob, fpre, cs, cd, hh, nn, yv);
#pragma unroll
      for (int d = 0; d < Q; ++d) ab[k][d] = yv[d] + (T(1) - A.gamma) * ab[k][d];
      if (ri[k] >= 0) {
#pragma unroll
        for (int d = 0; d < Q; ++d) A.rbar[d * Ni + ri[k]] = A.rbar[d * Ni + ri[k]] - yv[d];
      }
      const T fb0 = M.B0 + div_const(M.B5, T(6), T(1) / T(6));
      const T fb1 = M.B1 + div_const(M.B6, T(6), T(1) / T(6));
      lf[c] = fb0;
      lf[TP + c] = fb1;
      msv[k] = T(0);
      if constexpr (GRADS) {
        if (out[k]) {
          T f[Q];
#pragma unroll
          for (int d = 0; d < Q; ++d) f[d] = A.state[s][d * N + cell[k]];
          double a, b;
          elast_param_bar<T>(M, f, lam[k], mu[k], cs, a, b);
          gl[k] += a;
          gm[k] += b;
          T ms = f[0] + f[1] + f[2] + f[3];
          msv[k] = ms + T(2) * (f[4] + f[5] + f[6] + f[7]);
        }
      }
    }
    __syncthreads();
    // C: transposed coupling force
#pragma unroll
    for (int k = 0; k < ADJ_CPT; ++k) {
      const int c = threadIdx.x + k * ADJ_TILE_THREADS;
      T pb[9];
#pragma unroll
      for (int d = 0; d < 9; ++d) {
        const int j = c + (D::cx(d) * TILE + D::cy(d));
        const T cw = A.m3s * T(D::wd(d));
        T v = T(0);
        if (D::cx(d) != 0) v = v + (D::cx(d) > 0 ? cw : -cw) * lf[j];
        if (D::cy(d) != 0) v = v + (D::cy(d) > 0 ? cw : -cw) * lf[TP + j];
        if constexpr (HAS_GBC) {
          const int sn = lgs[j];
          if (sn >= 0 && A.gbc.code_ps != nullptr && A.gbc.code_ps[(size_t)sn * 9 + d] != 0) v = T(0);
        }
        pb[d] = v;
      }
      if constexpr (HAS_GBC) {
        if (gslot[k] >= 0 && A.gbc.code_ps != nullptr) {
#pragma unroll
          for (int d = 0; d < 9; ++d) {
            const int cc = A.gbc.code_ps[(size_t)gslot[k] * 9 + d];
            if (cc != 0 && cc != XB_RULE_ZERO) {
              const T cw = A.m3s * T(D::wd(d));
              T a = T(0);
              if (D::cx(d) != 0) a = a + (D::cx(d) > 0 ? cw : -cw) * lf[c];
              if (D::cy(d) != 0) a = a + (D::cy(d) > 0 ? cw : -cw) * lf[TP + c];
              const int kk = (cc > 0 ? cc : -cc) - 1;
#pragma unroll
              for (int e = 0; e < 9; ++e)
                if (e == kk) pb[e] = cc > 0 ? pb[e] + a : pb[e] - a;
            }
          }
        }
      }
      T q = pb[0];
#pragma unroll
      for (int d = 1; d < 9; ++d) q = q + pb[d];
      T kk = A.kdiv;
      double dk = 0.0;
      if (A.lamf != nullptr) {
        const T lm = lam[k] + mu[k];
        kk = elast_kdiv<T>(lam[k], mu[k]);
        const double l3 = 3.0 * (double)lm + 1.0;
        dk = 4.5 / (l3 * l3);
      }
      const T kq = kk * q;
#pragma unroll
      for (int d = 0; d < Q; ++d) ab[k][d] = ab[k][d] + (d < 4 ? kq : T(2) * kq);
      if (out[k]) {
        if constexpr (GRADS) {
          const double g = (double)q * (double)msv[k] * dk;
          gl[k] += g;
          gm[k] += g;
        }
        bb0[k] += (double)lf[c];
        bb1[k] += (double)lf[TP + c];
      }
    }
    // no barrier here: the next sweep's A writes lv, whose last reads (B) precede the
    // barrier before C, and its B writes lf only after the barrier that follows A,
    // which every thread reaches after its reads of lf in C
  }
#pragma unroll
  for (int k = 0; k < ADJ_CPT; ++k) {
    if (!out[k]) continue;
#pragma unroll
    for (int d = 0; d < Q; ++d) A.ab_out[d * N + cell[k]] = ab[k][d];
    if (GRADS) {
      A.g_lam[cell[k]] = gl[k];
      A.g_mu[cell[k]] = gm[k];
    }
    if (A.base_bar != nullptr) {
      A.base_bar[cell[k]] = bb0[k];
      A.base_bar[N + cell[k]] = bb1[k];
    }
  }
}

template <typename T>
int smooth_adjoint_tiles_t(const xb_elast_mg_adjoint_args *a, int steps, double scale, const xb_bc_table *gbc,
                           const void *const *states, double *base_bar, hipStream_t st) {
  TileAdjDev<T> A{};
  A.nx = a->nx; A.ny = a->ny; A.steps = steps;
  A.ix0 = a->ix0; A.ix1 = a->ix1; A.iy0 = a->iy0; A.iy1 = a->iy1;
  A.lam = (T)a->lam; A.mu = (T)a->mu; A.gamma = (T)a->gamma;
  A.m3s = T(-3 * scale);
  A.kdiv = T(-1.5 / (3 * (a->lam + a->mu) + 1));
  A.lamf = static_cast<const T *>(a->lam_field);
  A.muf = static_cast<const T *>(a->mu_field);
  A.ab_in = static_cast<const T *>(a->abar_in);
  A.ab_out = static_cast<T *>(a->abar_out);
  A.rbar = static_cast<T *>(a->rbar);
  for (int s = 0; s < steps; ++s) A.state[s] = states != nullptr ? static_cast<const T *>(states[s]) : nullptr;
  A.g_lam = a->grad_lam;
  A.g_mu = a->grad_mu;
  A.base_bar = base_bar;
  A.bc = bcdev<T>(a->bc);
  A.gbc = bcdev<T>(gbc);
  const int inner = TILE - 4 * steps;
  A.tiles_y = cdiv(a->ny, inner);
  const dim3 gr((unsigned)(cdiv(a->nx, inner) * A.tiles_y)), bl(ADJ_TILE_THREADS);
  const bool hb = A.bc.slot != nullptr, hg = A.gbc.slot != nullptr, gr_ = a->grad_lam != nullptr;
#define XB_ADJ_TILE(HB, HG, GR) hipLaunchKernelGGL((k_tile_smooth_adjoint<T, HB, HG, GR>), gr, bl, 0, st, A)
  if (gr_) {
    if (hb && hg) XB_ADJ_TILE(true, true, true);
    else if (hb) XB_ADJ_TILE(true, false, true);
    else if (hg) XB_ADJ_TILE(false, true, true);
    else XB_ADJ_TILE(false, false, true);
  } else {
    if (hb && hg) XB_ADJ_TILE(true, true, false);
    else if (hb) XB_ADJ_TILE(true, false, false);
    else if (hg) XB_ADJ_TILE(false, true, false);
    else XB_ADJ_TILE(false, false, false);
  }
#undef XB_ADJ_TILE
  return check_launch("xb_elast_smooth_adjoint_tiles");
}

// ---------------------------------------------------------------------------
// Coarse tail of a V-cycle in one workgroup.  Below a few thousand cells a level's
// sweep is bound by launch and load latency (a wavefront walks its rows one after
// another), so the levels L0.. of a V-cycle run as phases of one block separated by
// barriers: per sweep, every cell collides (phase A, post-collision populations into
// a scratch array), then every cell pulls from its neighbours and applies its rules
// (phase B).  The per-cell arithmetic is that of sweep_item, in the same order, so
// the result is bit-identical to the level-by-level launches.
constexpr int TAIL_THREADS = 512;
constexpr int TAIL_MAX_LEVELS = 16;

template <typename T> struct TailLevelDev {
  int nx, ny, ix0, ix1, iy0, iy1;
  BCDev<T> bc, gbc;
  T *f[2], *div[2];
  T *resid, *pc;
};

template <typename T> struct TailDev {
  T cs, cd, h, n, kdiv, gamma, omg, m3s;
};

// phase A of a sweep of level L from state s: pc = PC(C(f - r; F)), F the coupling force
template <typename T, bool COUPLED>
__device__ void tail_collide(const TailLevelDev<T> &L, const TailDev<T> &P, int s, bool has_r) {
  using D = DarcyColl<T>;
  ElastColl<T> coll;
  coll.cs = P.cs; coll.cd = P.cd; coll.h = P.h; coll.n = P.n;
  coll.has_b = COUPLED;
  const int nx = L.nx, ny = L.ny, N = nx * ny;
  const int nyi = L.iy1 - L.iy0;
  const int Ni = (L.ix1 - L.ix0) * nyi;
  const T *f = L.f[s];
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const int x = i / ny, y = i % ny;
    T pre[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) pre[d] = f[d * N + i];
    if (has_r && x >= L.ix0 && x < L.ix1 && y >= L.iy0 && y < L.iy1) {
      const int ci = (x - L.ix0) * nyi + (y - L.iy0);
#pragma unroll
      for (int d = 0; d < 8; ++d) pre[d] = pre[d] - L.resid[d * Ni + ci];
    }
    typename ElastColl<T>::Aux ax{T(0), T(0), T(0), T(0), T(0), T(0)};
    if constexpr (COUPLED) {
      const T *dv = L.div[s];
      T qs[9];
#pragma unroll
      for (int d = 0; d < 9; ++d) qs[d] = dv[wrapi(x - D::cx(d), nx) * ny + wrapi(y - D::cy(d), ny)];
      const int gs = L.gbc.slot != nullptr ? L.gbc.slot[i] : -1;
      if (gs >= 0) {
        T src[9];
#pragma unroll
        for (int d = 0; d < 9; ++d) src[d] = qs[0];
        apply_rules<T, 9>(qs, src, L.gbc.code_ps + (size_t)gs * 9, L.gbc.const_ps + (size_t)gs * 9);
      }
      stencil_grad<T>(qs, P.m3s, ax.b0, ax.b1);
    }
    T pc[8];
    coll.collide(pre, ax, pc);
    const int sl = L.bc.slot != nullptr ? L.bc.slot[i] : -1;
    if (sl >= 0 && L.bc.code_pc != nullptr)
      apply_rules<T, 8>(pc, pre, L.bc.code_pc + (size_t)sl * 8, L.bc.const_pc + (size_t)sl * 8);
#pragma unroll
    for (int d = 0; d < 8; ++d) L.pc[d * N + i] = pc[d];
  }
}

// the streamed, post-streaming-ruled populations of cell (x, y)
template <typename T>
__device__ __forceinline__ void tail_pull(const TailLevelDev<T> &L, int x, int y, T (&v)[8]) {
  using C = ElastColl<T>;
  const int nx = L.nx, ny = L.ny, N = nx * ny;
  const int i = x * ny + y;
#pragma unroll
  for (int d = 0; d < 8; ++d) v[d] = L.pc[d * N + wrapi(x - C::cx(d), nx) * ny + wrapi(y - C::cy(d), ny)];
  const int sl = L.bc.slot != nullptr ? L.bc.slot[i] : -1;
  if (sl >= 0 && L.bc.code_ps != nullptr) {
    T own[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) own[d] = L.pc[d * N + i];
    apply_rules<T, 8>(v, own, L.bc.code_ps + (size_t)sl * 8, L.bc.const_ps + (size_t)sl * 8);
  }
}

// damped smoothing sweep of level L from state s into state 1 - s
template <typename T, bool COUPLED>
__device__ void tail_smooth(const TailLevelDev<T> &L, const TailDev<T> &P, int s, bool has_r) {
  tail_collide<T, COUPLED>(L, P, s, has_r);
  __syncthreads();
  const int N = L.nx * L.ny;
  const T *f = L.f[s];
  T *o = L.f[1 - s];
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    T v[8], w[8];
    tail_pull<T>(L, i / L.ny, i % L.ny, v);
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      w[d] = P.omg * f[d * N + i] + P.gamma * v[d];
      o[d * N + i] = w[d];
    }
    if constexpr (COUPLED) L.div[1 - s][i] = elast_div_of<T>(w, P.kdiv);
  }
  __syncthreads();
}

// residual f - S(f - r) of level L, summed over 2 x 2 blocks of its interior box into
// the residual of level Lc (pairs along y first, as the sweep)
template <typename T, bool COUPLED>
__device__ void tail_residual(const TailLevelDev<T> &L, const TailLevelDev<T> &Lc, const TailDev<T> &P, int s,
                              bool has_r) {
  tail_collide<T, COUPLED>(L, P, s, has_r);
  __syncthreads();
  const int N = L.nx * L.ny;
  const T *f = L.f[s];
  const int cnyi = (L.iy1 - L.iy0) >> 1;
  const int Nc = ((L.ix1 - L.ix0) >> 1) * cnyi;
  for (int t = threadIdx.x; t < Nc; t += blockDim.x) {
    const int x0 = L.ix0 + 2 * (t / cnyi), y0 = L.iy0 + 2 * (t % cnyi);
    T pr[2][8];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      T r0[8], r1[8];
      tail_pull<T>(L, x0 + a, y0, r0);
      tail_pull<T>(L, x0 + a, y0 + 1, r1);
      const int i0 = (x0 + a) * L.ny + y0;
#pragma unroll
      for (int d = 0; d < 8; ++d) {
        const T ra = f[d * N + i0] - r0[d];
        const T rb = f[d * N + i0 + 1] - r1[d];
        pr[a][d] = ra + rb;
      }
    }
#pragma unroll
    for (int d = 0; d < 8; ++d) Lc.resid[d * Nc + t] = pr[0][d] + pr[1][d];
  }
  __syncthreads();
}

template <typename T, bool COUPLED>
__device__ void tail_fill_zero(const TailLevelDev<T> &L) {
  const int N = L.nx * L.ny;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
#pragma unroll
    for (int d = 0; d < 8; ++d) L.f[0][d * N + i] = T(0);
    if constexpr (COUPLED) L.div[0][i] = T(0);
  }
}

// f(L) += injection of f(Lc) over L's interior box; the carried divergence follows
template <typename T, bool COUPLED>
__device__ void tail_prolong(const TailLevelDev<T> &L, const TailLevelDev<T> &Lc, const TailDev<T> &P, int s) {
  const int N = L.nx * L.ny, Nc = Lc.nx * Lc.ny;
  const int nyi = L.iy1 - L.iy0;
  const int n = (L.ix1 - L.ix0) * nyi;
  T *f = L.f[s];
  const T *c = Lc.f[0];
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const int xi = t / nyi, yi = t % nyi;
    const int cell = (L.ix0 + xi) * L.ny + (L.iy0 + yi);
    const int cc = (L.ix0 + (xi >> 1)) * Lc.ny + (L.iy0 + (yi >> 1));
    T v[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      v[d] = f[d * N + cell] + c[d * Nc + cc];
      f[d * N + cell] = v[d];
    }
    if constexpr (COUPLED) L.div[s][cell] = elast_div_of<T>(v, P.kdiv);
  }
  __syncthreads();
}

// V-cycle of levels 0 .. nlev-1 of the table (level 0 takes its residual as input and
// starts from zero); the result is level 0's state 0
template <typename T, bool COUPLED>
__global__ void __launch_bounds__(TAIL_THREADS) k_vcycle_tail(const TailLevelDev<T> *__restrict__ lv, int nlev, int k,
                                                              TailDev<T> P) {
  unsigned cur = 0; // bit l: the state buffer of level l
  tail_fill_zero<T, COUPLED>(lv[0]);
  __syncthreads();
  for (int l = 0; l < nlev; ++l) {
    const TailLevelDev<T> &L = lv[l];
    for (int j = 0; j < k; ++j) {
      tail_smooth<T, COUPLED>(L, P, (cur >> l) & 1, true);
      cur ^= 1u << l;
    }
    if (l + 1 < nlev) {
      tail_residual<T, COUPLED>(L, lv[l + 1], P, (cur >> l) & 1, true);
      tail_fill_zero<T, COUPLED>(lv[l + 1]);
      cur &= ~(1u << (l + 1));
      __syncthreads();
    }
  }
  for (int l = nlev - 1; l >= 0; --l) {
    const TailLevelDev<T> &L = lv[l];
    if (l + 1 < nlev) tail_prolong<T, COUPLED>(L, lv[l + 1], P, (cur >> l) & 1);
    for (int j = 0; j < k; ++j) {
      tail_smooth<T, COUPLED>(L, P, (cur >> l) & 1, true);
      cur ^= 1u << l;
    }
  }
}

// The same V-cycle with every level's states, scratch, residual and rule tables in LDS:
// the levels' descriptors are re-pointed into an LDS arena, the rule tables and the
// first level's residual are copied in, and the first level's state 0 is copied out at
// the end.  The phases then wait on LDS instead of global memory (the per-cell
// arithmetic is unchanged, so the result is bit-identical).  Used when the levels fit
// TAIL_LDS_BYTES (tail_lds_bytes; the plan records the size, 0 when they do not).
constexpr int TAIL_LDS_BYTES = 128 * 1024;

__host__ __device__ inline size_t tail_rnd8(size_t b) { return (b + 7) & ~size_t(7); }

// LDS bytes of one level's arrays: states f0 f1 and the scratch (8 N each), the
// divergence carry (2 N, coupled), the residual (8 Ni), the rule tables present
struct TailLdsDims {
  int N, Ni, bc_slot, bc_ns, bc_pc, bc_ps, gb_slot, gb_ns;
};

__host__ __device__ inline size_t tail_level_lds_bytes(const TailLdsDims &d, size_t ts, bool coupled) {
  size_t b = 3 * tail_rnd8((size_t)8 * d.N * ts) + tail_rnd8((size_t)8 * d.Ni * ts);
  if (coupled) b += 2 * tail_rnd8((size_t)d.N * ts);
  if (d.bc_slot) b += tail_rnd8((size_t)d.N * 4);
  if (d.bc_pc) b += tail_rnd8((size_t)d.bc_ns * 8) + tail_rnd8((size_t)d.bc_ns * 8 * ts);
  if (d.bc_ps) b += tail_rnd8((size_t)d.bc_ns * 8) + tail_rnd8((size_t)d.bc_ns * 8 * ts);
  if (coupled && d.gb_slot) b += tail_rnd8((size_t)d.N * 4) + tail_rnd8((size_t)d.gb_ns * 9) +
                                 tail_rnd8((size_t)d.gb_ns * 9 * ts);
  return b;
}

template <typename T> __device__ TailLdsDims tail_dims(const TailLevelDev<T> &L) {
  TailLdsDims d;
  d.N = L.nx * L.ny;
  d.Ni = (L.ix1 - L.ix0) * (L.iy1 - L.iy0);
  d.bc_slot = L.bc.slot != nullptr;
  d.bc_ns = L.bc.nslots;
  d.bc_pc = L.bc.slot != nullptr && L.bc.code_pc != nullptr;
  d.bc_ps = L.bc.slot != nullptr && L.bc.code_ps != nullptr;
  d.gb_slot = L.gbc.slot != nullptr;
  d.gb_ns = L.gbc.nslots;
  return d;
}

template <typename U> __device__ __forceinline__ void lds_copy(U *dst, const U *src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

template <typename T, bool COUPLED>
__global__ void __launch_bounds__(TAIL_THREADS) k_vcycle_tail_lds(const TailLevelDev<T> *__restrict__ lv, int nlev,
                                                                  int k, TailDev<T> P) {
  __shared__ double arena_raw[TAIL_LDS_BYTES / sizeof(double)];
  // the re-pointed descriptors (raw storage: the descriptor type has default member
  // initialisers, which __shared__ variables may not run)
  __shared__ double ls_raw[(sizeof(TailLevelDev<T>) * TAIL_MAX_LEVELS + 7) / 8];
  TailLevelDev<T> *Ls = reinterpret_cast<TailLevelDev<T> *>(ls_raw);
  char *arena = reinterpret_cast<char *>(arena_raw);
  if (threadIdx.x == 0) {
    size_t off = 0;
    auto carve = [&](size_t bytes) {
      char *p = arena + off;
      off += tail_rnd8(bytes);
      return p;
    };
    for (int l = 0; l < nlev; ++l) {
      const TailLevelDev<T> &G = lv[l];
      TailLevelDev<T> D = G;
      const TailLdsDims d = tail_dims(G);
      D.f[0] = reinterpret_cast<T *>(carve((size_t)8 * d.N * sizeof(T)));
      D.f[1] = reinterpret_cast<T *>(carve((size_t)8 * d.N * sizeof(T)));
      D.pc = reinterpret_cast<T *>(carve((size_t)8 * d.N * sizeof(T)));
      D.resid = reinterpret_cast<T *>(carve((size_t)8 * d.Ni * sizeof(T)));
      if (COUPLED) {
        D.div[0] = reinterpret_cast<T *>(carve((size_t)d.N * sizeof(T)));
        D.div[1] = reinterpret_cast<T *>(carve((size_t)d.N * sizeof(T)));
      }
      if (d.bc_slot) D.bc.slot = reinterpret_cast<int32_t *>(carve((size_t)d.N * 4));
      if (d.bc_pc) {
        D.bc.code_pc = reinterpret_cast<int8_t *>(carve((size_t)d.bc_ns * 8));
        D.bc.const_pc = reinterpret_cast<T *>(carve((size_t)d.bc_ns * 8 * sizeof(T)));
      } else {
        D.bc.code_pc = nullptr;
      }
      if (d.bc_ps) {
        D.bc.code_ps = reinterpret_cast<int8_t *>(carve((size_t)d.bc_ns * 8));
        D.bc.const_ps = reinterpret_cast<T *>(carve((size_t)d.bc_ns * 8 * sizeof(T)));
      } else {
        D.bc.code_ps = nullptr;
      }
      if (COUPLED && d.gb_slot) {
        D.gbc.slot = reinterpret_cast<int32_t *>(carve((size_t)d.N * 4));
        D.gbc.code_ps = reinterpret_cast<int8_t *>(carve((size_t)d.gb_ns * 9));
        D.gbc.const_ps = reinterpret_cast<T *>(carve((size_t)d.gb_ns * 9 * sizeof(T)));
      }
      Ls[l] = D;
    }
  }
  __syncthreads();
  for (int l = 0; l < nlev; ++l) {
    const TailLevelDev<T> &G = lv[l];
    const TailLevelDev<T> &D = Ls[l];
    const TailLdsDims d = tail_dims(G);
    if (d.bc_slot) lds_copy(const_cast<int32_t *>(D.bc.slot), G.bc.slot, d.N);
    if (d.bc_pc) {
      lds_copy(const_cast<int8_t *>(D.bc.code_pc), G.bc.code_pc, d.bc_ns * 8);
      lds_copy(const_cast<T *>(D.bc.const_pc), G.bc.const_pc, d.bc_ns * 8);
    }
    if (d.bc_ps) {
      lds_copy(const_cast<int8_t *>(D.bc.code_ps), G.bc.code_ps, d.bc_ns * 8);
      lds_copy(const_cast<T *>(D.bc.const_ps), G.bc.const_ps, d.bc_ns * 8);
    }
    if (COUPLED && d.gb_slot) {
      lds_copy(const_cast<int32_t *>(D.gbc.slot), G.gbc.slot, d.N);
      lds_copy(const_cast<int8_t *>(D.gbc.code_ps), G.gbc.code_ps, d.gb_ns * 9);
      lds_copy(const_cast<T *>(D.gbc.const_ps), G.gbc.const_ps, d.gb_ns * 9);
    }
    if (l == 0) lds_copy(D.resid, G.resid, 8 * d.Ni);
  }
  __syncthreads();
  unsigned cur = 0; // bit l: the state buffer of level l
  tail_fill_zero<T, COUPLED>(Ls[0]);
  __syncthreads();
  for (int l = 0; l < nlev; ++l) {
    const TailLevelDev<T> &L = Ls[l];
    for (int j = 0; j < k; ++j) {
      tail_smooth<T, COUPLED>(L, P, (cur >> l) & 1, true);
      cur ^= 1u << l;
    }
    if (l + 1 < nlev) {
      tail_residual<T, COUPLED>(L, Ls[l + 1], P, (cur >> l) & 1, true);
      tail_fill_zero<T, COUPLED>(Ls[l + 1]);
      cur &= ~(1u << (l + 1));
      __syncthreads();
    }
  }
  for (int l = nlev - 1; l >= 0; --l) {
    const TailLevelDev<T> &L = Ls[l];
    if (l + 1 < nlev) tail_prolong<T, COUPLED>(L, Ls[l + 1], P, (cur >> l) & 1);
    for (int j = 0; j < k; ++j) {
      tail_smooth<T, COUPLED>(L, P, (cur >> l) & 1, true);
      cur ^= 1u << l;
    }
  }
  // the result: the first level's state (cur bit 0 is back at 0 after 2 k sweeps)
  lds_copy(lv[0].f[(cur & 1)], Ls[0].f[(cur & 1)], 8 * lv[0].nx * lv[0].ny);
}

template <typename T> TailDev<T> tail_params(const xb_tail_plan *p) {
  const double lm = p->lam + p->mu;
  double g = 1.0 / (12.0 * lm - 4.0);
  if (std::fabs(g) > 100.0) g = 0.0;
  TailDev<T> P;
  P.cs = T((3 * lm - 1) / (3 * lm + 1));
  P.cd = T((6 * p->mu - 1) / (6 * p->mu + 1));
  P.h = T(0.25 + 0.5 * g);
  P.n = T(-0.25 * g);
  P.kdiv = T(-1.5 / (3 * (p->lam + p->mu) + 1));
  P.gamma = T(p->gamma);
  P.omg = T(1.0 - p->gamma);
  P.m3s = T(-3 * p->nl_scale);
  return P;
}

} // namespace

#if XB_IN_PART(7)
int xb_part_elast_batch_step(const xb_batch_plan *p, const int32_t *active, hipStream_t st) {
  return p->dtype == XB_F64 ? launch_elast_batch<double, XB_MODE_STEP>(p, active, st)
                            : launch_elast_batch<float, XB_MODE_STEP>(p, active, st);
}
#endif
#if XB_IN_PART(8)
int xb_part_elast_batch_smooth(const xb_batch_plan *p, const int32_t *active, hipStream_t st) {
  return p->dtype == XB_F64 ? launch_elast_batch<double, XB_MODE_SMOOTH>(p, active, st)
                            : launch_elast_batch<float, XB_MODE_SMOOTH>(p, active, st);
}
#endif
#if XB_IN_PART(9)
int xb_part_elast_batch_residual(const xb_batch_plan *p, const int32_t *active, hipStream_t st) {
  return p->dtype == XB_F64 ? launch_elast_batch<double, XB_MODE_RESIDUAL>(p, active, st)
                            : launch_elast_batch<float, XB_MODE_RESIDUAL>(p, active, st);
}
#endif
#if XB_IN_PART(10)
int xb_part_smoothn_batch(const xb_batch_plan *p, const int32_t *active, hipStream_t st) {
  return p->dtype == XB_F64 ? launch_smoothn_batch<double>(p, active, st) : launch_smoothn_batch<float>(p, active, st);
}
#endif

// ============================================================================
extern "C" {

#if XB_IN_PART(0)
int xb_abi_version(void) { return XB_ABI_VERSION; }

const char *xb_last_error(void) { return xb_err_msg; }

int64_t xb_elast_partials_count(const xb_elast_args *a) {
  if (a == nullptr) return -1;
  SweepGeom g;
  elast_geom(a, g);
  return g.nitems;
}

#endif
#if XB_IN_PART(1)
int xb_elast_step(const xb_elast_args *a, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  int32_t arith = 0;
  if (const int rm = a != nullptr ? split_storage(a->dtype, arith) : 0) {
    // buffers and arithmetic in `arith`, narrower storage: the sweep with its stores rounded
    xb_elast_args b = *a;
    b.dtype = arith;
    if (int r = validate_elast(&b, "xb_elast_step")) return r;
    return arith == XB_F64 ? elast_step_t<double>(&b, st, rm) : elast_step_t<float>(&b, st, rm);
  }
  if (int r = validate_elast(a, "xb_elast_step")) return r;
  return a->dtype == XB_F64 ? elast_step_t<double>(a, st) : elast_step_t<float>(a, st);
}

int xb_elast_step2(const xb_elast_args *a, void *stream) {
  if (int r = validate_elast(a, "xb_elast_step2")) return r;
  if (a->mode != XB_MODE_STEP || a->fpc_out != nullptr || a->resid != nullptr || a->div_in != nullptr)
    return fail(XB_EUNSUPPORTED, "xb_elast_step2: plain steps only (no post-collision output, residual or "
                "coupling force)");
  if (a->nx < 2 || a->ny < 2) return fail(XB_EINVAL, "xb_elast_step2: grid must be at least 2 x 2");
  if (a->f_in == a->f_out) return fail(XB_EINVAL, "xb_elast_step2: f_in and f_out must differ");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const SweepGeom g = make_geom(a->nx, a->ny, 0, 0, a->nx, a->ny, 0, a->nx, 0, a->ny, a->rows, STRIP2, 1);
  if (a->dtype == XB_F64) {
    ElastColl<double> c;
    NLDev<double> nl;
    elast_coll<double>(a, c, nl);
    return launch_sweep2<double>(g, c, a->f_in, a->f_out, a->bc, st);
  }
  ElastColl<float> c;
  NLDev<float> nl;
  elast_coll<float>(a, c, nl);
  return launch_sweep2<float>(g, c, a->f_in, a->f_out, a->bc, st);
}

#endif
#if XB_IN_PART(0)
int xb_elast_prolong_add(int32_t nx, int32_t ny, int32_t dtype, void *f, const void *coarse, int32_t cnx, int32_t cny,
                         int32_t ix0, int32_t ix1, int32_t iy0, int32_t iy1, void *stream) {
  const int rmode = split_storage(dtype, dtype); // narrower storage: the sums rounded
  if (int r = check_dtype(dtype)) return r;
  if (f == nullptr || coarse == nullptr) return fail(XB_EINVAL, "xb_elast_prolong_add: null pointer");
  if (ix0 < 0 || ix1 > nx || iy0 < 0 || iy1 > ny || ix0 >= ix1 || iy0 >= iy1 || ((ix1 - ix0) & 1) || ((iy1 - iy0) & 1))
    return fail(XB_EINVAL, "xb_elast_prolong_add: bad interior box");
  if (cnx != ix0 + (ix1 - ix0) / 2 + (nx - ix1) || cny != iy0 + (iy1 - iy0) / 2 + (ny - iy1))
    return fail(XB_EINVAL, "xb_elast_prolong_add: coarse grid %d x %d does not match the fine box", cnx, cny);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 gr = ew_grid((int64_t)(ix1 - ix0) * (iy1 - iy0));
  if (dtype == XB_F64)
    hipLaunchKernelGGL(k_prolong_add<double>, gr, dim3(256), 0, st, nx, ny, (double *)f, (const double *)coarse, cnx,
                       cny, ix0, ix1, iy0, iy1, rmode);
  else
    hipLaunchKernelGGL(k_prolong_add<float>, gr, dim3(256), 0, st, nx, ny, (float *)f, (const float *)coarse, cnx, cny,
                       ix0, ix1, iy0, iy1, rmode);
  return check_launch("xb_elast_prolong_add");
}

int xb_elast_prolong_add_div(int32_t nx, int32_t ny, int32_t dtype, double lam, double mu, void *f,
                             const void *coarse, int32_t cnx, int32_t cny, int32_t ix0, int32_t ix1, int32_t iy0,
                             int32_t iy1, void *div, const void *lam_field, const void *mu_field, void *stream) {
  const int rmode = split_storage(dtype, dtype); // narrower storage: the sums rounded
  if (int r = check_dtype(dtype)) return r;
  if (f == nullptr || coarse == nullptr || div == nullptr)
    return fail(XB_EINVAL, "xb_elast_prolong_add_div: null pointer");
  if ((lam_field == nullptr) != (mu_field == nullptr))
    return fail(XB_EINVAL, "xb_elast_prolong_add_div: lam_field and mu_field go together");
  if (ix0 < 0 || ix1 > nx || iy0 < 0 || iy1 > ny || ix0 >= ix1 || iy0 >= iy1 || ((ix1 - ix0) & 1) || ((iy1 - iy0) & 1))
    return fail(XB_EINVAL, "xb_elast_prolong_add_div: bad interior box");
  if (cnx != ix0 + (ix1 - ix0) / 2 + (nx - ix1) || cny != iy0 + (iy1 - iy0) / 2 + (ny - iy1))
    return fail(XB_EINVAL, "xb_elast_prolong_add_div: coarse grid %d x %d does not match the fine box", cnx, cny);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 gr = ew_grid((int64_t)(ix1 - ix0) * (iy1 - iy0));
  const double k = -1.5 / (3 * (lam + mu) + 1);
  if (dtype == XB_F64)
    hipLaunchKernelGGL(k_prolong_add_div<double>, gr, dim3(256), 0, st, nx, ny, (double *)f, (const double *)coarse,
                       cnx, cny, ix0, ix1, iy0, iy1, k, (double *)div, (const double *)lam_field,
                       (const double *)mu_field, rmode);
  else
    hipLaunchKernelGGL(k_prolong_add_div<float>, gr, dim3(256), 0, st, nx, ny, (float *)f, (const float *)coarse, cnx,
                       cny, ix0, ix1, iy0, iy1, (float)k, (float *)div, (const float *)lam_field,
                       (const float *)mu_field, rmode);
  return check_launch("xb_elast_prolong_add_div");
}

int xb_elast_macro(int32_t nx, int32_t ny, int32_t dtype, double lam, double mu, const void *f, const void *force_a,
                   const void *force_b, void *eta_out, void *sigma_out, const void *lam_field, const void *mu_field,
                   void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (nx <= 0 || ny <= 0 || f == nullptr) return fail(XB_EINVAL, "xb_elast_macro: bad arguments");
  if ((lam_field == nullptr) != (mu_field == nullptr))
    return fail(XB_EINVAL, "xb_elast_macro: lam_field and mu_field go together");
  const int64_t N = (int64_t)nx * ny;
  const double cs = -1.5 * (lam + mu) / (3 * (lam + mu) + 1);
  const double cd = -3 * mu / (6 * mu + 1);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == XB_F64)
    hipLaunchKernelGGL(k_elast_macro<double>, ew_grid(N), dim3(256), 0, st, N, cs, cd, (const double *)f,
                       (const double *)force_a, (const double *)force_b, (double *)eta_out, (double *)sigma_out,
                       (const double *)lam_field, (const double *)mu_field);
  else
    hipLaunchKernelGGL(k_elast_macro<float>, ew_grid(N), dim3(256), 0, st, N, (float)cs, (float)cd, (const float *)f,
                       (const float *)force_a, (const float *)force_b, (float *)eta_out, (float *)sigma_out,
                       (const float *)lam_field, (const float *)mu_field);
  return check_launch("xb_elast_macro");
}

int xb_elast_divergence(int32_t nx, int32_t ny, int32_t dtype, double lam, double mu, const void *f, const void *f_sub,
                        void *div_out, const void *lam_field, const void *mu_field, void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (nx <= 0 || ny <= 0 || f == nullptr || div_out == nullptr) return fail(XB_EINVAL, "xb_elast_divergence: bad arguments");
  if ((lam_field == nullptr) != (mu_field == nullptr))
    return fail(XB_EINVAL, "xb_elast_divergence: lam_field and mu_field go together");
  const int64_t N = (int64_t)nx * ny;
  const double k = -1.5 / (3 * (lam + mu) + 1);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == XB_F64)
    hipLaunchKernelGGL(k_elast_div<double>, ew_grid(N), dim3(256), 0, st, N, k, (const double *)f,
                       (const double *)f_sub, (double *)div_out, (const double *)lam_field, (const double *)mu_field);
  else
    hipLaunchKernelGGL(k_elast_div<float>, ew_grid(N), dim3(256), 0, st, N, (float)k, (const float *)f,
                       (const float *)f_sub, (float *)div_out, (const float *)lam_field, (const float *)mu_field);
  return check_launch("xb_elast_divergence");
}

int xb_elast_equilibrium(int32_t nx, int32_t ny, int32_t dtype, double lam, double mu, const void *eta,
                         const void *sigma, const void *force_a, const void *force_b, void *f_out,
                         const void *lam_field, const void *mu_field, void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (nx <= 0 || ny <= 0 || eta == nullptr || sigma == nullptr || f_out == nullptr)
    return fail(XB_EINVAL, "xb_elast_equilibrium: bad arguments");
  if ((lam_field == nullptr) != (mu_field == nullptr))
    return fail(XB_EINVAL, "xb_elast_equilibrium: lam_field and mu_field go together");
  const int64_t N = (int64_t)nx * ny;
  double g = 1.0 / (12.0 * (lam + mu) - 4.0);
  if (std::fabs(g) > 100.0) g = 0.0;
  const double cs = -1.0 - 1 / (3 * (lam + mu));
  const double cd = -1.0 - 1 / (6 * mu);
  const double h = 0.25 + 0.5 * g, n = -0.25 * g;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == XB_F64)
    hipLaunchKernelGGL(k_elast_equilibrium<double>, ew_grid(N), dim3(256), 0, st, N, cs, cd, h, n, (const double *)eta,
                       (const double *)sigma, (const double *)force_a, (const double *)force_b, (double *)f_out,
                       (const double *)lam_field, (const double *)mu_field);
  else
    hipLaunchKernelGGL(k_elast_equilibrium<float>, ew_grid(N), dim3(256), 0, st, N, (float)cs, (float)cd, (float)h,
                       (float)n, (const float *)eta, (const float *)sigma, (const float *)force_a,
                       (const float *)force_b, (float *)f_out, (const float *)lam_field, (const float *)mu_field);
  return check_launch("xb_elast_equilibrium");
}

#endif
#if XB_IN_PART(2)
int xb_darcy_step(const xb_darcy_args *a, void *stream) {
  if (a == nullptr) return fail(XB_EINVAL, "xb_darcy_step: null args");
  int32_t dtype = 0;
  const int rmode = split_storage(a->dtype, dtype); // narrower storage: the stores rounded
  if (int r = check_dtype(dtype)) return r;
  if (a->nx <= 0 || a->ny <= 0 || a->f_in == nullptr || a->f_out == nullptr)
    return fail(XB_EINVAL, "xb_darcy_step: bad arguments");
  if (int r = check_bc(a->bc, false)) return r;
  if (a->rows != 0 && (a->rows < 1 || a->rows > 64 || (a->rows & (a->rows - 1))))
    return fail(XB_EINVAL, "xb_darcy_step: rows must be 0 or a power of two in [1, 64]");
  hipStream_t st = static_cast<hipStream_t>(stream);
  SweepGeom g = make_geom(a->nx, a->ny, 0, 0, a->nx, a->ny, 0, a->nx, 0, a->ny, a->rows, STRIP, 1);
  g.rmode = rmode;
  if (dtype == XB_F64) {
    DarcyColl<double> c;
    c.omega = a->omega;
    c.sfac = 1 - 0.5 * a->omega;
    c.om_field = static_cast<const double *>(a->omega_field);
    c.sa = static_cast<const double *>(a->source_a);
    c.sb = static_cast<const double *>(a->source_b);
    return launch_sweep<double, DarcyColl<double>, XB_MODE_STEP>(g, c, a->f_in, a->f_out, a->fpc_out, nullptr, a->bc,
                                                                 1.0, nullptr, nullptr, st);
  }
  DarcyColl<float> c;
  c.omega = (float)a->omega;
  c.sfac = (float)(1 - 0.5 * a->omega);
  c.om_field = static_cast<const float *>(a->omega_field);
  c.sa = static_cast<const float *>(a->source_a);
  c.sb = static_cast<const float *>(a->source_b);
  return launch_sweep<float, DarcyColl<float>, XB_MODE_STEP>(g, c, a->f_in, a->f_out, a->fpc_out, nullptr, a->bc, 1.0,
                                                             nullptr, nullptr, st);
}

int xb_darcy_step2(const xb_darcy_args *a, void *stream) {
  if (a == nullptr) return fail(XB_EINVAL, "xb_darcy_step2: null args");
  if (int r = check_dtype(a->dtype)) return r;
  if (a->nx < 2 || a->ny < 2 || a->f_in == nullptr || a->f_out == nullptr)
    return fail(XB_EINVAL, "xb_darcy_step2: bad arguments (grid at least 2 x 2)");
  if (a->f_in == a->f_out) return fail(XB_EINVAL, "xb_darcy_step2: f_in and f_out must differ");
  if (a->fpc_out != nullptr) return fail(XB_EUNSUPPORTED, "xb_darcy_step2: no post-collision output");
  if (int r = check_bc(a->bc, false)) return r;
  if (a->rows != 0 && (a->rows < 1 || a->rows > 64 || (a->rows & (a->rows - 1))))
    return fail(XB_EINVAL, "xb_darcy_step2: rows must be 0 or a power of two in [1, 64]");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const SweepGeom g = make_geom(a->nx, a->ny, 0, 0, a->nx, a->ny, 0, a->nx, 0, a->ny, a->rows, STRIP2, 1);
  if (a->dtype == XB_F64) {
    DarcyColl<double> c;
    c.omega = a->omega;
    c.sfac = 1 - 0.5 * a->omega;
    c.om_field = static_cast<const double *>(a->omega_field);
    c.sa = static_cast<const double *>(a->source_a);
    c.sb = static_cast<const double *>(a->source_b);
    return launch_sweep2<double>(g, c, a->f_in, a->f_out, a->bc, st);
  }
  DarcyColl<float> c;
  c.omega = (float)a->omega;
  c.sfac = (float)(1 - 0.5 * a->omega);
  c.om_field = static_cast<const float *>(a->omega_field);
  c.sa = static_cast<const float *>(a->source_a);
  c.sb = static_cast<const float *>(a->source_b);
  return launch_sweep2<float>(g, c, a->f_in, a->f_out, a->bc, st);
}

#endif
#if XB_IN_PART(0)
int xb_darcy_macro(int32_t nx, int32_t ny, int32_t dtype, double omega, const void *omega_field, const void *f,
                   const void *source_a, const void *source_b, void *p_out, void *grad_out, void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (nx <= 0 || ny <= 0 || f == nullptr) return fail(XB_EINVAL, "xb_darcy_macro: bad arguments");
  const int64_t N = (int64_t)nx * ny;
  const double m3om = -3 * omega;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == XB_F64)
    hipLaunchKernelGGL(k_darcy_macro<double>, ew_grid(N), dim3(256), 0, st, N, m3om, (const double *)omega_field,
                       (const double *)f, (const double *)source_a, (const double *)source_b, (double *)p_out,
                       (double *)grad_out);
  else
    hipLaunchKernelGGL(k_darcy_macro<float>, ew_grid(N), dim3(256), 0, st, N, (float)m3om, (const float *)omega_field,
                       (const float *)f, (const float *)source_a, (const float *)source_b, (float *)p_out,
                       (float *)grad_out);
  return check_launch("xb_darcy_macro");
}

int xb_darcy_gradient(int32_t nx, int32_t ny, int32_t dtype, const void *q, double scale, const xb_bc_table *bc,
                      const void *base, double a, double b, void *out, void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (nx <= 0 || ny <= 0 || q == nullptr || out == nullptr) return fail(XB_EINVAL, "xb_darcy_gradient: bad arguments");
  if (int r = check_bc(bc, true)) return r;
  const int64_t N = (int64_t)nx * ny;
  const double m3s = -3 * scale;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == XB_F64)
    hipLaunchKernelGGL(k_darcy_gradient<double>, ew_grid(N), dim3(256), 0, st, nx, ny, (const double *)q, m3s,
                       bcdev<double>(bc), (const double *)base, a, b, (double *)out);
  else
    hipLaunchKernelGGL(k_darcy_gradient<float>, ew_grid(N), dim3(256), 0, st, nx, ny, (const float *)q, (float)m3s,
                       bcdev<float>(bc), (const float *)base, (float)a, (float)b, (float *)out);
  return check_launch("xb_darcy_gradient");
}

int xb_lincomb(int64_t n, int32_t dtype, double a, const void *x, double b, const void *y, void *out, void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (n < 0 || x == nullptr || out == nullptr) return fail(XB_EINVAL, "xb_lincomb: bad arguments");
  if (n == 0) return XB_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == XB_F64)
    hipLaunchKernelGGL(k_lincomb<double>, ew_grid(n), dim3(256), 0, st, n, a, (const double *)x, b, (const double *)y,
                       (double *)out);
  else
    hipLaunchKernelGGL(k_lincomb<float>, ew_grid(n), dim3(256), 0, st, n, (float)a, (const float *)x, (float)b,
                       (const float *)y, (float *)out);
  return check_launch("xb_lincomb");
}

int xb_fill(int64_t n, int32_t dtype, double value, void *out, void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (n < 0 || (n > 0 && out == nullptr)) return fail(XB_EINVAL, "xb_fill: bad arguments");
  if (n == 0) return XB_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == XB_F64)
    hipLaunchKernelGGL(k_fill<double>, ew_grid(n), dim3(256), 0, st, n, value, (double *)out);
  else
    hipLaunchKernelGGL(k_fill<float>, ew_grid(n), dim3(256), 0, st, n, (float)value, (float *)out);
  return check_launch("xb_fill");
}

int xb_sqdiff_partials(int64_t n, int32_t dtype, const void *a, const void *b, double *partials, int32_t max_partials,
                       int32_t *count, void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (n <= 0 || a == nullptr || partials == nullptr || max_partials <= 0 || count == nullptr)
    return fail(XB_EINVAL, "xb_sqdiff_partials: bad arguments");
  int64_t blocks = (n + 255) / 256;
  blocks = std::min<int64_t>(blocks, max_partials);
  blocks = std::min<int64_t>(blocks, 2048);
  *count = (int32_t)blocks;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == XB_F64)
    hipLaunchKernelGGL(k_sqdiff<double>, dim3((unsigned)blocks), dim3(256), 0, st, n, (const double *)a,
                       (const double *)b, partials);
  else
    hipLaunchKernelGGL(k_sqdiff<float>, dim3((unsigned)blocks), dim3(256), 0, st, n, (const float *)a,
                       (const float *)b, partials);
  return check_launch("xb_sqdiff_partials");
}

int xb_norm_finalize(const double *partials, int32_t n, double *out, void *stream) {
  if (partials == nullptr || out == nullptr || n <= 0) return fail(XB_EINVAL, "xb_norm_finalize: bad arguments");
  hipLaunchKernelGGL(k_norm_finalize, dim3(1), dim3(NORM_THREADS), 0, static_cast<hipStream_t>(stream), partials, n, out);
  return check_launch("xb_norm_finalize");
}

#endif
#if XB_IN_PART(3)
int64_t xb_batch_table_bytes(int32_t kind, int32_t dtype, const void *jobs, int32_t njobs) {
  if (check_dtype(dtype) != XB_OK || jobs == nullptr || njobs <= 0) return -1;
  if (kind < XB_BATCH_ELAST_STEP || kind > XB_BATCH_ELAST_SMOOTHN) return -1;
  int64_t need = -1;
  const int r = guarded("xb_batch_table_bytes", [&] {
    return dtype == XB_F64 ? batch_prepare_t<double>(kind, jobs, nullptr, njobs, nullptr, 0, nullptr, nullptr, &need)
                           : batch_prepare_t<float>(kind, jobs, nullptr, njobs, nullptr, 0, nullptr, nullptr, &need);
  });
  return r == XB_OK ? need : -1;
}

int xb_batch_prepare(int32_t kind, int32_t dtype, const void *jobs, const int32_t *inst, int32_t njobs, void *table,
                     int64_t table_bytes, xb_batch_plan *plan, void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (jobs == nullptr || njobs <= 0 || plan == nullptr) return fail(XB_EINVAL, "xb_batch_prepare: bad arguments");
  if (kind < XB_BATCH_ELAST_STEP || kind > XB_BATCH_ELAST_SMOOTHN)
    return fail(XB_EINVAL, "xb_batch_prepare: bad kind %d", kind);
  if ((kind == XB_BATCH_ELAST_STEP || kind == XB_BATCH_ELAST_SMOOTHN) &&
      static_cast<const xb_elast_args *>(jobs)[0].dtype != dtype)
    return fail(XB_EINVAL, "xb_batch_prepare: job dtype differs from the batch dtype");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return guarded("xb_batch_prepare", [&] {
    return dtype == XB_F64 ? batch_prepare_t<double>(kind, jobs, inst, njobs, table, table_bytes, plan, st, nullptr)
                           : batch_prepare_t<float>(kind, jobs, inst, njobs, table, table_bytes, plan, st, nullptr);
  });
}

int xb_batch_launch(const xb_batch_plan *plan, const int32_t *active, void *stream) {
  if (plan == nullptr || plan->table == nullptr) return fail(XB_EINVAL, "xb_batch_launch: no plan");
  if (plan->kind < XB_BATCH_ELAST_STEP || plan->kind > XB_BATCH_ELAST_SMOOTHN || plan->kind == XB_BATCH_NORM)
    return fail(XB_EINVAL, "xb_batch_launch: plan kind %d is not launchable", plan->kind);
  if (plan->nblocks == 0) return XB_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  return plan->dtype == XB_F64 ? batch_launch_t<double>(plan, active, st) : batch_launch_t<float>(plan, active, st);
}

int xb_batch_compact(const xb_batch_plan *plans, int32_t nplans, const int32_t *active, void *stream) {
  // `plans` is a device array: its contents are read by the kernel, not here
  if (plans == nullptr || nplans <= 0 || nplans > 65535)
    return fail(XB_EINVAL, "xb_batch_compact: needs a device array of 1 to 65535 plans");
  hipLaunchKernelGGL(k_batch_compact, dim3(nplans), dim3(256), 0, static_cast<hipStream_t>(stream), plans, active);
  return check_launch("xb_batch_compact");
}

int xb_batch_converge(const xb_batch_plan *plan, int32_t ninst, const xb_converge_state *state, double tol,
                      int32_t max_cycles, void *stream) {
  if (plan == nullptr || plan->kind != XB_BATCH_NORM || plan->table == nullptr)
    return fail(XB_EINVAL, "xb_batch_converge: needs a norm plan");
  if (state == nullptr || state->r0 == nullptr || state->rlast == nullptr || state->cycles == nullptr ||
      state->active == nullptr || state->status == nullptr || state->nactive == nullptr)
    return fail(XB_EINVAL, "xb_batch_converge: incomplete state");
  if (ninst <= 0 || max_cycles <= 0 || !(tol >= 0.0)) return fail(XB_EINVAL, "xb_batch_converge: bad arguments");
  const NormJob *jobs = reinterpret_cast<const NormJob *>(static_cast<const char *>(plan->table) + plan->jobs_offset);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (plan->njobs > 0)
    hipLaunchKernelGGL(k_batch_converge, dim3(plan->njobs), dim3(NORM_THREADS), 0, st, jobs, tol, max_cycles,
                       state->r0, state->rlast, state->cycles, state->active, state->status, state->rref);
  hipLaunchKernelGGL(k_count_active, dim3(1), dim3(NORM_THREADS), 0, st, state->active, ninst, state->nactive);
  return check_launch("xb_batch_converge");
}

#endif
#if XB_IN_PART(4)
int xb_elast_adjoint(const xb_elast_adjoint_args *a, void *stream) {
  if (a == nullptr) return fail(XB_EINVAL, "xb_elast_adjoint: null args");
  if (int r = check_dtype(a->dtype)) return r;
  if (a->nx <= 0 || a->ny <= 0 || a->abar_in == nullptr || a->abar_out == nullptr)
    return fail(XB_EINVAL, "xb_elast_adjoint: bad arguments");
  if (a->abar_in == a->abar_out) return fail(XB_EINVAL, "xb_elast_adjoint: abar_in and abar_out must differ");
  if ((a->grad_lam != nullptr || a->grad_mu != nullptr) && a->f_in == nullptr)
    return fail(XB_EINVAL, "xb_elast_adjoint: parameter gradients need f_in");
  if ((a->lam_field == nullptr) != (a->mu_field == nullptr))
    return fail(XB_EINVAL, "xb_elast_adjoint: lam_field and mu_field go together");
  if (int r = check_bc(a->bc, false)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool has_bc = a->bc != nullptr && a->bc->cell_slot != nullptr;
  const dim3 gr = ew_grid((int64_t)a->nx * a->ny);
  if (a->dtype == XB_F64) {
    ElastAdjDev<double> A{a->nx, a->ny, a->lam, a->mu, (const double *)a->lam_field, (const double *)a->mu_field,
                          (const double *)a->abar_in, (const double *)a->f_in, (double *)a->abar_out,
                          (const double *)a->abar_add, a->grad_lam, a->grad_mu, a->grad_force, bcdev<double>(a->bc)};
    if (has_bc) hipLaunchKernelGGL((k_elast_adjoint<double, true>), gr, dim3(256), 0, st, A);
    else hipLaunchKernelGGL((k_elast_adjoint<double, false>), gr, dim3(256), 0, st, A);
  } else {
    ElastAdjDev<float> A{a->nx, a->ny, (float)a->lam, (float)a->mu, (const float *)a->lam_field,
                         (const float *)a->mu_field, (const float *)a->abar_in, (const float *)a->f_in,
                         (float *)a->abar_out, (const float *)a->abar_add, a->grad_lam, a->grad_mu, a->grad_force,
                         bcdev<float>(a->bc)};
    if (has_bc) hipLaunchKernelGGL((k_elast_adjoint<float, true>), gr, dim3(256), 0, st, A);
    else hipLaunchKernelGGL((k_elast_adjoint<float, false>), gr, dim3(256), 0, st, A);
  }
  return check_launch("xb_elast_adjoint");
}

int xb_darcy_adjoint(const xb_darcy_adjoint_args *a, void *stream) {
  if (a == nullptr) return fail(XB_EINVAL, "xb_darcy_adjoint: null args");
  if (int r = check_dtype(a->dtype)) return r;
  if (a->nx <= 0 || a->ny <= 0 || a->abar_in == nullptr || a->abar_out == nullptr)
    return fail(XB_EINVAL, "xb_darcy_adjoint: bad arguments");
  if (a->abar_in == a->abar_out) return fail(XB_EINVAL, "xb_darcy_adjoint: abar_in and abar_out must differ");
  if (a->grad_omega != nullptr && a->f_in == nullptr)
    return fail(XB_EINVAL, "xb_darcy_adjoint: the relaxation-rate gradient needs f_in");
  if (int r = check_bc(a->bc, false)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool has_bc = a->bc != nullptr && a->bc->cell_slot != nullptr;
  const dim3 gr = ew_grid((int64_t)a->nx * a->ny);
  if (a->dtype == XB_F64) {
    DarcyAdjDev<double> A{a->nx, a->ny, a->omega, (const double *)a->omega_field, (const double *)a->source_a,
                          (const double *)a->source_b, (const double *)a->abar_in, (const double *)a->f_in,
                          (double *)a->abar_out, (const double *)a->abar_add, a->grad_omega, a->grad_source,
                          bcdev<double>(a->bc)};
    if (has_bc) hipLaunchKernelGGL((k_darcy_adjoint<double, true>), gr, dim3(256), 0, st, A);
    else hipLaunchKernelGGL((k_darcy_adjoint<double, false>), gr, dim3(256), 0, st, A);
  } else {
    DarcyAdjDev<float> A{a->nx, a->ny, (float)a->omega, (const float *)a->omega_field, (const float *)a->source_a,
                         (const float *)a->source_b, (const float *)a->abar_in, (const float *)a->f_in,
                         (float *)a->abar_out, (const float *)a->abar_add, a->grad_omega, a->grad_source,
                         bcdev<float>(a->bc)};
    if (has_bc) hipLaunchKernelGGL((k_darcy_adjoint<float, true>), gr, dim3(256), 0, st, A);
    else hipLaunchKernelGGL((k_darcy_adjoint<float, false>), gr, dim3(256), 0, st, A);
  }
  return check_launch("xb_darcy_adjoint");
}

#endif
#if XB_IN_PART(3)
int xb_batch_norms(const xb_batch_plan *plan, double *out, void *stream) {
  if (plan == nullptr || plan->kind != XB_BATCH_NORM || plan->table == nullptr || out == nullptr)
    return fail(XB_EINVAL, "xb_batch_norms: needs a norm plan and an output");
  const NormJob *jobs = reinterpret_cast<const NormJob *>(static_cast<const char *>(plan->table) + plan->jobs_offset);
  if (plan->njobs > 0)
    hipLaunchKernelGGL(k_batch_norms, dim3(plan->njobs), dim3(NORM_THREADS), 0, static_cast<hipStream_t>(stream), jobs,
                       out);
  return check_launch("xb_batch_norms");
}

#endif
#if XB_IN_PART(4)
int xb_elast_mg_adjoint(const xb_elast_mg_adjoint_args *a, void *stream) {
  if (a == nullptr) return fail(XB_EINVAL, "xb_elast_mg_adjoint: null args");
  if (int r = check_dtype(a->dtype)) return r;
  if (a->mode != XB_ADJ_SMOOTH && a->mode != XB_ADJ_RESIDUAL)
    return fail(XB_EINVAL, "xb_elast_mg_adjoint: unknown mode %d", a->mode);
  if (a->nx <= 0 || a->ny <= 0 || a->abar_in == nullptr || a->abar_out == nullptr)
    return fail(XB_EINVAL, "xb_elast_mg_adjoint: bad arguments");
  if (a->abar_in == a->abar_out || (a->abar_add != nullptr && a->abar_add == a->abar_out))
    return fail(XB_EINVAL, "xb_elast_mg_adjoint: abar_out must differ from abar_in and abar_add");
  if (a->ix0 < 0 || a->ix1 > a->nx || a->iy0 < 0 || a->iy1 > a->ny || a->ix0 >= a->ix1 || a->iy0 >= a->iy1)
    return fail(XB_EINVAL, "xb_elast_mg_adjoint: bad interior box");
  if (a->mode == XB_ADJ_RESIDUAL && (((a->ix1 - a->ix0) & 1) || ((a->iy1 - a->iy0) & 1)))
    return fail(XB_EINVAL, "xb_elast_mg_adjoint: the restricted box must have even sides");
  if ((a->lam_field == nullptr) != (a->mu_field == nullptr))
    return fail(XB_EINVAL, "xb_elast_mg_adjoint: lam_field and mu_field go together");
  if ((a->grad_lam == nullptr) != (a->grad_mu == nullptr))
    return fail(XB_EINVAL, "xb_elast_mg_adjoint: grad_lam and grad_mu go together");
  if (a->grad_lam != nullptr && (a->f_in == nullptr || a->lam_field == nullptr))
    return fail(XB_EINVAL, "xb_elast_mg_adjoint: parameter gradients need f_in and the Lame fields");
  if (int r = check_bc(a->bc, false)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (a->dtype == XB_F64) return elast_mg_adjoint_t<double>(a, st);
  return elast_mg_adjoint_t<float>(a, st);
}

int xb_elast_smooth_adjoint_small(const xb_elast_mg_adjoint_args *a, int32_t sweeps, double scale,
                                  const xb_bc_table *grad_bc, void *tmp, void *stream) {
  if (a == nullptr) return fail(XB_EINVAL, "xb_elast_smooth_adjoint_small: null args");
  if (int r = check_dtype(a->dtype)) return r;
  if (a->mode != XB_ADJ_SMOOTH) return fail(XB_EINVAL, "xb_elast_smooth_adjoint_small: needs XB_ADJ_SMOOTH");
  if (a->nx <= 0 || a->ny <= 0 || (int64_t)a->nx * a->ny > ADJ_SMALL_MAX_CELLS)
    return fail(XB_EINVAL, "xb_elast_smooth_adjoint_small: grid %d x %d (at most %d cells)", a->nx, a->ny,
                ADJ_SMALL_MAX_CELLS);
  if (sweeps < 1) return fail(XB_EINVAL, "xb_elast_smooth_adjoint_small: sweeps must be >= 1");
  if (a->abar_in == nullptr || a->abar_out == nullptr || tmp == nullptr || a->force_bar == nullptr)
    return fail(XB_EINVAL, "xb_elast_smooth_adjoint_small: null abar_in, abar_out, tmp or force_bar");
  if (a->abar_in == a->abar_out || tmp == a->abar_in || tmp == a->abar_out)
    return fail(XB_EINVAL, "xb_elast_smooth_adjoint_small: abar_in, abar_out and tmp must differ");
  if (a->abar_add != nullptr || a->grad_lam != nullptr || a->grad_mu != nullptr)
    return fail(XB_EUNSUPPORTED, "xb_elast_smooth_adjoint_small: no abar_add or parameter derivatives");
  if (a->ix0 < 0 || a->ix1 > a->nx || a->iy0 < 0 || a->iy1 > a->ny || a->ix0 >= a->ix1 || a->iy0 >= a->iy1)
    return fail(XB_EINVAL, "xb_elast_smooth_adjoint_small: bad interior box");
  if ((a->lam_field == nullptr) != (a->mu_field == nullptr))
    return fail(XB_EINVAL, "xb_elast_smooth_adjoint_small: lam_field and mu_field go together");
  if (int r = check_bc(a->bc, false)) return r;
  if (int r = check_bc(grad_bc, true)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (a->dtype == XB_F64) return smooth_adjoint_small_t<double>(a, sweeps, scale, grad_bc, tmp, st);
  return smooth_adjoint_small_t<float>(a, sweeps, scale, grad_bc, tmp, st);
}

int xb_elast_smooth_adjoint_tiles(const xb_elast_mg_adjoint_args *a, int32_t steps, double scale,
                                  const xb_bc_table *grad_bc, const void *const *states, double *base_bar,
                                  void *stream) {
  const char *who = "xb_elast_smooth_adjoint_tiles";
  if (a == nullptr) return fail(XB_EINVAL, "%s: null args", who);
  if (int r = check_dtype(a->dtype)) return r;
  if (a->mode != XB_ADJ_SMOOTH) return fail(XB_EINVAL, "%s: needs XB_ADJ_SMOOTH", who);
  if (a->nx <= 0 || a->ny <= 0) return fail(XB_EINVAL, "%s: bad grid %d x %d", who, a->nx, a->ny);
  if (steps < 1 || steps > ADJ_TILE_MAX_STEPS)
    return fail(XB_EINVAL, "%s: steps must be in [1, %d]", who, ADJ_TILE_MAX_STEPS);
  if (a->abar_in == nullptr || a->abar_out == nullptr || a->abar_in == a->abar_out)
    return fail(XB_EINVAL, "%s: abar_in and abar_out must be given and differ", who);
  if (a->abar_add != nullptr || a->force_bar != nullptr)
    return fail(XB_EUNSUPPORTED, "%s: no abar_add; the force adjoint stays in the tile (force_bar null)", who);
  if (a->ix0 < 0 || a->ix1 > a->nx || a->iy0 < 0 || a->iy1 > a->ny || a->ix0 >= a->ix1 || a->iy0 >= a->iy1)
    return fail(XB_EINVAL, "%s: bad interior box", who);
  if ((a->lam_field == nullptr) != (a->mu_field == nullptr))
    return fail(XB_EINVAL, "%s: lam_field and mu_field go together", who);
  if ((a->grad_lam == nullptr) != (a->grad_mu == nullptr))
    return fail(XB_EINVAL, "%s: grad_lam and grad_mu go together", who);
  if (a->grad_lam != nullptr) {
    if (a->lam_field == nullptr || states == nullptr || a->resid != nullptr)
      return fail(XB_EINVAL, "%s: parameter gradients need the Lame fields and the sweeps' states (and no "
                  "residual)", who);
    for (int s = 0; s < steps; ++s)
      if (states[s] == nullptr) return fail(XB_EINVAL, "%s: null state of sweep %d", who, s);
  }
  if (a->rbar != nullptr && (a->rbar == a->abar_in || a->rbar == a->abar_out))
    return fail(XB_EINVAL, "%s: rbar must not alias abar_in or abar_out", who);
  if (int r = check_bc(a->bc, false)) return r;
  if (int r = check_bc(grad_bc, true)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (a->dtype == XB_F64) return smooth_adjoint_tiles_t<double>(a, steps, scale, grad_bc, states, base_bar, st);
  return smooth_adjoint_tiles_t<float>(a, steps, scale, grad_bc, states, base_bar, st);
}

int xb_elast_coupling_adjoint(int32_t nx, int32_t ny, int32_t dtype, double lam, double mu, const void *lam_field,
                              const void *mu_field, const void *force_bar, double scale, const xb_bc_table *grad_bc,
                              void *abar, const void *f_in, double *grad_lam, double *grad_mu, double *base_bar,
                              void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (nx <= 0 || ny <= 0 || force_bar == nullptr || abar == nullptr)
    return fail(XB_EINVAL, "xb_elast_coupling_adjoint: bad arguments");
  if ((lam_field == nullptr) != (mu_field == nullptr))
    return fail(XB_EINVAL, "xb_elast_coupling_adjoint: lam_field and mu_field go together");
  if ((grad_lam == nullptr) != (grad_mu == nullptr))
    return fail(XB_EINVAL, "xb_elast_coupling_adjoint: grad_lam and grad_mu go together");
  if (grad_lam != nullptr && (f_in == nullptr || lam_field == nullptr))
    return fail(XB_EINVAL, "xb_elast_coupling_adjoint: parameter gradients need f_in and the Lame fields");
  if (int r = check_bc(grad_bc, true)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const double k = -1.5 / (3 * (lam + mu) + 1);
  if (dtype == XB_F64)
    return grad_adjoint_t<double>(nx, ny, scale, 1.0, k, force_bar, nullptr, abar, f_in, lam_field, mu_field,
                                  grad_lam, grad_mu, base_bar, grad_bc, true, st);
  return grad_adjoint_t<float>(nx, ny, scale, 1.0, k, force_bar, nullptr, abar, f_in, lam_field, mu_field, grad_lam,
                               grad_mu, base_bar, grad_bc, true, st);
}

int xb_darcy_gradient_adjoint(int32_t nx, int32_t ny, int32_t dtype, const void *gbar, double scale,
                              const xb_bc_table *bc, double b, void *qbar, void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (nx <= 0 || ny <= 0 || gbar == nullptr || qbar == nullptr)
    return fail(XB_EINVAL, "xb_darcy_gradient_adjoint: bad arguments");
  if (int r = check_bc(bc, true)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == XB_F64)
    return grad_adjoint_t<double>(nx, ny, scale, b, 0.0, gbar, qbar, nullptr, nullptr, nullptr, nullptr, nullptr,
                                  nullptr, nullptr, bc, false, st);
  return grad_adjoint_t<float>(nx, ny, scale, b, 0.0, gbar, qbar, nullptr, nullptr, nullptr, nullptr, nullptr,
                               nullptr, nullptr, bc, false, st);
}

int xb_elast_restrict_sum(int32_t nx, int32_t ny, int32_t dtype, const void *f, void *coarse, int32_t cnx,
                          int32_t cny, int32_t ix0, int32_t ix1, int32_t iy0, int32_t iy1, void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (f == nullptr || coarse == nullptr) return fail(XB_EINVAL, "xb_elast_restrict_sum: null pointer");
  if (ix0 < 0 || ix1 > nx || iy0 < 0 || iy1 > ny || ix0 >= ix1 || iy0 >= iy1 || ((ix1 - ix0) & 1) || ((iy1 - iy0) & 1))
    return fail(XB_EINVAL, "xb_elast_restrict_sum: bad interior box");
  if (cnx != ix0 + (ix1 - ix0) / 2 + (nx - ix1) || cny != iy0 + (iy1 - iy0) / 2 + (ny - iy1))
    return fail(XB_EINVAL, "xb_elast_restrict_sum: coarse grid %d x %d does not match the fine box", cnx, cny);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 gr = ew_grid((int64_t)cnx * cny);
  if (dtype == XB_F64)
    hipLaunchKernelGGL(k_restrict_sum<double>, gr, dim3(256), 0, st, nx, ny, (const double *)f, (double *)coarse, cnx,
                       cny, ix0, ix1, iy0, iy1);
  else
    hipLaunchKernelGGL(k_restrict_sum<float>, gr, dim3(256), 0, st, nx, ny, (const float *)f, (float *)coarse, cnx,
                       cny, ix0, ix1, iy0, iy1);
  return check_launch("xb_elast_restrict_sum");
}

int xb_elast_macro_adjoint(int32_t nx, int32_t ny, int32_t dtype, double lam, double mu, const void *f,
                           const void *eta_bar, const void *sigma_bar, void *fbar, const void *lam_field,
                           const void *mu_field, double *grad_lam, double *grad_mu, void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (nx <= 0 || ny <= 0 || fbar == nullptr) return fail(XB_EINVAL, "xb_elast_macro_adjoint: bad arguments");
  if ((lam_field == nullptr) != (mu_field == nullptr) || (grad_lam == nullptr) != (grad_mu == nullptr))
    return fail(XB_EINVAL, "xb_elast_macro_adjoint: fields and gradients go in pairs");
  if (grad_lam != nullptr && (f == nullptr || lam_field == nullptr))
    return fail(XB_EINVAL, "xb_elast_macro_adjoint: parameter gradients need f and the Lame fields");
  const int64_t N = (int64_t)nx * ny;
  hipStream_t st = static_cast<hipStream_t>(stream);
  double *gl = sigma_bar != nullptr ? grad_lam : nullptr, *gm = sigma_bar != nullptr ? grad_mu : nullptr;
  if (dtype == XB_F64)
    hipLaunchKernelGGL(k_elast_macro_adjoint<double>, ew_grid(N), dim3(256), 0, st, N, lam, mu, (const double *)f,
                       (const double *)eta_bar, (const double *)sigma_bar, (double *)fbar,
                       (const double *)lam_field, (const double *)mu_field, gl, gm);
  else
    hipLaunchKernelGGL(k_elast_macro_adjoint<float>, ew_grid(N), dim3(256), 0, st, N, (float)lam, (float)mu,
                       (const float *)f, (const float *)eta_bar, (const float *)sigma_bar, (float *)fbar,
                       (const float *)lam_field, (const float *)mu_field, gl, gm);
  return check_launch("xb_elast_macro_adjoint");
}

int xb_elast_divergence_adjoint(int32_t nx, int32_t ny, int32_t dtype, double lam, double mu, const void *f,
                                const void *f_sub, const void *div_bar, void *fbar, void *fsub_bar,
                                const void *lam_field, const void *mu_field, double *grad_lam, double *grad_mu,
                                void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (nx <= 0 || ny <= 0 || div_bar == nullptr) return fail(XB_EINVAL, "xb_elast_divergence_adjoint: bad arguments");
  if ((lam_field == nullptr) != (mu_field == nullptr) || (grad_lam == nullptr) != (grad_mu == nullptr))
    return fail(XB_EINVAL, "xb_elast_divergence_adjoint: fields and gradients go in pairs");
  if (grad_lam != nullptr && (f == nullptr || lam_field == nullptr))
    return fail(XB_EINVAL, "xb_elast_divergence_adjoint: parameter gradients need f and the Lame fields");
  const int64_t N = (int64_t)nx * ny;
  const double k = -1.5 / (3 * (lam + mu) + 1);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == XB_F64)
    hipLaunchKernelGGL(k_elast_div_adjoint<double>, ew_grid(N), dim3(256), 0, st, N, k, (const double *)f,
                       (const double *)f_sub, (const double *)div_bar, (double *)fbar, (double *)fsub_bar,
                       (const double *)lam_field, (const double *)mu_field, grad_lam, grad_mu);
  else
    hipLaunchKernelGGL(k_elast_div_adjoint<float>, ew_grid(N), dim3(256), 0, st, N, (float)k, (const float *)f,
                       (const float *)f_sub, (const float *)div_bar, (float *)fbar, (float *)fsub_bar,
                       (const float *)lam_field, (const float *)mu_field, grad_lam, grad_mu);
  return check_launch("xb_elast_divergence_adjoint");
}

int xb_darcy_macro_adjoint(int32_t nx, int32_t ny, int32_t dtype, double omega, const void *omega_field, const void *f,
                           const void *p_bar, const void *grad_bar, void *fbar, void *source_bar, double *grad_omega,
                           void *stream) {
  if (int r = check_dtype(dtype)) return r;
  if (nx <= 0 || ny <= 0) return fail(XB_EINVAL, "xb_darcy_macro_adjoint: bad arguments");
  if (grad_omega != nullptr && grad_bar != nullptr && f == nullptr)
    return fail(XB_EINVAL, "xb_darcy_macro_adjoint: the relaxation-rate gradient needs f");
  const int64_t N = (int64_t)nx * ny;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == XB_F64)
    hipLaunchKernelGGL(k_darcy_macro_adjoint<double>, ew_grid(N), dim3(256), 0, st, N, omega,
                       (const double *)omega_field, (const double *)f, (const double *)p_bar,
                       (const double *)grad_bar, (double *)fbar, (double *)source_bar, grad_omega);
  else
    hipLaunchKernelGGL(k_darcy_macro_adjoint<float>, ew_grid(N), dim3(256), 0, st, N, (float)omega,
                       (const float *)omega_field, (const float *)f, (const float *)p_bar, (const float *)grad_bar,
                       (float *)fbar, (float *)source_bar, grad_omega);
  return check_launch("xb_darcy_macro_adjoint");
}

#endif
#if XB_IN_PART(0)
int64_t xb_vcycle_tail_bytes(int32_t nlevels) {
  if (nlevels < 1 || nlevels > TAIL_MAX_LEVELS) return 0;
  return (int64_t)sizeof(TailLevelDev<double>) * nlevels;
}

int xb_vcycle_tail_prepare(int32_t dtype, double lam, double mu, double gamma, int32_t k_smooth, double nl_scale,
                           int32_t coupled, const xb_tail_level *levels, int32_t nlevels, void *table,
                           int64_t table_bytes, xb_tail_plan *plan) {
  if (int r = check_dtype(dtype)) return r;
  if (levels == nullptr || plan == nullptr || table == nullptr || nlevels < 1 || nlevels > TAIL_MAX_LEVELS ||
      k_smooth < 1)
    return fail(XB_EINVAL, "xb_vcycle_tail_prepare: bad arguments");
  if (table_bytes < xb_vcycle_tail_bytes(nlevels))
    return fail(XB_EINVAL, "xb_vcycle_tail_prepare: table of %lld bytes, %lld needed", (long long)table_bytes,
                (long long)xb_vcycle_tail_bytes(nlevels));
  for (int l = 0; l < nlevels; ++l) {
    const xb_tail_level &L = levels[l];
    if (L.nx <= 0 || L.ny <= 0 || L.f0 == nullptr || L.f1 == nullptr || L.resid == nullptr || L.scratch == nullptr)
      return fail(XB_EINVAL, "xb_vcycle_tail_prepare: level %d: bad arguments", l);
    if (coupled && (L.div0 == nullptr || L.div1 == nullptr))
      return fail(XB_EINVAL, "xb_vcycle_tail_prepare: level %d: the coupling force needs div0 and div1", l);
    if (L.ix0 < 0 || L.ix1 > L.nx || L.iy0 < 0 || L.iy1 > L.ny || L.ix0 >= L.ix1 || L.iy0 >= L.iy1)
      return fail(XB_EINVAL, "xb_vcycle_tail_prepare: level %d: bad interior box", l);
    if (int r = check_bc(L.bc, false)) return r;
    if (int r = check_bc(L.grad_bc, true)) return r;
    if (l + 1 < nlevels) {
      const xb_tail_level &C = levels[l + 1];
      if (((L.ix1 - L.ix0) & 1) || ((L.iy1 - L.iy0) & 1) || C.ix1 - C.ix0 != (L.ix1 - L.ix0) / 2 ||
          C.iy1 - C.iy0 != (L.iy1 - L.iy0) / 2 || C.ix0 != L.ix0 || C.iy0 != L.iy0 ||
          C.nx != L.ix0 + (L.ix1 - L.ix0) / 2 + (L.nx - L.ix1) || C.ny != L.iy0 + (L.iy1 - L.iy0) / 2 + (L.ny - L.iy1))
        return fail(XB_EINVAL, "xb_vcycle_tail_prepare: level %d does not coarsen level %d", l + 1, l);
    }
  }
  auto fill = [&](auto tag) {
    using T = decltype(tag);
    std::vector<TailLevelDev<T>> v((size_t)nlevels);
    for (int l = 0; l < nlevels; ++l) {
      const xb_tail_level &L = levels[l];
      TailLevelDev<T> &D = v[(size_t)l];
      D.nx = L.nx; D.ny = L.ny; D.ix0 = L.ix0; D.ix1 = L.ix1; D.iy0 = L.iy0; D.iy1 = L.iy1;
      D.bc = bcdev<T>(L.bc);
      D.gbc = bcdev<T>(L.grad_bc);
      D.f[0] = static_cast<T *>(L.f0);
      D.f[1] = static_cast<T *>(L.f1);
      D.div[0] = static_cast<T *>(L.div0);
      D.div[1] = static_cast<T *>(L.div1);
      D.resid = static_cast<T *>(L.resid);
      D.pc = static_cast<T *>(L.scratch);
    }
    const hipError_t e = hipMemcpy(table, v.data(), sizeof(TailLevelDev<T>) * v.size(), hipMemcpyHostToDevice);
    return e == hipSuccess ? XB_OK : fail(XB_EHIP, "xb_vcycle_tail_prepare: %s", hipGetErrorString(e));
  };
  if (int r = guarded("xb_vcycle_tail_prepare", [&] { return dtype == XB_F64 ? fill(double(0)) : fill(float(0)); }))
    return r;
  plan->dtype = dtype;
  plan->nlevels = nlevels;
  plan->k_smooth = k_smooth;
  plan->coupled = coupled != 0;
  plan->lam = lam;
  plan->mu = mu;
  plan->gamma = gamma;
  plan->nl_scale = nl_scale;
  plan->table = table;
  // the LDS form when every level's arrays and rule tables fit its arena
  size_t lds = 0;
  const size_t ts = dtype == XB_F64 ? sizeof(double) : sizeof(float);
  for (int l = 0; l < nlevels; ++l) {
    const xb_tail_level &L = levels[l];
    TailLdsDims d;
    d.N = L.nx * L.ny;
    d.Ni = (L.ix1 - L.ix0) * (L.iy1 - L.iy0);
    d.bc_slot = L.bc != nullptr && L.bc->cell_slot != nullptr;
    d.bc_ns = L.bc != nullptr ? L.bc->nslots : 0;
    d.bc_pc = d.bc_slot && L.bc->code_pc != nullptr;
    d.bc_ps = d.bc_slot && L.bc->code_ps != nullptr;
    d.gb_slot = L.grad_bc != nullptr && L.grad_bc->cell_slot != nullptr;
    d.gb_ns = L.grad_bc != nullptr ? L.grad_bc->nslots : 0;
    lds += tail_level_lds_bytes(d, ts, coupled != 0);
  }
  plan->lds_bytes = lds <= (size_t)TAIL_LDS_BYTES ? (int64_t)lds : 0;
  return XB_OK;
}

int xb_vcycle_tail(const xb_tail_plan *plan, void *stream) {
  if (plan == nullptr || plan->table == nullptr || plan->nlevels < 1 || plan->nlevels > TAIL_MAX_LEVELS)
    return fail(XB_EINVAL, "xb_vcycle_tail: needs a prepared plan");
  if (int r = check_dtype(plan->dtype)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 gr(1), bl(TAIL_THREADS);
  if (plan->lds_bytes > 0) {
    if (plan->dtype == XB_F64) {
      const auto *t = static_cast<const TailLevelDev<double> *>(plan->table);
      if (plan->coupled)
        hipLaunchKernelGGL((k_vcycle_tail_lds<double, true>), gr, bl, 0, st, t, plan->nlevels, plan->k_smooth,
                           tail_params<double>(plan));
      else
        hipLaunchKernelGGL((k_vcycle_tail_lds<double, false>), gr, bl, 0, st, t, plan->nlevels, plan->k_smooth,
                           tail_params<double>(plan));
    } else {
      const auto *t = static_cast<const TailLevelDev<float> *>(plan->table);
      if (plan->coupled)
        hipLaunchKernelGGL((k_vcycle_tail_lds<float, true>), gr, bl, 0, st, t, plan->nlevels, plan->k_smooth,
                           tail_params<float>(plan));
      else
        hipLaunchKernelGGL((k_vcycle_tail_lds<float, false>), gr, bl, 0, st, t, plan->nlevels, plan->k_smooth,
                           tail_params<float>(plan));
    }
    return check_launch("xb_vcycle_tail");
  }
  if (plan->dtype == XB_F64) {
    const auto *t = static_cast<const TailLevelDev<double> *>(plan->table);
    if (plan->coupled)
      hipLaunchKernelGGL((k_vcycle_tail<double, true>), gr, bl, 0, st, t, plan->nlevels, plan->k_smooth,
                         tail_params<double>(plan));
    else
      hipLaunchKernelGGL((k_vcycle_tail<double, false>), gr, bl, 0, st, t, plan->nlevels, plan->k_smooth,
                         tail_params<double>(plan));
  } else {
    const auto *t = static_cast<const TailLevelDev<float> *>(plan->table);
    if (plan->coupled)
      hipLaunchKernelGGL((k_vcycle_tail<float, true>), gr, bl, 0, st, t, plan->nlevels, plan->k_smooth,
                         tail_params<float>(plan));
    else
      hipLaunchKernelGGL((k_vcycle_tail<float, false>), gr, bl, 0, st, t, plan->nlevels, plan->k_smooth,
                         tail_params<float>(plan));
  }
  return check_launch("xb_vcycle_tail");
}

#endif
#if XB_IN_PART(2)
int xb_darcy_stepn(const xb_darcy_args *a, int32_t steps, void *stream) {
  if (a == nullptr) return fail(XB_EINVAL, "xb_darcy_stepn: null args");
  if (int r = check_dtype(a->dtype)) return r;
  if (a->nx < 1 || a->ny < 1 || a->f_in == nullptr || a->f_out == nullptr)
    return fail(XB_EINVAL, "xb_darcy_stepn: bad arguments");
  if (steps < 1 || steps > TILE_MAX_STEPS) return fail(XB_EINVAL, "xb_darcy_stepn: steps must be in [1, %d]", TILE_MAX_STEPS);
  if (a->f_in == a->f_out) return fail(XB_EINVAL, "xb_darcy_stepn: f_in and f_out must differ");
  if (a->fpc_out != nullptr) return fail(XB_EUNSUPPORTED, "xb_darcy_stepn: no post-collision output");
  if (int r = check_bc(a->bc, false)) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (a->dtype == XB_F64) {
    DarcyColl<double> c;
    c.omega = a->omega;
    c.sfac = 1 - 0.5 * a->omega;
    c.om_field = static_cast<const double *>(a->omega_field);
    c.sa = static_cast<const double *>(a->source_a);
    c.sb = static_cast<const double *>(a->source_b);
    return launch_tile_steps<double>(a->nx, a->ny, steps, c, a->f_in, a->f_out, a->bc, a->rows, st);
  }
  DarcyColl<float> c;
  c.omega = (float)a->omega;
  c.sfac = (float)(1 - 0.5 * a->omega);
  c.om_field = static_cast<const float *>(a->omega_field);
  c.sa = static_cast<const float *>(a->source_a);
  c.sb = static_cast<const float *>(a->source_b);
  return launch_tile_steps<float>(a->nx, a->ny, steps, c, a->f_in, a->f_out, a->bc, a->rows, st);
}

#endif
#if XB_IN_PART(5)
int xb_elast_stepn(const xb_elast_args *a, int32_t steps, void *stream) {
  if (a == nullptr) return fail(XB_EINVAL, "xb_elast_stepn: null args");
  {
    // rows is the tile block size here (any value), not the sweep launch shape
    xb_elast_args b = *a;
    b.rows = 0;
    if (int r = validate_elast(&b, "xb_elast_stepn")) return r;
  }
  if (a->mode != XB_MODE_STEP || a->fpc_out != nullptr || a->resid != nullptr || a->div_in != nullptr)
    return fail(XB_EUNSUPPORTED, "xb_elast_stepn: plain steps only (no post-collision output, residual or "
                "coupling force)");
  if (steps < 1 || steps > TILE_MAX_STEPS) return fail(XB_EINVAL, "xb_elast_stepn: steps must be in [1, %d]", TILE_MAX_STEPS);
  if (a->f_in == a->f_out) return fail(XB_EINVAL, "xb_elast_stepn: f_in and f_out must differ");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (a->dtype == XB_F64) {
    ElastColl<double> c;
    NLDev<double> nl;
    elast_coll<double>(a, c, nl);
    return launch_tile_steps<double>(a->nx, a->ny, steps, c, a->f_in, a->f_out, a->bc, a->rows, st);
  }
  ElastColl<float> c;
  NLDev<float> nl;
  elast_coll<float>(a, c, nl);
  return launch_tile_steps<float>(a->nx, a->ny, steps, c, a->f_in, a->f_out, a->bc, a->rows, st);
}

#endif
#if XB_IN_PART(6)
int xb_elast_smoothn(const xb_elast_args *a, int32_t steps, void *stream) {
  if (int r = check_smoothn(a, steps, "xb_elast_smoothn")) return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (a->dtype == XB_F64) return elast_smoothn_t<double>(a, steps, st);
  return elast_smoothn_t<float>(a, steps, st);
}

#endif
} // extern "C"

#if XB_IN_PART(1)
// ---------------------------------------------------------------------------
// Allocation-free residual sweep.  The squared-norm partials of the residual sweep
// (one per wavefront item) live in this file-scope device array instead of a
// caller workspace: 8 B x XB_STATIC_PARTIALS = 1 MiB, held by every GPU the
// library runs on until the process ends.  Every call on one GPU shares that
// GPU's copy; the sweep writes each partial the finalisation reads, so nothing of
// it has to start at zero, and calls that overlap on one GPU (different streams,
// or host threads) race on it.
static __device__ __align__(256) double scratch[XB_STATIC_PARTIALS];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// Device address of the current device's copy of `scratch`, looked up once per
// device (with that device current).  In host code the array's name is a host-side
// stand-in, never a device address: kernels and copies use this pointer.
int scratch_address(double **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCRATCH_MAX_DEVICES)
    return fail(XB_EHIP, "xb_elast_residual_noalloc: no usable current device (%d)", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(XB_EHIP, "xb_elast_residual_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return XB_OK;
}

} // namespace

extern "C" {

int xb_elast_residual_noalloc(const xb_elast_args *a, double *norm_out, void *stream) {
  if (a == nullptr || norm_out == nullptr) return fail(XB_EINVAL, "xb_elast_residual_noalloc: null argument");
  if (a->mode != XB_MODE_RESIDUAL) return fail(XB_EINVAL, "xb_elast_residual_noalloc: residual mode only");
  if (a->norm_partials != nullptr)
    return fail(XB_EINVAL, "xb_elast_residual_noalloc: norm_partials must be null (the partials live in the "
                           "library's static scratch)");
  double *partials = nullptr;
  if (int r = guarded("xb_elast_residual_noalloc", [&] { return scratch_address(&partials); })) return r;
  xb_elast_args b = *a;
  b.norm_partials = partials;
  if (int r = validate_elast(&b, "xb_elast_residual_noalloc")) return r;
  SweepGeom g;
  elast_geom(&b, g);
  if ((int64_t)g.nitems > (int64_t)XB_STATIC_PARTIALS)
    return fail(XB_EINVAL, "xb_elast_residual_noalloc: the sweep has %d wavefront items, the static scratch holds "
                "%d partials (more rows per wavefront, or xb_elast_step with a workspace)", g.nitems,
                (int)XB_STATIC_PARTIALS);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (int r = b.dtype == XB_F64 ? elast_step_t<double>(&b, st) : elast_step_t<float>(&b, st)) return r;
  hipLaunchKernelGGL(k_norm_finalize, dim3(1), dim3(NORM_THREADS), 0, st, partials, g.nitems, norm_out);
  return check_launch("xb_elast_residual_noalloc");
}

} // extern "C"
#endif // XB_IN_PART(1)
