// Device-side building blocks of the gfx950 lattice-Boltzmann kernels:
// wavefront geometry, lane shifts, boundary-rule evaluation and the per-cell
// collision operators of the D2Q8 elasticity and D2Q9 Darcy lattices.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "xlb_biot.h"

namespace xb {

constexpr int WAVE = 64;             // CDNA wavefront
constexpr int STRIP = WAVE - 2;      // output cells per wavefront along y (2 halo lanes)
constexpr int WAVES_PER_BLOCK = 4;
constexpr int BLOCK = WAVE * WAVES_PER_BLOCK;

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Value held by lane-1 / lane+1 of the wavefront (edge lanes get their own value,
// which the sweep never consumes).
template <typename T> __device__ __forceinline__ T from_prev_lane(T v) { return __shfl_up(v, 1u, WAVE); }
template <typename T> __device__ __forceinline__ T from_next_lane(T v) { return __shfl_down(v, 1u, WAVE); }

// Blocks are dealt round-robin to the 8 XCDs; renumber them so that consecutive
// logical blocks (neighbouring strips, row blocks and tiles, which share halo cells)
// land on the same XCD and therefore the same L2.  XCD x (blocks b = x mod 8) takes
// the run of logical blocks [x q + min(x, r), ...) of length q + (x < r), where
// nb = 8 q + r: a bijection of [0, nb) for every grid size.
__device__ __forceinline__ int xcd_block_id() {
  const int nb = gridDim.x;
  const int b = blockIdx.x;
  const int x = b & 7, q = nb >> 3, r = nb & 7;
  return x * q + min(x, r) + (b >> 3);
}

template <typename T> struct BCDev {
  const int32_t *slot = nullptr;
  const int8_t *code_pc = nullptr;
  const int8_t *code_ps = nullptr;
  const T *const_pc = nullptr;
  const T *const_ps = nullptr;
  int nslots = 0;
};

// Rule tables of at most LDS_BC_SLOTS slots are copied into LDS by each wavefront
// of a sweep that meets a rule cell (into its own region, the first time one of
// its rows has one): the rule reads then go through the LDS counter and never
// force a wait on the populations still in flight (a vector-memory wait counts
// loads in issue order, so waiting for a rule load would also wait for every row
// prefetched after it), and wavefronts without rule cells do no extra work.  Sweep
// kernels take the rule mode as a template argument:
//   0 no table, 1 table read from global memory, 2 table staged in LDS.
constexpr int LDS_BC_SLOTS = 32;

template <typename T, int Q> struct BCShared {
  int8_t code_pc[LDS_BC_SLOTS * Q];
  int8_t code_ps[LDS_BC_SLOTS * Q];
  T const_pc[LDS_BC_SLOTS * Q];
  T const_ps[LDS_BC_SLOTS * Q];
};

// Copy of a small rule table into one wavefront's LDS region, by all lanes of that
// wavefront (no barrier: LDS operations of one wavefront complete in order).
template <typename T, int Q>
__device__ __forceinline__ void stage_rules_wave(const BCDev<T> &bc, BCShared<T, Q> &sh) {
  const int n = bc.nslots * Q;
  for (int k = (int)(threadIdx.x & (WAVE - 1)); k < n; k += WAVE) {
    sh.code_pc[k] = bc.code_pc != nullptr ? bc.code_pc[k] : int8_t(0);
    sh.const_pc[k] = bc.const_pc != nullptr ? bc.const_pc[k] : T(0);
    sh.code_ps[k] = bc.code_ps != nullptr ? bc.code_ps[k] : int8_t(0);
    sh.const_ps[k] = bc.const_ps != nullptr ? bc.const_ps[k] : T(0);
  }
}

inline int bc_mode(const int32_t *slot, int nslots) { return slot == nullptr ? 0 : (nslots <= LDS_BC_SLOTS ? 2 : 1); }

// v[d] <- rule(v[d], src) + const[d]  (rule codes of xlb_biot.h).  Selection by
// compare chains keeps src in registers (no dynamic register indexing).
template <typename T, int Q>
__device__ __forceinline__ void apply_rules(T (&v)[Q], const T (&src)[Q], const int8_t *code, const T *cst) {
  T out[Q];
#pragma unroll
  for (int d = 0; d < Q; ++d) {
    const int c = code[d];
    T base = v[d];
    if (c == XB_RULE_ZERO) {
      base = T(0);
    } else if (c != 0) {
      const int k = (c > 0 ? c : -c) - 1;
      T s = src[0];
#pragma unroll
      for (int j = 1; j < Q; ++j) s = (k == j) ? src[j] : s;
      base = c > 0 ? s : -s;
    }
    out[d] = base + cst[d];
  }
#pragma unroll
  for (int d = 0; d < Q; ++d) v[d] = out[d];
}

// ---------------------------------------------------------------------------
// D2Q8 elasticity: moment relaxation (reference src/elasticity.py:959-1000).
// Directions c = (1,0) (0,1) (-1,0) (0,-1) (1,1) (-1,1) (-1,-1) (1,-1).

// Relaxation rates and m_to_f entries from the Lame parameters of one cell
// (elasticity.py:805-848; g is set to 0 where |g| > 100, :811-813).
template <typename T> __device__ __forceinline__ void elast_coef(T lam, T mu, T &cs, T &cd, T &h, T &n) {
  const T lm = lam + mu;
  T g = T(1) / (T(12) * lm - T(4));
  if (fabs(g) > T(100)) g = T(0);
  cs = (T(3) * lm - T(1)) / (T(3) * lm + T(1));
  cd = (T(6) * mu - T(1)) / (T(6) * mu + T(1));
  h = T(0.25) + T(0.5) * g;
  n = T(-0.25) * g;
}

// a / b for a constant b, correctly rounded without a division instruction sequence:
// q = RN(a RN(1/b)) is within one ulp of a / b, the fma gives the exact remainder
// r = a - b q, and RN(q + r RN(1/b)) is then RN(a / b) (Markstein's theorem; finite a,
// no underflow).  Bit-identical to the division for b = 3 and 6 in FP64 and FP32
// (tests/test_division.py checks it on 10^8 samples); about a third of the FP64
// instructions of the division.
template <typename T> __device__ __forceinline__ T div_const(T a, T b, T inv_b) {
  const T q = a * inv_b;
  const T r = fma(-q, b, a);
  return fma(r, inv_b, q);
}

template <typename T> struct ElastColl {
  static constexpr int Q = 8;
  __host__ __device__ static constexpr int cx(int d) { return d == 0 || d == 4 || d == 7 ? 1 : (d == 2 || d == 5 || d == 6 ? -1 : 0); }
  __host__ __device__ static constexpr int cy(int d) { return d == 1 || d == 4 || d == 5 ? 1 : (d == 3 || d == 6 || d == 7 ? -1 : 0); }
  T cs, cd, h, n;       // relaxation of m_s, of m_d and m_11; m_to_f entries h, n
  const T *fa = nullptr; // body force (2, N)
  const T *fb = nullptr; // external force (2, N)
  const T *lamf = nullptr; // per-cell Lame parameters (N each; both or neither):
  const T *muf = nullptr;  // then cs, cd, h, n are formed per cell
  size_t N = 0;
  bool has_b = false;    // an external force takes part (given, or formed in the sweep)

  // Per-cell operands besides the populations, loaded ahead of the collision so that
  // the sweep can keep them in flight with the populations of the same row.
  struct Aux {
    T a0, a1, b0, b1, lam, mu;
  };
  __device__ __forceinline__ Aux load_aux(size_t cell) const {
    Aux x{T(0), T(0), T(0), T(0), T(0), T(0)};
    if (fa != nullptr) { x.a0 = fa[cell]; x.a1 = fa[N + cell]; }
    if (fb != nullptr) { x.b0 = fb[cell]; x.b1 = fb[N + cell]; }
    if (lamf != nullptr) { x.lam = lamf[cell]; x.mu = muf[cell]; }
    return x;
  }

  // divergence coefficient of the cell: per cell with Lame fields, else k0
  __device__ __forceinline__ T kdiv(const Aux &x, T k0) const {
    return lamf != nullptr ? T(-1.5) / (T(3) * (x.lam + x.mu) + T(1)) : k0;
  }

  __device__ __forceinline__ void collide(const T (&f)[8], const Aux &x, T (&o)[8]) const {
    T cs_ = cs, cd_ = cd, h_ = h, n_ = n;
    if (lamf != nullptr) elast_coef<T>(x.lam, x.mu, cs_, cd_, h_, n_);
    const T m10 = f[0] - f[2] + f[4] - f[5] - f[6] + f[7];
    const T m01 = f[1] - f[3] + f[4] + f[5] - f[6] - f[7];
    const T ms = f[0] + f[1] + f[2] + f[3] + T(2) * (f[4] + f[5] + f[6] + f[7]);
    const T md = f[0] - f[1] + f[2] - f[3];
    const T m11 = f[4] - f[5] + f[6] - f[7];
    constexpr T third = T(1) / T(3), sixth = T(1) / T(6);
    const T m10_3 = div_const(m10, T(3), third), m01_3 = div_const(m01, T(3), third);
    T M0 = m10, M1 = m01, M5 = m10_3, M6 = m01_3;
    if (fa != nullptr || has_b) {
      // reference order: body force + external force
      const T F0 = (fa != nullptr && has_b) ? x.a0 + x.b0 : (fa != nullptr ? x.a0 : x.b0);
      const T F1 = (fa != nullptr && has_b) ? x.a1 + x.b1 : (fa != nullptr ? x.a1 : x.b1);
      M0 = m10 + F0;
      M1 = m01 + F1;
      M5 = m10_3 + div_const(F0, T(6), sixth);
      M6 = m01_3 + div_const(F1, T(6), sixth);
    }
    const T M2 = ms * cs_, M3 = md * cd_, M4 = m11 * cd_;
    const T hM2 = h_ * M2, nM2 = n_ * M2;
    o[0] = T(0.5) * M0 + hM2 + T(0.25) * M3 - T(0.5) * M5;
    o[1] = T(0.5) * M1 + hM2 - T(0.25) * M3 - T(0.5) * M6;
    o[2] = -T(0.5) * M0 + hM2 + T(0.25) * M3 + T(0.5) * M5;
    o[3] = -T(0.5) * M1 + hM2 - T(0.25) * M3 + T(0.5) * M6;
    o[4] = nM2 + T(0.25) * M4 + T(0.25) * M5 + T(0.25) * M6;
    o[5] = nM2 - T(0.25) * M4 - T(0.25) * M5 + T(0.25) * M6;
    o[6] = nM2 + T(0.25) * M4 - T(0.25) * M5 - T(0.25) * M6;
    o[7] = nM2 - T(0.25) * M4 + T(0.25) * M5 - T(0.25) * M6;
  }
};

// ---------------------------------------------------------------------------
// D2Q9 BGK advection-diffusion-reaction at zero velocity (reference
// src/darcy.py:106-129).  Directions in the reference order
// cx = 0 0 0 1 -1 1 -1 1 -1, cy = 0 1 -1 0 1 -1 0 1 -1.
template <typename T> struct DarcyColl {
  static constexpr int Q = 9;
  __host__ __device__ static constexpr int cx(int d) { return d == 3 || d == 5 || d == 7 ? 1 : (d == 4 || d == 6 || d == 8 ? -1 : 0); }
  __host__ __device__ static constexpr int cy(int d) { return d == 1 || d == 4 || d == 7 ? 1 : (d == 2 || d == 5 || d == 8 ? -1 : 0); }
  __host__ __device__ static constexpr double wd(int d) {
    return d == 0 ? 4.0 / 9.0 : ((cx(d) != 0 && cy(d) != 0) ? 1.0 / 36.0 : 1.0 / 9.0);
  }
  T omega;
  T sfac;                      // (1 - omega/2) of the scalar rate, formed in double on the host
  const T *om_field = nullptr; // (N) per-cell relaxation rate
  const T *sa = nullptr;       // static source (N)
  const T *sb = nullptr;       // external source (N)

  struct Aux {
    T sa, sb, om;
  };
  __device__ __forceinline__ Aux load_aux(size_t cell) const {
    Aux x{T(0), T(0), omega};
    if (sa != nullptr) x.sa = sa[cell];
    if (sb != nullptr) x.sb = sb[cell];
    if (om_field != nullptr) x.om = om_field[cell];
    return x;
  }

  __device__ __forceinline__ void collide(const T (&f)[9], const Aux &x, T (&o)[9]) const {
    T rho = f[0];
#pragma unroll
    for (int d = 1; d < 9; ++d) rho = rho + f[d];
    const bool has_s = sa != nullptr || sb != nullptr;
    T S = T(0);
    if (has_s) {
      S = (sa != nullptr && sb != nullptr) ? x.sa + x.sb : (sa != nullptr ? x.sa : x.sb);
      rho = rho + T(0.5) * S;
    }
    const T om = x.om;
    const T sf = om_field != nullptr ? T(1) - T(0.5) * om : sfac;
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      const T w = T(wd(d));
      const T fneq = f[d] - rho * w;
      T v = f[d] - om * fneq;
      if (has_s) v = v + S * (sf * w);
      o[d] = v;
    }
  }
};

// ---------------------------------------------------------------------------
// Shared by the stand-alone kernels and the fused sweep so both paths do the same
// arithmetic in the same order.

// divergence coefficient k = -1.5 / (3 (lam + mu) + 1) of one cell with per-cell
// Lame parameters (reference elasticity.py:1040-1058)
template <typename T> __device__ __forceinline__ T elast_kdiv(T lam, T mu) {
  return T(-1.5) / (T(3) * (lam + mu) + T(1));
}

// divergence of D2Q8 populations: k * m_s (reference elasticity.py:1040-1058)
template <typename T> __device__ __forceinline__ T elast_div_of(const T (&v)[8], T k) {
  return k * (v[0] + v[1] + v[2] + v[3] + T(2) * (v[4] + v[5] + v[6] + v[7]));
}

// lattice gradient sum_d qs_d ((-3 scale) w_d) c_d over the D2Q9 stencil
// (reference darcy.py:75-104, the final matmul)
template <typename T> __device__ __forceinline__ void stencil_grad(const T (&qs)[9], T m3s, T &g0, T &g1) {
  using C = DarcyColl<T>;
  g0 = T(0);
  g1 = T(0);
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    const T cw = m3s * T(C::wd(d));
    if (C::cx(d) != 0) g0 = g0 + qs[d] * (C::cx(d) > 0 ? cw : -cw);
    if (C::cy(d) != 0) g1 = g1 + qs[d] * (C::cy(d) > 0 ? cw : -cw);
  }
}

// Fused Biot coupling force of the elasticity sweep
template <typename T> struct NLDev {
  const T *div_in = nullptr;
  T *div_out = nullptr;
  T m3s = T(0);   // -3 * scale
  T kdiv = T(0);  // -1.5 / (3 (lam + mu) + 1)
  BCDev<T> gbc;   // Darcy wall rules of the gradient
  const T *base = nullptr;
};

} // namespace xb
