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
// logical blocks (neighbouring strips and row blocks, which share halo rows)
// land on the same XCD and therefore the same L2.
__device__ __forceinline__ int xcd_block_id() {
  const int nb = gridDim.x;
  const int b = blockIdx.x;
  if ((nb & 7) != 0) return b;
  return (b & 7) * (nb >> 3) + (b >> 3);
}

template <typename T> struct BCDev {
  const int32_t *slot = nullptr;
  const int8_t *code_pc = nullptr;
  const int8_t *code_ps = nullptr;
  const T *const_pc = nullptr;
  const T *const_ps = nullptr;
};

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
template <typename T> struct ElastColl {
  static constexpr int Q = 8;
  __host__ __device__ static constexpr int cx(int d) { return d == 0 || d == 4 || d == 7 ? 1 : (d == 2 || d == 5 || d == 6 ? -1 : 0); }
  __host__ __device__ static constexpr int cy(int d) { return d == 1 || d == 4 || d == 5 ? 1 : (d == 3 || d == 6 || d == 7 ? -1 : 0); }
  T cs, cd, h, n;       // relaxation of m_s, of m_d and m_11; m_to_f entries h, n
  const T *fa = nullptr; // body force (2, N)
  const T *fb = nullptr; // external force (2, N)
  size_t N = 0;

  __device__ __forceinline__ void collide(const T (&f)[8], T (&o)[8], size_t cell) const {
    const T m10 = f[0] - f[2] + f[4] - f[5] - f[6] + f[7];
    const T m01 = f[1] - f[3] + f[4] + f[5] - f[6] - f[7];
    const T ms = f[0] + f[1] + f[2] + f[3] + T(2) * (f[4] + f[5] + f[6] + f[7]);
    const T md = f[0] - f[1] + f[2] - f[3];
    const T m11 = f[4] - f[5] + f[6] - f[7];
    T M0 = m10, M1 = m01, M5 = m10 / T(3), M6 = m01 / T(3);
    if (fa != nullptr || fb != nullptr) {
      T F0, F1;
      if (fa != nullptr && fb != nullptr) {
        F0 = fa[cell] + fb[cell];
        F1 = fa[N + cell] + fb[N + cell];
      } else {
        const T *F = fa != nullptr ? fa : fb;
        F0 = F[cell];
        F1 = F[N + cell];
      }
      M0 = m10 + F0;
      M1 = m01 + F1;
      M5 = m10 / T(3) + F0 / T(6);
      M6 = m01 / T(3) + F1 / T(6);
    }
    const T M2 = ms * cs, M3 = md * cd, M4 = m11 * cd;
    const T hM2 = h * M2, nM2 = n * M2;
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

  __device__ __forceinline__ void collide(const T (&f)[9], T (&o)[9], size_t cell) const {
    T rho = f[0];
#pragma unroll
    for (int d = 1; d < 9; ++d) rho = rho + f[d];
    const bool has_s = sa != nullptr || sb != nullptr;
    T S = T(0);
    if (has_s) {
      S = (sa != nullptr && sb != nullptr) ? sa[cell] + sb[cell] : (sa != nullptr ? sa[cell] : sb[cell]);
      rho = rho + T(0.5) * S;
    }
    const T om = om_field != nullptr ? om_field[cell] : omega;
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

} // namespace xb
