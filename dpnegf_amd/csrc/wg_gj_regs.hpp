// Register-resident blocked Gauss-Jordan elimination (FP64, gfx950 wave64).
//
// The augmented matrix (N rows, N <= 64 a multiple of 16) is split into
// 16-column tiles; each tile is owned by one wave and held in registers in the
// accumulator layout of the 16x16x4 f64 matrix-core instruction: lane
// (g = lane/16, c = lane%16) holds column c of the tile at rows
// 16 I + g + 4 r in element e = 4 I + r of a vector of N/4 doubles (re and im
// separately).  Only the current panel (4 columns), the rank-4 multipliers and
// a few scalars pass through LDS, so a panel costs one workgroup barrier and
// the trailing update runs entirely on register operands:
//
//   sweeper (owner of the panel's tile)  : copies the 4 panel columns to LDS,
//       sweeps them with one row per lane (the sequential pivot steps), and
//       publishes -Fm (matrix-core A operand), the pivot rows pr[k], the
//       4 x 4 lower factor L and the pivot reciprocals; barrier;
//   every owner, for its own columns     : gathers the 4 pivot-row values of
//       each column (uniform register index + lane permute), forward-
//       substitutes u_k = (z_k - sum_{k'<k} L[k][k'] u_k') / d_k, writes u_k
//       over pivot row pr[k] and adds (-Fm) u with four matrix-core steps.
//
// This is the same arithmetic as the LDS-resident blocked form (pivot choice,
// multipliers and the blocked-LU style update of the unreduced pivot rows);
// only the storage differs.
#pragma once
#include "wg_linalg.hpp"

namespace negf {

template <int N>
struct RegGJ {
  static_assert(N % 16 == 0 && N <= WAVE, "register Gauss-Jordan: N multiple of 16, <= 64");
  static constexpr int NT16 = N / 16;  // row tiles
  static constexpr int E = N / 4;      // elements per lane and tile column
  typedef double vE __attribute__((ext_vector_type(E)));
  typedef double v4 __attribute__((ext_vector_type(4)));

  // What one panel publishes.
  struct Slot {
    cx<double> FM[NT16][WAVE];  // -Fm, A-operand fragment order
    cx<double> L[4][4];
    cx<double> dinv[4];
    int pr[4];
  };
  struct Shared {
    Slot slot[2];                  // by panel parity
    cx<double> PAN[N][4];          // the panel's columns, one row per entry
    int piv[N];                    // piv[j]: row that pivots column j
    int pinv[N];                   // pinv[piv[j]] = j
    int singular;
  };

  // Row held by element e of a lane in row group g.
  static __device__ __forceinline__ int row_of(int e, int g) {
    return 16 * (e >> 2) + g + 4 * (e & 3);
  }

  static __device__ __forceinline__ void cmfma(vE& xr, vE& xi, int I, cx<double> a,
                                               cx<double> b) {
    v4 cr = {xr[4 * I], xr[4 * I + 1], xr[4 * I + 2], xr[4 * I + 3]};
    v4 ci = {xi[4 * I], xi[4 * I + 1], xi[4 * I + 2], xi[4 * I + 3]};
    cr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.re, cr, 0, 0, 0);
    cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.im, b.im, cr, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.im, ci, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.re, ci, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      xr[4 * I + r] = cr[r];
      xi[4 * I + r] = ci[r];
    }
  }

  // Sweeper, step 1: the panel's columns c0..c0+3 of the calling wave's tile
  // to PAN (row-major), then a wave-local fence.
  static __device__ __forceinline__ void put_panel(const vE& xr, const vE& xi, int c0,
                                                   Shared* sh) {
    const int lane = threadIdx.x & (WAVE - 1), g = lane >> 4, c = lane & 15;
    if (c >= c0 && c < c0 + 4) {
#pragma unroll
      for (int e = 0; e < E; ++e) sh->PAN[row_of(e, g)][c - c0] = cx<double>{xr[e], xi[e]};
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }

  // Sweeper, step 2: the four pivot steps of panel j0 on PAN, one row per
  // lane.  `used`: this lane's row has pivoted in an earlier panel.
  static __device__ __forceinline__ void sweep(int j0, Slot* sl, bool used, Shared* sh) {
    const int lane = threadIdx.x & (WAVE - 1);
    cx<double> v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = lane < N ? sh->PAN[lane][c] : cx<double>{0, 0};
    sweep_row<4>(v, j0, sl, used, sh);
  }

  // The four pivot steps of panel j0 on a row held by this lane (lane = row):
  // v[0..3] are the panel's columns, v[4..NC-1] further columns that receive
  // the same row operations.  Pivot = largest |re| + |im| among unused rows,
  // compared in single precision to 17 mantissa bits, lowest row on ties (a
  // DPP wave maximum).
  //
  // Each step eliminates with the standard multiplier, row -= (f / d) prow,
  // and leaves the pivot row unscaled; the pivot rows are scaled by 1 / d
  // once at the end of the panel (an unscaled row r satisfies
  // r = d * (the scaled row) through every later step, so the result is the
  // Gauss-Jordan one).  Per column and step that is four FMAs.
  //
  // Publishes -Fm (Fm: the multipliers of the scaled rows, zero for a pivot
  // row's own and earlier steps), the pivot rows, L (the pivot rows' values at
  // earlier steps, never scaled: they were not pivots yet), the reciprocals,
  // piv / pinv and the singular flag.
  template <int NC>
  static __device__ __forceinline__ void sweep_row(cx<double> (&v)[NC], int j0, Slot* sl,
                                                   bool used, Shared* sh) {
    sweep_row<NC>([&](int c) { return v[c]; }, [&](int c, cx<double> x) { v[c] = x; }, j0, sl,
                  used, sh);
  }
  // Same, with column c of the row read and written through get(c) /
  // set(c, x) (c a compile-time constant after unrolling), so that a row held
  // in the tile's own registers is swept in place.
  template <int NC, class GET, class SET>
  static __device__ __forceinline__ void sweep_row(GET get, SET set, int j0, Slot* sl, bool used,
                                                   Shared* sh) {
    const int lane = threadIdx.x & (WAVE - 1);
    const bool row_ok = lane < N;
    cx<double> F[4], dinv[4];
    int pr[4];
    int pstep = 4;
    bool sing = false;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      unsigned key = 0;
      if (row_ok && !used) {
        const float kf = (float)pivot_key(get(jj));
        key = (((__float_as_uint(kf) >> 6) + 1u) << 6) | (unsigned)(WAVE - 1 - lane);
      }
      key = wave_max_u32(key);
      int p = WAVE - 1 - (int)(key & 63u);
      p = p < N ? p : N - 1;
      pr[jj] = p;
      const bool isp = (lane == p);
      if (isp) {
        used = true;
        pstep = jj;
#pragma unroll
        for (int k2 = 0; k2 < jj; ++k2) sl->L[jj][k2] = F[k2];
      }
      F[jj] = get(jj);
      const cx<double> d = bcast_lane(F[jj], p);
      dinv[jj] = recip_nr(d);
      sing |= (d.re == 0.0 && d.im == 0.0);
      // multiplier f / d; the pivot row is left as it is
      const cx<double> m0 = F[jj] * dinv[jj];
      const cx<double> mj = {isp ? 0.0 : m0.re, isp ? 0.0 : m0.im};
#pragma unroll
      for (int c = jj + 1; c < NC; ++c) {
        const cx<double> vc = get(c);
        const cx<double> pc = bcast_lane(vc, p);
        set(c, cx<double>{fma(-mj.re, pc.re, fma(mj.im, pc.im, vc.re)),
                          fma(-mj.re, pc.im, fma(-mj.im, pc.re, vc.im))});
      }
    }
    // scale the pivot rows: by 1 / d of their step, for the further columns
    // and for the multipliers of later steps
    cx<double> s = {1.0, 0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (pstep == k) s = dinv[k];
#pragma unroll
    for (int c = 4; c < NC; ++c) set(c, get(c) * s);
    if (row_ok) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const cx<double> fm = (pstep < 4 && k <= pstep) ? cx<double>{0, 0} : F[k] * s;
        sl->FM[lane >> 4][(lane & 15) + 16 * k] = -fm;
      }
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        sl->pr[k] = pr[k];
        sl->dinv[k] = dinv[k];
        sh->piv[j0 + k] = pr[k];
        sh->pinv[pr[k]] = j0 + k;
      }
      if (sing) sh->singular = 1;
    }
  }

  // The calling wave's tile (16 columns) to one row per lane through the
  // 17 N-entry LDS buffer tb (rows padded to 17 elements: with 16, every lane
  // of the row reads hit the same four banks and took 2.1-2.6k cycles a
  // tile): v[0..3] = columns 0..3 of row `lane`, the other twelve columns
  // into xr / xi (E >= 12), as (xr[m], xi[m]) = column 4 + m.  Lanes >= N
  // get zeros (they read row N - 1 and drop it: no load under a lane mask).
  static constexpr int TB_LD = 17;
  static __device__ __forceinline__ void tile_to_rows(vE& xr, vE& xi, cx<double> (&v)[4],
                                                      cx<double>* tb) {
    static_assert(E >= 12, "row-per-lane tile needs twelve register pairs");
    const int lane = threadIdx.x & (WAVE - 1), g = lane >> 4, c = lane & 15;
#pragma unroll
    for (int e = 0; e < E; ++e) tb[row_of(e, g) * TB_LD + c] = cx<double>{xr[e], xi[e]};
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bool ok = lane < N;
    const cx<double>* rowp = tb + (ok ? lane : N - 1) * TB_LD;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const cx<double> t = rowp[k];
      v[k] = cx<double>{ok ? t.re : 0.0, ok ? t.im : 0.0};
    }
#pragma unroll
    for (int m = 0; m < 12; ++m) {
      const cx<double> t = rowp[4 + m];
      xr[m] = ok ? t.re : 0.0;
      xi[m] = ok ? t.im : 0.0;
    }
  }

  // After the panel's barrier, every wave: marks the panel's pivot rows as
  // used (lane = row).
  static __device__ __forceinline__ bool mark_used(bool used, const Slot* sl) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int k = 0; k < 4; ++k) used |= (lane == sl->pr[k]);
    return used;
  }

  // After the panel's barrier, an owner: elimination step on its tile.
  static __device__ __forceinline__ void update(vE& xr, vE& xi, const Slot* sl) {
    const int lane = threadIdx.x & (WAVE - 1), g = lane >> 4, c = lane & 15;
    int pe[4], pg[4];
    cx<double> z[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int p = __builtin_amdgcn_readfirstlane(sl->pr[k]);
      pe[k] = 4 * (p >> 4) + ((p & 15) >> 2);
      pg[k] = p & 3;
      const double zr = xr[pe[k]], zi = xi[pe[k]];
      const int src = 16 * pg[k] + c;
      z[k] = cx<double>{__shfl(zr, src, WAVE), __shfl(zi, src, WAVE)};
    }
    cx<double> u[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      cx<double> acc = z[k];
#pragma unroll
      for (int k2 = 0; k2 < k; ++k2) acc = acc - sl->L[k][k2] * u[k2];
      u[k] = acc * sl->dinv[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const bool own = (g == pg[k]);
      const double cr = xr[pe[k]], ci = xi[pe[k]];
      xr[pe[k]] = own ? u[k].re : cr;
      xi[pe[k]] = own ? u[k].im : ci;
    }
    // B operand: lane group g supplies step g.  Scalar selects: a conditional
    // between two array elements selects their addresses and leaves u[] in
    // scratch memory.
    const bool g1 = g & 1, g2 = g & 2;
    const double b01r = g1 ? u[1].re : u[0].re, b01i = g1 ? u[1].im : u[0].im;
    const double b23r = g1 ? u[3].re : u[2].re, b23i = g1 ? u[3].im : u[2].im;
    const cx<double> b = {g2 ? b23r : b01r, g2 ? b23i : b01i};
#pragma unroll
    for (int I = 0; I < NT16; ++I) cmfma(xr, xi, I, sl->FM[I][lane], b);
  }
};

}  // namespace negf
