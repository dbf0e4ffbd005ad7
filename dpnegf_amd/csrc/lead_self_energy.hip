// Batched Lopez-Sancho decimation + lead self-energy, FP64, one workgroup per
// (E,k) point.
//
// Replaces _surface_green_scipy_core / _surface_green_numba_core
// (dpnegf/negf/surface_green.py:102-159, 169-226) and the Sigma assembly of
// selfEnergy (surface_green.py:72-86), which the reference runs per (E,k) in a
// joblib process pool (dpnegf/negf/lead_property.py:599-683).
//
// Per iteration, as the reference (:189-204):
//   [X_a | X_b] = W \ [alpha | beta],  W = zS - eps   (solve, not W^-1 times)
//   eps   += alpha X_b + beta X_a
//   eps_s += beta X_a
//   alpha <- alpha X_a,  beta <- beta X_b
// until max_ij |alpha_ij| + |beta_ij| < stop (1e-40 in the reference, :208).
// Then g_s = (zS - eps_s)^-1, the Dyson residual of :211-220, and
//   Sigma = (e S_DL - H_DL) g_s (e S_DL^H - H_DL^H)      (:86, real-axis e)
// or the no-coupling forms of :72-80.
//
// The solve is Gauss-Jordan elimination on [W | alpha | beta] (forward stable
// like the reference's LU solve; forming W^-1 and multiplying instead loses
// the Dyson residual near band edges at small eta).  Matrices are padded to
// NMAX with an identity block in W and zero rows/columns in alpha and beta,
// which leaves the n x n results exactly as without padding.
//
// Layout: 3 NMAX/16 waves; wave w owns 16-column tile w of [W | alpha | beta]
// in registers (the register-resident elimination of wg_gj_regs.hpp), so the
// solve moves only panels through LDS (at NMAX = 48 each panel is published
// with a flag instead of a workgroup barrier; see the solve).  After it the alpha / beta owners hold
// X = [X_a | X_b] (rows in pivot order), which goes to LDS in matrix-core
// B-operand fragment order; alpha and beta stay in LDS as A-operand fragments
// from the previous iteration.  Then, with no data leaving the owners:
//   alpha tile j owner:  alpha X_a (column tile j) -> its registers (next RHS)
//   beta  tile j owner:  eps_s += beta X_a (stored), then beta X_b -> its
//                        registers
//   W     tile j owner:  D += alpha X_b (stored), with D = eps - eps_s; then
//                        W = zS - eps_s - D for the next solve
// so eps = eps_s + D receives beta X_a + alpha X_b.  D lives in the workspace
// and eps_s in the g_s output.
#include <type_traits>

#include "rgf_kernels.hpp"
#include "wg_gj_regs.hpp"

namespace negf {

// The small, hot structures come first: LDS instructions address up to 64 KB
// through their immediate offset, beyond that each address needs a register.
template <int N>
struct LsLoopShared {
  static constexpr int NT16 = N / 16, KS = N / 4;
  int pflag[KS];  // panel p published in iteration pflag[p] & 0xffff; bit 16 on panel 0: stop
  int nvote;      // stop-test votes written so far (NW per iteration)
  int abort;      // a wait ran past its limit: the point is given up
  typename RegGJ<N>::Shared gj;
  cx<double> Aa[NT16 * KS][WAVE];      // alpha, A-operand fragments (row tile, k-step)
  cx<double> Ab[NT16 * KS][WAVE];      // beta
  cx<double> XS[KS * 2 * NT16][WAVE];  // [X_a | X_b], B-operand fragments (k-step, col tile)
};
template <int N>
struct LsPostShared {
  cx<double> AUG[N * (3 * N + 1)];
  cx<double> SP[N * (N + 1)];
  GJBScratch<double, N> gjb;
};
template <int N>
struct LsShared {
  double red[16];
  union {
    LsLoopShared<N> loop;
    LsPostShared<N> post;
  };
};

template <int NMAX> constexpr int ls_threads() { return 3 * (NMAX / 16) * WAVE; }
// Polls (each followed by s_sleep 1, about 64 clocks) before a wait for a
// published panel or for the stop votes gives up: about 17M clocks, some
// thousand times the longest wait of a working sweep.
constexpr int SPIN_LIMIT = 1 << 18;

template <int NMAX, bool ORTHO>
__global__ __launch_bounds__(ls_threads<NMAX>()) void lead_self_energy_kernel(LeadArgs a) {
  typedef double R;
  typedef RegGJ<NMAX> GJ;
  typedef typename GJ::vE vE;
  constexpr int N = NMAX;
  constexpr int NTH = ls_threads<NMAX>();
  constexpr int NW = NTH / WAVE;
  constexpr int NT16 = N / 16;
  constexpr int KS = N / 4;
  constexpr int E = GJ::E;
  constexpr int LDG = 3 * N + 1;
  constexpr int LDP = N + 1;
  constexpr int NTILE = (N / 16) * (N / 16);
  constexpr int MAXT1 = (NTILE + NW - 1) / NW;
  // tile_to_rows scratch (17 N elements) in the X fragments' space, after the
  // KS panel slots, 256-byte aligned
  constexpr int T2R_OFF =
      ((KS * (int)sizeof(typename GJ::Slot) + 255) / 256) * 256 / (int)sizeof(cx<R>);
  static_assert(NT16 != 3 || T2R_OFF + GJ::TB_LD * N <= KS * 2 * NT16 * WAVE,
                "panel slots + scratch in XS");
  static_assert(NW <= 16, "reduction slots");
  __shared__ LsShared<N> sh;
  double* const red = sh.red;
  cx<R>* AUG = sh.post.AUG;
  cx<R>* SP = sh.post.SP;
  GJBScratch<R, N>& gjb = sh.post.gjb;
  auto& L = sh.loop;

  const int pt = blockIdx.x;
  if (pt >= a.P) return;
  const int n = a.n, nr = a.nred;
  const int sys = a.sys_idx ? a.sys_idx[pt] : 0;
  if (sys < 0 || sys >= a.nsys) {  // uniform per workgroup, before any barrier
    if (threadIdx.x == 0) {
      if (a.iters) a.iters[pt] = 0;
      if (a.status) a.status[pt] = STATUS_BAD_INDEX;
    }
    return;
  }
  const cx<R>* H = a.lead + (int64_t)sys * a.lead_stride;
  const cx<R>* h01 = H + n * n;
  const cx<R>* S = h01 + n * n;
  const cx<R>* s01 = S + n * n;
  const cx<R>* hDL = s01 + n * n;
  const cx<R>* sDL = hDL + nr * n;
  const cx<R> e = a.e[pt];
  const cx<R> z = {e.re, e.im + a.eta};
  cx<R>* eps = a.eps + (int64_t)pt * n * n;
  cx<R>* epss = a.gs + (int64_t)pt * n * n;

  auto zS = [&](int i, int k) -> cx<R> {
    if (ORTHO) return (i == k) ? z : cx<R>{0, 0};
    return z * S[i * n + k];
  };
  auto zs01 = [&](int i, int k) -> cx<R> {
    if (ORTHO) return cx<R>{0, 0};
    return z * s01[i * n + k];
  };
  // Padded copy of the n x n matrix f(i, k) into AUG[:, :N] (identity outside).
  auto stage = [&](auto f) __attribute__((always_inline)) {
    for (int x = threadIdx.x; x < N * N; x += NTH) {
      const int i = x / N, k = x % N;
      AUG[i * LDG + k] = (i < n && k < n) ? f(i, k) : cx<R>{i == k ? R(1) : R(0), 0};
    }
  };

  const int wave = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  const int g = lane >> 4, c = lane & 15;
  // Column tile of the wave: kind 0 W, 1 alpha, 2 beta.  With 9 waves dealt
  // round-robin over the 4 SIMDs, waves 0, 4, 8 share one SIMD.  Products:
  // W and alpha owners 144 matrix-core steps per iteration, beta owners 288.
  // Elimination: alpha and beta owners update every panel, W tile j sweeps
  // its own four panels and is updated only before them.  The SIMDs hold
  // {W0, alpha0, alpha1}, {W1, beta0}, {W2, beta1}, {alpha2, beta2}: 432
  // product steps each, and at most two panel updates at a time.
  int kind = wave / NT16, jt = wave % NT16;
  if (NT16 == 3) {
    kind = wave < 3 ? 0 : (wave == 3 || wave == 4 || wave == 8) ? 1 : 2;
    jt = kind == 0 ? wave : kind == 2 ? wave - 5 : (wave == 4 ? 0 : wave == 8 ? 1 : 2);
  }
  const int col = 16 * jt + c;
  // A-operand fragment slot of element el (row 16 I + g + 4 r, column col):
  // fragment I * KS + col / 4, lane g + 4 r + 16 (col % 4), stored at that
  // lane XOR 4 ((col / 4) % 4) so that the owners' scattered writes spread
  // over all LDS banks (a reader of k-step kk applies the same XOR).
  auto fslot = [&](int el) -> int {
    const int I = el >> 2, r = el & 3;
    return (I * KS + (col >> 2)) * WAVE + ((g + 4 * r + 16 * (col & 3)) ^ (4 * ((col >> 2) & 3)));
  };
  cx<R>* const Aa = &L.Aa[0][0];
  cx<R>* const Ab = &L.Ab[0][0];
  vE xr, xi;
  // alpha0 = h10 - z s10, beta0 = h01 - z s01, eps = eps_s = H, W = zS - H  (:170-177);
  // the workspace holds D = eps - eps_s (zero here) and the g_s output eps_s
#pragma unroll
  for (int el = 0; el < E; ++el) {
    const int row = GJ::row_of(el, g);
    const bool in = row < n && col < n;
    cx<R> v = {0, 0};
    if (kind == 0) {
      v = cx<R>{row == col ? R(1) : R(0), 0};
      if (in) {
        const cx<R> h = H[row * n + col];
        eps[row * n + col] = cx<R>{0, 0};
        epss[row * n + col] = h;
        v = zS(row, col) - h;
      }
    } else if (kind == 1) {
      if (in) {
        // s10 = s01^H enters as z * conj(s01[k][i]); z itself is not conjugated
        v = conj(h01[col * n + row]);
        if (!ORTHO) v = v - z * conj(s01[col * n + row]);
      }
      Aa[fslot(el)] = v;
    } else {
      if (in) v = h01[row * n + col] - zs01(row, col);
      Ab[fslot(el)] = v;
    }
    xr[el] = v.re;
    xi[el] = v.im;
  }
  if (threadIdx.x < N) L.gj.pinv[threadIdx.x] = threadIdx.x;
  if (threadIdx.x < KS) L.pflag[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    L.gj.singular = 0;
    L.nvote = 0;
    L.abort = 0;
  }
  __syncthreads();

  int status = 0;
  int it = 0;
  // The stop test of iteration it is decided after the first panel barrier of
  // iteration it + 1 (below the iteration limit): the first panel's sweep
  // starts while the votes are written, and a workgroup barrier is saved.  A
  // stop there discards that panel; the iteration's results (eps, eps_s) are
  // complete, and `sing` holds the singular flag of the solves that counted.
  bool voted = false, converged = false;
  int sing = 0;
  auto stop_votes = [&]() -> bool {
    double anyw = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) anyw += red[w];
    return anyw == 0.0;
  };
  while (true) {
    ++it;
    // ---- [X_a | X_b] = W \ [alpha | beta] ----
    bool used = false;
    if constexpr (NT16 == 3) {
      // The sweeping wave of W tile ts copies its tile to one row per lane
      // (v = the first panel, xr / xi = the other twelve columns) and sweeps
      // its four panels there, applying each pivot step to its remaining
      // columns itself; it needs no matrix-core update of its own tile.  Each
      // panel is published into its own slot (in the X fragments' space, free
      // during the solve) with a flag, and the other waves apply the panels in
      // order as their flags appear: no workgroup barrier per panel, so the
      // sweeper runs its four panels back to back and the owner of the next
      // W tile starts as soon as it has applied them.  Waits are bounded: past
      // the limit a wave sets `abort` and the point is given up (status
      // STATUS_SYNC_TIMEOUT) instead of the workgroup hanging.
      typedef typename GJ::Slot Slot;
      Slot* const slots = reinterpret_cast<Slot*>(&L.XS[0][0]);
      cx<R>* const t2r = &L.XS[0][0] + T2R_OFF;
      auto wait_for = [&](const int* f, int want) -> int {
        int v = 0;
        for (int n = 0;; ++n) {
          v = __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if ((v & 0xffff) == want) break;
          if (__hip_atomic_load(&L.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) ||
              n >= SPIN_LIMIT) {
            if (lane == 0) __hip_atomic_store(&L.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            v = 1 << 16;  // stop
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        return v;
      };
      bool stop = false;
      // Applies published panels [p0, p1) to the wave's tile; false on a stop.
      auto apply_panels = [&](int p0, int p1) -> bool {
        for (int pn = p0; pn < p1; ++pn) {
          const int f = wait_for(&L.pflag[pn], it);
          if (f >> 16) return false;
          if (kind == 0) used = GJ::mark_used(used, &slots[pn]);
          GJ::update(xr, xi, &slots[pn]);
        }
        return true;
      };
      if (kind == 0) {
        // W owner of tile jt: the panels of the tiles left of it, then its own
        if (!apply_panels(0, 4 * jt)) {
          stop = true;
        } else {
          cx<R> v[4];
          GJ::tile_to_rows(xr, xi, v, t2r);
          // Panel q of the tile (q a compile-time constant: the row length).
          auto sweep_panel = [&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            constexpr int NC = 16 - 4 * q;  // panel + further columns of the tile
            const int pn = 4 * jt + q;
            cx<R> row[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
              const int src = m + 4 * q - 4;  // register pair of column 4 q + m (v for the first panel)
              if (src < 0) row[m] = v[m < 4 ? m : 0];
              else row[m] = cx<R>{xr[src], xi[src]};
            }
            GJ::template sweep_row<NC>(row, 4 * pn, &slots[pn], used, &L.gj);
#pragma unroll
            for (int m = 0; m < NC; ++m) {
              const int dst = m + 4 * q - 4;
              if (dst >= 0) {
                xr[dst] = row[m].re;
                xi[dst] = row[m].im;
              }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            used = GJ::mark_used(used, &slots[pn]);
            // the first panel carries the previous iteration's stop decision
            if (pn == 0 && it > 1) {
              wait_for(&L.nvote, (NW * (it - 1)) & 0xffff);
              stop = stop_votes() || L.abort;
            }
            __builtin_amdgcn_wave_barrier();
            if (lane == 0)
              __hip_atomic_store(&L.pflag[pn], it | (stop ? 1 << 16 : 0), __ATOMIC_RELEASE,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
          };
          sweep_panel(std::integral_constant<int, 0>{});
          if (!stop) {
            sweep_panel(std::integral_constant<int, 1>{});
            sweep_panel(std::integral_constant<int, 2>{});
            sweep_panel(std::integral_constant<int, 3>{});
          }
        }
      } else {
        stop = !apply_panels(0, KS);
      }
      // every panel applied (and the slots free) before X takes their space;
      // the stop decision and an abort are uniform from here on
      __syncthreads();
      if (L.abort) {
        status |= STATUS_SYNC_TIMEOUT;
        --it;
        break;
      }
      converged = stop;
    } else {
      for (int p = 0; p < KS; ++p) {
        const int j0 = 4 * p, buf = p & 1;
        if (kind == 0 && jt == (j0 >> 4)) {
          GJ::put_panel(xr, xi, j0 & 15, &L.gj);
          GJ::sweep(j0, &L.gj.slot[buf], used, &L.gj);
        }
        __syncthreads();
        if (p == 0 && voted && stop_votes()) {
          converged = true;
          break;
        }
        used = GJ::mark_used(used, &L.gj.slot[buf]);
        // W tiles wholly left of the next panel are no longer needed
        if (kind != 0 || 16 * jt + 15 >= j0 + 4) GJ::update(xr, xi, &L.gj.slot[buf]);
      }
    }
    if (converged) {  // at iteration it - 1
      --it;
      break;
    }
    // The lane's element offsets (row * n + col of element el) and the
    // constants of W = zS - eps are formed from a copy of n the compiler
    // cannot see through, once per group of accesses: formed once, a dozen
    // addresses per matrix stay live across the products, are spilled, and
    // every reload's wait for scratch memory also waits for the loads and
    // stores in flight.
    struct TileIdx {
      int n, g, col;
    };
    auto fresh_n = [&]() {
      TileIdx t = {n, g, col};
      asm volatile("" : "+s"(t.n), "+v"(t.g), "+v"(t.col));
      return t;
    };
    auto tile_off = [&](const TileIdx& t, int el) { return GJ::row_of(el, t.g) * t.n + t.col; };
    auto tile_in = [&](const TileIdx& t, int el) {
      return GJ::row_of(el, t.g) < t.n && t.col < t.n;
    };
    // X (register row i is row pinv[i] of X) to B-operand fragments
    if (kind != 0) {
      const int jc = (kind - 1) * NT16 + jt;
#pragma unroll
      for (int el = 0; el < E; ++el) {
        const int k = L.gj.pinv[GJ::row_of(el, g)];
        L.XS[(k >> 2) * 2 * NT16 + jc][16 * (k & 3) + c] = cx<R>{xr[el], xi[el]};
      }
    }
    // The products' accumulators start from the stored sums, loaded before
    // the barrier, so the loads overlap the other waves' fragment writes:
    //   beta owner, column tile j:  eps_s[:, j] += beta X_a[:, j], then the new
    //                               beta[:, j] = beta X_b[:, j];
    //   W owner:                    D[:, j] += alpha X_b[:, j] with
    //                               D = eps - eps_s, so eps = eps_s + D gets
    //                               beta X_a + alpha X_b;
    //   alpha owner:                the new alpha[:, j] = alpha X_a[:, j].
    // Per SIMD that is 432 matrix-core steps either way (three alpha owners,
    // or a W and a beta owner), and no wave waits for memory mid-product.
    // (the X tile has left the registers: they take the loaded sums).  The
    // loads are unconditional (a load under a lane mask is waited for inside
    // its branch): outside the n x n matrix a lane loads element 0, a finite
    // value that stays in the padding of the accumulator tile (the padded
    // rows of alpha, beta and columns of X are zero, and only in-range
    // elements are stored or enter W).
    if (kind != 1) {
      const cx<R>* src = __builtin_amdgcn_readfirstlane(kind) == 0 ? eps : epss;
      const TileIdx q = fresh_n();
#pragma unroll
      for (int el = 0; el < E; ++el) {
        const cx<R> d = src[tile_in(q, el) ? tile_off(q, el) : 0];
        xr[el] = d.re;
        xi[el] = d.im;
      }
    }
    __syncthreads();
    // ---- products ----
    // acc(I) += sum_kk A(I, kk) B(kk), operands of step kk + 1 loaded before
    // the matrix-core steps of kk are issued
    auto gemm_acc = [&](const cx<R>* A, const cx<R>* B) __attribute__((always_inline)) {
      constexpr int BS = 2 * NT16 * WAVE;
      cx<R> av[NT16], bv = B[lane];
#pragma unroll
      for (int I = 0; I < NT16; ++I) av[I] = A[I * KS * WAVE + lane];  // k-step 0: no XOR
#pragma unroll
      for (int kk = 0; kk < KS; ++kk) {
        cx<R> na[NT16], nb;
        if (kk + 1 < KS) {
          nb = B[(kk + 1) * BS + lane];
#pragma unroll
          for (int I = 0; I < NT16; ++I)
            na[I] = A[(I * KS + kk + 1) * WAVE + (lane ^ (4 * ((kk + 1) & 3)))];
        }
#pragma unroll
        for (int I = 0; I < NT16; ++I) GJ::cmfma(xr, xi, I, av[I], bv);
        if (kk + 1 < KS) {
          bv = nb;
#pragma unroll
          for (int I = 0; I < NT16; ++I) av[I] = na[I];
        }
      }
    };
    const cx<R>* XS = &L.XS[0][0];
    if (kind == 0) {
      gemm_acc(Aa, XS + (NT16 + jt) * WAVE);
      const TileIdx q = fresh_n();
#pragma unroll
      for (int el = 0; el < E; ++el)
        if (tile_in(q, el)) eps[tile_off(q, el)] = cx<R>{xr[el], xi[el]};
    } else {
      if (kind == 2) {
        gemm_acc(Ab, XS + jt * WAVE);
        const TileIdx q = fresh_n();
#pragma unroll
        for (int el = 0; el < E; ++el)
          if (tile_in(q, el)) epss[tile_off(q, el)] = cx<R>{xr[el], xi[el]};
      }
#pragma unroll
      for (int el = 0; el < E; ++el) {
        xr[el] = 0;
        xi[el] = 0;
      }
      // alpha X_a (the new alpha) or beta X_b (the new beta)
      gemm_acc(kind == 1 ? Aa : Ab, XS + ((kind - 1) * NT16 + jt) * WAVE);
    }
    __syncthreads();
    sing = L.gj.singular;  // no solve writes it again before this iteration's vote is read
    if (kind != 0) {
      cx<R>* A = kind == 1 ? Aa : Ab;
#pragma unroll
      for (int el = 0; el < E; ++el) A[fslot(el)] = cx<R>{xr[el], xi[el]};
    } else {
      // W = zS - eps_s - D for the next solve; eps_s[:, j] was stored by the
      // beta owner of tile j before the barrier (unconditional loads, as above)
      const TileIdx q = fresh_n();
      double one = 1.0, zr = z.re, zi = z.im;
      asm volatile("" : "+v"(one), "+v"(zr), "+v"(zi));
      cx<R> es[E];
#pragma unroll
      for (int el = 0; el < E; ++el) es[el] = epss[tile_in(q, el) ? tile_off(q, el) : 0];
#pragma unroll
      for (int el = 0; el < E; ++el) {
        const int row = GJ::row_of(el, g);
        const bool in = tile_in(q, el);
        const int off = in ? tile_off(q, el) : 0;
        const cx<R> zq = {zr, zi};
        const cx<R> zs = ORTHO ? (row == col ? zq : cx<R>{0, 0}) : zq * S[off];
        const cx<R> w = zs - (es[el] + cx<R>{xr[el], xi[el]});
        xr[el] = in ? w.re : (row == col ? one : 0.0);
        xi[el] = in ? w.im : 0.0;
      }
    }
    if (threadIdx.x < N) L.gj.pinv[threadIdx.x] = threadIdx.x;
    __syncthreads();
    // Stop when max_ij |alpha_ij| + |beta_ij| < stop  (:205); the alpha owners
    // read beta's fragments.  Only the decision is needed: a wave vote, with
    // the hardware square root (approximate, but near the threshold the terms
    // shrink by many orders of magnitude per iteration).  NaN keeps
    // iterating, as the reference's comparison does.
    bool go_on = false;
    if (kind == 1) {
#pragma unroll
      for (int el = 0; el < E; ++el) {
        const int row = GJ::row_of(el, g);
        const cx<R> bv = Ab[fslot(el)];
        const double sa = __builtin_amdgcn_sqrt(xr[el] * xr[el] + xi[el] * xi[el]);
        const double sb = __builtin_amdgcn_sqrt(bv.re * bv.re + bv.im * bv.im);
        if (row < n && col < n) go_on |= !(sa + sb < a.stop);
      }
    }
    go_on = __any(go_on);
    if (lane == 0) red[wave] = go_on ? 1.0 : 0.0;
    if (NT16 == 3 && lane == 0)
      __hip_atomic_fetch_add(&L.nvote, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (it >= 101) {  // the limit: decide now (surface_green.py:222-224)
      __syncthreads();
      if (!stop_votes()) status |= STATUS_LS_NOCONV;
      break;
    }
    voted = true;
  }
  if (sing) status |= STATUS_SINGULAR;
  __syncthreads();

  // g_s = (zS - eps_s)^-1, written over eps_s (the g_s output)   (:209)
  stage([&](int i, int k) { return zS(i, k) - epss[i * n + k]; });
  __syncthreads();
  status |= wg_gj_inverse_blocked<R, NTH, N, LDG>(
      AUG, [&](int i, int k, cx<R> v) { if (i < n && k < n) epss[i * n + k] = v; }, &gjb);
  const cx<R>* gs = epss;
  // residual: max |(zS - H - (z s01 - h01) g_s (z s10 - h10)) g_s - I|   (:211-212)
  wg_gemm<R, NW, MAXT1>(
      n, n, n, [&](int i, int k) { return zs01(i, k) - h01[i * n + k]; },
      [&](int k, int j) { return gs[k * n + j]; },
      [&](int i, int j, cx<R> v) { SP[i * LDP + j] = v; });
  __syncthreads();
  wg_gemm<R, NW, MAXT1>(
      n, n, n, [&](int i, int k) { return SP[i * LDP + k]; },
      [&](int k, int j) {
        cx<R> s10 = ORTHO ? cx<R>{0, 0} : z * conj(s01[j * n + k]);
        return s10 - conj(h01[j * n + k]);
      },
      [&](int i, int j, cx<R> v) { AUG[i * LDG + N + j] = zS(i, j) - H[i * n + j] - v; });
  __syncthreads();
  double rm = 0.0;
  wg_gemm<R, NW, MAXT1>(
      n, n, n, [&](int i, int k) { return AUG[i * LDG + N + k]; },
      [&](int k, int j) { return gs[k * n + j]; },
      [&](int i, int j, cx<R> v) {
        if (i == j) v.re -= 1.0;
        rm = fmax(rm, cabs(v));
      });
  rm = block_max(rm, red);
  if (rm >= 3.0e-5) status |= STATUS_LS_RESID_FAIL;
  else if (rm > 1.0e-8) status |= STATUS_LS_RESID_WARN;

  cx<R>* sig = a.sigma + (int64_t)pt * nr * nr;
  if (a.mode == 0) {
    // Y = g_s (e S_DL^H - H_DL^H) [n x nr];  Sigma = (e S_DL - H_DL) Y
    wg_gemm<R, NW, MAXT1>(
        n, nr, n, [&](int i, int k) { return gs[i * n + k]; },
        [&](int k, int j) { return e * conj(sDL[j * n + k]) - conj(hDL[j * n + k]); },
        [&](int i, int j, cx<R> v) { AUG[i * LDG + N + j] = v; });
    __syncthreads();
    wg_gemm<R, NW, MAXT1>(
        nr, nr, n, [&](int i, int k) { return e * sDL[i * n + k] - hDL[i * n + k]; },
        [&](int k, int j) { return AUG[k * LDG + N + j]; },
        [&](int i, int j, cx<R> v) { sig[i * nr + j] = v; });
  } else {
    // g_s^-1 by a second inverse; mode 1: (e S - H) - g_s^-1, mode 2: g_s^-1
    const int mode = a.mode;
    stage([&](int i, int k) { return gs[i * n + k]; });
    __syncthreads();
    status |= wg_gj_inverse_blocked<R, NTH, N, LDG>(
        AUG,
        [&](int i, int k, cx<R> v) {
          if (i < n && k < n) {
            if (mode == 1) sig[i * n + k] = e * S[i * n + k] - H[i * n + k] - v;
            else sig[i * n + k] = v;
          }
        },
        &gjb);
  }
  if (threadIdx.x == 0) {
    if (a.iters) a.iters[pt] = it;
    if (a.status) a.status[pt] = status;
    if (a.resid) a.resid[pt] = rm;
  }
}

template <int NMAX>
static hipError_t launch_ls(const LeadArgs& a, bool ortho, hipStream_t st) {
  dim3 grid(a.P), block(ls_threads<NMAX>());
  if (ortho)
    hipLaunchKernelGGL((lead_self_energy_kernel<NMAX, true>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((lead_self_energy_kernel<NMAX, false>), grid, block, 0, st, a);
  return hipGetLastError();
}

int64_t lead_self_energy_workspace(int n, int P) { return (int64_t)P * n * n; }

hipError_t launch_lead_self_energy(const LeadArgs& a, bool ortho, hipStream_t st) {
  const int m = a.n > a.nred ? a.n : a.nred;
  if (m <= 16) return launch_ls<16>(a, ortho, st);
  if (m <= 32) return launch_ls<32>(a, ortho, st);
  if (m <= 48) return launch_ls<48>(a, ortho, st);
  return hipErrorInvalidValue;
}

}  // namespace negf
