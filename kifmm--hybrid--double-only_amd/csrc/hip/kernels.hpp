/*
 * kernels.hpp -- gfx950 kernels of the KIFMM evaluation path (FP64 plain path).
 *
 * Stage map (reference kernel it replaces, fmm/cuda.cu):
 *   k_p2p_ulist ....... ulist_eval__gpu             (604-663)   U-list P2P
 *   k_s2m ............. up_eval__gpu, first half    (108-190)   sources -> UC check potential
 *   k_gemm_cols ....... the dgemv loops of up_eval__gpu, up_eval__gpu_non_leaf,
 *                       down_eval__gpu_depth_2 / _non_leaves    (430-528, 1682-1801)
 *                       as FP64 MFMA GEMMs over a column list
 *   k_dft_fwd ......... compute_fft_src__gpu_eval + cuFFT D2Z   (718-871)
 *   k_trans ........... compute_fft_trans__gpu_eval + D2Z x343  (874-963)
 *   k_hadamard ........ vlist_calc__gpu_eval                    (1200-1240)
 *   k_dft_inv ......... scale_nrmfc + cuFFT Z2D + regVal2SamVal (965-1196)
 *   k_wx<., false> .... wlist_eval__gpu                         (2070-2197)
 *   k_wx<., true> ..... xlist_eval__gpu                         (2380-2493)
 *   k_l2t ............. down_eval__gpu_leaves                   (1804-1872)
 *
 * The FP64 U-list gives a wave 8 targets held in scalar registers, its 64
 * lanes stride over a source tile staged in LDS, and the 8 per-lane partial
 * sums are combined by a 6-step butterfly (recursive halving) across the
 * wave.  The W / X lists and the L2T give each lane a target and read the
 * sources as wave-wide broadcasts.  Sources are (x, y, z, w) double4
 * records; tiles are padded with zero-weight points far outside the root
 * box so no per-pair bounds checks are needed.
 */
#pragma once
#include <hip/hip_runtime.h>

namespace kifmm_dev {

constexpr double OOFP_R = 0.079577471545947667884441881686257181; /* 1/(4 pi) */
constexpr double FAR = 1.0e8; /* coordinate of padding sources */

typedef double v4d __attribute__((ext_vector_type(4)));

/* x^(-1/2): the hardware estimate (v_rsq_f64, relative error below 2^-24,
 * measured max 5.2e-8) refined by one Newton step, e = 1 - x y^2,
 * y <- y + y e / 2, which leaves a relative error of (3/2) e0^2 <= 2e-14
 * (4e-15 for the measured estimate): far below the 1e-12 parity bar, and
 * one instruction per pair cheaper than a third-order step. */
/* 2 / sqrt(x): the v_rsq_f64 estimate y (relative error <= 2^-24.2 on this
 * chip) and one Newton step written as y (3 - x y^2) = (2 / sqrt(x)) (1 -
 * 3e^2/2 + ...): three VALU instructions after the rsq.  The factor 2 is
 * left in; pair sums built from it are halved once at the end (an exact
 * power-of-two scale), which saves a multiply per pair. */
__device__ __forceinline__ double rsqrt2_fp64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  return y * __builtin_fma(-x, y * y, 3.0);
}
__device__ __forceinline__ double rsqrt_fp64(double x) { return 0.5 * rsqrt2_fp64(x); }

/* 2 / r, and 1 / r; MASK: r = 0 (coincident points) gives 0 */
template <bool MASK>
__device__ __forceinline__ double inv_r2(double r2) {
  double y = rsqrt2_fp64(r2);
  if (MASK) y = (r2 > 0.0) ? y : 0.0;
  return y;
}
template <bool MASK>
__device__ __forceinline__ double inv_r(double r2) {
  return 0.5 * inv_r2<MASK>(r2);
}
constexpr double HALF_OOFP_R = 0.5 * OOFP_R; /* scales sums of 2 w / r */

/* 8 partial sums per lane -> lane l holds the wave total of target (l >> 3)
 * in lanes with (l & 7) == 0.  Fixed order, deterministic. */
__device__ __forceinline__ double wave_reduce8(const double acc[8], int lane) {
  double a[4], b[2], c;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const bool hi = lane & 32;
    double send = hi ? acc[i] : acc[i + 4];
    double keep = hi ? acc[i + 4] : acc[i];
    a[i] = keep + __shfl_xor(send, 32);
  }
#pragma unroll
  for (int i = 0; i < 2; i++) {
    const bool hi = lane & 16;
    double send = hi ? a[i] : a[i + 2];
    double keep = hi ? a[i + 2] : a[i];
    b[i] = keep + __shfl_xor(send, 16);
  }
  {
    const bool hi = lane & 8;
    double send = hi ? b[0] : b[1];
    double keep = hi ? b[1] : b[0];
    c = keep + __shfl_xor(send, 8);
  }
  c += __shfl_xor(c, 4);
  c += __shfl_xor(c, 2);
  c += __shfl_xor(c, 1);
  return c;
}

/* acc[t] += sum over LDS sources [0, ns) (ns multiple of 64) of 2 w / |x_t - s| */
template <bool MASK>
__device__ __forceinline__ void eval8_lds(const double4* __restrict__ src, int ns, const double tx[8],
                                          const double ty[8], const double tz[8], double acc[8],
                                          int lane) {
  for (int s = lane; s < ns; s += 64) {
    const double4 p = src[s];
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const double dx = tx[t] - p.x, dy = ty[t] - p.y, dz = tz[t] - p.z;
      const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
      acc[t] = __builtin_fma(p.w, inv_r2<MASK>(r2), acc[t]);
    }
  }
}

/* ------------------------------------------------------------------ U-list */

/* U-list work item = {source range set, first target, target count, leaf},
 * up to 4 * 8 * G targets of one target leaf.  A block of 4 waves takes one
 * item: target group g (8 targets) belongs to wave g % 4, so a wave carries
 * up to G groups of accumulators and every staged source tile serves all
 * the item's targets.  The sources are the leaf's U-list, a list of merged
 * ranges {beg, num} of the gathered source array, walked in order.  (The 8
 * targets of a group are wave-uniform and live in SGPRs.  Per pair the
 * difference form spends 3 sub + 3 for r^2, v_rsq_f64 + 3 for the Newton
 * step and 1 FMA; the Gram form, below, 1 add + 3 FMA for r^2.)
 * Per target, the sum is the same for every item size and launch form:
 * the tiles depend only on the ranges, the lane order and the reduction
 * are fixed. */
constexpr int P2P_TILE = 512;
constexpr int P2P_G = 4;
constexpr int P2P_MAXT = 4 * 8 * P2P_G;
/* The FP64 Gram-form kernels take items of up to 160 targets (5 groups per
 * wave, 120 VGPRs: still 4 waves per SIMD): a leaf of 129-160 targets is
 * then one item instead of two, which saves a source walk and evens out
 * the groups per wave. */
constexpr int P2P_G5 = 5;
constexpr int P2P_MAXT5 = 4 * 8 * P2P_G5;

/* Walks an item's source ranges {beg, num} one tile (P2P_TILE sources) at a
 * time, in order.  The walk state is block-uniform; thread tid fetches the
 * sources of tile slots tid and tid + 256 into registers, so a tile's loads
 * are in flight while the previous tile is computed.  Returns the tile's
 * source count (0 when the ranges are used up). */
constexpr int P2P_PER = P2P_TILE / 256;
struct RangeWalk {
  int ri, ro, rlast;
};
__device__ __forceinline__ int walk_tile(RangeWalk& w, const int2* __restrict__ rng,
                                         const double4* __restrict__ src4, int tid, double4 (&pre)[P2P_PER]) {
  int fill = 0, idx[P2P_PER];
#pragma unroll
  for (int m = 0; m < P2P_PER; m++) idx[m] = -1;
  while (fill < P2P_TILE && w.ri < w.rlast) {
    const int2 r = rng[w.ri];
    const int take = min(r.y - w.ro, P2P_TILE - fill);
#pragma unroll
    for (int m = 0; m < P2P_PER; m++) {
      const int sl = tid + 256 * m - fill;
      if (sl >= 0 && sl < take) idx[m] = r.x + w.ro + sl;
    }
    fill += take;
    w.ro += take;
    if (w.ro == r.y) {
      w.ri++;
      w.ro = 0;
    }
  }
#pragma unroll
  for (int m = 0; m < P2P_PER; m++)
    if (idx[m] >= 0) pre[m] = src4[idx[m]];
  return fill;
}

/* FP64 U-list on the Gram form (used when the close pairs are listed, see
 * k_close_near).  Coordinates are centred on the target leaf and scaled by
 * sc = GRAM_SC / h (a power of two: exact), so a pair is close, r <
 * CLOSE_FRAC h, exactly when its scaled distance r' is below 2.  Per pair
 *     r'^2 = (|t|^2 + |s|^2) + (-2t) . s              (1 add, 3 FMA)
 *     v = clamp01(y (3 - r'^2 y^2)),  y = rsq(|r'^2|)  (rsq, mul, FMA, mul)
 * then 1 FMA into the sum: 9 VALU instructions against 11 for the
 * difference form, and v = 2 / r' for every r' >= 2.  For those pairs the
 * Gram r'^2 is within about eps (2|t| + r')^2 of the exact value, with |t|
 * <= 16 sqrt(3) here: below 2e-13 relative at r' = 2 and falling with r',
 * whatever the size of the source leaf.  Each listed close pair then adds
 * w (2 / r'_exact - v) when its target's sum is written, with v recomputed
 * by the same instructions (gram_pair) from the same operands, so the
 * clamped or cancelled main-loop term is replaced exactly (r = 0 with MASK:
 * 2 / r' taken as 0).  The clamp keeps every main-loop term in [0, 1]:
 * r'^2 = 0 gives 0 * inf = NaN in the Newton step, which the clamp maps to 0. */
constexpr double CLOSE_FRAC = 0.125;
constexpr double GRAM_SC = 2.0 / CLOSE_FRAC;
constexpr double GRAM_FAR = 1.0e6; /* scaled coordinate of padding sources */

/* (x', y', z', |s'|^2) of point p for leaf centre g and scale sc */
__device__ __forceinline__ double4 gram_src(double4 p, double4 g, double sc) {
  const double x = (p.x - g.x) * sc, y = (p.y - g.y) * sc, z = (p.z - g.z) * sc;
  return make_double4(x, y, z, __builtin_fma(x, x, __builtin_fma(y, y, z * z)));
}
/* the main-loop term of target record a = (-2x', -2y', -2z', |t'|^2) and
 * source record s = gram_src(...) */
__device__ __forceinline__ double gram_pair(double4 a, double4 s) {
  const double r2 = __builtin_fma(a.x, s.x, __builtin_fma(a.y, s.y, __builtin_fma(a.z, s.z, a.w + s.w)));
  const double y = __builtin_amdgcn_rsq(__builtin_fabs(r2));
  const double v = y * __builtin_fma(-r2, y * y, 3.0);
  return __builtin_fmin(__builtin_fmax(v, 0.0), 1.0);
}

/* acc[t] += w * gram_pair(a[t], s) over LDS sources [0, ns) (ns a multiple of 64) */
__device__ __forceinline__ void eval8_gram(const double4* __restrict__ src, const double* __restrict__ sw, int ns,
                                           const double4 (&a)[8], double acc[8], int lane) {
#pragma unroll 2
  for (int s = lane; s < ns; s += 64) {
    const double4 p = src[s];
    const double w = sw[s];
#pragma unroll
    for (int t = 0; t < 8; t++) acc[t] = __builtin_fma(w, gram_pair(a[t], p), acc[t]);
  }
}

/* Scratch of the Gram form: the scaled source weights of a tile, the close-
 * pair sums of an item's targets (LDS), and the per-target records
 * (-2t', |t'|^2) with the close-pair CSR list (device memory). */
struct GramArgs {
  double* sw;
  double* sc_sum;
  const double4* __restrict__ geo;
  const double4* __restrict__ trg_gram;
  const int* __restrict__ cptr;
  const int* __restrict__ cidx;
};

template <bool MASK, int G, bool GRAM>
__device__ __forceinline__ void p2p_item(double4* tile, const int4 it, int rot, const int* __restrict__ rng_ptr,
                                         const int2* __restrict__ rng, const double4* __restrict__ src4,
                                         const double4* __restrict__ trg4, double* __restrict__ pot,
                                         const GramArgs& ga, int accum) {
  const int tid = threadIdx.x, lane = tid & 63;
  /* groups g = wave' + 4k with wave' the wave index rotated by the item:
   * when the group count is not a multiple of 4 the short wave moves from
   * SIMD to SIMD instead of always idling the same one */
  const int wave = (__builtin_amdgcn_readfirstlane(tid >> 6) + rot) & 3;
  const int ngroups = (it.z + 7) >> 3;
  double4 g = make_double4(0.0, 0.0, 0.0, 1.0);
  double sc = 1.0;
  if constexpr (GRAM) {
    g = ga.geo[it.w];
    sc = GRAM_SC / g.w;
  }
  double acc[G][8];
#pragma unroll
  for (int k = 0; k < G; k++)
#pragma unroll
    for (int t = 0; t < 8; t++) acc[k][t] = 0.0;
  int ri = rng_ptr[it.x], ro = 0;
  const int rlast = rng_ptr[it.x + 1];
  while (ri < rlast) {
    /* fill the tile: block-uniform walk over the ranges (every range has
     * num >= 1, so each pass consumes at least one source).  Loading the
     * next tile ahead into registers (as the mixed kernel does) costs this
     * kernel spills and a wave of occupancy: measured slower. */
    __syncthreads();
    int fill = 0;
    while (fill < P2P_TILE && ri < rlast) {
      const int2 r = rng[ri];
      const int take = min(r.y - ro, P2P_TILE - fill);
      for (int k = tid; k < take; k += 256) {
        const double4 q = src4[r.x + ro + k];
        if constexpr (GRAM) {
          tile[fill + k] = gram_src(q, g, sc);
          ga.sw[fill + k] = q.w;
        } else {
          tile[fill + k] = q;
        }
      }
      fill += take;
      ro += take;
      if (ro == r.y) {
        ri++;
        ro = 0;
      }
    }
    const int ns = (fill + 63) & ~63;
    for (int k = fill + tid; k < ns; k += 256) {
      if constexpr (GRAM) {
        tile[k] = make_double4(GRAM_FAR, GRAM_FAR, GRAM_FAR, 3.0 * GRAM_FAR * GRAM_FAR);
        ga.sw[k] = 0.0;
      } else {
        tile[k] = make_double4(FAR, FAR, FAR, 0.0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < G; k++) {
      const int grp = wave + 4 * k;
      if (grp < ngroups) {
        /* slots past the count re-read the item's last target (in range) */
        if constexpr (GRAM) {
          double4 a[8];
#pragma unroll
          for (int t = 0; t < 8; t++) a[t] = ga.trg_gram[it.y + min(8 * grp + t, it.z - 1)];
          eval8_gram(tile, ga.sw, ns, a, acc[k], lane);
        } else {
          double tx[8], ty[8], tz[8];
#pragma unroll
          for (int t = 0; t < 8; t++) {
            const double4 q = trg4[it.y + min(8 * grp + t, it.z - 1)];
            tx[t] = q.x; ty[t] = q.y; tz[t] = q.z;
          }
          eval8_lds<MASK>(tile, ns, tx, ty, tz, acc[k], lane);
        }
      }
    }
  }
  if constexpr (GRAM) {
    /* close-pair terms: a group's pairs are one contiguous run of cidx
     * (targets in order), spread over the 64 lanes and summed per target by
     * a segmented scan over the lanes (fixed order: deterministic) */
    const double isc = g.w / GRAM_SC; /* 1 / sc, exact */
#pragma unroll
    for (int k = 0; k < G; k++) {
      const int grp = wave + 4 * k;
      if (grp < ngroups) {
        const int t0 = 8 * grp, nt = min(8, it.z - t0), tg0 = it.y + t0;
        if (lane < 8) ga.sc_sum[t0 + lane] = 0.0;
        const int kbeg = ga.cptr[tg0], kend = ga.cptr[tg0 + nt];
        const int ej = ga.cptr[tg0 + min(lane & 7, nt - 1) + 1]; /* end of target lane&7's run */
        for (int kb = kbeg; kb < kend; kb += 64) {
          const int kk = kb + lane;
          const bool live = kk < kend;
          int tl = 0; /* the target of pair kk: the number of runs that end at or before kk */
#pragma unroll
          for (int j = 0; j < 8; j++) tl += __builtin_amdgcn_readlane(ej, j) <= kk;
          tl = live ? tl : 8;
          double v = 0.0;
          if (live) {
            const double4 q = src4[ga.cidx[kk]];
            const double4 tp = trg4[tg0 + tl];
            const double dx = tp.x - q.x, dy = tp.y - q.y, dz = tp.z - q.z;
            const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
            v = q.w * (inv_r2<MASK>(r2) * isc - gram_pair(ga.trg_gram[tg0 + tl], gram_src(q, g, sc)));
          }
#pragma unroll
          for (int off = 1; off < 64; off <<= 1) {
            const double u = __shfl_up(v, off, 64);
            const int ut = __shfl_up(tl, off, 64);
            if (lane >= off && ut == tl) v += u;
          }
          const int nx = __shfl_down(tl, 1, 64);
          if (live && (lane == 63 || nx != tl)) ga.sc_sum[t0 + tl] += v; /* one writer per target and chunk */
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < G; k++) {
    const int grp = wave + 4 * k;
    if (grp < ngroups) {
      const double sum = wave_reduce8(acc[k], lane);
      const int t = 8 * grp + (lane >> 3);
      if ((lane & 7) == 0 && t < it.z) {
        const double v = GRAM ? (HALF_OOFP_R * sc) * (sum + ga.sc_sum[t]) : HALF_OOFP_R * sum;
        pot[it.y + t] = accum ? v + pot[it.y + t] : v;
      }
    }
  }
}

/* One block per item of at most 32 G targets. */
template <bool MASK, int G, bool GRAM>
__global__ __launch_bounds__(256) void k_p2p_ulist(const int4* __restrict__ items, int n_items, int per_xcd,
                                                   const int* __restrict__ rng_ptr, const int2* __restrict__ rng,
                                                   const double4* __restrict__ src4,
                                                   const double4* __restrict__ trg4, double* __restrict__ pot,
                                                   const double4* __restrict__ geo,
                                                   const double4* __restrict__ trg_gram,
                                                   const int* __restrict__ cptr, const int* __restrict__ cidx,
                                                   int accum) {
  __shared__ double4 tile[P2P_TILE];
  __shared__ double sw[GRAM ? P2P_TILE : 1];
  __shared__ double sc_sum[GRAM ? 32 * G : 1];
  /* workgroups go round-robin to the 8 XCDs: give each XCD a contiguous run
   * of items (neighbouring leaves) so their shared sources stay in its L2 */
  const int item = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (item >= n_items) return; /* whole block, before any barrier */
  const GramArgs ga{sw, sc_sum, geo, trg_gram, cptr, cidx};
  p2p_item<MASK, G, GRAM>(tile, items[item], item, rng_ptr, rng, src4, trg4, pot, ga, accum);
}

/* The same with items of at most 64 targets (G = 2 groups per wave): the
 * compiler is held to 6 waves per SIMD, which these 16 accumulators per
 * lane allow without spills (71 VGPRs on the Gram form, against 104 when it
 * is left free): 5.89 -> 5.64 ms at 4M points.  (For the 128-target items
 * more waves did not pay: 8 waves x 2 groups per block ran 5.75 ms against
 * 5.61 for 4 x 4.) */
template <bool MASK, bool GRAM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_p2p_ulist64(
    const int4* __restrict__ items, int n_items, int per_xcd, const int* __restrict__ rng_ptr,
    const int2* __restrict__ rng, const double4* __restrict__ src4, const double4* __restrict__ trg4,
    double* __restrict__ pot, const double4* __restrict__ geo, const double4* __restrict__ trg_gram,
    const int* __restrict__ cptr, const int* __restrict__ cidx, int accum) {
  __shared__ double4 tile[P2P_TILE];
  __shared__ double sw[GRAM ? P2P_TILE : 1];
  __shared__ double sc_sum[GRAM ? P2P_MAXT : 1];
  const int item = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (item >= n_items) return; /* whole block, before any barrier */
  const GramArgs ga{sw, sc_sum, geo, trg_gram, cptr, cidx};
  p2p_item<MASK, P2P_G / 2, GRAM>(tile, items[item], item, rng_ptr, rng, src4, trg4, pot, ga, accum);
}

/* Persistent form over the same items (<= 32 G targets): a fixed grid
 * of a few blocks per CU; the blocks of XCD x draw the items of x's
 * contiguous run one at a time from counter x (zeroed before the launch),
 * so uneven items even out inside each XCD.  Every block leaves the loop
 * at its first draw past the run's end; the counters only grow. */
template <bool MASK, int G, bool GRAM>
__global__ __launch_bounds__(256) void k_p2p_ulist_persist(const int4* __restrict__ items, int n_items, int per_xcd,
                                                           const int* __restrict__ rng_ptr,
                                                           const int2* __restrict__ rng,
                                                           const double4* __restrict__ src4,
                                                           const double4* __restrict__ trg4, double* __restrict__ pot,
                                                           int* __restrict__ counters,
                                                           const double4* __restrict__ geo,
                                                           const double4* __restrict__ trg_gram,
                                                           const int* __restrict__ cptr, const int* __restrict__ cidx,
                                                           int accum) {
  __shared__ double4 tile[P2P_TILE];
  __shared__ double sw[GRAM ? P2P_TILE : 1];
  __shared__ double sc_sum[GRAM ? 32 * G : 1];
  __shared__ int draw;
  const int x = (int)(blockIdx.x & 7);
  const int lo = x * per_xcd, n = min(n_items - lo, per_xcd);
  const GramArgs ga{sw, sc_sum, geo, trg_gram, cptr, cidx};
  for (;;) {
    if (threadIdx.x == 0) draw = atomicAdd(counters + x, 1);
    __syncthreads();
    const int k = draw;
    __syncthreads(); /* every thread has read draw before the next draw overwrites it */
    if (k >= n) break;
    p2p_item<MASK, G, GRAM>(tile, items[lo + k], lo + k, rng_ptr, rng, src4, trg4, pot, ga, accum);
  }
}

/* The Gram form's per-target records (-2t', |t'|^2), t' = (t - c) sc for the
 * target's leaf (centre c, sc = GRAM_SC / h): fixed for the handle's
 * lifetime, written once at create time.  One block per U-list item. */
static __global__ __launch_bounds__(128) void k_trg_gram(const int4* __restrict__ items,
                                                        const double4* __restrict__ geo,
                                                        const double4* __restrict__ trg4,
                                                        double4* __restrict__ trg_gram) {
  const int4 it = items[blockIdx.x];
  const int k = threadIdx.x;
  if (k >= it.z) return;
  const double4 g = geo[it.w];
  const double4 s = gram_src(trg4[it.y + k], g, GRAM_SC / g.w);
  trg_gram[it.y + k] = make_double4(-2.0 * s.x, -2.0 * s.y, -2.0 * s.z, s.w);
}

static __global__ void k_fill_int(int* __restrict__ p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

/* ------------------------------------------------- U-list, mixed precision */

typedef float v4f __attribute__((ext_vector_type(4)));

/* Mixed-precision U-list (NP <= 4).  Coordinates are centred on the
 * target leaf and scaled by 1/h (h = its half-width, a power of two, so
 * the scaling is exact); the distances of a 16 x 16 (target x source) tile
 * come from one v_mfma_f32_16x16x4f32 on the Gram form
 *     r^2 = |t|^2 + [-2tx, -2ty, -2tz, 1] . [sx, sy, sz, |s|^2]
 * (A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane
 * j + 16(i/4), element i%4: measured by test_mfma_f32_operand_layout, C
 * preloaded with |t|^2), then w * v_rsq_f32(r^2) is summed in FP32 over
 * a source tile and folded into FP64 accumulators.  The FP32 Gram r^2 has
 * an absolute error near 4 eps32 (r + 2 sqrt(3))^2, so pairs closer than
 * MIX_TAU h (Gram r^2 < MIX_TAU^2) are recomputed in FP64 from the FP64
 * centred coordinates and added to the FP64 sums (r = 0 exactly 0 for
 * coincident points); every other pair's 1/r is within ~2e-4 relative, a
 * contribution below 1e-6 of a target's U-list sum. */
constexpr float MIX_TAU2 = 0.01f;
constexpr double MIX_FAR = 1.0e6;

template <bool MASK, int G>
__device__ __forceinline__ void mixed_chunks(const float4* __restrict__ sB, const float* __restrict__ sW,
                                             const double4* __restrict__ sS, const double4* const* sTg,
                                             int nchunks, const float* a, const v4f* c, double (*acc)[4],
                                             int lane) {
  const int j = lane & 15, k4 = lane >> 4;
  const float* sBf = (const float*)sB;
  float part[G][4];
#pragma unroll
  for (int q = 0; q < G; q++)
#pragma unroll
    for (int e = 0; e < 4; e++) part[q][e] = 0.0f;
  for (int ch = 0; ch < nchunks; ch++) {
    const int s = ch * 16 + j;
    const float b = sBf[4 * s + k4];
    const float w = sW[s];
    v4f d[G];
#pragma unroll
    for (int q = 0; q < G; q++) d[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q], b, c[q], 0, 0, 0);
    bool close = false;
#pragma unroll
    for (int q = 0; q < G; q++)
      close |= (d[q][0] < MIX_TAU2) | (d[q][1] < MIX_TAU2) | (d[q][2] < MIX_TAU2) | (d[q][3] < MIX_TAU2);
    if (__any(close)) {
      if (close) {
        const double4 sp = sS[s];
#pragma unroll
        for (int q = 0; q < G; q++)
#pragma unroll
          for (int e = 0; e < 4; e++) {
            if (d[q][e] < MIX_TAU2) {
              const double4 tp = sTg[q][4 * k4 + e];
              const double dx = tp.x - sp.x, dy = tp.y - sp.y, dz = tp.z - sp.z;
              acc[q][e] = __builtin_fma(sp.w, inv_r<MASK>(__builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx))),
                                        acc[q][e]);
              d[q][e] = __builtin_inff();
            }
          }
      }
    }
#pragma unroll
    for (int q = 0; q < G; q++)
#pragma unroll
      for (int e = 0; e < 4; e++) part[q][e] = __builtin_fmaf(w, __builtin_amdgcn_rsqf(d[q][e]), part[q][e]);
  }
#pragma unroll
  for (int q = 0; q < G; q++)
#pragma unroll
    for (int e = 0; e < 4; e++) acc[q][e] += (double)part[q][e];
}

template <bool MASK>
__global__ __launch_bounds__(256) void k_p2p_ulist_mixed(const int4* __restrict__ items, int n_items, int per_xcd,
                                                         const int* __restrict__ rng_ptr,
                                                         const int2* __restrict__ rng,
                                                         const double4* __restrict__ geo,
                                                         const double4* __restrict__ src4,
                                                         const double4* __restrict__ trg4,
                                                         double* __restrict__ pot, int accum) {
  __shared__ float4 sB[P2P_TILE];
  __shared__ float sW[P2P_TILE];
  __shared__ double4 sS[P2P_TILE];
  __shared__ double4 sT[P2P_MAXT];
  const int item = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (item >= n_items) return; /* whole block, before any barrier */
  const int4 it = items[item];
  const double4 g = geo[it.w];
  const double ih = 1.0 / g.w;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ngroups = (it.z + 15) >> 4;
  /* targets in leaf units; slots past the count repeat the last target */
  for (int k = tid; k < ngroups * 16; k += 256) {
    const double4 q = trg4[it.y + min(k, it.z - 1)];
    const double x = (q.x - g.x) * ih, y = (q.y - g.y) * ih, z = (q.z - g.z) * ih;
    sT[k] = make_double4(x, y, z, __builtin_fma(x, x, __builtin_fma(y, y, z * z)));
  }
  __syncthreads();
  float a[2];
  v4f c[2];
  double acc[2][4];
  const double4* tg[2];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int grp = min(wave + 4 * q, ngroups - 1);
    tg[q] = sT + 16 * grp;
    const double4 t = tg[q][lane & 15];
    const int k4 = lane >> 4;
    a[q] = (float)(k4 == 0 ? -2.0 * t.x : k4 == 1 ? -2.0 * t.y : k4 == 2 ? -2.0 * t.z : 1.0);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      c[q][e] = (float)tg[q][4 * k4 + e].w;
      acc[q][e] = 0.0;
    }
  }
  int ri = rng_ptr[it.x], ro = 0;
  const int rlast = rng_ptr[it.x + 1];
  while (ri < rlast) {
    __syncthreads();
    int fill = 0;
    while (fill < P2P_TILE && ri < rlast) {
      const int2 r = rng[ri];
      const int take = min(r.y - ro, P2P_TILE - fill);
      for (int k = tid; k < take; k += 256) {
        const double4 q = src4[r.x + ro + k];
        const double x = (q.x - g.x) * ih, y = (q.y - g.y) * ih, z = (q.z - g.z) * ih;
        sS[fill + k] = make_double4(x, y, z, q.w);
        sB[fill + k] = make_float4((float)x, (float)y, (float)z, (float)__builtin_fma(x, x, __builtin_fma(y, y, z * z)));
        sW[fill + k] = (float)q.w;
      }
      fill += take;
      ro += take;
      if (ro == r.y) {
        ri++;
        ro = 0;
      }
    }
    const int ns = (fill + 15) & ~15;
    for (int k = fill + tid; k < ns; k += 256) {
      sS[k] = make_double4(MIX_FAR, MIX_FAR, MIX_FAR, 0.0);
      sB[k] = make_float4((float)MIX_FAR, (float)MIX_FAR, (float)MIX_FAR, (float)(3.0 * MIX_FAR * MIX_FAR));
      sW[k] = 0.0f;
    }
    __syncthreads();
    if (wave + 4 < ngroups) mixed_chunks<MASK, 2>(sB, sW, sS, tg, ns >> 4, a, c, acc, lane);
    else if (wave < ngroups) mixed_chunks<MASK, 1>(sB, sW, sS, tg, ns >> 4, a, c, acc, lane);
  }
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int grp = wave + 4 * q;
    if (grp < ngroups) {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        double v = acc[q][e];
        v += __shfl_xor(v, 8, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 1, 64);
        const int t = 16 * grp + 4 * (lane >> 4) + e;
        if ((lane & 15) == 0 && t < it.z) {
          const double u = OOFP_R * ih * v;
          pot[it.y + t] = accum ? u + pot[it.y + t] : u;
        }
      }
    }
  }
}

/* Mixed-precision U-list with the close pairs listed ahead (NP <= 4).
 * Coordinates are centred on the target leaf (between the two leaves of a
 * sibling pair) and scaled by 1 / (tau h) with tau = MIX_CLOSE = 1/8 (a
 * power of two: exact), so a pair is close, r < tau h, exactly when its
 * scaled distance r' is below 1.
 *
 * Distances: one v_mfma_f32_16x16x32_bf16 per 16 x 16 tile on the Gram form
 *     r'^2 = |t|^2 + sum_k A[i][k] B[k][j]
 * with every FP32 operand split exactly into three bf16 parts (v = h + m + l,
 * split3), k = 9c + 3a + b carrying (-2 t_c)_a (s_c)_b for the 3 axes c,
 * k = 27..29 carrying 1 * (|s|^2)_b and k = 30, 31 zero: the FP32 Gram
 * form's accuracy at twice the FP32 MFMA rate (A[i][k] in lane
 * i + 16(k/8) element k%8, B[k][j] in lane j + 16(k/8), D as for the FP32
 * MFMA; pinned by test_mfma_bf16_operand_layout).  LDS holds B as
 * [k/8][source] 16-byte rows, one conflict-free read per lane and chunk.
 *
 * Every pair then adds w * min(1, 1 / r') in FP32, the min and a sign guard
 * being the free clamp and |x| modifiers of v_rsq_f32: no per-pair test.
 * The pairs with r' < 1 in FP64 were found once per plan (k_close_near:
 * per target, a CSR list of source indices); when a target's sum is
 * written, each adds w (1 / r' - 1) in FP64 (w (0 - 1) for r = 0 with
 * MASK), which replaces the clamped w by the FP64 term.  Pairs near r' = 1
 * whose FP32 and FP64 sides disagree are off by the FP32 error of a pair
 * at distance tau h (~1e-4 relative to that pair, as for any FP32 pair). */
constexpr double MIX_CLOSE = CLOSE_FRAC;
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

/* FP32(v) = h + m + l exactly: each remainder is exact in FP32 and the
 * last has at most 8 significant bits.  The empty asm keeps the compiler
 * from rounding v to bf16 straight from FP64 (its f64 -> f32 round-to-odd
 * emulation costs 9 VALU instructions per value, and an h that is not
 * bf16(f) leaves up to 9 bits for l) */
__device__ __forceinline__ void split3(double v, __bf16& h, __bf16& m, __bf16& l) {
  float f = (float)v;
  __asm__("" : "+v"(f));
  h = (__bf16)f;
  const float r1 = f - (float)h;
  m = (__bf16)r1;
  l = (__bf16)(r1 - (float)m);
}

/* B rows of one source at scaled coordinates (x, y, z), |s|^2 = ss */
__device__ __forceinline__ void gram_b_rows(double x, double y, double z, double ss, v8bf (*sB)[P2P_TILE], int s) {
  __bf16 p[4][3];
  split3(x, p[0][0], p[0][1], p[0][2]);
  split3(y, p[1][0], p[1][1], p[1][2]);
  split3(z, p[2][0], p[2][1], p[2][2]);
  split3(ss, p[3][0], p[3][1], p[3][2]);
#pragma unroll
  for (int kb = 0; kb < 4; kb++) {
    v8bf v;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int k = 8 * kb + e;
      v[e] = k < 27 ? p[k / 9][k % 3] : k < 30 ? p[3][k - 27] : (__bf16)0.0f;
    }
    sB[kb][s] = v;
  }
}

/* A row block kb (k = 8kb .. 8kb + 7) of a target at scaled (x, y, z) */
__device__ __forceinline__ v8bf gram_a_block(double x, double y, double z, int kb) {
  __bf16 p[3][3];
  split3(-2.0 * x, p[0][0], p[0][1], p[0][2]);
  split3(-2.0 * y, p[1][0], p[1][1], p[1][2]);
  split3(-2.0 * z, p[2][0], p[2][1], p[2][2]);
  v8bf out;
#pragma unroll
  for (int kc = 0; kc < 4; kc++) {
    v8bf v;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int k = 8 * kc + e;
      v[e] = k < 27 ? p[k / 9][(k % 9) / 3] : k < 30 ? (__bf16)1.0f : (__bf16)0.0f;
    }
    if (kb == kc) out = v;
  }
  return out;
}

/* chunks [ch_lo, ch_hi) of 16 sources of the staged tile (an even count) */
template <int G>
__device__ __forceinline__ void mixed_tau_chunks(const v8bf (*sB)[P2P_TILE], const float* __restrict__ sW,
                                                 int ch_lo, int ch_hi, const v8bf* a, const v4f* c,
                                                 double (*acc)[4], int lane) {
  const int j = lane & 15;
  const v8bf* __restrict__ col = sB[lane >> 4];
  float part[G][4];
#pragma unroll
  for (int q = 0; q < G; q++)
#pragma unroll
    for (int e = 0; e < 4; e++) part[q][e] = 0.0f;
  for (int ch = ch_lo; ch < ch_hi; ch += 2) {
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int s = (ch + u) * 16 + j;
      const v8bf b = col[s];
      const float w = sW[s];
#pragma unroll
      for (int q = 0; q < G; q++) {
        const v4f d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[q], b, c[q], 0, 0, 0);
#pragma unroll
        for (int e = 0; e < 4; e++)
          part[q][e] = __builtin_fmaf(
              w, __builtin_amdgcn_fmed3f(__builtin_amdgcn_rsqf(__builtin_fabsf(d[e])), 0.0f, 1.0f), part[q][e]);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < G; q++)
#pragma unroll
    for (int e = 0; e < 4; e++) acc[q][e] += (double)part[q][e];
}

/* close pairs per target up to which the per-target form sums them */
constexpr int MIX_FEW = 16;

/* FP64 close-pair terms of one 16-target group (slots t0 .. t0 + 15 of sT,
 * nt of them real, the first at target index tg0): sC[t0 + t] = sum over
 * the listed pairs of target t of w (1 / r' - 1), in the scaled frame (g,
 * sc).  A group's pairs are one contiguous run of cidx (targets in order).
 * When no target of the group has more than MIX_FEW of them, four lanes
 * serve each target (pairs s, s + 4, ... of its run on lane 16 s + t) and
 * add their sums by two lane exchanges; otherwise the pairs are spread over
 * the 64 lanes whatever their distribution over the 16 targets and summed
 * per target by a segmented scan over the lanes.  Both orders are fixed:
 * deterministic.  One wave. */
template <bool MASK>
__device__ __forceinline__ void mixed_close_group(const double4* __restrict__ sT, double* __restrict__ sC, int t0,
                                                  int nt, int tg0, const int* __restrict__ cptr,
                                                  const int* __restrict__ cidx, const double4* __restrict__ src4,
                                                  double4 g, double sc, int lane) {
  const int tj = lane & 15;
  const int sj = cptr[tg0 + min(tj, nt - 1)];
  const int ej = cptr[tg0 + min(tj, nt - 1) + 1]; /* end of target lane&15's run */
  int most = tj < nt ? ej - sj : 0;
#pragma unroll
  for (int off = 1; off < 16; off <<= 1) most = max(most, __shfl_xor(most, off, 64));
  if (__builtin_amdgcn_readfirstlane(most) <= MIX_FEW) {
    double v = 0.0;
    if (tj < nt) {
      const double4 tp = sT[t0 + tj];
      for (int k = sj + (lane >> 4); k < ej; k += 4) {
        const double4 sp = src4[cidx[k]];
        const double dx = tp.x - (sp.x - g.x) * sc, dy = tp.y - (sp.y - g.y) * sc, dz = tp.z - (sp.z - g.z) * sc;
        const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
        v += sp.w * (inv_r<MASK>(r2) - 1.0);
      }
    }
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (lane < 16) sC[t0 + lane] = v;
    return;
  }
  if (lane < 16) sC[t0 + lane] = 0.0;
  const int kbeg = cptr[tg0], kend = cptr[tg0 + nt];
  for (int kb = kbeg; kb < kend; kb += 64) {
    const int k = kb + lane;
    const bool live = k < kend;
    int tl = 0; /* the target of pair k: the number of runs that end at or before k */
#pragma unroll
    for (int j = 0; j < 16; j++) tl += __builtin_amdgcn_readlane(ej, j) <= k;
    tl = live ? tl : 16;
    double v = 0.0;
    if (live) {
      const double4 sp = src4[cidx[k]];
      const double4 tp = sT[t0 + tl];
      const double dx = tp.x - (sp.x - g.x) * sc, dy = tp.y - (sp.y - g.y) * sc, dz = tp.z - (sp.z - g.z) * sc;
      const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
      v = sp.w * (inv_r<MASK>(r2) - 1.0);
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const double u = __shfl_up(v, off, 64);
      const int ut = __shfl_up(tl, off, 64);
      if (lane >= off && ut == tl) v += u;
    }
    const int nx = __shfl_down(tl, 1, 64);
    if (live && (lane == 63 || nx != tl)) sC[t0 + tl] += v; /* one writer per target and chunk */
  }
}

/* Groups of 16 targets per wave: items of up to P2P_MAXT5 targets (10
 * groups) give a wave at most 3. */
constexpr int MIX_G = 3;
static_assert(4 * 16 * MIX_G >= P2P_MAXT5, "mixed U-list groups");

/* mixed_tau_chunks over the first nch chunks of the tile for the wave's
 * groups q with bit q of M set (a compile-time set, so that the operand
 * arrays stay in registers) */
template <int M>
__device__ __forceinline__ void mixed_tau_masked(const v8bf (*sB)[P2P_TILE], const float* __restrict__ sW, int nch,
                                                 const v8bf (&a)[MIX_G], const v4f (&c)[MIX_G],
                                                 double (&acc)[MIX_G][4], int lane) {
  const int j = lane & 15;
  const v8bf* __restrict__ col = sB[lane >> 4];
  float part[MIX_G][4];
#pragma unroll
  for (int q = 0; q < MIX_G; q++)
#pragma unroll
    for (int e = 0; e < 4; e++) part[q][e] = 0.0f;
  for (int ch = 0; ch < nch; ch += 2) {
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int s = (ch + u) * 16 + j;
      const v8bf b = col[s];
      const float w = sW[s];
#pragma unroll
      for (int q = 0; q < MIX_G; q++) {
        if (((M >> q) & 1) != 0) {
          const v4f d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[q], b, c[q], 0, 0, 0);
#pragma unroll
          for (int e = 0; e < 4; e++)
            part[q][e] = __builtin_fmaf(
                w, __builtin_amdgcn_fmed3f(__builtin_amdgcn_rsqf(__builtin_fabsf(d[e])), 0.0f, 1.0f), part[q][e]);
        }
      }
    }
  }
  /* every group, the others adding their zero partial sums: the same code
   * in every instantiation (per-group adds under the mask let the compiler
   * merge them across the switch below into stores at a variable index,
   * which puts acc in scratch memory) */
#pragma unroll
  for (int q = 0; q < MIX_G; q++)
#pragma unroll
    for (int e = 0; e < 4; e++) acc[q][e] += (double)part[q][e];
}
__device__ __forceinline__ void mixed_tau_range(const v8bf (*sB)[P2P_TILE], const float* __restrict__ sW, int nch,
                                                int mask, const v8bf (&a)[MIX_G], const v4f (&c)[MIX_G],
                                                double (&acc)[MIX_G][4], int lane) {
  switch (mask) {
    case 1: mixed_tau_masked<1>(sB, sW, nch, a, c, acc, lane); break;
    case 2: mixed_tau_masked<2>(sB, sW, nch, a, c, acc, lane); break;
    case 3: mixed_tau_masked<3>(sB, sW, nch, a, c, acc, lane); break;
    case 4: mixed_tau_masked<4>(sB, sW, nch, a, c, acc, lane); break;
    case 6: mixed_tau_masked<6>(sB, sW, nch, a, c, acc, lane); break;
    case 7: mixed_tau_masked<7>(sB, sW, nch, a, c, acc, lane); break;
    default: break; /* 0: the wave has no group in this segment */
  }
}
template <bool MASK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_p2p_ulist_mixed_tau(const int4* __restrict__ items, int n_items, int per_xcd,
                                                             const int* __restrict__ rng_ptr,
                                                             const int2* __restrict__ rng,
                                                             const double4* __restrict__ geo,
                                                             const double4* __restrict__ src4,
                                                             const double4* __restrict__ trg4,
                                                             const int* __restrict__ cptr,
                                                             const int* __restrict__ cidx,
                                                             double* __restrict__ pot, int accum) {
  __shared__ v8bf sB[4][P2P_TILE];
  __shared__ float sW[P2P_TILE];
  __shared__ double4 sT[P2P_MAXT5];
  const int item = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (item >= n_items) return; /* whole block, before any barrier */
  const int4 it = items[item];
  /* a leaf (nB = 0) or two sibling leaves A = it.w and B = it.w + 1 whose
   * targets follow each other: groups 0 .. gA - 1 are A's, the rest B's,
   * and the frame is centred between the two leaves */
  const int nA = it.z & 0xffff, nB = it.z >> 16;
  const int gA = (nA + 15) >> 4, ngroups = gA + ((nB + 15) >> 4);
  double4 g = geo[it.w];
  if (nB > 0) {
    const double4 g2 = geo[it.w + 1];
    g.x = 0.5 * (g.x + g2.x);
    g.y = 0.5 * (g.y + g2.y);
    g.z = 0.5 * (g.z + g2.z);
  }
  const double sc = (1.0 / MIX_CLOSE) / g.w; /* 1 / (tau h), a power of two */
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = (__builtin_amdgcn_readfirstlane(tid >> 6) + item) & 3; /* rotated as in p2p_item */
  /* scaled targets; slots past a leaf's count repeat its last target */
  for (int k = tid; k < ngroups * 16; k += 256) {
    const int tix = k < 16 * gA ? min(k, nA - 1) : nA + min(k - 16 * gA, nB - 1);
    const double4 q = trg4[it.y + tix];
    const double x = (q.x - g.x) * sc, y = (q.y - g.y) * sc, z = (q.z - g.z) * sc;
    sT[k] = make_double4(x, y, z, __builtin_fma(x, x, __builtin_fma(y, y, z * z)));
  }
  __syncthreads();
  v8bf a[MIX_G];
  v4f c[MIX_G];
  double acc[MIX_G][4];
#pragma unroll
  for (int q = 0; q < MIX_G; q++) {
#pragma unroll
    for (int e = 0; e < 4; e++) acc[q][e] = 0.0;
    a[q] = v8bf{};
    c[q] = v4f{};
    /* operands only for the groups this wave owns (wave-uniform): a wave of
     * a <= 64-target item has one, and the bf16 splitting of an A block is
     * about 90 VALU instructions per lane */
    if (wave + 4 * q < ngroups) {
      const int grp = wave + 4 * q;
      const double4 t = sT[16 * grp + (lane & 15)];
      const int k4 = lane >> 4;
      a[q] = gram_a_block(t.x, t.y, t.z, k4);
#pragma unroll
      for (int e = 0; e < 4; e++) c[q][e] = (float)sT[16 * grp + 4 * k4 + e].w;
    }
  }
  /* the wave's groups wave + 4q: q < qa are A's, qa <= q < qn B's */
  const int qn = wave < ngroups ? (ngroups - wave + 3) >> 2 : 0;
  const int qa = wave < gA ? (gA - wave + 3) >> 2 : 0;
  /* source segments, each its own run of sets in rng_ptr and its own tiles:
   * a leaf's U-list; or for a pair the boxes in both U-lists (every group),
   * those only in A's (A's groups) and those only in B's (B's groups) */
  const int nseg = nB > 0 ? 3 : 1;
  int seg = 0;
  RangeWalk walk{rng_ptr[it.x], 0, rng_ptr[it.x + 1]};
  double4 pre[P2P_PER];
  auto next_tile = [&](int& tseg) {
    int f = walk_tile(walk, rng, src4, tid, pre);
    while (f == 0 && seg + 1 < nseg) {
      seg++;
      walk = RangeWalk{rng_ptr[it.x + seg], 0, rng_ptr[it.x + seg + 1]};
      f = walk_tile(walk, rng, src4, tid, pre);
    }
    tseg = seg;
    return f;
  };
  int tseg = 0;
  int fill = next_tile(tseg);
  while (fill > 0) {
    __syncthreads(); /* the previous tile's readers are done */
#pragma unroll
    for (int m = 0; m < P2P_PER; m++) {
      const int k = tid + 256 * m;
      if (k < fill) {
        const double4 q = pre[m];
        const double x = (q.x - g.x) * sc, y = (q.y - g.y) * sc, z = (q.z - g.z) * sc;
        gram_b_rows(x, y, z, __builtin_fma(x, x, __builtin_fma(y, y, z * z)), sB, k);
        sW[k] = (float)q.w;
      }
    }
    const int ns = (fill + 31) & ~31; /* whole chunk pairs (P2P_TILE is a multiple of 32) */
    for (int k = fill + tid; k < ns; k += 256) {
      gram_b_rows(MIX_FAR, MIX_FAR, MIX_FAR, 3.0 * MIX_FAR * MIX_FAR, sB, k);
      sW[k] = 0.0f;
    }
    __syncthreads();
    int nseg_next = 0;
    const int next = next_tile(nseg_next); /* loads in flight during the tile */
    const int q0 = tseg == 2 ? qa : 0, q1 = tseg == 1 ? qa : qn; /* wave-uniform */
    mixed_tau_range(sB, sW, ns >> 4, ((1 << q1) - 1) & ~((1 << q0) - 1), a, c, acc, lane);
    fill = next;
    tseg = nseg_next;
  }
  /* FP64 close-pair terms of the wave's target groups.  The per-target sums
   * reuse the source tile's LDS once every wave is done with it (keeps the
   * block within 40 KB: 4 blocks per CU). */
  static_assert(sizeof(sB) >= P2P_MAXT5 * sizeof(double), "close-pair sums in the tile");
  __syncthreads();
  double* const sC = reinterpret_cast<double*>(&sB[0][0]);
  /* group grp's first target (offset from it.y) and target count */
  auto grp_first = [&](int grp) { return grp < gA ? 16 * grp : nA + 16 * (grp - gA); };
  auto grp_count = [&](int grp) { return grp < gA ? min(16, nA - 16 * grp) : min(16, nB - 16 * (grp - gA)); };
#pragma unroll
  for (int q = 0; q < MIX_G; q++) {
    const int grp = wave + 4 * q;
    if (grp < ngroups)
      mixed_close_group<MASK>(sT, sC, 16 * grp, grp_count(grp), it.y + grp_first(grp), cptr, cidx, src4, g, sc,
                              lane);
  }
  const double scale = OOFP_R * sc; /* 1/(4 pi) sum w / r = 1/(4 pi tau h) sum w / r' */
#pragma unroll
  for (int q = 0; q < MIX_G; q++) {
    const int grp = wave + 4 * q;
    if (grp < ngroups) {
      const int first = grp_first(grp), cnt = grp_count(grp);
#pragma unroll
      for (int e = 0; e < 4; e++) {
        double v = acc[q][e];
        v += __shfl_xor(v, 8, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 1, 64);
        const int r = 4 * (lane >> 4) + e; /* the target's row in the group */
        if ((lane & 15) == 0 && r < cnt) {
          const double u = scale * (v + sC[16 * grp + r]);
          double* const o = pot + it.y + first + r;
          *o = accum ? u + *o : u;
        }
      }
    }
  }
}

/* Mixed U-list, source-split form (items of <= P2P_MAXT targets).  The
 * arithmetic is the tau kernel's (bf16 Gram MFMA distances, clamped FP32
 * rsq, listed FP64 close-pair terms); the work is divided differently.  The
 * item's 16-target groups form sets of up to MIXS_G = 2 groups, and the
 * waves of a set split every staged source tile between them (contiguous
 * chunk ranges), each running all of the set's groups: a 64-target leaf is
 * two sets of two waves, a 16-target one a set of four.  So a wave reads
 * each B operand once for up to 2 MFMAs, where the tau kernel gives a wave
 * of a 64-target leaf a single group.  (Sets of 4 groups, one per 64
 * targets, measured 40 % slower at 64 targets per leaf: 3 waves per SIMD
 * for their registers.)  The partial sums of a set's waves are added per
 * target in wave order through LDS, then the close-pair terms.  Sums differ
 * from the tau kernel's at rounding level (other partial sums), and are
 * bit-identical from run to run. */
constexpr int MIXS_G = 2;
template <bool MASK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_p2p_ulist_mixed_split(
    const int4* __restrict__ items, int n_items, int per_xcd, const int* __restrict__ rng_ptr,
    const int2* __restrict__ rng, const double4* __restrict__ geo, const double4* __restrict__ src4,
    const double4* __restrict__ trg4, const int* __restrict__ cptr, const int* __restrict__ cidx,
    double* __restrict__ pot, int accum) {
  __shared__ v8bf sB[4][P2P_TILE];
  __shared__ float sW[P2P_TILE];
  __shared__ double4 sT[P2P_MAXT];
  const int item = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (item >= n_items) return; /* whole block, before any barrier */
  const int4 it = items[item];
  const double4 g = geo[it.w];
  const double sc = (1.0 / MIX_CLOSE) / g.w; /* 1 / (tau h), a power of two */
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ngroups = (it.z + 15) >> 4;                   /* 1 .. 8 */
  const int nsets = (ngroups + MIXS_G - 1) / MIXS_G;      /* 1 .. 4 */
  const int wps = nsets == 1 ? 4 : nsets == 2 ? 2 : 1;    /* waves per set: the source split */
  const int set = wave / wps, part = wave % wps;
  const int g0 = MIXS_G * set, gn = min(MIXS_G, ngroups - g0); /* gn < 1: a wave with no set (3 sets) */
  for (int k = tid; k < ngroups * 16; k += 256) {
    const double4 q = trg4[it.y + min(k, it.z - 1)];
    const double x = (q.x - g.x) * sc, y = (q.y - g.y) * sc, z = (q.z - g.z) * sc;
    sT[k] = make_double4(x, y, z, __builtin_fma(x, x, __builtin_fma(y, y, z * z)));
  }
  __syncthreads();
  v8bf a[MIXS_G];
  v4f c[MIXS_G];
  double acc[MIXS_G][4];
#pragma unroll
  for (int q = 0; q < MIXS_G; q++) {
    const int grp = min(g0 + q, ngroups - 1);
    const double4 t = sT[16 * grp + (lane & 15)];
    const int k4 = lane >> 4;
    a[q] = gram_a_block(t.x, t.y, t.z, k4);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      c[q][e] = (float)sT[16 * grp + 4 * k4 + e].w;
      acc[q][e] = 0.0;
    }
  }
  RangeWalk walk{rng_ptr[it.x], 0, rng_ptr[it.x + 1]};
  double4 pre[P2P_PER];
  int fill = walk_tile(walk, rng, src4, tid, pre);
  while (fill > 0) {
    __syncthreads(); /* the previous tile's readers are done */
#pragma unroll
    for (int m = 0; m < P2P_PER; m++) {
      const int k = tid + 256 * m;
      if (k < fill) {
        const double4 q = pre[m];
        const double x = (q.x - g.x) * sc, y = (q.y - g.y) * sc, z = (q.z - g.z) * sc;
        gram_b_rows(x, y, z, __builtin_fma(x, x, __builtin_fma(y, y, z * z)), sB, k);
        sW[k] = (float)q.w;
      }
    }
    const int ns = (fill + 31) & ~31; /* whole chunk pairs */
    for (int k = fill + tid; k < ns; k += 256) {
      gram_b_rows(MIX_FAR, MIX_FAR, MIX_FAR, 3.0 * MIX_FAR * MIX_FAR, sB, k);
      sW[k] = 0.0f;
    }
    __syncthreads();
    const int next = walk_tile(walk, rng, src4, tid, pre); /* loads in flight during the tile */
    const int npairs = ns >> 5; /* chunk pairs of the tile, split among the set's waves */
    const int lo = 2 * (npairs * part / wps), hi = 2 * (npairs * (part + 1) / wps);
    if (gn == 2) mixed_tau_chunks<2>(sB, sW, lo, hi, a, c, acc, lane);
    else if (gn == 1) mixed_tau_chunks<1>(sB, sW, lo, hi, a, c, acc, lane);
    fill = next;
  }
  /* the tile's LDS now holds the per-wave partial sums red[wave][target] and
   * the close-pair sums sC[target] */
  static_assert(sizeof(sB) >= (4 + 1) * P2P_MAXT * sizeof(double), "partial sums in the tile");
  __syncthreads();
  double* const red = reinterpret_cast<double*>(&sB[0][0]);
  double* const sC = red + 4 * P2P_MAXT;
#pragma unroll
  for (int q = 0; q < MIXS_G; q++) {
    if (q < gn) {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        double v = acc[q][e];
        v += __shfl_xor(v, 8, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 1, 64);
        const int t = 16 * (g0 + q) + 4 * (lane >> 4) + e;
        if ((lane & 15) == 0) red[part * P2P_MAXT + t] = v;
      }
    }
  }
  /* FP64 close-pair terms, group grp by wave grp % 4 (as in the tau kernel) */
  for (int grp = wave; grp < ngroups; grp += 4)
    mixed_close_group<MASK>(sT, sC, 16 * grp, min(16, it.z - 16 * grp), it.y + 16 * grp, cptr, cidx, src4, g, sc,
                            lane);
  __syncthreads();
  const double scale = OOFP_R * sc;
  for (int t = tid; t < it.z; t += 256) {
    double v = red[t]; /* the target's set: its waves' partial sums in wave order */
    for (int p_ = 1; p_ < wps; p_++) v += red[p_ * P2P_MAXT + t];
    const double u = scale * (v + sC[t]);
    pot[it.y + t] = accum ? u + pot[it.y + t] : u;
  }
}

/* The close pairs of the U-list (geometry only, found once per plan): for
 * each target of an item, the sources of its leaf's U-list boxes at FP64
 * scaled distance r' < 1 (r < MIX_CLOSE h, h the target leaf's half-width),
 * in U-list order and, inside a box, in source order: the order of the
 * U-list kernels' range walk.  One wave per item (<= P2P_MAXT5 targets).
 * Only boxes that some target lies within r' < 1 (+ margin) of are read:
 * for a target inside its own leaf that is the leaf itself, plus the face
 * neighbours of the targets within h/8 of a face and rarely an edge or
 * corner neighbour, so a uniform leaf of q points costs about 1.3 q^2 pair
 * tests instead of the 27 q^2 of the whole U-list.  Each candidate pair is
 * classified by exactly the expression of the all-pairs search it replaces
 * (scaled differences, then fma(dz, dz, fma(dy, dy, dx * dx)) < 1), so the
 * list is the same.  FILL = false writes the count per target; FILL = true
 * writes the source indices from the exclusive offsets cptr.  Items in the
 * U-list kernels' XCD-contiguous order, so neighbouring leaves share an L2. */
constexpr int CLOSE_TC = (P2P_MAXT5 + 63) / 64; /* target chunks of 64 per item */
template <bool FILL>
__global__ __launch_bounds__(64) void k_close_near(const int4* __restrict__ items, int n_items, int per_xcd,
                                                   const int* __restrict__ u_ptr, const int* __restrict__ u_idx,
                                                   const int* __restrict__ sbeg, const int* __restrict__ snum,
                                                   const double4* __restrict__ geo,
                                                   const double4* __restrict__ src4,
                                                   const double4* __restrict__ trg4, int* __restrict__ cnt,
                                                   const int* __restrict__ cptr, int* __restrict__ cidx) {
  __shared__ double4 sT[64 * CLOSE_TC];
  __shared__ int run[64 * CLOSE_TC];
  const int item = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (item >= n_items) return; /* whole block (one wave) */
  const int4 it = items[item];
  const int lane = threadIdx.x;
  const double4 g = geo[it.w];
  const double sc = (1.0 / MIX_CLOSE) / g.w; /* a power of two: exact */
#pragma unroll
  for (int c = 0; c < CLOSE_TC; c++) {
    const int k = lane + 64 * c;
    double4 t = make_double4(0.0, 0.0, 0.0, 0.0);
    if (k < it.z) {
      const double4 q = trg4[it.y + k];
      t = make_double4((q.x - g.x) * sc, (q.y - g.y) * sc, (q.z - g.z) * sc, 0.0);
    }
    sT[k] = t;
    run[k] = (FILL && k < it.z) ? cptr[it.y + k] : 0;
  }
  __syncthreads();
  for (int u = u_ptr[it.w]; u < u_ptr[it.w + 1]; u++) {
    const int s_node = u_idx[u];
    const int ns = snum[s_node];
    if (ns <= 0) continue;
    const int sb = sbeg[s_node];
    const double4 gs = geo[s_node];
    /* the box of s_node in the item's scaled coordinates */
    const double cx = (gs.x - g.x) * sc, cy = (gs.y - g.y) * sc, cz = (gs.z - g.z) * sc, hw = gs.w * sc;
    unsigned long long near[CLOSE_TC];
    unsigned long long any = 0;
#pragma unroll
    for (int c = 0; c < CLOSE_TC; c++) {
      const int k = lane + 64 * c;
      const double4 t = sT[k];
      const double ex = __builtin_fmax(__builtin_fabs(t.x - cx) - hw, 0.0);
      const double ey = __builtin_fmax(__builtin_fabs(t.y - cy) - hw, 0.0);
      const double ez = __builtin_fmax(__builtin_fabs(t.z - cz) - hw, 0.0);
      /* a point of the box lies at least this far from the target; the
       * margin covers the rounding of the scaled coordinates */
      near[c] = __ballot(k < it.z && ex * ex + ey * ey + ez * ez < 1.0 + 1e-6);
      any |= near[c];
    }
    if (!any) continue;
    for (int base = 0; base < ns; base += 64) {
      const bool sl = base + lane < ns;
      const int si = sb + base + lane;
      double sx = 1e30, sy = 1e30, sz = 1e30;
      if (sl) {
        const double4 q = src4[si];
        sx = (q.x - g.x) * sc; sy = (q.y - g.y) * sc; sz = (q.z - g.z) * sc;
      }
#pragma unroll
      for (int c = 0; c < CLOSE_TC; c++) {
        unsigned long long m = near[c];
        while (m) {
          const int b = __builtin_ctzll(m);
          m &= m - 1;
          const int k = 64 * c + b;
          const double4 t = sT[k];
          const double dx = t.x - sx, dy = t.y - sy, dz = t.z - sz;
          const bool close = sl && __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx)) < 1.0;
          const unsigned long long cm = __ballot(close);
          if (cm) {
            const int r = run[k];
            if (FILL && close) {
              const int pos = r + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(cm >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((unsigned)cm, 0u));
              if (pos < cptr[it.y + k + 1]) cidx[pos] = si; /* the count pass sized it */
            }
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) run[k] = r + __popcll(cm);
            __builtin_amdgcn_wave_barrier();
          }
        }
      }
    }
  }
  __syncthreads();
  if (!FILL)
#pragma unroll
    for (int c = 0; c < CLOSE_TC; c++) {
      const int k = lane + 64 * c;
      if (k < it.z) cnt[it.y + k] = run[k];
    }
}

/* Exclusive prefix sums of n counts into out[0..n] (out[n] = the total), in
 * a fixed order: k_scan_blocks sums SCAN_PER consecutive counts per block
 * (int64), k_scan_tops scans those block sums in one block, k_scan_out writes
 * each block's exclusive sums from its offset.  n <= SCAN_PER * SCAN_PER. */
constexpr int SCAN_T = 1024, SCAN_V = 16, SCAN_PER = SCAN_T * SCAN_V;
__device__ __forceinline__ long long block_exclusive(long long v, long long* sh, long long* total) {
  /* inclusive scan over the block by Hillis-Steele in LDS, then shift */
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_T; off <<= 1) {
    const long long u = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += u;
    __syncthreads();
  }
  const long long inc = sh[t];
  *total = sh[SCAN_T - 1];
  __syncthreads();
  return inc - v;
}
static __global__ __launch_bounds__(SCAN_T) void k_scan_blocks(const int* __restrict__ in, long long n,
                                                               long long* __restrict__ bsum) {
  __shared__ long long sh[SCAN_T];
  const long long b0 = (long long)blockIdx.x * SCAN_PER + (long long)threadIdx.x * SCAN_V;
  long long s = 0;
  for (int k = 0; k < SCAN_V; k++)
    if (b0 + k < n) s += in[b0 + k];
  long long tot;
  block_exclusive(s, sh, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}
static __global__ __launch_bounds__(SCAN_T) void k_scan_tops(long long* __restrict__ bsum, int nb,
                                                             long long* __restrict__ total) {
  __shared__ long long sh[SCAN_T];
  const int b0 = threadIdx.x * SCAN_V;
  long long v[SCAN_V], s = 0;
  for (int k = 0; k < SCAN_V; k++) {
    v[k] = b0 + k < nb ? bsum[b0 + k] : 0;
    s += v[k];
  }
  long long tot;
  long long run = block_exclusive(s, sh, &tot);
  for (int k = 0; k < SCAN_V; k++)
    if (b0 + k < nb) {
      bsum[b0 + k] = run;
      run += v[k];
    }
  if (threadIdx.x == 0) *total = tot;
}
/* out is written as int: the caller checks the total against its cap first */
static __global__ __launch_bounds__(SCAN_T) void k_scan_out(const int* __restrict__ in, long long n,
                                                            const long long* __restrict__ boff,
                                                            const long long* __restrict__ total,
                                                            int* __restrict__ out) {
  __shared__ long long sh[SCAN_T];
  const long long b0 = (long long)blockIdx.x * SCAN_PER + (long long)threadIdx.x * SCAN_V;
  long long s = 0;
  for (int k = 0; k < SCAN_V; k++)
    if (b0 + k < n) s += in[b0 + k];
  long long tot;
  long long run = boff[blockIdx.x] + block_exclusive(s, sh, &tot);
  for (int k = 0; k < SCAN_V; k++)
    if (b0 + k < n) {
      out[b0 + k] = (int)run;
      run += in[b0 + k];
    }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = (int)*total;
}

/* ------------------------------------------------------------------ S2M */

/* S2M check potentials (up_eval__gpu, fmm/cuda.cu:108-190 first half): one
 * wave per source leaf.  In leaf units (centre c, radius h of the node:
 * s' = (s - c) / h, a power-of-two scale, exact) the check surface is the
 * same table for every leaf (rows x, y, z, |p|^2 of the radius-3 box), and
 * the leaf's sources, staged S2M_TILE at a time in LDS as (s', |s'|^2) and
 * w, are read as broadcasts.  The wave is G phases of 64 / G lanes: phase g
 * takes the sources k with k mod G = g, and a lane holds P check points (i,
 * i + 64 / G, ...), so one LDS read of a source (40 bytes per lane) feeds P
 * pairs; the phases' sums meet in a lane shuffle at the end.  (G, P) =
 * (4, 7), (2, 5), (1, 5) for 98, 152, 296 check points (NP = 3, 4, 6):
 * 88-95 % of the lane slots busy.  Per pair the Gram form
 *     r'^2 = (|p|^2 + |s'|^2) + (-2p) . s'        (1 add, 3 FMA)
 * then the rsq, the Newton step and the FMA: 9 VALU instructions against
 * 11 for coordinate differences.  The check points lie at least 2 leaf
 * units from the leaf's points (|p| <= 3 sqrt(3), |s'| <= sqrt(3)), so the
 * Gram r'^2 is within a few eps (|p| + |s'|)^2 / 4 relative: ~1e-15. */
constexpr int S2M_TILE = 128;
template <int G, int P>
static __global__ __launch_bounds__(64) void k_s2m(const int4* __restrict__ items, /* {node, src_beg, src_num, col} */
                                                   const double4* __restrict__ geo, /* node centre xyz + radius */
                                                   const double* __restrict__ surf_uc, /* 4 x ld, radius-3 box */
                                                   int n_uc, int ld_uc_surf,
                                                   const double4* __restrict__ src4,
                                                   double* __restrict__ chk, int ld_chk) {
  constexpr int L = 64 / G;
  static_assert(S2M_TILE % G == 0, "whole phases per tile");
  __shared__ double4 sS[S2M_TILE];
  __shared__ double sW[S2M_TILE];
  const int4 it = items[blockIdx.x];
  const int lane = threadIdx.x, g = lane / L, i = lane % L;
  const double4 c = geo[it.x];
  const double ih = 1.0 / c.w;
  double ax[P], ay[P], az[P], pp[P], acc[P];
#pragma unroll
  for (int p = 0; p < P; p++) {
    const int j = min(i + L * p, n_uc - 1);
    ax[p] = -2.0 * surf_uc[j];
    ay[p] = -2.0 * surf_uc[ld_uc_surf + j];
    az[p] = -2.0 * surf_uc[2 * ld_uc_surf + j];
    pp[p] = surf_uc[3 * ld_uc_surf + j];
    acc[p] = 0.0;
  }
  for (int base = 0; base < it.z; base += S2M_TILE) {
    const int n = min(S2M_TILE, it.z - base);
    const int nG = (n + G - 1) / G * G; /* padded with weight-0 copies of the tile's last source */
    __syncthreads(); /* the previous tile's readers are done */
    for (int k = lane; k < nG; k += 64) {
      const double4 q = src4[it.y + base + min(k, n - 1)];
      const double x = (q.x - c.x) * ih, y = (q.y - c.y) * ih, z = (q.z - c.z) * ih;
      sS[k] = make_double4(x, y, z, __builtin_fma(x, x, __builtin_fma(y, y, z * z)));
      sW[k] = k < n ? q.w : 0.0;
    }
    __syncthreads();
#pragma unroll 2
    for (int k = g; k < nG; k += G) {
      const double4 q = sS[k];
      const double w = sW[k];
#pragma unroll
      for (int p = 0; p < P; p++) {
        const double r2 = __builtin_fma(ax[p], q.x, __builtin_fma(ay[p], q.y, __builtin_fma(az[p], q.z, pp[p] + q.w)));
        acc[p] = __builtin_fma(w, rsqrt2_fp64(r2), acc[p]);
      }
    }
  }
#pragma unroll
  for (int p = 0; p < P; p++)
    for (int m = L; m < 64; m *= 2) acc[p] += __shfl_xor(acc[p], m, 64);
  if (g == 0)
#pragma unroll
    for (int p = 0; p < P; p++)
      if (i + L * p < n_uc) chk[(size_t)it.w * ld_chk + i + L * p] = (HALF_OOFP_R * ih) * acc[p];
}

/* Mixed-precision S2M and L2T (NP <= 4): the same lane-per-target scheme
 * in FP32, in units of the leaf's half-width h centred on the leaf (the
 * check and equivalent surfaces are unit-box surfaces there, shared by all
 * leaves, rows x, y, z, |p|^2), summed in FP32 per lane and scaled by 1/h
 * into the FP64 check values or potentials.  Per pair the Gram form
 *     r^2 = (|p|^2 + |s|^2) + (-2p) . s          (1 add, 3 FMA)
 * then v_rsq_f32 and the FMA with the weight: 5 VALU instructions and the
 * rsq, against 7 for coordinate differences.  The surfaces lie at least 2
 * leaf units (max norm) from the leaf's points, |p| <= 3 sqrt(3) and |s| <=
 * sqrt(3), so the Gram r^2 carries at most ~12 FP32 roundings of r^2, ~1e-6
 * relative in 1 / r (the FP32 sum over the leaf's points is of the same
 * order); the pinv applies that follow stay FP64. */
/* FP32 leaf-local source records for the mixed S2M: (s - c) / h and w for
 * every source of every source leaf (one block per leaf). */
static __global__ __launch_bounds__(256) void k_src_local32(const int4* __restrict__ items,
                                                            const double4* __restrict__ geo,
                                                            const double4* __restrict__ src4,
                                                            float4* __restrict__ src32) {
  const int4 it = items[blockIdx.x];
  const double4 g = geo[it.x];
  const double ih = 1.0 / g.w;
  for (int k = threadIdx.x; k < it.z; k += 256) {
    const double4 s = src4[it.y + k];
    src32[it.y + k] = make_float4((float)((s.x - g.x) * ih), (float)((s.y - g.y) * ih), (float)((s.z - g.z) * ih),
                                  (float)s.w);
  }
}

/* LPB consecutive leaves per wave in the mixed S2M and L2T (1 for leaves
 * of ~100 points and more; 4 for small leaves, whose per-block cost and
 * surface loads are then shared).  S2M: the FP64 kernel's wave layout (G
 * source phases, P check points per lane) in FP32; the leaf's sources are
 * staged in LDS as (s, |s|^2) and w, read as broadcasts, and the phases'
 * FP32 sums meet in FP64. */
constexpr int S2M_TILE32 = 256;
template <int LPB, int G, int P>
__global__ __launch_bounds__(64) void k_s2m_mixed(const int4* __restrict__ items, int n_items,
                                                  const double4* __restrict__ geo,
                                                  const float* __restrict__ surf_uc32, int n_uc,
                                                  int ld_uc_surf, const float4* __restrict__ src32,
                                                  double* __restrict__ chk, int ld_chk) {
  constexpr int L = 64 / G;
  static_assert(S2M_TILE32 % G == 0, "whole phases per tile");
  __shared__ float4 sS[S2M_TILE32];
  __shared__ float sW[S2M_TILE32];
  const int lane = threadIdx.x, g = lane / L, i = lane % L;
  float ax[P], ay[P], az[P], pp[P];
#pragma unroll
  for (int p = 0; p < P; p++) {
    const int j = min(i + L * p, n_uc - 1);
    ax[p] = -2.0f * surf_uc32[j];
    ay[p] = -2.0f * surf_uc32[ld_uc_surf + j];
    az[p] = -2.0f * surf_uc32[2 * ld_uc_surf + j];
    pp[p] = surf_uc32[3 * ld_uc_surf + j];
  }
  for (int r = 0; r < LPB; r++) {
    const int item = blockIdx.x * LPB + r;
    if (item >= n_items) return; /* the same for every lane */
    const int4 it = items[item];
    const double ih = 1.0 / geo[it.x].w;
    float acc[P];
#pragma unroll
    for (int p = 0; p < P; p++) acc[p] = 0.0f;
    for (int base = 0; base < it.z; base += S2M_TILE32) {
      const int n = min(S2M_TILE32, it.z - base);
      const int nG = (n + G - 1) / G * G; /* padded with weight-0 copies of the tile's last source */
      __syncthreads(); /* the previous tile's (or leaf's) readers are done */
      for (int k = lane; k < nG; k += 64) {
        const float4 q = src32[it.y + base + min(k, n - 1)];
        sS[k] = make_float4(q.x, q.y, q.z, __builtin_fmaf(q.x, q.x, __builtin_fmaf(q.y, q.y, q.z * q.z)));
        sW[k] = k < n ? q.w : 0.0f;
      }
      __syncthreads();
#pragma unroll 2
      for (int k = g; k < nG; k += G) {
        const float4 q = sS[k];
        const float w = sW[k];
#pragma unroll
        for (int p = 0; p < P; p++) {
          const float r2 = __builtin_fmaf(ax[p], q.x, __builtin_fmaf(ay[p], q.y, __builtin_fmaf(az[p], q.z, pp[p] + q.w)));
          acc[p] = __builtin_fmaf(w, __builtin_amdgcn_rsqf(r2), acc[p]);
        }
      }
    }
#pragma unroll
    for (int p = 0; p < P; p++) {
      double t = (double)acc[p];
      for (int m = L; m < 64; m *= 2) t += __shfl_xor(t, m, 64);
      if (g == 0 && i + L * p < n_uc) chk[(size_t)it.w * ld_chk + i + L * p] = OOFP_R * ih * t;
    }
  }
}

/* L2T: lane = target of an item (<= 64 targets of one leaf); the DE surface
 * (rows x, y, z, |p|^2) is read as wave-uniform values, the leaf's
 * densities once per item into LDS as FP32. */
template <int LPB>
__global__ __launch_bounds__(64) void k_l2t_mixed(const int4* __restrict__ items, int n_items,
                                                         const double4* __restrict__ geo,
                                                         const float* __restrict__ surf_de, int ld_surf, int n_e,
                                                         const double* __restrict__ dden, int ld_e,
                                                         const double4* __restrict__ trg4, double* __restrict__ pot) {
  __shared__ float sD[64];
  for (int r = 0; r < LPB; r++) {
    const int item = blockIdx.x * LPB + r;
    if (item >= n_items) return; /* the same for the whole (one-wave) block */
    const int4 it = items[item];
    const int k = threadIdx.x;
    const double4 q = trg4[it.y + min(k, it.z - 1)];
    const double4 g = geo[it.x];
    const double ih = 1.0 / g.w;
    const float tx = (float)((q.x - g.x) * ih), ty = (float)((q.y - g.y) * ih), tz = (float)((q.z - g.z) * ih);
    const float ax = -2.0f * tx, ay = -2.0f * ty, az = -2.0f * tz;
    const float tt = __builtin_fmaf(tx, tx, __builtin_fmaf(ty, ty, tz * tz));
    __builtin_amdgcn_wave_barrier(); /* the previous item's reads of sD are done (one wave) */
    if (k < n_e) sD[k] = (float)dden[(size_t)it.x * ld_e + k];
    __builtin_amdgcn_wave_barrier();
    float acc0 = 0.0f, acc1 = 0.0f;
    int j = 0;
    for (; j + 2 <= n_e; j += 2) {
      const float r0 = __builtin_fmaf(surf_de[j], ax, __builtin_fmaf(surf_de[ld_surf + j], ay,
                                      __builtin_fmaf(surf_de[2 * ld_surf + j], az, tt + surf_de[3 * ld_surf + j])));
      const float r1 = __builtin_fmaf(surf_de[j + 1], ax, __builtin_fmaf(surf_de[ld_surf + j + 1], ay,
                                      __builtin_fmaf(surf_de[2 * ld_surf + j + 1], az, tt + surf_de[3 * ld_surf + j + 1])));
      acc0 = __builtin_fmaf(sD[j], __builtin_amdgcn_rsqf(r0), acc0);
      acc1 = __builtin_fmaf(sD[j + 1], __builtin_amdgcn_rsqf(r1), acc1);
    }
    if (j < n_e) {
      const float r0 = __builtin_fmaf(surf_de[j], ax, __builtin_fmaf(surf_de[ld_surf + j], ay,
                                      __builtin_fmaf(surf_de[2 * ld_surf + j], az, tt + surf_de[3 * ld_surf + j])));
      acc0 = __builtin_fmaf(sD[j], __builtin_amdgcn_rsqf(r0), acc0);
    }
    if (k < it.z) pot[it.y + k] += OOFP_R * ih * ((double)acc0 + (double)acc1);
  }
}

/* ------------------------------------------------------------ GEMM (MFMA) */

/* out[:, j] = (acc ? out[:, j] : 0) + 2^sexp[j] * A[M x K] * in[:, j]
 * A column-major with lda rows (zero-padded to a multiple of 16) and K a
 * multiple of 4 (zero-padded columns).  Column j reads K doubles at
 * in + in_off[j] and writes M doubles at out + out_off[j].
 * A block of four waves owns 16*RT rows (row group blockIdx.y) and 64*CT
 * columns; wave w owns column tiles w*CT .. w*CT+CT-1 and runs all RT row
 * tiles against them (RT*CT v_mfma_f64_16x16x4f64 accumulators).  K is
 * walked in chunks of GEMM_KC staged in LDS; the next chunk is loaded into
 * registers while the current one is multiplied.  Operand layout (measured
 * on gfx950 by test_mfma_f64_operand_layout): A[i][k] in lane i + 16k,
 * B[k][j] in lane j + 16k, D[i][j] in lane j + 16(i%4), element i/4.
 * With ws != nullptr the unscaled partial product of K-slice blockIdx.z
 * goes to ws[(z * ncols + j) * ldw + row], which k_gemm_reduce then adds in
 * slice order. */
constexpr int GEMM_KC = 16;
template <int RT, int CT>
__global__ __launch_bounds__(256) void k_gemm_cols(int M, int K, const double* __restrict__ A, int lda,
                                                   int ncols, const int* __restrict__ in_off,
                                                   const int* __restrict__ out_off,
                                                   const int* __restrict__ sexp,
                                                   const double* __restrict__ in,
                                                   double* __restrict__ out, int accumulate,
                                                   int ksplit, double* __restrict__ ws, int ldw) {
  constexpr int ROWS = 16 * RT, NC = 64 * CT;
  /* row strides chosen so the 4 k-rows a wave reads hit disjoint bank halves */
  constexpr int LA = ROWS + (((16 - ROWS) % 32) + 32) % 32;
  constexpr int LB = NC + 16;
  constexpr int AE = GEMM_KC * ROWS / 256, BE = GEMM_KC * NC / 256;
  static_assert(GEMM_KC * ROWS % 256 == 0 && GEMM_KC * NC % 256 == 0, "chunk split");
  __shared__ double As[GEMM_KC][LA];
  __shared__ double Bs[GEMM_KC][LB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int col0 = blockIdx.x * NC, row0 = blockIdx.y * ROWS;
  const int kbeg = blockIdx.z * ksplit, kend = min(K, kbeg + ksplit);
  v4d acc[RT][CT];
#pragma unroll
  for (int r = 0; r < RT; r++)
#pragma unroll
    for (int c = 0; c < CT; c++) acc[r][c] = (v4d){0.0, 0.0, 0.0, 0.0};
  double ra[AE], rb[BE];
  auto load = [&](int k0) {
#pragma unroll
    for (int e = 0; e < AE; e++) {
      const int idx = tid + 256 * e, k = idx / ROWS, i = idx % ROWS;
      ra[e] = (k0 + k < kend && row0 + i < lda) ? A[(size_t)(row0 + i) + (size_t)(k0 + k) * lda] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < BE; e++) {
      const int idx = tid + 256 * e, c = idx / GEMM_KC, k = idx % GEMM_KC, col = col0 + c;
      rb[e] = (col < ncols && k0 + k < kend) ? in[(size_t)in_off[col] + k0 + k] : 0.0;
    }
  };
  load(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += GEMM_KC) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < AE; e++) {
      const int idx = tid + 256 * e;
      As[idx / ROWS][idx % ROWS] = ra[e];
    }
#pragma unroll
    for (int e = 0; e < BE; e++) {
      const int idx = tid + 256 * e;
      Bs[idx % GEMM_KC][idx / GEMM_KC] = rb[e];
    }
    __syncthreads();
    if (k0 + GEMM_KC < kend) load(k0 + GEMM_KC);
#pragma unroll
    for (int kk = 0; kk < GEMM_KC; kk += 4) {
      double b[CT];
#pragma unroll
      for (int c = 0; c < CT; c++) b[c] = Bs[kk + (lane >> 4)][(wave * CT + c) * 16 + (lane & 15)];
#pragma unroll
      for (int r = 0; r < RT; r++) {
        const double a = As[kk + (lane >> 4)][r * 16 + (lane & 15)];
#pragma unroll
        for (int c = 0; c < CT; c++) acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[c], acc[r][c], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < CT; c++) {
    const int col = col0 + (wave * CT + c) * 16 + (lane & 15);
    if (col >= ncols) continue;
    double* o = ws ? ws + ((size_t)blockIdx.z * ncols + col) * ldw : out + out_off[col];
    const double s = ws ? 1.0 : ldexp(1.0, sexp[col]);
    const bool add = !ws && accumulate;
#pragma unroll
    for (int r = 0; r < RT; r++) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int row = row0 + r * 16 + (lane >> 4) + 4 * i;
        if (row < M) {
          const double v = s * acc[r][c][i];
          o[row] = add ? o[row] + v : v;
        }
      }
    }
  }
}

/* Sums the nsplit partial products of a split-K GEMM in split order and
 * applies scale / accumulate as k_gemm_cols does.  A launch of its own: a
 * last-block reduction inside the GEMM needs a device-scope fence per block,
 * which on gfx950 writes back the XCD's L2 and cost more than this launch. */
static __global__ __launch_bounds__(256) void k_gemm_reduce(int M, int ncols, int nsplit,
                                                            const double* __restrict__ ws, int ldw,
                                                            const int* __restrict__ out_off,
                                                            const int* __restrict__ sexp, double* __restrict__ out,
                                                            int accumulate) {
  const size_t n = (size_t)ncols * M;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int col = (int)(e / M), row = (int)(e % M);
    double v = 0.0;
    for (int z = 0; z < nsplit; z++) v += ws[((size_t)z * ncols + col) * ldw + row];
    double* o = out + out_off[col] + row;
    const double t = ldexp(1.0, sexp[col]) * v;
    *o = accumulate ? *o + t : t;
  }
}

/* acc[u] = sum over steps st < nsteps, in order, of the 16 x 16 x 4 FP64
 * MFMA of A value lda(st, u) and B value ldb(st): the A values (operator
 * entries from global memory) of KB steps are loaded before their MFMAs,
 * so a chain of nsteps waits on about nsteps / KB load latencies instead
 * of nsteps. */
template <int KB, int TPW, typename LA, typename LB>
__device__ __forceinline__ void mfma_chain(int nsteps, LA lda, LB ldb, v4d (&acc)[TPW]) {
  for (int sb = 0; sb < nsteps; sb += KB) {
    double a[KB][TPW];
#pragma unroll
    for (int st = 0; st < KB; st++)
#pragma unroll
      for (int u = 0; u < TPW; u++) a[st][u] = sb + st < nsteps ? lda(sb + st, u) : 0.0;
#pragma unroll
    for (int st = 0; st < KB; st++) {
      if (sb + st >= nsteps) break;
      const double b = ldb(sb + st);
#pragma unroll
      for (int u = 0; u < TPW; u++) acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[st][u], b, acc[u], 0, 0, 0);
    }
  }
}

/* One level of the downward pass with its two operator applies fused
 * (down_eval__gpu_non_leaves, fmm/cuda.cu:1682-1801): a block owns up to 16
 * parents at depth d - 1 (in_off[j] = their downward densities, out_off[j]
 * = their first child's check values) and one child octant o = blockIdx.y.
 *   phase 1  T = dchk[child] + 2^(d-1) DE2DC[o] dden[parent]   (L2L; written
 *            back to dchk for every child)
 *   phase 2  dden[child] = 2^-d DC2DE T                         (DC2DE, for the
 *            children that hold targets)
 * Both are v_mfma_f64_16x16x4f64 products over the 16 parent columns (the
 * operand layout of k_gemm_cols), K walked in steps of 4 in order, with the
 * parents' densities and T staged in LDS: one launch per level where the
 * column GEMMs took two, and the check values never make a round trip
 * through HBM between them.  Rows RT * 16 >= n_e; wave w takes row tiles
 * w, w + 4, ... of each phase. */
template <int RT>
__global__ __launch_bounds__(256) void k_l2l_dc2de(int ne, int lde, const double* __restrict__ DE2DC, int lda_l2l,
                                                   const double* __restrict__ DC2DE, int lda_e, int npar,
                                                   const int* __restrict__ in_off, const int* __restrict__ out_off,
                                                   const int* __restrict__ tag, int d, double* __restrict__ dchk,
                                                   double* __restrict__ dden) {
  constexpr int ROWS = 16 * RT;
  constexpr int LDP = 160 + 1; /* >= lde for NP <= 6; odd: conflict-free column reads */
  constexpr int LDT = ROWS + 1;
  __shared__ double Dp[16][LDP];
  __shared__ double T[16][LDT];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int o = blockIdx.y, j0 = blockIdx.x * 16;
  const int jn = min(16, npar - j0);
  for (int e = tid; e < 16 * lde; e += 256) {
    const int j = e / lde, k = e % lde;
    Dp[j][k] = j < jn ? dden[(size_t)in_off[j0 + j] + k] : 0.0;
  }
  __syncthreads();
  const int jl = lane & 15, kl = lane >> 4;
  const double s1 = ldexp(1.0, d - 1), s2 = ldexp(1.0, -d);
  const bool live = jl < jn;
  const size_t child_off = live ? (size_t)out_off[j0 + jl] + (size_t)o * lde : 0; /* this lane's column */
  /* a wave's row tiles share each B value and keep independent
   * accumulator chains in flight (mfma_chain: operator loads 8 steps ahead) */
  constexpr int TPW = (RT + 3) / 4;
  /* this lane's check values, loaded before the products (their latency
   * hidden behind phase 1) */
  double cv[TPW][4];
#pragma unroll
  for (int u = 0; u < TPW; u++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = 16 * (wave + 4 * u) + kl + 4 * i;
      cv[u][i] = (live && wave + 4 * u < RT && row < ne) ? dchk[child_off + row] : 0.0;
    }
  const bool trg = live && (tag[child_off / lde] & 2) != 0; /* KIFMM_TRGNODE */
  v4d acc[TPW];
#pragma unroll
  for (int u = 0; u < TPW; u++) acc[u] = (v4d){0.0, 0.0, 0.0, 0.0};
  mfma_chain<8>(
      lde / 4,
      [&](int st, int u) {
        const int row_a = 16 * (wave + 4 * u) + (lane & 15);
        return row_a < ne ? DE2DC[(size_t)(4 * st + kl) * lda_l2l + (size_t)o * lde + row_a] : 0.0;
      },
      [&](int st) { return Dp[jl][4 * st + kl]; }, acc);
#pragma unroll
  for (int u = 0; u < TPW; u++) {
    const int r = wave + 4 * u;
    if (r >= RT) continue;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = 16 * r + kl + 4 * i;
      double t = 0.0;
      if (live && row < ne) {
        t = cv[u][i] + s1 * acc[u][i];
        dchk[child_off + row] = t;
      }
      T[jl][row] = t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < TPW; u++) acc[u] = (v4d){0.0, 0.0, 0.0, 0.0};
  mfma_chain<8>(
      lde / 4,
      [&](int st, int u) {
        const int row_a = min(16 * (wave + 4 * u) + (lane & 15), lda_e - 1); /* rows past RT tiles: unused */
        return DC2DE[(size_t)(4 * st + kl) * lda_e + row_a];
      },
      [&](int st) { return T[jl][4 * st + kl]; }, acc);
#pragma unroll
  for (int u = 0; u < TPW; u++) {
    const int r = wave + 4 * u;
    if (r >= RT) continue;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = 16 * r + kl + 4 * i;
      if (trg && row < ne) dden[child_off + row] = s2 * acc[u][i];
    }
  }
}

/* One level of the upward pass with its two operator applies fused
 * (up_eval__gpu_non_leaf + up_calc__gpu_reduction, fmm/cuda.cu:430-566): a
 * block owns up to 16 parents at depth d (in_off[j] = their 8 children's
 * upward densities, 8 * lde contiguous doubles; out_off[j] = their own).
 *   phase 1  chk = 2^(d+1) UE2UC_stack (children's densities)   (UC check values)
 *   phase 2  uden[parent] = 2^-d UC2UE chk
 * v_mfma_f64_16x16x4f64 over the 16 parent columns, K in steps of 4 in
 * order (the column GEMM's order, so its sums are bit-identical to two
 * unsplit k_gemm_cols launches), the children's densities staged through LDS
 * 64 k at a time (the next chunk loaded during the current one) and chk kept
 * in LDS between the phases.  RT1 / RT2: row
 * tiles of the two operators (lda_uc / 16, lda_e / 16). */
constexpr int M2M_KC = 64;
template <int RT1, int RT2>
__global__ __launch_bounds__(256) void k_m2m(int ne, int lde, int lduc, const double* __restrict__ UE2UC, int lda_uc,
                                             const double* __restrict__ UC2UE, int lda_e, int npar,
                                             const int* __restrict__ in_off, const int* __restrict__ out_off, int d,
                                             double* __restrict__ uden) {
  constexpr int LDC = 16 * RT1 + 1; /* odd: conflict-free column reads */
  __shared__ double Bs[16][M2M_KC + 1];
  __shared__ double C[16][LDC];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int j0 = blockIdx.x * 16, jn = min(16, npar - j0);
  const int jl = lane & 15, kl = lane >> 4;
  const int K = 8 * lde;
  constexpr int TPW1 = (RT1 + 3) / 4, TPW2 = (RT2 + 3) / 4;
  v4d acc[TPW1];
#pragma unroll
  for (int u = 0; u < TPW1; u++) acc[u] = (v4d){0.0, 0.0, 0.0, 0.0};
  /* the children's densities of the next K chunk are loaded into
   * registers while the current chunk is multiplied */
  constexpr int BPT = 16 * M2M_KC / 256;
  double rb[BPT];
  size_t boff[BPT];
#pragma unroll
  for (int m = 0; m < BPT; m++) {
    const int j = (tid + 256 * m) / M2M_KC;
    boff[m] = j < jn ? (size_t)in_off[j0 + j] : 0;
  }
  auto load_b = [&](int kc) {
    const int kn = min(M2M_KC, K - kc);
#pragma unroll
    for (int m = 0; m < BPT; m++) {
      const int e = tid + 256 * m, j = e / M2M_KC, k = e % M2M_KC;
      rb[m] = (j < jn && k < kn) ? uden[boff[m] + kc + k] : 0.0;
    }
  };
  load_b(0);
  for (int kc = 0; kc < K; kc += M2M_KC) {
    const int kn = min(M2M_KC, K - kc); /* a multiple of 4 (8 lde) */
    __syncthreads(); /* the previous chunk's readers are done */
#pragma unroll
    for (int m = 0; m < BPT; m++) {
      const int e = tid + 256 * m;
      Bs[e / M2M_KC][e % M2M_KC] = rb[m];
    }
    __syncthreads();
    if (kc + M2M_KC < K) load_b(kc + M2M_KC);
    mfma_chain<8>(
        kn / 4,
        [&](int st, int u) {
          const int row_a = min(16 * (wave + 4 * u) + (lane & 15), lda_uc - 1); /* rows past RT1 tiles: unused */
          return UE2UC[(size_t)(kc + 4 * st + kl) * lda_uc + row_a];
        },
        [&](int st) { return Bs[jl][4 * st + kl]; }, acc);
  }
  const double s1 = ldexp(1.0, d + 1), s2 = ldexp(1.0, -d);
#pragma unroll
  for (int u = 0; u < TPW1; u++) {
    const int r = wave + 4 * u;
    if (r >= RT1) continue;
#pragma unroll
    for (int i = 0; i < 4; i++) C[jl][16 * r + kl + 4 * i] = s1 * acc[u][i];
  }
  __syncthreads();
  v4d acc2[TPW2];
#pragma unroll
  for (int u = 0; u < TPW2; u++) acc2[u] = (v4d){0.0, 0.0, 0.0, 0.0};
  mfma_chain<8>(
      lduc / 4,
      [&](int st, int u) {
        const int row_a = min(16 * (wave + 4 * u) + (lane & 15), lda_e - 1);
        return UC2UE[(size_t)(4 * st + kl) * lda_e + row_a];
      },
      [&](int st) { return C[jl][4 * st + kl]; }, acc2);
  if (jl >= jn) return;
  double* const out = uden + out_off[j0 + jl];
#pragma unroll
  for (int u = 0; u < TPW2; u++) {
    const int r = wave + 4 * u;
    if (r >= RT2) continue;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = 16 * r + kl + 4 * i;
      if (row < ne) out[row] = s2 * acc2[u][i];
    }
  }
}

/* Column GEMM for a short operator over many columns: M <= 64 rows, K <=
 * GEMM_SK and a multiple of 8 (the evaluator's leaf UC2UE at NP <= 4, 56 x
 * 152 over one column per leaf).  out[:, j] = (accumulate ? out[:, j] : 0)
 * + 2^sexp[j] A in[:, j], as k_gemm_cols.  A block of 8 waves stages the
 * whole operator in LDS once ([k][row], up to 80 KB: 2 blocks per CU) and
 * each wave then streams column tiles of 16 on its own, with no barrier:
 * tile b * 8 + w, then every gridDim.x * 8 further.  MFMA step s = 2t + u
 * (v_mfma_f64_16x16x4f64) takes k = 8t + 2q + u in k-slot q, so lane j +
 * 16q reads its column's doubles 8t + 2q, 8t + 2q + 1 with one 16-byte
 * load (the 4 lanes of a column read 64 contiguous bytes per group of 8
 * k), and the A operand from LDS the same k; the next group's loads are in
 * flight during the current group's MFMAs.  k_gemm_cols runs into its
 * K-chunk loop's latency on these shapes (256 blocks of 10 chunks at 32768
 * columns).  The sum order over k differs from k_gemm_cols': rounding-level
 * differences. */
constexpr int GEMM_SK = 160;
template <int RT>
__global__ __launch_bounds__(512) void k_gemm_stream(int M, int K, const double* __restrict__ A, int lda,
                                                     int ncols, const int* __restrict__ in_off,
                                                     const int* __restrict__ out_off,
                                                     const int* __restrict__ sexp,
                                                     const double* __restrict__ in,
                                                     double* __restrict__ out, int accumulate) {
  constexpr int ROWS = 16 * RT;
  constexpr int U = 4; /* groups of 8 k (16-byte loads per lane) in flight */
  __shared__ double As[GEMM_SK * ROWS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int e = tid; e < K * ROWS; e += 512) {
    const int k = e / ROWS, i = e - k * ROWS;
    As[e] = A[(size_t)i + (size_t)k * lda];
  }
  __syncthreads();
  const int j = lane & 15, q = lane >> 4;
  const int ntiles = (ncols + 15) >> 4, kt = K >> 3;
  for (int tile = blockIdx.x * 8 + wave; tile < ntiles; tile += gridDim.x * 8) {
    const int col = tile * 16 + j;
    const bool live = col < ncols;
    const double* __restrict__ b = in + (live ? (size_t)in_off[col] : 0) + 2 * q;
    v4d acc[RT];
#pragma unroll
    for (int r = 0; r < RT; r++) acc[r] = (v4d){0.0, 0.0, 0.0, 0.0};
    double2 cur[U], nxt[U];
    auto fetch = [&](double2 (&v)[U], int t0) {
#pragma unroll
      for (int u = 0; u < U; u++)
        v[u] = (live && t0 + u < kt) ? *reinterpret_cast<const double2*>(b + 8 * (t0 + u)) : make_double2(0.0, 0.0);
    };
    fetch(cur, 0);
    for (int t0 = 0; t0 < kt; t0 += U) {
      if (t0 + U < kt) fetch(nxt, t0 + U);
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (t0 + u < kt) { /* wave-uniform */
          const double* a0 = As + (8 * (t0 + u) + 2 * q) * ROWS + j;
#pragma unroll
          for (int r = 0; r < RT; r++) acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[16 * r], cur[u].x, acc[r], 0, 0, 0);
#pragma unroll
          for (int r = 0; r < RT; r++)
            acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[ROWS + 16 * r], cur[u].y, acc[r], 0, 0, 0);
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) cur[u] = nxt[u];
    }
    if (!live) continue;
    /* D[i][j] in lane j + 16 (i % 4), element i / 4 */
    double* const o = out + out_off[col];
    const double sc = ldexp(1.0, sexp[col]);
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int row = 16 * r + q + 4 * e;
        if (row < M) {
          const double v = sc * acc[r][e];
          o[row] = accumulate ? o[row] + v : v;
        }
      }
  }
}

/* Launch shape of one column-list GEMM (host side, shared by the evaluator
 * and the kernel tests): the row-tile count picks the instantiation, and K
 * is split (partials reduced in a fixed order) when there are too few
 * blocks to fill the chip.  Short operators over many columns take
 * k_gemm_stream (stream = 1): its grid is col_blocks blocks of 8 waves.
 * The choice depends on the shape alone, so every stream layout gives the
 * same bits.  (Beside a kernel that fills the LDS, as the U-list on the
 * second stream of the overlap layouts does, the 80 KB blocks wait for CUs
 * to drain, 2 - 11 ms against 70 us alone; the autotuner times the layouts
 * with it.) */
struct GemmShape {
  int rt, ct, row_groups, col_blocks, nsplit, ksplit, ldw, stream;
};
inline GemmShape gemm_shape(int ncols, int M, int K) {
  GemmShape s;
  s.stream = 0;
  /* from 8192 columns: 64 blocks and more (at 4096 columns its 32 blocks lost to the split-K column GEMM) */
  if (M <= 64 && K <= GEMM_SK && K % 8 == 0 && ncols >= 8192) {
    s.stream = 1;
    s.rt = (M + 15) / 16;
    s.ct = 1;
    s.row_groups = 1;
    s.col_blocks = min((ncols + 127) / 128, 512); /* at most 2 blocks per CU, the rest by striding */
    s.nsplit = 1;
    s.ksplit = K;
    s.ldw = (M + 7) & ~7;
    return s;
  }
  const int tiles = (M + 15) / 16;
  /* short operators: two 64-column tiles per block; otherwise one, with the
   * rows split into groups of <= 10 tiles.  When the inputs are short
   * (K <= 2M: re-reading them per row group is cheap) the rows are split
   * further, down to 4 tiles, while the launch has fewer than 1024 blocks
   * (4 per CU); long inputs are split along K instead (below). */
  s.ct = tiles <= 4 ? 2 : 1;
  const int cb = (ncols + 64 * s.ct - 1) / (64 * s.ct);
  s.row_groups = (tiles + 9) / 10;
  while (s.ct == 1 && K <= 2 * M && cb * s.row_groups < 1024 && (tiles + s.row_groups) / (s.row_groups + 1) >= 4)
    s.row_groups++;
  const int per = (tiles + s.row_groups - 1) / s.row_groups;
  s.rt = s.ct == 2 ? (per <= 2 ? 2 : 4) : per <= 4 ? 4 : per <= 5 ? 5 : per <= 7 ? 7 : 10;
  s.col_blocks = (ncols + 64 * s.ct - 1) / (64 * s.ct);
  const int blocks = s.col_blocks * s.row_groups;
  const int kchunks = (K + GEMM_KC - 1) / GEMM_KC;
  int ns = 1;
  if (blocks < 256) ns = min(min(16, kchunks), (512 + blocks - 1) / blocks);
  ns = max(ns, 1);
  const int kper = (kchunks + ns - 1) / ns;
  s.ksplit = kper * GEMM_KC;
  s.nsplit = (K + s.ksplit - 1) / s.ksplit;
  s.ldw = (M + 7) & ~7;
  return s;
}

/* Launches k_gemm_cols (and k_gemm_reduce when K is split; ws must then
 * hold nsplit * ncols * ldw doubles). */
inline hipError_t launch_gemm_cols(hipStream_t st, const GemmShape& sh, int M, int K, const double* A, int lda,
                                   int ncols, const int* in_off, const int* out_off, const int* sexp,
                                   const double* in, double* out, int accumulate, double* ws) {
  if (sh.stream) {
#define KIFMM_GEMM_STREAM(R)                                                                                   \
  if (sh.rt == R)                                                                                              \
    hipLaunchKernelGGL(k_gemm_stream<R>, dim3(sh.col_blocks), dim3(512), 0, st, M, K, A, lda, ncols, in_off,   \
                       out_off, sexp, in, out, accumulate);
    KIFMM_GEMM_STREAM(1)
    else KIFMM_GEMM_STREAM(2)
    else KIFMM_GEMM_STREAM(3)
    else KIFMM_GEMM_STREAM(4)
    else return hipErrorInvalidValue;
#undef KIFMM_GEMM_STREAM
    return hipGetLastError();
  }
  const dim3 grid(sh.col_blocks, sh.row_groups, sh.nsplit);
  double* w = sh.nsplit > 1 ? ws : nullptr;
  const int ks = sh.nsplit > 1 ? sh.ksplit : K;
#define KIFMM_GEMM_CASE(R, C)                                                                                  \
  if (sh.rt == R && sh.ct == C)                                                                                \
    hipLaunchKernelGGL((k_gemm_cols<R, C>), grid, dim3(256), 0, st, M, K, A, lda, ncols, in_off, out_off, sexp, \
                       in, out, accumulate, ks, w, sh.ldw);
  KIFMM_GEMM_CASE(2, 2)
  else KIFMM_GEMM_CASE(4, 2)
  else KIFMM_GEMM_CASE(4, 1)
  else KIFMM_GEMM_CASE(5, 1)
  else KIFMM_GEMM_CASE(7, 1)
  else KIFMM_GEMM_CASE(10, 1)
  else return hipErrorInvalidValue;
#undef KIFMM_GEMM_CASE
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || sh.nsplit == 1) return e;
  const size_t n = (size_t)ncols * M;
  hipLaunchKernelGGL(k_gemm_reduce, dim3((int)min((size_t)2048, (n + 255) / 256)), dim3(256), 0, st, M, ncols,
                     sh.nsplit, ws, sh.ldw, out_off, sexp, out, accumulate);
  return hipGetLastError();
}

/* ---------------------------------------------------------- V-list DFTs */

/* Grid geometry for accuracy NP: N = 2 NP points per axis, H = N/2 + 1
 * half-spectrum length along x, F = N*N*H complex frequencies, the surface
 * sits at grid offset NP/2 (fmm/trans.cc:335-355,394-416). */
/* Storage of the V-list spectra: double2 on the FP64 path, float2 on the
 * mixed path (FP32 complex M2L).  R is the arithmetic type of the Hadamard
 * multiply-accumulate; the DFTs compute in FP64 and round on store. */
template <typename ST>
struct Spec;
template <>
struct Spec<double2> {
  typedef double R;
  __device__ static double2 mk(double a, double b) { return make_double2(a, b); }
};
template <>
struct Spec<float2> {
  typedef float R;
  __device__ static float2 mk(double a, double b) { return make_float2((float)a, (float)b); }
};
__device__ __forceinline__ double fm(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fm(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <int NP>
struct Grid {
  static constexpr int N = 2 * NP, H = NP + 1, F = N * N * H, OFF = NP / 2;
  static constexpr int NE = NP * NP * NP - (NP - 2) * (NP - 2) * (NP - 2);
};

/* Forward transform of the 8 sibling equivalent densities of one child
 * block: scatter 2^depth * uden onto the (2NP)^3 grid (only the NP^3 cube
 * at offset NP/2 is non-zero, so the three axis passes are pruned to it)
 * and produce the r2c spectrum [kz][ky][kx<=N/2] (FFTW layout), stored
 * interleaved by octant: X[blk][f][oct].  Each pass gives a thread one
 * column: its NP inputs read once from LDS, all N (x pass: H) outputs of
 * the column summed with the twiddles held in registers (the loops are
 * unrolled, so every twiddle index is a constant).  The LDS planes are
 * octant-fastest, [..][oct], so the lanes of a wave touch consecutive
 * entries, and the z pass writes each kz plane of X contiguously.  Per
 * output the sum runs over the input index in increasing order. */
template <int NP, typename ST = double2>
__global__ __launch_bounds__(512) void k_dft_fwd(const int* __restrict__ blocks, /* first child node per block */
                                                 const int* __restrict__ xslot,  /* X slot per block */
                                                 const int* __restrict__ depth,
                                                 const double* __restrict__ uden, int ld_e,
                                                 const double2* __restrict__ tw, /* W^m, m < N */
                                                 const int* __restrict__ cube2surf, /* NP^3: sample id or -1 */
                                                 ST* __restrict__ X) {
  using G = Grid<NP>;
  constexpr int N = G::N, H = G::H, F = G::F, NE = G::NE, OFF = G::OFF;
  __shared__ double gs[NE][8];
  __shared__ double2 A[NP][NP][H][8]; /* [iz][iy][kx][oct] */
  __shared__ double2 B[NP][N][H][8];  /* [iz][ky][kx][oct] */
  __shared__ int c2s[NP * NP * NP];
  const int tid = threadIdx.x;
  const int first = blocks[blockIdx.x];
  double2 w[N];
#pragma unroll
  for (int m = 0; m < N; m++) w[m] = tw[m];
  for (int e = tid; e < NP * NP * NP; e += blockDim.x) c2s[e] = cube2surf[e];
  for (int e = tid; e < 8 * NE; e += blockDim.x) {
    const int o = e / NE, j = e % NE, node = first + o;
    gs[j][o] = ldexp(uden[(size_t)node * ld_e + j], depth[node]);
  }
  __syncthreads();
  /* x pass: column (iz, iy, oct), real input x in [OFF, OFF + NP) -> kx <= N/2 */
  for (int e = tid; e < NP * NP * 8; e += blockDim.x) {
    const int o = e & 7, iy = (e >> 3) % NP, iz = (e >> 3) / NP;
    double v[NP];
#pragma unroll
    for (int ix = 0; ix < NP; ix++) {
      const int s = c2s[(ix * NP + iy) * NP + iz];
      v[ix] = (s >= 0) ? gs[s][o] : 0.0;
    }
#pragma unroll
    for (int kx = 0; kx < H; kx++) {
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int ix = 0; ix < NP; ix++) {
        const double2 t = w[(kx * (ix + OFF)) % N];
        re = __builtin_fma(v[ix], t.x, re);
        im = __builtin_fma(v[ix], t.y, im);
      }
      A[iz][iy][kx][o] = make_double2(re, im);
    }
  }
  __syncthreads();
  /* y pass: column (iz, kx, oct) */
  for (int e = tid; e < NP * H * 8; e += blockDim.x) {
    const int o = e & 7, kx = (e >> 3) % H, iz = (e >> 3) / H;
    double2 a[NP];
#pragma unroll
    for (int iy = 0; iy < NP; iy++) a[iy] = A[iz][iy][kx][o];
#pragma unroll
    for (int ky = 0; ky < N; ky++) {
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int iy = 0; iy < NP; iy++) {
        const double2 t = w[(ky * (iy + OFF)) % N];
        re = __builtin_fma(a[iy].x, t.x, __builtin_fma(-a[iy].y, t.y, re));
        im = __builtin_fma(a[iy].x, t.y, __builtin_fma(a[iy].y, t.x, im));
      }
      B[iz][ky][kx][o] = make_double2(re, im);
    }
  }
  __syncthreads();
  /* z pass: column (ky, kx, oct) = thread e; output (kz, ky, kx, oct) at
   * X[slot][kz * N * H * 8 + e] */
  ST* out = X + (size_t)xslot[blockIdx.x] * F * 8;
  for (int e = tid; e < N * H * 8; e += blockDim.x) {
    const int o = e & 7, kk = e >> 3; /* kk = ky * H + kx */
    double2 b[NP];
#pragma unroll
    for (int iz = 0; iz < NP; iz++) b[iz] = B[iz][kk / H][kk % H][o];
#pragma unroll
    for (int kz = 0; kz < N; kz++) {
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int iz = 0; iz < NP; iz++) {
        const double2 t = w[(kz * (iz + OFF)) % N];
        re = __builtin_fma(b[iz].x, t.x, __builtin_fma(-b[iz].y, t.y, re));
        im = __builtin_fma(b[iz].x, t.y, __builtin_fma(b[iz].y, t.x, im));
      }
      out[(size_t)kz * N * H * 8 + e] = Spec<ST>::mk(re, im);
    }
  }
}

/* Translation spectra (compute_fft_trans__gpu_eval + D2Z, fmm/cuda.cu:874-963):
 * for id = (i1+3) + 7(i2+3) + 49(i3+3) with max|i| > 1,
 * tt[g] = 1/(4 pi |2 i - RP_g|) on the wrapped (2NP)^3 grid, full r2c.
 * Output layout T[f][id] so that a wave reads one frequency's 343 entries
 * as uniform (scalar) loads. */
template <int NP, typename ST = double2>
__global__ __launch_bounds__(256) void k_trans(const double2* __restrict__ tw, ST* __restrict__ T) {
  using G = Grid<NP>;
  constexpr int N = G::N, H = G::H, F = G::F;
  __shared__ double g[N * N * N];
  __shared__ double2 a[N * N * H];
  __shared__ double2 b[N * N * H];
  __shared__ double2 w[N];
  const int id = blockIdx.x, tid = threadIdx.x;
  const int i1 = id % 7 - 3, i2 = (id / 7) % 7 - 3, i3 = id / 49 - 3;
  const bool valid = abs(i1) > 1 || abs(i2) > 1 || abs(i3) > 1;
  if (!valid) {
    for (int f = tid; f < F; f += blockDim.x) T[(size_t)f * 343 + id] = Spec<ST>::mk(0.0, 0.0);
    return;
  }
  if (tid < N) w[tid] = tw[tid];
  const double step = 2.0 / (NP - 1);
  for (int e = tid; e < N * N * N; e += blockDim.x) {
    const int i = e % N, j = (e / N) % N, k = e / (N * N);
    const int gi = i < NP ? i : i - N, gj = j < NP ? j : j - N, gk = k < NP ? k : k - N;
    const double dx = i1 * 2.0 - gi * step, dy = i2 * 2.0 - gj * step, dz = i3 * 2.0 - gk * step;
    g[e] = OOFP_R / sqrt(dx * dx + dy * dy + dz * dz);
  }
  __syncthreads();
  for (int e = tid; e < N * N * H; e += blockDim.x) {
    const int kx = e % H, yz = e / H;
    double re = 0.0, im = 0.0;
    for (int x = 0; x < N; x++) {
      const double v = g[yz * N + x];
      const double2 t = w[(kx * x) % N];
      re = __builtin_fma(v, t.x, re);
      im = __builtin_fma(v, t.y, im);
    }
    a[e] = make_double2(re, im);
  }
  __syncthreads();
  for (int e = tid; e < N * N * H; e += blockDim.x) {
    const int kx = e % H, ky = (e / H) % N, z = e / (H * N);
    double re = 0.0, im = 0.0;
    for (int y = 0; y < N; y++) {
      const double2 v = a[(z * N + y) * H + kx];
      const double2 t = w[(ky * y) % N];
      re = __builtin_fma(v.x, t.x, __builtin_fma(-v.y, t.y, re));
      im = __builtin_fma(v.x, t.y, __builtin_fma(v.y, t.x, im));
    }
    b[e] = make_double2(re, im);
  }
  __syncthreads();
  for (int e = tid; e < N * N * H; e += blockDim.x) {
    const int kx = e % H, ky = (e / H) % N, kz = e / (H * N);
    double re = 0.0, im = 0.0;
    for (int z = 0; z < N; z++) {
      const double2 v = b[(z * N + ky) * H + kx];
      const double2 t = w[(kz * z) % N];
      re = __builtin_fma(v.x, t.x, __builtin_fma(-v.y, t.y, re));
      im = __builtin_fma(v.x, t.y, __builtin_fma(v.y, t.x, im));
    }
    T[(size_t)e * 343 + id] = Spec<ST>::mk(re, im);
  }
}

/* Spectral M2L for all 8 children of a parent at once.  The V-list of child
 * s of parent p is every child c of p's 26 same-depth neighbours o that is
 * not adjacent to s (fmm/input.cc:314-336).  With child offset
 * d = 2o + (c - s) in [-3,3]^3 the translation id is
 * (dx+3) + 7(dy+3) + 49(dz+3) (fmm/evaluate--cuda.cu:1713-1723), i.e. a
 * per-neighbour base plus a per-(c,s) constant.  Adjacent pairs
 * (max|d| < 2) hit the 27 near-field ids, whose spectra are exactly zero,
 * so every (c,s) pair is computed unconditionally and adds exactly 0 where
 * the reference has no V entry.  Missing neighbours use the all-zero X
 * slot 0.
 *
 * Block = one group of up to 64 parents (lane = parent) at one frequency f
 * (block-uniform, so translation entries are scalar loads).  The
 * union of the group's neighbour slots is staged in LDS once (216 slots
 * for an aligned 4x4x4 group of parents instead of 64 x 26 reads from L2);
 * wave h computes target children s in [4h, 4h+4). */
constexpr int HAD_CAP = 232; /* LDS slots per group */

/* Block order of the Hadamard launches (1-D grid of 8 per_xcd blocks, per_xcd
 * = ceil(n_groups * n_fblocks / 8)): the XCDs take consecutive blocks round
 * robin, so block b runs on XCD b % 8, and v = (b % 8) per_xcd + b / 8 gives
 * each XCD one contiguous run of v = frequency block * n_groups + group.  All
 * the groups of a frequency block then run on one XCD, and the source
 * spectra of that block, which the groups' neighbour unions share, are
 * fetched into one L2 instead of eight.  Returns -1 past the last pair. */
__device__ __forceinline__ int had_block(int n_groups, int n_fblocks, int per_xcd) {
  const int v = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  return v < n_groups * n_fblocks ? v : -1;
}
template <int NP, int HALF, typename ST>
__device__ __forceinline__ void hadamard_half(const ST (*Xs)[9], const unsigned char (*sloc)[64], int lane,
                                              const ST* __restrict__ Tf, typename Spec<ST>::R yr[4],
                                              typename Spec<ST>::R yi[4]) {
  typedef typename Spec<ST>::R R;
#pragma unroll 1
  for (int o = 0; o < 27; o++) {
    if (o == 13) continue; /* the parent itself: its children are all adjacent */
    const int ox = o / 9 - 1, oy = (o / 3) % 3 - 1, oz = o % 3 - 1;
    const ST* __restrict__ To = Tf + (2 * ox + 3) + 7 * (2 * oy + 3) + 49 * (2 * oz + 3);
    const ST* xs = Xs[sloc[o][lane]];
    R xr[8], xi[8];
#pragma unroll
    for (int c = 0; c < 8; c++) { const ST v = xs[c]; xr[c] = v.x; xi[c] = v.y; }
#pragma unroll
    for (int c = 0; c < 8; c++) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int sidx = 4 * HALF + q;
        const int ex = ((c >> 2) & 1) - ((sidx >> 2) & 1);
        const int ey = ((c >> 1) & 1) - ((sidx >> 1) & 1);
        const int ez = (c & 1) - (sidx & 1);
        const ST t = To[ex + 7 * ey + 49 * ez];
        yr[q] = fm(t.x, xr[c], fm(-t.y, xi[c], yr[q]));
        yi[q] = fm(t.x, xi[c], fm(t.y, xr[c], yi[q]));
      }
    }
  }
}

/* One block = one parent group (up to 64 parents, a lane each) and
 * FB consecutive frequencies (4, 8 or 16, chosen by the autotuner).  The union of the group's neighbour
 * slots is staged in LDS per frequency; the next frequency's spectra are
 * loaded into registers while the current one is multiplied, and the
 * parents' 27 union indices are staged once per block. */
template <int NP, typename ST, int FB>
__global__ __launch_bounds__(128) void k_hadamard(const int4* __restrict__ groups, /* {first parent, count, union beg, union size} */
                                                  const int* __restrict__ uslots,  /* X slots of each union */
                                                  const unsigned char* __restrict__ loc, /* 27 union indices per parent */
                                                  const int* __restrict__ yslot,   /* Y slot per parent */
                                                  const ST* __restrict__ T,
                                                  const ST* __restrict__ X,
                                                  ST* __restrict__ Y, int n_groups, int per_xcd) {
  typedef typename Spec<ST>::R R;
  using G = Grid<NP>;
  constexpr int F = G::F;
  static_assert(F % FB == 0, "frequency blocks");
  constexpr int PER = (HAD_CAP * 8 + 127) / 128;
  __shared__ ST Xs[HAD_CAP][9];
  __shared__ unsigned char sloc[27][64];
  const int v = had_block(n_groups, F / FB, per_xcd);
  if (v < 0) return; /* whole block, before any barrier */
  const int4 g = groups[v % n_groups];
  const int f0 = (v / n_groups) * FB;
  const int tid = threadIdx.x, lane = tid & 63;
  const int half = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ne8 = g.w * 8;
  int us[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int e = tid + 128 * k;
    us[k] = e < ne8 ? uslots[g.z + (e >> 3)] : 0;
  }
  ST pre[PER];
  auto fetch = [&](int f) {
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int e = tid + 128 * k;
      pre[k] = e < ne8 ? X[((size_t)us[k] * F + f) * 8 + (e & 7)] : Spec<ST>::mk(0.0, 0.0);
    }
  };
  fetch(f0);
  /* union indices of the block's parents, [neighbour][lane]; lanes past the
   * group's count repeat its last parent */
  for (int e = tid; e < 27 * 64; e += 128) {
    const int l = e / 27, o = e - 27 * l;
    sloc[o][l] = loc[(size_t)(g.x + min(l, g.y - 1)) * 27 + o];
  }
  const int p = g.x + min(lane, g.y - 1);
  const int ys = yslot[p];
  for (int fi = 0; fi < FB; fi++) {
    const int f = f0 + fi;
    __syncthreads(); /* the previous frequency's readers are done with Xs */
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int e = tid + 128 * k;
      if (e < ne8) Xs[e >> 3][e & 7] = pre[k];
    }
    __syncthreads();
    if (fi + 1 < FB) fetch(f + 1);
    const ST* __restrict__ Tf = T + (size_t)f * 343;
    R yr[4] = {0, 0, 0, 0}, yi[4] = {0, 0, 0, 0};
    if (half == 0)
      hadamard_half<NP, 0, ST>(Xs, sloc, lane, Tf, yr, yi);
    else
      hadamard_half<NP, 1, ST>(Xs, sloc, lane, Tf, yr, yi);
    if (lane < g.y) {
      ST* yb = Y + ((size_t)ys * F + f) * 8 + 4 * half;
#pragma unroll
      for (int q = 0; q < 4; q++) yb[q] = Spec<ST>::mk(yr[q], yi[q]);
    }
  }
}

/* MFMA form of the Hadamard.  For one frequency and one neighbour
 * direction o every parent applies the same 8 x 8 complex matrix
 * T_o[c][c'] = T[f][2o + 3 + c' - c] (target child c, source child c') to
 * its neighbour's 8 child spectra.  In real form [[Tr, -Ti], [Ti, Tr]]
 * (16 x 16, rows: Re Y[c] then Im Y[c]; columns: Re X[c'] then Im X[c'])
 * over 16 parents that is four 16x16x4 MFMAs, k = 4m + kk with source
 * child (4m + kk) & 7, real parts for m < 2 and imaginary for m >= 2.
 * Wave w owns parents 16w .. 16w + 15 of the group: lane j = lane & 15 is
 * row i of A and parent column j of B, kk = lane >> 4.  Per direction a
 * lane reads two translation values (A) and two neighbour spectra (B)
 * from LDS; the group's neighbour-slot union and the frequency's 343
 * translation values are staged per frequency, the next frequency's
 * prefetched into registers meanwhile.  A lane's A operand is Re or Im of a
 * translation value for the real-part MFMAs and -Im or Re for the
 * imaginary-part ones, by the lane's row.  In FP32 the translation values
 * are staged twice, as (Re, Im) in TA and (-Im, Re) in TB, so that the
 * operand is one LDS read at a lane-fixed offset, with no select or negation
 * on the VALU, whose issue adds to the 35-cycle FP32 MFMAs' (-15 % at 16M
 * points, NP 4).  The FP64 form keeps one (Re, Im) copy and selects: its
 * fully unrolled direction loop prefetches the LDS reads of later directions
 * at 128 VGPRs, which the 64-cycle FP64 MFMAs need more than the saved VALU
 * slots (the staged-twice form measured 5 % slower there).  MFMA operand
 * layouts as pinned by test_mfma_f64_operand_layout /
 * test_mfma_f32_operand_layout. */
__device__ __forceinline__ v4d mfma4(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename R>
struct Acc4;
template <>
struct Acc4<double> {
  typedef v4d T;
};
template <>
struct Acc4<float> {
  typedef v4f T;
};

template <int NP, typename ST, int FB>
__global__ __launch_bounds__(256) void k_hadamard_mfma(const int4* __restrict__ groups,
                                                       const int* __restrict__ uslots,
                                                       const unsigned char* __restrict__ loc,
                                                       const int* __restrict__ yslot, const ST* __restrict__ T,
                                                       const ST* __restrict__ X, ST* __restrict__ Y, int n_groups,
                                                       int per_xcd) {
  typedef typename Spec<ST>::R R;
  typedef typename Acc4<R>::T V;
  using G = Grid<NP>;
  constexpr int F = G::F;
  static_assert(F % FB == 0, "frequency blocks");
  constexpr int PER = (HAD_CAP * 8 + 255) / 256;
  constexpr int PERT = (343 + 255) / 256;
  __shared__ ST Xs[HAD_CAP][9];
  constexpr bool F32 = sizeof(R) == 4;
  /* FP32: TA = (Re, Im) then TB = (-Im, Re), 2 x 343 float2; FP64: Ts = (Re, Im) */
  __shared__ ST Tsh[F32 ? 2 * 343 : 343];
  R* const TA = reinterpret_cast<R*>(Tsh);
  R* const TB = TA + 2 * 343;
  __shared__ unsigned char sloc[27][64];
  const int v = had_block(n_groups, F / FB, per_xcd);
  if (v < 0) return; /* whole block, before any barrier */
  const int4 g = groups[v % n_groups];
  const int f0 = (v / n_groups) * FB;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ne8 = g.w * 8;
  int us[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int e = tid + 256 * k;
    us[k] = e < ne8 ? uslots[g.z + (e >> 3)] : 0;
  }
  ST pre[PER], pret[PERT];
  auto fetch = [&](int f) {
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int e = tid + 256 * k;
      pre[k] = e < ne8 ? X[((size_t)us[k] * F + f) * 8 + (e & 7)] : Spec<ST>::mk(0.0, 0.0);
    }
#pragma unroll
    for (int k = 0; k < PERT; k++) {
      const int e = tid + 256 * k;
      pret[k] = e < 343 ? T[(size_t)f * 343 + e] : Spec<ST>::mk(0.0, 0.0);
    }
  };
  fetch(f0);
  /* union indices of the block's parents, [neighbour][parent]; parents past
   * the group's count repeat its last parent */
  for (int e = tid; e < 27 * 64; e += 256) {
    const int l = e / 27, o = e - 27 * l;
    sloc[o][l] = loc[(size_t)(g.x + min(l, g.y - 1)) * 27 + o];
  }
  const int j = lane & 15, kk = lane >> 4;
  const int pj = 16 * wave + j;
  const bool active = 16 * wave < g.y; /* wave-uniform */
  /* row i = j of A: target child c = j & 7, real part for j < 8; the two
   * source children this lane supplies are kk and kk + 4 */
  const int c = j & 7;
  const bool rrow = j < 8;
  const int sel = rrow ? 0 : 1; /* rows 0-7: Re Y (A = Re T, -Im T); rows 8-15: Im Y (A = Im T, Re T) */
  auto base = [&](int cs) {
    return (3 + ((cs >> 2) & 1) - ((c >> 2) & 1)) + 7 * (3 + ((cs >> 1) & 1) - ((c >> 1) & 1)) +
           49 * (3 + (cs & 1) - (c & 1));
  };
  const int b0 = base(kk), b1 = base(kk + 4);
  /* FP32: this lane's A operands for source children kk and kk + 4, before
   * the direction offset */
  const R* const a0r = TA + 2 * b0 + sel;
  const R* const a1r = TA + 2 * b1 + sel;
  const int ys = yslot[g.x + min(pj, g.y - 1)];
  for (int fi = 0; fi < FB; fi++) {
    const int f = f0 + fi;
    __syncthreads(); /* the previous frequency's readers are done with Xs, Ts */
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int e = tid + 256 * k;
      if (e < ne8) Xs[e >> 3][e & 7] = pre[k];
    }
#pragma unroll
    for (int k = 0; k < PERT; k++) {
      const int e = tid + 256 * k;
      if (e < 343) {
        if constexpr (F32) {
          TA[2 * e] = pret[k].x;
          TA[2 * e + 1] = pret[k].y;
          TB[2 * e] = -pret[k].y;
          TB[2 * e + 1] = pret[k].x;
        } else {
          Tsh[e] = pret[k];
        }
      }
    }
    __syncthreads();
    if (fi + 1 < FB) fetch(f + 1);
    if (!active) continue;
    /* four independent accumulator chains: no MFMA waits on the one just
     * issued */
    V accr0 = {0, 0, 0, 0}, acci0 = {0, 0, 0, 0}, accr1 = {0, 0, 0, 0}, acci1 = {0, 0, 0, 0};
    if constexpr (!F32) {
#pragma unroll 2
      for (int o = 0; o < 27; o++) {
        if (o == 13) continue; /* the parent itself: its children are all adjacent */
        const int os = 2 * (o / 9 - 1) + 14 * ((o / 3) % 3 - 1) + 98 * (o % 3 - 1);
        const ST t0 = Tsh[b0 + os], t1 = Tsh[b1 + os];
        const ST* xs = Xs[sloc[o][pj]];
        const ST x0 = xs[kk], x1 = xs[kk + 4];
        accr0 = mfma4(rrow ? t0.x : t0.y, x0.x, accr0);
        acci0 = mfma4(rrow ? -t0.y : t0.x, x0.y, acci0);
        accr1 = mfma4(rrow ? t1.x : t1.y, x1.x, accr1);
        acci1 = mfma4(rrow ? -t1.y : t1.x, x1.y, acci1);
      }
    } else {
    /* the 26 directions (ox, oy, oz) in the order o = 9 (ox + 1) + 3 (oy +
     * 1) + oz + 1; only the z loop is unrolled: unrolling all of them would
     * hoist the LDS reads of many directions into registers and halve the
     * occupancy */
#pragma unroll 1
    for (int ox = -1; ox <= 1; ox++)
#pragma unroll 1
      for (int oy = -1; oy <= 1; oy++) {
        const int oxy = 9 * (ox + 1) + 3 * (oy + 1);
        const int osxy = 2 * (2 * ox + 14 * oy); /* in R elements */
#pragma unroll
        for (int oz = -1; oz <= 1; oz++) {
          if (ox == 0 && oy == 0 && oz == 0) continue; /* the parent itself: its children are all adjacent */
          const ST* xs = Xs[sloc[oxy + oz + 1][pj]];
          const ST x0 = xs[kk], x1 = xs[kk + 4];
          const R* const p0 = a0r + osxy + 2 * 98 * oz;
          const R* const p1 = a1r + osxy + 2 * 98 * oz;
          accr0 = mfma4(p0[0], x0.x, accr0);
          acci0 = mfma4(p0[2 * 343], x0.y, acci0);
          accr1 = mfma4(p1[0], x1.x, accr1);
          acci1 = mfma4(p1[2 * 343], x1.y, acci1);
        }
      }
    }
    const V acc = (accr0 + accr1) + (acci0 + acci1);
    ST* yb = Y + ((size_t)ys * F + f) * 8;
    if constexpr (sizeof(R) == 8) {
      /* D[i][j] in lane j + 16(i % 4), element i / 4: rows kk, kk+4 (Re)
       * and kk+8, kk+12 (Im) = children kk and kk + 4 */
      if (pj < g.y) {
        yb[kk] = Spec<ST>::mk(acc[0], acc[2]);
        yb[kk + 4] = Spec<ST>::mk(acc[1], acc[3]);
      }
    } else {
      /* D[i][j] in lane j + 16(i / 4), element i % 4: lanes kk < 2 hold Re
       * of children 4kk + e, lanes kk + 2 (the same column j) their Im */
      R im[4];
#pragma unroll
      for (int e = 0; e < 4; e++) im[e] = __shfl_xor(acc[e], 32, 64);
      if (pj < g.y && kk < 2)
#pragma unroll
        for (int e = 0; e < 4; e++) yb[4 * kk + e] = Spec<ST>::mk(acc[e], im[e]);
    }
  }
}

/* Inverse transform of the 8 sibling spectra of one child block, pruned to
 * the NP^3 cube holding the check surface; normalised by 1/N^3 and added
 * into the downward check values (compute_ifft_trg, fmm/omp_sse_block.cc:482-533:
 * FFTW c2r semantics, the kx = 0 and kx = N/2 imaginary parts are ignored).
 * The z and y passes give a thread one column (inputs read once, all NP
 * outputs with the twiddles in registers, unrolled); the LDS planes are
 * octant-fastest.  Per output the sum runs over the input index in
 * increasing order. */
template <int NP, typename ST = double2>
__global__ __launch_bounds__(512) void k_dft_inv(const int* __restrict__ blocks, const int* __restrict__ yslot,
                                                 const ST* __restrict__ Y, const double2* __restrict__ tw,
                                                 const int* __restrict__ surf2cube, /* NE: (ix*NP+iy)*NP+iz */
                                                 double* __restrict__ dchk, int ld_e) {
  using G = Grid<NP>;
  constexpr int N = G::N, H = G::H, F = G::F, NE = G::NE, OFF = G::OFF;
  __shared__ double2 B[NP][N][H][8];  /* [iz][ky][kx][oct] */
  __shared__ double2 A[NP][NP][H][8]; /* [iz][iy][kx][oct] */
  __shared__ double2 ws[N];           /* for the x pass, whose twiddle index is not a constant */
  const int tid = threadIdx.x;
  const int first = blocks[blockIdx.x];
  double2 w[N];
#pragma unroll
  for (int m = 0; m < N; m++) w[m] = tw[m];
  if (tid < N) ws[tid] = tw[tid];
  const ST* in = Y + (size_t)yslot[blockIdx.x] * F * 8;
  /* z pass: column (ky, kx, oct) = thread e, reading Y[slot][kz * N * H * 8 + e] */
  for (int e = tid; e < N * H * 8; e += blockDim.x) {
    const int o = e & 7, kk = e >> 3; /* kk = ky*H + kx */
    double re[NP], im[NP];
#pragma unroll
    for (int iz = 0; iz < NP; iz++) { re[iz] = 0.0; im[iz] = 0.0; }
#pragma unroll
    for (int kz = 0; kz < N; kz++) {
      const ST vs = in[(size_t)kz * N * H * 8 + e];
      const double2 v = make_double2(vs.x, vs.y);
#pragma unroll
      for (int iz = 0; iz < NP; iz++) {
        const double2 t = w[(kz * (iz + OFF)) % N]; /* conj applied below */
        re[iz] = __builtin_fma(v.x, t.x, __builtin_fma(v.y, t.y, re[iz]));
        im[iz] = __builtin_fma(v.y, t.x, __builtin_fma(-v.x, t.y, im[iz]));
      }
    }
    const int ky = kk / H, kx = kk % H;
#pragma unroll
    for (int iz = 0; iz < NP; iz++) B[iz][ky][kx][o] = make_double2(re[iz], im[iz]);
  }
  __syncthreads();
  /* y pass: column (iz, kx, oct) */
  for (int e = tid; e < NP * H * 8; e += blockDim.x) {
    const int o = e & 7, kx = (e >> 3) % H, iz = (e >> 3) / H;
    double2 v[N];
#pragma unroll
    for (int ky = 0; ky < N; ky++) v[ky] = B[iz][ky][kx][o];
#pragma unroll
    for (int iy = 0; iy < NP; iy++) {
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int ky = 0; ky < N; ky++) {
        const double2 t = w[(ky * (iy + OFF)) % N];
        re = __builtin_fma(v[ky].x, t.x, __builtin_fma(v[ky].y, t.y, re));
        im = __builtin_fma(v[ky].y, t.x, __builtin_fma(-v[ky].x, t.y, im));
      }
      A[iz][iy][kx][o] = make_double2(re, im);
    }
  }
  __syncthreads();
  /* x pass (c2r) at the surface samples */
  const double nrm = 1.0 / (double)(N * N * N);
  for (int e = tid; e < 8 * NE; e += blockDim.x) {
    const int o = e / NE, j = e % NE;
    const int cube = surf2cube[j];
    const int ix = cube / (NP * NP), iy = (cube / NP) % NP, iz = cube % NP;
    const double2 (*r)[8] = A[iz][iy];
    const int xg = ix + OFF;
    double s = r[0][o].x + ((xg & 1) ? -r[N / 2][o].x : r[N / 2][o].x);
    double acc = 0.0;
#pragma unroll
    for (int kx = 1; kx < N / 2; kx++) {
      const double2 t = ws[(kx * xg) % N];
      acc = __builtin_fma(r[kx][o].x, t.x, __builtin_fma(r[kx][o].y, t.y, acc));
    }
    s = __builtin_fma(2.0, acc, s);
    dchk[(size_t)(first + o) * ld_e + j] += nrm * s;
  }
}

/* ------------------------------------------------- W, X and L2T evaluation */

/* W- and X-list evaluation (wlist_eval__gpu, fmm/cuda.cu:2070-2197;
 * xlist_eval__gpu, fmm/cuda.cu:2380-2493): one wave per item, a lane per
 * target point, the point in registers for the whole item and its sum in
 * four lane-private accumulators (an item of at most 32 or 16 points runs
 * 2 or 4 source phases, one per lane group, added by a lane shuffle at the
 * end).  Item = two int4:
 *   a = {node t, first point, count <= 64, kind}: kind 0 the targets of
 *       leaf t (trg4 from tbeg[t]), kind 1 the DC check surface c + r SP of
 *       node t, results into dchk (X-list nodes with n_e or more targets);
 *   b = {first entry, end entry, partial slot or -1, 0}: a chunk of t's
 *       list (lidx), cut at entry boundaries.
 * The sources of entry s are leaf points (every X-list entry; W-list leaves
 * with fewer than n_e sources) or the UE surface c_s + r_s SP with the
 * densities uden[s] (other W-list nodes), staged WX_TILE at a time in LDS
 * and read as wave-wide broadcasts, the next tile loaded into registers
 * while the current one is summed.  A list longer than one chunk is split:
 * its chunks write their sums to part[slot][lane] and k_wx_reduce adds them
 * to the output in chunk order (deterministic). */
constexpr int WX_TILE = 256;
constexpr int WX_PER = WX_TILE / 64;
constexpr int WX_CHUNK_SRC = 1024; /* sources per chunk (an entry is never split) */

struct WxWalk {
  int k, kend, base; /* entry, end of the chunk, offset within the entry */
};

/* Loads the next tile (up to WX_TILE sources of one entry, padded with
 * zero-weight far points) into registers; returns its count, 0 at the end. */
template <bool XL>
__device__ __forceinline__ int wx_fetch(WxWalk& w, const int* __restrict__ lidx, const int* __restrict__ child,
                                        const int* __restrict__ snum, const int* __restrict__ sbeg,
                                        const double4* __restrict__ geo, const double* __restrict__ surf_ue,
                                        int ld_surf, int n_e, const double* __restrict__ uden, int ld_e,
                                        const double4* __restrict__ src4, int lane, double4 (&pre)[WX_PER]) {
  int s = 0, n = 0;
  bool direct = true;
  for (; w.k < w.kend; w.k++, w.base = 0) { /* skips entries without sources */
    s = lidx[w.k];
    direct = XL || (child[s] == -1 && snum[s] < n_e);
    n = direct ? snum[s] : n_e;
    if (w.base < n) break;
  }
  if (w.k >= w.kend) return 0;
  const int cnt = min(WX_TILE, n - w.base);
  if (direct) {
    const double4* __restrict__ p = src4 + sbeg[s] + w.base;
#pragma unroll
    for (int m = 0; m < WX_PER; m++) {
      const int i = lane + 64 * m;
      pre[m] = i < cnt ? p[i] : make_double4(FAR, FAR, FAR, 0.0);
    }
  } else {
    const double4 g = geo[s];
    const double* __restrict__ d = uden + (size_t)s * ld_e + w.base;
#pragma unroll
    for (int m = 0; m < WX_PER; m++) {
      const int i = lane + 64 * m, j = w.base + i;
      pre[m] = i < cnt ? make_double4(g.x + g.w * surf_ue[j], g.y + g.w * surf_ue[ld_surf + j],
                                      g.z + g.w * surf_ue[2 * ld_surf + j], d[i])
                       : make_double4(FAR, FAR, FAR, 0.0);
    }
  }
  w.base += cnt;
  if (w.base >= n) {
    w.k++;
    w.base = 0;
  }
  return cnt;
}

/* acc[u] += 2 w / |x - p| over the staged sources j0 + u, j0 = first,
 * first + stride, ... < n4 (n4, first and stride multiples of 4) */
template <bool MASK>
__device__ __forceinline__ void wx_sum_tile(const double4* tile, int n4, int first, int stride, double tx, double ty,
                                            double tz, double (&acc)[4]) {
  for (int j = first; j < n4; j += stride) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const double4 p = tile[j + u];
      const double dx = tx - p.x, dy = ty - p.y, dz = tz - p.z;
      const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
      acc[u] = __builtin_fma(p.w, inv_r2<MASK>(r2), acc[u]);
    }
  }
}

/* XL = false: W-list (lidx = w_idx), XL = true: X-list (lidx = x_idx) */
template <bool MASK, bool XL>
__global__ __launch_bounds__(64) void k_wx(const int4* __restrict__ items, const int* __restrict__ lidx,
                                           const int* __restrict__ child, const int* __restrict__ snum,
                                           const int* __restrict__ sbeg, const int* __restrict__ tbeg,
                                           const double4* __restrict__ geo, const double* __restrict__ surf_ue,
                                           const double* __restrict__ surf_dc, int ld_surf, int n_e,
                                           const double* __restrict__ uden, int ld_e,
                                           const double4* __restrict__ src4, const double4* __restrict__ trg4,
                                           double* __restrict__ pot, double* __restrict__ dchk,
                                           double* __restrict__ part) {
  __shared__ double4 tile[WX_TILE];
  const int4 a = items[2 * blockIdx.x], b = items[2 * blockIdx.x + 1];
  const int lane = threadIdx.x;
  /* an item of at most 32 (16) points splits the wave into 2 (4) source
   * phases of 64 / G lanes: lane = phase * (64 / G) + point, phase ph takes
   * the source quads 4 (ph + G i); the phases' sums meet in a lane shuffle */
  const int lg = a.z <= 16 ? 2 : a.z <= 32 ? 1 : 0; /* log2 G, item-uniform */
  const int tl = lane & ((64 >> lg) - 1), ph = lane >> (6 - lg);
  const int pt = a.y + min(tl, a.z - 1);
  double tx, ty, tz;
  if (a.w) {
    const double4 g = geo[a.x];
    tx = g.x + g.w * surf_dc[pt];
    ty = g.y + g.w * surf_dc[ld_surf + pt];
    tz = g.z + g.w * surf_dc[2 * ld_surf + pt];
  } else {
    const double4 q = trg4[tbeg[a.x] + pt];
    tx = q.x; ty = q.y; tz = q.z;
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  WxWalk w{b.x, b.y, 0};
  double4 pre[WX_PER];
  int cnt = wx_fetch<XL>(w, lidx, child, snum, sbeg, geo, surf_ue, ld_surf, n_e, uden, ld_e, src4, lane, pre);
  while (cnt > 0) {
    __syncthreads(); /* the previous tile's reads are done (one wave: an LDS ordering point) */
#pragma unroll
    for (int m = 0; m < WX_PER; m++) tile[lane + 64 * m] = pre[m];
    __syncthreads();
    const int next = wx_fetch<XL>(w, lidx, child, snum, sbeg, geo, surf_ue, ld_surf, n_e, uden, ld_e, src4, lane, pre);
    if (a.w) wx_sum_tile<false>(tile, (cnt + 3) & ~3, 4 * ph, 4 << lg, tx, ty, tz, acc);
    else wx_sum_tile<MASK>(tile, (cnt + 3) & ~3, 4 * ph, 4 << lg, tx, ty, tz, acc);
    cnt = next;
  }
  double sum = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  for (int m = 64 >> lg; m < 64; m <<= 1) sum += __shfl_xor(sum, m, 64);
  if (ph == 0 && tl < a.z) {
    if (b.z >= 0) part[(size_t)b.z * 64 + tl] = sum;
    else if (a.w) dchk[(size_t)a.x * ld_e + a.y + tl] += HALF_OOFP_R * sum;
    else pot[tbeg[a.x] + a.y + tl] += HALF_OOFP_R * sum;
  }
}

/* Adds the chunk sums of a split W / X item to its output, in chunk order.
 * Item = {node t, first point, count, kind} and {first slot, slots, 0, 0}. */
static __global__ __launch_bounds__(64) void k_wx_reduce(const int4* __restrict__ items,
                                                        const int* __restrict__ tbeg, int ld_e,
                                                        const double* __restrict__ part,
                                                        double* __restrict__ pot, double* __restrict__ dchk) {
  const int4 a = items[2 * blockIdx.x], b = items[2 * blockIdx.x + 1];
  const int lane = threadIdx.x;
  if (lane >= a.z) return;
  double s = 0.0;
  for (int c = 0; c < b.y; c++) s += part[(size_t)(b.x + c) * 64 + lane];
  if (a.w) dchk[(size_t)a.x * ld_e + a.y + lane] += HALF_OOFP_R * s;
  else pot[tbeg[a.x] + a.y + lane] += HALF_OOFP_R * s;
}
/* L2T (down_eval__gpu_leaves, fmm/cuda.cu:1804-1872): item = {leaf, first
 * target, count <= 64}, one wave; lane = target, the leaf's DE surface
 * points c + r * SP_DE[j] with densities dden[leaf][j] are wave-uniform. */
static __global__ __launch_bounds__(64) void k_l2t(const int4* __restrict__ items,
                                                   const double4* __restrict__ geo,
                                                   const double* __restrict__ surf_de, int ld_surf, int n_e,
                                                   const double* __restrict__ dden, int ld_e,
                                                   const double4* __restrict__ trg4, double* __restrict__ pot) {
  /* in leaf units (t - c) / h the DE surface is the same unit-box surface
   * for every leaf, read as wave-uniform values; the sum is scaled by 1/h
   * (h a power of two: exact) */
  const int4 it = items[blockIdx.x];
  const int k = threadIdx.x;
  const double4 q = trg4[it.y + min(k, it.z - 1)];
  const double4 g = geo[it.x];
  const double ih = 1.0 / g.w;
  const double tx = (q.x - g.x) * ih, ty = (q.y - g.y) * ih, tz = (q.z - g.z) * ih;
  const double ax = -2.0 * tx, ay = -2.0 * ty, az = -2.0 * tz, tt = __builtin_fma(tx, tx, __builtin_fma(ty, ty, tz * tz));
  const double* __restrict__ den = dden + (size_t)it.x * ld_e;
  /* Gram form (surface points >= 1.27 leaf units from the leaf's targets):
   * r^2 = |t|^2 + |p|^2 - 2 t.p, |p|^2 from the surface table's row 3 */
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  int j = 0;
  for (; j + 4 <= n_e; j += 4) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const double r2 = __builtin_fma(surf_de[j + u], ax, __builtin_fma(surf_de[ld_surf + j + u], ay,
                                      __builtin_fma(surf_de[2 * ld_surf + j + u], az, tt + surf_de[3 * ld_surf + j + u])));
      acc[u] = __builtin_fma(den[j + u], inv_r2<false>(r2), acc[u]);
    }
  }
  for (; j < n_e; j++) {
    const double r2 = __builtin_fma(surf_de[j], ax, __builtin_fma(surf_de[ld_surf + j], ay,
                                    __builtin_fma(surf_de[2 * ld_surf + j], az, tt + surf_de[3 * ld_surf + j])));
    acc[0] = __builtin_fma(den[j], inv_r2<false>(r2), acc[0]);
  }
  if (k < it.z) pot[it.y + k] += HALF_OOFP_R * ih * ((acc[0] + acc[1]) + (acc[2] + acc[3]));
}

/* Sampled direct sum for validation (validate, fmm/driver.cc:141-206, on
 * the GPU): one block per target point, 256 lanes stride over all sources,
 * FP64 w / sqrt(r^2) (IEEE division and square root, r = 0 excluded) with
 * Neumaier-compensated lane sums, combined by lane 0 in lane order (also
 * compensated): the result is independent of the source order to about
 * one rounding. */
__device__ __forceinline__ void neumaier(double& s, double& c, double v) {
  const double t = s + v;
  c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
  s = t;
}

static __global__ __launch_bounds__(256) void k_direct(int nt, const double* __restrict__ tx,
                                                       const double* __restrict__ ty, const double* __restrict__ tz,
                                                       const double4* __restrict__ src4, int n,
                                                       double* __restrict__ out) {
  __shared__ double ss[256], sc[256];
  const int tid = threadIdx.x;
  for (int t = blockIdx.x; t < nt; t += gridDim.x) {
    const double x = tx[t], y = ty[t], z = tz[t];
    double s = 0.0, c = 0.0;
    for (int j = tid; j < n; j += 256) {
      const double4 p = src4[j];
      const double dx = x - p.x, dy = y - p.y, dz = z - p.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 != 0.0) neumaier(s, c, p.w / sqrt(r2));
    }
    ss[tid] = s;
    sc[tid] = c;
    __syncthreads();
    if (tid == 0) {
      double S = 0.0, Cc = 0.0;
      for (int k = 0; k < 256; k++) {
        neumaier(S, Cc, ss[k]);
        Cc += sc[k];
      }
      out[t] = OOFP_R * (S + Cc);
    }
    __syncthreads();
  }
}

static __global__ void k_fill(double* __restrict__ p, size_t n, double v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

/* pot = u + pot over [0, n): the U-list sums (u) written on the second
 * stream joined with the far field (W, X, L2T) accumulated in pot, in the
 * same operand order as the U-list kernels' own accumulating write */
static __global__ void k_join_pot(double* __restrict__ pot, const double* __restrict__ u, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) pot[i] = u[i] + pot[i];
}

/* (x, y, z, w) records from coordinate arrays; w = 0 without a density array */
static __global__ void k_pack4(double4* __restrict__ out, const double* __restrict__ x, const double* __restrict__ y,
                               const double* __restrict__ z, const double* __restrict__ w, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = make_double4(x[i], y[i], z[i], w ? w[i] : 0.0);
}

/* new densities (leaf order) into the w field of the source records */
static __global__ void k_set_w(double4* __restrict__ src4, const double* __restrict__ w, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) src4[i].w = w[i];
}

}  // namespace kifmm_dev
