/*
 * kifmm_oracle.cc -- CPU oracle for the KIFMM evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY: a restatement of the reference CPU path used
 * as the checker by tests/, the smoke check and bench.py's cpu_baseline.
 * The product library never links or calls it.
 *
 * What is restated, and from where (reference file:line):
 *   generators ............ fmm/input.cc:29-84, fmm/driver.cc:244-267
 *   source octree ......... fmm/input.cc:203-298 (BFS level order, split while srcNum > q)
 *   target binning ........ fmm/input.cc:386-460
 *   U/V/W/X lists ......... fmm/input.cc:303-381 (sorted, SRCNODE-filtered)
 *   gather ................ fmm/evaluate.cc:13-96
 *   surfaces / grid ....... fmm/trans.cc:13-88
 *   operators + pinv ...... fmm/trans.cc:143-264, fmm/vecmatop.cc:34-78, fmm/svdrep.cc:10-66
 *                           (the SVD is a one-sided Jacobi written here; MKL DGESVD is absent)
 *   FFT (r2c/c2r) ......... FFTW layout [z][y][x/2+1], unnormalised inverse
 *                           (fmm/trans.cc:271-416); separable DFT written here
 *   evaluation ............ fmm/omp_sse_block.cc:60-887, fmm/evaluate--basic.cc:30-87
 *   validate .............. fmm/driver.cc:141-206
 *
 * Deliberate differences from the reference, each a documented defect there:
 *   - r == 0 pairs are skipped (fmm/omp_sse.cc:294-295); omp_sse_block
 *     returns NaN for them (fmm/omp_sse_block.cc:297).
 *   - every accumulator is zeroed (fmm/driver.cc:318-322 allocates without zeroing).
 *   - the up pass walks every level bottom-up; the reference hard-codes
 *     level 3 as the end of its UPBS-blocked part (fmm/omp_sse_block.cc:209,225),
 *     which gives the same result for trees of depth >= 2.
 */
#include "kifmm_oracle.h"

#include <emmintrin.h>
#include <xmmintrin.h>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <queue>
#include <set>
#include <string>
#include <vector>

namespace {

const double OOFP_R = 1.0 / (4.0 * M_PI);
enum { SRCNODE = 1, TRGNODE = 2 };
enum { UE = 0, UC = 1, DE = 2, DC = 3 };

struct Node {
  int parent, child, depth, tag;
  int path[3];
  int srcBeg = 0, srcNum = 0, trgBeg = 0, trgNum = 0;
  std::vector<int> srcIdx, trgIdx, U, V, W, X;
  Node(int p, int c, int px, int py, int pz, int d) : parent(p), child(c), depth(d), tag(0) {
    path[0] = px; path[1] = py; path[2] = pz;
  }
};

struct Surf {
  std::vector<double> x, y, z;
  int n = 0;
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/* ----------------------------------------------------------------- geometry */

/* fmm/input.cc:113-137: centre of a node, exact dyadic arithmetic */
void node_center(const Node& n, double c[3]) {
  const double tmp = (double)(1 << n.depth);
  for (int d = 0; d < 3; d++) c[d] = -1.0 + (2.0 * 1.0) * (n.path[d] + 0.5) / tmp;
}
double node_radius(const Node& n) { return 1.0 / (double)(1 << n.depth); }

/* fmm/trans.cc:31-62 */
Surf sampos(int np, double R) {
  Surf s;
  s.n = np * np * np - (np - 2) * (np - 2) * (np - 2);
  s.x.resize(s.n); s.y.resize(s.n); s.z.resize(s.n);
  const double step = 2.0 / (np - 1);
  int cnt = 0;
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++)
        if (i == 0 || i == np - 1 || j == 0 || j == np - 1 || k == 0 || k == np - 1) {
          s.x[cnt] = R * (-1.0 + i * step);
          s.y[cnt] = R * (-1.0 + j * step);
          s.z[cnt] = R * (-1.0 + k * step);
          cnt++;
        }
  return s;
}

/* fmm/trans.cc:64-88 */
Surf regpos(int np, double R) {
  Surf s;
  const int N = 2 * np;
  s.n = N * N * N;
  s.x.resize(s.n); s.y.resize(s.n); s.z.resize(s.n);
  const double step = 2.0 / (np - 1);
  int cnt = 0;
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        int gi = (i < np) ? i : i - N, gj = (j < np) ? j : j - N, gk = (k < np) ? k : k - N;
        s.x[cnt] = R * gi * step;
        s.y[cnt] = R * gj * step;
        s.z[cnt] = R * gk * step;
        cnt++;
      }
  return s;
}

/* grid offsets of the surface samples (fmm/trans.cc:335-355,394-416) */
std::vector<int> sample_grid_offsets(int np) {
  std::vector<int> off;
  const int N = 2 * np;
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++)
        if (i == 0 || i == np - 1 || j == 0 || j == np - 1 || k == 0 || k == np - 1)
          off.push_back((k + np / 2) * N * N + (j + np / 2) * N + (i + np / 2));
  return off;
}

/* ------------------------------------------------------------- the kernel */

/* When set, every exact-kernel source sum runs in reverse order: a
 * summation-order perturbation used to measure how much two correct FP64
 * evaluations can differ (the conditioning floor of the pinv applies). */
int g_reverse_order = 0;

/* ulist__direct_evaluation (fmm/omp_sse_block.cc:266-327), pot += 1/4pi sum w/r */
void direct_eval(int nt, const double* tx, const double* ty, const double* tz, double* pot,
                 int ns, const double* sx, const double* sy, const double* sz,
                 const double* w, int mode) {
  int i = 0;
  if (mode == ORACLE_KERNEL_REF) {
    const __m128d oofp = _mm_set1_pd(OOFP_R);
    for (; i + 2 <= nt; i += 2) {
      __m128d TX = _mm_loadu_pd(tx + i), TY = _mm_loadu_pd(ty + i), TZ = _mm_loadu_pd(tz + i);
      __m128d TD = _mm_setzero_pd();
      for (int j = 0; j < ns; j++) {
        __m128d SD = _mm_load1_pd(w + j);
        __m128d DX = _mm_sub_pd(TX, _mm_load1_pd(sx + j));
        __m128d DY = _mm_sub_pd(TY, _mm_load1_pd(sy + j));
        __m128d DZ = _mm_sub_pd(TZ, _mm_load1_pd(sz + j));
        __m128d R = _mm_mul_pd(DX, DX);
        R = _mm_add_pd(R, _mm_mul_pd(DY, DY));
        R = _mm_add_pd(R, _mm_mul_pd(DZ, DZ));
        __m128d zero = _mm_cmpeq_pd(R, _mm_setzero_pd());
        __m128d xh = _mm_mul_pd(_mm_set1_pd(0.5), R);
        /* rsqrtps of the FP32 conversion, then one FP64 Newton step */
        R = _mm_cvtps_pd(_mm_rsqrt_ps(_mm_cvtpd_ps(R)));
        __m128d rnr = _mm_mul_pd(R, R);
        rnr = _mm_mul_pd(rnr, xh);
        rnr = _mm_sub_pd(_mm_set1_pd(1.5), rnr);
        R = _mm_mul_pd(rnr, R);
        R = _mm_andnot_pd(zero, R);
        TD = _mm_add_pd(TD, _mm_mul_pd(SD, R));
      }
      TD = _mm_mul_pd(TD, oofp);
      TD = _mm_add_pd(TD, _mm_loadu_pd(pot + i));
      _mm_storeu_pd(pot + i, TD);
    }
  }
  for (; i < nt; i++) {
    double td = 0.0;
    for (int jj = 0; jj < ns; jj++) {
      const int j = g_reverse_order ? ns - 1 - jj : jj;
      double x = tx[i] - sx[j], y = ty[i] - sy[j], z = tz[i] - sz[j];
      double r2 = x * x + y * y + z * z;
      if (r2 != 0.0) td += w[j] / std::sqrt(r2);
    }
    pot[i] += OOFP_R * td;
  }
}

/* kernel() (fmm/omp_sse_block.cc:329-382): mat[i*tn+j] = 1/(4 pi |x_i - y_j|) */
void kernel_mat(int sn, int tn, const double* x1, const double* x2, const double* x3,
                const double* y1, const double* y2, const double* y3, double* mat, int mode) {
  for (int i = 0; i < sn; i++) {
    int j = 0;
    if (mode == ORACLE_KERNEL_REF) {
      __m128d SX = _mm_load1_pd(x1 + i), SY = _mm_load1_pd(x2 + i), SZ = _mm_load1_pd(x3 + i);
      for (; j + 2 <= tn; j += 2) {
        __m128d DX = _mm_sub_pd(SX, _mm_loadu_pd(y1 + j));
        __m128d DY = _mm_sub_pd(SY, _mm_loadu_pd(y2 + j));
        __m128d DZ = _mm_sub_pd(SZ, _mm_loadu_pd(y3 + j));
        __m128d R = _mm_mul_pd(DX, DX);
        R = _mm_add_pd(R, _mm_mul_pd(DY, DY));
        R = _mm_add_pd(R, _mm_mul_pd(DZ, DZ));
        __m128d xh = _mm_mul_pd(_mm_set1_pd(0.5), R);
        R = _mm_cvtps_pd(_mm_rsqrt_ps(_mm_cvtpd_ps(R)));
        __m128d rnr = _mm_mul_pd(R, R);
        rnr = _mm_mul_pd(rnr, xh);
        rnr = _mm_sub_pd(_mm_set1_pd(1.5), rnr);
        R = _mm_mul_pd(rnr, R);
        R = _mm_mul_pd(R, _mm_set1_pd(OOFP_R));
        _mm_storeu_pd(mat + (size_t)i * tn + j, R);
      }
    }
    for (; j < tn; j++) {
      double dx = x1[i] - y1[j], dy = x2[i] - y2[j], dz = x3[i] - y3[j];
      double dr = std::sqrt(dx * dx + dy * dy + dz * dz);
      mat[(size_t)i * tn + j] = (dr != 0.0) ? OOFP_R / dr : 0.0;
    }
  }
}

/* y (+)= A x, A column-major [m x n] (MKL dgemv 'N' as used in fmm/vecmatop.cc:15-22) */
void gemv(int m, int n, const double* A, const double* x, double* y, bool accumulate) {
  std::vector<double> t(m, 0.0);
  for (int j = 0; j < n; j++) {
    const double xj = x[j];
    const double* a = A + (size_t)j * m;
    for (int i = 0; i < m; i++) t[i] += a[i] * xj;
  }
  for (int i = 0; i < m; i++) y[i] = accumulate ? y[i] + t[i] : t[i];
}

/* ------------------------------------------------------------------- SVD */

/* One-sided (Hestenes) Jacobi SVD of M [m x n] column-major, m >= n.
 * Returns singular values descending, U [m x n], V [n x n]. */
void jacobi_svd(const double* M, int m, int n, std::vector<double>& S, std::vector<double>& U,
                std::vector<double>& V) {
  std::vector<double> A(M, M + (size_t)m * n);
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) V[(size_t)i * n + i] = 1.0;
  const double tol = 1e-15;
  for (int sweep = 0; sweep < 80; sweep++) {
    int rotations = 0;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double* ap = &A[(size_t)p * m];
        double* aq = &A[(size_t)q * m];
        double alpha = 0, beta = 0, gamma = 0;
        for (int i = 0; i < m; i++) {
          alpha += ap[i] * ap[i];
          beta += aq[i] * aq[i];
          gamma += ap[i] * aq[i];
        }
        if (std::fabs(gamma) <= tol * std::sqrt(alpha * beta) || gamma == 0.0) continue;
        rotations++;
        double zeta = (beta - alpha) / (2.0 * gamma);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        for (int i = 0; i < m; i++) {
          double a = ap[i], b = aq[i];
          ap[i] = c * a - s * b;
          aq[i] = s * a + c * b;
        }
        double* vp = &V[(size_t)p * n];
        double* vq = &V[(size_t)q * n];
        for (int i = 0; i < n; i++) {
          double a = vp[i], b = vq[i];
          vp[i] = c * a - s * b;
          vq[i] = s * a + c * b;
        }
      }
    if (rotations == 0) break;
  }
  std::vector<double> s(n);
  for (int j = 0; j < n; j++) {
    double nrm = 0;
    for (int i = 0; i < m; i++) nrm += A[(size_t)j * m + i] * A[(size_t)j * m + i];
    s[j] = std::sqrt(nrm);
  }
  std::vector<int> order(n);
  for (int j = 0; j < n; j++) order[j] = j;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return s[a] > s[b]; });
  S.resize(n);
  U.assign((size_t)m * n, 0.0);
  std::vector<double> V2((size_t)n * n);
  for (int jj = 0; jj < n; jj++) {
    int j = order[jj];
    S[jj] = s[j];
    for (int i = 0; i < m; i++) U[(size_t)jj * m + i] = s[j] > 0 ? A[(size_t)j * m + i] / s[j] : 0.0;
    for (int i = 0; i < n; i++) V2[(size_t)jj * n + i] = V[(size_t)j * n + i];
  }
  V.swap(V2);
}

/* pinv with relative cutoff eps*sigma_0 keeping sigma >= cutoff
 * (fmm/svdrep.cc:21-39, fmm/vecmatop.cc:34-78).  R is [n x m] column-major. */
void pinv(const double* M, int m, int n, double eps, double* R) {
  std::vector<double> S, U, V;
  if (m >= n) {
    jacobi_svd(M, m, n, S, U, V);
  } else {
    std::vector<double> MT((size_t)m * n);
    for (int i = 0; i < m; i++)
      for (int j = 0; j < n; j++) MT[(size_t)i * n + j] = M[(size_t)j * m + i];
    std::vector<double> S2, U2, V2;
    jacobi_svd(MT.data(), n, m, S2, U2, V2);
    /* M^T = U2 S V2^T  =>  M = V2 S U2^T */
    S = S2; U = V2; V = U2;
    int k = m;
    std::vector<double> Ut((size_t)m * k), Vt((size_t)n * k);
    for (int j = 0; j < k; j++) {
      for (int i = 0; i < m; i++) Ut[(size_t)j * m + i] = U[(size_t)j * m + i];
      for (int i = 0; i < n; i++) Vt[(size_t)j * n + i] = V[(size_t)j * n + i];
    }
    U.swap(Ut); V.swap(Vt);
  }
  const int k = std::min(m, n);
  const double cutoff = eps * S[0];
  int r = 0;
  while (r < k && std::fabs(S[r]) >= cutoff) r++;
  for (size_t i = 0; i < (size_t)n * m; i++) R[i] = 0.0;
  /* R = V_r S_r^-1 U_r^T, R[a + b*n] = sum_l V[a,l] U[b,l] / S[l] */
  for (int l = 0; l < r; l++) {
    const double inv = 1.0 / S[l];
    const double* v = &V[(size_t)l * n];
    const double* u = &U[(size_t)l * m];
    for (int b = 0; b < m; b++) {
      const double ub = u[b] * inv;
      double* col = R + (size_t)b * n;
      for (int a = 0; a < n; a++) col[a] += v[a] * ub;
    }
  }
}

/* ------------------------------------------------------------------- FFT */

struct Dft {
  int N = 0;
  std::vector<std::complex<double>> w; /* w[m] = exp(-2 pi i m / N) */
  explicit Dft(int n) : N(n), w(n) {
    for (int m = 0; m < n; m++) {
      long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
      w[m] = std::complex<double>((double)cosl(a), (double)(-sinl(a)));
    }
  }
};

/* r2c: in[z][y][x] (x fastest) -> out[z][y][kx], kx in [0,N/2], interleaved complex */
void rfft3(const Dft& F, const double* in, std::complex<double>* out) {
  const int N = F.N, H = N / 2 + 1;
  std::vector<std::complex<double>> a((size_t)N * N * H), b((size_t)N * N * H);
  for (int z = 0; z < N; z++)
    for (int y = 0; y < N; y++)
      for (int kx = 0; kx < H; kx++) {
        std::complex<double> s = 0;
        for (int x = 0; x < N; x++) s += in[(z * N + y) * N + x] * F.w[(kx * x) % N];
        a[(z * N + y) * H + kx] = s;
      }
  for (int z = 0; z < N; z++)
    for (int ky = 0; ky < N; ky++)
      for (int kx = 0; kx < H; kx++) {
        std::complex<double> s = 0;
        for (int y = 0; y < N; y++) s += a[(z * N + y) * H + kx] * F.w[(ky * y) % N];
        b[(z * N + ky) * H + kx] = s;
      }
  for (int kz = 0; kz < N; kz++)
    for (int ky = 0; ky < N; ky++)
      for (int kx = 0; kx < H; kx++) {
        std::complex<double> s = 0;
        for (int z = 0; z < N; z++) s += b[(z * N + ky) * H + kx] * F.w[(kz * z) % N];
        out[(kz * N + ky) * H + kx] = s;
      }
}

/* c2r, unnormalised (FFTW convention: imaginary parts of the kx = 0 and
 * kx = N/2 planes are dropped by the last-dimension c2r) */
void irfft3(const Dft& F, const std::complex<double>* in, double* out) {
  const int N = F.N, H = N / 2 + 1;
  std::vector<std::complex<double>> a((size_t)N * N * H), b((size_t)N * N * H);
  for (int z = 0; z < N; z++)
    for (int ky = 0; ky < N; ky++)
      for (int kx = 0; kx < H; kx++) {
        std::complex<double> s = 0;
        for (int kz = 0; kz < N; kz++) s += in[(kz * N + ky) * H + kx] * std::conj(F.w[(kz * z) % N]);
        a[(z * N + ky) * H + kx] = s;
      }
  for (int z = 0; z < N; z++)
    for (int y = 0; y < N; y++)
      for (int kx = 0; kx < H; kx++) {
        std::complex<double> s = 0;
        for (int ky = 0; ky < N; ky++) s += a[(z * N + ky) * H + kx] * std::conj(F.w[(ky * y) % N]);
        b[(z * N + y) * H + kx] = s;
      }
  for (int z = 0; z < N; z++)
    for (int y = 0; y < N; y++)
      for (int x = 0; x < N; x++) {
        const std::complex<double>* r = &b[(z * N + y) * H];
        double s = r[0].real() + ((x & 1) ? -r[N / 2].real() : r[N / 2].real());
        for (int kx = 1; kx < N / 2; kx++) s += 2.0 * (r[kx] * std::conj(F.w[(kx * x) % N])).real();
        out[(z * N + y) * N + x] = s;
      }
}

}  // namespace

/* ------------------------------------------------------------- the problem */

struct oracle_problem {
  int n = 0, np = 0, q = 0;
  unsigned short rng[3];
  std::vector<double> sx, sy, sz, den, tx, ty, tz; /* original order */
  std::vector<Node> nodes;
  std::vector<int> level; /* nodeLevel */
  std::vector<double> gsx, gsy, gsz, gden, gtx, gty, gtz, gpot;
  std::vector<int> src_perm, trg_perm;
  Surf SP[4], RP;
  int n_e = 0, n_uc = 0;
  bool have_ops = false;
  std::vector<double> uc2ue, ue2uc, dc2de, de2dc;
  std::vector<double> uden, dchk, dden;
};

namespace {

int getbits(int v, int b) { return (v >> b) & 1; }
int child_of(const std::vector<Node>& nv, int id, int a, int b, int c) { return nv[id].child + a * 4 + b * 2 + c; }

/* fmm/input.cc:152-165 */
int findgnt(const std::vector<Node>& nv, int wntdepth, const int wntpath[3]) {
  int cur = 0;
  int left[3] = {wntpath[0], wntpath[1], wntpath[2]};
  while (nv[cur].depth < wntdepth && nv[cur].child != -1) {
    int dif = wntdepth - nv[cur].depth;
    int tmp = 1 << (dif - 1);
    int ch[3];
    for (int d = 0; d < 3; d++) {
      ch[d] = left[d] / tmp;
      left[d] -= ch[d] * tmp;
    }
    cur = child_of(nv, cur, ch[0], ch[1], ch[2]);
  }
  return cur;
}

/* fmm/input.cc:170-185 */
bool adjacent(const std::vector<Node>& nv, int a, int b) {
  int md = std::max(nv[a].depth, nv[b].depth);
  int ar = 1 << (md - nv[a].depth), br = 1 << (md - nv[b].depth);
  int radius = ar + br, linf = 0;
  bool le = true;
  for (int d = 0; d < 3; d++) {
    int ac = (2 * nv[a].path[d] + 1) * ar, bc = (2 * nv[b].path[d] + 1) * br;
    int dif = std::abs(ac - bc);
    le = le && dif <= radius;
    linf = std::max(linf, dif);
  }
  return le && linf == radius;
}

/* fmm/input.cc:303-381 */
void compute_lists(std::vector<Node>& nv, int id) {
  std::set<int> Us, Vs, Ws, Xs;
  Node& cur = nv[id];
  if (cur.parent != -1) {
    const Node& par = nv[cur.parent];
    const int maxi = 1 << cur.depth;
    for (int i = -2; i < 4; i++)
      for (int j = -2; j < 4; j++)
        for (int k = -2; k < 4; k++) {
          int tp[3] = {2 * par.path[0] + i, 2 * par.path[1] + j, 2 * par.path[2] + k};
          bool in = true, same = true;
          for (int d = 0; d < 3; d++) {
            in = in && tp[d] >= 0 && tp[d] < maxi;
            same = same && tp[d] == cur.path[d];
          }
          if (!in || same) continue;
          int res = findgnt(nv, cur.depth, tp);
          bool adj = adjacent(nv, res, id);
          if (nv[res].depth < cur.depth) {
            if (adj) {
              if (cur.child == -1) Us.insert(res);
            } else {
              Xs.insert(res);
            }
          }
          if (nv[res].depth == cur.depth) {
            if (!adj) {
              Vs.insert(res);
            } else if (cur.child == -1) {
              std::queue<int> rest;
              rest.push(res);
              while (!rest.empty()) {
                int f = rest.front();
                rest.pop();
                if (!adjacent(nv, f, id)) {
                  Ws.insert(f);
                } else if (nv[f].child == -1) {
                  Us.insert(f);
                } else {
                  for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++)
                      for (int c = 0; c < 2; c++) rest.push(child_of(nv, f, a, b, c));
                }
              }
            }
          }
        }
  }
  if (cur.child == -1) Us.insert(id);
  for (int s : Us) if (nv[s].tag & SRCNODE) cur.U.push_back(s);
  for (int s : Vs) if (nv[s].tag & SRCNODE) cur.V.push_back(s);
  for (int s : Ws) if (nv[s].tag & SRCNODE) cur.W.push_back(s);
  for (int s : Xs) if (nv[s].tag & SRCNODE) cur.X.push_back(s);
}

void build_trees(oracle_problem* p) {
  std::vector<Node>& nv = p->nodes;
  const int N = p->n;
  /* source tree, fmm/input.cc:203-298 */
  std::vector<std::vector<int>> vecId;
  std::vector<int> srcNum;
  nv.emplace_back(-1, -1, 0, 0, 0, 0);
  vecId.emplace_back();
  for (int k = 0; k < N; k++) vecId[0].push_back(k);
  srcNum.push_back(N);
  int arr_beg = 0, arr_end = 1;
  p->level.push_back(arr_beg);
  while (arr_beg < arr_end) {
    int arr_count = arr_end;
    for (int k = arr_beg; k < arr_end; k++) {
      if (srcNum[k] > p->q) {
        nv[k].child = arr_count;
        arr_count += 8;
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
              Node ch(k, -1, 2 * nv[k].path[0] + a, 2 * nv[k].path[1] + b, 2 * nv[k].path[2] + c,
                      nv[k].depth + 1);
              nv.push_back(ch);
              vecId.emplace_back();
              srcNum.push_back(0);
            }
        double cc[3];
        node_center(nv[k], cc);
        for (int id : vecId[k]) {
          int a = p->sx[id] >= cc[0], b = p->sy[id] >= cc[1], c = p->sz[id] >= cc[2];
          vecId[child_of(nv, k, a, b, c)].push_back(id);
        }
        std::vector<int>().swap(vecId[k]);
        for (int o = 0; o < 8; o++) srcNum[nv[k].child + o] = (int)vecId[nv[k].child + o].size();
      }
    }
    arr_beg = arr_end;
    arr_end = arr_count;
    p->level.push_back(arr_beg);
  }
  int sum = 0;
  for (size_t i = 0; i < nv.size(); i++)
    if (srcNum[i] > 0) {
      nv[i].tag |= SRCNODE;
      if (nv[i].child == -1) {
        nv[i].srcBeg = sum;
        nv[i].srcNum = srcNum[i];
        sum += srcNum[i];
        nv[i].srcIdx = vecId[i];
      }
    }
  /* target tree, fmm/input.cc:386-460 */
  std::vector<std::vector<int>> tId(nv.size());
  std::vector<int> trgNum(nv.size(), 0);
  for (int k = 0; k < N; k++) tId[0].push_back(k);
  trgNum[0] = N;
  for (size_t i = 0; i < nv.size(); i++) {
    if (nv[i].child != -1) {
      double cc[3];
      node_center(nv[i], cc);
      for (int id : tId[i]) {
        int a = p->tx[id] >= cc[0], b = p->ty[id] >= cc[1], c = p->tz[id] >= cc[2];
        tId[child_of(nv, (int)i, a, b, c)].push_back(id);
      }
      std::vector<int>().swap(tId[i]);
      for (int o = 0; o < 8; o++) trgNum[nv[i].child + o] = (int)tId[nv[i].child + o].size();
    }
  }
  sum = 0;
  for (size_t i = 0; i < nv.size(); i++)
    if (trgNum[i] > 0) {
      nv[i].tag |= TRGNODE;
      if (nv[i].child == -1) {
        nv[i].trgBeg = sum;
        nv[i].trgNum = trgNum[i];
        sum += trgNum[i];
        nv[i].trgIdx = tId[i];
      }
    }
#pragma omp parallel for schedule(dynamic, 64)
  for (long i = 0; i < (long)nv.size(); i++)
    if (nv[i].tag & TRGNODE) compute_lists(nv, (int)i);

  /* gather, fmm/evaluate.cc:13-52 */
  p->gsx.assign(N, 0); p->gsy.assign(N, 0); p->gsz.assign(N, 0); p->gden.assign(N, 0);
  p->gtx.assign(N, 0); p->gty.assign(N, 0); p->gtz.assign(N, 0); p->gpot.assign(N, 0);
  p->src_perm.assign(N, -1); p->trg_perm.assign(N, -1);
  for (const Node& nd : nv) {
    for (int k = 0; k < nd.srcNum; k++) {
      int o = nd.srcIdx[k];
      p->gsx[nd.srcBeg + k] = p->sx[o];
      p->gsy[nd.srcBeg + k] = p->sy[o];
      p->gsz[nd.srcBeg + k] = p->sz[o];
      p->gden[nd.srcBeg + k] = p->den[o];
      p->src_perm[nd.srcBeg + k] = o;
    }
    for (int k = 0; k < nd.trgNum; k++) {
      int o = nd.trgIdx[k];
      p->gtx[nd.trgBeg + k] = p->tx[o];
      p->gty[nd.trgBeg + k] = p->ty[o];
      p->gtz[nd.trgBeg + k] = p->tz[o];
      p->trg_perm[nd.trgBeg + k] = o;
    }
  }
}

void setup_surfaces(oracle_problem* p) {
  const int np = p->np;
  /* trans_setup, fmm/trans.cc:13-26 */
  p->SP[UE] = sampos(np, 1.0);
  p->SP[UC] = sampos(np + 2, 3.0);
  p->SP[DE] = sampos(np, 3.0);
  p->SP[DC] = sampos(np, 1.0);
  p->RP = regpos(np, 1.0);
  p->n_e = p->SP[UE].n;
  p->n_uc = p->SP[UC].n;
}

/* gaussmix: 8 centres uniform in [0.2,0.8]^3 (drawn once; a separate
 * target set is a second draw from the same mixture), then per point a
 * cluster index and a Box-Muller pair per two coordinates; clipped to
 * [0,1] and mapped by x' = 2x - 1.  (No reference counterpart.) */
void draw_centres(unsigned short* st, double cen[8][3]) {
  for (int c = 0; c < 8; c++)
    for (int d = 0; d < 3; d++) cen[c][d] = 0.2 + 0.6 * erand48(st);
}

void draw_gaussmix(unsigned short* st, const double cen[8][3], int N, double* x, double* y, double* z) {
  const double sigma = 0.02;
  for (int i = 0; i < N; i++) {
    int c = (int)(erand48(st) * 8.0);
    if (c > 7) c = 7;
    double g[4];
    for (int h = 0; h < 2; h++) {
      double u1 = erand48(st), u2 = erand48(st);
      double r = std::sqrt(-2.0 * std::log(1.0 - u1));
      g[2 * h] = r * std::cos(2.0 * M_PI * u2);
      g[2 * h + 1] = r * std::sin(2.0 * M_PI * u2);
    }
    double v[3];
    for (int d = 0; d < 3; d++) {
      double t = cen[c][d] + sigma * g[d];
      t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
      v[d] = 2.0 * t - 1.0;
    }
    x[i] = v[0]; y[i] = v[1]; z[i] = v[2];
  }
}

bool generate(oracle_problem* p, const char* dist, long seed, int coincide) {
  const int N = p->n;
  unsigned short* st = p->rng;
  /* srand48(seed) state, as used by erand48 */
  st[0] = 0x330E;
  st[1] = (unsigned short)(seed & 0xFFFF);
  st[2] = (unsigned short)((seed >> 16) & 0xFFFF);
  p->sx.resize(N); p->sy.resize(N); p->sz.resize(N); p->den.resize(N);
  p->tx.resize(N); p->ty.resize(N); p->tz.resize(N);
  std::string d(dist);
  bool sep_targets = true;
  if (d == "uniform") {
    for (int i = 0; i < N; i++) {
      p->sx[i] = 2.0 * erand48(st) - 1.0;
      p->sy[i] = 2.0 * erand48(st) - 1.0;
      p->sz[i] = 2.0 * erand48(st) - 1.0;
    }
    if (!coincide) {
      for (int i = 0; i < N; i++) {
        p->tx[i] = 2.0 * erand48(st) - 1.0;
        p->ty[i] = 2.0 * erand48(st) - 1.0;
        p->tz[i] = 2.0 * erand48(st) - 1.0;
      }
    }
    sep_targets = !coincide;
  } else if (d == "ellipseUniformAngles") {
    /* fmm/input.cc:61-79: targets are always a separate draw */
    const double r = 0.49, c0 = 0.5, c1 = 0.5, c2 = 0.5;
    for (int pass = 0; pass < 2; pass++) {
      double* X = pass ? p->tx.data() : p->sx.data();
      double* Y = pass ? p->ty.data() : p->sy.data();
      double* Z = pass ? p->tz.data() : p->sz.data();
      for (int i = 0; i < N; i++) {
        double phi = 2 * M_PI * erand48(st);
        double theta = M_PI * erand48(st);
        X[i] = c0 + 0.25 * r * std::sin(theta) * std::cos(phi);
        Y[i] = c1 + 0.25 * r * std::sin(theta) * std::sin(phi);
        Z[i] = c2 + r * std::cos(theta);
      }
    }
  } else if (d == "ellipsoid") {
    for (int pass = 0; pass < (coincide ? 1 : 2); pass++) {
      double* X = pass ? p->tx.data() : p->sx.data();
      double* Y = pass ? p->ty.data() : p->sy.data();
      double* Z = pass ? p->tz.data() : p->sz.data();
      for (int i = 0; i < N; i++) {
        double phi = 2 * M_PI * erand48(st);
        double theta = M_PI * erand48(st);
        X[i] = std::sin(theta) * std::cos(phi);
        Y[i] = std::sin(theta) * std::sin(phi);
        Z[i] = 0.25 * std::cos(theta);
      }
    }
    sep_targets = !coincide;
  } else if (d == "gaussmix") {
    double cen[8][3];
    draw_centres(st, cen);
    draw_gaussmix(st, cen, N, p->sx.data(), p->sy.data(), p->sz.data());
    if (!coincide) draw_gaussmix(st, cen, N, p->tx.data(), p->ty.data(), p->tz.data());
    sep_targets = !coincide;
  } else {
    return false;
  }
  if (!sep_targets) {
    p->tx = p->sx; p->ty = p->sy; p->tz = p->sz;
  }
  for (int i = 0; i < N; i++) p->den[i] = erand48(st);
  return true;
}

/* ---------------------------------------------------------------- phases */

void build_ops(oracle_problem* p, int mode) {
  const int ne = p->n_e, nuc = p->n_uc;
  const Surf& ue = p->SP[UE];
  const Surf& uc = p->SP[UC];
  const Surf& de = p->SP[DE];
  const Surf& dc = p->SP[DC];
  const double eps = std::pow(0.1, p->np + 1); /* alt(), fmm/trans.cc:120-124 */
  std::vector<double> K((size_t)ne * nuc);
  /* compute_UC2UE_mat, fmm/trans.cc:143-168 */
  kernel_mat(ne, nuc, ue.x.data(), ue.y.data(), ue.z.data(), uc.x.data(), uc.y.data(), uc.z.data(),
             K.data(), mode);
  p->uc2ue.assign((size_t)ne * nuc, 0.0);
  pinv(K.data(), nuc, ne, eps, p->uc2ue.data());
  /* compute_UE2UC_mat, fmm/trans.cc:174-203 */
  p->ue2uc.assign((size_t)8 * nuc * ne, 0.0);
  std::vector<double> cx(nuc), cy(nuc), cz(nuc), dx(ne), dy(ne), dz(ne);
  for (int o = 0; o < 8; o++) {
    int ix = getbits(o, 2), iy = getbits(o, 1), iz = getbits(o, 0);
    for (int j = 0; j < nuc; j++) { cx[j] = 2.0 * uc.x[j]; cy[j] = 2.0 * uc.y[j]; cz[j] = 2.0 * uc.z[j]; }
    for (int j = 0; j < ne; j++) {
      dx[j] = ue.x[j] + (2 * ix - 1);
      dy[j] = ue.y[j] + (2 * iy - 1);
      dz[j] = ue.z[j] + (2 * iz - 1);
    }
    kernel_mat(ne, nuc, dx.data(), dy.data(), dz.data(), cx.data(), cy.data(), cz.data(),
               &p->ue2uc[(size_t)o * nuc * ne], mode);
  }
  /* compute_DC2DE_mat, fmm/trans.cc:208-233 */
  std::vector<double> K2((size_t)ne * ne);
  kernel_mat(ne, ne, de.x.data(), de.y.data(), de.z.data(), dc.x.data(), dc.y.data(), dc.z.data(),
             K2.data(), mode);
  p->dc2de.assign((size_t)ne * ne, 0.0);
  pinv(K2.data(), ne, ne, eps, p->dc2de.data());
  /* compute_DE2DC_mat, fmm/trans.cc:237-264 */
  p->de2dc.assign((size_t)8 * ne * ne, 0.0);
  std::vector<double> ex(ne), ey(ne), ez(ne);
  for (int o = 0; o < 8; o++) {
    int ix = getbits(o, 2), iy = getbits(o, 1), iz = getbits(o, 0);
    for (int j = 0; j < ne; j++) {
      ex[j] = 0.5 * dc.x[j] + (ix - 0.5);
      ey[j] = 0.5 * dc.y[j] + (iy - 0.5);
      ez[j] = 0.5 * dc.z[j] + (iz - 0.5);
    }
    kernel_mat(ne, ne, de.x.data(), de.y.data(), de.z.data(), ex.data(), ey.data(), ez.data(),
               &p->de2dc[(size_t)o * ne * ne], mode);
  }
  p->have_ops = true;
}

struct Timer {
  double t0;
  Timer() : t0(now_s()) {}
  double lap() { double t = now_s(), d = t - t0; t0 = t; return d; }
};

/* up_compute, fmm/omp_sse_block.cc:60-158 */
void up_compute(oracle_problem* p, int i, int mode, std::vector<double>& cx, std::vector<double>& cy,
                std::vector<double>& cz, std::vector<double>& pot) {
  const std::vector<Node>& nv = p->nodes;
  const int ne = p->n_e, nuc = p->n_uc;
  const Node& nd = nv[i];
  if (!(nd.tag & SRCNODE)) return;
  double* uden = &p->uden[(size_t)i * ne];
  std::vector<double> tmp(ne, 0.0);
  std::fill(pot.begin(), pot.end(), 0.0);
  if (nd.child == -1) {
    double c[3];
    node_center(nd, c);
    double r = node_radius(nd);
    for (int j = 0; j < nuc; j++) {
      cx[j] = c[0] + r * p->SP[UC].x[j];
      cy[j] = c[1] + r * p->SP[UC].y[j];
      cz[j] = c[2] + r * p->SP[UC].z[j];
    }
    direct_eval(nuc, cx.data(), cy.data(), cz.data(), pot.data(), nd.srcNum, &p->gsx[nd.srcBeg],
                &p->gsy[nd.srcBeg], &p->gsz[nd.srcBeg], &p->gden[nd.srcBeg], mode);
  } else {
    std::vector<double> cd(ne);
    for (int o = 0; o < 8; o++) {
      int ch = nd.child + o;
      if (!(nv[ch].tag & SRCNODE)) continue;
      const double s = std::ldexp(1.0, nv[ch].depth);
      for (int j = 0; j < ne; j++) cd[j] = p->uden[(size_t)ch * ne + j] * s;
      gemv(nuc, ne, &p->ue2uc[(size_t)o * nuc * ne], cd.data(), pot.data(), true);
    }
  }
  gemv(ne, nuc, p->uc2ue.data(), pot.data(), tmp.data(), false);
  const double s = std::ldexp(1.0, -nd.depth);
  for (int j = 0; j < ne; j++) uden[j] += tmp[j] * s;
}

/* up_calc__cpu, fmm/omp_sse_block.cc:161-232 (UPBS-blocked bottom level) */
void up_pass(oracle_problem* p, int mode) {
  const std::vector<Node>& nv = p->nodes;
  const std::vector<int>& lv = p->level;
  const int L = (int)lv.size(); /* levels 0..L-2, lv[L-1] == n_nodes */
  const int nuc = p->n_uc;
  auto run_range = [&](int beg, int end) {
#pragma omp parallel
    {
      std::vector<double> cx(nuc), cy(nuc), cz(nuc), pot(nuc);
#pragma omp for schedule(guided)
      for (int i = beg; i < end; i++) up_compute(p, i, mode, cx, cy, cz, pot);
    }
  };
  const char* s = getenv("UPBS");
  int bz = s ? atoi(s) : 1000;
  if (bz <= 0) bz = 1000;
  if (L >= 5) {
    const int leaf_nodes = lv[L - 1] - lv[L - 2];
    int nbz = (int)std::ceil((double)leaf_nodes / bz);
    if (nbz < 1) nbz = 1;
    int beg = 0, end = 0;
    for (int k = 0; k < nbz; k++)
      for (int j = L - 1; j > 3; j--) {
        if (j == L - 1) {
          beg = lv[j - 1] + k * bz;
          end = std::min(beg + bz, lv[j]);
        } else {
          beg = (k == 0) ? lv[j - 1] : nv[beg - 1].parent;
          end = (k == nbz - 1) ? lv[j] : nv[end - 1].parent;
        }
        run_range(beg, end);
      }
    for (int j = 3; j > 0; j--) run_range(lv[j - 1], lv[j]);
  } else {
    for (int j = L - 1; j > 0; j--) run_range(lv[j - 1], lv[j]);
  }
}

void fft_src(oracle_problem* p, std::vector<std::complex<double>>& eden, const Dft& F) {
  const std::vector<Node>& nv = p->nodes;
  const int ne = p->n_e, N = F.N, H = N / 2 + 1, eff = N * N * H;
  const std::vector<int> off = sample_grid_offsets(p->np);
  eden.assign((size_t)nv.size() * eff, 0.0);
#pragma omp parallel
  {
    std::vector<double> reg((size_t)N * N * N);
#pragma omp for schedule(static)
    for (long i = 0; i < (long)nv.size(); i++) {
      if (!(nv[i].tag & SRCNODE)) continue;
      std::fill(reg.begin(), reg.end(), 0.0);
      const double s = std::ldexp(1.0, nv[i].depth); /* fmm/trans.cc:308-333 */
      for (int j = 0; j < ne; j++) reg[off[j]] = p->uden[(size_t)i * ne + j] * s;
      rfft3(F, reg.data(), &eden[(size_t)i * eff]);
    }
  }
}

void translation_spectra(int np, int mode, const Dft& F, std::vector<std::complex<double>>& T) {
  const int N = F.N, H = N / 2 + 1, eff = N * N * H;
  Surf rp = regpos(np, 1.0);
  T.assign((size_t)343 * eff, 0.0);
#pragma omp parallel for schedule(static)
  for (int id = 0; id < 343; id++) {
    int i1 = id % 7 - 3, i2 = (id / 7) % 7 - 3, i3 = id / 49 - 3;
    if (!(std::abs(i1) > 1 || std::abs(i2) > 1 || std::abs(i3) > 1)) continue;
    double px = i1 * 2.0, py = i2 * 2.0, pz = i3 * 2.0; /* fmm/omp_sse_block.cc:455-466 */
    std::vector<double> tt(rp.n);
    kernel_mat(1, rp.n, &px, &py, &pz, rp.x.data(), rp.y.data(), rp.z.data(), tt.data(), mode);
    rfft3(F, tt.data(), &T[(size_t)id * eff]);
  }
}

/* vlist_calc__cpu, fmm/omp_sse_block.cc:537-621 */
void vlist_pass(oracle_problem* p, int mode, int nthreads) {
  std::vector<Node>& nv = p->nodes;
  const int ne = p->n_e, np = p->np, N = 2 * np, H = N / 2 + 1, eff = N * N * H;
  Dft F(N);
  std::vector<std::complex<double>> eden, T;
  fft_src(p, eden, F);
  translation_spectra(np, mode, F, T);
  std::vector<std::complex<double>> eval((size_t)nv.size() * eff, 0.0);
  /* work_division_V, fmm/evaluate.cc:143-176 */
  const int nn = (int)nv.size();
  std::vector<int> lo(nn + 1, 0), Ti(nthreads + 1, 0);
  for (int i = 0; i < nn; i++) lo[i + 1] = lo[i] + (int)nv[i].V.size();
  {
    int split = lo[nn] / nthreads + 1, cutoff = split, ct = 1;
    for (int i = 0; i < nn; i++)
      if (lo[i] > cutoff && ct < nthreads) { Ti[ct++] = i; cutoff += split; }
    for (int t = ct; t < nthreads; t++) Ti[t] = nn;
    Ti[nthreads] = nn;
    for (int t = 1; t <= nthreads; t++) Ti[t] = std::max(Ti[t], Ti[t - 1]);
  }
  const char* s = getenv("BS");
  int bs = s ? atoi(s) : 49;
  if (bs <= 0) bs = 49;
#pragma omp parallel num_threads(nthreads)
  {
    const int tid = omp_get_thread_num();
    const int tvc = 343;
    const int nb = std::max(1, tvc / bs);
    for (int b = 0; b < nb; b++) {
      const int st = b * bs, en = (b == nb - 1) ? tvc : (b + 1) * bs;
      for (int i = Ti[tid]; i < Ti[tid + 1]; i++) {
        if (!((nv[i].tag & TRGNODE) && !nv[i].V.empty())) continue;
        double tc[3], vc[3];
        node_center(nv[i], tc);
        const double D = 2.0 * node_radius(nv[i]);
        std::complex<double>* y = &eval[(size_t)i * eff];
        for (int src : nv[i].V) {
          node_center(nv[src], vc);
          int idx[3];
          for (int d = 0; d < 3; d++) idx[d] = (int)std::round((vc[d] - tc[d]) / D);
          const int id = (idx[0] + 3) + (idx[1] + 3) * 7 + (idx[2] + 3) * 49;
          if (id < st || id >= en) continue;
          const std::complex<double>* x = &eden[(size_t)src * eff];
          const std::complex<double>* t = &T[(size_t)id * eff];
          /* pointwise_mult, fmm/omp_sse_block.cc:627-658: z += t * x */
          for (int f = 0; f < eff; f++) {
            double zr = y[f].real(), zi = y[f].imag();
            zr = zr + t[f].real() * x[f].real();
            zi = zi + t[f].real() * x[f].imag();
            zr = zr + (-(t[f].imag() * x[f].imag()));
            zi = zi + t[f].imag() * x[f].real();
            y[f] = std::complex<double>(zr, zi);
          }
        }
      }
    }
  }
  /* compute_ifft_trg, fmm/omp_sse_block.cc:482-533 */
  const std::vector<int> off = sample_grid_offsets(np);
  const double nrmfc = 1.0 / (double)(N * N * N);
#pragma omp parallel
  {
    std::vector<double> reg((size_t)N * N * N);
#pragma omp for schedule(static)
    for (long i = 0; i < nn; i++) {
      if (!((nv[i].tag & TRGNODE) && !nv[i].V.empty())) continue;
      std::complex<double>* y = &eval[(size_t)i * eff];
      for (int f = 0; f < eff; f++) y[f] *= nrmfc;
      irfft3(F, y, reg.data());
      for (int j = 0; j < ne; j++) p->dchk[(size_t)i * ne + j] += reg[off[j]];
    }
  }
}

/* wlist_calc__cpu, fmm/omp_sse_block.cc:663-707 */
void wlist_pass(oracle_problem* p, int mode) {
  const std::vector<Node>& nv = p->nodes;
  const int ne = p->n_e;
#pragma omp parallel
  {
    std::vector<double> x(ne), y(ne), z(ne);
#pragma omp for schedule(guided)
    for (long i = 0; i < (long)nv.size(); i++) {
      const Node& nd = nv[i];
      if (!(nd.tag & TRGNODE) || nd.child != -1) continue;
      for (int src : nd.W) {
        const Node& sn = nv[src];
        if (sn.child == -1 && sn.srcNum < ne) {
          direct_eval(nd.trgNum, &p->gtx[nd.trgBeg], &p->gty[nd.trgBeg], &p->gtz[nd.trgBeg],
                      &p->gpot[nd.trgBeg], sn.srcNum, &p->gsx[sn.srcBeg], &p->gsy[sn.srcBeg],
                      &p->gsz[sn.srcBeg], &p->gden[sn.srcBeg], mode);
        } else {
          double c[3];
          node_center(sn, c);
          double r = node_radius(sn);
          for (int j = 0; j < ne; j++) {
            x[j] = c[0] + r * p->SP[UE].x[j];
            y[j] = c[1] + r * p->SP[UE].y[j];
            z[j] = c[2] + r * p->SP[UE].z[j];
          }
          direct_eval(nd.trgNum, &p->gtx[nd.trgBeg], &p->gty[nd.trgBeg], &p->gtz[nd.trgBeg],
                      &p->gpot[nd.trgBeg], ne, x.data(), y.data(), z.data(),
                      &p->uden[(size_t)src * ne], mode);
        }
      }
    }
  }
}

/* xlist_calc__cpu, fmm/omp_sse_block.cc:712-756 */
void xlist_pass(oracle_problem* p, int mode) {
  const std::vector<Node>& nv = p->nodes;
  const int ne = p->n_e;
#pragma omp parallel
  {
    std::vector<double> x(ne), y(ne), z(ne);
#pragma omp for schedule(static)
    for (long i = 0; i < (long)nv.size(); i++) {
      const Node& nd = nv[i];
      if (!(nd.tag & TRGNODE)) continue;
      for (int src : nd.X) {
        const Node& sn = nv[src];
        if (nd.child == -1 && nd.trgNum < ne) {
          direct_eval(nd.trgNum, &p->gtx[nd.trgBeg], &p->gty[nd.trgBeg], &p->gtz[nd.trgBeg],
                      &p->gpot[nd.trgBeg], sn.srcNum, &p->gsx[sn.srcBeg], &p->gsy[sn.srcBeg],
                      &p->gsz[sn.srcBeg], &p->gden[sn.srcBeg], mode);
        } else {
          double c[3];
          node_center(nd, c);
          double r = node_radius(nd);
          for (int j = 0; j < ne; j++) {
            x[j] = c[0] + r * p->SP[DC].x[j];
            y[j] = c[1] + r * p->SP[DC].y[j];
            z[j] = c[2] + r * p->SP[DC].z[j];
          }
          direct_eval(ne, x.data(), y.data(), z.data(), &p->dchk[(size_t)i * ne], sn.srcNum,
                      &p->gsx[sn.srcBeg], &p->gsy[sn.srcBeg], &p->gsz[sn.srcBeg],
                      &p->gden[sn.srcBeg], mode);
        }
      }
    }
  }
}

/* down_calc__cpu, fmm/omp_sse_block.cc:761-868 */
void down_pass(oracle_problem* p, int mode) {
  const std::vector<Node>& nv = p->nodes;
  const std::vector<int>& lv = p->level;
  const int ne = p->n_e;
  for (size_t j = 0; j + 1 < lv.size(); j++) {
#pragma omp parallel
    {
      std::vector<double> tmp(ne), x(ne), y(ne), z(ne);
#pragma omp for schedule(guided)
      for (long i = lv[j]; i < lv[j + 1]; i++) {
        const Node& nd = nv[i];
        if (!(nd.tag & TRGNODE)) continue;
        double* tval = &p->dchk[(size_t)i * ne];
        if (nd.depth >= 3) {
          const int par = nd.parent;
          int o = (nd.path[0] - 2 * nv[par].path[0]) * 4 + (nd.path[1] - 2 * nv[par].path[1]) * 2 +
                  (nd.path[2] - 2 * nv[par].path[2]);
          const double s = std::ldexp(1.0, nv[par].depth);
          for (int k = 0; k < ne; k++) tmp[k] = p->dden[(size_t)par * ne + k] * s;
          gemv(ne, ne, &p->de2dc[(size_t)o * ne * ne], tmp.data(), tval, true);
        }
        if (nd.depth >= 2) {
          double* dd = &p->dden[(size_t)i * ne];
          gemv(ne, ne, p->dc2de.data(), tval, tmp.data(), false);
          const double s = std::ldexp(1.0, -nd.depth);
          for (int k = 0; k < ne; k++) dd[k] = dd[k] + tmp[k] * s;
        }
        if (nd.child == -1) {
          double c[3];
          node_center(nd, c);
          double r = node_radius(nd);
          for (int k = 0; k < ne; k++) {
            x[k] = c[0] + r * p->SP[DE].x[k];
            y[k] = c[1] + r * p->SP[DE].y[k];
            z[k] = c[2] + r * p->SP[DE].z[k];
          }
          direct_eval(nd.trgNum, &p->gtx[nd.trgBeg], &p->gty[nd.trgBeg], &p->gtz[nd.trgBeg],
                      &p->gpot[nd.trgBeg], ne, x.data(), y.data(), z.data(),
                      &p->dden[(size_t)i * ne], mode);
        }
      }
    }
  }
}

/* ulist_calc__cpu, fmm/omp_sse_block.cc:236-261 with work_division_U (fmm/evaluate.cc:100-141) */
void ulist_pass(oracle_problem* p, int mode, int nthreads) {
  const std::vector<Node>& nv = p->nodes;
  const int nn = (int)nv.size();
  std::vector<int> Ti(nthreads + 1, 0);
  unsigned long long sum = 0;
  for (int i = 0; i < nn; i++)
    for (int s : nv[i].U) sum += (unsigned long long)nv[i].trgNum * nv[s].srcNum;
  unsigned long long split = sum / nthreads + 1, cutoff = split;
  int ct = 1;
  sum = 0;
  for (int i = 0; i < nn; i++)
    for (int s : nv[i].U) {
      sum += (unsigned long long)nv[i].trgNum * nv[s].srcNum;
      if (sum > cutoff && ct < nthreads) { Ti[ct++] = i; cutoff += split; }
    }
  for (int t = ct; t <= nthreads; t++) Ti[t] = nn;
  for (int t = 1; t <= nthreads; t++) Ti[t] = std::max(Ti[t], Ti[t - 1]);
#pragma omp parallel num_threads(nthreads)
  {
    const int tid = omp_get_thread_num();
    for (int i = Ti[tid]; i < Ti[tid + 1]; i++) {
      const Node& nd = nv[i];
      for (int s : nd.U) {
        const Node& sn = nv[s];
        direct_eval(nd.trgNum, &p->gtx[nd.trgBeg], &p->gty[nd.trgBeg], &p->gtz[nd.trgBeg],
                    &p->gpot[nd.trgBeg], sn.srcNum, &p->gsx[sn.srcBeg], &p->gsy[sn.srcBeg],
                    &p->gsz[sn.srcBeg], &p->gden[sn.srcBeg], mode);
      }
    }
  }
}

}  // namespace

/* ================================================================ C API */

extern "C" {

oracle_problem* oracle_create(int n, const char* dist, int pts_max, int np, long seed, int coincide) {
  if (n <= 0 || pts_max <= 0 || np < 2 || !dist) return nullptr;
  oracle_problem* p = new oracle_problem();
  p->n = n; p->np = np; p->q = pts_max;
  if (!generate(p, dist, seed, coincide)) { delete p; return nullptr; }
  build_trees(p);
  setup_surfaces(p);
  return p;
}

oracle_problem* oracle_create_from_points(int n, const double* sx, const double* sy, const double* sz,
                                          const double* den, const double* tx, const double* ty,
                                          const double* tz, int pts_max, int np) {
  if (n <= 0 || pts_max <= 0 || np < 2) return nullptr;
  oracle_problem* p = new oracle_problem();
  p->n = n; p->np = np; p->q = pts_max;
  p->rng[0] = 0x330E; p->rng[1] = 26; p->rng[2] = 0;
  p->sx.assign(sx, sx + n); p->sy.assign(sy, sy + n); p->sz.assign(sz, sz + n);
  p->den.assign(den, den + n);
  p->tx.assign(tx, tx + n); p->ty.assign(ty, ty + n); p->tz.assign(tz, tz + n);
  build_trees(p);
  setup_surfaces(p);
  return p;
}

void oracle_destroy(oracle_problem* p) { delete p; }

int oracle_num_nodes(const oracle_problem* p) { return (int)p->nodes.size(); }
int oracle_num_levels(const oracle_problem* p) { return (int)p->level.size(); }
int oracle_level_starts(const oracle_problem* p, int* out) {
  for (size_t i = 0; i < p->level.size(); i++) out[i] = p->level[i];
  return 0;
}

int oracle_tree_arrays(const oracle_problem* p, int* parent, int* child, int* depth, int* tag,
                       int* path3, int* src_beg, int* src_num, int* trg_beg, int* trg_num) {
  for (size_t i = 0; i < p->nodes.size(); i++) {
    const Node& nd = p->nodes[i];
    parent[i] = nd.parent; child[i] = nd.child; depth[i] = nd.depth; tag[i] = nd.tag;
    for (int d = 0; d < 3; d++) path3[3 * i + d] = nd.path[d];
    src_beg[i] = nd.srcBeg; src_num[i] = nd.srcNum; trg_beg[i] = nd.trgBeg; trg_num[i] = nd.trgNum;
  }
  return 0;
}

static const std::vector<int>& list_of(const Node& nd, int which) {
  switch (which) {
    case ORACLE_LIST_U: return nd.U;
    case ORACLE_LIST_V: return nd.V;
    case ORACLE_LIST_W: return nd.W;
    default: return nd.X;
  }
}

int oracle_list_nnz(const oracle_problem* p, int which) {
  long s = 0;
  for (const Node& nd : p->nodes) s += (long)list_of(nd, which).size();
  return (int)s;
}

int oracle_list_csr(const oracle_problem* p, int which, int* ptr, int* idx) {
  ptr[0] = 0;
  for (size_t i = 0; i < p->nodes.size(); i++) {
    const std::vector<int>& l = list_of(p->nodes[i], which);
    for (size_t k = 0; k < l.size(); k++) idx[ptr[i] + k] = l[k];
    ptr[i + 1] = ptr[i] + (int)l.size();
  }
  return 0;
}

int oracle_points(const oracle_problem* p, double* sx, double* sy, double* sz, double* den,
                  double* tx, double* ty, double* tz) {
  const size_t n = p->n;
  std::memcpy(sx, p->sx.data(), n * 8); std::memcpy(sy, p->sy.data(), n * 8);
  std::memcpy(sz, p->sz.data(), n * 8); std::memcpy(den, p->den.data(), n * 8);
  std::memcpy(tx, p->tx.data(), n * 8); std::memcpy(ty, p->ty.data(), n * 8);
  std::memcpy(tz, p->tz.data(), n * 8);
  return 0;
}

int oracle_perms(const oracle_problem* p, int* src_perm, int* trg_perm) {
  std::memcpy(src_perm, p->src_perm.data(), (size_t)p->n * 4);
  std::memcpy(trg_perm, p->trg_perm.data(), (size_t)p->n * 4);
  return 0;
}

int oracle_build_operators(oracle_problem* p, int kernel_mode) {
  build_ops(p, kernel_mode);
  return 0;
}

int oracle_set_operators(oracle_problem* p, const double* uc2ue, const double* ue2uc,
                         const double* dc2de, const double* de2dc) {
  const size_t ne = p->n_e, nuc = p->n_uc;
  p->uc2ue.assign(uc2ue, uc2ue + ne * nuc);
  p->ue2uc.assign(ue2uc, ue2uc + 8 * nuc * ne);
  p->dc2de.assign(dc2de, dc2de + ne * ne);
  p->de2dc.assign(de2dc, de2dc + 8 * ne * ne);
  p->have_ops = true;
  return 0;
}

int oracle_get_operators(const oracle_problem* p, double* uc2ue, double* ue2uc, double* dc2de,
                         double* de2dc) {
  if (!p->have_ops) return -1;
  std::memcpy(uc2ue, p->uc2ue.data(), p->uc2ue.size() * 8);
  std::memcpy(ue2uc, p->ue2uc.data(), p->ue2uc.size() * 8);
  std::memcpy(dc2de, p->dc2de.data(), p->dc2de.size() * 8);
  std::memcpy(de2dc, p->de2dc.data(), p->de2dc.size() * 8);
  return 0;
}

int oracle_evaluate(oracle_problem* p, int mode, int nthreads, double* pot_orig, double* times) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  omp_set_num_threads(nthreads);
  const size_t nn = p->nodes.size(), ne = p->n_e;
  double t[8] = {0};
  Timer tm;
  const double t_start = now_s();
  if (!p->have_ops) build_ops(p, mode); /* the reference builds them inside up/down */
  p->uden.assign(nn * ne, 0.0);
  p->dchk.assign(nn * ne, 0.0);
  p->dden.assign(nn * ne, 0.0);
  std::fill(p->gpot.begin(), p->gpot.end(), 0.0);
  tm.lap();
  up_pass(p, mode);
  t[0] = tm.lap();
  ulist_pass(p, mode, nthreads);
  t[1] = tm.lap();
  vlist_pass(p, mode, nthreads);
  t[2] = tm.lap();
  wlist_pass(p, mode);
  t[3] = tm.lap();
  xlist_pass(p, mode);
  t[4] = tm.lap();
  down_pass(p, mode);
  t[5] = tm.lap();
  /* copy_trg_val__cpu, fmm/omp_sse_block.cc:873-887 */
  for (int k = 0; k < p->n; k++) pot_orig[p->trg_perm[k]] = p->gpot[k];
  t[6] = tm.lap();
  t[7] = now_s() - t_start;
  if (times) std::memcpy(times, t, sizeof t);
  return 0;
}

int oracle_intermediates(const oracle_problem* p, double* uden, double* dchk, double* dden) {
  if (uden) std::memcpy(uden, p->uden.data(), p->uden.size() * 8);
  if (dchk) std::memcpy(dchk, p->dchk.data(), p->dchk.size() * 8);
  if (dden) std::memcpy(dden, p->dden.data(), p->dden.size() * 8);
  return 0;
}

double oracle_validate(oracle_problem* p, const double* pot_orig, int nv, int* chk_ids,
                       double* direct_out) {
  std::vector<int> ids(nv);
  for (int k = 0; k < nv; k++) ids[k] = (int)std::floor(erand48(p->rng) * p->n);
  std::vector<double> chk(nv, 0.0);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nv; k++) {
    const int t = ids[k];
    double x = p->tx[t], y = p->ty[t], z = p->tz[t], pot = 0.0;
    direct_eval(1, &x, &y, &z, &pot, p->n, p->sx.data(), p->sy.data(), p->sz.data(),
                p->den.data(), ORACLE_KERNEL_EXACT);
    chk[k] = pot;
  }
  double vn = 0, en = 0;
  for (int k = 0; k < nv; k++) {
    double d = chk[k] - pot_orig[ids[k]];
    vn += pot_orig[ids[k]] * pot_orig[ids[k]];
    en += d * d;
  }
  if (chk_ids) for (int k = 0; k < nv; k++) chk_ids[k] = ids[k];
  if (direct_out) for (int k = 0; k < nv; k++) direct_out[k] = chk[k];
  return std::sqrt(en) / std::sqrt(vn);
}

int oracle_surface(int np, double R, double* xyz) {
  Surf s = sampos(np, R);
  for (int i = 0; i < s.n; i++) { xyz[i] = s.x[i]; xyz[s.n + i] = s.y[i]; xyz[2 * s.n + i] = s.z[i]; }
  return s.n;
}

int oracle_regpos(int np, double R, double* xyz) {
  Surf s = regpos(np, R);
  for (int i = 0; i < s.n; i++) { xyz[i] = s.x[i]; xyz[s.n + i] = s.y[i]; xyz[2 * s.n + i] = s.z[i]; }
  return s.n;
}

int oracle_pinv(const double* M, int m, int n, double eps, double* R) {
  pinv(M, m, n, eps, R);
  return 0;
}

int oracle_svd_values(const double* M, int m, int n, double* s) {
  std::vector<double> S, U, V;
  if (m >= n) {
    jacobi_svd(M, m, n, S, U, V);
  } else {
    std::vector<double> MT((size_t)m * n);
    for (int i = 0; i < m; i++)
      for (int j = 0; j < n; j++) MT[(size_t)i * n + j] = M[(size_t)j * m + i];
    jacobi_svd(MT.data(), n, m, S, U, V);
  }
  for (size_t i = 0; i < S.size(); i++) s[i] = S[i];
  return (int)S.size();
}

int oracle_rfft3(int N, const double* in, double* out) {
  Dft F(N);
  rfft3(F, in, reinterpret_cast<std::complex<double>*>(out));
  return 0;
}

int oracle_irfft3(int N, const double* in, double* out) {
  Dft F(N);
  irfft3(F, reinterpret_cast<const std::complex<double>*>(in), out);
  return 0;
}

int oracle_translation_spectra(int np, int mode, double* out) {
  Dft F(2 * np);
  std::vector<std::complex<double>> T;
  translation_spectra(np, mode, F, T);
  std::memcpy(out, T.data(), T.size() * 16);
  return 0;
}

int oracle_direct(int nt, const double* tx, const double* ty, const double* tz, int ns,
                  const double* sx, const double* sy, const double* sz, const double* w, int mode,
                  double* pot) {
  for (int i = 0; i < nt; i++) pot[i] = 0.0;
  direct_eval(nt, tx, ty, tz, pot, ns, sx, sy, sz, w, mode);
  return 0;
}

void oracle_set_reverse_order(int on) { g_reverse_order = on ? 1 : 0; }

int oracle_drand48(long seed, int k, double* out) {
  unsigned short st[3] = {0x330E, (unsigned short)(seed & 0xFFFF), (unsigned short)((seed >> 16) & 0xFFFF)};
  for (int i = 0; i < k; i++) out[i] = erand48(st);
  return 0;
}

}  // extern "C"
