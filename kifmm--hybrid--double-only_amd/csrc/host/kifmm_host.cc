/*
 * kifmm_host.cc -- host side of the MI355X KIFMM evaluator.
 *
 * Point generation, source octree, target binning, U/V/W/X lists, gather,
 * the translation operators and the sampled validation.  These keep the
 * reference's CPU-side semantics exactly (identical node order, lists and
 * gathered order, so the GPU path is a drop-in behind the same tree):
 *   generators     fmm/input.cc:29-84, fmm/driver.cc:244-267
 *   source octree  fmm/input.cc:203-298   (BFS level order, split while count > q,
 *                                          children contiguous, octant 4x+2y+z,
 *                                          stable point order inside a node)
 *   target binning fmm/input.cc:386-460
 *   lists          fmm/input.cc:303-381
 *   gather         fmm/evaluate.cc:13-96   (leaf ranges assigned in BFS order)
 *   operators      fmm/trans.cc:13-264, pinv fmm/vecmatop.cc:34-78 + fmm/svdrep.cc:10-66
 *   validate       fmm/driver.cc:141-206
 * The implementation is this library's own: index-range octree with a
 * stable counting sort per node (no per-node vectors), sorted-vector lists,
 * and a one-sided Jacobi SVD carried in 80-bit long double.
 */
#include "kifmm_host.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <strings.h>
#include <vector>

#include "kifmm_internal.h"

namespace kifmm {

thread_local std::string g_last_error;

void set_error(const std::string& s) { g_last_error = s; }

double wall_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/* ------------------------------------------------------------ drand48 */
/* POSIX drand48: X <- (0x5DEECE66D * X + 0xB) mod 2^48, value X / 2^48;
 * srand48(s) sets X = (s << 16) | 0x330E. */
struct Rand48 {
  uint64_t x = 0;
  void seed(long s) { x = (((uint64_t)(s & 0xFFFFFFFFu)) << 16) | 0x330Eu; }
  double next() {
    x = (0x5DEECE66DULL * x + 0xBULL) & 0xFFFFFFFFFFFFULL;
    return std::ldexp((double)x, -48);
  }
};

}  // namespace kifmm

using namespace kifmm;

struct kifmm_problem {
  int n = 0, np = 6, q = 0, coincident = 0;
  Rand48 rng;
  std::vector<double> sx, sy, sz, den, tx, ty, tz; /* input order */
  /* tree */
  std::vector<int> level, parent, child, depth, tag, path, sbeg, snum, tbeg, tnum;
  std::vector<int> sperm, tperm; /* gathered position -> input index */
  std::vector<int> lptr[4], lidx[4];
  std::vector<double> gsx, gsy, gsz, gden, gtx, gty, gtz;
  /* operators */
  int n_e = 0, n_uc = 0;
  bool have_ops = false;
  std::vector<double> uc2ue, ue2uc, dc2de, de2dc;
  double t_gen = 0, t_tree = 0, t_ops = 0;
};

namespace {

int env_positive(const char* name, int def) {
  const char* s = getenv(name);
  if (!s) return def;
  long t = (long)atof(s);
  if (t <= 0) {
    fprintf(stderr, "*** WARNING: %s (=%ld) should be >= 1 ***\n    Using default value: %d\n", name, t, def);
    return def;
  }
  return (int)t;
}

/* ------------------------------------------------------------ generation */

void draw_uniform(Rand48& r, int n, double* x, double* y, double* z) {
  for (int i = 0; i < n; i++) {
    x[i] = 2.0 * r.next() - 1.0;
    y[i] = 2.0 * r.next() - 1.0;
    z[i] = 2.0 * r.next() - 1.0;
  }
}

/* angles drawn phi first, then theta, as fmm/input.cc:65-70 */
void draw_ellipse(Rand48& r, int n, double cx, double cy, double cz, double ax, double ay,
                  double az, double* x, double* y, double* z) {
  for (int i = 0; i < n; i++) {
    double phi = 2 * M_PI * r.next();
    double theta = M_PI * r.next();
    x[i] = cx + ax * std::sin(theta) * std::cos(phi);
    y[i] = cy + ay * std::sin(theta) * std::sin(phi);
    z[i] = cz + az * std::cos(theta);
  }
}

/* Mixture of 8 isotropic Gaussians, sigma 0.02 in [0,1]^3: centres uniform
 * in [0.2,0.8]^3, drawn once per problem (targets, when separate, are a
 * second draw from the same mixture); per point a cluster draw and two
 * Box-Muller pairs (the fourth normal is discarded); clipped to [0,1],
 * mapped by x' = 2x - 1. */
struct Centres {
  double c[8][3];
};

Centres draw_centres(Rand48& r) {
  Centres k;
  for (int c = 0; c < 8; c++)
    for (int d = 0; d < 3; d++) k.c[c][d] = 0.2 + 0.6 * r.next();
  return k;
}

void draw_gaussmix(Rand48& r, const Centres& k, int n, double* x, double* y, double* z) {
  for (int i = 0; i < n; i++) {
    int c = (int)(r.next() * 8.0);
    if (c > 7) c = 7;
    double g[4];
    for (int h = 0; h < 2; h++) {
      double u1 = r.next(), u2 = r.next();
      double rad = std::sqrt(-2.0 * std::log(1.0 - u1));
      g[2 * h] = rad * std::cos(2.0 * M_PI * u2);
      g[2 * h + 1] = rad * std::sin(2.0 * M_PI * u2);
    }
    double v[3];
    for (int d = 0; d < 3; d++) {
      double t = k.c[c][d] + 0.02 * g[d];
      t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
      v[d] = 2.0 * t - 1.0;
    }
    x[i] = v[0]; y[i] = v[1]; z[i] = v[2];
  }
}

bool generate(kifmm_problem* p, const char* dist, long seed, int coincide) {
  const int n = p->n;
  p->rng.seed(seed);
  p->sx.resize(n); p->sy.resize(n); p->sz.resize(n); p->den.resize(n);
  p->tx.resize(n); p->ty.resize(n); p->tz.resize(n);
  bool separate;
  if (!strcmp(dist, "uniform")) {
    draw_uniform(p->rng, n, p->sx.data(), p->sy.data(), p->sz.data());
    separate = !coincide;
    if (separate) draw_uniform(p->rng, n, p->tx.data(), p->ty.data(), p->tz.data());
  } else if (!strcmp(dist, "ellipseUniformAngles")) {
    const double r = 0.49;
    draw_ellipse(p->rng, n, 0.5, 0.5, 0.5, 0.25 * r, 0.25 * r, r, p->sx.data(), p->sy.data(), p->sz.data());
    draw_ellipse(p->rng, n, 0.5, 0.5, 0.5, 0.25 * r, 0.25 * r, r, p->tx.data(), p->ty.data(), p->tz.data());
    separate = true;
  } else if (!strcmp(dist, "ellipsoid")) {
    draw_ellipse(p->rng, n, 0.0, 0.0, 0.0, 1.0, 1.0, 0.25, p->sx.data(), p->sy.data(), p->sz.data());
    separate = !coincide;
    if (separate)
      draw_ellipse(p->rng, n, 0.0, 0.0, 0.0, 1.0, 1.0, 0.25, p->tx.data(), p->ty.data(), p->tz.data());
  } else if (!strcmp(dist, "gaussmix")) {
    const Centres k = draw_centres(p->rng);
    draw_gaussmix(p->rng, k, n, p->sx.data(), p->sy.data(), p->sz.data());
    separate = !coincide;
    if (separate) draw_gaussmix(p->rng, k, n, p->tx.data(), p->ty.data(), p->tz.data());
  } else {
    return false;
  }
  if (!separate) {
    p->tx = p->sx; p->ty = p->sy; p->tz = p->sz;
  }
  p->coincident = separate ? 0 : 1;
  for (int i = 0; i < n; i++) p->den[i] = p->rng.next();
  return true;
}

/* ------------------------------------------------------------ octree */

inline double box_center(int pathc, int dep) {
  return -1.0 + (2.0 * 1.0) * (pathc + 0.5) / (double)(1 << dep); /* fmm/input.cc:123-137 */
}

/* Stable partition of perm[b,e) into 8 octants of the box centred at c;
 * returns the 9 octant offsets. */
void split_range(int* perm, int* tmp, int b, int e, const double* x, const double* y,
                 const double* z, const double c[3], int off[9]) {
  int cnt[8] = {0};
  for (int i = b; i < e; i++) {
    int k = perm[i];
    int o = ((x[k] >= c[0]) << 2) | ((y[k] >= c[1]) << 1) | (z[k] >= c[2]);
    tmp[i] = o;
    cnt[o]++;
  }
  off[0] = b;
  for (int o = 0; o < 8; o++) off[o + 1] = off[o] + cnt[o];
  int pos[8];
  for (int o = 0; o < 8; o++) pos[o] = off[o];
  /* tmp[b,e) holds octants; scatter perm through a second pass */
  std::vector<int> out(e - b);
  for (int i = b; i < e; i++) out[pos[tmp[i]]++ - b] = perm[i];
  std::memcpy(perm + b, out.data(), sizeof(int) * (e - b));
}

/* deepest level a tree may reach: box paths are ints and box centres
 * -1 + (2 path + 1) 2^-depth stay exact doubles well within it */
constexpr int MAX_LEVELS = 30;

bool build_tree(kifmm_problem* p) {
  const int n = p->n;
  std::vector<int>& par = p->parent;
  std::vector<int>& chd = p->child;
  std::vector<int>& dep = p->depth;
  std::vector<int>& pth = p->path;
  std::vector<int> rb, re; /* source ranges per node */
  p->sperm.resize(n);
  for (int i = 0; i < n; i++) p->sperm[i] = i;
  std::vector<int> tmp(n);
  par.assign(1, -1); chd.assign(1, -1); dep.assign(1, 0); pth.assign(3, 0);
  rb.assign(1, 0); re.assign(1, n);
  p->level.assign(1, 0);
  int beg = 0, end = 1;
  while (beg < end) {
    /* child slots are handed out in node order (fmm/input.cc:228-273) */
    std::vector<int> first(end - beg, -1);
    int next = end;
    for (int k = beg; k < end; k++)
      if (re[k] - rb[k] > p->q) { first[k - beg] = next; next += 8; }
    if (next > end && p->level.size() >= (size_t)MAX_LEVELS) {
      set_error("tree build: more than 30 levels (too many coincident points for the points per box)");
      return false;
    }
    par.resize(next); chd.resize(next); dep.resize(next); pth.resize(3 * (size_t)next);
    rb.resize(next); re.resize(next);
#pragma omp parallel for schedule(dynamic, 16)
    for (int k = beg; k < end; k++) {
      const int f = first[k - beg];
      if (f < 0) continue;
      chd[k] = f;
      const double c[3] = {box_center(pth[3 * k], dep[k]), box_center(pth[3 * k + 1], dep[k]),
                           box_center(pth[3 * k + 2], dep[k])};
      int off[9];
      split_range(p->sperm.data(), tmp.data(), rb[k], re[k], p->sx.data(), p->sy.data(), p->sz.data(), c, off);
      for (int o = 0; o < 8; o++) {
        const int ch = f + o;
        par[ch] = k; chd[ch] = -1; dep[ch] = dep[k] + 1;
        pth[3 * ch] = 2 * pth[3 * k] + ((o >> 2) & 1);
        pth[3 * ch + 1] = 2 * pth[3 * k + 1] + ((o >> 1) & 1);
        pth[3 * ch + 2] = 2 * pth[3 * k + 2] + (o & 1);
        rb[ch] = off[o]; re[ch] = off[o + 1];
      }
    }
    beg = end;
    end = next;
    p->level.push_back(beg);
  }
  const int nn = (int)par.size();
  p->tag.assign(nn, 0);
  p->sbeg.assign(nn, 0); p->snum.assign(nn, 0); p->tbeg.assign(nn, 0); p->tnum.assign(nn, 0);
  std::vector<int> sgath(n);
  int sum = 0;
  for (int i = 0; i < nn; i++) {
    const int cnt = re[i] - rb[i];
    if (cnt > 0) {
      p->tag[i] |= KIFMM_SRCNODE;
      if (chd[i] == -1) {
        p->sbeg[i] = sum; p->snum[i] = cnt;
        for (int k = 0; k < cnt; k++) sgath[sum + k] = p->sperm[rb[i] + k];
        sum += cnt;
      }
    }
  }
  p->sperm.swap(sgath);
  /* targets binned into the source tree (fmm/input.cc:386-460) */
  std::vector<int> tperm(n), tb(nn, 0), te(nn, 0);
  for (int i = 0; i < n; i++) tperm[i] = i;
  te[0] = n;
  for (size_t L = 0; L + 1 < p->level.size(); L++) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int k = p->level[L]; k < p->level[L + 1]; k++) {
      if (chd[k] == -1) continue;
      const double c[3] = {box_center(pth[3 * k], dep[k]), box_center(pth[3 * k + 1], dep[k]),
                           box_center(pth[3 * k + 2], dep[k])};
      int off[9];
      split_range(tperm.data(), tmp.data(), tb[k], te[k], p->tx.data(), p->ty.data(), p->tz.data(), c, off);
      for (int o = 0; o < 8; o++) { tb[chd[k] + o] = off[o]; te[chd[k] + o] = off[o + 1]; }
    }
  }
  std::vector<int> tgath(n);
  sum = 0;
  for (int i = 0; i < nn; i++) {
    const int cnt = te[i] - tb[i];
    if (cnt > 0) {
      p->tag[i] |= KIFMM_TRGNODE;
      if (chd[i] == -1) {
        p->tbeg[i] = sum; p->tnum[i] = cnt;
        for (int k = 0; k < cnt; k++) tgath[sum + k] = tperm[tb[i] + k];
        sum += cnt;
      }
    }
  }
  p->tperm.swap(tgath);
  return true;
}

/* ------------------------------------------------------------ lists */

/* deepest node on the way to box (d, path) with depth <= d (fmm/input.cc:152-165) */
inline int find_box(const kifmm_problem* p, int d, const int want[3]) {
  int cur = 0;
  while (p->depth[cur] < d && p->child[cur] != -1) {
    const int sh = d - p->depth[cur] - 1;
    const int o = (((want[0] >> sh) & 1) << 2) | (((want[1] >> sh) & 1) << 1) | ((want[2] >> sh) & 1);
    cur = p->child[cur] + o;
  }
  return cur;
}

/* boxes touch (share a face, edge or corner) without overlapping (fmm/input.cc:170-185) */
inline bool touching(const kifmm_problem* p, int a, int b) {
  const int md = std::max(p->depth[a], p->depth[b]);
  const int ra = 1 << (md - p->depth[a]), rbb = 1 << (md - p->depth[b]);
  const int rad = ra + rbb;
  int linf = 0;
  for (int d = 0; d < 3; d++) {
    const int dif = std::abs((2 * p->path[3 * a + d] + 1) * ra - (2 * p->path[3 * b + d] + 1) * rbb);
    if (dif > rad) return false;
    linf = std::max(linf, dif);
  }
  return linf == rad;
}

void node_lists(const kifmm_problem* p, int t, std::vector<int> out[4]) {
  for (int k = 0; k < 4; k++) out[k].clear();
  const bool leaf = p->child[t] == -1;
  const int d = p->depth[t];
  const int pa = p->parent[t];
  std::vector<int> queue;
  if (pa != -1) {
    const int lim = 1 << d;
    for (int i = -2; i < 4; i++)
      for (int j = -2; j < 4; j++)
        for (int k = -2; k < 4; k++) {
          const int w[3] = {2 * p->path[3 * pa] + i, 2 * p->path[3 * pa + 1] + j, 2 * p->path[3 * pa + 2] + k};
          if (w[0] < 0 || w[1] < 0 || w[2] < 0 || w[0] >= lim || w[1] >= lim || w[2] >= lim) continue;
          if (w[0] == p->path[3 * t] && w[1] == p->path[3 * t + 1] && w[2] == p->path[3 * t + 2]) continue;
          const int r = find_box(p, d, w);
          const bool adj = touching(p, r, t);
          if (p->depth[r] < d) {
            if (!adj) out[3].push_back(r);
            else if (leaf) out[0].push_back(r);
          } else if (!adj) {
            out[1].push_back(r);
          } else if (leaf) {
            queue.assign(1, r);
            for (size_t h = 0; h < queue.size(); h++) {
              const int f = queue[h];
              if (!touching(p, f, t)) out[2].push_back(f);
              else if (p->child[f] == -1) out[0].push_back(f);
              else for (int o = 0; o < 8; o++) queue.push_back(p->child[f] + o);
            }
          }
        }
  }
  if (leaf) out[0].push_back(t);
  for (int k = 0; k < 4; k++) {
    std::vector<int>& v = out[k];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    v.erase(std::remove_if(v.begin(), v.end(), [&](int s) { return !(p->tag[s] & KIFMM_SRCNODE); }), v.end());
  }
}

void build_lists(kifmm_problem* p) {
  const int nn = (int)p->parent.size();
  std::vector<std::vector<int>> per[4];
  for (int k = 0; k < 4; k++) per[k].assign(nn, std::vector<int>());
#pragma omp parallel
  {
    std::vector<int> tmp[4];
#pragma omp for schedule(dynamic, 64)
    for (int t = 0; t < nn; t++) {
      if (!(p->tag[t] & KIFMM_TRGNODE)) continue;
      node_lists(p, t, tmp);
      for (int k = 0; k < 4; k++) per[k][t] = tmp[k];
    }
  }
  for (int k = 0; k < 4; k++) {
    p->lptr[k].assign(nn + 1, 0);
    for (int t = 0; t < nn; t++) p->lptr[k][t + 1] = p->lptr[k][t] + (int)per[k][t].size();
    p->lidx[k].resize(p->lptr[k][nn]);
    for (int t = 0; t < nn; t++)
      std::copy(per[k][t].begin(), per[k][t].end(), p->lidx[k].begin() + p->lptr[k][t]);
  }
}

void gather(kifmm_problem* p) {
  const int n = p->n;
  p->gsx.resize(n); p->gsy.resize(n); p->gsz.resize(n); p->gden.resize(n);
  p->gtx.resize(n); p->gty.resize(n); p->gtz.resize(n);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; k++) {
    const int s = p->sperm[k], t = p->tperm[k];
    p->gsx[k] = p->sx[s]; p->gsy[k] = p->sy[s]; p->gsz[k] = p->sz[s]; p->gden[k] = p->den[s];
    p->gtx[k] = p->tx[t]; p->gty[k] = p->ty[t]; p->gtz[k] = p->tz[t];
  }
}

bool finish(kifmm_problem* p) {
  double t0 = wall_seconds();
  if (!build_tree(p)) return false;
  build_lists(p);
  gather(p);
  p->t_tree = wall_seconds() - t0;
  p->n_e = p->np * p->np * p->np - (p->np - 2) * (p->np - 2) * (p->np - 2);
  p->n_uc = (p->np + 2) * (p->np + 2) * (p->np + 2) - p->np * p->np * p->np;
  return true;
}

/* ------------------------------------------------------------ operators */

struct Surface {
  int n = 0;
  std::vector<double> x, y, z;
};

/* fmm/trans.cc:31-62: boundary nodes of an np^3 grid on [-R,R]^3, x slowest */
Surface surface(int np, double R) {
  Surface s;
  const double step = 2.0 / (np - 1);
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++) {
        if (!(i == 0 || i == np - 1 || j == 0 || j == np - 1 || k == 0 || k == np - 1)) continue;
        s.x.push_back(R * (-1.0 + i * step));
        s.y.push_back(R * (-1.0 + j * step));
        s.z.push_back(R * (-1.0 + k * step));
      }
  s.n = (int)s.x.size();
  return s;
}

/* exact FP64 Laplace matrix: mat[i*tn + j] = 1/(4 pi |x_i - y_j|) (fmm/omp_sse_block.cc:329-382) */
void laplace_matrix(const Surface& src, const Surface& trg, double* mat) {
  const double c = 1.0 / (4.0 * M_PI);
  for (int i = 0; i < src.n; i++)
    for (int j = 0; j < trg.n; j++) {
      const double dx = src.x[i] - trg.x[j], dy = src.y[i] - trg.y[j], dz = src.z[i] - trg.z[j];
      mat[(size_t)i * trg.n + j] = c / std::sqrt(dx * dx + dy * dy + dz * dz);
    }
}

/* Truncated pseudo-inverse R [n x m] of M [m x n] (column-major) keeping
 * singular values >= eps * sigma_max.  One-sided Jacobi on the columns of
 * M (or of M^T when m < n), accumulated in long double. */
void pinv_jacobi(const double* M, int m, int n, double eps, double* R) {
  const bool tr = m < n;
  const int rows = tr ? n : m, cols = tr ? m : n;
  std::vector<long double> A((size_t)rows * cols), V((size_t)cols * cols, 0.0L);
  for (int c = 0; c < cols; c++)
    for (int r = 0; r < rows; r++)
      A[(size_t)c * rows + r] = tr ? M[(size_t)r * m + c] : M[(size_t)c * m + r];
  for (int c = 0; c < cols; c++) V[(size_t)c * cols + c] = 1.0L;
  for (int sweep = 0; sweep < 60; sweep++) {
    long double off = 0.0L;
    for (int a = 0; a < cols - 1; a++)
      for (int b = a + 1; b < cols; b++) {
        long double* ca = &A[(size_t)a * rows];
        long double* cb = &A[(size_t)b * rows];
        long double al = 0, be = 0, ga = 0;
        for (int r = 0; r < rows; r++) { al += ca[r] * ca[r]; be += cb[r] * cb[r]; ga += ca[r] * cb[r]; }
        if (ga == 0.0L) continue;
        const long double rel = fabsl(ga) / sqrtl(al * be);
        off = std::max(off, rel);
        if (rel < 1e-17L) continue;
        const long double zeta = (be - al) / (2.0L * ga);
        const long double t = (zeta >= 0 ? 1.0L : -1.0L) / (fabsl(zeta) + sqrtl(1.0L + zeta * zeta));
        const long double cs = 1.0L / sqrtl(1.0L + t * t), sn = cs * t;
        for (int r = 0; r < rows; r++) {
          const long double u = ca[r], v = cb[r];
          ca[r] = cs * u - sn * v;
          cb[r] = sn * u + cs * v;
        }
        long double* va = &V[(size_t)a * cols];
        long double* vb = &V[(size_t)b * cols];
        for (int r = 0; r < cols; r++) {
          const long double u = va[r], v = vb[r];
          va[r] = cs * u - sn * v;
          vb[r] = sn * u + cs * v;
        }
      }
    if (off < 1e-17L) break;
  }
  std::vector<long double> s(cols);
  for (int c = 0; c < cols; c++) {
    long double nr = 0;
    for (int r = 0; r < rows; r++) nr += A[(size_t)c * rows + r] * A[(size_t)c * rows + r];
    s[c] = sqrtl(nr);
  }
  long double smax = 0;
  for (int c = 0; c < cols; c++) smax = std::max(smax, s[c]);
  const long double cut = (long double)eps * smax;
  /* A = U S (columns), so M' = U S V^T and pinv(M') = V S^-1 U^T = V S^-2 A^T */
  std::vector<long double> acc((size_t)n * m, 0.0L);
  for (int c = 0; c < cols; c++) {
    if (!(s[c] >= cut) || s[c] == 0.0L) continue;
    const long double w = 1.0L / (s[c] * s[c]);
    const long double* u = &A[(size_t)c * rows]; /* = s_c * U_c */
    const long double* v = &V[(size_t)c * cols];
    if (!tr) {
      /* pinv(M) [n x m] = sum_c V_c (A_c)^T / s_c^2 */
      for (int j = 0; j < m; j++) {
        const long double uj = u[j] * w;
        long double* col = &acc[(size_t)j * n];
        for (int i = 0; i < n; i++) col[i] += v[i] * uj;
      }
    } else {
      /* M^T = U S V^T  =>  pinv(M) = U S^-1 V^T = sum_c A_c V_c^T / s_c^2 */
      for (int j = 0; j < m; j++) {
        const long double vj = v[j] * w;
        long double* col = &acc[(size_t)j * n];
        for (int i = 0; i < n; i++) col[i] += u[i] * vj;
      }
    }
  }
  for (size_t i = 0; i < acc.size(); i++) R[i] = (double)acc[i];
}

/* The four operator sets depend on NP only: built once per NP per process
 * and copied into each problem that asks for them. */
struct OperatorSet {
  std::vector<double> uc2ue, ue2uc, dc2de, de2dc;
};
std::mutex g_ops_mutex;
std::map<int, std::shared_ptr<const OperatorSet>> g_ops_cache;

void compute_operators(kifmm_problem* p);

void build_operators(kifmm_problem* p) {
  const double t0 = wall_seconds();
  std::shared_ptr<const OperatorSet> hit;
  {
    std::lock_guard<std::mutex> lock(g_ops_mutex);
    auto it = g_ops_cache.find(p->np);
    if (it != g_ops_cache.end()) hit = it->second;
  }
  if (!hit) {
    compute_operators(p);
    auto set = std::make_shared<OperatorSet>();
    set->uc2ue = p->uc2ue; set->ue2uc = p->ue2uc; set->dc2de = p->dc2de; set->de2dc = p->de2dc;
    std::lock_guard<std::mutex> lock(g_ops_mutex);
    g_ops_cache.emplace(p->np, std::move(set));
  } else {
    p->uc2ue = hit->uc2ue; p->ue2uc = hit->ue2uc; p->dc2de = hit->dc2de; p->de2dc = hit->de2dc;
  }
  p->have_ops = true;
  p->t_ops = wall_seconds() - t0;
}

/* UC2UE, UE2UC, DC2DE, DE2DC for p->np (fmm/trans.cc:143-264) */
void compute_operators(kifmm_problem* p) {
  const int np = p->np;
  const Surface ue = surface(np, 1.0), uc = surface(np + 2, 3.0), de = surface(np, 3.0), dc = surface(np, 1.0);
  const int ne = ue.n, nuc = uc.n;
  const double eps = std::pow(0.1, np + 1); /* alt(), fmm/trans.cc:120-124 */
  p->uc2ue.assign((size_t)ne * nuc, 0.0);
  p->ue2uc.assign((size_t)8 * nuc * ne, 0.0);
  p->dc2de.assign((size_t)ne * ne, 0.0);
  p->de2dc.assign((size_t)8 * ne * ne, 0.0);
#pragma omp parallel sections
  {
#pragma omp section
    {
      std::vector<double> K((size_t)ne * nuc);
      laplace_matrix(ue, uc, K.data()); /* [n_uc x n_e] column-major */
      pinv_jacobi(K.data(), nuc, ne, eps, p->uc2ue.data());
    }
#pragma omp section
    {
      std::vector<double> K((size_t)ne * ne);
      laplace_matrix(de, dc, K.data());
      pinv_jacobi(K.data(), ne, ne, eps, p->dc2de.data());
    }
#pragma omp section
    {
      for (int o = 0; o < 8; o++) {
        const int ix = (o >> 2) & 1, iy = (o >> 1) & 1, iz = o & 1;
        Surface c2 = uc, d2 = ue;
        for (int j = 0; j < nuc; j++) { c2.x[j] = 2.0 * uc.x[j]; c2.y[j] = 2.0 * uc.y[j]; c2.z[j] = 2.0 * uc.z[j]; }
        for (int j = 0; j < ne; j++) {
          d2.x[j] = ue.x[j] + (2 * ix - 1); d2.y[j] = ue.y[j] + (2 * iy - 1); d2.z[j] = ue.z[j] + (2 * iz - 1);
        }
        laplace_matrix(d2, c2, &p->ue2uc[(size_t)o * nuc * ne]);
        Surface c3 = dc;
        for (int j = 0; j < ne; j++) {
          c3.x[j] = 0.5 * dc.x[j] + (ix - 0.5); c3.y[j] = 0.5 * dc.y[j] + (iy - 0.5); c3.z[j] = 0.5 * dc.z[j] + (iz - 0.5);
        }
        laplace_matrix(de, c3, &p->de2dc[(size_t)o * ne * ne]);
      }
    }
  }
}

kifmm_problem* checked_create(kifmm_problem* p) {
  try {
    if (!finish(p)) {
      delete p;
      return nullptr;
    }
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed while building the tree");
    delete p;
    return nullptr;
  }
  return p;
}

}  // namespace

extern "C" {

const char* kifmm_last_error(void) { return g_last_error.c_str(); }

int kifmm_getenv_accuracy(void) { return env_positive("NP", 6); }
int kifmm_getenv_validate(void) { return env_positive("NV", 1000); }
int kifmm_getenv_block_size(void) { return env_positive("BS", 49); }
int kifmm_getenv_upbs(void) { return env_positive("UPBS", 1000); }
int kifmm_getenv_coincide(void) {
  const char* s = getenv("C");
  if (!s) return 0;
  if (!strcasecmp(s, "y") || !strcasecmp(s, "yes") || !strcasecmp(s, "t") || !strcasecmp(s, "true")) return 1;
  int v = 0;
  if (sscanf(s, "%d", &v) != 1) return 0;
  return v ? 1 : 0;
}
long kifmm_getenv_seed(void) {
  const char* s = getenv("SEED");
  if (!s) return 26;
  return atol(s);
}

kifmm_problem* kifmm_problem_create(int n, const char* distribution, int pts_max,
                                    const kifmm_problem_opts* opts) {
  kifmm_problem* p = kifmm::problem_generate(n, distribution, pts_max, opts);
  return p ? checked_create(p) : nullptr;
}

kifmm_problem* kifmm_problem_from_points(int n, const double* sx, const double* sy, const double* sz,
                                         const double* den, const double* tx, const double* ty,
                                         const double* tz, int pts_max, int np) {
  kifmm_problem* p = kifmm::problem_points(n, sx, sy, sz, den, tx, ty, tz, pts_max, np);
  return p ? checked_create(p) : nullptr;
}

}  // extern "C"

/* Points of a new problem without its tree (the generators of
 * kifmm_problem_create), for builders outside this file. */
kifmm_problem* kifmm::problem_generate(int n, const char* distribution, int pts_max, const kifmm_problem_opts* opts) {
  if (n <= 0 || pts_max <= 0 || !distribution) {
    set_error("kifmm_problem_create: n and pts_max must be positive");
    return nullptr;
  }
  kifmm_problem* p = new (std::nothrow) kifmm_problem();
  if (!p) { set_error("host allocation failed"); return nullptr; }
  p->n = n;
  p->q = pts_max;
  p->np = (opts && opts->np > 0) ? opts->np : kifmm_getenv_accuracy();
  if (p->np < 3) { set_error("NP must be >= 3"); delete p; return nullptr; }
  const long seed = (opts && opts->seed >= 0) ? opts->seed : kifmm_getenv_seed();
  const int coincide = (opts && opts->coincide >= 0) ? opts->coincide : kifmm_getenv_coincide();
  if (opts && opts->nthreads > 0) omp_set_num_threads(opts->nthreads);
  const double t0 = wall_seconds();
  try {
    if (!generate(p, distribution, seed, coincide)) {
      set_error(std::string("invalid distribution '") + distribution +
                "' (uniform, ellipseUniformAngles, ellipsoid, gaussmix)");
      delete p;
      return nullptr;
    }
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed while generating points");
    delete p;
    return nullptr;
  }
  p->t_gen = wall_seconds() - t0;
  return p;
}

kifmm_problem* kifmm::problem_points(int n, const double* sx, const double* sy, const double* sz, const double* den,
                                     const double* tx, const double* ty, const double* tz, int pts_max, int np) {
  if (n <= 0 || pts_max <= 0 || np < 3 || !sx || !sy || !sz || !den || !tx || !ty || !tz) {
    set_error("kifmm_problem_from_points: invalid argument");
    return nullptr;
  }
  for (int i = 0; i < n; i++) {
    const double v[6] = {sx[i], sy[i], sz[i], tx[i], ty[i], tz[i]};
    for (double c : v)
      if (!(c >= -1.0 && c <= 1.0)) {
        set_error("kifmm_problem_from_points: coordinates must lie in [-1,1]");
        return nullptr;
      }
  }
  kifmm_problem* p = new (std::nothrow) kifmm_problem();
  if (!p) { set_error("host allocation failed"); return nullptr; }
  p->n = n; p->q = pts_max; p->np = np;
  p->rng.seed(26);
  p->sx.assign(sx, sx + n); p->sy.assign(sy, sy + n); p->sz.assign(sz, sz + n);
  p->den.assign(den, den + n);
  p->tx.assign(tx, tx + n); p->ty.assign(ty, ty + n); p->tz.assign(tz, tz + n);
  p->coincident = (sx == tx && sy == ty && sz == tz) ? 1 : 0;
  return p;
}

void kifmm::problem_coords(const kifmm_problem* p, const double* s3[3], const double* t3[3], int* n, int* q) {
  s3[0] = p->sx.data(); s3[1] = p->sy.data(); s3[2] = p->sz.data();
  t3[0] = p->tx.data(); t3[1] = p->ty.data(); t3[2] = p->tz.data();
  *n = p->n;
  *q = p->q;
}

void kifmm::problem_install(kifmm_problem* p, TreeArrays&& t, double seconds) {
  p->level.swap(t.level); p->parent.swap(t.parent); p->child.swap(t.child); p->depth.swap(t.depth);
  p->tag.swap(t.tag); p->path.swap(t.path); p->sbeg.swap(t.sbeg); p->snum.swap(t.snum);
  p->tbeg.swap(t.tbeg); p->tnum.swap(t.tnum); p->sperm.swap(t.sperm); p->tperm.swap(t.tperm);
  for (int k = 0; k < 4; k++) { p->lptr[k].swap(t.lptr[k]); p->lidx[k].swap(t.lidx[k]); }
  const double t0 = wall_seconds();
  gather(p);
  p->t_tree = seconds + (wall_seconds() - t0);
  p->n_e = p->np * p->np * p->np - (p->np - 2) * (p->np - 2) * (p->np - 2);
  p->n_uc = (p->np + 2) * (p->np + 2) * (p->np + 2) - p->np * p->np * p->np;
}

extern "C" {

/* Point files: 8-byte magic "KIFMMPT1", int64 n, int32 has_targets,
 * int32 reserved (0), then n little-endian doubles each of sx, sy, sz, den
 * and, when has_targets, tx, ty, tz.  Coordinates in the root box
 * [-1,1]^3; without targets the sources are the targets (C = 1). */
static const char kPtsMagic[8] = {'K', 'I', 'F', 'M', 'M', 'P', 'T', '1'};

int kifmm_points_write(const char* path, int n, const double* sx, const double* sy, const double* sz,
                       const double* den, const double* tx, const double* ty, const double* tz) {
  if (!path || n <= 0 || !sx || !sy || !sz || !den || ((tx || ty || tz) && !(tx && ty && tz))) {
    set_error("kifmm_points_write: invalid argument");
    return KIFMM_EINVAL;
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) { set_error(std::string("kifmm_points_write: cannot open ") + path); return KIFMM_EINVAL; }
  const long long nn = n;
  const int hdr[2] = {tx ? 1 : 0, 0};
  bool ok = std::fwrite(kPtsMagic, 1, 8, f) == 8 && std::fwrite(&nn, sizeof nn, 1, f) == 1 &&
            std::fwrite(hdr, sizeof(int), 2, f) == 2;
  const double* arrs[7] = {sx, sy, sz, den, tx, ty, tz};
  for (int k = 0; k < (tx ? 7 : 4) && ok; k++) ok = std::fwrite(arrs[k], sizeof(double), n, f) == (size_t)n;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) { set_error("kifmm_points_write: write failed"); return KIFMM_EINVAL; }
  return KIFMM_OK;
}

kifmm_problem* kifmm_problem_from_file(const char* path, int pts_max, int np) {
  if (!path) { set_error("kifmm_problem_from_file: null path"); return nullptr; }
  FILE* f = std::fopen(path, "rb");
  if (!f) { set_error(std::string("kifmm_problem_from_file: cannot open ") + path); return nullptr; }
  char magic[8];
  long long nn = 0;
  int hdr[2] = {0, 0};
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kPtsMagic, 8) == 0 &&
            std::fread(&nn, sizeof nn, 1, f) == 1 && std::fread(hdr, sizeof(int), 2, f) == 2 && nn > 0 &&
            nn <= (1LL << 31) - 1 && (hdr[0] == 0 || hdr[0] == 1);
  const int n = ok ? (int)nn : 0;
  std::vector<double> a[7];
  for (int k = 0; k < (hdr[0] ? 7 : 4) && ok; k++) {
    a[k].resize(n);
    ok = std::fread(a[k].data(), sizeof(double), n, f) == (size_t)n;
  }
  std::fclose(f);
  if (!ok) { set_error(std::string("kifmm_problem_from_file: malformed or truncated file ") + path); return nullptr; }
  if (!hdr[0]) return kifmm_problem_from_points(n, a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[0].data(),
                                                a[1].data(), a[2].data(), pts_max, np);
  return kifmm_problem_from_points(n, a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data(), a[5].data(),
                                   a[6].data(), pts_max, np);
}

void kifmm_problem_destroy(kifmm_problem* p) { delete p; }

int kifmm_problem_view(const kifmm_problem* p, kifmm_tree_view* v) {
  if (!p || !v) { set_error("kifmm_problem_view: null argument"); return KIFMM_EINVAL; }
  v->np = p->np;
  v->n_nodes = (int)p->parent.size();
  v->n_levels = (int)p->level.size();
  v->level_start = p->level.data();
  v->parent = p->parent.data(); v->child = p->child.data(); v->depth = p->depth.data();
  v->tag = p->tag.data(); v->path = p->path.data();
  v->src_beg = p->sbeg.data(); v->src_num = p->snum.data();
  v->trg_beg = p->tbeg.data(); v->trg_num = p->tnum.data();
  v->u_ptr = p->lptr[0].data(); v->u_idx = p->lidx[0].data();
  v->v_ptr = p->lptr[1].data(); v->v_idx = p->lidx[1].data();
  v->w_ptr = p->lptr[2].data(); v->w_idx = p->lidx[2].data();
  v->x_ptr = p->lptr[3].data(); v->x_idx = p->lidx[3].data();
  v->n_src = p->n; v->n_trg = p->n;
  v->sx = p->gsx.data(); v->sy = p->gsy.data(); v->sz = p->gsz.data(); v->sden = p->gden.data();
  v->tx = p->gtx.data(); v->ty = p->gty.data(); v->tz = p->gtz.data();
  v->coincident = p->coincident;
  return KIFMM_OK;
}

int kifmm_problem_operators(kifmm_problem* p, kifmm_operators* o) {
  if (!p || !o) { set_error("kifmm_problem_operators: null argument"); return KIFMM_EINVAL; }
  if (!p->have_ops) {
    try {
      build_operators(p);
    } catch (const std::bad_alloc&) {
      set_error("host allocation failed while building operators");
      return KIFMM_ENOMEM;
    }
  }
  o->np = p->np; o->n_e = p->n_e; o->n_uc = p->n_uc;
  o->uc2ue = p->uc2ue.data(); o->ue2uc = p->ue2uc.data();
  o->dc2de = p->dc2de.data(); o->de2dc = p->de2dc.data();
  return KIFMM_OK;
}

int kifmm_problem_scatter(const kifmm_problem* p, const double* pot_leaf, double* pot_orig) {
  if (!p || !pot_leaf || !pot_orig) { set_error("kifmm_problem_scatter: null argument"); return KIFMM_EINVAL; }
  for (int k = 0; k < p->n; k++) pot_orig[p->tperm[k]] = pot_leaf[k];
  return KIFMM_OK;
}

int kifmm_problem_points(const kifmm_problem* p, double* sx, double* sy, double* sz, double* den,
                         double* tx, double* ty, double* tz) {
  if (!p) { set_error("kifmm_problem_points: null problem"); return KIFMM_EINVAL; }
  const size_t b = sizeof(double) * p->n;
  if (sx) std::memcpy(sx, p->sx.data(), b);
  if (sy) std::memcpy(sy, p->sy.data(), b);
  if (sz) std::memcpy(sz, p->sz.data(), b);
  if (den) std::memcpy(den, p->den.data(), b);
  if (tx) std::memcpy(tx, p->tx.data(), b);
  if (ty) std::memcpy(ty, p->ty.data(), b);
  if (tz) std::memcpy(tz, p->tz.data(), b);
  return KIFMM_OK;
}

}  // extern "C"

int kifmm::sample_targets(kifmm_problem* p, int nv, int* ids, double* xyz) {
  if (!p || nv <= 0 || !ids || !xyz) return KIFMM_EINVAL;
  for (int k = 0; k < nv; k++) ids[k] = (int)std::floor(p->rng.next() * p->n);
  for (int k = 0; k < nv; k++) {
    xyz[k] = p->tx[ids[k]];
    xyz[nv + k] = p->ty[ids[k]];
    xyz[2 * nv + k] = p->tz[ids[k]];
  }
  return KIFMM_OK;
}

extern "C" {

double kifmm_problem_validate(kifmm_problem* p, const double* pot_orig, int nv, int* chk_ids,
                              double* direct_out) {
  if (!p || !pot_orig || nv <= 0) { set_error("kifmm_problem_validate: invalid argument"); return -1.0; }
  std::vector<int> ids(nv);
  for (int k = 0; k < nv; k++) ids[k] = (int)std::floor(p->rng.next() * p->n);
  std::vector<double> direct(nv);
  const double c = 1.0 / (4.0 * M_PI);
#pragma omp parallel for schedule(dynamic, 4)
  for (int k = 0; k < nv; k++) {
    const double x = p->tx[ids[k]], y = p->ty[ids[k]], z = p->tz[ids[k]];
    double s = 0.0;
    for (int j = 0; j < p->n; j++) {
      const double dx = x - p->sx[j], dy = y - p->sy[j], dz = z - p->sz[j];
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 != 0.0) s += p->den[j] / std::sqrt(r2);
    }
    direct[k] = c * s;
  }
  double vn = 0, en = 0;
  for (int k = 0; k < nv; k++) {
    const double f = pot_orig[ids[k]];
    vn += f * f;
    en += (direct[k] - f) * (direct[k] - f);
  }
  if (chk_ids) std::memcpy(chk_ids, ids.data(), sizeof(int) * nv);
  if (direct_out) std::memcpy(direct_out, direct.data(), sizeof(double) * nv);
  return std::sqrt(en) / std::sqrt(vn);
}

int kifmm_problem_setup_times(const kifmm_problem* p, double* t) {
  if (!p || !t) return KIFMM_EINVAL;
  t[0] = p->t_gen; t[1] = p->t_tree; t[2] = p->t_ops;
  return KIFMM_OK;
}

int kifmm_surface(int np, double radius, double* xyz) {
  if (np < 2 || !xyz) return KIFMM_EINVAL;
  Surface s = surface(np, radius);
  for (int i = 0; i < s.n; i++) { xyz[i] = s.x[i]; xyz[s.n + i] = s.y[i]; xyz[2 * s.n + i] = s.z[i]; }
  return s.n;
}

int kifmm_pinv(const double* M, int m, int n, double eps, double* R) {
  if (!M || !R || m <= 0 || n <= 0) return KIFMM_EINVAL;
  pinv_jacobi(M, m, n, eps, R);
  return KIFMM_OK;
}

int kifmm_drand48(long seed, int k, double* out) {
  if (!out || k < 0) return KIFMM_EINVAL;
  Rand48 r;
  r.seed(seed);
  for (int i = 0; i < k; i++) out[i] = r.next();
  return KIFMM_OK;
}

}  // extern "C"
