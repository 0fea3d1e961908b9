/*
 * kifmm_gpu_tree.hip -- the source/target octree, the U/V/W/X interaction
 * lists and the gather permutations, built on the GPU.
 *
 * Same tree, lists and point order as the host builder (which follows
 * src_tree / trg_tree / the list construction of fmm/input.cc:203-460):
 *   - nodes in BFS level order, the 8 children of a node contiguous (octant
 *     4x + 2y + z), a node split while it holds more than q sources;
 *   - targets binned into the source tree level by level;
 *   - leaf point ranges in node order, points inside a leaf in input order;
 *   - per target node, U (adjacent source leaves, itself when a leaf), V
 *     (children of the parent's colleagues that do not touch it), W and X
 *     (the coarse/fine non-adjacent pairs), ascending, unique, source nodes
 *     only.
 * Level by level: split flags and an exclusive scan hand out child slots in
 * node order; every point finds its octant against the box centre (the
 * same exact dyadic arithmetic as the host) and moves to the child, counts
 * via integer atomics (order-free).  The gather order is a stable radix
 * sort of the input index by final leaf.  Lists: a counting pass, a fill
 * pass writing (node, entry) keys, one radix sort per list, then unique +
 * source-node filter + compaction.  The host calls return once everything
 * is copied back into the problem.
 */
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "kifmm_hip.h"
#include "../host/kifmm_internal.h"

using kifmm::set_error;

namespace {

constexpr int MAX_DEPTH = 30;
constexpr int STACK = 8 * MAX_DEPTH + 8;

__device__ __forceinline__ double centre(int pathc, int dep) {
  return -1.0 + 2.0 * (pathc + 0.5) / (double)(1 << dep); /* exact: dyadic */
}

/* points whose node lies in [beg, end) and was split move to a child, and
 * count into it.  The children of the level are nodes c0 .. c0 + nc - 1.
 * Up to ASSIGN_LDS of them (the top levels, where a few counters take every
 * point: the first level's 16M increments of 8 counters took 62 ms) the
 * block counts into an LDS histogram over its grid-stride share of the
 * points and adds each nonzero bin with one atomic; deeper levels add per
 * point.  Integer counts: the same either way. */
constexpr int ASSIGN_LDS = 4096;
__global__ __launch_bounds__(256) void k_assign(int n, const double* __restrict__ x, const double* __restrict__ y,
                                                const double* __restrict__ z, int* __restrict__ node, int beg, int end,
                                                const int* __restrict__ child, const int* __restrict__ path,
                                                const int* __restrict__ depth, int* __restrict__ count, int c0, int nc) {
  __shared__ int hist[ASSIGN_LDS];
  const bool local = nc <= ASSIGN_LDS; /* block-uniform */
  if (local) {
    for (int e = threadIdx.x; e < nc; e += blockDim.x) hist[e] = 0;
    __syncthreads();
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int k = node[i];
    if (k < beg || k >= end) continue;
    const int c = child[k];
    if (c < 0) continue;
    const int d = depth[k];
    const int o = ((x[i] >= centre(path[3 * k], d)) << 2) | ((y[i] >= centre(path[3 * k + 1], d)) << 1) |
                  (z[i] >= centre(path[3 * k + 2], d));
    node[i] = c + o;
    if (local) atomicAdd(&hist[c + o - c0], 1);
    else atomicAdd(&count[c + o], 1);
  }
  if (local) {
    __syncthreads();
    for (int e = threadIdx.x; e < nc; e += blockDim.x)
      if (hist[e] != 0) atomicAdd(&count[c0 + e], hist[e]);
  }
}
/* grid of k_assign: with the LDS histogram few enough blocks that its flush
 * stays small against the points */
inline int assign_blocks(int n, int nc) {
  const int b = (int)std::max<long long>(1, ((long long)n + 255) / 256);
  return nc <= ASSIGN_LDS ? std::min(b, nc <= 512 ? 1024 : 256) : b;
}

__global__ void k_split_flags(int beg, int end, int q, const int* __restrict__ count, int* __restrict__ flag) {
  const int k = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (k < end) flag[k - beg] = count[k] > q ? 1 : 0;
}

__global__ void k_children(int beg, int end, const int* __restrict__ flag, const int* __restrict__ rank,
                           int* __restrict__ parent, int* __restrict__ child, int* __restrict__ depth,
                           int* __restrict__ path, int* __restrict__ count, int* __restrict__ tcount) {
  const int k = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= end || !flag[k - beg]) return;
  const int f = end + 8 * rank[k - beg];
  child[k] = f;
  for (int o = 0; o < 8; o++) {
    const int c = f + o;
    parent[c] = k;
    child[c] = -1;
    depth[c] = depth[k] + 1;
    path[3 * c] = 2 * path[3 * k] + ((o >> 2) & 1);
    path[3 * c + 1] = 2 * path[3 * k + 1] + ((o >> 1) & 1);
    path[3 * c + 2] = 2 * path[3 * k + 2] + (o & 1);
    count[c] = 0;
    tcount[c] = 0;
  }
}

/* per node: tag, and the leaf range sizes to be scanned */
__global__ void k_tags(int nn, const int* __restrict__ child, const int* __restrict__ count,
                       const int* __restrict__ tcount, int* __restrict__ tag, int* __restrict__ sl,
                       int* __restrict__ tl) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nn) return;
  tag[k] = (count[k] > 0 ? KIFMM_SRCNODE : 0) | (tcount[k] > 0 ? KIFMM_TRGNODE : 0);
  const bool leaf = child[k] == -1;
  sl[k] = leaf ? count[k] : 0;
  tl[k] = leaf ? tcount[k] : 0;
}

__global__ void k_iota(int n, int* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

/* ---------------------------------------------------------------- lists */

struct TreeDev {
  const int *parent, *child, *depth, *path, *tag;
};

__device__ __forceinline__ int find_box(const TreeDev& T, int d, int w0, int w1, int w2) {
  int cur = 0;
  while (T.depth[cur] < d && T.child[cur] != -1) {
    const int sh = d - T.depth[cur] - 1;
    const int o = (((w0 >> sh) & 1) << 2) | (((w1 >> sh) & 1) << 1) | ((w2 >> sh) & 1);
    cur = T.child[cur] + o;
  }
  return cur;
}

__device__ __forceinline__ bool touching(const TreeDev& T, int a, int b) {
  const int da = T.depth[a], db = T.depth[b];
  const int md = max(da, db);
  const int ra = 1 << (md - da), rb = 1 << (md - db), rad = ra + rb;
  int linf = 0;
  for (int d = 0; d < 3; d++) {
    const int dif = abs((2 * T.path[3 * a + d] + 1) * ra - (2 * T.path[3 * b + d] + 1) * rb);
    if (dif > rad) return false;
    linf = max(linf, dif);
  }
  return linf == rad;
}

/* The host builder's list enumeration for target node t.  FILL = false
 * counts entries (duplicates included) per list, FILL = true writes
 * (t, entry) keys from the offsets the counts produced. */
template <bool FILL>
__global__ void k_lists(int nn, TreeDev T, int* __restrict__ cnt, const long long* __restrict__ off,
                        unsigned long long* __restrict__ keys0, unsigned long long* __restrict__ keys1,
                        unsigned long long* __restrict__ keys2, unsigned long long* __restrict__ keys3,
                        int* __restrict__ overflow) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nn) return;
  int c[4] = {0, 0, 0, 0};
  unsigned long long* keys[4] = {keys0, keys1, keys2, keys3};
  long long pos[4] = {0, 0, 0, 0};
  if (FILL)
    for (int l = 0; l < 4; l++) pos[l] = off[(size_t)l * nn + t];
  auto push = [&](int l, int s) {
    if (FILL) keys[l][pos[l] + c[l]] = ((unsigned long long)t << 32) | (unsigned)s;
    c[l]++;
  };
  if (T.tag[t] & KIFMM_TRGNODE) {
    const bool leaf = T.child[t] == -1;
    const int d = T.depth[t], pa = T.parent[t];
    if (pa != -1) {
      const int lim = 1 << d;
      int st[STACK];
      for (int i = -2; i < 4; i++)
        for (int j = -2; j < 4; j++)
          for (int k = -2; k < 4; k++) {
            const int w0 = 2 * T.path[3 * pa] + i, w1 = 2 * T.path[3 * pa + 1] + j, w2 = 2 * T.path[3 * pa + 2] + k;
            if (w0 < 0 || w1 < 0 || w2 < 0 || w0 >= lim || w1 >= lim || w2 >= lim) continue;
            if (w0 == T.path[3 * t] && w1 == T.path[3 * t + 1] && w2 == T.path[3 * t + 2]) continue;
            const int r = find_box(T, d, w0, w1, w2);
            const bool adj = touching(T, r, t);
            if (T.depth[r] < d) {
              if (!adj) push(3, r);
              else if (leaf) push(0, r);
            } else if (!adj) {
              push(1, r);
            } else if (leaf) {
              int sp = 0;
              st[sp++] = r;
              while (sp > 0) {
                const int f = st[--sp];
                if (!touching(T, f, t)) push(2, f);
                else if (T.child[f] == -1) push(0, f);
                else if (sp + 8 <= STACK)
                  for (int o = 7; o >= 0; o--) st[sp++] = T.child[f] + o;
                else
                  atomicOr(overflow, 1);
              }
            }
          }
    }
    if (leaf) push(0, t);
  }
  if (!FILL)
    for (int l = 0; l < 4; l++) cnt[(size_t)l * nn + t] = c[l];
}

/* keep the first of each run of equal keys whose entry is a source node */
__global__ void k_keep(long long m, const unsigned long long* __restrict__ keys, const int* __restrict__ tag,
                       int* __restrict__ keep, int* __restrict__ per_node) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const unsigned long long k = keys[i];
  const int s = (int)(k & 0xffffffffu), t = (int)(k >> 32);
  const int kp = (i == 0 || keys[i - 1] != k) && (tag[s] & KIFMM_SRCNODE);
  keep[i] = kp;
  if (kp) atomicAdd(&per_node[t], 1);
}

__global__ void k_compact(long long m, const unsigned long long* __restrict__ keys, const int* __restrict__ keep,
                          const int* __restrict__ pos, int* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && keep[i]) out[pos[i]] = (int)(keys[i] & 0xffffffffu);
}

inline int blocks(long long n, int t) { return (int)std::max<long long>(1, (n + t - 1) / t); }

/* ------------------------------------------------------------ host side */

struct Dev {
  std::vector<void*> ptrs;
  ~Dev() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  T* alloc(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return (T*)p;
  }
  void release(void* p) {
    auto it = std::find(ptrs.begin(), ptrs.end(), p);
    if (it != ptrs.end()) {
      (void)hipFree(p);
      ptrs.erase(it);
    }
  }
};

#define GT_TRY(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      set_error(std::string("GPU tree build: ") + #expr + ": " + hipGetErrorString(e_)); \
      return false;                                                                     \
    }                                                                                   \
  } while (0)
#define GT_NEED(p)                                             \
  do {                                                         \
    if (!(p)) {                                                \
      set_error("GPU tree build: device allocation failed");  \
      return false;                                            \
    }                                                          \
  } while (0)

template <typename T>
bool grow(Dev& D, T*& p, size_t have, size_t want, hipStream_t st) {
  T* q = D.alloc<T>(want);
  GT_NEED(q);
  if (p && have) GT_TRY(hipMemcpyAsync(q, p, sizeof(T) * have, hipMemcpyDeviceToDevice, st));
  if (p) {
    GT_TRY(hipStreamSynchronize(st));
    D.release(p);
  }
  p = q;
  return true;
}

bool scan(Dev& D, const int* in, int* out, int n, hipStream_t st) {
  size_t tb = 0;
  GT_TRY(rocprim::exclusive_scan(nullptr, tb, in, out, 0, (size_t)n, rocprim::plus<int>(), st));
  void* tmp = D.alloc<char>(tb);
  GT_NEED(tmp);
  GT_TRY(rocprim::exclusive_scan(tmp, tb, in, out, 0, (size_t)n, rocprim::plus<int>(), st));
  GT_TRY(hipStreamSynchronize(st));
  D.release(tmp);
  return true;
}

/* stable order of 0..n-1 by key (the final leaf of each point) */
bool sort_by_leaf(Dev& D, const int* key, int n, int nbits, int* perm, hipStream_t st) {
  int* iota = D.alloc<int>(n);
  int* kout = D.alloc<int>(n);
  GT_NEED(iota);
  GT_NEED(kout);
  hipLaunchKernelGGL(k_iota, dim3(blocks(n, 256)), dim3(256), 0, st, n, iota);
  size_t tb = 0;
  GT_TRY(rocprim::radix_sort_pairs(nullptr, tb, key, kout, iota, perm, n, 0u, (unsigned)nbits, st));
  void* tmp = D.alloc<char>(tb);
  GT_NEED(tmp);
  GT_TRY(rocprim::radix_sort_pairs(tmp, tb, key, kout, iota, perm, n, 0u, (unsigned)nbits, st));
  GT_TRY(hipStreamSynchronize(st));
  D.release(tmp);
  D.release(iota);
  D.release(kout);
  return true;
}

template <typename T>
bool down(std::vector<T>& v, const T* d, size_t n, hipStream_t st) {
  v.resize(n);
  if (n) GT_TRY(hipMemcpyAsync(v.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost, st));
  return true;
}

bool build(int device, kifmm_problem* p, double* seconds) {
  const double t0 = kifmm::wall_seconds();
  const double* s3[3];
  const double* t3[3];
  int n = 0, q = 0;
  kifmm::problem_coords(p, s3, t3, &n, &q);
  GT_TRY(hipSetDevice(device));
  hipStream_t st = nullptr;
  GT_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } sg{st};
  Dev D;
  double* xyz[6];
  for (int a = 0; a < 6; a++) {
    xyz[a] = D.alloc<double>(n);
    GT_NEED(xyz[a]);
    GT_TRY(hipMemcpyAsync(xyz[a], a < 3 ? s3[a] : t3[a - 3], sizeof(double) * n, hipMemcpyHostToDevice, st));
  }
  int* node = D.alloc<int>(n);
  int* tnode = D.alloc<int>(n);
  GT_NEED(node);
  GT_NEED(tnode);
  GT_TRY(hipMemsetAsync(node, 0, sizeof(int) * n, st));
  GT_TRY(hipMemsetAsync(tnode, 0, sizeof(int) * n, st));
  /* node arrays, grown as levels are added */
  size_t cap = 0;
  int *parent = nullptr, *child = nullptr, *depth = nullptr, *path = nullptr, *count = nullptr, *tcount = nullptr;
  auto ensure = [&](size_t want) -> bool {
    if (want <= cap) return true;
    const size_t nc = std::max(want, 2 * cap + 64);
    if (!grow(D, parent, cap, nc, st) || !grow(D, child, cap, nc, st) || !grow(D, depth, cap, nc, st) ||
        !grow(D, path, 3 * cap, 3 * nc, st) || !grow(D, count, cap, nc, st) || !grow(D, tcount, cap, nc, st))
      return false;
    cap = nc;
    return true;
  };
  if (!ensure(1)) return false;
  const int root[4] = {-1, -1, 0, n};
  GT_TRY(hipMemcpyAsync(parent, &root[0], sizeof(int), hipMemcpyHostToDevice, st));
  GT_TRY(hipMemcpyAsync(child, &root[1], sizeof(int), hipMemcpyHostToDevice, st));
  GT_TRY(hipMemcpyAsync(depth, &root[2], sizeof(int), hipMemcpyHostToDevice, st));
  GT_TRY(hipMemsetAsync(path, 0, sizeof(int) * 3, st));
  GT_TRY(hipMemcpyAsync(count, &root[3], sizeof(int), hipMemcpyHostToDevice, st));
  GT_TRY(hipMemcpyAsync(tcount, &root[3], sizeof(int), hipMemcpyHostToDevice, st));
  kifmm::TreeArrays T;
  T.level.assign(1, 0);
  int beg = 0, end = 1, lv = 0;
  for (;;) {
    const int m = end - beg;
    int* flag = D.alloc<int>(m);
    int* rank = D.alloc<int>(m);
    GT_NEED(flag);
    GT_NEED(rank);
    hipLaunchKernelGGL(k_split_flags, dim3(blocks(m, 256)), dim3(256), 0, st, beg, end, q, count, flag);
    if (!scan(D, flag, rank, m, st)) return false;
    int last[2];
    GT_TRY(hipMemcpyAsync(&last[0], flag + m - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    GT_TRY(hipMemcpyAsync(&last[1], rank + m - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    GT_TRY(hipStreamSynchronize(st));
    const int nsplit = last[0] + last[1];
    T.level.push_back(end);
    if (nsplit == 0) break;
    if (lv + 1 >= MAX_DEPTH) {
      set_error("GPU tree build: more than 30 levels (too many coincident points for the points per box)");
      return false;
    }
    const int next = end + 8 * nsplit;
    if (!ensure((size_t)next)) return false;
    hipLaunchKernelGGL(k_children, dim3(blocks(m, 256)), dim3(256), 0, st, beg, end, flag, rank, parent, child, depth,
                       path, count, tcount);
    hipLaunchKernelGGL(k_assign, dim3(assign_blocks(n, next - end)), dim3(256), 0, st, n, xyz[0], xyz[1], xyz[2],
                       node, beg, end, child, path, depth, count, end, next - end);
    GT_TRY(hipGetLastError());
    D.release(flag);
    D.release(rank);
    beg = end;
    end = next;
    lv++;
  }
  const int nn = end;
  /* targets binned into the source tree, level by level */
  for (size_t L = 0; L + 2 < T.level.size(); L++) {
    const int c0 = T.level[L + 1], nc = T.level[L + 2] - T.level[L + 1]; /* the children of level L */
    hipLaunchKernelGGL(k_assign, dim3(assign_blocks(n, nc)), dim3(256), 0, st, n, xyz[3], xyz[4], xyz[5], tnode,
                       T.level[L], T.level[L + 1], child, path, depth, tcount, c0, nc);
  }
  int* tag = D.alloc<int>(nn);
  int* sl = D.alloc<int>(nn);
  int* tl = D.alloc<int>(nn);
  int* sb = D.alloc<int>(nn);
  int* tb = D.alloc<int>(nn);
  GT_NEED(tag);
  GT_NEED(sl);
  GT_NEED(tl);
  GT_NEED(sb);
  GT_NEED(tb);
  hipLaunchKernelGGL(k_tags, dim3(blocks(nn, 256)), dim3(256), 0, st, nn, child, count, tcount, tag, sl, tl);
  GT_TRY(hipGetLastError());
  if (!scan(D, sl, sb, nn, st) || !scan(D, tl, tb, nn, st)) return false;
  int nbits = 1;
  while ((1 << nbits) < nn) nbits++;
  int* sperm = D.alloc<int>(n);
  int* tperm = D.alloc<int>(n);
  GT_NEED(sperm);
  GT_NEED(tperm);
  if (!sort_by_leaf(D, node, n, nbits, sperm, st) || !sort_by_leaf(D, tnode, n, nbits, tperm, st)) return false;
  /* lists: count, offsets, fill, sort, unique + source filter, compact */
  TreeDev TD{parent, child, depth, path, tag};
  int* cnt = D.alloc<int>(4 * (size_t)nn);
  int* ovf = D.alloc<int>(1);
  GT_NEED(cnt);
  GT_NEED(ovf);
  GT_TRY(hipMemsetAsync(ovf, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_lists<false>, dim3(blocks(nn, 64)), dim3(64), 0, st, nn, TD, cnt, (const long long*)nullptr,
                     nullptr, nullptr, nullptr, nullptr, ovf);
  GT_TRY(hipGetLastError());
  std::vector<int> hcnt(4 * (size_t)nn);
  GT_TRY(hipMemcpyAsync(hcnt.data(), cnt, sizeof(int) * 4 * (size_t)nn, hipMemcpyDeviceToHost, st));
  GT_TRY(hipStreamSynchronize(st));
  std::vector<long long> hoff(4 * (size_t)nn);
  long long tot[4] = {0, 0, 0, 0};
  for (int l = 0; l < 4; l++)
    for (int t = 0; t < nn; t++) {
      hoff[(size_t)l * nn + t] = tot[l];
      tot[l] += hcnt[(size_t)l * nn + t];
    }
  long long* off = D.alloc<long long>(4 * (size_t)nn);
  GT_NEED(off);
  GT_TRY(hipMemcpyAsync(off, hoff.data(), sizeof(long long) * 4 * (size_t)nn, hipMemcpyHostToDevice, st));
  unsigned long long* keys[4];
  for (int l = 0; l < 4; l++) {
    keys[l] = D.alloc<unsigned long long>((size_t)tot[l]);
    GT_NEED(keys[l]);
  }
  hipLaunchKernelGGL(k_lists<true>, dim3(blocks(nn, 64)), dim3(64), 0, st, nn, TD, cnt, off, keys[0], keys[1],
                     keys[2], keys[3], ovf);
  GT_TRY(hipGetLastError());
  int hovf = 0;
  GT_TRY(hipMemcpyAsync(&hovf, ovf, sizeof(int), hipMemcpyDeviceToHost, st));
  GT_TRY(hipStreamSynchronize(st));
  if (hovf) {
    set_error("GPU tree build: list traversal stack exceeded");
    return false;
  }
  for (int l = 0; l < 4; l++) {
    const long long m = tot[l];
    T.lptr[l].assign((size_t)nn + 1, 0);
    if (m == 0) continue;
    unsigned long long* ks = D.alloc<unsigned long long>((size_t)m);
    GT_NEED(ks);
    size_t tbytes = 0;
    GT_TRY(rocprim::radix_sort_keys(nullptr, tbytes, keys[l], ks, (int)m, 0u, (unsigned)(32 + nbits), st));
    void* tmp = D.alloc<char>(tbytes);
    GT_NEED(tmp);
    GT_TRY(rocprim::radix_sort_keys(tmp, tbytes, keys[l], ks, (int)m, 0u, (unsigned)(32 + nbits), st));
    int* keep = D.alloc<int>((size_t)m);
    int* pos = D.alloc<int>((size_t)m);
    int* per = D.alloc<int>((size_t)nn);
    GT_NEED(keep);
    GT_NEED(pos);
    GT_NEED(per);
    GT_TRY(hipMemsetAsync(per, 0, sizeof(int) * nn, st));
    hipLaunchKernelGGL(k_keep, dim3(blocks(m, 256)), dim3(256), 0, st, m, ks, tag, keep, per);
    GT_TRY(hipGetLastError());
    if (!scan(D, keep, pos, (int)m, st)) return false;
    int lastk[2];
    GT_TRY(hipMemcpyAsync(&lastk[0], keep + m - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    GT_TRY(hipMemcpyAsync(&lastk[1], pos + m - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    GT_TRY(hipStreamSynchronize(st));
    const int kept = lastk[0] + lastk[1];
    int* idx = D.alloc<int>((size_t)kept);
    GT_NEED(idx);
    hipLaunchKernelGGL(k_compact, dim3(blocks(m, 256)), dim3(256), 0, st, m, ks, keep, pos, idx);
    GT_TRY(hipGetLastError());
    std::vector<int> hper;
    if (!down(hper, per, (size_t)nn, st) || !down(T.lidx[l], idx, (size_t)kept, st)) return false;
    GT_TRY(hipStreamSynchronize(st));
    for (int t = 0; t < nn; t++) T.lptr[l][t + 1] = T.lptr[l][t] + hper[t];
    for (void* pp : {(void*)ks, tmp, (void*)keep, (void*)pos, (void*)per, (void*)idx}) D.release(pp);
  }
  /* tree arrays back to the host */
  std::vector<int> cnt_h, tcnt_h, sb_h, tb_h;
  if (!down(T.parent, parent, nn, st) || !down(T.child, child, nn, st) || !down(T.depth, depth, nn, st) ||
      !down(T.path, path, 3 * (size_t)nn, st) || !down(T.tag, tag, nn, st) || !down(cnt_h, count, nn, st) ||
      !down(tcnt_h, tcount, nn, st) || !down(sb_h, sb, nn, st) || !down(tb_h, tb, nn, st) ||
      !down(T.sperm, sperm, n, st) || !down(T.tperm, tperm, n, st))
    return false;
  GT_TRY(hipStreamSynchronize(st));
  T.sbeg.assign(nn, 0);
  T.snum.assign(nn, 0);
  T.tbeg.assign(nn, 0);
  T.tnum.assign(nn, 0);
  for (int k = 0; k < nn; k++) {
    if (T.child[k] != -1) continue;
    if (cnt_h[k] > 0) {
      T.sbeg[k] = sb_h[k];
      T.snum[k] = cnt_h[k];
    }
    if (tcnt_h[k] > 0) {
      T.tbeg[k] = tb_h[k];
      T.tnum[k] = tcnt_h[k];
    }
  }
  kifmm::problem_install(p, std::move(T), kifmm::wall_seconds() - t0);
  *seconds = kifmm::wall_seconds() - t0;
  return true;
}

kifmm_problem* finish_gpu(int device, kifmm_problem* p) {
  if (!p) return nullptr;
  double sec = 0.0;
  bool ok = false;
  try {
    ok = build(device, p, &sec);
  } catch (const std::bad_alloc&) {
    set_error("GPU tree build: host allocation failed");
  }
  if (!ok) {
    kifmm_problem_destroy(p);
    return nullptr;
  }
  return p;
}

}  // namespace

extern "C" {

kifmm_problem* kifmm_problem_create_gpu(int device, int n, const char* distribution, int pts_max,
                                        const kifmm_problem_opts* opts) {
  return finish_gpu(device, kifmm::problem_generate(n, distribution, pts_max, opts));
}

kifmm_problem* kifmm_problem_from_points_gpu(int device, int n, const double* sx, const double* sy, const double* sz,
                                             const double* den, const double* tx, const double* ty, const double* tz,
                                             int pts_max, int np) {
  return finish_gpu(device, kifmm::problem_points(n, sx, sy, sz, den, tx, ty, tz, pts_max, np));
}

}  // extern "C"
