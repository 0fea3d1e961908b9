/*
 * kifmm_gpu.hip -- create / evaluate / destroy of the MI355X KIFMM evaluator.
 *
 * kifmm_gpu_create is the counterpart of preproc() (fmm/evaluate--cuda.cu:1441-2239):
 * it derives the device work lists from the flattened tree and uploads
 * everything once.  kifmm_gpu_evaluate is the counterpart of run()
 * (fmm/run--cuda.cu:27-345): up (S2M + M2M), U, V (forward DFT,
 * translation spectra, spectral multiply-add, inverse DFT), W, X, down
 * (L2L + DC2DE + L2T) and the copy back, in that order, on one stream.
 *
 * Device layout (HBM):
 *   src4 / trg4   gathered points as (x, y, z, w) double4 records
 *   uden/dchk/dden  node-major vectors, ld_e = n_e rounded up to 8 doubles,
 *                 zero padding
 *   chk           UC check potentials of one batch of columns (ld_uc)
 *   X / Y         source / target spectra, slot-major [slot][f][octant]:
 *                 slot 1 + b holds the 8 children of child block b (nodes
 *                 1+8b .. 8+8b); slot 0 is all zero and stands for a missing
 *                 neighbour
 *   T             translation spectra, frequency-major [f][343]
 *   operators     padded column-major, UE2UC and DE2DC stacked by octant
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "kifmm_hip.h"
#include "../host/kifmm_internal.h"
#include "kernels.hpp"

using namespace kifmm_dev;
using kifmm::set_error;

/* Device scratch of kifmm_gpu_direct_noalloc: the staged coordinates of up
 * to KIFMM_DIRECT_NOALLOC_MAX points and their sums.  Every device the
 * library runs on holds its own copy until the process ends.  In host code
 * the array's name is not a device address: the per-device address comes
 * from hipGetSymbolAddress (scratch_address, called by kifmm_gpu_create with
 * the handle's device current) and only that address is passed on. */
static __device__ __align__(256) double scratch[4 * KIFMM_DIRECT_NOALLOC_MAX];

namespace {

inline int rup(int x, int m) { return (x + m - 1) / m * m; }

struct GemmList {
  int n = 0;
  int* in_off = nullptr;
  int* out_off = nullptr;
  int* sexp = nullptr;
};

struct LevelOps {
  GemmList m2m_a, m2m_b; /* UE2UC stack -> chk, UC2UE -> uden */
  GemmList l2l, dc2de;
};

}  // namespace

struct kifmm_gpu {
  int dev = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[16] = {};
  /* KIFMM_GPU_OVERLAP: the U-list on a second stream (default priority)
   * alongside up and V; fork / join events and the U-list's own timing */
  int overlap = 0;
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr, su0 = nullptr, su1 = nullptr;
  int np = 0, n_e = 0, n_uc = 0, ld_e = 0, ld_uc = 0, lda_e = 0, lda_uc = 0, lda_l2l = 0, ld_surf = 0;
  int N = 0, H = 0, F = 0;
  int n_nodes = 0, n_src = 0, n_trg = 0, coincident = 0, max_depth = 0, precision = 0;
  std::vector<void*> allocs;
  /* device data */
  double4 *src4 = nullptr, *trg4 = nullptr, *geo = nullptr;
  double *pot = nullptr, *uden = nullptr, *dchk = nullptr, *dden = nullptr, *chk = nullptr;
  double *UC2UE = nullptr, *UE2UC = nullptr, *DC2DE = nullptr, *DE2DC = nullptr;
  double *surf_ue = nullptr, *surf_uc = nullptr, *surf_de = nullptr, *surf_dc = nullptr;
  double2 *X = nullptr, *Y = nullptr, *T = nullptr, *tw = nullptr;
  /* mixed path: FP32 spectra, FP32 unit surfaces, FP32 leaf-local sources */
  float2 *Xf = nullptr, *Yf = nullptr, *Tf = nullptr;
  float *surf_uc32 = nullptr, *surf_de32 = nullptr;
  float4* src32 = nullptr;
  int *cube2surf = nullptr, *surf2cube = nullptr;
  int *child = nullptr, *snum = nullptr, *sbeg = nullptr, *tbeg = nullptr, *depth = nullptr, *tag = nullptr;
  int *widx = nullptr, *xidx = nullptr;
  int *u_ptr = nullptr, *u_idx = nullptr; /* U-list CSR (node ids), for the close-pair search */
  /* work lists */
  int4* p2p_items = nullptr; int n_p2p = 0;
  int4* p2p_items64 = nullptr; int n_p2p64 = 0;
  int4* p2p_items5 = nullptr; int n_p2p5 = 0; /* <= P2P_MAXT5 targets: the FP64 Gram-form kernels */
  int4* p2p_itemsP = nullptr; int n_p2pP = 0; /* sibling pairs and single leaves: the mixed tau kernel */
  int* rng_ptr = nullptr; int2* rng = nullptr;
  int* p2p_counters = nullptr; /* 8, one per XCD: persistent U-list */
  int persist_grid[2] = {0, 0}; /* blocks of the persistent U-list form: difference, Gram */
  /* launch shapes (kifmm_gpu_set_tuning / kifmm_gpu_autotune) */
  int mix_lpb = 1; /* leaves per block of the mixed S2M / L2T */
  int p2p_form = KIFMM_P2P_STATIC;
  int had_fb = 8;
  int had_form = KIFMM_HAD_MFMA;
  int up_form = KIFMM_LEVEL_AUTO, down_form = KIFMM_LEVEL_AUTO;
  /* mixed path: per target, the U-list sources closer than MIX_CLOSE h */
  int* cptr = nullptr; int* cidx = nullptr;
  int close_listed = 0;
  double4* trg_gram = nullptr; /* FP64 Gram-form U-list: (-2t', |t'|^2) per target, when close_listed */
  long long n_close = 0;
  int4* s2m_items = nullptr; int n_s2m = 0;
  GemmList s2m_gemm;
  std::vector<LevelOps> lv;
  int* vblocks = nullptr; int* vslot = nullptr; int n_vblocks = 0;
  int* hslot = nullptr; int n_hpar = 0;
  int4* hgroups = nullptr; int n_hgroups = 0;
  int* huslots = nullptr;
  unsigned char* hloc = nullptr;
  int4* witems = nullptr; int n_witems = 0; /* k_wx items, two int4 each */
  int4* xitems = nullptr; int n_xitems = 0;
  int4* wred = nullptr; int n_wred = 0;     /* k_wx_reduce items of split lists */
  int4* xred = nullptr; int n_xred = 0;
  double* wxpart = nullptr;                 /* chunk sums of split lists, 64 per slot */
  int4* l2t_items = nullptr; int n_l2t = 0;
  size_t chk_cols = 0;
  double* ws = nullptr;
  size_t ws_len = 0;
  kifmm_work_counts counts{};
  /* wall ms of kifmm_gpu_create: host plan, upload, close-pair search,
   * other device set-up (translation spectra, Gram records), whole call */
  double create_ms[5] = {0, 0, 0, 0, 0};
  /* pinned host staging of the potentials (device -> host copy at PCIe rate) */
  double* pinned = nullptr;
  /* HIP graph of the whole launch sequence (KIFMM_GPU_GRAPH) */
  int use_graph = 0;
  bool capturing = false;
  hipGraph_t graph[2] = {nullptr, nullptr}; /* [0]: the far-field sequence */
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  /* copy-out: target cuts of the U-list chunks (leaf starts, 0 .. n_trg),
   * the first item of each chunk per item list, the copy stream (default
   * priority) and the events that order chunks, copies and joins */
  std::vector<int> cuts;
  std::vector<int> chunk_beg[4];
  hipStream_t cstream = nullptr;
  std::vector<hipEvent_t> ch_ev, cd_ev;
  hipEvent_t far_ev = nullptr, sdone_ev = nullptr, cdone_ev = nullptr;
  /* the U-list's own sums when it runs on the second stream (overlap layout) */
  double* potU = nullptr;
  /* staging of kifmm_gpu_set_densities */
  double* den_stage = nullptr;
  /* this device's address of the file-scope scratch array */
  double* scratch = nullptr;
};

namespace {

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) {                                                                \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                        \
      return KIFMM_EDEVICE;                                                                \
    }                                                                                      \
  } while (0)

template <typename T>
int dev_upload(kifmm_gpu* g, T** dst, const T* src, size_t n) {
  *dst = nullptr;
  const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    set_error("hipMalloc of " + std::to_string(bytes) + " bytes failed");
    return KIFMM_ENOMEM;
  }
  g->allocs.push_back(p);
  *dst = (T*)p;
  /* on the handle's (non-blocking) stream, so that the kernels that fill
   * the buffer at create time run after the clear: a null-stream memset is
   * not ordered with them */
  if (src && n) HIP_TRY(hipMemcpyAsync(p, src, n * sizeof(T), hipMemcpyHostToDevice, g->stream));
  else HIP_TRY(hipMemsetAsync(p, 0, bytes, g->stream));
  if (src && n) HIP_TRY(hipStreamSynchronize(g->stream)); /* src may be a temporary */
  return KIFMM_OK;
}

template <typename T>
int dev_zero(kifmm_gpu* g, T** dst, size_t n) {
  return dev_upload<T>(g, dst, nullptr, n);
}

#define TRY(expr)            \
  do {                       \
    int r_ = (expr);         \
    if (r_ != KIFMM_OK) return r_; \
  } while (0)

/* boundary nodes of an np^3 grid on [-R,R]^3, x slowest (fmm/trans.cc:31-62):
 * rows x, y, z, then |p|^2 (row 3, for the Gram-form L2T) */
void host_surface(int np, double R, std::vector<double>& xyz, int ld) {
  const double step = 2.0 / (np - 1);
  int cnt = 0;
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      for (int k = 0; k < np; k++) {
        if (!(i == 0 || i == np - 1 || j == 0 || j == np - 1 || k == 0 || k == np - 1)) continue;
        const double x = R * (-1.0 + i * step), y = R * (-1.0 + j * step), z = R * (-1.0 + k * step);
        xyz[cnt] = x;
        xyz[ld + cnt] = y;
        xyz[2 * ld + cnt] = z;
        xyz[3 * ld + cnt] = x * x + y * y + z * z;
        cnt++;
      }
}

int find_box(const kifmm_tree_view* t, int d, const int want[3]) {
  int cur = 0;
  while (t->depth[cur] < d && t->child[cur] != -1) {
    const int sh = d - t->depth[cur] - 1;
    const int o = (((want[0] >> sh) & 1) << 2) | (((want[1] >> sh) & 1) << 1) | ((want[2] >> sh) & 1);
    cur = t->child[cur] + o;
  }
  return cur;
}

/* The per-device address of `scratch`, looked up once per device (the
 * device must be current) and kept. */
int scratch_address(int dev, double** out) {
  static std::mutex mu;
  static std::vector<double*> addr;
  std::lock_guard<std::mutex> lock(mu);
  if (dev < 0) {
    set_error("scratch_address: invalid device");
    return KIFMM_EINVAL;
  }
  if ((size_t)dev >= addr.size()) addr.resize(dev + 1, nullptr);
  if (!addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return KIFMM_EDEVICE;
    }
    addr[dev] = (double*)p;
  }
  *out = addr[dev];
  return KIFMM_OK;
}

/* Structural checks of a caller-supplied tree view: every index a kernel
 * derives from it must land inside the arrays it addresses. */
int check_view(const kifmm_tree_view* t) {
  const int nn = t->n_nodes;
  auto bad = [](const std::string& m) {
    set_error("kifmm_gpu_create: malformed tree view: " + m);
    return KIFMM_EINVAL;
  };
  if (t->n_levels < 2 || t->level_start[0] != 0 || t->level_start[t->n_levels - 1] != nn)
    return bad("level_start");
  for (int i = 0; i < nn; i++) {
    const int c = t->child[i];
    if (c != -1 && (c < 1 || c + 8 > nn || (c - 1) % 8 != 0)) return bad("child index");
    if (t->depth[i] < 0 || t->depth[i] > 30) return bad("depth");
    if (t->depth[i] + 2 > t->n_levels) return bad("depth vs levels");
    const int lim = 1 << t->depth[i];
    for (int k = 0; k < 3; k++)
      if (t->path[3 * i + k] < 0 || t->path[3 * i + k] >= lim) return bad("path");
    if (t->src_num[i] < 0 || t->src_beg[i] < 0 || (long)t->src_beg[i] + t->src_num[i] > t->n_src) return bad("source range");
    if (t->trg_num[i] < 0 || t->trg_beg[i] < 0 || (long)t->trg_beg[i] + t->trg_num[i] > t->n_trg) return bad("target range");
    if ((t->src_num[i] > 0 || t->trg_num[i] > 0) && c != -1) return bad("points on an inner node");
  }
  for (int d = 0; d + 1 < t->n_levels; d++)
    if (t->level_start[d] > t->level_start[d + 1]) return bad("level order");
  const int* ptrs[4] = {t->u_ptr, t->v_ptr, t->w_ptr, t->x_ptr};
  const int* idxs[4] = {t->u_idx, t->v_idx, t->w_idx, t->x_idx};
  for (int L = 0; L < 4; L++) {
    if (ptrs[L][0] != 0) return bad("list pointer");
    for (int i = 0; i < nn; i++) {
      if (ptrs[L][i + 1] < ptrs[L][i]) return bad("list pointer order");
      for (int k = ptrs[L][i]; k < ptrs[L][i + 1]; k++)
        if (idxs[L][k] < 0 || idxs[L][k] >= nn) return bad("list entry");
    }
  }
  return KIFMM_OK;
}

/* Everything the device needs, derived on the host from the tree view.
 * make_plan() touches no device; upload() copies a checked plan. */
struct HostGemm {
  std::vector<int> in, out, sexp;
};
struct HostLevel {
  HostGemm m2m_a, m2m_b, l2l, dc2de;
};
struct Plan {
  int np = 0, n_e = 0, n_uc = 0, ld_e = 0, ld_uc = 0, lda_e = 0, lda_uc = 0, lda_l2l = 0, ld_surf = 0;
  int N = 0, H = 0, F = 0, nblk = 0, max_depth = 0;
  std::vector<int4> p2p_items, p2p_items64, p2p_items5, p2p_itemsP;
  /* copy-out chunks: target cut points (leaf starts, 0 .. n_trg) and, per
   * item list (<= P2P_MAXT, <= P2P_MAXT / 2, <= P2P_MAXT5, pairs), the
   * first item of each chunk */
  std::vector<int> cuts;
  std::vector<int> chunk_beg[4];
  std::vector<int> rng_ptr;
  std::vector<int2> rng;
  std::vector<int4> s2m_items;
  HostGemm s2m_gemm;
  std::vector<HostLevel> lv;
  size_t chk_cols = 0;
  size_t ws_len = 0;
  std::vector<int> vblocks, vslot, hnbr, hslot, wleaves, xnodes, l2t;
  std::vector<int4> witems, xitems, wred, xred; /* two int4 per item */
  int wx_slots = 0, w_slots = 0, x_slots = 0;     /* partial slots of split W / X lists */
  std::vector<int4> l2t_items;
  std::vector<int4> hgroups;
  std::vector<int> huslots;
  std::vector<unsigned char> hloc;
  kifmm_work_counts wc{};
};

bool offsets_ok(const std::vector<int>& off, long len_read, long buf_len) {
  for (int o : off)
    if (o < 0 || (long)o + len_read > buf_len) return false;
  return true;
}

/* Targets of a U-list work item: it.z holds n_A | n_B << 16, with n_B = 0
 * but for the sibling pairs of the pair list. */
inline int item_targets(const int4& it) { return (it.z & 0xffff) + (it.z >> 16); }

/* The mixed tau kernel's item list.  Two sibling leaves A and B = A + 1
 * (octants 2m and 2m + 1 of one parent, so adjacent) at one depth, whose
 * targets follow each other and fill at most P2P_MAXT5 / 16 groups of 16
 * (each leaf's own groups), are one item: three source sets follow each
 * other in rng_ptr (the boxes in both U-lists, in A's order; those only in
 * A's; those only in B's) and the item is (first set, A's first target,
 * n_A | n_B << 16, A).  Every other leaf keeps its items of the P2P_MAXT5
 * list (its own set, n_B = 0).  A pair stages each shared source once for
 * the targets of both leaves: 36 boxes of sources instead of 2 x 27 for
 * same-size neighbours. */
int pair_items(const kifmm_tree_view* t, Plan& P) {
  const int nn = t->n_nodes;
  auto trg_leaf = [&](int i) { return i >= 0 && i < nn && t->child[i] == -1 && t->trg_num[i] > 0; };
  std::vector<char> mark(nn, 0);
  std::vector<int> sh, oa, ob;
  auto add_set = [&](const std::vector<int>& boxes) {
    for (int s : boxes) {
      const int b = t->src_beg[s], n = t->src_num[s];
      if (n <= 0) continue;
      if ((int)P.rng.size() > P.rng_ptr.back() && P.rng.back().x + P.rng.back().y == b) P.rng.back().y += n;
      else P.rng.push_back(make_int2(b, n));
    }
    P.rng_ptr.push_back((int)P.rng.size());
  };
  const std::vector<int4>& L5 = P.p2p_items5; /* in leaf order */
  size_t k5 = 0;
  P.p2p_itemsP.clear();
  for (int i = 0; i < nn; i++) {
    if (!trg_leaf(i)) continue;
    const int j = i + 1, nA = t->trg_num[i];
    const bool pair = i >= 1 && (i - 1) % 2 == 0 && trg_leaf(j) && (i - 1) / 8 == (j - 1) / 8 &&
                      t->depth[i] == t->depth[j] && t->trg_beg[j] == t->trg_beg[i] + nA &&
                      (nA + 15) / 16 + (t->trg_num[j] + 15) / 16 <= P2P_MAXT5 / 16;
    while (k5 < L5.size() && L5[k5].w < i) k5++;
    if (!pair) {
      while (k5 < L5.size() && L5[k5].w == i) P.p2p_itemsP.push_back(L5[k5++]);
      continue;
    }
    sh.clear();
    oa.clear();
    ob.clear();
    for (int k = t->u_ptr[j]; k < t->u_ptr[j + 1]; k++) mark[t->u_idx[k]] = 1;
    for (int k = t->u_ptr[i]; k < t->u_ptr[i + 1]; k++) {
      const int s = t->u_idx[k];
      if (mark[s]) {
        sh.push_back(s);
        mark[s] = 2;
      } else {
        oa.push_back(s);
      }
    }
    for (int k = t->u_ptr[j]; k < t->u_ptr[j + 1]; k++) {
      const int s = t->u_idx[k];
      if (mark[s] == 1) ob.push_back(s);
      mark[s] = 0;
    }
    const int set = (int)P.rng_ptr.size() - 1;
    add_set(sh);
    add_set(oa);
    add_set(ob);
    P.p2p_itemsP.push_back(make_int4(set, t->trg_beg[i], nA | (t->trg_num[j] << 16), i));
    P.wc.pair_items += 1.0;
    i = j; /* B is done */
  }
  for (const int4& it : P.p2p_itemsP) {
    const int nA = it.z & 0xffff, nB = it.z >> 16, sets = nB ? 3 : 1;
    if (it.x < 0 || it.x + sets >= (int)P.rng_ptr.size() || nA < 1 ||
        (nA + 15) / 16 + (nB + 15) / 16 > P2P_MAXT5 / 16 || it.y < 0 || (long)it.y + nA + nB > t->n_trg ||
        it.w < 0 || it.w + (nB ? 1 : 0) >= nn) {
      set_error("internal: U-list pair item out of range");
      return KIFMM_EINVAL;
    }
  }
  return KIFMM_OK;
}

int make_plan(const kifmm_tree_view* t, const kifmm_operators* ops, Plan& P) {
  TRY(check_view(t));
  const int np = t->np;
  P.np = np;
  P.n_e = ops->n_e;
  P.n_uc = ops->n_uc;
  P.ld_e = rup(P.n_e, 8);
  P.ld_uc = rup(P.n_uc, 8);
  P.lda_e = rup(P.n_e, 16);
  P.lda_uc = rup(P.n_uc, 16);
  P.lda_l2l = rup(8 * P.ld_e, 16);
  P.ld_surf = rup(std::max(P.n_uc, P.n_e), 8);
  P.N = 2 * np;
  P.H = np + 1;
  P.F = P.N * P.N * P.H;
  const int nn = t->n_nodes, ne = P.n_e, nuc = P.n_uc, lde = P.ld_e, lduc = P.ld_uc;
  P.nblk = (nn - 1) / 8;
  int maxd = 0;
  for (int i = 0; i < nn; i++) maxd = std::max(maxd, t->depth[i]);
  P.max_depth = maxd;
  kifmm_work_counts& wc = P.wc;
  std::memset(&wc, 0, sizeof wc);
  auto fail = [](const char* m) {
    set_error(std::string("internal: ") + m);
    return KIFMM_EINVAL;
  };

  /* U-list: merged source ranges per target leaf; a leaf's targets split
   * into the fewest items of <= P2P_MAXT, balanced, in multiples of 16
   * (also of <= P2P_MAXT / 2 and <= P2P_MAXT5 for the other launch shapes) */
  P.rng_ptr.assign(1, 0);
  for (int i = 0; i < nn; i++) {
    if (t->child[i] != -1 || t->trg_num[i] <= 0) continue;
    const int set = (int)P.rng_ptr.size() - 1;
    for (int k = t->u_ptr[i]; k < t->u_ptr[i + 1]; k++) {
      const int s = t->u_idx[k];
      const int b = t->src_beg[s], n = t->src_num[s];
      if (n <= 0) continue;
      wc.u_pairs += (double)n * t->trg_num[i];
      if ((int)P.rng.size() > P.rng_ptr.back() && P.rng.back().x + P.rng.back().y == b) P.rng.back().y += n;
      else P.rng.push_back(make_int2(b, n));
    }
    P.rng_ptr.push_back((int)P.rng.size());
    const int n = t->trg_num[i];
    for (int cap : {P2P_MAXT, P2P_MAXT / 2, P2P_MAXT5}) { /* the item lists of the launch shapes */
      const int nit = (n + cap - 1) / cap;
      const int per = (((n + nit - 1) / nit) + 15) & ~15;
      std::vector<int4>& out = cap == P2P_MAXT ? P.p2p_items : cap == P2P_MAXT5 ? P.p2p_items5 : P.p2p_items64;
      for (int c = 0; c < n; c += per) out.push_back(make_int4(set, t->trg_beg[i] + c, std::min(per, n - c), i));
    }
  }
  /* k_p2p_ulist reads trg4[it.y .. it.y+it.z-1] and src4 inside each range */
  for (int cap : {P2P_MAXT, P2P_MAXT / 2, P2P_MAXT5})
    for (const int4& it : cap == P2P_MAXT ? P.p2p_items : cap == P2P_MAXT5 ? P.p2p_items5 : P.p2p_items64)
      if (it.x < 0 || it.x + 1 >= (int)P.rng_ptr.size() || it.z < 1 || it.z > cap || it.y < 0 ||
          (long)it.y + it.z > t->n_trg || it.w < 0 || it.w >= nn)
        return fail("U-list work item out of range");
  TRY(pair_items(t, P));
  for (const int2& r : P.rng)
    if (r.x < 0 || r.y < 1 || (long)r.x + r.y > t->n_src) return fail("U-list source range out of range");
  /* Copy-out chunks.  The U-list runs last, chunk by chunk, and each
   * chunk's potentials go to the host while the next chunks compute.  A
   * chunk is a run of whole leaves whose targets are one contiguous range:
   * the item lists cover [0, n_trg) in item order with no gap (checked), and
   * cuts fall at leaf starts, where every list has an item boundary.  The
   * chunks shrink towards the end so that the copy left after the last
   * chunk is short.  Small inputs are one chunk. */
  {
    const std::vector<int4>* lists[4] = {&P.p2p_items, &P.p2p_items64, &P.p2p_items5, &P.p2p_itemsP};
    bool tiled = t->n_trg > 0;
    for (const std::vector<int4>* L : lists) {
      long next = 0;
      for (const int4& it : *L) {
        if (it.y != next) tiled = false;
        next = (long)it.y + item_targets(it);
      }
      if (next != t->n_trg) tiled = false;
    }
    P.cuts.assign(1, 0);
    if (tiled && t->n_trg >= (1 << 18)) {
      static const double frac[] = {0.30, 0.55, 0.75, 0.88, 0.95};
      for (double f : frac) {
        const long want = (long)(f * t->n_trg);
        /* the first leaf start at or after want that starts an item of
         * the pair list (then every list has an item boundary there) */
        int cut = -1;
        for (const int4& it : P.p2p_itemsP)
          if (it.y >= want && it.y == t->trg_beg[it.w]) {
            cut = it.y;
            break;
          }
        if (cut > P.cuts.back() && cut < t->n_trg) P.cuts.push_back(cut);
      }
    }
    P.cuts.push_back(t->n_trg);
    for (int l = 0; l < 4; l++) {
      const std::vector<int4>& L = *lists[l];
      P.chunk_beg[l].clear();
      size_t k = 0;
      for (size_t c = 0; c + 1 < P.cuts.size(); c++) {
        while (k < L.size() && L[k].y < P.cuts[c]) k++;
        if (k < L.size() && L[k].y != P.cuts[c] && c > 0) return fail("copy-out chunk does not start an item");
        P.chunk_beg[l].push_back((int)k);
      }
      P.chunk_beg[l].push_back((int)L.size());
    }
  }

  /* S2M + UC2UE over source leaves */
  for (int i = 0; i < nn; i++) {
    if (t->child[i] != -1 || !(t->tag[i] & KIFMM_SRCNODE) || t->src_num[i] <= 0) continue;
    const int col = (int)P.s2m_items.size();
    P.s2m_items.push_back(make_int4(i, t->src_beg[i], t->src_num[i], col));
    P.s2m_gemm.in.push_back(col * lduc);
    P.s2m_gemm.out.push_back(i * lde);
    P.s2m_gemm.sexp.push_back(-t->depth[i]);
    wc.s2m_pairs += (double)nuc * t->src_num[i];
    wc.gemm_flops += 2.0 * ne * nuc;
  }
  P.chk_cols = P.s2m_items.size();
  for (const int4& it : P.s2m_items)
    if (it.y < 0 || it.z < 1 || (long)it.y + it.z > t->n_src) return fail("S2M source range out of range");

  /* per-level M2M, L2L and DC2DE column lists */
  P.lv.assign(maxd + 1, HostLevel());
  for (int d = 0; d <= maxd; d++) {
    HostLevel& L = P.lv[d];
    int col = 0;
    for (int i = t->level_start[d]; i < t->level_start[d + 1]; i++) {
      if (t->depth[i] != d) return fail("level_start does not match depth");
      /* upward densities of depth 0 and 1 are never read: V-lists start at
       * depth 2 and W-list boxes lie at depth >= 2 */
      if (d >= 2 && t->child[i] != -1 && (t->tag[i] & KIFMM_SRCNODE)) {
        L.m2m_a.in.push_back(t->child[i] * lde); L.m2m_a.out.push_back(col * lduc); L.m2m_a.sexp.push_back(d + 1);
        L.m2m_b.in.push_back(col * lduc); L.m2m_b.out.push_back(i * lde); L.m2m_b.sexp.push_back(-d);
        col++;
        wc.gemm_flops += 2.0 * nuc * 8 * ne + 2.0 * ne * nuc;
      }
      if (d + 1 >= 3 && t->child[i] != -1 && (t->tag[i] & KIFMM_TRGNODE)) {
        L.l2l.in.push_back(i * lde); L.l2l.out.push_back(t->child[i] * lde); L.l2l.sexp.push_back(d);
        wc.gemm_flops += 2.0 * 8 * ne * ne;
      }
      if (d >= 2 && (t->tag[i] & KIFMM_TRGNODE)) {
        L.dc2de.in.push_back(i * lde); L.dc2de.out.push_back(i * lde); L.dc2de.sexp.push_back(-d);
        wc.gemm_flops += 2.0 * ne * ne;
      }
    }
    P.chk_cols = std::max(P.chk_cols, (size_t)col);
  }
  /* the fused downward kernel (k_l2l_dc2de) of depth d >= 3 runs over the
   * L2L parents at d - 1 and writes the densities of their target
   * children: exactly the DC2DE list of depth d */
  for (int d = 3; d <= maxd; d++) {
    size_t kids = 0;
    for (int off : P.lv[d - 1].l2l.out) {
      const int c0 = off / lde;
      for (int o = 0; o < 8; o++) kids += (t->tag[c0 + o] & KIFMM_TRGNODE) != 0;
    }
    if (kids != P.lv[d].dc2de.in.size()) return fail("L2L parents do not cover the DC2DE nodes of a level");
  }
  /* split-K workspace: the largest nsplit * ncols * ldw of any launch */
  {
    auto need = [&](size_t ncols, int M, int K) {
      if (!ncols) return;
      const GemmShape sh = gemm_shape((int)ncols, M, K);
      if (sh.nsplit > 1) P.ws_len = std::max(P.ws_len, (size_t)sh.nsplit * ncols * sh.ldw);
    };
    need(P.s2m_gemm.in.size(), ne, lduc);
    for (const HostLevel& L : P.lv) {
      need(L.m2m_a.in.size(), nuc, 8 * lde);
      need(L.m2m_b.in.size(), ne, lduc);
      need(L.l2l.in.size(), 8 * lde, lde);
      need(L.dc2de.in.size(), ne, lde);
    }
  }
  /* GEMM columns: inputs read K doubles, outputs write M doubles */
  {
    const long nodes = (long)nn * lde, chk = (long)std::max<size_t>(P.chk_cols, 1) * lduc;
    bool ok = offsets_ok(P.s2m_gemm.in, lduc, chk) && offsets_ok(P.s2m_gemm.out, ne, nodes);
    for (const HostLevel& L : P.lv)
      ok = ok && offsets_ok(L.m2m_a.in, 8L * lde, nodes) && offsets_ok(L.m2m_a.out, nuc, chk) &&
           offsets_ok(L.m2m_b.in, lduc, chk) && offsets_ok(L.m2m_b.out, ne, nodes) &&
           offsets_ok(L.l2l.in, lde, nodes) && offsets_ok(L.l2l.out, 8L * lde, nodes) &&
           offsets_ok(L.dc2de.in, lde, nodes) && offsets_ok(L.dc2de.out, ne, nodes);
    if (!ok) return fail("GEMM column offset out of range");
  }

  /* V-list: child blocks below parents of depth >= 1 and their 27 neighbour slots */
  for (int i = 0; i < nn; i++) {
    if (t->child[i] == -1 || t->depth[i] < 1) continue;
    const int blk = (t->child[i] - 1) / 8;
    P.vblocks.push_back(t->child[i]);
    P.vslot.push_back(blk + 1);
    P.hslot.push_back(blk + 1);
    const int d = t->depth[i], lim = 1 << d;
    for (int o = 0; o < 27; o++) {
      const int off[3] = {o / 9 - 1, (o / 3) % 3 - 1, o % 3 - 1};
      int w[3], slot = 0;
      bool in = true;
      for (int k = 0; k < 3; k++) {
        w[k] = t->path[3 * i + k] + off[k];
        in = in && w[k] >= 0 && w[k] < lim;
      }
      if (in) {
        const int c = find_box(t, d, w);
        if (t->depth[c] == d && t->child[c] != -1) slot = (t->child[c] - 1) / 8 + 1;
      }
      P.hnbr.push_back(slot);
    }
  }
  for (size_t k = 0; k < P.vblocks.size(); k++)
    if (P.vblocks[k] < 1 || P.vblocks[k] + 8 > nn || P.vslot[k] < 1 || P.vslot[k] > P.nblk || P.hslot[k] < 1 ||
        P.hslot[k] > P.nblk)
      return fail("V-list block out of range");
  for (int v : P.hnbr)
    if (v < 0 || v > P.nblk) return fail("V-list neighbour slot out of range");
  /* Hadamard groups: runs of consecutive parents whose neighbour-slot union
   * fits the LDS staging area; 64 parents (one lane each) when possible */
  {
    const int np_ = (int)P.hslot.size();
    P.hloc.assign((size_t)np_ * 27, 0);
    std::vector<int> uni;
    int first = 0;
    while (first < np_) {
      int cnt = std::min(64, np_ - first);
      for (;;) {
        uni.assign(P.hnbr.begin() + (size_t)first * 27, P.hnbr.begin() + (size_t)(first + cnt) * 27);
        std::sort(uni.begin(), uni.end());
        uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
        if ((int)uni.size() <= HAD_CAP || cnt == 1) break;
        cnt = std::max(1, cnt / 2);
      }
      if ((int)uni.size() > HAD_CAP) return fail("Hadamard neighbour union exceeds the LDS capacity");
      P.hgroups.push_back(make_int4(first, cnt, (int)P.huslots.size(), (int)uni.size()));
      for (int k = 0; k < cnt * 27; k++) {
        const int slot = P.hnbr[(size_t)first * 27 + k];
        P.hloc[(size_t)first * 27 + k] = (unsigned char)(std::lower_bound(uni.begin(), uni.end(), slot) - uni.begin());
      }
      P.huslots.insert(P.huslots.end(), uni.begin(), uni.end());
      first += cnt;
    }
    for (const int4& gq : P.hgroups)
      if (gq.y < 1 || gq.y > 64 || gq.x + gq.y > np_ || gq.w > HAD_CAP || gq.z + gq.w > (int)P.huslots.size())
        return fail("Hadamard group out of range");
  }
  wc.v_entries = t->v_ptr[nn];
  wc.v_cmac_ref = 0.5 * (2.0 * P.F) * wc.v_entries;
  wc.v_cmac = 26.0 * 64.0 * P.F * (double)P.hslot.size();
  wc.fft_bytes = (double)P.vblocks.size() * 8 * (2.0 * 8 * ne + 2.0 * 16 * P.F);
  wc.hadamard_bytes = (double)P.hslot.size() * 8 * 16.0 * P.F * 2 + 343.0 * 16 * P.F;

  /* W, X and L2T work */
  for (int i = 0; i < nn; i++) {
    const bool leaf = t->child[i] == -1;
    const bool trg = (t->tag[i] & KIFMM_TRGNODE) != 0;
    if (trg && leaf && t->w_ptr[i + 1] > t->w_ptr[i]) {
      P.wleaves.push_back(i);
      for (int k = t->w_ptr[i]; k < t->w_ptr[i + 1]; k++) {
        const int s = t->w_idx[k];
        const bool direct = t->child[s] == -1 && t->src_num[s] < ne;
        wc.w_pairs += (double)t->trg_num[i] * (direct ? t->src_num[s] : ne);
      }
    }
    if (trg && t->x_ptr[i + 1] > t->x_ptr[i]) {
      P.xnodes.push_back(i);
      const bool direct = leaf && t->trg_num[i] < ne;
      for (int k = t->x_ptr[i]; k < t->x_ptr[i + 1]; k++)
        wc.x_pairs += (double)t->src_num[t->x_idx[k]] * (direct ? t->trg_num[i] : ne);
    }
    if (trg && leaf && t->trg_num[i] > 0 && t->depth[i] >= 2) {
      P.l2t.push_back(i);
      wc.l2t_pairs += (double)ne * t->trg_num[i];
      for (int c = 0; c < t->trg_num[i]; c += 64)
        P.l2t_items.push_back(make_int4(i, t->trg_beg[i] + c, std::min(64, t->trg_num[i] - c), 0));
    }
  }
  /* W / X items (k_wx): per target node, its list cut at entry boundaries
   * into chunks of at most WX_CHUNK_SRC sources (an entry larger than that is
   * a chunk of its own), and its points into blocks of 64; one item per
   * (block, chunk).  A node with several chunks gets a partial slot per item
   * and a reduce item per block.  Items run longest chunk first. */
  auto wx_plan = [&](const std::vector<int>& nodes, const int* ptr, const int* idx, bool xl, std::vector<int4>& items,
                     std::vector<int4>& red) {
    std::vector<std::pair<long, std::array<int4, 2>>> work;
    for (int i : nodes) {
      const int kind = xl && !(t->child[i] == -1 && t->trg_num[i] < ne) ? 1 : 0;
      const int npts = kind ? ne : t->trg_num[i];
      std::vector<int> cut{ptr[i]};
      std::vector<long> cost;
      long acc = 0;
      for (int k = ptr[i]; k < ptr[i + 1]; k++) {
        const int s_ = idx[k];
        const long n = xl || (t->child[s_] == -1 && t->src_num[s_] < ne) ? t->src_num[s_] : ne;
        if (acc > 0 && acc + n > WX_CHUNK_SRC) {
          cut.push_back(k);
          cost.push_back(acc);
          acc = 0;
        }
        acc += n;
      }
      cut.push_back(ptr[i + 1]);
      cost.push_back(acc);
      const int nch = (int)cost.size();
      for (int c = 0; c < npts; c += 64) {
        const int4 a = make_int4(i, c, std::min(64, npts - c), kind);
        const int slot0 = nch > 1 ? P.wx_slots : -1;
        for (int h = 0; h < nch; h++)
          work.push_back({cost[h], {a, make_int4(cut[h], cut[h + 1], nch > 1 ? P.wx_slots + h : -1, 0)}});
        if (nch > 1) {
          red.push_back(a);
          red.push_back(make_int4(slot0, nch, 0, 0));
          P.wx_slots += nch;
        }
      }
    }
    std::stable_sort(work.begin(), work.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
    for (const auto& w : work) {
      items.push_back(w.second[0]);
      items.push_back(w.second[1]);
    }
  };
  P.wx_slots = 0;
  wx_plan(P.wleaves, t->w_ptr, t->w_idx, false, P.witems, P.wred);
  P.w_slots = P.wx_slots;
  P.wx_slots = 0;
  wx_plan(P.xnodes, t->x_ptr, t->x_idx, true, P.xitems, P.xred);
  P.x_slots = P.wx_slots;
  /* k_wx clamps point indices to [first, first + count - 1] of its item and
   * reads trg4 from tbeg[t] (kind 0) or the n_e-point DC surface (kind 1) */
  auto wx_check = [&](const std::vector<int4>& v, const int* ptr, int slots, bool reduce) {
    for (size_t m = 0; m < v.size(); m += 2) {
      const int4 a = v[m], b = v[m + 1];
      if (a.x < 0 || a.x >= nn || a.z < 1 || a.z > 64 || a.y < 0 || a.y + a.z > (a.w ? ne : t->trg_num[a.x]))
        return false;
      if (!reduce && (b.x < ptr[a.x] || b.y > ptr[a.x + 1] || b.x >= b.y || b.z >= slots)) return false;
      if (reduce && (b.x < 0 || b.y < 2 || b.x + b.y > slots)) return false;
    }
    return true;
  };
  if (!wx_check(P.witems, t->w_ptr, P.w_slots, false) || !wx_check(P.wred, t->w_ptr, P.w_slots, true))
    return fail("W-list item out of range");
  if (!wx_check(P.xitems, t->x_ptr, P.x_slots, false) || !wx_check(P.xred, t->x_ptr, P.x_slots, true))
    return fail("X-list item out of range");
  for (int i : P.l2t)
    if (t->trg_num[i] < 1) return fail("L2T leaf without targets");
  /* k_l2t reads trg4[it.y .. it.y + it.z - 1] and the leaf's DE densities */
  for (const int4& it : P.l2t_items)
    if (it.x < 0 || it.x >= nn || it.z < 1 || it.z > 64 || it.y < 0 || (long)it.y + it.z > t->n_trg)
      return fail("L2T work item out of range");
  if (nuc > 320) return fail("NP too large for the S2M block");
  return KIFMM_OK;
}

int upload_gemm(kifmm_gpu* g, GemmList& L, const HostGemm& h) {
  L.n = (int)h.in.size();
  TRY(dev_upload(g, &L.in_off, h.in.data(), h.in.size()));
  TRY(dev_upload(g, &L.out_off, h.out.data(), h.out.size()));
  TRY(dev_upload(g, &L.sexp, h.sexp.data(), h.sexp.size()));
  return KIFMM_OK;
}

int build(kifmm_gpu* g, const kifmm_tree_view* t, const kifmm_operators* ops, const Plan& P) {
  const int np = P.np, nn = t->n_nodes, ne = P.n_e, nuc = P.n_uc, lde = P.ld_e, lduc = P.ld_uc;
  g->np = np; g->n_e = ne; g->n_uc = nuc; g->ld_e = lde; g->ld_uc = lduc;
  g->lda_e = P.lda_e; g->lda_uc = P.lda_uc; g->lda_l2l = P.lda_l2l; g->ld_surf = P.ld_surf;
  g->N = P.N; g->H = P.H; g->F = P.F;
  g->n_nodes = nn; g->n_src = t->n_src; g->n_trg = t->n_trg;
  g->coincident = t->coincident;
  g->max_depth = P.max_depth;
  g->counts = P.wc;

  /* points: the caller's coordinate arrays go up as they are and are
   * packed into (x, y, z, w) records on the device (k_pack4) */
  {
    TRY(dev_zero(g, &g->src4, (size_t)t->n_src));
    TRY(dev_zero(g, &g->trg4, (size_t)t->n_trg));
    TRY(dev_zero(g, &g->pot, (size_t)t->n_trg));
    const size_t m = (size_t)std::max(std::max(t->n_src, t->n_trg), 1);
    double* stage = nullptr;
    if (hipMalloc(&stage, 4 * m * sizeof(double)) != hipSuccess) {
      set_error("hipMalloc of the point staging buffer failed");
      return KIFMM_ENOMEM;
    }
    struct Free {
      double* p;
      ~Free() { (void)hipFree(p); }
    } free_stage{stage};
    const int blocks = (int)std::min<size_t>(4096, (m + 255) / 256);
    const double* srcs[4] = {t->sx, t->sy, t->sz, t->sden};
    const double* trgs[3] = {t->tx, t->ty, t->tz};
    if (t->n_src > 0) {
      for (int k = 0; k < 4; k++)
        HIP_TRY(hipMemcpyAsync(stage + k * m, srcs[k], sizeof(double) * t->n_src, hipMemcpyHostToDevice, g->stream));
      hipLaunchKernelGGL(k_pack4, dim3(blocks), dim3(256), 0, g->stream, g->src4, stage, stage + m, stage + 2 * m,
                         stage + 3 * m, (size_t)t->n_src);
      HIP_TRY(hipGetLastError());
    }
    if (t->n_trg > 0) {
      HIP_TRY(hipStreamSynchronize(g->stream)); /* the stage is reused */
      for (int k = 0; k < 3; k++)
        HIP_TRY(hipMemcpyAsync(stage + k * m, trgs[k], sizeof(double) * t->n_trg, hipMemcpyHostToDevice, g->stream));
      hipLaunchKernelGGL(k_pack4, dim3(blocks), dim3(256), 0, g->stream, g->trg4, stage, stage + m, stage + 2 * m,
                         nullptr, (size_t)t->n_trg);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(g->stream));
  }
  /* node geometry (fmm/input.cc:113-137) and node arrays */
  {
    std::vector<double4> geo(nn);
    for (int i = 0; i < nn; i++) {
      const double sc = (double)(1 << t->depth[i]);
      geo[i] = make_double4(-1.0 + (2.0 * 1.0) * (t->path[3 * i] + 0.5) / sc,
                            -1.0 + (2.0 * 1.0) * (t->path[3 * i + 1] + 0.5) / sc,
                            -1.0 + (2.0 * 1.0) * (t->path[3 * i + 2] + 0.5) / sc, 1.0 / sc);
    }
    TRY(dev_upload(g, &g->geo, geo.data(), geo.size()));
    TRY(dev_upload(g, &g->child, t->child, nn));
    TRY(dev_upload(g, &g->snum, t->src_num, nn));
    TRY(dev_upload(g, &g->sbeg, t->src_beg, nn));
    TRY(dev_upload(g, &g->tbeg, t->trg_beg, nn));
    TRY(dev_upload(g, &g->depth, t->depth, nn));
    TRY(dev_upload(g, &g->tag, t->tag, nn));
    TRY(dev_upload(g, &g->widx, t->w_idx, t->w_ptr[nn]));
    TRY(dev_upload(g, &g->xidx, t->x_idx, t->x_ptr[nn]));
    TRY(dev_upload(g, &g->u_ptr, t->u_ptr, (size_t)nn + 1));
    TRY(dev_upload(g, &g->u_idx, t->u_idx, t->u_ptr[nn]));
    TRY(dev_zero(g, &g->uden, (size_t)nn * lde));
    TRY(dev_zero(g, &g->dchk, (size_t)nn * lde));
    TRY(dev_zero(g, &g->dden, (size_t)nn * lde));
  }
  /* operators, padded column-major */
  {
    std::vector<double> a((size_t)g->lda_e * lduc, 0.0);
    for (int k = 0; k < nuc; k++)
      for (int i = 0; i < ne; i++) a[(size_t)k * g->lda_e + i] = ops->uc2ue[(size_t)k * ne + i];
    TRY(dev_upload(g, &g->UC2UE, a.data(), a.size()));
    std::vector<double> b((size_t)g->lda_uc * 8 * lde, 0.0);
    for (int o = 0; o < 8; o++)
      for (int k = 0; k < ne; k++)
        for (int i = 0; i < nuc; i++)
          b[(size_t)(o * lde + k) * g->lda_uc + i] = ops->ue2uc[(size_t)o * nuc * ne + (size_t)k * nuc + i];
    TRY(dev_upload(g, &g->UE2UC, b.data(), b.size()));
    std::vector<double> c((size_t)g->lda_e * lde, 0.0);
    for (int k = 0; k < ne; k++)
      for (int i = 0; i < ne; i++) c[(size_t)k * g->lda_e + i] = ops->dc2de[(size_t)k * ne + i];
    TRY(dev_upload(g, &g->DC2DE, c.data(), c.size()));
    std::vector<double> d((size_t)g->lda_l2l * lde, 0.0);
    for (int o = 0; o < 8; o++)
      for (int k = 0; k < ne; k++)
        for (int i = 0; i < ne; i++)
          d[(size_t)k * g->lda_l2l + o * lde + i] = ops->de2dc[(size_t)o * ne * ne + (size_t)k * ne + i];
    TRY(dev_upload(g, &g->DE2DC, d.data(), d.size()));
  }
  /* unit surfaces (fmm/trans.cc:13-26) and DFT tables */
  {
    const int ls = g->ld_surf;
    std::vector<double> s(4 * (size_t)ls, 0.0);
    host_surface(np, 1.0, s, ls); TRY(dev_upload(g, &g->surf_ue, s.data(), s.size()));
    std::fill(s.begin(), s.end(), 0.0);
    host_surface(np + 2, 3.0, s, ls); TRY(dev_upload(g, &g->surf_uc, s.data(), s.size()));
    if (g->precision == KIFMM_PREC_MIXED) {
      std::vector<float> f(s.begin(), s.end());
      TRY(dev_upload(g, &g->surf_uc32, f.data(), f.size()));
    }
    std::fill(s.begin(), s.end(), 0.0);
    host_surface(np, 3.0, s, ls); TRY(dev_upload(g, &g->surf_de, s.data(), s.size()));
    if (g->precision == KIFMM_PREC_MIXED) {
      std::vector<float> f(s.begin(), s.end());
      TRY(dev_upload(g, &g->surf_de32, f.data(), f.size()));
    }
    std::fill(s.begin(), s.end(), 0.0);
    host_surface(np, 1.0, s, ls); TRY(dev_upload(g, &g->surf_dc, s.data(), s.size()));
    const int N = g->N;
    std::vector<double2> w(N);
    for (int m = 0; m < N; m++) {
      const long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)N;
      w[m] = make_double2((double)cosl(a), (double)(-sinl(a)));
    }
    TRY(dev_upload(g, &g->tw, w.data(), w.size()));
    std::vector<int> c2s(np * np * np, -1), s2c;
    for (int i = 0; i < np; i++)
      for (int j = 0; j < np; j++)
        for (int k = 0; k < np; k++)
          if (i == 0 || i == np - 1 || j == 0 || j == np - 1 || k == 0 || k == np - 1) {
            c2s[(i * np + j) * np + k] = (int)s2c.size();
            s2c.push_back((i * np + j) * np + k);
          }
    TRY(dev_upload(g, &g->cube2surf, c2s.data(), c2s.size()));
    TRY(dev_upload(g, &g->surf2cube, s2c.data(), s2c.size()));
  }
  /* work lists */
  g->n_p2p = (int)P.p2p_items.size();
  TRY(dev_upload(g, &g->p2p_items, P.p2p_items.data(), P.p2p_items.size()));
  g->n_p2p64 = (int)P.p2p_items64.size();
  TRY(dev_upload(g, &g->p2p_items64, P.p2p_items64.data(), P.p2p_items64.size()));
  g->n_p2p5 = (int)P.p2p_items5.size();
  TRY(dev_upload(g, &g->p2p_items5, P.p2p_items5.data(), P.p2p_items5.size()));
  g->n_p2pP = (int)P.p2p_itemsP.size();
  TRY(dev_upload(g, &g->p2p_itemsP, P.p2p_itemsP.data(), P.p2p_itemsP.size()));
  TRY(dev_zero(g, &g->p2p_counters, 8 * std::max<size_t>(P.cuts.size(), 1)));
  g->cuts = P.cuts;
  for (int l = 0; l < 4; l++) g->chunk_beg[l] = P.chunk_beg[l];
  TRY(dev_zero(g, &g->potU, (size_t)t->n_trg));
  TRY(dev_zero(g, &g->den_stage, (size_t)t->n_src));
  {
    int ncu = 0, per_cu = 0;
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, g->dev));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_p2p_ulist_persist<false, P2P_G, false>, 256, 0));
    g->persist_grid[0] = std::max(8, (ncu * std::max(per_cu, 1) + 7) / 8 * 8);
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_p2p_ulist_persist<false, P2P_G5, true>, 256, 0));
    g->persist_grid[1] = std::max(8, (ncu * std::max(per_cu, 1) + 7) / 8 * 8);
  }
  TRY(dev_upload(g, &g->rng_ptr, P.rng_ptr.data(), P.rng_ptr.size()));
  TRY(dev_upload(g, &g->rng, P.rng.data(), P.rng.size()));
  g->n_s2m = (int)P.s2m_items.size();
  TRY(dev_upload(g, &g->s2m_items, P.s2m_items.data(), P.s2m_items.size()));
  TRY(upload_gemm(g, g->s2m_gemm, P.s2m_gemm));
  g->lv.assign(P.max_depth + 1, LevelOps());
  for (int d = 0; d <= P.max_depth; d++) {
    TRY(upload_gemm(g, g->lv[d].m2m_a, P.lv[d].m2m_a));
    TRY(upload_gemm(g, g->lv[d].m2m_b, P.lv[d].m2m_b));
    TRY(upload_gemm(g, g->lv[d].l2l, P.lv[d].l2l));
    TRY(upload_gemm(g, g->lv[d].dc2de, P.lv[d].dc2de));
  }
  g->chk_cols = P.chk_cols;
  TRY(dev_zero(g, &g->chk, std::max<size_t>(P.chk_cols, 1) * lduc));
  g->ws_len = P.ws_len;
  TRY(dev_zero(g, &g->ws, std::max<size_t>(P.ws_len, 1)));
  g->n_vblocks = (int)P.vblocks.size();
  g->n_hpar = (int)P.hslot.size();
  TRY(dev_upload(g, &g->vblocks, P.vblocks.data(), P.vblocks.size()));
  TRY(dev_upload(g, &g->vslot, P.vslot.data(), P.vslot.size()));
  g->n_hgroups = (int)P.hgroups.size();
  TRY(dev_upload(g, &g->hgroups, P.hgroups.data(), P.hgroups.size()));
  TRY(dev_upload(g, &g->huslots, P.huslots.data(), P.huslots.size()));
  TRY(dev_upload(g, &g->hloc, P.hloc.data(), P.hloc.size()));
  TRY(dev_upload(g, &g->hslot, P.hslot.data(), P.hslot.size()));
  const size_t spec = (size_t)(P.nblk + 1) * g->F * 8;
  if (g->precision == KIFMM_PREC_MIXED) {
    TRY(dev_zero(g, &g->Xf, spec));
    TRY(dev_zero(g, &g->Yf, spec));
    TRY(dev_zero(g, &g->Tf, (size_t)g->F * 343));
    TRY(dev_zero(g, &g->src32, (size_t)std::max(g->n_src, 1)));
  } else {
    TRY(dev_zero(g, &g->X, spec));
    TRY(dev_zero(g, &g->Y, spec));
    TRY(dev_zero(g, &g->T, (size_t)g->F * 343));
  }
  g->n_witems = (int)P.witems.size() / 2;
  g->n_xitems = (int)P.xitems.size() / 2;
  g->n_wred = (int)P.wred.size() / 2;
  g->n_xred = (int)P.xred.size() / 2;
  g->n_l2t = (int)P.l2t_items.size();
  TRY(dev_upload(g, &g->witems, P.witems.data(), P.witems.size()));
  TRY(dev_upload(g, &g->xitems, P.xitems.data(), P.xitems.size()));
  TRY(dev_upload(g, &g->wred, P.wred.data(), P.wred.size()));
  TRY(dev_upload(g, &g->xred, P.xred.data(), P.xred.size()));
  TRY(dev_zero(g, &g->wxpart, (size_t)std::max(P.w_slots, P.x_slots) * 64));
  TRY(dev_upload(g, &g->l2t_items, P.l2t_items.data(), P.l2t_items.size()));
  return KIFMM_OK;
}

/* FP32 leaf-local source records of the mixed S2M ((s - c) / h, w): they
 * depend on the positions, fixed for the handle, and the densities, so they
 * are written here at create time and again by kifmm_gpu_set_densities, not
 * in every evaluation. */
int mixed_sources(kifmm_gpu* g) {
  if (g->n_s2m > 0)
    hipLaunchKernelGGL(k_src_local32, dim3(g->n_s2m), dim3(256), 0, g->stream, g->s2m_items, g->geo, g->src4,
                       g->src32);
  HIP_TRY(hipGetLastError());
  return KIFMM_OK;
}

/* Mixed path, once per evaluator: leaves per block for the mixed S2M and
 * L2T, and the FP32 source records. */
int mixed_setup(kifmm_gpu* g) {
  g->mix_lpb = (g->n_s2m > 0 && (double)g->n_src / g->n_s2m < 100.0) ? 4 : 1;
  TRY(mixed_sources(g));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return KIFMM_OK;
}

/* Close pairs of the U-list (k_close_near): a count pass, the exclusive
 * offsets on the device (k_scan_*), then the fill pass.  Only the total
 * comes back to the host, to size the index array.  An input whose list
 * would exceed 16 pairs per target on average is not listed: the mixed path
 * then keeps the in-loop test (k_p2p_ulist_mixed), the FP64 path the
 * difference form. */
int close_list(kifmm_gpu* g) {
  const long long n = g->n_trg;
  const int nb = (int)std::max<long long>(1, (n + SCAN_PER - 1) / SCAN_PER);
  if (nb > SCAN_PER) return KIFMM_OK; /* beyond the scan's reach: not listed */
  struct Tmp {
    std::vector<void*> p;
    ~Tmp() {
      for (void* q : p) (void)hipFree(q);
    }
  } tmp;
  auto tmp_alloc = [&](size_t bytes) -> void* {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(bytes, 8)) != hipSuccess) return nullptr;
    tmp.p.push_back(q);
    return q;
  };
  int* cnt = (int*)tmp_alloc(sizeof(int) * std::max<long long>(n, 1));
  long long* bsum = (long long*)tmp_alloc(sizeof(long long) * (nb + 1));
  if (!cnt || !bsum) {
    set_error("hipMalloc failed for the close-pair search");
    return KIFMM_ENOMEM;
  }
  long long* total = bsum + nb;
  const int per5 = (g->n_p2p5 + 7) / 8;
  if (g->n_p2p5 > 0)
    hipLaunchKernelGGL(k_close_near<false>, dim3(8 * per5), dim3(64), 0, g->stream, g->p2p_items5, g->n_p2p5, per5,
                       g->u_ptr, g->u_idx, g->sbeg, g->snum, g->geo, g->src4, g->trg4, cnt, nullptr, nullptr);
  HIP_TRY(hipGetLastError());
  if (n > 0) hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(SCAN_T), 0, g->stream, cnt, n, bsum);
  hipLaunchKernelGGL(k_scan_tops, dim3(1), dim3(SCAN_T), 0, g->stream, bsum, n > 0 ? nb : 0, total);
  HIP_TRY(hipGetLastError());
  long long tot = 0;
  HIP_TRY(hipMemcpyAsync(&tot, total, sizeof tot, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  const long long cap = std::min<long long>(std::max(16LL * n, 1LL << 20), INT_MAX - 1);
  if (tot > cap) return KIFMM_OK; /* close_listed stays 0 */
  TRY(dev_zero(g, &g->cptr, (size_t)n + 1));
  TRY(dev_zero(g, &g->cidx, (size_t)std::max<long long>(tot, 1)));
  if (n > 0)
    hipLaunchKernelGGL(k_scan_out, dim3(nb), dim3(SCAN_T), 0, g->stream, cnt, n, bsum, total, g->cptr);
  if (g->n_p2p5 > 0)
    hipLaunchKernelGGL(k_close_near<true>, dim3(8 * per5), dim3(64), 0, g->stream, g->p2p_items5, g->n_p2p5, per5,
                       g->u_ptr, g->u_idx, g->sbeg, g->snum, g->geo, g->src4, g->trg4, nullptr, g->cptr, g->cidx);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->stream));
  g->n_close = tot;
  g->close_listed = 1;
  return KIFMM_OK;
}

/* The 343 translation spectra (k_trans; compute_fft_trans__gpu_eval,
 * fmm/cuda.cu:874-963, which the reference recomputes every run): they
 * depend only on NP, so they are formed once per handle. */
template <int NP>
int trans_setup_np(kifmm_gpu* g) {
  if (g->precision == KIFMM_PREC_MIXED)
    hipLaunchKernelGGL((k_trans<NP, float2>), dim3(343), dim3(256), 0, g->stream, g->tw, g->Tf);
  else
    hipLaunchKernelGGL((k_trans<NP, double2>), dim3(343), dim3(256), 0, g->stream, g->tw, g->T);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->stream));
  return KIFMM_OK;
}
int trans_setup(kifmm_gpu* g) {
  switch (g->np) {
    case 3: return trans_setup_np<3>(g);
    case 4: return trans_setup_np<4>(g);
    case 6: return trans_setup_np<6>(g);
    default: set_error("NP must be 3, 4 or 6 on the GPU"); return KIFMM_EUNSUP;
  }
}

/* FP64 path with listed close pairs: the U-list runs on the Gram form
 * (k_p2p_ulist<., ., true>), whose per-target records are written here. */
int gram_setup(kifmm_gpu* g) {
  TRY(dev_zero(g, &g->trg_gram, (size_t)std::max(g->n_trg, 1)));
  if (g->n_p2p > 0)
    hipLaunchKernelGGL(k_trg_gram, dim3(g->n_p2p), dim3(128), 0, g->stream, g->p2p_items, g->geo, g->trg4,
                       g->trg_gram);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->stream));
  return KIFMM_OK;
}

#define LAUNCH_CHECK()                                                           \
  do {                                                                           \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess) {                                                      \
      set_error(std::string("kernel launch: ") + hipGetErrorString(e_));         \
      return KIFMM_EDEVICE;                                                      \
    }                                                                            \
  } while (0)

#define REC(e)                                                   \
  do {                                                           \
    if (!g->capturing) HIP_TRY(hipEventRecord((e), g->stream)); \
  } while (0)

int gemm(kifmm_gpu* g, const GemmList& L, int M, int K, const double* A, int lda, const double* in, double* out,
         int accumulate) {
  if (L.n == 0) return KIFMM_OK;
  const GemmShape sh = gemm_shape(L.n, M, K);
  if (sh.nsplit > 1 && (size_t)sh.nsplit * L.n * sh.ldw > g->ws_len) {
    set_error("internal: GEMM workspace too small");
    return KIFMM_EINVAL;
  }
  const hipError_t e = launch_gemm_cols(g->stream, sh, M, K, A, lda, L.n, L.in_off, L.out_off, L.sexp, in, out,
                                        accumulate, g->ws);
  if (e != hipSuccess) {
    set_error(std::string("GEMM launch: ") + hipGetErrorString(e));
    return KIFMM_EDEVICE;
  }
  return KIFMM_OK;
}

/* Per level, M2M (UE2UC stack, then UC2UE) and L2L + DC2DE run either as
 * two column GEMMs or as one fused kernel.  A fused block runs every k step
 * of both operators for its 16 columns with 4 waves: fine for the square
 * n_e x n_e operators of the downward pass and for the UE2UC stack (K =
 * 8 n_e) up to NP = 4, but at NP = 6 (K = 1216) that serial chain per block
 * is longer than the column GEMMs (64 columns a block, K split over blocks
 * on short levels).  The automatic choice is fused down, fused up at
 * NP <= 4 and GEMMs up at NP = 6; KIFMM_LEVEL_GEMM / _FUSED force one form. */
bool fused_up(const kifmm_gpu* g) {
  return g->up_form == KIFMM_LEVEL_FUSED || (g->up_form == KIFMM_LEVEL_AUTO && g->np <= 4);
}
bool fused_down(const kifmm_gpu* g) {
  return g->down_form != KIFMM_LEVEL_GEMM; /* AUTO: the fused kernel at every NP */
}

int m2m(kifmm_gpu* g, const GemmList& A, const GemmList& B, int d) {
  const dim3 grid((A.n + 15) / 16);
  const int rt1 = g->lda_uc / 16, rt2 = g->lda_e / 16;
#define KIFMM_M2M_CASE(R1, R2)                                                                                    \
  if (rt1 == R1 && rt2 == R2)                                                                                     \
    hipLaunchKernelGGL((k_m2m<R1, R2>), grid, dim3(256), 0, g->stream, g->n_e, g->ld_e, g->ld_uc, g->UE2UC,        \
                       g->lda_uc, g->UC2UE, g->lda_e, A.n, A.in_off, B.out_off, d, g->uden);
  KIFMM_M2M_CASE(7, 2)
  else KIFMM_M2M_CASE(10, 4)
  else KIFMM_M2M_CASE(19, 10)
  else {
    set_error("internal: no fused upward kernel for these operator sizes");
    return KIFMM_EINVAL;
  }
#undef KIFMM_M2M_CASE
  LAUNCH_CHECK();
  return KIFMM_OK;
}

/* L2L of the parents at depth d - 1 and DC2DE of their children, fused */
int l2l_dc2de(kifmm_gpu* g, const GemmList& L, int d) {
  if (L.n == 0) return KIFMM_OK;
  const dim3 grid((L.n + 15) / 16, 8);
  switch ((g->n_e + 15) / 16) {
#define KIFMM_DOWN_CASE(R)                                                                                        \
  case R:                                                                                                         \
    hipLaunchKernelGGL(k_l2l_dc2de<R>, grid, dim3(256), 0, g->stream, g->n_e, g->ld_e, g->DE2DC, g->lda_l2l,       \
                       g->DC2DE, g->lda_e, L.n, L.in_off, L.out_off, g->tag, d, g->dchk, g->dden);                 \
    break;
    KIFMM_DOWN_CASE(2)
    KIFMM_DOWN_CASE(4)
    KIFMM_DOWN_CASE(10)
#undef KIFMM_DOWN_CASE
    default:
      set_error("internal: no fused downward kernel for this n_e");
      return KIFMM_EINVAL;
  }
  LAUNCH_CHECK();
  return KIFMM_OK;
}

int fill(kifmm_gpu* g, double* p, size_t n) {
  const int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(k_fill, dim3(std::max(blocks, 1)), dim3(256), 0, g->stream, p, n, 0.0);
  LAUNCH_CHECK();
  return KIFMM_OK;
}

template <int NP, typename ST>
int vlist_phase(kifmm_gpu* g, ST* X, ST* Y, ST* T, hipEvent_t e1, hipEvent_t e2, hipEvent_t e3) {
  if (g->n_vblocks > 0)
    hipLaunchKernelGGL((k_dft_fwd<NP, ST>), dim3(g->n_vblocks), dim3(512), 0, g->stream, g->vblocks, g->vslot,
                       g->depth, g->uden, g->ld_e, g->tw, g->cube2surf, X);
  LAUNCH_CHECK();
  REC(e1);
  /* the translation spectra T depend only on NP: computed once at create
   * time (trans_setup), so this sub-phase is empty */
  REC(e2);
  /* XCD-contiguous block order (had_block): 8 per_xcd blocks for the
   * n_hgroups x F / fb (group, frequency block) pairs */
  const int had_per_xcd = (int)(((long)g->n_hgroups * (g->F / g->had_fb) + 7) / 8);
  if (g->n_hgroups > 0)
    switch (g->had_fb * 2 + g->had_form) {
#define KIFMM_HAD_CASE(FBV)                                                                                    \
  case 2 * FBV:                                                                                                \
    hipLaunchKernelGGL((k_hadamard<NP, ST, FBV>), dim3(8 * had_per_xcd), dim3(128), 0, g->stream, g->hgroups,  \
                       g->huslots, g->hloc, g->hslot, T, X, Y, g->n_hgroups, had_per_xcd);                    \
    break;                                                                                                     \
  case 2 * FBV + 1:                                                                                            \
    hipLaunchKernelGGL((k_hadamard_mfma<NP, ST, FBV>), dim3(8 * had_per_xcd), dim3(256), 0, g->stream,         \
                       g->hgroups, g->huslots, g->hloc, g->hslot, T, X, Y, g->n_hgroups, had_per_xcd);        \
    break;
      KIFMM_HAD_CASE(4)
      KIFMM_HAD_CASE(8)
      KIFMM_HAD_CASE(16)
#undef KIFMM_HAD_CASE
      default:
        set_error("internal: unknown Hadamard shape");
        return KIFMM_EINVAL;
    }
  LAUNCH_CHECK();
  REC(e3);
  if (g->n_vblocks > 0)
    hipLaunchKernelGGL((k_dft_inv<NP, ST>), dim3(g->n_vblocks), dim3(512), 0, g->stream, g->vblocks, g->vslot, Y,
                       g->tw, g->surf2cube, g->dchk, g->ld_e);
  LAUNCH_CHECK();
  return KIFMM_OK;
}

template <int NP>
int vlist_phase(kifmm_gpu* g, hipEvent_t e1, hipEvent_t e2, hipEvent_t e3) {
  if (g->precision == KIFMM_PREC_MIXED) return vlist_phase<NP, float2>(g, g->Xf, g->Yf, g->Tf, e1, e2, e3);
  return vlist_phase<NP, double2>(g, g->X, g->Y, g->T, e1, e2, e3);
}

/* The item list the U-list launch uses: the mixed tau kernel takes sibling
 * pairs and single leaves (list 3), the FP64 Gram-form kernels items of
 * <= P2P_MAXT5 targets (list 2), the STATIC64 shape items of <= 64 (list
 * 1), the difference form, the mixed source-split form and the mixed
 * in-loop test items of <= P2P_MAXT (list 0). */
int u_list_index(const kifmm_gpu* g) {
  if (g->precision == KIFMM_PREC_MIXED) return g->close_listed && g->p2p_form != KIFMM_P2P_SPLIT ? 3 : 0;
  if (g->p2p_form == KIFMM_P2P_STATIC64) return 1;
  return g->trg_gram ? 2 : 0;
}

/* U-list over the items of copy-out chunks [c0, c1) on stream st, writing
 * out[t] = u (accum = 0) or out[t] = u + out[t] (accum = 1).  The
 * persistent form's per-XCD counters of the launch are slot c0. */
template <bool MASK>
int launch_u(kifmm_gpu* g, hipStream_t st, int c0, int c1, double* out, int accum) {
  const int l = u_list_index(g);
  const int4* list = l == 0 ? g->p2p_items : l == 1 ? g->p2p_items64 : l == 2 ? g->p2p_items5 : g->p2p_itemsP;
  const int beg = g->chunk_beg[l][c0], end = g->chunk_beg[l][c1];
  const int chunk = c0;
  const int n = end - beg, per = (n + 7) / 8;
  if (n <= 0) return KIFMM_OK;
  const int4* items = list + beg;
  if (g->precision == KIFMM_PREC_MIXED) {
    if (g->close_listed && g->p2p_form == KIFMM_P2P_SPLIT)
      hipLaunchKernelGGL(k_p2p_ulist_mixed_split<MASK>, dim3(8 * per), dim3(256), 0, st, items, n, per, g->rng_ptr,
                         g->rng, g->geo, g->src4, g->trg4, g->cptr, g->cidx, out, accum);
    else if (g->close_listed)
      hipLaunchKernelGGL(k_p2p_ulist_mixed_tau<MASK>, dim3(8 * per), dim3(256), 0, st, items, n, per, g->rng_ptr,
                         g->rng, g->geo, g->src4, g->trg4, g->cptr, g->cidx, out, accum);
    else
      hipLaunchKernelGGL(k_p2p_ulist_mixed<MASK>, dim3(8 * per), dim3(256), 0, st, items, n, per, g->rng_ptr, g->rng,
                         g->geo, g->src4, g->trg4, out, accum);
  } else if (g->p2p_form == KIFMM_P2P_STATIC64) {
    if (g->trg_gram)
      hipLaunchKernelGGL((k_p2p_ulist64<MASK, true>), dim3(8 * per), dim3(256), 0, st, items, n, per, g->rng_ptr,
                         g->rng, g->src4, g->trg4, out, g->geo, g->trg_gram, g->cptr, g->cidx, accum);
    else
      hipLaunchKernelGGL((k_p2p_ulist64<MASK, false>), dim3(8 * per), dim3(256), 0, st, items, n, per, g->rng_ptr,
                         g->rng, g->src4, g->trg4, out, g->geo, g->trg_gram, g->cptr, g->cidx, accum);
  } else if (g->p2p_form == KIFMM_P2P_PERSISTENT) {
    int* counters = g->p2p_counters + 8 * std::max(chunk, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, counters, 8, 0);
    if (g->trg_gram)
      hipLaunchKernelGGL((k_p2p_ulist_persist<MASK, P2P_G5, true>), dim3(g->persist_grid[1]), dim3(256), 0, st, items,
                         n, per, g->rng_ptr, g->rng, g->src4, g->trg4, out, counters, g->geo, g->trg_gram, g->cptr,
                         g->cidx, accum);
    else
      hipLaunchKernelGGL((k_p2p_ulist_persist<MASK, P2P_G, false>), dim3(g->persist_grid[0]), dim3(256), 0, st, items,
                         n, per, g->rng_ptr, g->rng, g->src4, g->trg4, out, counters, g->geo, g->trg_gram, g->cptr,
                         g->cidx, accum);
  } else {
    if (g->trg_gram)
      hipLaunchKernelGGL((k_p2p_ulist<MASK, P2P_G5, true>), dim3(8 * per), dim3(256), 0, st, items, n, per,
                         g->rng_ptr, g->rng, g->src4, g->trg4, out, g->geo, g->trg_gram, g->cptr, g->cidx, accum);
    else
      hipLaunchKernelGGL((k_p2p_ulist<MASK, P2P_G, false>), dim3(8 * per), dim3(256), 0, st, items, n, per,
                         g->rng_ptr, g->rng, g->src4, g->trg4, out, g->geo, g->trg_gram, g->cptr, g->cidx, accum);
  }
  LAUNCH_CHECK();
  return KIFMM_OK;
}

int launch_u(kifmm_gpu* g, hipStream_t st, int c0, int c1, double* out, int accum) {
  return g->coincident ? launch_u<true>(g, st, c0, c1, out, accum) : launch_u<false>(g, st, c0, c1, out, accum);
}

/* Copy-out chunks whose U-list runs on the second stream from the start of
 * the evaluation (alongside the far field) rather than after it: none in
 * the default layout (tuning.overlap 0) and the sequential one (4), all of
 * them in the overlap layout (1), the first one or two in the split
 * layouts (2, 3 and their sequential forms 5, 6). */
int early_chunks(const kifmm_gpu* g, unsigned mask) {
  if (!(mask & KIFMM_PHASE_U)) return 0;
  const int nch = (int)g->cuts.size() - 1;
  switch (g->overlap) {
    case 1: return nch;
    case 2: case 5: return std::min(1, nch);
    case 3: case 6: return std::min(2, nch);
    default: return 0;
  }
}

template <bool MASK>
int point_phases_wx(kifmm_gpu* g, unsigned mask, hipEvent_t ew) {
  if ((mask & KIFMM_PHASE_W) && g->n_witems > 0) {
    hipLaunchKernelGGL((k_wx<MASK, false>), dim3(g->n_witems), dim3(64), 0, g->stream, g->witems, g->widx, g->child,
                       g->snum, g->sbeg, g->tbeg, g->geo, g->surf_ue, g->surf_dc, g->ld_surf, g->n_e, g->uden,
                       g->ld_e, g->src4, g->trg4, g->pot, g->dchk, g->wxpart);
    if (g->n_wred > 0)
      hipLaunchKernelGGL(k_wx_reduce, dim3(g->n_wred), dim3(64), 0, g->stream, g->wred, g->tbeg, g->ld_e, g->wxpart,
                         g->pot, g->dchk);
  }
  LAUNCH_CHECK();
  REC(ew);
  if ((mask & KIFMM_PHASE_X) && g->n_xitems > 0) {
    hipLaunchKernelGGL((k_wx<MASK, true>), dim3(g->n_xitems), dim3(64), 0, g->stream, g->xitems, g->xidx, g->child,
                       g->snum, g->sbeg, g->tbeg, g->geo, g->surf_ue, g->surf_dc, g->ld_surf, g->n_e, g->uden,
                       g->ld_e, g->src4, g->trg4, g->pot, g->dchk, g->wxpart);
    if (g->n_xred > 0)
      hipLaunchKernelGGL(k_wx_reduce, dim3(g->n_xred), dim3(64), 0, g->stream, g->xred, g->tbeg, g->ld_e, g->wxpart,
                         g->pot, g->dchk);
  }
  LAUNCH_CHECK();
  return KIFMM_OK;
}

/* Phase events (recorded outside graph capture): E_START before the
 * sequence, then one after each phase in launch order. */
enum { E_START = 0, E_UP, E_FFT, E_TRANS, E_HAD, E_IFFT, E_W, E_X, E_DOWN, E_U, E_END };

/* device -> host copy of potentials [b, e) to dst + b */
int copy_out(kifmm_gpu* g, hipStream_t st, double* dst, int b, int e) {
  if (e > b)
    HIP_TRY(hipMemcpyAsync(dst + b, g->pot + b, sizeof(double) * (size_t)(e - b), hipMemcpyDeviceToHost, st));
  return KIFMM_OK;
}

/* One evaluation is three launch sequences on the handle's streams:
 *   enqueue_side  the U-list of the early copy-out chunks (early_chunks) on
 *                 the second stream from the start, into potU, alongside
 *                 the far field;
 *   enqueue_far   the fills, up, V, W, X and down on the main stream (what a
 *                 graph handle captures and replays);
 *   enqueue_near  the join pot = potU + pot of the early chunks, then the
 *                 U-list of the other chunks, alternating between the two
 *                 compute streams, each chunk's potentials copied to the
 *                 host (dst, when non-null) on the copy stream while the
 *                 later chunks compute.
 * The U-list's accumulating write (u + pot) and the join (potU + pot) add
 * the same two terms, so every layout gives the same bits. */
int enqueue_side(kifmm_gpu* g, unsigned mask) {
  const int e = early_chunks(g, mask);
  if (e == 0) return KIFMM_OK;
  HIP_TRY(hipEventRecord(g->fork_ev, g->stream)); /* after the previous evaluation's join read potU */
  HIP_TRY(hipStreamWaitEvent(g->side, g->fork_ev, 0));
  HIP_TRY(hipEventRecord(g->su0, g->side));
  TRY(launch_u(g, g->side, 0, e, g->potU, 0));
  HIP_TRY(hipEventRecord(g->su1, g->side));
  HIP_TRY(hipEventRecord(g->join_ev, g->side));
  return KIFMM_OK;
}

int enqueue_far(kifmm_gpu* g, unsigned mask) {
  hipEvent_t* ev = g->ev;
  const int lde = g->ld_e, lduc = g->ld_uc;
  /* zero what this call accumulates into (fill kernels, so the sequence
   * stays capturable): pot (W, X, L2T and the U-list add into it) and dchk
   * (V, X, L2L) always; uden and dden only when their phase, which assigns
   * every entry that is read, is left out */
  TRY(fill(g, g->pot, (size_t)g->n_trg));
  if (!(mask & KIFMM_PHASE_UP)) TRY(fill(g, g->uden, (size_t)g->n_nodes * lde));
  TRY(fill(g, g->dchk, (size_t)g->n_nodes * lde));
  if (!(mask & KIFMM_PHASE_DOWN)) TRY(fill(g, g->dden, (size_t)g->n_nodes * lde));
  /* up: S2M check potentials, UC2UE, then M2M bottom-up */
  if ((mask & KIFMM_PHASE_UP) && g->n_s2m > 0) {
    /* sources never sit on the check surface: no r = 0 mask */
    if (g->precision == KIFMM_PREC_MIXED) {
      /* the FP32 leaf-local source records (mixed_sources) */
      /* (phases, check points per lane) as in the FP64 S2M; NP <= 4 here (98 or 152 check points) */
      if (g->n_uc > 160) {
        set_error("mixed S2M: more than 160 check points");
        return KIFMM_EUNSUP;
      }
      const bool small = g->n_uc <= 112;
      auto s2m = g->mix_lpb == 4 ? (small ? k_s2m_mixed<4, 4, 7> : k_s2m_mixed<4, 2, 5>)
                                 : (small ? k_s2m_mixed<1, 4, 7> : k_s2m_mixed<1, 2, 5>);
      hipLaunchKernelGGL(s2m, dim3((g->n_s2m + g->mix_lpb - 1) / g->mix_lpb), dim3(64), 0, g->stream, g->s2m_items,
                         g->n_s2m, g->geo, g->surf_uc32, g->n_uc, g->ld_surf, g->src32, g->chk, lduc);
    } else {
      /* (phases, check points per lane) by the check surface size: 98, 152, 296 for NP 3, 4, 6 */
      auto s2m = g->n_uc <= 112 ? k_s2m<4, 7> : g->n_uc <= 160 ? k_s2m<2, 5> : k_s2m<1, 5>;
      if (g->n_uc > 320) {
        set_error("S2M: more than 320 check points");
        return KIFMM_EUNSUP;
      }
      hipLaunchKernelGGL(s2m, dim3(g->n_s2m), dim3(64), 0, g->stream, g->s2m_items, g->geo, g->surf_uc, g->n_uc,
                         g->ld_surf, g->src4, g->chk, lduc);
    }
    LAUNCH_CHECK();
    TRY(gemm(g, g->s2m_gemm, g->n_e, lduc, g->UC2UE, g->lda_e, g->chk, g->uden, 0));
  }
  for (int d = g->max_depth - 1; d >= 0 && (mask & KIFMM_PHASE_UP); d--) {
    if (g->lv[d].m2m_a.n == 0) continue;
    if (fused_up(g)) {
      TRY(m2m(g, g->lv[d].m2m_a, g->lv[d].m2m_b, d));
    } else {
      TRY(gemm(g, g->lv[d].m2m_a, g->n_uc, 8 * lde, g->UE2UC, g->lda_uc, g->uden, g->chk, 0));
      TRY(gemm(g, g->lv[d].m2m_b, g->n_e, lduc, g->UC2UE, g->lda_e, g->chk, g->uden, 0));
    }
  }
  REC(ev[E_UP]);
  /* V-list */
  if (mask & KIFMM_PHASE_V) switch (g->np) {
    case 3: TRY(vlist_phase<3>(g, ev[E_FFT], ev[E_TRANS], ev[E_HAD])); break;
    case 4: TRY(vlist_phase<4>(g, ev[E_FFT], ev[E_TRANS], ev[E_HAD])); break;
    case 6: TRY(vlist_phase<6>(g, ev[E_FFT], ev[E_TRANS], ev[E_HAD])); break;
    default: set_error("NP must be 3, 4 or 6 on the GPU"); return KIFMM_EUNSUP;
  }
  else {
    REC(ev[E_FFT]);
    REC(ev[E_TRANS]);
    REC(ev[E_HAD]);
  }
  REC(ev[E_IFFT]);
  /* W and X */
  if (mask & (KIFMM_PHASE_W | KIFMM_PHASE_X))
    TRY(g->coincident ? point_phases_wx<true>(g, mask, ev[E_W]) : point_phases_wx<false>(g, mask, ev[E_W]));
  else
    REC(ev[E_W]);
  REC(ev[E_X]);
  /* down: DC2DE at depth 2, then per level the fused L2L + DC2DE
   * (k_l2l_dc2de over the parents at depth d - 1), then L2T */
  for (int d = 2; d <= g->max_depth && (mask & KIFMM_PHASE_DOWN); d++) {
    if (d >= 3 && fused_down(g)) {
      TRY(l2l_dc2de(g, g->lv[d - 1].l2l, d));
      continue;
    }
    if (d >= 3) TRY(gemm(g, g->lv[d - 1].l2l, 8 * lde, lde, g->DE2DC, g->lda_l2l, g->dden, g->dchk, 1));
    TRY(gemm(g, g->lv[d].dc2de, g->n_e, lde, g->DC2DE, g->lda_e, g->dchk, g->dden, 0));
  }
  if ((mask & KIFMM_PHASE_DOWN) && g->n_l2t > 0) {
    if (g->precision == KIFMM_PREC_MIXED)
      hipLaunchKernelGGL(g->mix_lpb == 4 ? k_l2t_mixed<4> : k_l2t_mixed<1>, dim3((g->n_l2t + g->mix_lpb - 1) / g->mix_lpb),
                         dim3(64), 0, g->stream, g->l2t_items, g->n_l2t, g->geo, g->surf_de32,
                         g->ld_surf, g->n_e, g->dden, lde, g->trg4, g->pot);
    else
      hipLaunchKernelGGL(k_l2t, dim3(g->n_l2t), dim3(64), 0, g->stream, g->l2t_items, g->geo, g->surf_de, g->ld_surf,
                         g->n_e, g->dden, lde, g->trg4, g->pot);
  }
  LAUNCH_CHECK();
  REC(ev[E_DOWN]);
  return KIFMM_OK;
}

int enqueue_near(kifmm_gpu* g, unsigned mask, double* dst) {
  const int n = g->n_trg;
  if (!(mask & KIFMM_PHASE_U)) {
    HIP_TRY(hipEventRecord(g->ev[E_U], g->stream));
    if (dst) TRY(copy_out(g, g->stream, dst, 0, n));
    return KIFMM_OK;
  }
  const int nch = (int)g->cuts.size() - 1, e = early_chunks(g, mask);
  if (e > 0) { /* the early chunks: join, then their copy */
    const size_t m = (size_t)g->cuts[e];
    HIP_TRY(hipStreamWaitEvent(g->stream, g->join_ev, 0));
    const int blocks = (int)std::min<size_t>(4096, (m + 255) / 256);
    if (m > 0) hipLaunchKernelGGL(k_join_pot, dim3(std::max(blocks, 1)), dim3(256), 0, g->stream, g->pot, g->potU, m);
    LAUNCH_CHECK();
    if (dst) {
      HIP_TRY(hipEventRecord(g->ch_ev[0], g->stream));
      HIP_TRY(hipStreamWaitEvent(g->cstream, g->ch_ev[0], 0));
      TRY(copy_out(g, g->cstream, dst, 0, (int)m));
      for (int c = 0; c < e; c++) HIP_TRY(hipEventRecord(g->cd_ev[c], g->cstream));
    }
  }
  if (e < nch) { /* the other chunks after the far field */
    const int late = dst ? nch - e : 1; /* without a copy: one launch */
    /* on the two streams in turn (one chunk's last blocks run beside the
     * next chunk's first), or in layouts 4 - 6 one after another on the
     * main stream: the first copies then start sooner, which pays when the
     * copy back is long against the U-list */
    const bool pairs = late > 1 && g->overlap < 4;
    if (pairs) {
      HIP_TRY(hipEventRecord(g->far_ev, g->stream));
      HIP_TRY(hipStreamWaitEvent(g->side, g->far_ev, 0));
    }
    for (int k = 0; k < late; k++) {
      const int c = e + k;
      hipStream_t st = (pairs && (k & 1)) ? g->side : g->stream;
      TRY(launch_u(g, st, c, dst ? c + 1 : nch, g->pot, 1));
      if (dst) {
        HIP_TRY(hipEventRecord(g->ch_ev[c], st));
        HIP_TRY(hipStreamWaitEvent(g->cstream, g->ch_ev[c], 0));
        TRY(copy_out(g, g->cstream, dst, g->cuts[c], g->cuts[c + 1]));
        HIP_TRY(hipEventRecord(g->cd_ev[c], g->cstream));
      }
    }
    if (pairs) {
      HIP_TRY(hipEventRecord(g->sdone_ev, g->side));
      HIP_TRY(hipStreamWaitEvent(g->stream, g->sdone_ev, 0));
    }
  }
  HIP_TRY(hipEventRecord(g->ev[E_U], g->stream));
  if (dst) {
    HIP_TRY(hipEventRecord(g->cdone_ev, g->cstream));
    HIP_TRY(hipStreamWaitEvent(g->stream, g->cdone_ev, 0));
  }
  return KIFMM_OK;
}

/* pinned potentials -> the caller's buffer, a few host threads */
void host_copy(double* to, const double* from, size_t b, size_t e) {
  const size_t n = e - b;
  const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, n / (1 << 16)));
  if (nt == 1) {
    std::memcpy(to + b, from + b, sizeof(double) * n);
    return;
  }
  std::vector<std::thread> th;
  for (int k = 0; k < nt; k++) {
    const size_t x = b + n * k / nt, y = b + n * (k + 1) / nt;
    th.emplace_back([=] { std::memcpy(to + x, from + x, sizeof(double) * (y - x)); });
  }
  for (auto& t : th) t.join();
}

int evaluate(kifmm_gpu* g, unsigned mask, double* pot_leaf, kifmm_phase_times* times) {
  hipEvent_t* ev = g->ev;
  HIP_TRY(hipSetDevice(g->dev));
  /* a graph handle replays the far-field sequence (fills, up, V, W, X,
   * down) as one graph: HIP runs a graph's nodes in one order on the launch
   * stream, so the U-list chunks and their copies, whose point is to run
   * beside one another, stay outside it */
  const bool graph = g->use_graph && mask == KIFMM_PHASE_ALL;
  /* the copy lands in the handle's pinned buffer (the caller's own buffer
   * when it is that one; otherwise one host pass follows), or without
   * pinned memory straight into the caller's buffer */
  double* dst = pot_leaf ? (g->pinned ? g->pinned : pot_leaf) : nullptr;
  if (graph && !g->gexec[0]) {
    g->capturing = true;
    HIP_TRY(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    const int r = enqueue_far(g, mask);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(g->stream, &gr);
    g->capturing = false;
    if (r != KIFMM_OK) {
      if (gr) (void)hipGraphDestroy(gr);
      return r;
    }
    HIP_TRY(e);
    g->graph[0] = gr;
    HIP_TRY(hipGraphInstantiate(&g->gexec[0], gr, nullptr, nullptr, 0));
  }
  HIP_TRY(hipEventRecord(ev[E_START], g->stream));
  TRY(enqueue_side(g, mask));
  if (graph) {
    HIP_TRY(hipGraphLaunch(g->gexec[0], g->stream));
    for (int i = E_UP; i <= E_DOWN; i++) HIP_TRY(hipEventRecord(ev[i], g->stream));
  } else {
    TRY(enqueue_far(g, mask));
  }
  TRY(enqueue_near(g, mask, dst));
  HIP_TRY(hipEventRecord(ev[E_END], g->stream));
  double host_copy_ms = 0.0;
  if (pot_leaf && dst == g->pinned && pot_leaf != g->pinned) {
    const int nch = (int)g->cuts.size() - 1;
    const bool pipelined = (mask & KIFMM_PHASE_U) && nch > 1; /* cd_ev recorded for every chunk */
    double t0 = 0.0;
    if (pipelined) { /* each chunk as soon as its copy is done */
      for (int c = 0; c < nch; c++) {
        HIP_TRY(hipEventSynchronize(g->cd_ev[c]));
        if (c == nch - 1) t0 = kifmm::wall_seconds();
        host_copy(pot_leaf, g->pinned, (size_t)g->cuts[c], (size_t)g->cuts[c + 1]);
      }
    } else {
      HIP_TRY(hipEventSynchronize(ev[E_END]));
      t0 = kifmm::wall_seconds();
      host_copy(pot_leaf, g->pinned, 0, (size_t)g->n_trg);
    }
    host_copy_ms = (kifmm::wall_seconds() - t0) * 1e3; /* what is left after the device work */
  }
  HIP_TRY(hipEventSynchronize(ev[E_END]));
  if (times) {
    float ms[E_END + 1] = {0};
    for (int i = E_UP; i <= E_END; i++) HIP_TRY(hipEventElapsedTime(&ms[i], ev[i - 1], ev[i]));
    float tot = 0;
    HIP_TRY(hipEventElapsedTime(&tot, ev[E_START], ev[E_END]));
    std::memset(times, 0, sizeof *times);
    if (!graph) {
      times->up = ms[E_UP];
      times->v_fft_src = ms[E_FFT];
      times->v_trans = ms[E_TRANS];
      times->v_hadamard = ms[E_HAD];
      times->v_ifft = ms[E_IFFT];
      times->vlist = ms[E_FFT] + ms[E_TRANS] + ms[E_HAD] + ms[E_IFFT];
      times->wlist = ms[E_W];
      times->xlist = ms[E_X];
      times->down = ms[E_DOWN];
    }
    /* the U-list: its chunks after the far field, plus the span of the
     * early chunks on the second stream (which overlaps the far field) */
    times->ulist = times->p2p_kernel = ms[E_U];
    if (early_chunks(g, mask) > 0) {
      float mu = 0;
      HIP_TRY(hipEventElapsedTime(&mu, g->su0, g->su1));
      times->ulist += mu;
      times->p2p_kernel = early_chunks(g, mask) == (int)g->cuts.size() - 1 ? mu : times->ulist;
    }
    /* the copy left after the last kernel (the rest overlaps the U-list) */
    times->copy = ms[E_END] + host_copy_ms;
    times->total = tot + host_copy_ms;
  }
  return KIFMM_OK;
}

}  // namespace

extern "C" {

int kifmm_gpu_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n > 0 ? 1 : 0;
}

int kifmm_gpu_create(const kifmm_tree_view* tree, const kifmm_operators* ops, const kifmm_gpu_options* opts,
                     kifmm_gpu** out) {
  if (!tree || !ops || !out) {
    set_error("kifmm_gpu_create: null argument");
    return KIFMM_EINVAL;
  }
  *out = nullptr;
  if (tree->np != 3 && tree->np != 4 && tree->np != 6) {
    set_error("kifmm_gpu_create: NP must be 3, 4 or 6 on the GPU (got " + std::to_string(tree->np) + ")");
    return KIFMM_EUNSUP;
  }
  if (ops->np != tree->np || ops->n_e != tree->np * tree->np * tree->np - (tree->np - 2) * (tree->np - 2) * (tree->np - 2)) {
    set_error("kifmm_gpu_create: operators do not match the tree's NP");
    return KIFMM_EINVAL;
  }
  if (tree->n_nodes < 1 || tree->n_src < 0 || tree->n_trg < 0 || (tree->n_nodes - 1) % 8 != 0) {
    set_error("kifmm_gpu_create: malformed tree view");
    return KIFMM_EINVAL;
  }
  /* every index the kernels derive is checked on the host before any device call */
  const double t_plan0 = kifmm::wall_seconds();
  Plan P;
  try {
    const int pr = make_plan(tree, ops, P);
    if (pr != KIFMM_OK) return pr;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed in kifmm_gpu_create");
    return KIFMM_ENOMEM;
  }
  const double t_plan1 = kifmm::wall_seconds();
  kifmm_gpu* g = new (std::nothrow) kifmm_gpu();
  if (!g) {
    set_error("host allocation failed");
    return KIFMM_ENOMEM;
  }
  g->create_ms[0] = (t_plan1 - t_plan0) * 1e3;
  g->dev = opts ? opts->device : 0;
  g->precision = opts ? opts->precision : KIFMM_PREC_FP64;
  g->use_graph = opts ? (opts->flags & KIFMM_GPU_GRAPH) != 0 : 0;
  if (g->precision != KIFMM_PREC_FP64 && !(g->precision == KIFMM_PREC_MIXED && tree->np <= 4)) {
    set_error("kifmm_gpu_create: precision must be FP64, or MIXED with NP <= 4");
    delete g;
    return KIFMM_EUNSUP;
  }
  int r = KIFMM_OK;
  if (hipSetDevice(g->dev) != hipSuccess) {
    set_error("hipSetDevice failed");
    delete g;
    return KIFMM_EDEVICE;
  }
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    delete g;
    return KIFMM_EDEVICE;
  }
  for (auto& e : g->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      set_error("hipEventCreate failed");
      kifmm_gpu_destroy(g);
      return KIFMM_EDEVICE;
    }
  }
  g->overlap = opts ? (opts->flags & KIFMM_GPU_OVERLAP) != 0 : 0;
  /* the second stream exists on every handle: the layout is a tuning choice */
  if (hipStreamCreateWithFlags(&g->side, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&g->cstream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&g->fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&g->join_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&g->far_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&g->sdone_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&g->cdone_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreate(&g->su0) != hipSuccess || hipEventCreate(&g->su1) != hipSuccess) {
    set_error("kifmm_gpu_create: second stream, copy stream or their events could not be created");
    kifmm_gpu_destroy(g);
    return KIFMM_EDEVICE;
  }
  if (scratch_address(g->dev, &g->scratch) != KIFMM_OK) {
    kifmm_gpu_destroy(g);
    return KIFMM_EDEVICE;
  }
  g->ch_ev.assign(P.cuts.size(), nullptr);
  g->cd_ev.assign(P.cuts.size(), nullptr);
  for (size_t c = 0; c < P.cuts.size(); c++)
    if (hipEventCreateWithFlags(&g->ch_ev[c], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->cd_ev[c], hipEventDisableTiming) != hipSuccess) {
      set_error("kifmm_gpu_create: copy-out events could not be created");
      kifmm_gpu_destroy(g);
      return KIFMM_EDEVICE;
    }
  try {
    const double t0 = kifmm::wall_seconds();
    r = build(g, tree, ops, P);
    if (r == KIFMM_OK) r = hipStreamSynchronize(g->stream) == hipSuccess ? KIFMM_OK : KIFMM_EDEVICE;
    const double t1 = kifmm::wall_seconds();
    if (r == KIFMM_OK && !(opts && (opts->flags & KIFMM_GPU_NO_CLOSE_LIST))) r = close_list(g);
    const double t2 = kifmm::wall_seconds();
    if (r == KIFMM_OK) r = trans_setup(g);
    if (r == KIFMM_OK && g->precision == KIFMM_PREC_MIXED) r = mixed_setup(g);
    if (r == KIFMM_OK && g->precision == KIFMM_PREC_FP64 && g->close_listed) r = gram_setup(g);
    const double t3 = kifmm::wall_seconds();
    g->create_ms[1] = (t1 - t0) * 1e3;
    g->create_ms[2] = (t2 - t1) * 1e3;
    g->create_ms[3] = (t3 - t2) * 1e3;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed in kifmm_gpu_create");
    r = KIFMM_ENOMEM;
  }
  if (r != KIFMM_OK) {
    kifmm_gpu_destroy(g);
    return r;
  }
  /* pinned staging for the potentials; without it the copy goes straight
   * to the caller's (pageable) buffer */
  if (hipHostMalloc((void**)&g->pinned, sizeof(double) * std::max(g->n_trg, 1), hipHostMallocDefault) != hipSuccess)
    g->pinned = nullptr;
  if (hipDeviceSynchronize() != hipSuccess) {
    set_error("device synchronisation failed after upload");
    kifmm_gpu_destroy(g);
    return KIFMM_EDEVICE;
  }
  g->create_ms[4] = (kifmm::wall_seconds() - t_plan0) * 1e3;
  *out = g;
  return KIFMM_OK;
}

int kifmm_gpu_evaluate(kifmm_gpu* g, double* pot_leaf, kifmm_phase_times* times) {
  if (!g) {
    set_error("kifmm_gpu_evaluate: null handle");
    return KIFMM_EINVAL;
  }
  return evaluate(g, KIFMM_PHASE_ALL, pot_leaf, times);
}

int kifmm_gpu_evaluate_phases(kifmm_gpu* g, unsigned phases, double* pot_leaf, kifmm_phase_times* times) {
  if (!g) {
    set_error("kifmm_gpu_evaluate_phases: null handle");
    return KIFMM_EINVAL;
  }
  return evaluate(g, phases & KIFMM_PHASE_ALL, pot_leaf, times);
}

int kifmm_gpu_destroy(kifmm_gpu* g) {
  if (!g) return KIFMM_OK;
  (void)hipSetDevice(g->dev);
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  if (g->pinned) (void)hipHostFree(g->pinned);
  if (g->side) (void)hipStreamSynchronize(g->side);
  if (g->cstream) (void)hipStreamSynchronize(g->cstream);
  for (int k = 0; k < 2; k++) {
    if (g->gexec[k]) (void)hipGraphExecDestroy(g->gexec[k]);
    if (g->graph[k]) (void)hipGraphDestroy(g->graph[k]);
  }
  for (void* p : g->allocs) (void)hipFree(p);
  for (auto& e : g->ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : {g->fork_ev, g->join_ev, g->su0, g->su1, g->far_ev, g->sdone_ev, g->cdone_ev})
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : g->ch_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : g->cd_ev)
    if (e) (void)hipEventDestroy(e);
  if (g->cstream) (void)hipStreamDestroy(g->cstream);
  if (g->side) (void)hipStreamDestroy(g->side);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return KIFMM_OK;
}

int kifmm_gpu_direct(kifmm_gpu* g, int nt, const double* tx, const double* ty, const double* tz, double* out) {
  if (!g || nt <= 0 || !tx || !ty || !tz || !out) {
    set_error("kifmm_gpu_direct: invalid argument");
    return KIFMM_EINVAL;
  }
  HIP_TRY(hipSetDevice(g->dev));
  double* d = nullptr;
  HIP_TRY(hipMalloc(&d, sizeof(double) * 4 * (size_t)nt));
  auto done = [&](int r) {
    (void)hipFree(d);
    return r;
  };
  const size_t b = sizeof(double) * nt;
  if (hipMemcpyAsync(d, tx, b, hipMemcpyHostToDevice, g->stream) != hipSuccess ||
      hipMemcpyAsync(d + nt, ty, b, hipMemcpyHostToDevice, g->stream) != hipSuccess ||
      hipMemcpyAsync(d + 2 * (size_t)nt, tz, b, hipMemcpyHostToDevice, g->stream) != hipSuccess) {
    set_error("kifmm_gpu_direct: upload failed");
    return done(KIFMM_EDEVICE);
  }
  hipLaunchKernelGGL(k_direct, dim3(std::min(nt, 65535)), dim3(256), 0, g->stream, nt, d, d + nt, d + 2 * (size_t)nt,
                     g->src4, g->n_src, d + 3 * (size_t)nt);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(out, d + 3 * (size_t)nt, b, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
      hipStreamSynchronize(g->stream) != hipSuccess) {
    set_error("kifmm_gpu_direct: kernel or copy failed");
    return done(KIFMM_EDEVICE);
  }
  return done(KIFMM_OK);
}

int kifmm_gpu_direct_noalloc(kifmm_gpu* g, int nt, const double* tx, const double* ty, const double* tz,
                             double* out) {
  if (!g || nt <= 0 || !tx || !ty || !tz || !out) {
    set_error("kifmm_gpu_direct_noalloc: invalid argument");
    return KIFMM_EINVAL;
  }
  if (nt > KIFMM_DIRECT_NOALLOC_MAX) {
    set_error("kifmm_gpu_direct_noalloc: " + std::to_string(nt) + " points exceed the scratch of " +
              std::to_string(KIFMM_DIRECT_NOALLOC_MAX) + " (use kifmm_gpu_direct)");
    return KIFMM_EINVAL;
  }
  if (!g->scratch) {
    set_error("kifmm_gpu_direct_noalloc: no scratch address for this device");
    return KIFMM_EDEVICE;
  }
  HIP_TRY(hipSetDevice(g->dev));
  /* targets at d, d + nt, d + 2 nt, sums at d + 3 nt: inside 4 * MAX doubles;
   * the kernel assigns every sum, so nothing needs clearing */
  double* d = g->scratch;
  const size_t b = sizeof(double) * nt;
  HIP_TRY(hipMemcpyAsync(d, tx, b, hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipMemcpyAsync(d + nt, ty, b, hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipMemcpyAsync(d + 2 * (size_t)nt, tz, b, hipMemcpyHostToDevice, g->stream));
  hipLaunchKernelGGL(k_direct, dim3(std::min(nt, 65535)), dim3(256), 0, g->stream, nt, d, d + nt, d + 2 * (size_t)nt,
                     g->src4, g->n_src, d + 3 * (size_t)nt);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, d + 3 * (size_t)nt, b, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return KIFMM_OK;
}

double kifmm_problem_validate_gpu(kifmm_problem* p, kifmm_gpu* g, const double* pot_orig, int nv, int* chk_ids,
                                  double* direct_out) {
  if (!p || !g || !pot_orig || nv <= 0) {
    set_error("kifmm_problem_validate_gpu: invalid argument");
    return -1.0;
  }
  std::vector<int> ids(nv);
  std::vector<double> xyz(3 * (size_t)nv), direct(nv);
  if (kifmm::sample_targets(p, nv, ids.data(), xyz.data()) != KIFMM_OK) return -1.0;
  if (kifmm_gpu_direct(g, nv, xyz.data(), xyz.data() + nv, xyz.data() + 2 * (size_t)nv, direct.data()) != KIFMM_OK)
    return -1.0;
  double vn = 0.0, en = 0.0;
  for (int k = 0; k < nv; k++) {
    const double f = pot_orig[ids[k]];
    vn += f * f;
    en += (direct[k] - f) * (direct[k] - f);
  }
  if (chk_ids) std::memcpy(chk_ids, ids.data(), sizeof(int) * nv);
  if (direct_out) std::memcpy(direct_out, direct.data(), sizeof(double) * nv);
  return std::sqrt(en) / std::sqrt(vn);
}

int kifmm_gpu_plan(const kifmm_tree_view* tree, const kifmm_operators* ops, kifmm_work_counts* out) {
  if (!tree || !ops) {
    set_error("kifmm_gpu_plan: null argument");
    return KIFMM_EINVAL;
  }
  if (tree->np != 3 && tree->np != 4 && tree->np != 6) {
    set_error("kifmm_gpu_plan: NP must be 3, 4 or 6 on the GPU");
    return KIFMM_EUNSUP;
  }
  Plan P;
  int r = KIFMM_OK;
  try {
    r = make_plan(tree, ops, P);
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed in kifmm_gpu_plan");
    return KIFMM_ENOMEM;
  }
  if (r == KIFMM_OK && out) *out = P.wc;
  return r;
}

int kifmm_gpu_set_tuning(kifmm_gpu* g, const kifmm_gpu_tuning* t) {
  if (!g || !t) {
    set_error("kifmm_gpu_set_tuning: null argument");
    return KIFMM_EINVAL;
  }
  if (t->p2p_form < KIFMM_P2P_STATIC || t->p2p_form > KIFMM_P2P_SPLIT ||
      (t->hadamard_fb != 4 && t->hadamard_fb != 8 && t->hadamard_fb != 16) ||
      (t->hadamard_form != KIFMM_HAD_VALU && t->hadamard_form != KIFMM_HAD_MFMA) ||
      t->overlap < 0 || t->overlap > 6 || t->up_form < KIFMM_LEVEL_AUTO || t->up_form > KIFMM_LEVEL_FUSED ||
      t->down_form < KIFMM_LEVEL_AUTO || t->down_form > KIFMM_LEVEL_FUSED) {
    set_error("kifmm_gpu_set_tuning: p2p_form must be a KIFMM_P2P_* value, hadamard_fb 4, 8 or 16, "
              "hadamard_form a KIFMM_HAD_* value, overlap 0 to 6 and up_form / down_form KIFMM_LEVEL_* "
              "values");
    return KIFMM_EINVAL;
  }
  if (g->precision == KIFMM_PREC_MIXED && t->p2p_form != KIFMM_P2P_STATIC &&
      !(t->p2p_form == KIFMM_P2P_SPLIT && g->close_listed)) {
    set_error("kifmm_gpu_set_tuning: the mixed path has the STATIC and, with a close-pair list, the SPLIT "
              "U-list forms");
    return KIFMM_EUNSUP;
  }
  if (g->precision == KIFMM_PREC_FP64 && t->p2p_form == KIFMM_P2P_SPLIT) {
    set_error("kifmm_gpu_set_tuning: the SPLIT U-list form is the mixed path's");
    return KIFMM_EUNSUP;
  }
  if (t->p2p_form == g->p2p_form && t->hadamard_fb == g->had_fb && t->hadamard_form == g->had_form &&
      t->overlap == g->overlap && t->up_form == g->up_form && t->down_form == g->down_form)
    return KIFMM_OK;
  g->p2p_form = t->p2p_form;
  g->had_fb = t->hadamard_fb;
  g->had_form = t->hadamard_form;
  g->overlap = t->overlap;
  g->up_form = t->up_form;
  g->down_form = t->down_form;
  for (int k = 0; k < 2; k++) {
    if (g->gexec[k]) (void)hipGraphExecDestroy(g->gexec[k]);
    if (g->graph[k]) (void)hipGraphDestroy(g->graph[k]);
    g->gexec[k] = nullptr;
    g->graph[k] = nullptr;
  }
  return KIFMM_OK;
}

int kifmm_gpu_get_tuning(const kifmm_gpu* g, kifmm_gpu_tuning* t) {
  if (!g || !t) {
    set_error("kifmm_gpu_get_tuning: null argument");
    return KIFMM_EINVAL;
  }
  std::memset(t, 0, sizeof *t);
  t->p2p_form = g->p2p_form;
  t->hadamard_fb = g->had_fb;
  t->hadamard_form = g->had_form;
  t->overlap = g->overlap;
  t->up_form = g->up_form;
  t->down_form = g->down_form;
  return KIFMM_OK;
}

int kifmm_gpu_autotune(kifmm_gpu* g, int reps, kifmm_gpu_tuning* best, double* ms, int n_ms) {
  if (!g || reps < 1 || (ms && n_ms < 0)) {
    set_error("kifmm_gpu_autotune: null handle or reps < 1");
    return KIFMM_EINVAL;
  }
  double t[21];
  std::fill(t, t + 21, -1.0);
  kifmm_gpu_tuning cur{};
  kifmm_gpu_get_tuning(g, &cur);
  /* the fastest of reps timed runs after one warm-up: the U-list, the
   * Hadamard, or the up and the down phase (out[0], out[1]) */
  auto time_phase = [&](unsigned phase, double* out) -> int {
    kifmm_phase_times pt{};
    double m[2] = {0.0, 0.0};
    for (int r = 0; r <= reps; r++) {
      const int rc = evaluate(g, phase, nullptr, &pt);
      if (rc != KIFMM_OK) return rc;
      const double v[2] = {phase == KIFMM_PHASE_U ? pt.ulist : phase == KIFMM_PHASE_V ? pt.v_hadamard : pt.up,
                           pt.down};
      for (int k = 0; k < 2; k++)
        if (r == 1 || (r > 1 && v[k] < m[k])) m[k] = v[k]; /* run 0 warms up */
    }
    out[0] = m[0];
    if (phase == (KIFMM_PHASE_UP | KIFMM_PHASE_DOWN)) out[1] = m[1];
    return KIFMM_OK;
  };
  kifmm_gpu_tuning c = cur;
  /* the U-list forms of the handle's path, with their slot in t */
  std::vector<std::pair<int, int>> forms;
  if (g->precision == KIFMM_PREC_FP64)
    forms = {{KIFMM_P2P_STATIC, 0}, {KIFMM_P2P_PERSISTENT, 1}, {KIFMM_P2P_STATIC64, 2}};
  else if (g->close_listed)
    forms = {{KIFMM_P2P_STATIC, 0}, {KIFMM_P2P_SPLIT, 11}};
  else
    forms = {{KIFMM_P2P_STATIC, 0}};
  int best_form = KIFMM_P2P_STATIC, best_slot = 0;
  for (const auto& f : forms) {
    c.p2p_form = f.first;
    TRY(kifmm_gpu_set_tuning(g, &c));
    TRY(time_phase(KIFMM_PHASE_U, &t[f.second]));
    if (t[f.second] < t[best_slot]) {
      best_form = f.first;
      best_slot = f.second;
    }
  }
  c.p2p_form = best_form;
  const int fbs[3] = {4, 8, 16};
  int best_v = 0;
  for (int k = 0; k < 6; k++) { /* VALU fb 4, 8, 16, then MFMA fb 4, 8, 16 */
    c.hadamard_form = k / 3;
    c.hadamard_fb = fbs[k % 3];
    TRY(kifmm_gpu_set_tuning(g, &c));
    TRY(time_phase(KIFMM_PHASE_V, &t[3 + k]));
    if (t[3 + k] < t[3 + best_v]) best_v = k;
  }
  c.hadamard_form = best_v / 3;
  c.hadamard_fb = fbs[best_v % 3];
  /* up and down: the column GEMMs against the fused kernels, each phase on
   * its own (slots 14, 15: up; 16, 17: down) */
  for (int k = 0; k < 2; k++) {
    c.up_form = c.down_form = k == 0 ? KIFMM_LEVEL_GEMM : KIFMM_LEVEL_FUSED;
    TRY(kifmm_gpu_set_tuning(g, &c));
    double ud[2];
    TRY(time_phase(KIFMM_PHASE_UP | KIFMM_PHASE_DOWN, ud));
    t[14 + k] = ud[0];
    t[16 + k] = ud[1];
  }
  c.up_form = t[15] < t[14] ? KIFMM_LEVEL_FUSED : KIFMM_LEVEL_GEMM;
  c.down_form = t[17] < t[16] ? KIFMM_LEVEL_FUSED : KIFMM_LEVEL_GEMM;
  /* stream layout, on whole evaluations timed eagerly (a graph handle
   * would otherwise capture here) */
  const int use_graph = g->use_graph;
  g->use_graph = 0;
  /* slots of the layouts 0 .. 6 in t */
  const int lslot[7] = {9, 10, 12, 13, 18, 19, 20};
  int best_l = 0, rc = KIFMM_OK;
  for (int l = 0; l < 7 && rc == KIFMM_OK; l++) {
    c.overlap = l;
    rc = kifmm_gpu_set_tuning(g, &c);
    kifmm_phase_times pt{};
    for (int r = 0; r <= reps && rc == KIFMM_OK; r++) {
      rc = evaluate(g, KIFMM_PHASE_ALL, g->pinned, &pt); /* with the copy-out, as the layouts differ there */
      if (r == 1 || (r > 1 && pt.total < t[lslot[l]])) t[lslot[l]] = pt.total; /* run 0 warms up */
    }
    if (rc == KIFMM_OK && t[lslot[l]] < t[lslot[best_l]]) best_l = l;
  }
  g->use_graph = use_graph;
  TRY(rc);
  c.overlap = best_l;
  TRY(kifmm_gpu_set_tuning(g, &c));
  if (best) *best = c;
  if (ms)
    for (int k = 0; k < std::min(n_ms, 21); k++) ms[k] = t[k];
  return KIFMM_OK;
}

int kifmm_gpu_work_counts(const kifmm_gpu* g, kifmm_work_counts* out) {
  if (!g || !out) return KIFMM_EINVAL;
  *out = g->counts;
  out->close_pairs = g->close_listed ? (double)g->n_close : -1.0;
  out->copy_chunks = (double)std::max<size_t>(g->cuts.size(), 2) - 1.0;
  return KIFMM_OK;
}

double* kifmm_gpu_output(kifmm_gpu* g) { return g ? g->pinned : nullptr; }

int kifmm_gpu_set_densities(kifmm_gpu* g, const double* den_leaf) {
  if (!g || (!den_leaf && g->n_src > 0)) {
    set_error("kifmm_gpu_set_densities: invalid argument");
    return KIFMM_EINVAL;
  }
  if (g->n_src == 0) return KIFMM_OK;
  HIP_TRY(hipSetDevice(g->dev));
  HIP_TRY(hipMemcpyAsync(g->den_stage, den_leaf, sizeof(double) * (size_t)g->n_src, hipMemcpyHostToDevice,
                         g->stream));
  const int blocks = (int)std::min<size_t>(4096, ((size_t)g->n_src + 255) / 256);
  hipLaunchKernelGGL(k_set_w, dim3(blocks), dim3(256), 0, g->stream, g->src4, g->den_stage, (size_t)g->n_src);
  HIP_TRY(hipGetLastError());
  if (g->precision == KIFMM_PREC_MIXED) TRY(mixed_sources(g));
  HIP_TRY(hipStreamSynchronize(g->stream)); /* den_leaf may be released on return */
  return KIFMM_OK;
}

int kifmm_gpu_create_times(const kifmm_gpu* g, double* ms, int n) {
  if (!g || !ms || n < 0) {
    set_error("kifmm_gpu_create_times: invalid argument");
    return KIFMM_EINVAL;
  }
  for (int k = 0; k < std::min(n, 5); k++) ms[k] = g->create_ms[k];
  return KIFMM_OK;
}

long kifmm_gpu_close_list(const kifmm_gpu* g, int* ptr, long n_ptr, int* idx, long n_idx) {
  if (!g || n_ptr < 0 || n_idx < 0 || (n_ptr && !ptr) || (n_idx && !idx)) {
    set_error("kifmm_gpu_close_list: invalid argument");
    return KIFMM_EINVAL;
  }
  if (!g->close_listed) { /* KIFMM_GPU_NO_CLOSE_LIST, or too many pairs */
    set_error("kifmm_gpu_close_list: this handle has no close-pair list");
    return KIFMM_EUNSUP;
  }
  if (n_ptr > (long)g->n_trg + 1 || n_idx > g->n_close) {
    set_error("kifmm_gpu_close_list: count exceeds the list");
    return KIFMM_EINVAL;
  }
  if (hipSetDevice(g->dev) != hipSuccess || hipStreamSynchronize(g->stream) != hipSuccess ||
      (n_ptr && hipMemcpy(ptr, g->cptr, sizeof(int) * n_ptr, hipMemcpyDeviceToHost) != hipSuccess) ||
      (n_idx && hipMemcpy(idx, g->cidx, sizeof(int) * n_idx, hipMemcpyDeviceToHost) != hipSuccess)) {
    set_error("kifmm_gpu_close_list: copy failed");
    return KIFMM_EDEVICE;
  }
  return (long)g->n_close;
}

int kifmm_gpu_copy_intermediate(const kifmm_gpu* g, int which, double* out, long count) {
  if (!g || !out || count < 0) {
    set_error("kifmm_gpu_copy_intermediate: invalid argument");
    return KIFMM_EINVAL;
  }
  const double* src = nullptr;
  long avail = 0;
  switch (which) {
    case 0: src = g->uden; avail = (long)g->n_nodes * g->ld_e; break;
    case 1: src = g->dchk; avail = (long)g->n_nodes * g->ld_e; break;
    case 2: src = g->dden; avail = (long)g->n_nodes * g->ld_e; break;
    case 3: src = g->pot; avail = g->n_trg; break;
    default: set_error("kifmm_gpu_copy_intermediate: which must be 0..3"); return KIFMM_EINVAL;
  }
  if (count > avail) {
    set_error("kifmm_gpu_copy_intermediate: count exceeds the array");
    return KIFMM_EINVAL;
  }
  if (hipSetDevice(g->dev) != hipSuccess || hipStreamSynchronize(g->stream) != hipSuccess ||
      hipMemcpy(out, src, sizeof(double) * count, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("kifmm_gpu_copy_intermediate: copy failed");
    return KIFMM_EDEVICE;
  }
  return KIFMM_OK;
}

}  // extern "C"
