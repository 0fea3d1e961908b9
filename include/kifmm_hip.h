/*
 * kifmm_hip.h -- GPU evaluator of the MI355X KIFMM path (C ABI).
 *
 * The reference plugs an evaluator in at link time through
 *   FMMWrapper_t* preproc (AllNodes*)   fmm/evaluate.h:18, body fmm/evaluate--cuda.cu:1441-2239
 *   int run (FMMWrapper_t*)             fmm/evaluate.h:23, body fmm/run--cuda.cu:27-345
 * kifmm_gpu_create replaces preproc (flatten + upload, library-owned device
 * memory); kifmm_gpu_evaluate replaces run (up, U, V, W, X, down, copy;
 * phase timing as fmm/run--cuda.cu:323-340 with the copy actually timed);
 * kifmm_gpu_destroy frees what create allocated (the reference never frees).
 *
 * All entry points return 0 on success or a negative KIFMM_E* code, never
 * exit.  A handle is not thread-safe; several handles may coexist.  Each
 * handle owns three HIP streams at default priority on its device: the
 * main one, a second compute stream and a copy stream.
 */
#ifndef KIFMM_HIP_H
#define KIFMM_HIP_H

#include "kifmm_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kifmm_gpu kifmm_gpu;

enum {
  KIFMM_PREC_FP64 = 0, /* plain FP64 path */
  KIFMM_PREC_MIXED = 1 /* NP <= 4: U-list distances in FP32 Gram form on MFMA, FP32 rsqrt,
                          FP64 terms for pairs closer than 1/8 of the target leaf's
                          half-width, FP64 accumulation; FP32 S2M / L2T and V-list
                          spectra; FP64 pseudo-inverse applies, M2M / L2L and W / X */
};

enum {
  KIFMM_GPU_GRAPH = 1, /* capture the far-field launch sequence (fills, up, V, W, X, down)
                          in one HIP graph at the first full evaluate and replay it
                          afterwards; the U-list chunks and their copies back, which run
                          beside one another on three streams, stay outside the graph
                          (phase times then report the U-list, the copy and the total) */
  KIFMM_GPU_NO_CLOSE_LIST = 2, /* do not list the U-list pairs closer than 1/8 of the target
                                  leaf's half-width at create time: the mixed path then tests
                                  for them inside its loop, the FP64 path forms r^2 from
                                  coordinate differences (11 instead of 9 instructions per
                                  pair).  The list is also skipped by itself when it would
                                  exceed 16 pairs per target on average. */
  KIFMM_GPU_MIXED_INLINE = KIFMM_GPU_NO_CLOSE_LIST, /* the earlier name of the same flag */
  KIFMM_GPU_OVERLAP = 4 /* start with the U-list on the handle's second stream (default
                           priority) alongside up, V, W, X and down, into a buffer of its own
                           that one pass adds to the far field at the end
                           (kifmm_gpu_tuning.overlap; the autotuner may change it).  The
                           potentials are bit-identical to the default layout, but the copy
                           back can no longer overlap the U-list.  Phase times then overlap:
                           ulist is the U-list's own span. */
};

/* Launch shapes of a handle, the counterpart of the reference's hand-set
 * blocking knobs BS and UPBS (fmm/util.c:118-163, read by the up pass at
 * fmm/omp_sse_block.cc:203): chosen per input by kifmm_gpu_autotune. */
enum {
  KIFMM_P2P_STATIC = 0,     /* one block per U-list item of <= 128 targets (default) */
  KIFMM_P2P_PERSISTENT = 1, /* a few blocks per CU drawing the same items from per-XCD queues */
  KIFMM_P2P_STATIC64 = 2,   /* one block per item of <= 64 targets */
  KIFMM_P2P_SPLIT = 3       /* mixed path only: items of <= 128 targets, up to 2 groups of 16
                               targets per wave, each source tile split among the waves that
                               share a set of groups (pays on small leaves) */
};
typedef struct kifmm_gpu_tuning {
  int p2p_form;      /* KIFMM_P2P_*: STATIC, PERSISTENT or STATIC64 on the FP64 path;
                        STATIC or SPLIT on the mixed path */
  int hadamard_fb;   /* frequencies per V-list Hadamard block: 4, 8 (default) or 16 */
  int hadamard_form; /* KIFMM_HAD_*: V-list Hadamard on MFMA (default) or on the VALU */
  int overlap;       /* stream layout of the U-list's copy-out chunks: 0 all of them last,
                        after the far field, on the two streams in turn, their copies back
                        overlapping the later chunks; 1 all of them on the handle's second
                        stream from the start, alongside the far field (initially 1 with
                        KIFMM_GPU_OVERLAP); 2, 3 the first one or two chunks there, the
                        rest last; 4 all of them last, one after another on the main stream
                        (the first copies start sooner: pays when the copy back is long
                        against the U-list); 5, 6 as 2, 3 with the rest one after another */
  int up_form;       /* KIFMM_LEVEL_*: how each level's M2M (UE2UC stack, then UC2UE) runs */
  int down_form;     /* KIFMM_LEVEL_*: how each level's L2L + DC2DE runs */
  int reserved[2];
} kifmm_gpu_tuning;
enum {
  KIFMM_LEVEL_AUTO = 0,  /* default: the fused kernel for L2L + DC2DE, and for M2M at NP <= 4
                            (column GEMMs for M2M at NP = 6): the faster forms measured on
                            uniform and clustered inputs */
  KIFMM_LEVEL_GEMM = 1,  /* two column GEMMs per level (K split over blocks when the level has
                            few columns, then a reduction launch) */
  KIFMM_LEVEL_FUSED = 2  /* one kernel per level: both operator applies of 16 columns per
                            block, the intermediate vectors kept in LDS */
};
enum {
  KIFMM_HAD_VALU = 0, /* a lane per parent, two waves per 64 parents */
  KIFMM_HAD_MFMA = 1  /* 16 parents per 16x16x4 MFMA tile, four waves per 64 parents (default) */
};

typedef struct kifmm_gpu_options {
  int device;    /* HIP device ordinal */
  int precision; /* KIFMM_PREC_* */
  int flags;     /* KIFMM_GPU_* */
  int reserved[5];
} kifmm_gpu_options;

/* Milliseconds, measured with HIP events on the handle's stream.  The
 * phases run in the order up, V, W, X, down, U-list. */
typedef struct kifmm_phase_times {
  double up;        /* S2M + M2M                          */
  double ulist;     /* U-list P2P (the last phase)        */
  double vlist;     /* sum of the four V sub-phases       */
  double v_fft_src; /* surface -> grid scatter + forward DFT */
  double v_trans;   /* translation spectra (316 vectors): formed once at
                       create time, so ~0 here               */
  double v_hadamard;/* spectral multiply-accumulate       */
  double v_ifft;    /* inverse DFT + check-surface gather */
  double wlist;
  double xlist;
  double down;      /* L2L + DC2DE + L2T                  */
  double copy;      /* device -> host potentials: the part left after the last
                       kernel (the copies of earlier U-list chunks overlap the later
                       chunks), plus the host pass into a caller buffer that is not
                       kifmm_gpu_output()'s */
  double total;     /* first event to last event (+ that host pass) */
  double p2p_kernel;/* the U-list kernel alone            */
  double reserved[3];
} kifmm_phase_times;

/* Algorithmic work counts of the uploaded problem (for roofline figures). */
typedef struct kifmm_work_counts {
  double u_pairs;      /* sum over U entries of trg_num * src_num  */
  double s2m_pairs;    /* n_uc * src_num over source leaves        */
  double l2t_pairs;    /* n_e * trg_num over target leaves         */
  double w_pairs, x_pairs;
  double v_entries;    /* V-list nnz                                */
  double v_cmac;       /* complex multiply-adds in the Hadamard kernel as launched */
  double v_cmac_ref;   /* (eff/2) * V nnz, the reference count      */
  double gemm_flops;   /* 2*m*k per applied operator column         */
  double fft_bytes;    /* HBM bytes of the forward + inverse DFT kernels */
  double hadamard_bytes; /* compulsory HBM bytes of the Hadamard kernel */
  double close_pairs;  /* U-list pairs closer than 1/8 of the target leaf's half-width,
                          listed at create time; -1 when not listed (KIFMM_GPU_NO_CLOSE_LIST
                          or too many) */
  double copy_chunks;  /* U-list chunks whose copies back overlap the later chunks */
  double pair_items;   /* sibling-leaf pairs among the mixed U-list's work items */
  double reserved[2];
} kifmm_work_counts;

int kifmm_gpu_create(const kifmm_tree_view* tree, const kifmm_operators* ops,
                     const kifmm_gpu_options* opts, kifmm_gpu** out);
/* Evaluate the potentials.  pot_leaf (host, n_trg doubles, gathered target
 * order) may be NULL: the result then stays in device memory and the copy
 * phase is skipped.  The copy goes to the handle's pinned host buffer,
 * chunk by chunk while the U-list (the last phase) still computes; with
 * pot_leaf == kifmm_gpu_output(g) that is all, any other pot_leaf gets one
 * host pass from it (pipelined chunk by chunk in eager mode).  times may be
 * NULL. */
int kifmm_gpu_evaluate(kifmm_gpu* g, double* pot_leaf, kifmm_phase_times* times);
int kifmm_gpu_destroy(kifmm_gpu* g);

/* The handle's pinned host buffer of n_trg potentials (gathered target
 * order), valid until kifmm_gpu_destroy; NULL if pinned memory could not be
 * allocated.  Passing it to kifmm_gpu_evaluate copies the result at PCIe
 * rate with no host pass: the counterpart of the host array that
 * copy_trg_val__gpu fills (fmm/cuda.cu:2543-2565).  Each evaluation
 * overwrites it. */
double* kifmm_gpu_output(kifmm_gpu* g);

/* New source densities (n_src doubles, gathered source order) for the
 * handle's fixed geometry, for repeated evaluations with changing
 * densities (in the reference a new AllNodes and preproc).  Everything
 * create derives ahead of time (lists, close pairs, Gram records,
 * translation spectra) depends on the points only; evaluations read the
 * densities afresh.  Blocks until the upload is done. */
int kifmm_gpu_set_densities(kifmm_gpu* g, const double* den_leaf);

/* Wall milliseconds of the handle's kifmm_gpu_create: [0] the host plan
 * (work lists and their bounds checks), [1] the upload, [2] the close-pair
 * search, [3] the other device set-up (translation spectra, Gram records),
 * [4] the whole call.  n entries (up to 5) are written. */
int kifmm_gpu_create_times(const kifmm_gpu* g, double* ms, int n);

/* The close-pair list for tests: n_ptr <= n_trg + 1 offsets and n_idx
 * source indices (gathered order), per target in U-list order.  Returns
 * the number of listed pairs, or KIFMM_EUNSUP when the handle has none. */
long kifmm_gpu_close_list(const kifmm_gpu* g, int* ptr, long n_ptr, int* idx, long n_idx);

/* Phase selection, the counterpart of the reference's separately callable
 * phase functions up_calc__gpu, ulist_calc__gpu, vlist_calc__gpu,
 * wlist_calc__gpu, xlist_calc__gpu, down_calc__gpu (fmm/evaluate.h:42-57).
 * Accumulators are zeroed first; a phase whose inputs were not produced in
 * the same call sees zeros. */
enum {
  KIFMM_PHASE_UP = 1,
  KIFMM_PHASE_U = 2,
  KIFMM_PHASE_V = 4,
  KIFMM_PHASE_W = 8,
  KIFMM_PHASE_X = 16,
  KIFMM_PHASE_DOWN = 32,
  KIFMM_PHASE_ALL = 63
};
int kifmm_gpu_evaluate_phases(kifmm_gpu* g, unsigned phases, double* pot_leaf, kifmm_phase_times* times);

int kifmm_gpu_work_counts(const kifmm_gpu* g, kifmm_work_counts* out);
/* Derives and bounds-checks the device work lists of a tree view on the
 * host only (no device call) -- the checks kifmm_gpu_create runs before
 * it uploads anything.  out may be NULL. */
int kifmm_gpu_plan(const kifmm_tree_view* tree, const kifmm_operators* ops, kifmm_work_counts* out);
/* Test access to device intermediates after an evaluate, node-major with
 * n_e doubles per node: which = 0 upward equivalent densities,
 * 1 downward check values, 2 downward equivalent densities, 3 potentials
 * (n_trg doubles, gathered order). */
int kifmm_gpu_copy_intermediate(const kifmm_gpu* g, int which, double* out, long count);

/* Set or read the launch shapes and the stream layout.  The U-list forms,
 * the Hadamard block sizes and the layouts give bit-identical potentials;
 * the two Hadamard forms sum in different orders, and so do the two
 * mixed U-list forms (rounding-level differences).  Setting drops a
 * captured graph (the next full evaluate captures again).  EINVAL for an
 * unknown value, EUNSUP for a U-list form the handle's path does not have
 * (PERSISTENT / STATIC64 on the mixed path, SPLIT on the FP64 path or
 * without a close-pair list). */
int kifmm_gpu_set_tuning(kifmm_gpu* g, const kifmm_gpu_tuning* t);
int kifmm_gpu_get_tuning(const kifmm_gpu* g, kifmm_gpu_tuning* t);
/* Time every U-list form of the handle's path on the U phase, every
 * Hadamard form and block size on the V phase, both level forms on the up
 * and on the down phase, then the stream layouts on whole (eager)
 * evaluations with the copy back of this handle's own input, reps timed
 * runs each after one warm-up, keep the fastest of each (minimum over runs)
 * and set it.  best (may be NULL) receives the choice; ms (may be NULL, n_ms
 * entries, up to 21) receives the times in the order STATIC, PERSISTENT,
 * STATIC64, VALU fb 4, 8, 16, MFMA fb 4, 8, 16, layout 0, layout 1, SPLIT,
 * layout 2, layout 3, up as column GEMMs, up fused, down as column GEMMs,
 * down fused, layouts 4, 5, 6; -1 for a candidate not timed. */
int kifmm_gpu_autotune(kifmm_gpu* g, int reps, kifmm_gpu_tuning* best, double* ms, int n_ms);

/* 1 if a HIP device is usable from this process. */
int kifmm_gpu_available(void);

/* Direct FP64 sum phi(t) = 1/(4 pi) sum_s w_s / |t - s| (r = 0 excluded) at
 * nt host-given points over all the handle's sources, on its device, with
 * compensated summation.  Replaces the O(NV N) host loop of validate
 * (fmm/driver.cc:141-206). */
int kifmm_gpu_direct(kifmm_gpu* g, int nt, const double* tx, const double* ty, const double* tz, double* out);

/* kifmm_gpu_direct without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid it).  No counterpart in the reference.  The targets and sums
 * are staged in a file-scope device array of the library holding at most
 * KIFMM_DIRECT_NOALLOC_MAX points (512 KiB per device, kept until the
 * process ends); larger nt is rejected with KIFMM_EINVAL.  Every call on
 * one device shares that array: calls that overlap on one device (other
 * threads, other handles) race.  Blocks until the sums are in out. */
#define KIFMM_DIRECT_NOALLOC_MAX 16384
int kifmm_gpu_direct_noalloc(kifmm_gpu* g, int nt, const double* tx, const double* ty, const double* tz,
                             double* out);

/* kifmm_problem_validate with the direct sums on the GPU: the same sample
 * (the problem's generator stream continued) and the same relative L2
 * error; g must hold p's sources.  Returns the error, or -1 on failure. */
double kifmm_problem_validate_gpu(kifmm_problem* p, kifmm_gpu* g, const double* pot_orig, int nv, int* chk_ids,
                                  double* direct_out);

/* kifmm_problem_create / kifmm_problem_from_points with the octree, the
 * U/V/W/X lists and the gather permutations built on GPU `device`.  The
 * result is identical, array for array, to the host build, which follows
 * the preprocessing of fmm/input.cc:203-460 (preproc, fmm/evaluate.h:18-20);
 * only kifmm_problem_setup_times' tree time differs.  Fails (NULL,
 * kifmm_last_error) on invalid input, on a device error, or when the tree
 * would need more than 30 levels (coincident points beyond pts_max). */
kifmm_problem* kifmm_problem_create_gpu(int device, int n, const char* distribution, int pts_max,
                                        const kifmm_problem_opts* opts);
kifmm_problem* kifmm_problem_from_points_gpu(int device, int n, const double* sx, const double* sy, const double* sz,
                                             const double* den, const double* tx, const double* ty, const double* tz,
                                             int pts_max, int np);

#ifdef __cplusplus
}
#endif
#endif
