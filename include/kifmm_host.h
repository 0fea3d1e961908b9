/*
 * kifmm_host.h -- host side of the MI355X KIFMM evaluator (C ABI).
 *
 * These entry points replace the reference's CPU-side setup that feeds the
 * GPU evaluation path:
 *   kifmm_problem_create ..... create() + densities + src_tree() + trg_tree()
 *                              + get_input()   (fmm/input.cc:29-84,203-460,
 *                              fmm/driver.cc:244-299, fmm/evaluate.cc:57-96)
 *   kifmm_problem_operators .. compute_UC2UE_mat / UE2UC / DC2DE / DE2DC
 *                              (fmm/trans.cc:143-264; pinv fmm/vecmatop.cc:34-78)
 *   kifmm_problem_scatter .... copy_trg_val's scatter to input order
 *                              (fmm/omp_sse_block.cc:873-887, set_value
 *                              fmm/evaluate.cc:192-202)
 *   kifmm_problem_validate ... validate() (fmm/driver.cc:141-206)
 *   kifmm_getenv_* ........... getenv__accuracy / __validate / __coincide /
 *                              __block_size / __upbs (fmm/util.h:8-15)
 *
 * Coordinates live in the reference root box [-1,1]^3.  Error convention:
 * functions return 0 (or a non-negative count) on success and a negative
 * code on failure; kifmm_last_error() describes the last failure on the
 * calling thread.  Nothing here exits the process.
 */
#ifndef KIFMM_HOST_H
#define KIFMM_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

#define KIFMM_SRCNODE 1 /* LET_SRCNODE, fmm/input.h:10-13 */
#define KIFMM_TRGNODE 2 /* LET_TRGNODE */

enum {
  KIFMM_OK = 0,
  KIFMM_EINVAL = -1,  /* bad argument */
  KIFMM_ENOMEM = -2,  /* host or device allocation failed */
  KIFMM_EDEVICE = -3, /* HIP runtime error */
  KIFMM_EUNSUP = -4   /* unsupported configuration (e.g. NP not in {3,4,6} on the GPU) */
};

typedef struct kifmm_problem kifmm_problem;

/* Flattened, read-only view of the octree, interaction lists and gathered
 * points: everything the reference's preproc() reads from AllNodes
 * (fmm/node.h:28-78, fmm/input.h:17-51).  Node order is the reference's
 * BFS level order with the 8 children of a node contiguous, octant index
 * 4*x + 2*y + z.  Lists are CSR over all nodes, ascending node ids. */
typedef struct kifmm_tree_view {
  int np;                 /* NP accuracy parameter */
  int n_nodes;
  int n_levels;           /* entries in level_start (= max depth + 2) */
  const int* level_start; /* nodeLevel: first node of each depth, last = n_nodes */
  const int* parent;      /* -1 for the root */
  const int* child;       /* first child, -1 for a leaf */
  const int* depth;
  const int* tag;         /* KIFMM_SRCNODE | KIFMM_TRGNODE */
  const int* path;        /* 3 ints per node: integer box coordinates at its depth */
  const int* src_beg;     /* leaf source range in the gathered arrays (0 for non-leaves) */
  const int* src_num;
  const int* trg_beg;     /* leaf target range */
  const int* trg_num;
  const int* u_ptr; const int* u_idx; /* U-list */
  const int* v_ptr; const int* v_idx; /* V-list */
  const int* w_ptr; const int* w_idx; /* W-list */
  const int* x_ptr; const int* x_idx; /* X-list */
  int n_src, n_trg;
  const double* sx; const double* sy; const double* sz; const double* sden; /* leaf order */
  const double* tx; const double* ty; const double* tz;                     /* leaf order */
  int coincident;         /* 1 when targets are the sources (r == 0 pairs occur) */
} kifmm_tree_view;

/* Translation operators for the unit box (fmm/trans.cc:143-264), FP64,
 * column-major.  n_e = np^3 - (np-2)^3, n_uc = (np+2)^3 - np^3. */
typedef struct kifmm_operators {
  int np, n_e, n_uc;
  const double* uc2ue; /* [n_e  x n_uc]      pseudo-inverse, cutoff 0.1^(np+1) */
  const double* ue2uc; /* 8 x [n_uc x n_e]   octant o at offset o*n_uc*n_e     */
  const double* dc2de; /* [n_e  x n_e]       pseudo-inverse                     */
  const double* de2dc; /* 8 x [n_e  x n_e]                                     */
} kifmm_operators;

typedef struct kifmm_problem_opts {
  int np;       /* <= 0: NP environment variable, default 6 */
  long seed;    /* < 0:  SEED environment variable, default 26 (fmm/driver.cc:244) */
  int coincide; /* < 0:  C environment variable, default 0 */
  int nthreads; /* <= 0: all host threads */
} kifmm_problem_opts;

/* Generate points ("uniform", "ellipseUniformAngles" as the reference;
 * "ellipsoid": semi-axes (1,1,0.25) about 0, i.e. the [0,1]^3 ellipsoid
 * (0.5,0.5,0.125) mapped by x' = 2x-1; "gaussmix": 8 Gaussian clusters,
 * sigma 0.04 after mapping), build the trees and lists, gather. */
kifmm_problem* kifmm_problem_create(int n, const char* distribution, int pts_max,
                                    const kifmm_problem_opts* opts);
/* Same from caller points (original order, root box [-1,1]^3).  When
 * tx == sx && ty == sy && tz == sz the sets are marked coincident. */
kifmm_problem* kifmm_problem_from_points(int n, const double* sx, const double* sy,
                                         const double* sz, const double* den,
                                         const double* tx, const double* ty,
                                         const double* tz, int pts_max, int np);
void kifmm_problem_destroy(kifmm_problem* p);

/* Point files (no counterpart in the reference, which only generates its
 * inputs, fmm/driver.cc:264-267): 8-byte magic "KIFMMPT1", int64 n, int32
 * has_targets, int32 0, then n little-endian doubles of sx, sy, sz, den and,
 * with targets, tx, ty, tz; coordinates in [-1,1]^3.  tx = ty = tz = NULL
 * writes a sources-are-targets file. */
int kifmm_points_write(const char* path, int n, const double* sx, const double* sy, const double* sz,
                       const double* den, const double* tx, const double* ty, const double* tz);
kifmm_problem* kifmm_problem_from_file(const char* path, int pts_max, int np);

int kifmm_problem_view(const kifmm_problem* p, kifmm_tree_view* out);
/* Builds (once) and returns the operators; the arrays live in the problem. */
int kifmm_problem_operators(kifmm_problem* p, kifmm_operators* out);
/* pot_leaf (gathered target order) -> pot_orig (input order) */
int kifmm_problem_scatter(const kifmm_problem* p, const double* pot_leaf, double* pot_orig);
/* Original-order points; any pointer may be NULL. */
int kifmm_problem_points(const kifmm_problem* p, double* sx, double* sy, double* sz,
                         double* den, double* tx, double* ty, double* tz);
/* Sampled relative L2 error of pot_orig against the FP64 direct sum at nv
 * targets drawn by continuing the generator stream; r == 0 excluded.
 * chk_ids / direct_out may be NULL.  Returns a negative value on error. */
double kifmm_problem_validate(kifmm_problem* p, const double* pot_orig, int nv,
                              int* chk_ids, double* direct_out);
/* Seconds spent in generation, tree+lists, operators (last call). */
int kifmm_problem_setup_times(const kifmm_problem* p, double* times3);

/* Environment surface (fmm/util.c:53-222); SEED is new (default 26). */
int kifmm_getenv_accuracy(void);   /* NP,   default 6    */
int kifmm_getenv_validate(void);   /* NV,   default 1000 */
int kifmm_getenv_coincide(void);   /* C,    default 0    */
int kifmm_getenv_block_size(void); /* BS,   default 49   */
int kifmm_getenv_upbs(void);       /* UPBS, default 1000 */
long kifmm_getenv_seed(void);      /* SEED, default 26   */

/* Operator building blocks, exposed for tests. */
int kifmm_surface(int np, double radius, double* xyz); /* 3*n, returns n */
int kifmm_pinv(const double* M, int m, int n, double eps, double* R);
int kifmm_drand48(long seed, int k, double* out);

const char* kifmm_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
