/*
 * kifmm_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a CPU restatement of the
 * reference KIFMM CPU path (fmm/omp_sse_block.cc, fmm/input.cc,
 * fmm/trans.cc, fmm/driver.cc).  Only tests/, the smoke check and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.
 * Nothing in the product library links, loads or calls it.
 *
 * Kernel modes (the reference CPU kernel is not exact FP64):
 *   ORACLE_KERNEL_EXACT = 0  w / sqrt(r^2), FP64 (fmm/naive.cc:179-185)
 *   ORACLE_KERNEL_REF   = 1  SSE rsqrtps + one FP64 Newton step
 *                            (fmm/omp_sse_block.cc:286-296, reals_aligned.h:61)
 * Both skip r == 0 pairs (fmm/omp_sse.cc:294-295 convention).
 */
#ifndef KIFMM_ORACLE_H
#define KIFMM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_problem oracle_problem;

enum { ORACLE_KERNEL_EXACT = 0, ORACLE_KERNEL_REF = 1 };
enum { ORACLE_LIST_U = 0, ORACLE_LIST_V = 1, ORACLE_LIST_W = 2, ORACLE_LIST_X = 3 };

/* create + densities + src_tree + trg_tree + lists + get_input
 * (fmm/driver.cc:244-299).  dist: "uniform", "ellipseUniformAngles",
 * "ellipsoid" (semi-axes 1,1,0.25 about 0: the [0,1]^3 ellipsoid
 * (0.5,0.5,0.125) mapped by x'=2x-1), "gaussmix". */
oracle_problem* oracle_create(int n, const char* dist, int pts_max, int np,
                              long seed, int coincide);
/* same, from caller-supplied points in the reference root box [-1,1]^3 */
oracle_problem* oracle_create_from_points(int n, const double* sx, const double* sy,
                                          const double* sz, const double* den,
                                          const double* tx, const double* ty,
                                          const double* tz, int pts_max, int np);
void oracle_destroy(oracle_problem* p);

int oracle_num_nodes(const oracle_problem* p);
int oracle_num_levels(const oracle_problem* p); /* nodeLevel.size() */
int oracle_level_starts(const oracle_problem* p, int* out);
int oracle_tree_arrays(const oracle_problem* p, int* parent, int* child, int* depth,
                       int* tag, int* path3, int* src_beg, int* src_num,
                       int* trg_beg, int* trg_num);
/* CSR of list `which` over all nodes: ptr[n_nodes+1], idx[ptr[n_nodes]] */
int oracle_list_csr(const oracle_problem* p, int which, int* ptr, int* idx);
int oracle_list_nnz(const oracle_problem* p, int which);
/* original-order points (what create() drew) */
int oracle_points(const oracle_problem* p, double* sx, double* sy, double* sz,
                  double* den, double* tx, double* ty, double* tz);
/* leaf-ordered (gathered) point index maps: src_perm[k] = original index */
int oracle_perms(const oracle_problem* p, int* src_perm, int* trg_perm);

/* Operators (fmm/trans.cc:143-264).  Sizes: n_e = np^3-(np-2)^3,
 * n_uc = (np+2)^3-np^3.  Column-major:
 *   uc2ue [n_e x n_uc], ue2uc 8 x [n_uc x n_e], dc2de [n_e x n_e],
 *   de2dc 8 x [n_e x n_e]; octant index = z + 2y + 4x. */
int oracle_build_operators(oracle_problem* p, int kernel_mode);
int oracle_set_operators(oracle_problem* p, const double* uc2ue, const double* ue2uc,
                         const double* dc2de, const double* de2dc);
int oracle_get_operators(const oracle_problem* p, double* uc2ue, double* ue2uc,
                         double* dc2de, double* de2dc);

/* Full evaluation (fmm/evaluate--basic.cc:30-87).  pot_orig[n] in the
 * original target order.  times[8] = up,U,V,W,X,down,copy,total seconds. */
int oracle_evaluate(oracle_problem* p, int kernel_mode, int nthreads,
                    double* pot_orig, double* times);
/* Phase-by-phase access for tests: leaf-order intermediate arrays.
 * uden [n_nodes*n_e], dchk [n_nodes*n_e], dden [n_nodes*n_e]. */
int oracle_intermediates(const oracle_problem* p, double* uden, double* dchk,
                         double* dden);

/* validate() (fmm/driver.cc:141-206): NV targets drawn by continuing the
 * generator stream; exact FP64 direct sum with r == 0 excluded and
 * zeroed accumulators.  chk_ids[nv] receives the sampled target ids. */
double oracle_validate(oracle_problem* p, const double* pot_orig, int nv, int* chk_ids,
                       double* direct_out);

/* primitives, exposed for pinning against closed forms and numpy */
int oracle_surface(int np, double R, double* xyz);        /* 3*n, x then y then z */
int oracle_regpos(int np, double R, double* xyz);         /* 3*(2np)^3 */
int oracle_pinv(const double* M, int m, int n, double eps, double* R);
int oracle_svd_values(const double* M, int m, int n, double* s);
int oracle_rfft3(int N, const double* in, double* out);   /* [z][y][x] -> [z][y][x/2+1] cplx */
int oracle_irfft3(int N, const double* in, double* out);  /* unnormalised c2r */
int oracle_translation_spectra(int np, int kernel_mode, double* out); /* 343 x eff */
int oracle_direct(int nt, const double* tx, const double* ty, const double* tz,
                  int ns, const double* sx, const double* sy, const double* sz,
                  const double* w, int kernel_mode, double* pot);
int oracle_drand48(long seed, int k, double* out);
/* Reverse every exact-kernel source sum (a summation-order perturbation
 * for measuring the conditioning floor); process-wide, default off. */
void oracle_set_reverse_order(int on);

#ifdef __cplusplus
}
#endif
#endif
