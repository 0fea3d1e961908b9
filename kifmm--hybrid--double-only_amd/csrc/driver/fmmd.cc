/*
 * fmmd.cc -- the fmmd--hip command line, a drop-in for fmmd--cuda.
 *
 * Same surface as the reference driver (fmm/driver.cc:211-373):
 *   fmmd--hip <#pts> {uniform,ellipseUniformAngles,ellipsoid,gaussmix} <pts-per-box>
 * environment NP (3, 4, 6), NV (validation samples), C (sources =
 * targets), SEED (default 26, the reference's srand48(26)).
 * Additions: distribution "file:<path>" reads a KIFMMPT1 point file (the
 * count argument is then ignored); PREC=mixed selects the mixed-precision
 * path (NP <= 4); GRAPH=1 replays the evaluation as one HIP graph;
 * OVERLAP=1 runs the U-list on a second stream alongside the other phases.
 * Steps: generate + trees + lists + gather, operators, kifmm_gpu_create
 * (preproc), one untimed evaluation (loads the kernels), kifmm_gpu_evaluate
 * (run, with the copy back timed),
 * scatter to input order, validate, print the reference's estimates.
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "kifmm_hip.h"

static int tv_count(kifmm_problem* p) {
  kifmm_tree_view v;
  return kifmm_problem_view(p, &v) == KIFMM_OK ? v.n_src : 0;
}

static void usage(const char* prog) { fprintf(stderr, "usage: %s <N> <distribution> <pts/box> \n", prog); }

int main(int argc, char** argv) {
  if (argc != 4) {
    usage(argv[0]);
    return -1;
  }
  int n = atoi(argv[1]);
  const char* dist = argv[2];
  const int q = atoi(argv[3]);
  kifmm_problem_opts po = {-1, -1, -1, 0};
  fprintf(stderr, "Building trees and interaction lists...\n");
  kifmm_problem* p = strncmp(dist, "file:", 5) == 0 ? kifmm_problem_from_file(dist + 5, q, kifmm_getenv_accuracy())
                                                    : kifmm_problem_create(n, dist, q, &po);
  if (!p) {
    fprintf(stderr, "error: %s\n", kifmm_last_error());
    return 1;
  }
  kifmm_tree_view tv;
  kifmm_operators ops;
  n = tv_count(p);
  fprintf(stderr, "Setup for translations....\n");
  if (kifmm_problem_view(p, &tv) || kifmm_problem_operators(p, &ops)) {
    fprintf(stderr, "error: %s\n", kifmm_last_error());
    return 1;
  }
  fprintf(stderr, "Preprocessing...\n");
  kifmm_gpu_options go;
  memset(&go, 0, sizeof go);
  const char* prec = getenv("PREC");
  if (prec && strcmp(prec, "mixed") == 0) go.precision = KIFMM_PREC_MIXED;
  const char* graph = getenv("GRAPH");
  if (graph && atoi(graph) != 0) go.flags |= KIFMM_GPU_GRAPH;
  const char* overlap = getenv("OVERLAP");
  if (overlap && atoi(overlap) != 0) go.flags |= KIFMM_GPU_OVERLAP;
  kifmm_gpu* g = nullptr;
  if (kifmm_gpu_create(&tv, &ops, &go, &g) != KIFMM_OK) {
    fprintf(stderr, "error: %s\n", kifmm_last_error());
    return 1;
  }
  fprintf(stderr, "Finished setup.\n");
  fprintf(stderr, "Evaluation...\n");
  /* the potentials come back into the handle's pinned buffer (the copy is
   * part of the evaluation, as copy_trg_val__gpu is part of run) */
  std::vector<double> own_leaf, pot(n);
  double* pot_leaf = kifmm_gpu_output(g);
  if (!pot_leaf) {
    own_leaf.resize(n);
    pot_leaf = own_leaf.data();
  }
  kifmm_phase_times t;
  /* one untimed evaluation first: the first launch of each kernel in a
   * process loads its code object, which is no part of an evaluation */
  if (kifmm_gpu_evaluate(g, pot_leaf, &t) != KIFMM_OK || kifmm_gpu_evaluate(g, pot_leaf, &t) != KIFMM_OK) {
    fprintf(stderr, "error: %s\n", kifmm_last_error());
    return 1;
  }
  kifmm_problem_scatter(p, pot_leaf, pot.data());
  /* the evaluation's own interval (first to last event, copy included):
   * phases may overlap (OVERLAP=1) and are not recorded inside a graph */
  const double tot = t.total * 1e-3;
  fprintf(stderr, "=== Statistics ===\n");
  fprintf(stdout, "Code: hip-gfx950\n");
  fprintf(stdout, "Floating-point word size: %lu bytes\n", sizeof(double));
  fprintf(stderr, "== Timing for FMM3d ==\n");
  const double ph[7] = {t.up, t.ulist, t.vlist, t.wlist, t.xlist, t.down, t.copy};
  const char* nm[7] = {"Up    ", "U list", "V list", "W list", "X list", "Down  ", "Copy  "};
  if (go.flags & KIFMM_GPU_GRAPH)
    fprintf(stderr, "  (one HIP graph: no per-phase times)\n");
  else
    for (int i = 0; i < 7; i++)
      fprintf(stderr, "  %s: %g secs (%.1f%%)\n", nm[i], ph[i] * 1e-3, tot > 0 ? ph[i] * 1e-3 / tot * 100 : 0.0);
  fprintf(stderr, "  ==> Total Execution Time: %g secs\n", tot);
  fprintf(stderr, "  ==> Throughput: %.2f Mpoints/s\n", tot > 0 ? n / tot * 1e-6 : 0.0);
  const int nv = kifmm_getenv_validate();
  const double rerr = kifmm_problem_validate(p, pot.data(), nv, nullptr, nullptr);
  kifmm_work_counts wc;
  kifmm_gpu_work_counts(g, &wc);
  const int eff = 2 * (2 * tv.np) * (2 * tv.np) * (tv.np + 1);
  double bytes_u = 0, bytes_v = 0;
  for (int i = 0; i < tv.n_nodes; i++) {
    bytes_u += 4.0 * 8 * tv.trg_num[i];
    for (int k = tv.u_ptr[i]; k < tv.u_ptr[i + 1]; k++) bytes_u += 4.0 * 8 * tv.src_num[tv.u_idx[k]];
    bytes_v += 8.0 * eff + 2.0 * 8 * eff * (tv.v_ptr[i + 1] - tv.v_ptr[i]);
  }
  const double flops_u = 11.0 * wc.u_pairs, flops_v = 4.0 * eff * wc.v_entries;
  fprintf(stdout, "Inputs: %d | %s | %d \n", n, dist, q);
  fprintf(stdout, "Estimated no. of flops (Ulist): %g\n", flops_u);
  fprintf(stdout, "Estimated no. of flops (Vlist): %g\n", flops_v);
  fprintf(stdout, "Lower-bound on maximum memory traffic (Ulist): %g GB\n", 1e-9 * bytes_u);
  fprintf(stdout, "Lower-bound on maximum memory traffic (Vlist): %g GB\n", 1e-9 * bytes_v);
  fprintf(stdout, "Lower-bound on average intensity (Ulist): %.2f flops / byte\n", flops_u / bytes_u);
  fprintf(stdout, "Lower-bound on average intensity (Vlist): %.2f flops / byte\n", flops_v / bytes_v);
  fprintf(stderr, "Relative error: %e\n", rerr);
  kifmm_gpu_destroy(g);
  kifmm_problem_destroy(p);
  return 0;
}
