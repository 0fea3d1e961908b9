/* Declarations shared by the host and device halves of libkifmm. */
#ifndef KIFMM_INTERNAL_H
#define KIFMM_INTERNAL_H
#include <string>
#include <vector>

#include "kifmm_host.h"

namespace kifmm {
void set_error(const std::string& s);
double wall_seconds();
/* validate()'s sample (fmm/driver.cc:141-206): nv target ids drawn by
 * continuing the problem's generator stream, their coordinates
 * (xyz = 3 x nv, x then y then z) and the source count */
int sample_targets(kifmm_problem* p, int nv, int* ids, double* xyz);

/* Tree, lists and gather permutations of a problem, as kifmm_tree_view
 * exposes them (node order, CSR lists over all nodes). */
struct TreeArrays {
  std::vector<int> level, parent, child, depth, tag, path, sbeg, snum, tbeg, tnum, sperm, tperm, lptr[4], lidx[4];
};
/* a problem's points without its tree: kifmm_problem_create's generators,
 * or caller points checked as kifmm_problem_from_points checks them */
kifmm_problem* problem_generate(int n, const char* distribution, int pts_max, const kifmm_problem_opts* opts);
kifmm_problem* problem_points(int n, const double* sx, const double* sy, const double* sz, const double* den,
                              const double* tx, const double* ty, const double* tz, int pts_max, int np);
/* input-order coordinates (sources, targets), count and points per box */
void problem_coords(const kifmm_problem* p, const double* s3[3], const double* t3[3], int* n, int* q);
/* install a tree built elsewhere, then gather the points into leaf order */
void problem_install(kifmm_problem* p, TreeArrays&& t, double seconds);
}  // namespace kifmm

#endif
