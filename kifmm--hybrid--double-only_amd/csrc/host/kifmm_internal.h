/* Declarations shared by the host and device halves of libkifmm. */
#ifndef KIFMM_INTERNAL_H
#define KIFMM_INTERNAL_H
#include <string>

struct kifmm_problem;

namespace kifmm {
void set_error(const std::string& s);
double wall_seconds();
/* validate()'s sample (fmm/driver.cc:141-206): nv target ids drawn by
 * continuing the problem's generator stream, their coordinates
 * (xyz = 3 x nv, x then y then z) and the source count */
int sample_targets(kifmm_problem* p, int nv, int* ids, double* xyz);
}  // namespace kifmm

#endif
