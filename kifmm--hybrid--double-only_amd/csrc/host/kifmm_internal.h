/* Declarations shared by the host and device halves of libkifmm. */
#ifndef KIFMM_INTERNAL_H
#define KIFMM_INTERNAL_H
#include <string>

namespace kifmm {
void set_error(const std::string& s);
double wall_seconds();
}  // namespace kifmm

#endif
