// internal.hpp — what the serving tier needs from a case handle (api.hip).
#pragma once

#include <cstdint>

#include "../../include/lbflow_hip.h"
#include "engine.hpp"

namespace lbf {

struct CaseServeInfo {
  int velocity_set, precision, op;
  bool aa;
  int n_scalars;
  long long N;
  int grid;
  hipStream_t stream;
  bool fresh;     // the next step starts from natural slots
  uint64_t step;  // completed steps (host view, synchronised)
  KArgs args;     // momentum step arguments, flags = store moments
};

// Fills `out` for a created, initialised and synchronised case.
int case_serve_info(lbf_case* c, CaseServeInfo* out);
// Records `steps` steps that another launcher ran on the case's arrays.
void case_note_external_steps(lbf_case* c, uint64_t steps);
// Sets the thread-local message lbf_last_error() returns; returns `code`.
int set_error(int code, const char* msg);

}  // namespace lbf
