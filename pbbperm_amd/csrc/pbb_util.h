// pbb_util.h -- host helpers with no device dependency, shared by every host
// source of the library (pbb_io.cpp defines them).  Not part of the C ABI.
#pragma once

#include "../../include/pbb_engine.h"

namespace pbb {

// record a message for pbb_last_error() and return code
int fail(int code, const char *fmt, ...);
// every job id 0..n-1 exactly once (the ids index the time matrix)
bool is_perm(const int *p, int n);
int validate(const pbb_config *c);
int check_shape(const int32_t *ptm, int n, int m);
// makespan of the first len jobs of seq (machine-major times, n jobs)
int host_partial_makespan(const int32_t *ptm, int n, int m, const int *seq, int len);
// Stream replays report on stderr every 500 served requests when the
// environment sets PBB_REPLAY_PROGRESS (a long stream runs for minutes).
void replay_progress(const char *what, int before, int after, int count);

}  // namespace pbb
