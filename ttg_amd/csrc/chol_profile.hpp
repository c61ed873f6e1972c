// Per-launch kernel timing with HIP events (used when a driver runs with
// TTG_CHOL_F_PROFILE): every timed launch is bracketed by two events on its
// stream and tagged with a kernel kind and its algorithmic flop count.
#pragma once

#include <hip/hip_runtime.h>

namespace ttgc {

enum ProfKind { PROF_UPDATE = 0, PROF_TRSM = 1, PROF_POTRF = 2, PROF_KINDS = 3 };

// Global switch (ttg_chol_profile_enable): drivers then launch eagerly and
// time every launch.
bool prof_enabled();

// Record the start event of a timed launch on s; returns a slot id.
int prof_begin(hipStream_t s);
// Record the end event and attach kind / flops to the slot.
void prof_end(int slot, hipStream_t s, int kind, double flops);

}  // namespace ttgc
