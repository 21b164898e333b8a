// Per-launch timing hook used by the kernel launchers. When a profiler is
// attached, every launch is bracketed by two events on the stream it runs on,
// tagged with its kind and its algorithmic flop count, so per-kernel average
// durations and achieved rates can be read back without an external tool.
#pragma once
#include <hip/hip_runtime.h>

namespace rsl {

enum LaunchKind : int {
  LK_ASSEMBLE = 0,
  LK_CMOD,
  LK_CDIV_SMALL,
  LK_PANEL_GETF2,
  LK_PANEL_REPLAY,
  LK_PANEL_SCHUR,
  LK_EMIT,
  LK_SOLVE_GATHER,
  LK_FWD_TRI,
  LK_FWD_UPDATE,
  LK_DSOLVE,
  LK_BWD_UPDATE,
  LK_BWD_TRI,
  LK_SOLVE_SCATTER,
  LK_RESIDUAL,
  LK_CMOD_SUM,
  LK_MAKE_G,
  LK_COUNT
};

struct LaunchProfiler {
  virtual ~LaunchProfiler() = default;
  virtual void begin(int kind, double flops, double bytes, hipStream_t st) = 0;
  virtual void end(hipStream_t st) = 0;
};

// RAII bracket around one launch; a null profiler costs one branch.
struct ProfScope {
  LaunchProfiler* p;
  hipStream_t st;
  ProfScope(LaunchProfiler* prof, int kind, double flops, double bytes, hipStream_t s) : p(prof), st(s) {
    if (p) p->begin(kind, flops, bytes, st);
  }
  ~ProfScope() {
    if (p) p->end(st);
  }
};

}  // namespace rsl
