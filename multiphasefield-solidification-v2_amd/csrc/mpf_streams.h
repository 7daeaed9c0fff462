// Streams and events of one step of the resident solver (HIP side only).
#pragma once
#include <hip/hip_runtime.h>

namespace mpf {

// Besides the solver stream.  side == null runs every kernel on the solver
// stream: one after another (fused == 0), or the interface and many-phase
// passes as one launch (fused == 1).
struct StepStreams {
    int fused;
    hipStream_t side;      // many-phase kernel
    hipStream_t side2;     // interface cells of several solids
    hipEvent_t ev_merged;  // merge pass done (fork)
    hipEvent_t ev_general; // joins of the side streams
    hipEvent_t ev_front;
};

} // namespace mpf
