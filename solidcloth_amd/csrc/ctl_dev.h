// Device-side start of a substep in the solver control block.
#pragma once

#include "kernels.h"

namespace sc {

// Called by one thread of the first kernel of a substep: the finished
// substep's self-collision counts go into the step's peaks (so that a step
// that overflowed knows how far to grow), then the per-substep counters and
// error maxima are cleared.
__device__ __forceinline__ void begin_substep_ctl(SolverCtl* ctl)
{
    ctl->selfPeak[0] = max(ctl->selfPeak[0], ctl->selfCount[0]);
    ctl->selfPeak[1] = max(ctl->selfPeak[1], ctl->selfCount[1]);
    ctl->selfPeak[2] = max(ctl->selfPeak[2], ctl->candSlots[0]);
    ctl->selfPeak[3] = max(ctl->selfPeak[3], ctl->candSlots[1]);
    ctl->done = 0;
    ctl->k = 0;
    ctl->cb = 0;
    ctl->cv = 0;
    ctl->ce = 0;
    ctl->ct = 0;
    ctl->sct = 0;
    ctl->sce = 0;
    ctl->cl = 0;
    ctl->selfCount[0] = 0;
    ctl->selfCount[1] = 0;
    ctl->candSlots[0] = 0;
    ctl->candSlots[1] = 0;
    ctl->iterations = 0;
}

// Called by threads 0 .. kErrParts - 1 of the first kernel of a substep.
__device__ __forceinline__ void begin_substep_parts(SolverCtl* ctl, int t)
{
    ctl->errPart[0][t * kErrStride] = 0;
    ctl->errPart[1][t * kErrStride] = 0;
}

// This workgroup's part of error `which` (0 bend, 1 vertex contacts).
__device__ __forceinline__ unsigned* err_part(SolverCtl* ctl, int which)
{
    return &ctl->errPart[which][(blockIdx.x % kErrParts) * kErrStride];
}

}  // namespace sc
