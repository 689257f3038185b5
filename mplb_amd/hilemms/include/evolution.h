// evolution.h -- the time loop (reference Src/evolution.h:70-129).  The
// templates keep the reference signatures; given StreamCollision they hand
// the whole loop to libmplbhip (graph-captured steps, residual fused into
// the check step), any other cycle function is called step by step.
#ifndef EVOLUTION3D_H_
#define EVOLUTION3D_H_
#include "field.h"
#include "flowfield.h"
#include "model.h"
#include "type.h"

void StreamCollision(const Real time);
void SwapStreamCollision(const Real time);
void Iterate(const SizeType steps, const SizeType checkPointPeriod,
             const SizeType start = 0);
void Iterate(const Real convergenceCriteria, const SizeType checkPointPeriod,
             const SizeType start = 0);

namespace hilemms_detail {
bool IsStreamCollision(void (*cycle)(Real));
void CheckpointUpdate(SizeType iter);
Real CheckpointResidual(SizeType iter, SizeType checkPeriod);
void Finish();
}  // namespace hilemms_detail

template <typename T>
void Iterate(void (*cycle)(T), const SizeType steps, const SizeType checkPointPeriod,
             const SizeType start = 0) {
    if (hilemms_detail::IsStreamCollision(cycle)) {
        Iterate(steps, checkPointPeriod, start);
        return;
    }
    ops_printf("Starting the iteration...\n");
    for (SizeType iter = start; iter < start + steps; iter++) {
        const Real time{iter * TimeStep()};
        cycle(time);
        if (((iter + 1) % checkPointPeriod) == 0) hilemms_detail::CheckpointUpdate(iter + 1);
    }
    hilemms_detail::Finish();
}

template <typename T>
void Iterate(void (*cycle)(T), const Real convergenceCriteria,
             const SizeType checkPointPeriod, const SizeType start = 0) {
    if (hilemms_detail::IsStreamCollision(cycle)) {
        Iterate(convergenceCriteria, checkPointPeriod, start);
        return;
    }
    SizeType iter{start};
    Real residualError{1};
    do {
        const Real time{iter * TimeStep()};
        cycle(time);
        iter = iter + 1;
        if ((iter % checkPointPeriod) == 0)
            residualError = hilemms_detail::CheckpointResidual(iter, checkPointPeriod);
    } while (residualError >= convergenceCriteria);
    hilemms_detail::Finish();
}
#endif
