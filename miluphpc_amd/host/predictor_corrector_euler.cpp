// predictor_corrector_euler.cpp -- integrate(step) as predictor_corrector_euler.cpp: while
// the sub-step end time is not reached: rhs(step, true, true); setTimeStep (CFL, limited
// by the maximum time step and the remaining sub-step time); predictor; rhs(step, false,
// false) on the predicted state; corrector; t += dt.  Without the learned h, one sub-step
// is one mhpc_pc_euler_step (the same kernels, no host wait between them).
#include "predictor_corrector_euler.h"

#include <algorithm>

PredictorCorrectorEuler::PredictorCorrectorEuler(const SimulationParameters &sp, integer numParticles,
                                                 const Material &material)
    : Miluphpc(sp, numParticles, material) {}

PredictorCorrectorEuler::PredictorCorrectorEuler(const SimulationParameters &sp, const Material &material)
    : Miluphpc(sp, material) {}

PredictorCorrectorEuler::~PredictorCorrectorEuler() = default;

void PredictorCorrectorEuler::integrate(int step) {
    SimulationTimeHandler &t = simulationTimeHandler;
    while (t.h_currentTime < t.h_subEndTime) {
        subStep++;
        const uint32_t flags = rhsFlags(true);
        if (!(flags & MHPC_RHS_LEARNED_H)) {
            double dt = 0.0;
            check(mhpc_pc_euler_step(ctx_, flags, simulationParameters.courantFactor,
                                     simulationParameters.forcesFactor, simulationParameters.maxTimeStep,
                                     (double)(t.h_subEndTime - t.h_currentTime), &dt, &stats_, nullptr),
                  "integrate");
            t.h_dt = dt;
            lastDt_ = dt;
            t.h_currentTime += dt;
            continue;
        }
        rhs(step, true, true);
        double dt = 0.0;
        check(mhpc_time_step(ctx_, simulationParameters.courantFactor, simulationParameters.forcesFactor,
                             simulationParameters.maxTimeStep, &dt, nullptr),
              "setTimeStep");
        dt = std::min(dt, (double)(t.h_subEndTime - t.h_currentTime));
        t.h_dt = dt;
        lastDt_ = dt;
        check(mhpc_pc_predictor(ctx_, dt), "predictor");
        rhs(step, false, false);
        check(mhpc_pc_corrector(ctx_, dt), "corrector");
        t.h_currentTime += dt;
    }
}
