// predictor_corrector_euler.h -- the reference's PredictorCorrectorEuler integrator
// (include/integrator/predictor_corrector_euler.h, src/integrator/predictor_corrector_euler.cpp):
// rhs() with self-gravity, CFL time step, predictor, rhs() on the predicted state without
// self-gravity, corrector.
#pragma once

#include "miluphpc.h"

class PredictorCorrectorEuler : public Miluphpc {
public:
    PredictorCorrectorEuler(const SimulationParameters &simulationParameters, integer numParticles,
                            const Material &material = Material());
    // particles from simulationParameters.inputFile
    explicit PredictorCorrectorEuler(const SimulationParameters &simulationParameters,
                                     const Material &material = Material());
    ~PredictorCorrectorEuler() override;
    void integrate(int step = 0) override;
    real lastTimeStep() const { return lastDt_; }

private:
    real lastDt_ = 0.0;
};
