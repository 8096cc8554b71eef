// leapfrog.h -- the reference's Leapfrog integrator (include/integrator/leapfrog.h,
// src/integrator/leapfrog.cpp:21-117): kick-drift-kick on the gravitational accelerations.
#pragma once

#include "miluphpc.h"

class Leapfrog : public Miluphpc {
public:
    Leapfrog(const SimulationParameters &simulationParameters, integer numParticles,
             const Material &material = Material());
    // particles from simulationParameters.inputFile (Leapfrog(SimulationParameters), leapfrog.cpp:3-12)
    explicit Leapfrog(const SimulationParameters &simulationParameters, const Material &material = Material());
    ~Leapfrog() override;
    void integrate(int step = 0) override;
};
