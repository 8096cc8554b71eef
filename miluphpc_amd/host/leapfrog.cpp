// leapfrog.cpp -- Leapfrog::integrate as leapfrog.cpp:21-117: first rhs() at step 0, then
// per sub-step updateX (x += dt (v + dt/2 g), g_old = g), rhs(), updateV
// (v += dt/2 (g + g_old)), t += dt, until the sub-step end time.
#include "leapfrog.h"

Leapfrog::Leapfrog(const SimulationParameters &sp, integer numParticles, const Material &material)
    : Miluphpc(sp, numParticles, material) {
    particleHandler->initLeapfrog();
}

Leapfrog::Leapfrog(const SimulationParameters &sp, const Material &material) : Miluphpc(sp, material) {
    particleHandler->initLeapfrog();
}

Leapfrog::~Leapfrog() = default;

void Leapfrog::integrate(int step) {
    SimulationTimeHandler &t = simulationTimeHandler;
    while (t.h_currentTime < t.h_subEndTime) {
        if (step == 0 && subStep == 0) rhs(step, true, true);
        subStep++;
        check(mhpc_leapfrog_update_x(ctx_, t.h_dt), "updateX");
        rhs(step, true, true);
        check(mhpc_leapfrog_update_v(ctx_, t.h_dt), "updateV");
        t.h_currentTime += t.h_dt;
    }
}
