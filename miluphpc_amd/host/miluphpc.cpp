// miluphpc.cpp -- see miluphpc.h.
#include "miluphpc.h"

#include <cmath>

static mhpc_params to_params(const SimulationParameters &sp, const Material &m) {
    mhpc_params p;
    mhpc_default_params(&p);
    p.theta = sp.theta;
    p.smoothing = sp.smoothing;
    p.mac = sp.gravityForceVersion <= 1 ? MHPC_MAC_SQRT_THETA : MHPC_MAC_CLASSIC;
    p.quadrupole = sp.quadrupole ? 1 : 0;
    p.curve = sp.sfcSelection == Curve::hilbert ? MHPC_CURVE_HILBERT : MHPC_CURVE_LEBESGUE;
    p.precision = sp.precision;
    p.sml_max = sp.smlMax;
    p.eos = m.eos.type;
    p.gamma = m.eos.polytropic_gamma;
    p.K = m.eos.polytropic_K;
    p.cs_const = m.eos.cs;
    p.alpha = m.artificialViscosity.alpha;
    p.beta = m.artificialViscosity.beta;
    p.n_target = m.interactions;
    p.band_lo = m.bandLow;
    p.band_hi = m.bandHigh;
    return p;
}

Miluphpc::Miluphpc(const SimulationParameters &sp, integer numParticles, const Material &material)
    : simulationParameters(sp), material(material), numParticles(numParticles),
      numParticlesLocal(numParticles), numNodes(2 * numParticles + 20000) {
    const mhpc_params p = to_params(sp, material);
    // node capacity as the reference sizes it (miluphpc.cpp:11): 2 N + 20000 nodes
    check(mhpc_create(&p, numParticles, numNodes - numParticles, &ctx_), "mhpc_create");
    check(mhpc_set_num_particles(ctx_, numParticles), "mhpc_set_num_particles");
    particleHandler = new ParticleHandler(numParticles, numNodes, ctx_);
    simulationTimeHandler.h_dt = sp.timeStep;
    simulationTimeHandler.h_endTime = sp.timeEnd;
    simulationTimeHandler.h_subEndTime = sp.timeEnd;
}

Miluphpc::~Miluphpc() {
    delete particleHandler;
    if (ctx_) mhpc_destroy(ctx_);
}

void Miluphpc::check(int status, const char *where) {
    if (status != MHPC_OK) {
        std::string msg = std::string(where) + ": " + mhpc_status_string(status);
        if (ctx_) msg += std::string(" (") + mhpc_last_error(ctx_) + ")";
        throw MhpcException(status, msg);
    }
}

real Miluphpc::rhs(int step, bool selfGravity, bool assignParticlesToProcess) {
    (void)step;
    (void)assignParticlesToProcess;
    uint32_t flags = 0;
    if (simulationParameters.gravitySim && selfGravity) flags |= MHPC_RHS_GRAVITY;
    if (simulationParameters.sphSim) flags |= MHPC_RHS_SPH;
    if (simulationParameters.sphSim && simulationParameters.variableSML) {
        flags |= MHPC_RHS_VARIABLE_SML;
        if (simulationParameters.speculativeLists) flags |= MHPC_RHS_SPECULATIVE_NN;
        if (simulationParameters.learnedSML) flags |= MHPC_RHS_LEARNED_H;
    }
    if (simulationParameters.calculateEnergy && (flags & MHPC_RHS_GRAVITY)) flags |= MHPC_RHS_POTENTIAL;
    check(mhpc_rhs(ctx_, flags, &stats_), "rhs");
    return (real)stats_.ms_total;
}

real Miluphpc::energy() {
    mhpc_diagnostics d;
    check(mhpc_compute_diagnostics(ctx_, &d), "energy");
    totalEnergy = d.kinetic + d.potential + d.internal;
    return totalEnergy;
}

real Miluphpc::angularMomentum() {
    mhpc_diagnostics d;
    check(mhpc_compute_diagnostics(ctx_, &d), "angularMomentum");
    totalAngularMomentum = d.angular_momentum_norm;
    return totalAngularMomentum;
}

void Miluphpc::afterIntegrationStep() {
    if (simulationParameters.calculateEnergy) energy();
    if (simulationParameters.calculateAngularMomentum) angularMomentum();
}
