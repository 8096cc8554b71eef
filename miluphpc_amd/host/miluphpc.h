// miluphpc.h -- the reference's Miluphpc class (include/miluphpc.h:74-513) as a drop-in
// facade over the C-ABI: same constructor role, the same rhs(step, selfGravity,
// assignParticlesToProcess) returning the accumulated kernel milliseconds, the abstract
// integrate(step), afterIntegrationStep() with energy and angular momentum.  One
// subdomain on one GPU: assignParticlesToProcess is accepted and has no effect.
#pragma once

#include <stdexcept>
#include <string>

#include "../../include/mhpc.h"
#include "material.h"
#include "parameter.h"
#include "particle_handler.h"

// minimal stand-in for SimulationTimeHandler (simulation_time_handler.h): host-side times
struct SimulationTimeHandler {
    real h_dt = 0.0;
    real h_startTime = 0.0;
    real h_subEndTime = 0.0;
    real h_endTime = 0.0;
    real h_currentTime = 0.0;
    real *h_dt_ptr() { return &h_dt; }
};

class MhpcException : public std::runtime_error {
public:
    int status;
    MhpcException(int status, const std::string &what) : std::runtime_error(what), status(status) {}
};

class Miluphpc {
public:
    SimulationParameters simulationParameters;
    Material material;
    integer numParticles;
    integer numParticlesLocal;
    integer numNodes;
    int subStep = 0;
    real h_searchRadius = 0.0;
    real totalEnergy = 0.0;
    real totalAngularMomentum = 0.0;

    ParticleHandler *particleHandler = nullptr;
    SimulationTimeHandler simulationTimeHandler;

    Miluphpc(const SimulationParameters &simulationParameters, integer numParticles,
             const Material &material = Material());
    // the reference's construction (miluphpc.cpp:3-93, prepareSimulation :300-385): the
    // particle count and the initial distribution come from simulationParameters.inputFile
    // (a directory of .npy datasets, see npy_io.h)
    explicit Miluphpc(const SimulationParameters &simulationParameters, const Material &material = Material());
    virtual ~Miluphpc();

    // Miluphpc::distributionFromFile (miluphpc.cpp:157-300): /m, /x (N x 3), /v (N x 3),
    // /materialId, /u (internal energy, SPH runs) and, when present, /rho and /sml (otherwise
    // every particle gets the material's sml), into the host arrays and then the device.
    // Throws std::runtime_error on a missing dataset or a size mismatch.
    void distributionFromFile(const std::string &filename);
    // Miluphpc::particles2file (miluphpc.cpp:3451-3760): the snapshot of the current state as
    // simulationParameters.directory + "ts%06d" (datasets time, x, v, key, m, proc, rho, p, e,
    // sml, noi, cs, and totalEnergy / totalAngularMomentum when those are computed).  Returns
    // the elapsed milliseconds.
    real particles2file(int step);
    // particle count of a distribution directory (its /m dataset)
    static integer particlesInFile(const std::string &filename);

    // Right-hand side (miluphpc.cpp:387-562).  Throws MhpcException on a device error
    // (where the reference exits or traps).
    real rhs(int step, bool selfGravity = true, bool assignParticlesToProcess = true);

    virtual void integrate(int step = 0) = 0;

    void afterIntegrationStep();
    real energy();
    real angularMomentum();

    const mhpc_rhs_stats &lastStats() const { return stats_; }
    mhpc_ctx *context() { return ctx_; }

protected:
    mhpc_ctx *ctx_ = nullptr;
    mhpc_rhs_stats stats_{};
    void check(int status, const char *where);
    // the MHPC_RHS_* flags of rhs(step, selfGravity, ...) under the simulation parameters
    uint32_t rhsFlags(bool selfGravity) const;
};
