// particle_handler.h -- the reference's ParticleHandler and IntegratedParticleHandler
// (include/particle_handler.h:16-455, :458-540) over the C-ABI: host arrays h_* (owned here)
// and device arrays d_* (owned by the mhpc context, in the caller's particle order), copied
// with copy*(To::Target); setPointer / resetPointer point the context's fields at a second
// particle set and back, as the predictor-corrector integrator of the reference does.
#pragma once

#include <stdexcept>
#include <string>

#include "../../include/mhpc.h"
#include "parameter.h"

class IntegratedParticleHandler;

class ParticleHandler {
public:
    bool leapfrog = false;
    integer numParticles;
    integer numNodes;

    real *h_mass;
    real *h_x, *h_vx, *h_ax, *h_g_ax;
    real *h_y, *h_vy, *h_ay, *h_g_ay;
    real *h_z, *h_vz, *h_az, *h_g_az;
    real *h_g_ax_old = nullptr, *h_g_ay_old = nullptr, *h_g_az_old = nullptr;
    real *h_sml, *h_e, *h_dedt, *h_u, *h_cs, *h_rho, *h_p, *h_drhodt, *h_dsmldt;
    integer *h_noi, *h_level, *h_materialId;
    idInteger *h_uid;

    real *d_mass;
    real *d_x, *d_vx, *d_ax, *d_g_ax;
    real *d_y, *d_vy, *d_ay, *d_g_ay;
    real *d_z, *d_vz, *d_az, *d_g_az;
    real *d_g_ax_old = nullptr, *d_g_ay_old = nullptr, *d_g_az_old = nullptr;
    real *d_sml, *d_e, *d_dedt, *d_u, *d_cs, *d_rho, *d_p, *d_drhodt, *d_dsmldt;
    integer *d_noi, *d_level;

    ParticleHandler(integer numParticles, integer numNodes, mhpc_ctx *ctx);
    ~ParticleHandler();

    // leapfrog's old gravitational accelerations (particle_handler.cpp:630-683): the context
    // always holds them; these expose the host and device arrays
    void initLeapfrog();
    void freeLeapfrog();

    // the device (or host) array of x, y, z or mass (particle_handler.cpp:685-740)
    template <typename T>
    T *&getEntry(Entry::Name entry, Execution::Location location = Execution::device);

    // every following rhs() reads and writes the integrated set's arrays; resetPointer goes
    // back to the particle set's own (particle_handler.cpp:742-900)
    void setPointer(IntegratedParticleHandler *integratedParticleHandler);
    void resetPointer();

    void copyMass(To::Target target = To::device, bool includePseudoParticles = false);
    void copyUid(To::Target target = To::device);
    void copyMatId(To::Target target = To::device);
    void copySML(To::Target target = To::device);
    void copyPosition(To::Target target = To::device, bool includePseudoParticles = false);
    void copyVelocity(To::Target target = To::device, bool includePseudoParticles = false);
    void copyAcceleration(To::Target target = To::device, bool includePseudoParticles = false);
    void copyDistribution(To::Target target = To::device, bool velocity = true,
                          bool acceleration = true, bool includePseudoParticles = false);
    void copySPH(To::Target target);

    // status of the last failed copy (MHPC_OK if none)
    int lastStatus() const { return status_; }

private:
    mhpc_ctx *ctx_;
    int status_ = MHPC_OK;
    void copyField(int field, void *host, To::Target target);
    void refreshPointers();
};

// The second particle set of the predictor-corrector integrator: device arrays of the
// integrated quantities, allocated through the context.
class IntegratedParticleHandler {
public:
    integer numParticles;
    integer numNodes;

    idInteger *d_uid = nullptr;
    real *d_x, *d_vx, *d_ax;
    real *d_y, *d_vy, *d_ay;
    real *d_z, *d_vz, *d_az;
    real *d_rho, *d_e, *d_dedt, *d_p, *d_cs, *d_sml, *d_drhodt, *d_dsmldt;

    IntegratedParticleHandler(integer numParticles, integer numNodes, mhpc_ctx *ctx);
    ~IntegratedParticleHandler();

    // (field id, array) pairs of the set, for setPointer
    static constexpr int kNumFields = 17;
    int fieldId(int k) const;
    real *fieldArray(int k) const;

private:
    mhpc_ctx *ctx_;
};
