// particle_handler.cpp -- see particle_handler.h.  Copies go through mhpc_upload /
// mhpc_download (particle_handler.cpp:207-1230 in the reference do the same with
// cudaMemcpy).  Pseudo-particles (cells) live inside the context and are never exposed
// as particle fields, so includePseudoParticles is accepted and has no effect.
#include "particle_handler.h"

#include <cstring>

template <typename T>
static T *dptr(mhpc_ctx *ctx, int field) {
    void *p = nullptr;
    mhpc_field_ptr(ctx, field, &p);
    return static_cast<T *>(p);
}

ParticleHandler::ParticleHandler(integer numParticles, integer numNodes, mhpc_ctx *ctx)
    : numParticles(numParticles), numNodes(numNodes), ctx_(ctx) {
    const size_t n = (size_t)numParticles;
    real **hr[] = {&h_mass, &h_x, &h_vx, &h_ax, &h_g_ax, &h_y, &h_vy, &h_ay, &h_g_ay, &h_z, &h_vz,
                   &h_az, &h_g_az, &h_sml, &h_e, &h_dedt, &h_u, &h_cs, &h_rho, &h_p, &h_drhodt,
                   &h_dsmldt};
    for (real **p : hr) {
        *p = new real[n];
        std::memset(*p, 0, sizeof(real) * n);
    }
    integer **hi[] = {&h_noi, &h_level, &h_materialId, &h_uid};
    for (integer **p : hi) {
        *p = new integer[n];
        std::memset(*p, 0, sizeof(integer) * n);
    }
    for (integer i = 0; i < numParticles; i++) h_uid[i] = i;
    d_mass = dptr<real>(ctx, MHPC_F_MASS);
    d_x = dptr<real>(ctx, MHPC_F_X);
    d_y = dptr<real>(ctx, MHPC_F_Y);
    d_z = dptr<real>(ctx, MHPC_F_Z);
    d_vx = dptr<real>(ctx, MHPC_F_VX);
    d_vy = dptr<real>(ctx, MHPC_F_VY);
    d_vz = dptr<real>(ctx, MHPC_F_VZ);
    d_ax = dptr<real>(ctx, MHPC_F_AX);
    d_ay = dptr<real>(ctx, MHPC_F_AY);
    d_az = dptr<real>(ctx, MHPC_F_AZ);
    d_g_ax = dptr<real>(ctx, MHPC_F_G_AX);
    d_g_ay = dptr<real>(ctx, MHPC_F_G_AY);
    d_g_az = dptr<real>(ctx, MHPC_F_G_AZ);
    d_sml = dptr<real>(ctx, MHPC_F_SML);
    d_e = dptr<real>(ctx, MHPC_F_E);
    d_dedt = dptr<real>(ctx, MHPC_F_DEDT);
    d_u = dptr<real>(ctx, MHPC_F_U);
    d_cs = dptr<real>(ctx, MHPC_F_CS);
    d_rho = dptr<real>(ctx, MHPC_F_RHO);
    d_p = dptr<real>(ctx, MHPC_F_P);
    d_drhodt = dptr<real>(ctx, MHPC_F_DRHODT);
    d_dsmldt = dptr<real>(ctx, MHPC_F_DSMLDT);
    d_noi = dptr<integer>(ctx, MHPC_F_NOI);
    d_level = dptr<integer>(ctx, MHPC_F_LEVEL);
}

ParticleHandler::~ParticleHandler() {
    real *hr[] = {h_mass, h_x, h_vx, h_ax, h_g_ax, h_y, h_vy, h_ay, h_g_ay, h_z, h_vz, h_az,
                  h_g_az, h_sml, h_e, h_dedt, h_u, h_cs, h_rho, h_p, h_drhodt, h_dsmldt};
    for (real *p : hr) delete[] p;
    integer *hi[] = {h_noi, h_level, h_materialId, h_uid};
    for (integer *p : hi) delete[] p;
}

void ParticleHandler::copyField(int field, void *host, To::Target target) {
    const int st = (target == To::device) ? mhpc_upload(ctx_, field, host, numParticles)
                                          : mhpc_download(ctx_, field, host, numParticles);
    if (st != MHPC_OK) status_ = st;
}

void ParticleHandler::copyMass(To::Target target, bool) { copyField(MHPC_F_MASS, h_mass, target); }

void ParticleHandler::copyUid(To::Target) {
    // particle identity is the index in the caller's order; uids stay on the host
}

void ParticleHandler::copyMatId(To::Target) {
    // single-material build: the material id is not needed on the device
}

void ParticleHandler::copySML(To::Target target) { copyField(MHPC_F_SML, h_sml, target); }

void ParticleHandler::copyPosition(To::Target target, bool) {
    copyField(MHPC_F_X, h_x, target);
    copyField(MHPC_F_Y, h_y, target);
    copyField(MHPC_F_Z, h_z, target);
}

void ParticleHandler::copyVelocity(To::Target target, bool) {
    copyField(MHPC_F_VX, h_vx, target);
    copyField(MHPC_F_VY, h_vy, target);
    copyField(MHPC_F_VZ, h_vz, target);
}

void ParticleHandler::copyAcceleration(To::Target target, bool) {
    copyField(MHPC_F_AX, h_ax, target);
    copyField(MHPC_F_AY, h_ay, target);
    copyField(MHPC_F_AZ, h_az, target);
    copyField(MHPC_F_G_AX, h_g_ax, target);
    copyField(MHPC_F_G_AY, h_g_ay, target);
    copyField(MHPC_F_G_AZ, h_g_az, target);
}

void ParticleHandler::copyDistribution(To::Target target, bool velocity, bool acceleration, bool) {
    copyMass(target);
    copyPosition(target);
    if (velocity) copyVelocity(target);
    if (acceleration) copyAcceleration(target);
}

void ParticleHandler::copySPH(To::Target target) {
    copyField(MHPC_F_SML, h_sml, target);
    copyField(MHPC_F_E, h_e, target);
    copyField(MHPC_F_RHO, h_rho, target);
    copyField(MHPC_F_P, h_p, target);
    copyField(MHPC_F_CS, h_cs, target);
    if (target == To::host) {
        copyField(MHPC_F_DEDT, h_dedt, target);
        copyField(MHPC_F_DRHODT, h_drhodt, target);
        copyField(MHPC_F_DSMLDT, h_dsmldt, target);
        copyField(MHPC_F_NOI, h_noi, target);
    }
}
