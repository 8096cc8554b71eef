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
    refreshPointers();
}

// the d_* members as the context currently has them
void ParticleHandler::refreshPointers() {
    mhpc_ctx *ctx = ctx_;
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
    if (leapfrog) {
        d_g_ax_old = dptr<real>(ctx, MHPC_F_G_AX_OLD);
        d_g_ay_old = dptr<real>(ctx, MHPC_F_G_AY_OLD);
        d_g_az_old = dptr<real>(ctx, MHPC_F_G_AZ_OLD);
    }
}

ParticleHandler::~ParticleHandler() {
    freeLeapfrog();
    real *hr[] = {h_mass, h_x, h_vx, h_ax, h_g_ax, h_y, h_vy, h_ay, h_g_ay, h_z, h_vz, h_az,
                  h_g_az, h_sml, h_e, h_dedt, h_u, h_cs, h_rho, h_p, h_drhodt, h_dsmldt};
    for (real *p : hr) delete[] p;
    integer *hi[] = {h_noi, h_level, h_materialId, h_uid};
    for (integer *p : hi) delete[] p;
}

void ParticleHandler::initLeapfrog() {
    if (leapfrog) return;
    leapfrog = true;
    const size_t n = (size_t)numParticles;
    h_g_ax_old = new real[n]();
    h_g_ay_old = new real[n]();
    h_g_az_old = new real[n]();
    refreshPointers();
}

void ParticleHandler::freeLeapfrog() {
    if (!leapfrog) return;
    leapfrog = false;
    delete[] h_g_ax_old;
    delete[] h_g_ay_old;
    delete[] h_g_az_old;
    h_g_ax_old = h_g_ay_old = h_g_az_old = nullptr;
    d_g_ax_old = d_g_ay_old = d_g_az_old = nullptr;
}

template <typename T>
T *&ParticleHandler::getEntry(Entry::Name entry, Execution::Location location) {
    static_assert(sizeof(T) == sizeof(real), "getEntry returns the FP64 arrays");
    const bool dev = location == Execution::device;
    switch (entry) {
        case Entry::x: return reinterpret_cast<T *&>(dev ? d_x : h_x);
        case Entry::y: return reinterpret_cast<T *&>(dev ? d_y : h_y);
        case Entry::z: return reinterpret_cast<T *&>(dev ? d_z : h_z);
        case Entry::mass: return reinterpret_cast<T *&>(dev ? d_mass : h_mass);
    }
    throw std::invalid_argument("ParticleHandler::getEntry: unknown entry");
}

template real *&ParticleHandler::getEntry<real>(Entry::Name, Execution::Location);

void ParticleHandler::setPointer(IntegratedParticleHandler *iph) {
    for (int k = 0; k < IntegratedParticleHandler::kNumFields; k++) {
        const int st = mhpc_bind_field(ctx_, iph->fieldId(k), iph->fieldArray(k));
        if (st != MHPC_OK) status_ = st;
    }
    refreshPointers();
}

void ParticleHandler::resetPointer() {
    for (int f = 0; f < MHPC_F_NOI; f++) {
        const int st = mhpc_bind_field(ctx_, f, nullptr);
        if (st != MHPC_OK) status_ = st;
    }
    refreshPointers();
}

void ParticleHandler::copyField(int field, void *host, To::Target target) {
    const int st = (target == To::device) ? mhpc_upload(ctx_, field, host, numParticles)
                                          : mhpc_download(ctx_, field, host, numParticles);
    if (st != MHPC_OK) status_ = st;
}

void ParticleHandler::copyMass(To::Target target, bool) { copyField(MHPC_F_MASS, h_mass, target); }

void ParticleHandler::copyUid(To::Target) {
    // a particle's identity is its index in the caller's order (one subdomain, no particle
    // leaves it): h_uid stays on the host, unchanged
}

void ParticleHandler::copyMatId(To::Target target) {
    // one material per context: every particle must carry material id 0
    if (target != To::device) return;
    for (integer i = 0; i < numParticles; i++)
        if (h_materialId[i] != 0) {
            status_ = MHPC_E_ARG;
            throw std::invalid_argument("ParticleHandler::copyMatId: particle " + std::to_string(i) +
                                        " has material id " + std::to_string(h_materialId[i]) +
                                        "; this build runs one material (id 0) per context");
        }
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
    if (leapfrog) {
        copyField(MHPC_F_G_AX_OLD, h_g_ax_old, target);
        copyField(MHPC_F_G_AY_OLD, h_g_ay_old, target);
        copyField(MHPC_F_G_AZ_OLD, h_g_az_old, target);
    }
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

// ---------------------------------------------------------------------------------------
static const int kIntegratedFields[IntegratedParticleHandler::kNumFields] = {
    MHPC_F_X, MHPC_F_VX, MHPC_F_AX, MHPC_F_Y, MHPC_F_VY, MHPC_F_AY, MHPC_F_Z, MHPC_F_VZ, MHPC_F_AZ,
    MHPC_F_RHO, MHPC_F_E, MHPC_F_DEDT, MHPC_F_P, MHPC_F_CS, MHPC_F_SML, MHPC_F_DRHODT, MHPC_F_DSMLDT};

IntegratedParticleHandler::IntegratedParticleHandler(integer numParticles, integer numNodes, mhpc_ctx *ctx)
    : numParticles(numParticles), numNodes(numNodes), ctx_(ctx) {
    real **arr[kNumFields] = {&d_x, &d_vx, &d_ax, &d_y, &d_vy, &d_ay, &d_z, &d_vz, &d_az,
                              &d_rho, &d_e, &d_dedt, &d_p, &d_cs, &d_sml, &d_drhodt, &d_dsmldt};
    for (real **a : arr) *a = nullptr;
    for (real **a : arr) {
        void *p = nullptr;
        const int st = mhpc_device_alloc(ctx, (int64_t)sizeof(real) * numParticles, &p);
        if (st != MHPC_OK) {
            this->~IntegratedParticleHandler();
            throw std::runtime_error(std::string("IntegratedParticleHandler: ") + mhpc_status_string(st));
        }
        *a = static_cast<real *>(p);
    }
}

IntegratedParticleHandler::~IntegratedParticleHandler() {
    real *arr[kNumFields] = {d_x, d_vx, d_ax, d_y, d_vy, d_ay, d_z, d_vz, d_az,
                             d_rho, d_e, d_dedt, d_p, d_cs, d_sml, d_drhodt, d_dsmldt};
    for (real *a : arr)
        if (a) mhpc_device_free(ctx_, a);
}

int IntegratedParticleHandler::fieldId(int k) const { return kIntegratedFields[k]; }

real *IntegratedParticleHandler::fieldArray(int k) const {
    real *arr[kNumFields] = {d_x, d_vx, d_ax, d_y, d_vy, d_ay, d_z, d_vz, d_az,
                             d_rho, d_e, d_dedt, d_p, d_cs, d_sml, d_drhodt, d_dsmldt};
    return arr[k];
}
