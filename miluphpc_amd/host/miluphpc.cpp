// miluphpc.cpp -- see miluphpc.h.
#include "miluphpc.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

#include "npy_io.h"

static mhpc_params to_params(const SimulationParameters &sp, const Material &m) {
    mhpc_params p;
    mhpc_default_params(&p);
    p.theta = sp.theta;
    p.smoothing = sp.smoothing;
    p.mac = sp.gravityForceVersion <= 1 ? MHPC_MAC_SQRT_THETA : MHPC_MAC_CLASSIC;
    p.quadrupole = sp.quadrupole ? 1 : 0;
    p.bh_accept = (sp.gravityForceVersion <= 1 || sp.groupWalk || sp.quadrupole) ? MHPC_BH_VOTE
                                                                                : MHPC_BH_PER_PARTICLE;
    p.curve = sp.sfcSelection == Curve::hilbert ? MHPC_CURVE_HILBERT : MHPC_CURVE_LEBESGUE;
    p.precision = sp.precision;
    p.sml_max = sp.smlMax;
    p.smoothing_kernel = sp.smoothingKernelSelection;
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

static std::string join(const std::string &dir, const std::string &name) {
    if (dir.empty()) return name;
    return dir.back() == '/' ? dir + name : dir + "/" + name;
}

static npy::Array read_dataset(const std::string &dir, const char *name) {
    npy::Array a;
    const std::string err = npy::read(join(dir, std::string(name) + ".npy"), a);
    if (!err.empty()) throw std::runtime_error("distributionFromFile: " + err);
    return a;
}

integer Miluphpc::particlesInFile(const std::string &filename) {
    const npy::Array m = read_dataset(filename, "m");
    if (m.shape.size() != 1 || m.shape[0] < 1 || m.shape[0] > 0x3fffffff)
        throw std::runtime_error("distributionFromFile: /m must be a non-empty vector");
    return (integer)m.shape[0];
}

Miluphpc::Miluphpc(const SimulationParameters &sp, const Material &material)
    : Miluphpc(sp, particlesInFile(sp.inputFile), material) {
    distributionFromFile(sp.inputFile);
}

void Miluphpc::distributionFromFile(const std::string &filename) {
    const int64_t n = numParticles;
    ParticleHandler *ph = particleHandler;
    auto vec = [&](const char *name) {
        npy::Array a = read_dataset(filename, name);
        if (a.shape.size() != 1 || a.shape[0] != n)
            throw std::runtime_error(std::string("distributionFromFile: /") + name + " must have " +
                                     std::to_string(n) + " entries");
        return a;
    };
    auto mat3 = [&](const char *name) {
        npy::Array a = read_dataset(filename, name);
        if (a.shape.size() != 2 || a.shape[0] != n || a.shape[1] != DIM)
            throw std::runtime_error(std::string("distributionFromFile: /") + name + " must be " +
                                     std::to_string(n) + " x 3");
        return a;
    };
    const npy::Array m = vec("m"), x = mat3("x"), v = mat3("v");
    for (int64_t i = 0; i < n; i++) {
        ph->h_uid[i] = (idInteger)i;
        ph->h_mass[i] = npy::as_double(m, i);
        ph->h_x[i] = npy::as_double(x, 3 * i);
        ph->h_y[i] = npy::as_double(x, 3 * i + 1);
        ph->h_z[i] = npy::as_double(x, 3 * i + 2);
        ph->h_vx[i] = npy::as_double(v, 3 * i);
        ph->h_vy[i] = npy::as_double(v, 3 * i + 1);
        ph->h_vz[i] = npy::as_double(v, 3 * i + 2);
    }
    ph->copyDistribution(To::device, true, false);
    if (simulationParameters.sphSim) {
        const npy::Array mid = vec("materialId"), u = vec("u");
        const bool has_rho = npy::exists(join(filename, "rho.npy"));
        const bool has_sml = npy::exists(join(filename, "sml.npy"));
        npy::Array rho, sml;
        if (has_rho) rho = vec("rho");
        if (has_sml) sml = vec("sml");
        for (int64_t i = 0; i < n; i++) {
            ph->h_materialId[i] = (integer)npy::as_int(mid, i);
            ph->h_e[i] = npy::as_double(u, i);
            ph->h_rho[i] = has_rho ? npy::as_double(rho, i) : 0.0;
            ph->h_sml[i] = has_sml ? npy::as_double(sml, i) : material.sml;
        }
        ph->copyMatId(To::device);
        ph->copySPH(To::device);
    }
    if (ph->lastStatus() != MHPC_OK) check(ph->lastStatus(), "distributionFromFile");
}

real Miluphpc::particles2file(int step) {
    const auto t0 = std::chrono::steady_clock::now();
    ParticleHandler *ph = particleHandler;
    const int64_t n = numParticles;
    ph->copyDistribution(To::host, true, false);
    if (simulationParameters.sphSim) ph->copySPH(To::host);
    std::vector<uint64_t> key((size_t)n);
    check(mhpc_download(ctx_, MHPC_F_KEY, key.data(), n), "particles2file: keys");
    check(ph->lastStatus(), "particles2file");
    char name[32];
    snprintf(name, sizeof(name), "ts%06d", step);
    const std::string dir = simulationParameters.directory.empty() ? std::string(name)
                                                                   : join(simulationParameters.directory, name);
    std::string err = npy::make_dir(dir);
    std::vector<double> x3((size_t)(3 * n)), v3((size_t)(3 * n));
    for (int64_t i = 0; i < n; i++) {
        x3[3 * i] = ph->h_x[i];
        x3[3 * i + 1] = ph->h_y[i];
        x3[3 * i + 2] = ph->h_z[i];
        v3[3 * i] = ph->h_vx[i];
        v3[3 * i + 1] = ph->h_vy[i];
        v3[3 * i + 2] = ph->h_vz[i];
    }
    const double time = simulationTimeHandler.h_currentTime;
    const std::vector<int32_t> proc((size_t)n, 0);
    auto put = [&](const char *ds, npy::Type t, std::vector<int64_t> shape, const void *data) {
        if (err.empty()) err = npy::write(join(dir, std::string(ds) + ".npy"), t, shape, data);
    };
    put("time", npy::Type::f8, {1}, &time);
    put("x", npy::Type::f8, {n, 3}, x3.data());
    put("v", npy::Type::f8, {n, 3}, v3.data());
    put("key", npy::Type::u8, {n}, key.data());
    put("m", npy::Type::f8, {n}, ph->h_mass);
    put("proc", npy::Type::i4, {n}, proc.data());
    if (simulationParameters.sphSim) {
        put("rho", npy::Type::f8, {n}, ph->h_rho);
        put("p", npy::Type::f8, {n}, ph->h_p);
        put("e", npy::Type::f8, {n}, ph->h_e);
        put("sml", npy::Type::f8, {n}, ph->h_sml);
        put("noi", npy::Type::i4, {n}, ph->h_noi);
        put("cs", npy::Type::f8, {n}, ph->h_cs);
    }
    if (simulationParameters.calculateEnergy) put("totalEnergy", npy::Type::f8, {1}, &totalEnergy);
    if (simulationParameters.calculateAngularMomentum)
        put("totalAngularMomentum", npy::Type::f8, {1}, &totalAngularMomentum);
    if (!err.empty()) throw std::runtime_error("particles2file: " + err);
    return (real)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
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
    check(mhpc_rhs(ctx_, rhsFlags(selfGravity), &stats_), "rhs");
    return (real)stats_.ms_total;
}

uint32_t Miluphpc::rhsFlags(bool selfGravity) const {
    uint32_t flags = 0;
    if (simulationParameters.gravitySim && selfGravity) flags |= MHPC_RHS_GRAVITY;
    if (simulationParameters.sphSim) flags |= MHPC_RHS_SPH;
    if (simulationParameters.sphSim && simulationParameters.variableSML) {
        flags |= MHPC_RHS_VARIABLE_SML;
        if (simulationParameters.speculativeLists) flags |= MHPC_RHS_SPECULATIVE_NN;
        if (simulationParameters.learnedSML) flags |= MHPC_RHS_LEARNED_H;
    }
    if (simulationParameters.calculateEnergy && (flags & MHPC_RHS_GRAVITY)) flags |= MHPC_RHS_POTENTIAL;
    return flags;
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
