// rhs_driver.cpp -- drives the C++ facade exactly as reference code would: construct an
// integrator (Leapfrog), fill particleHandler->h_*, copyDistribution(To::device), call
// rhs() / integrate(), copy back.  Writes inputs and outputs to a raw binary file so the
// caller can check them.
//
//   mhpc_rhs_driver gravity <n> <steps> <out>   Leapfrog steps on a uniform ball (FP64)
//   mhpc_rhs_driver sph <nside> <out>           one SPH rhs() on a jittered lattice
//   mhpc_rhs_driver io <sph|gravity> <in> <outdir> <steps>
//        the reference's run: integrator constructed from the initial distribution <in>
//        (a directory of .npy datasets), snapshot ts000000 after the first rhs(), then
//        <steps> integrate() calls, each followed by a snapshot ts%06d
//   mhpc_rhs_driver npycopy <in.npy> <out.npy>   reads and rewrites one .npy file (CPU only)
//   mhpc_rhs_driver pointer <nside> <outdir>
//        setPointer / resetPointer: one SPH rhs() on a second particle set (an
//        IntegratedParticleHandler holding a shifted copy of the lattice), then one on the
//        particle set itself; both results as snapshots (ts000001: second set, ts000002:
//        the particle set)
//
// File layout of gravity / sph: int64 n, then n doubles per field in the order printed.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "leapfrog.h"
#include "npy_io.h"
#include "predictor_corrector_euler.h"

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static double urand() {   // splitmix64 -> [0, 1)
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static void write_fields(const char *path, int64_t n, const std::vector<const double *> &f,
                         const char *names) {
    FILE *fp = fopen(path, "wb");
    if (!fp) {
        perror(path);
        exit(2);
    }
    fwrite(&n, sizeof(n), 1, fp);
    for (const double *p : f) fwrite(p, sizeof(double), (size_t)n, fp);
    fclose(fp);
    printf("fields: %s\n", names);
}

static int run_gravity(int n, int steps, const char *out) {
    SimulationParameters sp;
    sp.gravitySim = true;
    sp.sphSim = false;
    sp.theta = 0.5;
    sp.smoothing = 1e-6;
    sp.gravityForceVersion = 2;   // v1 acceptance form
    sp.groupWalk = true;          // ... in the group walk
    sp.sfcSelection = Curve::hilbert;
    sp.timeStep = 1e-3;
    sp.timeEnd = 1e-3 * steps;
    sp.calculateEnergy = true;
    sp.calculateAngularMomentum = true;
    Leapfrog sim(sp, n);
    ParticleHandler *ph = sim.particleHandler;
    std::vector<double> x0(n), y0(n), z0(n), vx0(n), vy0(n), vz0(n);
    for (int i = 0; i < n; i++) {
        double x, y, z;
        do {
            x = 2 * urand() - 1;
            y = 2 * urand() - 1;
            z = 2 * urand() - 1;
        } while (x * x + y * y + z * z > 1.0);
        ph->h_x[i] = x0[i] = x;
        ph->h_y[i] = y0[i] = y;
        ph->h_z[i] = z0[i] = z;
        ph->h_vx[i] = vx0[i] = -0.3 * y;
        ph->h_vy[i] = vy0[i] = 0.3 * x;
        ph->h_vz[i] = vz0[i] = 0.0;
        ph->h_mass[i] = 1.0 / n;
    }
    ph->copyDistribution(To::device, true, true);
    for (int s = 0; s < steps; s++) {
        sim.simulationTimeHandler.h_subEndTime = (s + 1) * sp.timeStep;
        sim.integrate(s);
        sim.afterIntegrationStep();
    }
    ph->copyDistribution(To::host, true, true);
    if (ph->lastStatus() != MHPC_OK) return 3;
    printf("energy %.17g angular_momentum %.17g rhs_ms %.3f cells %lld\n", sim.totalEnergy,
           sim.totalAngularMomentum, sim.lastStats().ms_total, (long long)sim.lastStats().num_cells);
    write_fields(out, n,
                 {x0.data(), y0.data(), z0.data(), vx0.data(), vy0.data(), vz0.data(), ph->h_mass,
                  ph->h_x, ph->h_y, ph->h_z, ph->h_vx, ph->h_vy, ph->h_vz, ph->h_g_ax, ph->h_g_ay,
                  ph->h_g_az},
                 "x0 y0 z0 vx0 vy0 vz0 m x y z vx vy vz g_ax g_ay g_az");
    return 0;
}

static int run_sph(int nside, const char *out) {
    const int n = nside * nside * nside;
    SimulationParameters sp;
    sp.gravitySim = false;
    sp.sphSim = true;
    sp.variableSML = true;
    sp.sfcSelection = Curve::lebesgue;
    Material mat;
    mat.interactions = 64;
    mat.eos.type = EquationOfStates::EOS_TYPE_IDEAL_GAS;
    mat.eos.polytropic_gamma = 5.0 / 3.0;
    Leapfrog sim(sp, n, mat);
    ParticleHandler *ph = sim.particleHandler;
    const double h0 = std::cbrt(3.0 * 64 / (4.0 * M_PI * n));
    int i = 0;
    for (int a = 0; a < nside; a++)
        for (int b = 0; b < nside; b++)
            for (int c = 0; c < nside; c++, i++) {
                ph->h_x[i] = -0.5 + (c + 0.5 + 0.3 * (urand() - 0.5)) / nside;
                ph->h_y[i] = -0.5 + (b + 0.5 + 0.3 * (urand() - 0.5)) / nside;
                ph->h_z[i] = -0.5 + (a + 0.5 + 0.3 * (urand() - 0.5)) / nside;
                ph->h_vx[i] = 0.1 * (urand() - 0.5);
                ph->h_vy[i] = 0.1 * (urand() - 0.5);
                ph->h_vz[i] = 0.1 * (urand() - 0.5);
                ph->h_mass[i] = 1.0 / n;
                ph->h_sml[i] = h0;
                ph->h_e[i] = 1.0 + urand();
                ph->h_p[i] = 0.5;
                ph->h_rho[i] = 1.0;
            }
    ph->copyDistribution(To::device, true, false);
    ph->copySPH(To::device);
    std::vector<double> sml0(ph->h_sml, ph->h_sml + n), p0(ph->h_p, ph->h_p + n);
    const real ms = sim.rhs(0, true, true);
    ph->copySPH(To::host);
    ph->copyAcceleration(To::host);
    if (ph->lastStatus() != MHPC_OK) return 3;
    printf("rhs_ms %.3f neighbours %lld\n", ms, (long long)sim.lastStats().n_neighbours);
    write_fields(out, n,
                 {ph->h_x, ph->h_y, ph->h_z, ph->h_vx, ph->h_vy, ph->h_vz, ph->h_mass, sml0.data(),
                  ph->h_e, p0.data(), ph->h_sml, ph->h_rho, ph->h_p, ph->h_cs, ph->h_ax, ph->h_ay,
                  ph->h_az, ph->h_dedt},
                 "x y z vx vy vz m sml0 e p0 sml rho p cs ax ay az dedt");
    return 0;
}

// read a .npy file and write it back (the reader and writer alone; no device needed)
static int run_npycopy(const char *in, const char *out) {
    npy::Array a;
    std::string err = npy::read(in, a);
    if (err.empty()) err = npy::write(out, a.type, a.shape, a.data.data());
    if (!err.empty()) {
        fprintf(stderr, "error: %s\n", err.c_str());
        return 4;
    }
    printf("copied %lld elements, %d dimensions\n", (long long)a.count(), (int)a.shape.size());
    return 0;
}

static int run_io(bool sph, const char *in, const char *outdir, int steps) {
    SimulationParameters sp;
    sp.inputFile = in;
    sp.directory = outdir;
    sp.gravitySim = !sph;
    sp.sphSim = sph;
    sp.variableSML = sph;
    sp.theta = 0.5;
    sp.smoothing = 1e-6;
    sp.sfcSelection = Curve::hilbert;
    sp.gravityForceVersion = 2;   // computeForces_v1: one walk per particle
    sp.timeStep = 1e-3;
    sp.calculateEnergy = !sph;
    sp.calculateAngularMomentum = true;
    Material mat;
    mat.interactions = 64;
    mat.sml = 0.1;
    Miluphpc *sim = nullptr;
    if (sph) sim = new PredictorCorrectorEuler(sp, mat);
    else sim = new Leapfrog(sp, mat);
    sim->rhs(0, true, true);
    sim->afterIntegrationStep();
    sim->particles2file(0);
    for (int s = 0; s < steps; s++) {
        sim->simulationTimeHandler.h_subEndTime = sim->simulationTimeHandler.h_currentTime + sp.timeStep;
        sim->integrate(s + 1);
        sim->afterIntegrationStep();
        sim->particles2file(s + 1);
    }
    printf("particles %d time %.17g\n", sim->numParticles, sim->simulationTimeHandler.h_currentTime);
    delete sim;
    return 0;
}

static int run_pointer(int nside, const char *outdir) {
    const int n = nside * nside * nside;
    SimulationParameters sp;
    sp.gravitySim = false;
    sp.sphSim = true;
    sp.variableSML = true;
    sp.directory = outdir;
    Material mat;
    PredictorCorrectorEuler sim(sp, n, mat);
    ParticleHandler *ph = sim.particleHandler;
    const double h0 = 2.5 / nside;   // about 64 neighbours on the lattice
    int i = 0;
    for (int a = 0; a < nside; a++)
        for (int b = 0; b < nside; b++)
            for (int c = 0; c < nside; c++, i++) {
                ph->h_x[i] = -0.5 + (c + 0.5 + 0.3 * (urand() - 0.5)) / nside;
                ph->h_y[i] = -0.5 + (b + 0.5 + 0.3 * (urand() - 0.5)) / nside;
                ph->h_z[i] = -0.5 + (a + 0.5 + 0.3 * (urand() - 0.5)) / nside;
                ph->h_vx[i] = 0.1 * (urand() - 0.5);
                ph->h_vy[i] = 0.1 * (urand() - 0.5);
                ph->h_vz[i] = 0.1 * (urand() - 0.5);
                ph->h_mass[i] = 1.0 / n;
                ph->h_sml[i] = h0;
                ph->h_e[i] = 1.0 + urand();
                ph->h_p[i] = 0.5;
                ph->h_rho[i] = 1.0;
            }
    // the particle set
    ph->copyDistribution(To::device, true, false);
    ph->copySPH(To::device);
    // the second set: the same particles shifted by +0.25 in x, uploaded through the bound arrays
    IntegratedParticleHandler iph(n, sim.numNodes, sim.context());
    ph->setPointer(&iph);
    if (ph->d_x != iph.d_x || ph->d_sml != iph.d_sml) return 5;
    for (int k = 0; k < n; k++) ph->h_x[k] += 0.25;
    ph->copyDistribution(To::device, true, false);
    ph->copySPH(To::device);
    sim.rhs(0, false, false);
    ph->copyDistribution(To::host, true, false);
    ph->copySPH(To::host);
    sim.particles2file(1);
    ph->resetPointer();
    if (ph->d_x == iph.d_x) return 6;
    // back on the particle set: its arrays were not touched by the rhs() above
    ph->copyDistribution(To::host, true, false);
    ph->copySPH(To::host);
    sim.rhs(0, false, false);
    sim.particles2file(2);
    printf("pointer ok\n");
    return ph->lastStatus() == MHPC_OK ? 0 : 3;
}

int main(int argc, char **argv) {
    try {
        if (argc == 5 && !strcmp(argv[1], "gravity")) return run_gravity(atoi(argv[2]), atoi(argv[3]), argv[4]);
        if (argc == 4 && !strcmp(argv[1], "sph")) return run_sph(atoi(argv[2]), argv[3]);
        if (argc == 6 && !strcmp(argv[1], "io")) return run_io(!strcmp(argv[2], "sph"), argv[3], argv[4], atoi(argv[5]));
        if (argc == 4 && !strcmp(argv[1], "pointer")) return run_pointer(atoi(argv[2]), argv[3]);
        if (argc == 4 && !strcmp(argv[1], "npycopy")) return run_npycopy(argv[2], argv[3]);
    } catch (const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 4;
    }
    fprintf(stderr, "usage: %s gravity <n> <steps> <out> | sph <nside> <out> | io <sph|gravity> <in> <outdir> <steps>"
                    " | pointer <nside> <outdir>\n", argv[0]);
    return 1;
}
