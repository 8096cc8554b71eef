// parameter.h -- host-side types of the drop-in facade.  Names follow the reference's
// include/parameter.h (real/integer/keyType typedefs :12-19, SimulationParameters
// :122-159, struct To :161-171, Execution :189-201, Curve :203-216, Entry :252-271) so
// calling code compiles against either.
#pragma once

#include <string>

typedef double real;
typedef int integer;
typedef unsigned long keyType;
typedef int idInteger;

#define MAX_LEVEL 21
#define DIM 3
#define POW_DIM 8

struct To {
    enum Target { host, device };
    Target t_;
    To(Target t) : t_(t) {}
    operator Target() const { return t_; }
};

namespace Curve {
enum Type { lebesgue = 0, hilbert = 1 };
}

struct Execution {
    enum Location { host, device };
    Location t_;
    Execution(Location t) : t_(t) {}
    operator Location() const { return t_; }
};

struct Entry {
    enum Name { x, y, z, mass };
    Name t_;
    Entry(Name t) : t_(t) {}
    operator Name() const { return t_; }
};

// Runtime parameters.  Fields of the reference that configure the multi-process, I/O or
// load-balancing machinery (outputRank, loadBalancing*, particlesSent2H5, domainListSize,
// removeParticles*) are kept so existing set-up code compiles; they have no effect in a
// single-subdomain build.
struct SimulationParameters {
    std::string directory;
    std::string logDirectory;
    int verbosity = 0;
    bool timeKernels = true;
    int numOutputFiles = 1;
    real timeStep = 1e-3;
    real maxTimeStep = -1.0;
    real timeEnd = 1e-3;
    bool loadBalancing = false;
    int loadBalancingInterval = 0;
    int loadBalancingBins = 0;
    std::string inputFile;
    std::string materialConfigFile;
    int outputRank = 0;
    bool performanceLog = false;
    bool particlesSent2H5 = false;
    int sfcSelection = Curve::lebesgue;
    int integratorSelection = 0;
    // gravity
    real theta = 0.5;
    real smoothing = 0.0;          // eps^2
    // the reference's kernel versions (miluphpc.cpp:2150-2280): 0 and 1 (computeForces_v2,
    // _v2_1) test s^2/theta <= d^2+eps^2, all members of a group of 64 key-sorted particles
    // voting; 2, 3, 4 (computeForces_v1, _v1_1, _v1_2) test (d^2+eps^2) theta^2 > s^2 per
    // particle, each particle walking the tree on its own
    int gravityForceVersion = 0;
    // build extension: versions 2-4 with the group walk (the v1 test, every member of the group
    // voting) instead of one walk per particle
    bool groupWalk = false;
    // SPH
    int smoothingKernelSelection = 1;   // cubic spline
    int sphFixedRadiusNNVersion = 0;
    bool removeParticles = false;
    int removeParticlesCriterion = 0;
    real removeParticlesDimension = 0.0;
    int bins = 0;
    bool calculateAngularMomentum = false;
    bool calculateEnergy = false;
    bool calculateCenterOfMass = false;
    real particleMemoryContingent = 1.0;
    int domainListSize = 0;
    // compile-time switches of parameter.h, runtime here
    bool gravitySim = true;        // GRAVITY_SIM
    bool sphSim = false;           // SPH_SIM
    bool variableSML = false;      // VARIABLE_SML (plain h-iteration)
    // COURANT_FACT / FORCES_FACT of parameter.h:111-113 (compile-time there)
    real courantFactor = 0.4;
    real forcesFactor = 0.2;
    // build extensions
    bool quadrupole = false;       // monopole + quadrupole far field
    int precision = 0;             // 0 FP64 pair sums, 1 FP32 + Kahan
    bool speculativeLists = true;  // neighbour lists from pass 0 of the h-iteration (same sets)
    bool learnedSML = false;       // MLP start for the h-iteration (mhpc_mlp_* must have trained it)
    real smlMax = 0.0;             // upper bound of h in the h-iteration (0: none)
};
