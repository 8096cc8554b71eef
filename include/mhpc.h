/*
 * mhpc.h -- C-ABI of the MI355X-native single-subdomain RHS of milupHPC.
 *
 * One context owns one subdomain on one GPU: device-resident SoA particle fields in the
 * caller's index order, the key-sorted working copies, the octree and the neighbour
 * lists, one hipStream_t and a sticky device error word.  All entry points return an
 * mhpc_status (0 = MHPC_OK) and never exit the process; stage entry points write the
 * elapsed milliseconds of their kernels (hipEvents on the context stream) to `ms`
 * when it is non-NULL, which keeps the reference's "Launch::f(...) returns time"
 * convention.  Contexts are not shared between host threads.
 *
 * Each declaration names the reference interface it replaces (paths relative to the
 * milupHPC source tree).
 */
#ifndef MHPC_H
#define MHPC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHPC_API_VERSION 1

/* Barnes-Hut particle groups: the acceptance test and the interaction list are shared by
   this many consecutive key-sorted particles, one wavefront (the reference's
   computeForces_v2 shares them per warp of 32, gravity.cu:726-903; computeForces_v1 per
   particle).  The accelerations are those of the oracle's group walk with this size. */
#define MHPC_BH_GROUP 64

typedef struct mhpc_ctx mhpc_ctx;

/* Error codes.  Replace gpuErrorcheck's exit (cuda_utilities.cuh:41-46), cudaTerminate
 * (cuda_utilities.cuh:70-74) and the capacity exits (miluphpc.cpp:974-981, 2001-2006,
 * 2049-2055, 2600-2606, 2802-2807). */
typedef enum {
    MHPC_OK = 0,
    MHPC_E_CAP_NODES = 1,        /* more cells than the node capacity */
    MHPC_E_CAP_NEIGHBOURS = 2,   /* a particle has >= max_nb neighbours, or 0 after h-iteration */
    MHPC_E_DEPTH = 3,            /* two particles share the 21-digit key and the octant of the next
                                    bisection (equal positions; tree.cu:949-963) */
    MHPC_E_RHO_NONPOSITIVE = 4,  /* density <= 0 (density.cu:68-70) */
    MHPC_E_STACK = 5,            /* traversal stack overflow */
    MHPC_E_ARG = 6,              /* bad argument */
    MHPC_E_HIP = 7,              /* HIP runtime error */
    MHPC_E_STATE = 8,            /* call order (e.g. forces before density) */
    MHPC_E_SCRATCH = 9           /* more elements than a no-allocation entry point's scratch holds */
} mhpc_status;

/* Space-filling curves (reference Curve::Type, subdomain.cuh). */
enum { MHPC_CURVE_LEBESGUE = 0, MHPC_CURVE_HILBERT = 1 };

/* Equations of state (reference EquationOfStates ids, material.cuh).  100 is a build
 * extension: isothermal gas with a constant caller-given sound speed, p = cs^2 rho. */
enum {
    MHPC_EOS_POLYTROPIC = 0,
    MHPC_EOS_ISOTHERMAL = 3,
    MHPC_EOS_IDEAL_GAS = 9,
    MHPC_EOS_LOCALLY_ISOTHERMAL = 12,
    MHPC_EOS_ISOTHERMAL_CONST_CS = 100
};

/* Barnes-Hut acceptance forms: s/d < theta (computeForces_v1, gravity.cu:252-396) or
 * s^2/theta <= d^2 + eps^2 (computeForces_v2, gravity.cu:726-903); mhpc_params.bh_accept
 * selects who tests. */
enum { MHPC_MAC_CLASSIC = 0, MHPC_MAC_SQRT_THETA = 1 };

/* Who applies the acceptance test (mhpc_params.bh_accept):
 *  BOX: once per group of MHPC_BH_GROUP consecutive key-sorted particles, with the distance
 *       from the cell to the group's bounding box (stricter: opens more cells);
 *  VOTE (default): every member of the group tests the cell with its own distance and the
 *       group accepts it only if all do -- computeForces_v2's warp vote (gravity.cu:836) over
 *       the group;
 *  PER_PARTICLE: computeForces_v1 itself (gravity.cu:252-396): every particle walks the tree
 *       on its own, in the reference's depth-first order, skipping itself; FP64 and monopole
 *       only (mac must be MHPC_MAC_CLASSIC, quadrupole 0; precision is ignored).  The
 *       reference's plain method, slow: one thread per particle. */
enum { MHPC_BH_BOX = 0, MHPC_BH_VOTE = 1, MHPC_BH_PER_PARTICLE = 2 };

/* Pair-sum precision. */
enum { MHPC_PREC_FP64 = 0, MHPC_PREC_FP32_KAHAN = 1 };

/* SPH smoothing kernels, support radius h (reference Smoothing::Kernel, parameter.h:177-180;
 * SPH::SmoothingKernel, kernel.cu:14-183, 3D).  The Wendland C4 gradient is the reference's
 * expression (kernel.cu:140), which is not the derivative of that kernel's W; it is kept
 * as written so that results equal the reference's. */
enum {
    MHPC_KERNEL_SPIKY = 0,
    MHPC_KERNEL_CUBIC_SPLINE = 1,
    MHPC_KERNEL_WENDLAND_C2 = 2,
    MHPC_KERNEL_WENDLAND_C4 = 3,
    MHPC_KERNEL_WENDLAND_C6 = 4
};

/* Runtime parameters.  Field names follow SimulationParameters (parameter.h:122-159)
 * and Material (material.cuh:123-160) where they have a counterpart. */
typedef struct {
    /* gravity (SimulationParameters::theta, ::smoothing = eps^2, ::gravityForceVersion) */
    double theta;
    double smoothing;
    int mac;
    int quadrupole;         /* 0: monopole (reference), 1: monopole + quadrupole */
    double G;               /* 1 in N-body units (SI_UNITS 0) */
    /* keys (SimulationParameters::sfcSelection) */
    int curve;
    /* one material (Material::eos, ::artificialViscosity, ::interactions) */
    int eos;
    double gamma;           /* eos.polytropic_gamma */
    double K;               /* eos.polytropic_K */
    double cs_const;        /* sound speed of MHPC_EOS_ISOTHERMAL_CONST_CS; 0: each particle keeps
                               its uploaded cs (MHPC_F_CS) */
    double alpha, beta;     /* Monaghan artificial viscosity */
    int n_target;           /* Material::interactions (target neighbour count) */
    int band_lo, band_hi;   /* accepted neighbour-count band of the h-iteration */
    int max_counts;         /* counts per particle of the h-iteration (reference: 7) */
    int max_nb;             /* MAX_NUM_INTERACTIONS (parameter.h, 180) */
    int sml_avg;            /* kernel length (h_i + h_j)/2 (VARIABLE_SML || INTEGRATE_SML) */
    int precision;          /* MHPC_PREC_* for gravity and SPH pair sums */
    int cs_before_p;        /* 1: sound speed from the previous p, then p (miluphpc.cpp:2938-2951) */
    double sml_max;         /* build extension: upper bound of h in the h-iteration; 0 = none (the
                               reference).  On a free boundary a particle that cannot reach the
                               band under d < min(h_i, h_j) otherwise grows h on every pass. */
    int bh_accept;          /* MHPC_BH_VOTE (default), MHPC_BH_BOX or MHPC_BH_PER_PARTICLE */
    int smoothing_kernel;   /* MHPC_KERNEL_* (SimulationParameters::smoothingKernelSelection);
                               default MHPC_KERNEL_CUBIC_SPLINE */
} mhpc_params;

/* Field ids for upload / download / device pointers (reference Entry::Name, particles.cuh).
 * All particle fields are in the caller's (original) index order. */
typedef enum {
    MHPC_F_X = 0, MHPC_F_Y, MHPC_F_Z,
    MHPC_F_VX, MHPC_F_VY, MHPC_F_VZ,
    MHPC_F_MASS, MHPC_F_SML, MHPC_F_E, MHPC_F_RHO, MHPC_F_P, MHPC_F_CS,
    MHPC_F_AX, MHPC_F_AY, MHPC_F_AZ,            /* SPH accelerations */
    MHPC_F_G_AX, MHPC_F_G_AY, MHPC_F_G_AZ,      /* gravitational accelerations */
    MHPC_F_G_AX_OLD, MHPC_F_G_AY_OLD, MHPC_F_G_AZ_OLD,
    MHPC_F_DRHODT, MHPC_F_DEDT, MHPC_F_DSMLDT,
    MHPC_F_U,                                   /* gravitational potential energy 1/2 m phi */
    MHPC_F_MUIJMAX,                             /* max mu_ij of the artificial viscosity */
    MHPC_F_CURLV,                               /* |curl v| (learned-h feature) */
    MHPC_F_GRADRHO,                             /* |grad rho| (learned-h feature) */
    MHPC_F_NOI,                                 /* int32: number of interactions */
    MHPC_F_KEY,                                 /* uint64: SFC key of each particle */
    MHPC_F_SORTED,                              /* int32: key-sorted order -> particle index */
    MHPC_F_LEVEL,                               /* int32: leaf level of each particle */
    MHPC_F_H_ITERS,                             /* int32: h-iteration counts of the last RHS */
    MHPC_F_COUNT_
} mhpc_field;

/* RHS flags (GRAVITY_SIM, SPH_SIM, VARIABLE_SML switches of parameter.h). */
enum {
    MHPC_RHS_GRAVITY = 1u << 0,
    MHPC_RHS_SPH = 1u << 1,
    MHPC_RHS_VARIABLE_SML = 1u << 2,   /* plain h-iteration before the neighbour search */
    MHPC_RHS_POTENTIAL = 1u << 3,      /* also accumulate the gravitational potential */
    MHPC_RHS_TIMING = 1u << 4,         /* per-phase hipEvent timing (adds no sync; without it
                                          only ms_total, ms_nn_walk, ms_density and ms_forces
                                          are measured) */
    MHPC_RHS_LEARNED_H = 1u << 5,      /* MLP predicts h; the plain iteration corrects */
    MHPC_RHS_SPECULATIVE_NN = 1u << 6, /* pass 0 of the h-iteration writes the neighbour
                                          lists; only particles whose list can change are
                                          re-listed (same sets; implied by LEARNED_H) */
    MHPC_RHS_FEATURES = 1u << 7        /* learned-h features only (curl v, grad rho, MLP
                                          inputs), for plain steps that produce training
                                          pairs (implied by LEARNED_H) */
};

typedef struct {
    float ms_total;
    float ms_bbox, ms_keys, ms_sort, ms_gather, ms_tree, ms_com;
    float ms_gravity, ms_mlp, ms_hiter, ms_nn, ms_density, ms_eos, ms_forces, ms_scatter;
    int64_t num_cells;
    int64_t n_cell_interactions;       /* BH: particle-cell interactions */
    int64_t n_part_interactions;       /* BH: particle-particle interactions */
    int64_t n_neighbours;              /* sum of noi */
    int64_t n_h_counts;                /* neighbour counts done by the h-iteration */
    int32_t max_noi;
    int32_t h_passes;                  /* h-iteration passes that counted at least one particle */
    float ms_nn_walk;                  /* the h-iteration's pass-0 group walk kernel alone (part of
                                          ms_hiter) */
} mhpc_rhs_stats;

/* -------- context (replaces the Miluphpc ctor allocations, miluphpc.cpp:3-93, and the
 * ParticleHandler / TreeHandler / HelperHandler allocations, particle_handler.cpp:3-205,
 * tree_handler.cpp:3-41, helper_handler.cpp) -------- */
void mhpc_default_params(mhpc_params *p);
int mhpc_create(const mhpc_params *p, int64_t cap_particles, int64_t cap_cells, mhpc_ctx **out);
int mhpc_destroy(mhpc_ctx *ctx);
int mhpc_set_params(mhpc_ctx *ctx, const mhpc_params *p);
int mhpc_get_params(mhpc_ctx *ctx, mhpc_params *p);
int mhpc_set_num_particles(mhpc_ctx *ctx, int64_t n);
int64_t mhpc_num_particles(mhpc_ctx *ctx);
const char *mhpc_last_error(mhpc_ctx *ctx);
const char *mhpc_status_string(int status);
int mhpc_synchronize(mhpc_ctx *ctx);

/* -------- data movement (ParticleHandler::copy*(To::Target), particle_handler.h:391-455) -- */
int mhpc_upload(mhpc_ctx *ctx, int field, const void *host, int64_t n);
int mhpc_download(mhpc_ctx *ctx, int field, void *host, int64_t n);
int mhpc_field_ptr(mhpc_ctx *ctx, int field, void **device_ptr);
/* Point an FP64 field at a caller-owned device array of at least the context's capacity
 * (device_ptr NULL: back to the context's own array); every later stage reads and writes that
 * array.  ParticleHandler::setPointer / resetPointer (particle_handler.cpp:742-900), which the
 * predictor-corrector integrator uses to run rhs() on the predicted particle set
 * (predictor_corrector_euler.cpp:119-133). */
int mhpc_bind_field(mhpc_ctx *ctx, int field, void *device_ptr);
/* Device memory for such arrays (IntegratedParticleHandler's allocations,
 * particle_handler.cpp:1076-1230): zeroed; mhpc_device_free refuses an array still bound. */
int mhpc_device_alloc(mhpc_ctx *ctx, int64_t bytes, void **device_ptr);
int mhpc_device_free(mhpc_ctx *ctx, void *device_ptr);
/* Blocking copies between host memory and such device memory, after the context stream's work
 * (and, being null-stream copies, after earlier null-stream work) -- the cudaMemcpy calls of
 * IntegratedParticleHandler's users. */
int mhpc_device_upload(mhpc_ctx *ctx, void *device_ptr, const void *host, int64_t bytes);
int mhpc_device_download(mhpc_ctx *ctx, void *host, const void *device_ptr, int64_t bytes);
/* device-to-device copy of n elements into a field, ordered on the context stream (for
 * callers whose inputs already live in HBM, e.g. restoring a state between RHS calls) */
int mhpc_upload_device(mhpc_ctx *ctx, int field, const void *device_src, int64_t n);
/* neighbour lists in the reference layout nnl[i * max_nb + k] (original indices) */
int mhpc_download_neighbours(mhpc_ctx *ctx, int32_t *nnl, int32_t *noi, int64_t n, int max_nb);
/* octree in key-sorted numbering: particles [0,n) (sorted positions), cells n + c */
int mhpc_tree_size(mhpc_ctx *ctx, int64_t *num_cells);
int mhpc_download_tree(mhpc_ctx *ctx, int32_t *cell_level, int32_t *cell_start,
                       int32_t *cell_end, int32_t *cell_child, double *cell_mass,
                       double *cell_x, double *cell_y, double *cell_z, double *cell_quad);
int mhpc_get_box(mhpc_ctx *ctx, double box[6]);

/* -------- RHS: Miluphpc::rhs(step, selfGravity, assign) (miluphpc.h:478,
 * miluphpc.cpp:387-562) -------- */
int mhpc_rhs(mhpc_ctx *ctx, uint32_t flags, mhpc_rhs_stats *stats);

/* -------- stages (the Kernel::Launch wrappers of the reference) -------- */
/* TreeNS::Kernel::Launch::computeBoundingBox (tree.cu:1949-1955) */
int mhpc_bbox(mhpc_ctx *ctx, double box[6], float *ms);
/* SubDomainKeyTreeNS::Kernel::Launch::getParticleKeys (subdomain.cu:1564-1570) */
int mhpc_keys(mhpc_ctx *ctx, float *ms);
/* TreeNS::Kernel::Launch::prepareSorting + HelperNS::sortArray (tree.cu:1404-1413,
 * helper.cu:150-167) */
int mhpc_sort(mhpc_ctx *ctx, float *ms);
/* TreeNS::Kernel::Launch::buildTree (tree.cu:1929-1937), on the key-sorted order */
int mhpc_build_tree(mhpc_ctx *ctx, float *ms);
/* TreeNS::Kernel::Launch::calculateCentersOfMass (tree.cu:1416-1497) for all levels */
int mhpc_centres_of_mass(mhpc_ctx *ctx, float *ms);
/* Gravity::Kernel::Launch::computeForces_v2 (gravity.cu:2652-2663) */
int mhpc_gravity(mhpc_ctx *ctx, int potential, float *ms);
/* SPH::Kernel::Launch::fixedRadiusNN_variableSML (sph.cu:3128-3135) */
int mhpc_h_iterate(mhpc_ctx *ctx, float *ms);
/* SPH::Kernel::Launch::fixedRadiusNN (sph.cu:3094-3101) */
int mhpc_neighbours(mhpc_ctx *ctx, float *ms);
/* SPH::Kernel::Launch::calculateDensity (density.cu:74-79) */
int mhpc_density(mhpc_ctx *ctx, float *ms);
/* SPH::Kernel::Launch::calculateSoundSpeed + calculatePressure (soundspeed.cu,
 * pressure.cu:93) */
int mhpc_eos(mhpc_ctx *ctx, float *ms);
/* SPH::Kernel::Launch::internalForces (internal_forces.cu:671-676) */
int mhpc_internal_forces(mhpc_ctx *ctx, float *ms);

/* -------- global sums: Miluphpc::energy / angularMomentum (miluphpc.cpp:577-677),
 * Physics::Kernel::kineticEnergy and calculateAngularMomentumBlockwise
 * (subdomain.cu:2428-2619); FP64, compensated, deterministic order -------- */
typedef struct {
    double kinetic;                /* sum 1/2 m v^2 */
    double potential;              /* sum u (u = 1/2 m phi from mhpc_rhs with MHPC_RHS_POTENTIAL) */
    double internal;               /* sum m e */
    double total;                  /* kinetic + potential + internal */
    double momentum[3];            /* sum m v */
    double angular_momentum[3];    /* sum m x cross v */
    double angular_momentum_norm;  /* |L| (the reference's totalAngularMomentum) */
    double mass;
    double com[3];                 /* centre of mass */
} mhpc_diagnostics;
int mhpc_compute_diagnostics(mhpc_ctx *ctx, mhpc_diagnostics *out);

/* -------- key sort without allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  mhpc_sort's result (a stable sort by key, ties in ascending
 * position: cub::DeviceRadixSort::SortPairs over keys and 0..n-1, helper.cu:150-167) on
 * caller-owned device arrays: keys[n] in, sorted_keys[n] and sorted[n] (the permutation) out,
 * queued on `stream` (a hipStream_t; NULL = the null stream) without synchronising.  It calls
 * no allocator: its working memory is one static array per GPU of about 26 MiB, held until
 * the process ends, so n is at most MHPC_NOALLOC_SORT_MAX (else MHPC_E_SCRATCH, nothing
 * queued).  Every call on one GPU uses that GPU's array: calls that overlap in time on one GPU
 * (different streams, host threads) race and must be serialised by the caller. */
#define MHPC_NOALLOC_SORT_MAX 2097152
int mhpc_sort_keys_noalloc(const uint64_t *keys, int64_t n, uint64_t *sorted_keys, int32_t *sorted,
                           void *stream);

/* -------- predictor-corrector Euler (PredictorCorrectorEulerNS, predictor_corrector_euler.cpp,
 * device_predictor_corrector_euler.cu:130-540) -------- */
/* CFL time step (Kernel::Launch::setTimeStep): dt = min(courant_fact * min h/cs,
 * forces_fact * min sqrt(h/|a+g|), courant_fact * min h/(cs + 1.2(alpha cs + beta mu_max)),
 * max h / max |v|, dt_max if > 0).  parts (may be NULL) receives the five terms. */
int mhpc_time_step(mhpc_ctx *ctx, double courant_fact, double forces_fact, double dt_max,
                   double *dt, double parts[5]);
/* Kernel::Launch::predictor: saves the state and writes the predicted one into the fields */
int mhpc_pc_predictor(mhpc_ctx *ctx, double dt);
/* Kernel::Launch::corrector: combines the saved state with the predicted derivatives */
int mhpc_pc_corrector(mhpc_ctx *ctx, double dt);
/* One sub-step of PredictorCorrectorEuler::integrate (predictor_corrector_euler.cpp): rhs
 * (flags), setTimeStep -- dt = min(courant_fact q1, forces_fact q0, courant_fact q2, max h /
 * max |v|, dt_max if > 0, t_left if > 0) as mhpc_time_step -- predictor, rhs (flags without
 * MHPC_RHS_GRAVITY), corrector; the same results as those five calls in turn, queued without
 * waiting in between (dt is computed and used on the device).  *dt receives the step; stats0 /
 * stats1 (may be NULL) the two RHS's statistics.  Flags with MHPC_RHS_LEARNED_H or
 * MHPC_RHS_FEATURES are refused (MHPC_E_ARG): use the separate calls. */
int mhpc_pc_euler_step(mhpc_ctx *ctx, uint32_t flags, double courant_fact, double forces_fact, double dt_max,
                       double t_left, double *dt, mhpc_rhs_stats *stats0, mhpc_rhs_stats *stats1);

/* -------- learned smoothing length (build extension, no reference counterpart): a
 * 7 -> 64 -> 64 -> 1 tanh MLP, BF16 weights on MFMA with FP32 accumulation, predicting
 * log(h_converged / h_previous); the plain h-iteration starts from the prediction and
 * has the last word (MHPC_RHS_LEARNED_H).  Parameters: flat FP32 vector of
 * MHPC_MLP_NPARAM = W1[64][8] (input 8 is padding), b1[64], W2[64][64], b2[64], W3[64], b3.
 * norm[18]: feature means [0..6], feature stds [8..14], target mean [16], target std [17]. */
#define MHPC_MLP_NPARAM 4801
int mhpc_mlp_init(mhpc_ctx *ctx, uint64_t seed);
int mhpc_mlp_set_params(mhpc_ctx *ctx, const float *params, int nparam);
int mhpc_mlp_get_params(mhpc_ctx *ctx, float *params, int nparam);
int mhpc_mlp_set_norm(mhpc_ctx *ctx, const float norm[18]);
int mhpc_mlp_get_norm(mhpc_ctx *ctx, float norm[18]);
/* features of every particle from the current fields (the previous RHS's outputs) and
 * the time step dt; also saves h as h_previous.  Raw (unnormalised) features if
 * norm has not been set. */
int mhpc_mlp_features(mhpc_ctx *ctx, double dt);
int mhpc_mlp_download_features(mhpc_ctx *ctx, float *X, int64_t n);
/* sml <- h_previous * exp(prediction) (clamped to a factor of 4) */
int mhpc_mlp_predict(mhpc_ctx *ctx, float *ms);
/* `iters` Adam steps on random subsets of `nsamples` pairs (features, log(sml /
 * h_previous)) -- the pairs the plain h-iteration just produced */
int mhpc_mlp_train(mhpc_ctx *ctx, int nsamples, int iters, float lr, uint64_t seed, float *loss);
/* test hooks: forward / loss gradient on caller-given normalised inputs and targets */
int mhpc_mlp_forward(mhpc_ctx *ctx, const float *X, int64_t n, float *Y);
int mhpc_mlp_gradient(mhpc_ctx *ctx, const float *X, const float *T, int64_t n, float *grad,
                      float *loss);

/* -------- integrator kernels (LeapfrogNS::Kernel::Launch::updateX / updateV,
 * device_leapfrog.cu:4-58) -------- */
int mhpc_leapfrog_update_x(mhpc_ctx *ctx, double dt);
int mhpc_leapfrog_update_v(mhpc_ctx *ctx, double dt);

#ifdef __cplusplus
}
#endif
#endif
