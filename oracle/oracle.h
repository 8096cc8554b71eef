/*
 * oracle.h -- CPU restatement (FP64, plain C + OpenMP) of milupHPC's single-subdomain
 * right-hand side, used ONLY as test infrastructure: the checker of the HIP path in
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  Nothing in
 * the product (miluphpc_amd/) links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * milupHPC source tree).  Functions that no golden vector, closed form, invariant or
 * brute-force check pins are marked "parity unpinned" in their comment; at present that
 * is only the gradient of the Wendland C4 kernel (orc_sph_kernel, ORC_KERNEL_WENDLAND_C4),
 * which restates the reference's expression but is not the derivative of its W; everything
 * else is pinned by the tests under tests/test_oracle_*.py.
 *
 * Node numbering of the oracle's octree: indices [0, n) are particles, [n, n+ncells)
 * are cells, and n is the root.  (The reference aliases the root with particle 0's
 * child slots, tree.cu:768-1082; that is a storage artefact, not semantics.)
 */
#ifndef MHPC_ORACLE_H
#define MHPC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_E_CAP_NODES = 1,
    ORC_E_CAP_NEIGHBOURS = 2,
    ORC_E_DEPTH = 3,
    ORC_E_RHO_NONPOSITIVE = 4,
    ORC_E_STACK = 5,
    ORC_E_ARG = 6
};

/* curve types (reference Curve::Type) */
enum { ORC_CURVE_LEBESGUE = 0, ORC_CURVE_HILBERT = 1 };

/* EOS types: reference EquationOfStates ids (material.cuh); 100 = isothermal gas with a
 * constant, caller-given sound speed (p = cs^2 rho), a build extension. */
enum { ORC_EOS_POLYTROPIC = 0, ORC_EOS_ISOTHERMAL = 3, ORC_EOS_IDEAL_GAS = 9,
       ORC_EOS_LOCALLY_ISOTHERMAL = 12, ORC_EOS_ISOTHERMAL_CONST_CS = 100 };

int orc_num_threads(void);
void orc_set_num_threads(int n);

/* ---- a1: bounding box (tree.cu:302-437, seeded with 0 by device_rhs.cu:25-32) ---- */
void orc_bbox(int64_t n, const double *x, const double *y, const double *z, double box[6]);

/* ---- a2: SFC keys (tree.cu:157-240, tree.cu:40-53, tables tree.cuh:52-118) ---- */
uint64_t orc_lebesgue_key(double x, double y, double z, const double box[6]);
uint64_t orc_lebesgue2hilbert(uint64_t lebesgue);
void orc_keys(int64_t n, const double *x, const double *y, const double *z,
              const double box[6], int curve, uint64_t *keys);

/* ---- a3: stable key sort (helper.cu:150-167 = cub SortPairs) ---- */
void orc_sort(int64_t n, const uint64_t *keys, int32_t *perm);

/* ---- a5: reference insertion octree (tree.cu:768-1082), serial ---- */
int orc_tree_build(int64_t n, const double *x, const double *y, const double *z,
                   const double box[6], int64_t cap_cells,
                   int32_t *child, int32_t *level, double *nx, double *ny, double *nz,
                   int64_t *ncells);

/* ---- a6: centres of mass, level by level (tree.cu:1416-1497, miluphpc.cpp:1205-1210) ---- */
void orc_com(int64_t n, int64_t ncells, const int32_t *child, const int32_t *level,
             double *nm, double *nx, double *ny, double *nz);

/* traceless quadrupole moments about each cell's COM (build extension for the
 * mixed-precision far field), q = [xx, xy, xz, yy, yz, zz] per node */
void orc_quadrupole(int64_t n, int64_t ncells, const int32_t *child, const int32_t *level,
                    const double *nm, const double *nx, const double *ny, const double *nz,
                    double *q);

/* ---- a7 / a7': Barnes-Hut ---- */
/* per-particle MAC, reference computeForces_v1 (gravity.cu:252-396) */
int orc_bh_v1(int64_t n, int64_t ncells, const int32_t *child, const double *nm,
              const double *nx, const double *ny, const double *nz, double radius,
              double theta, double smoothing, double *ax, double *ay, double *az, double *u);
/* group MAC over `group` consecutive particles of the key order: the semantics the
 * wavefront-cooperative HIP walk implements.  mac bit 0: 0 = s^2 < theta^2 (d^2+eps^2)
 * (v1 form, gravity.cu:319-330), 1 = s^2/theta <= d^2+eps^2 (v2 form, gravity.cu:798-841).
 * mac bit 1: 0 = d is the distance from the cell's COM to the group's bounding box (one
 * test per group); 1 = vote: every member tests the cell with its own distance and the
 * group accepts only if all members do (the v2 warp vote, gravity.cu:836, over the
 * group).  q may be NULL (monopole only). */
int orc_bh_group(int64_t n, const int32_t *perm, int group, int64_t ncells,
                 const int32_t *child, const int32_t *level, const double *nm,
                 const double *nx, const double *ny, const double *nz, const double *q,
                 const double box[6], double theta, double smoothing, int mac,
                 double *ax, double *ay, double *az, double *phi,
                 int64_t *n_cell_int, int64_t *n_part_int);
/* the group walk over the listed groups gsel[0..ngsel) only (group g = sorted positions
 * [g*group, (g+1)*group)); the other particles' outputs are left untouched */
int orc_bh_group_subset(int64_t n, const int32_t *perm, int group, const int32_t *child,
                        const int32_t *level, const double *nm, const double *nx, const double *ny,
                        const double *nz, const double *q, const double box[6], double theta,
                        double smoothing, int mac, int64_t ngsel, const int64_t *gsel, double *ax,
                        double *ay, double *az, double *phi, int64_t *n_cell_int, int64_t *n_part_int);
/* direct summation on a subset of targets (FP64, softened) */
void orc_direct(int64_t n, const double *x, const double *y, const double *z,
                const double *m, double smoothing, int64_t ntarget, const int32_t *targets,
                double *ax, double *ay, double *az, double *phi);

/* ---- a8: neighbour search (sph.cu:93-281); lists are n x max_nb, row-major; pair test
 * d^2 < h_i^2 and d^2 < h_j^2 with d^2 = fma(dx, dx, fma(dy, dy, dz*dz)) ---- */
int orc_nn_brute(int64_t n, const double *x, const double *y, const double *z,
                 const double *h, int max_nb, int32_t *nnl, int32_t *noi);
int orc_nn_grid(int64_t n, const double *x, const double *y, const double *z,
                const double *h, int max_nb, int32_t *nnl, int32_t *noi);
int orc_nn_refwalk(int64_t n, const double *x, const double *y, const double *z,
                   const double *h, const int32_t *child, const double *nx,
                   const double *ny, const double *nz, double diam, int max_nb,
                   int32_t *nnl, int32_t *noi);

/* ---- a9: plain h-iteration, Jacobi form of sph.cu:739-961; sml_max > 0 bounds h (a
 * build extension; 0 is the reference) ---- */
void orc_h_factor_table(int n_target, int tmax, double *table);
int orc_h_iterate(int64_t n, const double *x, const double *y, const double *z, double *h,
                  int n_target, int band_lo, int band_hi, int max_counts, int max_nb,
                  const double *table, int tmax, double sml_max, int32_t *iters, int32_t *noi);

/* ---- a10-a14: SPH (FP64) ---- */
/* smoothing kernels: the reference's Smoothing::Kernel (parameter.h:177-180) */
enum { ORC_KERNEL_SPIKY = 0, ORC_KERNEL_CUBIC_SPLINE = 1, ORC_KERNEL_WENDLAND_C2 = 2,
       ORC_KERNEL_WENDLAND_C4 = 3, ORC_KERNEL_WENDLAND_C6 = 4 };
void orc_cubic_spline(double r2, double dx, double dy, double dz, double h, double *W,
                      double dWdx[3], double *dWdr);
int orc_sph_kernel(int kind, double r2, double dx, double dy, double dz, double h, double *W,
                   double dWdx[3], double *dWdr);
int orc_density(int64_t n, const double *x, const double *y, const double *z,
                const double *m, const double *h, const int32_t *nnl, const int32_t *noi,
                int max_nb, int sml_avg, int kernel, double *rho);
void orc_soundspeed(int64_t n, int eos, double gamma, double K, const double *rho,
                    const double *p, double *cs);
void orc_pressure(int64_t n, int eos, double gamma, double K, const double *rho,
                  const double *e, const double *cs, double *p);
void orc_forces(int64_t n, const double *x, const double *y, const double *z,
                const double *vx, const double *vy, const double *vz, const double *m,
                const double *h, const double *rho, const double *p, const double *cs,
                const int32_t *nnl, const int32_t *noi, int max_nb, int sml_avg, int kernel,
                double alpha, double beta, double *ax, double *ay, double *az,
                double *drhodt, double *dedt, double *dsmldt, double *muijmax);

/* ---- a17: leapfrog kick-drift (device_leapfrog.cu:4-58) ---- */
void orc_leapfrog_update_x(int64_t n, double dt, double *x, double *y, double *z,
                           const double *vx, const double *vy, const double *vz,
                           const double *gx, const double *gy, const double *gz,
                           double *gx_old, double *gy_old, double *gz_old);
void orc_leapfrog_update_v(int64_t n, double dt, double *vx, double *vy, double *vz,
                           const double *gx, const double *gy, const double *gz,
                           const double *gx_old, const double *gy_old, const double *gz_old);

/* ---- predictor-corrector Euler (device_predictor_corrector_euler.cu:130-540) ----
 * s: x,y,z,vx,vy,vz,ax,ay,az,g_ax,g_ay,g_az,e,dedt (14 arrays); b: 9 saved arrays
 * x,y,z,vx,vy,vz,ax,ay,az + e,dedt (11) */
void orc_pc_predictor(int64_t n, double dt, double **s, double **b);
void orc_pc_corrector(int64_t n, double dt, double **s, double **b);
/* terms[5]: min sqrt(h/|a+g|), min h/cs, min h/(cs+1.2(alpha cs+beta mu)), max|v|, max h */
void orc_time_step_terms(int64_t n, const double *ax, const double *ay, const double *az,
                         const double *gx, const double *gy, const double *gz, const double *h,
                         const double *cs, const double *mu, const double *vx, const double *vy,
                         const double *vz, double alpha, double beta, double terms[5]);

/* ---- learned smoothing length ----
 * local-state features over the neighbour lists (no reference counterpart):
 * curl[3i..] = sum_j (m_j/rho_j) grad_i W_ij x (v_j - v_i),
 * gradrho[3i..] = sum_j m_j grad_i W_ij, W with the same (averaged) h as the forces */
void orc_sph_features(int64_t n, const double *x, const double *y, const double *z,
                      const double *vx, const double *vy, const double *vz, const double *m,
                      const double *h, const double *rho, const int32_t *nnl, const int32_t *noi,
                      int max_nb, int sml_avg, int kernel, double *curl, double *gradrho);
/* 7 -> H -> H -> 1 tanh MLP, FP64 forward/backward */
void orc_mlp_forward(int64_t nb, int nin, int nh, const double *X, const double *W1,
                     const double *b1, const double *W2, const double *b2, const double *W3,
                     double b3, double *Y, double *A1, double *A2);
double orc_mlp_backward(int64_t nb, int nin, int nh, const double *X, const double *T,
                        const double *W1, const double *b1, const double *W2,
                        const double *b2, const double *W3, double b3,
                        double *gW1, double *gb1, double *gW2, double *gb2, double *gW3,
                        double *gb3);

/* ---- host-core baseline: the composite passes above timed at C level (no Python in the
 * timed region), on the caller's arrays ---- */
/* One full SPH RHS per rep: h-iteration from h0 (h bounded by sml_max when > 0), neighbour
 * lists, density, sound speed and pressure (cs before p), internal forces, FP64.
 * seconds[r] = wall time of rep r. */
int orc_bench_sph_rhs(int64_t n, const double *x, const double *y, const double *z,
                      const double *vx, const double *vy, const double *vz, const double *m,
                      const double *h0, const double *e, const double *p0, const double *cs0,
                      int n_target, int band_lo, int band_hi, int max_counts, int max_nb, double sml_max,
                      int eos, double gamma, double K, double alpha, double beta, int reps, double *seconds);
/* Barnes-Hut RHS: bounding box, keys, sort, octree, centres of mass (+ quadrupole) over all
 * particles (seconds[0]), then the group walk over every `sample_every`-th group (seconds[1],
 * groups walked in *walked, interactions in counts[2] = {cell, particle}). */
int orc_bench_gravity(int64_t n, const double *x, const double *y, const double *z, const double *m,
                      double theta, double smoothing, int curve, int mac, int quad, int group,
                      int64_t sample_every, double seconds[2], int64_t *walked, int64_t counts[2]);

#ifdef __cplusplus
}
#endif
#endif
