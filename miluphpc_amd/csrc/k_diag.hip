// k_diag.hip -- global sums (Miluphpc::energy / angularMomentum, miluphpc.cpp:577-677;
// Physics::Kernel::kineticEnergy and calculateAngularMomentumBlockwise, subdomain.cu:
// 2428-2619).  FP64 with Neumaier compensation in every partial and in the combine, and a
// fixed two-stage order (per-block partials, then one block folds them in block order),
// so the result is deterministic.  The reference's blockwise reduction reads past n and
// reduces the input array instead of its shared buffer (subdomain.cu:2324-2393); this
// one does not inherit that.
#include "mhpc_internal.h"

#define DIAG_Q 14          // quantities
#define DIAG_BLOCKS 1024
#define DIAG_PARTIAL (DIAG_BLOCKS * 2 * DIAG_Q)   // doubles of per-block partials

struct Neu {
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double x) {
        const double t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x; else c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ void add(const Neu &o) {
        add(o.s);
        add(o.c);
    }
    __device__ __forceinline__ double value() const { return s + c; }
};

// quantities: 0 kinetic, 1 potential (sum u), 2 internal (sum m e), 3-5 momentum,
// 6-8 angular momentum sum m x cross v, 9 mass, 10-12 sum m x, 13 reserved
__global__ __launch_bounds__(256) void k_diag_partial(int64_t n, const double *__restrict__ x,
                                                      const double *__restrict__ y,
                                                      const double *__restrict__ z,
                                                      const double *__restrict__ vx,
                                                      const double *__restrict__ vy,
                                                      const double *__restrict__ vz,
                                                      const double *__restrict__ m,
                                                      const double *__restrict__ u,
                                                      const double *__restrict__ e,
                                                      double *__restrict__ partial) {
    Neu q[DIAG_Q];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double mi = m[i], xi = x[i], yi = y[i], zi = z[i];
        const double a = vx[i], b = vy[i], cz = vz[i];
        q[0].add(0.5 * mi * (a * a + b * b + cz * cz));
        q[1].add(u ? u[i] : 0.0);
        q[2].add(e ? mi * e[i] : 0.0);
        q[3].add(mi * a);
        q[4].add(mi * b);
        q[5].add(mi * cz);
        q[6].add(mi * (yi * cz - zi * b));
        q[7].add(mi * (zi * a - xi * cz));
        q[8].add(mi * (xi * b - yi * a));
        q[9].add(mi);
        q[10].add(mi * xi);
        q[11].add(mi * yi);
        q[12].add(mi * zi);
    }
    __shared__ double ss[DIAG_Q][256], sc[DIAG_Q][256];
    for (int k = 0; k < DIAG_Q; k++) {
        ss[k][threadIdx.x] = q[k].s;
        sc[k][threadIdx.x] = q[k].c;
    }
    __syncthreads();
    if (threadIdx.x < DIAG_Q) {
        const int k = threadIdx.x;
        Neu acc;
        for (int t = 0; t < (int)blockDim.x; t++) {
            acc.add(ss[k][t]);
            acc.add(sc[k][t]);
        }
        partial[(int64_t)blockIdx.x * 2 * DIAG_Q + 2 * k] = acc.s;
        partial[(int64_t)blockIdx.x * 2 * DIAG_Q + 2 * k + 1] = acc.c;
    }
}

// folds the block partials in block order and writes the diagnostics record (the derived
// quantities -- total, |L|, centre of mass -- with the same expressions for every caller)
__global__ void k_diag_final(const double *__restrict__ partial, int nb, mhpc_diagnostics *__restrict__ out) {
    __shared__ double q[DIAG_Q];
    const int k = threadIdx.x;
    if (k < DIAG_Q) {
        Neu acc;
        for (int b = 0; b < nb; b++) {
            acc.add(partial[(int64_t)b * 2 * DIAG_Q + 2 * k]);
            acc.add(partial[(int64_t)b * 2 * DIAG_Q + 2 * k + 1]);
        }
        q[k] = acc.value();
    }
    __syncthreads();
    if (k != 0) return;
    mhpc_diagnostics d;
    d.kinetic = q[0];
    d.potential = q[1];
    d.internal = q[2];
    d.total = q[0] + q[1] + q[2];
    for (int a = 0; a < 3; a++) {
        d.momentum[a] = q[3 + a];
        d.angular_momentum[a] = q[6 + a];
        d.com[a] = q[9] > 0.0 ? q[10 + a] / q[9] : 0.0;
    }
    d.angular_momentum_norm = sqrt(q[6] * q[6] + q[7] * q[7] + q[8] * q[8]);
    d.mass = q[9];
    *out = d;
}

static int diag_blocks(int64_t n) {
    int nb = (int)ceil_div(n, 256 * 4);
    if (nb < 1) nb = 1;
    if (nb > DIAG_BLOCKS) nb = DIAG_BLOCKS;
    return nb;
}

// device doubles the context keeps for the record (at least the five time-step terms)
int diag_num_quantities() { return (int)(sizeof(mhpc_diagnostics) / sizeof(double)) > 5
                                       ? (int)(sizeof(mhpc_diagnostics) / sizeof(double)) : 5; }
int diag_partial_size() { return DIAG_PARTIAL; }

static void diag_launch(hipStream_t stream, int64_t n, const double *x, const double *y, const double *z,
                        const double *vx, const double *vy, const double *vz, const double *m,
                        const double *u, const double *e, double *partial, mhpc_diagnostics *out) {
    const int nb = diag_blocks(n);
    hipLaunchKernelGGL(k_diag_partial, dim3(nb), dim3(256), 0, stream, n, x, y, z, vx, vy, vz, m, u, e, partial);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, stream, partial, nb, out);
}

void launch_diagnostics(mhpc_ctx *c, double *partial, mhpc_diagnostics *out) {
    double **f = c->f;
    diag_launch(c->stream, c->n, f[MHPC_F_X], f[MHPC_F_Y], f[MHPC_F_Z], f[MHPC_F_VX], f[MHPC_F_VY],
                f[MHPC_F_VZ], f[MHPC_F_MASS], f[MHPC_F_U], f[MHPC_F_E], partial, out);
}
