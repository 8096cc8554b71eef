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
        q[1].add(u[i]);
        q[2].add(mi * e[i]);
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

__global__ void k_diag_final(const double *__restrict__ partial, int nb, double *__restrict__ out) {
    const int k = threadIdx.x;
    if (k >= DIAG_Q) return;
    Neu acc;
    for (int b = 0; b < nb; b++) {
        acc.add(partial[(int64_t)b * 2 * DIAG_Q + 2 * k]);
        acc.add(partial[(int64_t)b * 2 * DIAG_Q + 2 * k + 1]);
    }
    out[k] = acc.value();
}

int diag_num_quantities() { return DIAG_Q; }
int diag_partial_size() { return DIAG_BLOCKS * 2 * DIAG_Q; }

void launch_diagnostics(mhpc_ctx *c, double *partial, double *out) {
    int nb = (int)ceil_div(c->n, 256 * 4);
    if (nb < 1) nb = 1;
    if (nb > DIAG_BLOCKS) nb = DIAG_BLOCKS;
    double **f = c->f;
    hipLaunchKernelGGL(k_diag_partial, dim3(nb), dim3(256), 0, c->stream, c->n, f[MHPC_F_X], f[MHPC_F_Y],
                       f[MHPC_F_Z], f[MHPC_F_VX], f[MHPC_F_VY], f[MHPC_F_VZ], f[MHPC_F_MASS], f[MHPC_F_U],
                       f[MHPC_F_E], partial);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, c->stream, partial, nb, out);
}
