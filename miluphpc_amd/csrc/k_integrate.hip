// k_integrate.hip -- predictor-corrector Euler and its CFL time step
// (PredictorCorrectorEulerNS::Kernel::predictor / corrector / setTimeStep,
// device_predictor_corrector_euler.cu:130-540), FP64 on the caller-order fields.
//
// The reference keeps the predicted state in a second particle set and swaps pointers
// (particle_handler.h:404-406).  Here the state at t is saved to backup arrays, the
// predicted state is written into the live fields (so rhs() runs on them unchanged), and
// the corrector combines the two.
#include "mhpc_internal.h"

// backup slots: what the corrector reads (h, rho and their rates are not integrated:
// INTEGRATE_SML and INTEGRATE_DENSITY are 0, so the predicted state's values stay live)
enum { PB_X, PB_Y, PB_Z, PB_VX, PB_VY, PB_VZ, PB_AX, PB_AY, PB_AZ, PB_E, PB_DEDT, PB_COUNT_ };

struct PCArgs {
    int64_t n;
    double dt;
    const double *dt_dev;   // when set, the time step is read from here (computed on the device)
    double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az, *gx, *gy, *gz;
    double *e, *dedt;
    double *b[PB_COUNT_];
};

// predictor (device_predictor_corrector_euler.cu:239-291): save the state, then
// x* = x + dt v, v* = v + dt (a + g), e* = max(e + dt de/dt, 1e-6); h, rho, p, cs carried.
__global__ __launch_bounds__(256) void k_pc_predictor(PCArgs A) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const double dt = A.dt_dev ? *A.dt_dev : A.dt;
    const double x = A.x[i], y = A.y[i], z = A.z[i], vx = A.vx[i], vy = A.vy[i], vz = A.vz[i];
    const double ax = A.ax[i], ay = A.ay[i], az = A.az[i], e = A.e[i], dedt = A.dedt[i];
    A.b[PB_X][i] = x; A.b[PB_Y][i] = y; A.b[PB_Z][i] = z;
    A.b[PB_VX][i] = vx; A.b[PB_VY][i] = vy; A.b[PB_VZ][i] = vz;
    A.b[PB_AX][i] = ax; A.b[PB_AY][i] = ay; A.b[PB_AZ][i] = az;
    A.b[PB_E][i] = e; A.b[PB_DEDT][i] = dedt;
    A.x[i] = x + dt * vx;
    A.vx[i] = vx + dt * (ax + A.gx[i]);
    A.y[i] = y + dt * vy;
    A.vy[i] = vy + dt * (ay + A.gy[i]);
    A.z[i] = z + dt * vz;
    A.vz[i] = vz + dt * (az + A.gz[i]);
    double ep = e + dt * dedt;
    if (ep < 1e-6) ep = 1e-6;
    A.e[i] = ep;
}

// corrector (device_predictor_corrector_euler.cu:130-236): the live fields hold the
// predicted state and its derivatives, the backup the state and derivatives at t.
__global__ __launch_bounds__(256) void k_pc_corrector(PCArgs A) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const double dt = A.dt_dev ? *A.dt_dev : A.dt;
    const double vxp = A.vx[i], vyp = A.vy[i], vzp = A.vz[i];
    const double axp = A.ax[i], ayp = A.ay[i], azp = A.az[i];
    const double vx0 = A.b[PB_VX][i], vy0 = A.b[PB_VY][i], vz0 = A.b[PB_VZ][i];
    const double ax0 = A.b[PB_AX][i], ay0 = A.b[PB_AY][i], az0 = A.b[PB_AZ][i];
    const double gx = A.gx[i], gy = A.gy[i], gz = A.gz[i];
    A.x[i] = A.b[PB_X][i] + dt / 2 * (vxp + vx0);
    A.vx[i] = vx0 + dt / 2 * (axp + ax0 + 2 * gx);
    A.ax[i] = 0.5 * (axp + ax0) + gx;
    A.y[i] = A.b[PB_Y][i] + dt / 2 * (vyp + vy0);
    A.vy[i] = vy0 + dt / 2 * (ayp + ay0 + 2 * gy);
    A.ay[i] = 0.5 * (ayp + ay0) + gy;
    A.z[i] = A.b[PB_Z][i] + dt / 2 * (vzp + vz0);
    A.vz[i] = vz0 + dt / 2 * (azp + az0 + 2 * gz);
    A.az[i] = 0.5 * (azp + az0) + gz;
    const double dedtp = A.dedt[i], dedt0 = A.b[PB_DEDT][i];
    double e = A.b[PB_E][i] + dt / 2 * (dedtp + dedt0);
    if (e < 1e-6) e = 1e-6;
    A.e[i] = e;
    A.dedt[i] = 0.5 * (dedtp + dedt0);
    // INTEGRATE_SML 0: h is the predicted state's (the h-iteration's) value, already live
}

// CFL time step (setTimeStep, device_predictor_corrector_euler.cu:315-540): per-particle
// minima of sqrt(h/|a+g|), h/cs and h/(cs + 1.2 (alpha cs + beta mu_max)); min/max are
// exact, so partials per block and one folding block give a deterministic result.
#define TS_BLOCKS 512
// also max |v| and max h for the search-radius limit dt <= 2 max(h) / (2 max|v|)
// (miluphpc.cpp:2428-2433 and device_predictor_corrector_euler.cu:515-518; the reference
// folds the per-block maxima of |v| with min, this one with max)
__global__ __launch_bounds__(256) void k_time_step_partial(int64_t n, const double *__restrict__ ax,
                                                           const double *__restrict__ ay,
                                                           const double *__restrict__ az,
                                                           const double *__restrict__ gx,
                                                           const double *__restrict__ gy,
                                                           const double *__restrict__ gz,
                                                           const double *__restrict__ sml,
                                                           const double *__restrict__ cs,
                                                           const double *__restrict__ mu,
                                                           const double *__restrict__ vx,
                                                           const double *__restrict__ vy,
                                                           const double *__restrict__ vz,
                                                           double alpha, double beta,
                                                           double *__restrict__ partial) {
    double forces = 1.7976931348623157e308, courant = forces, artvisc = forces;
    double vmax = 0.0, hmax = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double a0 = ax[i] + gx[i], a1 = ay[i] + gy[i], a2 = az[i] + gz[i];
        double t = a0 * a0;
        t += a1 * a1;
        t += a2 * a2;
        const double h = sml[i];
        forces = fmin(forces, sqrt(h / sqrt(t)));
        courant = fmin(courant, h / cs[i]);
        artvisc = fmin(artvisc, h / (cs[i] + 1.2 * (alpha * cs[i] + beta * mu[i])));
        vmax = fmax(vmax, sqrt(vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]));
        hmax = fmax(hmax, h);
    }
    forces = wave_min(forces);
    courant = wave_min(courant);
    artvisc = wave_min(artvisc);
    vmax = wave_max(vmax);
    hmax = wave_max(hmax);
    __shared__ double s[5][4];
    const int w = threadIdx.x / 64;
    if (wave_lane() == 0) {
        s[0][w] = forces;
        s[1][w] = courant;
        s[2][w] = artvisc;
        s[3][w] = vmax;
        s[4][w] = hmax;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double v = s[threadIdx.x][0];
        for (int k = 1; k < 4; k++) v = threadIdx.x < 3 ? fmin(v, s[threadIdx.x][k]) : fmax(v, s[threadIdx.x][k]);
        partial[5 * blockIdx.x + threadIdx.x] = v;
    }
}

// one wavefront folds the block partials: each lane a strided share, then the wave (minima and
// maxima are exact, so the result does not depend on the order).  With R.dt_out set, lane 0
// also applies the time-step rule of mhpc_time_step (the same operations in the same order)
// and writes dt there, for a predictor and corrector that read it on the device.
__global__ __launch_bounds__(64) void k_time_step_final(const double *__restrict__ partial, int nb,
                                                        double *__restrict__ out, DtRule R) {
    const int l = threadIdx.x;
    double v0 = 1.7976931348623157e308, v1 = v0, v2 = v0, v3 = 0.0, v4 = 0.0;
    for (int b = l; b < nb; b += 64) {
        v0 = fmin(v0, partial[5 * b]);
        v1 = fmin(v1, partial[5 * b + 1]);
        v2 = fmin(v2, partial[5 * b + 2]);
        v3 = fmax(v3, partial[5 * b + 3]);
        v4 = fmax(v4, partial[5 * b + 4]);
    }
    v0 = wave_min(v0);
    v1 = wave_min(v1);
    v2 = wave_min(v2);
    v3 = wave_max(v3);
    v4 = wave_max(v4);
    if (l == 0) {
        out[0] = v0; out[1] = v1; out[2] = v2; out[3] = v3; out[4] = v4;
        if (R.dt_out) *R.dt_out = dt_rule(R, v0, v1, v2, v3, v4);
    }
}

static PCArgs pc_args(mhpc_ctx *c, double dt, const double *dt_dev) {
    PCArgs a;
    double **f = c->f;
    a.n = c->n;
    a.dt = dt;
    a.dt_dev = dt_dev;
    a.x = f[MHPC_F_X]; a.y = f[MHPC_F_Y]; a.z = f[MHPC_F_Z];
    a.vx = f[MHPC_F_VX]; a.vy = f[MHPC_F_VY]; a.vz = f[MHPC_F_VZ];
    a.ax = f[MHPC_F_AX]; a.ay = f[MHPC_F_AY]; a.az = f[MHPC_F_AZ];
    a.gx = f[MHPC_F_G_AX]; a.gy = f[MHPC_F_G_AY]; a.gz = f[MHPC_F_G_AZ];
    a.e = f[MHPC_F_E]; a.dedt = f[MHPC_F_DEDT];
    for (int k = 0; k < PB_COUNT_; k++) a.b[k] = c->pc_backup + (int64_t)k * c->cap;
    return a;
}

int pc_backup_fields() { return PB_COUNT_; }

void launch_pc_predictor(mhpc_ctx *c, double dt, const double *dt_dev) {
    if (c->n <= 0) return;
    hipLaunchKernelGGL(k_pc_predictor, dim3((unsigned)ceil_div(c->n, 256)), dim3(256), 0, c->stream,
                       pc_args(c, dt, dt_dev));
}

void launch_pc_corrector(mhpc_ctx *c, double dt, const double *dt_dev) {
    if (c->n <= 0) return;
    hipLaunchKernelGGL(k_pc_corrector, dim3((unsigned)ceil_div(c->n, 256)), dim3(256), 0, c->stream,
                       pc_args(c, dt, dt_dev));
}

void launch_time_step(mhpc_ctx *c, double *partial, double *out, const DtRule &rule) {
    int nb = (int)ceil_div(c->n, 256 * 4);
    if (nb < 1) nb = 1;
    if (nb > TS_BLOCKS) nb = TS_BLOCKS;
    double **f = c->f;
    hipLaunchKernelGGL(k_time_step_partial, dim3(nb), dim3(256), 0, c->stream, c->n, f[MHPC_F_AX],
                       f[MHPC_F_AY], f[MHPC_F_AZ], f[MHPC_F_G_AX], f[MHPC_F_G_AY], f[MHPC_F_G_AZ],
                       f[MHPC_F_SML], f[MHPC_F_CS], f[MHPC_F_MUIJMAX], f[MHPC_F_VX], f[MHPC_F_VY],
                       f[MHPC_F_VZ], c->p.alpha, c->p.beta, partial);
    hipLaunchKernelGGL(k_time_step_final, dim3(1), dim3(64), 0, c->stream, partial, nb, out, rule);
}
