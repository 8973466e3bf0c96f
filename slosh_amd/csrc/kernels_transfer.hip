// Particle <-> grid transfers, grid update, particle update and rigid-body
// integration of one substep on gfx950.
//
//   k_p2g            one workgroup per active block: lane = associated cell,
//                    each wave accumulates its particles' contributions to one
//                    z slab of their 27-node stencils in registers and sums them
//                    into its own LDS slab tile (no atomics, LDS or global); the
//                    6x6x6 tile, the slabs summed per node, is written once to HBM. Same per-(node, particle) arithmetic as
//                    p2g_scatter_style.slang:84-148; only the summation order
//                    differs (and is fixed for a given particle order).
//   k_grid_update    one wave per block: node momentum = own tile + the halo
//                    parts of the 7 "-1" neighbours' tiles, then
//                    grid_update.slang:38-163 (gravity, clamp, collision
//                    cache / collide(), boundary projection).
//   k_g2p_update     one workgroup per block: 6x6x6 node halo in LDS, then per
//                    primary particle g2p.slang:150-265 fused with
//                    particle_update.slang:32-160. Outputs go to the other
//                    particle buffer at the particle's sorted slot, so particle
//                    storage follows the block order of the last substep.
//   k_g2p / k_particle_update  the same two stages as separate in-place
//                    kernels, used when a hook replaces or observes G2P.
//   k_rigid_update   rigid_impulses.slang:70-121.
#include <type_traits>

#include "launch.h"
#include "../../include/slosh_hip.h"

namespace {

constexpr int ASSOC[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1},
                             {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};

// ---------------------------------------------------------------------------
// Quadratic B-spline (kernel.slang:160-174).

__device__ inline void k_eval_all(float x, float *w) {
    w[0] = 0.5f * (1.5f - x) * (1.5f - x);
    w[1] = 0.75f - (x - 1.0f) * (x - 1.0f);
    w[2] = 0.5f * (x - 0.5f) * (x - 0.5f);
}

// ---------------------------------------------------------------------------
// 3x3 helpers on shader matrices stored row-major in float[9].

__device__ inline float det3(const float *a) {
#pragma clang fp contract(off)
    return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}

__device__ inline void mul3(const float *a, const float *b, float *r) {
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r[i * 3 + j] = a[i * 3 + 0] * b[0 * 3 + j] + a[i * 3 + 1] * b[1 * 3 + j] + a[i * 3 + 2] * b[2 * 3 + j];
}

// mul(transpose(a), b)
__device__ inline void mul3_tn(const float *a, const float *b, float *r) {
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r[i * 3 + j] = a[0 * 3 + i] * b[0 * 3 + j] + a[1 * 3 + i] * b[1 * 3 + j] + a[2 * 3 + i] * b[2 * 3 + j];
}

// 3x3 SVD: cyclic Jacobi eigen-decomposition of A^T A (Golub & Van Loan
// symmetric Schur rotations), eigenvalues sorted descending, then a Givens QR
// of A V (the layout of McAdams et al. 2011). U and V are rotations,
// sig0 >= sig1 >= |sig2| and sign(sig2) = sign(det A). T = float or double.
// Evaluated without multiply-add contraction, operation by operation as written,
// so that the FP32 instance rounds like a plain IEEE evaluation of the same steps.
template <typename T>
__device__ void svd3(const T *a, T *u, T *sig, T *v) {
#pragma clang fp contract(off)
    const T tol = sizeof(T) == 4 ? (T)1.0e-14 : (T)1.0e-30;
    T s[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) s[i][j] = a[0 * 3 + i] * a[0 * 3 + j] + a[1 * 3 + i] * a[1 * 3 + j] + a[2 * 3 + i] * a[2 * 3 + j];
    T vv[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma unroll 1
    for (int sweep = 0; sweep < 12; sweep++) {
        T off = s[0][1] * s[0][1] + s[0][2] * s[0][2] + s[1][2] * s[1][2];
        T dg = s[0][0] * s[0][0] + s[1][1] * s[1][1] + s[2][2] * s[2][2];
        if (!(off > tol * dg)) break;
#pragma unroll
        for (int e = 0; e < 3; e++) {
            const int p = e == 2 ? 1 : 0;
            const int q = e == 0 ? 1 : 2;
            T apq = s[p][q];
            if (apq == (T)0) continue;
            T theta = (s[q][q] - s[p][p]) / ((T)2 * apq);
            T t = (theta >= (T)0 ? (T)1 : (T)-1) / (fabs(theta) + sqrt(theta * theta + (T)1));
            T c = (T)1 / sqrt(t * t + (T)1);
            T sn = t * c;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                T skp = s[k][p], skq = s[k][q];
                s[k][p] = c * skp - sn * skq;
                s[k][q] = sn * skp + c * skq;
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                T spk = s[p][k], sqk = s[q][k];
                s[p][k] = c * spk - sn * sqk;
                s[q][k] = sn * spk + c * sqk;
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                T vkp = vv[k][p], vkq = vv[k][q];
                vv[k][p] = c * vkp - sn * vkq;
                vv[k][q] = sn * vkp + c * vkq;
            }
        }
    }
    T ev[3] = {s[0][0], s[1][1], s[2][2]};
#pragma unroll
    for (int pass = 0; pass < 2; pass++)
#pragma unroll
        for (int i = 0; i < 2; i++)
            if (ev[i] < ev[i + 1]) {
                T te = ev[i];
                ev[i] = ev[i + 1];
                ev[i + 1] = te;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    T tv = vv[k][i];
                    vv[k][i] = -vv[k][i + 1];
                    vv[k][i + 1] = tv;
                }
            }
    T b[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) b[i][j] = a[i * 3 + 0] * vv[0][j] + a[i * 3 + 1] * vv[1][j] + a[i * 3 + 2] * vv[2][j];
    T qm[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma unroll
    for (int e = 0; e < 3; e++) {
        const int r0 = e == 2 ? 1 : 0;
        const int r1 = e == 0 ? 1 : 2;
        const int col = r0;
        T x = b[r0][col], y = b[r1][col];
        T r = sqrt(x * x + y * y);
        T c = (T)1, sn = (T)0;
        if (r > (T)0) {
            c = x / r;
            sn = y / r;
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            T b0 = b[r0][k], b1 = b[r1][k];
            b[r0][k] = c * b0 + sn * b1;
            b[r1][k] = -sn * b0 + c * b1;
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            T q0 = qm[k][r0], q1 = qm[k][r1];
            qm[k][r0] = c * q0 + sn * q1;
            qm[k][r1] = -sn * q0 + c * q1;
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        sig[i] = b[i][i];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            u[i * 3 + j] = qm[i][j];
            v[i * 3 + j] = vv[i][j];
        }
    }
}

// U diag(s) V^T
template <typename T>
__device__ inline void svd_recompose(const T *u, const T *s, const T *v, T *r) {
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r[i * 3 + j] = (u[i * 3 + 0] * s[0]) * v[j * 3 + 0] + (u[i * 3 + 1] * s[1]) * v[j * 3 + 1] + (u[i * 3 + 2] * s[2]) * v[j * 3 + 2];
}

// ---------------------------------------------------------------------------
// Material models (shaders/slosh/models).

// neo_hookean_elasticity.slang:16-27
__device__ inline void nh_stress(float lambda, float mu, const float *f, float *st) {
#pragma clang fp contract(off)
    float j = fmaxf(det3(f), 1.0e-10f);
    float diag = lambda * logf(j) - mu;
    mul3_tn(f, f, st);
#pragma unroll
    for (int k = 0; k < 9; k++) st[k] *= mu;
    st[0] += diag;
    st[4] += diag;
    st[8] += diag;
}

// linear_elasticity.slang:135-148 (fixed-corotated)
__device__ inline void fc_stress(float lambda, float mu, const float *f, float *st) {
#pragma clang fp contract(off)
    float u[9], s[3], v[9], rc[9];
    svd3<float>(f, u, s, v);
    float j = s[0] * s[1] * s[2];
    float sm1[3] = {s[0] - 1.0f, s[1] - 1.0f, s[2] - 1.0f};
    float diag = lambda * (j - 1.0f) * j;
    svd_recompose(u, sm1, v, rc);
    mul3_tn(f, rc, st);
    float k2 = 2.0f * mu;
#pragma unroll
    for (int k = 0; k < 9; k++) st[k] *= k2;
    st[0] += diag;
    st[4] += diag;
    st[8] += diag;
}

// drucker_prager.slang:144-200: returns the projected deformation gradient in f,
// updates state = {plastic det, hardening, log volume gain}.
__device__ inline void dp_project(float *state, const float4 hp, float lambda, float mu, float *f) {
#pragma clang fp contract(off)
    if (lambda == 0.0f) return;
    float u[9], s[3], v[9];
    svd3<float>(f, u, s, v);
    float angle = hp.x + (hp.y * state[1] - hp.w) * expf(-hp.z * state[1]);
    float s_angle = sinf(angle);
    float alpha = sqrtf(2.0f / 3.0f) * (2.0f * s_angle) / (3.0f - s_angle);
    const float d = 3.0f;
    float strain[3], dev[3];
    float lvg = state[2];
#pragma unroll
    for (int i = 0; i < 3; i++) strain[i] = logf(s[i]) + lvg / d;
    float tr = strain[0] + strain[1] + strain[2];
#pragma unroll
    for (int i = 0; i < 3; i++) dev[i] = strain[i] - tr / d;
    float ns[3];
    float hard;
    if (tr > 0.0f || (dev[0] == 0.0f && dev[1] == 0.0f && dev[2] == 0.0f)) {
        ns[0] = ns[1] = ns[2] = 1.0f;
        hard = sqrtf(strain[0] * strain[0] + strain[1] * strain[1] + strain[2] * strain[2]);
    } else {
        float dn = sqrtf(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
        float gamma = dn + (d * lambda + 2.0f * mu) / (2.0f * mu) * tr * alpha;
        if (gamma <= 0.0f) return;
#pragma unroll
        for (int i = 0; i < 3; i++) ns[i] = expf(strain[i] - dev[i] * (gamma / dn));
        hard = gamma;
    }
    float prev_det = s[0] * s[1] * s[2];
    float new_det = ns[0] * ns[1] * ns[2];
    state[0] = state[0] * prev_det / new_det;
    state[2] = state[2] + logf(prev_det) - logf(new_det);
    state[1] = state[1] + hard;
    svd_recompose(u, ns, v, f);
}

// Refined variants of the SVD-dependent steps. In FP32, F - R and log(sigma)
// near sigma = 1 lose ~5 digits to cancellation, which is the size of the parity
// tolerance. Here the quantities that cancel are computed to FP64 accuracy: the
// polar rotation R by FP64 Newton iterations from F; the singular values as
// |F v_i| / |v_i| in FP64 from the FP32 SVD's right singular vectors (the error
// is second order in theirs); log(sigma) and exp of the small log-strain
// corrections through log1p / expm1 of FP64 differences from 1.
// Orthogonal polar factor r of a (a = r s, s symmetric positive definite) by
// Newton's iteration x <- (g x + x^-T / g) / 2 with determinant scaling
// g = det(x)^(-1/3) while det(x) is far from 1 (N. J. Higham, "Computing the
// polar decomposition -- with applications", SIAM J. Sci. Stat. Comput. 7(4),
// 1986). For det(a) > 0 this is U V^T of the SVD a = U diag(sig) V^T. Returns
// false (r undefined) if det(a) <= 0 or it does not converge.
__device__ inline bool polar_f64(const double *a, double *r) {
    double x[9];
#pragma unroll
    for (int k = 0; k < 9; k++) x[k] = a[k];
    bool close = false;
#pragma unroll 1
    for (int it = 0; it < 24; it++) {
        double c[9];
        c[0] = x[4] * x[8] - x[5] * x[7];
        c[1] = x[5] * x[6] - x[3] * x[8];
        c[2] = x[3] * x[7] - x[4] * x[6];
        c[3] = x[2] * x[7] - x[1] * x[8];
        c[4] = x[0] * x[8] - x[2] * x[6];
        c[5] = x[1] * x[6] - x[0] * x[7];
        c[6] = x[1] * x[5] - x[2] * x[4];
        c[7] = x[2] * x[3] - x[0] * x[5];
        c[8] = x[0] * x[4] - x[1] * x[3];
        const double det = x[0] * c[0] + x[1] * c[1] + x[2] * c[2];
        if (!(det > 0.0)) return false;
        // Any positive scale keeps the iteration convergent; FP32 is enough for it.
        const double g = fabs(det - 1.0) > 1.0e-3 ? (double)(1.0f / cbrtf((float)det)) : 1.0;
        const double h = 0.5 / (g * det);
        double diff = 0.0, norm = 0.0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const double nx = 0.5 * g * x[k] + h * c[k];
            diff = fmax(diff, fabs(nx - x[k]));
            norm = fmax(norm, fabs(nx));
            x[k] = nx;
        }
        // Quadratic convergence: once a step is below 1e-9 the next one is at round-off.
        if (close) {
#pragma unroll
            for (int k = 0; k < 9; k++) r[k] = x[k];
            return true;
        }
        close = diff <= 1.0e-9 * norm;
    }
    return false;
}

__device__ inline double det3d(const double *a) {
    return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// linear_elasticity.slang:135-148 in FP64: tau = 2 mu (F - R) F^T + lambda (J - 1) J I
// with R the polar rotation (U V^T of the SVD), J = det F.
__device__ inline void fc_stress_f64(float lambda, float mu, const float *f32, float *st) {
    double f[9], rc[9], j;
#pragma unroll
    for (int k = 0; k < 9; k++) f[k] = f32[k];
    double r[9];
    if (!polar_f64(f, r)) {
        // det F <= 0 (the particle is disabled by the update) or no convergence.
        fc_stress(lambda, mu, f32, st);
        return;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) rc[k] = f[k] - r[k];
    j = det3d(f);
    const double mu2 = 2.0 * (double)mu;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double r2 = mu2 * (f[0 * 3 + i] * rc[0 * 3 + k] + f[1 * 3 + i] * rc[1 * 3 + k] + f[2 * 3 + i] * rc[2 * 3 + k]);
            if (i == k) r2 += (double)lambda * (j - 1.0) * j;
            st[i * 3 + k] = (float)r2;
        }
}

// log(x) for x near 1 from the FP64 difference x - 1.
__device__ inline double log_near1(double x) { return (double)log1pf((float)(x - 1.0)); }

// Right singular vectors of f (the columns of v) to FP64 accuracy from FP32 ones
// (the columns of v32): one Newton-Schulz step v = w (3 I - w^T w) / 2 makes them
// orthonormal to FP64 round-off, then cyclic Jacobi sweeps on v^T f^T f v, whose
// off-diagonal part starts at FP32 round-off, converge quadratically (at most 4
// sweeps, also for repeated singular values). With FP32 vectors U diag(s') V^T
// carries ~1e-7 absolute error, which is the size of a plastic F - I step.
__device__ inline void refine_right_vectors(const double *f, const float *v32, double *v) {
    double w[9], g[9];
#pragma unroll
    for (int k = 0; k < 9; k++) w[k] = v32[k];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) g[i * 3 + j] = w[0 * 3 + i] * w[0 * 3 + j] + w[1 * 3 + i] * w[1 * 3 + j] + w[2 * 3 + i] * w[2 * 3 + j];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            v[i * 3 + j] = 0.5 * (w[i * 3 + 0] * ((j == 0 ? 3.0 : 0.0) - g[0 * 3 + j]) +
                                  w[i * 3 + 1] * ((j == 1 ? 3.0 : 0.0) - g[1 * 3 + j]) +
                                  w[i * 3 + 2] * ((j == 2 ? 3.0 : 0.0) - g[2 * 3 + j]));
    double b[9], a[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) b[i * 3 + j] = f[i * 3 + 0] * v[0 * 3 + j] + f[i * 3 + 1] * v[1 * 3 + j] + f[i * 3 + 2] * v[2 * 3 + j];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) a[i][j] = b[0 * 3 + i] * b[0 * 3 + j] + b[1 * 3 + i] * b[1 * 3 + j] + b[2 * 3 + i] * b[2 * 3 + j];
#pragma unroll 1
    for (int sweep = 0; sweep < 4; sweep++) {
        const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        const double dg = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
        if (!(off > 1.0e-30 * dg)) break;
#pragma unroll
        for (int e = 0; e < 3; e++) {
            const int p = e == 2 ? 1 : 0;
            const int q = e == 0 ? 1 : 2;
            const double apq = a[p][q];
            if (apq == 0.0) continue;
            const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
            const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double x = a[k][p], y = a[k][q];
                a[k][p] = c * x - sn * y;
                a[k][q] = sn * x + c * y;
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double x = a[p][k], y = a[q][k];
                a[p][k] = c * x - sn * y;
                a[q][k] = sn * x + c * y;
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double x = v[k * 3 + p], y = v[k * 3 + q];
                v[k * 3 + p] = c * x - sn * y;
                v[k * 3 + q] = sn * x + c * y;
            }
        }
    }
}

// drucker_prager.slang:144-200 (Klar et al. 2016 return mapping in log-strain
// space on the singular values), refined as described above.
__device__ inline void dp_project_f64(float *state, const float4 hp, float lambda_f, float mu_f, float *f32) {
    if (lambda_f == 0.0f) return;
    const double lambda = lambda_f, mu = mu_f;
    float u[9], s32[3], v[9];
    svd3<float>(f32, u, s32, v);
    double s[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double v0 = v[0 * 3 + i], v1 = v[1 * 3 + i], v2 = v[2 * 3 + i];
        const double w0 = (double)f32[0] * v0 + (double)f32[1] * v1 + (double)f32[2] * v2;
        const double w1 = (double)f32[3] * v0 + (double)f32[4] * v1 + (double)f32[5] * v2;
        const double w2 = (double)f32[6] * v0 + (double)f32[7] * v1 + (double)f32[8] * v2;
        s[i] = sqrt((w0 * w0 + w1 * w1 + w2 * w2) / (v0 * v0 + v1 * v1 + v2 * v2));
    }
    if (s32[2] < 0.0f) s[2] = -s[2];  // svd3 gives sig[2] the sign of det F
    const float q = state[1];
    const float angle = hp.x + (hp.y * q - hp.w) * expf(-hp.z * q);
    const float sa = sinf(angle);
    const double alpha = (double)(sqrtf(2.0f / 3.0f) * (2.0f * sa) / (3.0f - sa));
    const double d = 3.0;
    double strain[3], dev[3], ns[3], hard;
#pragma unroll
    for (int i = 0; i < 3; i++) strain[i] = log_near1(s[i]) + (double)state[2] / d;
    double tr = strain[0] + strain[1] + strain[2];
#pragma unroll
    for (int i = 0; i < 3; i++) dev[i] = strain[i] - tr / d;
    const double prev_det = s[0] * s[1] * s[2];
    if (tr > 0.0 || (dev[0] == 0.0 && dev[1] == 0.0 && dev[2] == 0.0)) {
        // Tension: sigma -> 1, so F becomes its polar rotation U V^T.
        hard = sqrt(strain[0] * strain[0] + strain[1] * strain[1] + strain[2] * strain[2]);
        double fd[9], r[9];
#pragma unroll
        for (int k = 0; k < 9; k++) fd[k] = f32[k];
        if (polar_f64(fd, r)) {
#pragma unroll
            for (int k = 0; k < 9; k++) f32[k] = (float)r[k];
        } else {
            const float one[3] = {1.0f, 1.0f, 1.0f};
            svd_recompose(u, one, v, f32);
        }
        state[0] = (float)((double)state[0] * prev_det);
        state[2] = (float)((double)state[2] + log_near1(prev_det));
        state[1] = (float)((double)state[1] + hard);
        return;
    }
    double dn = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
    double gamma = dn + (d * lambda + 2.0 * mu) / (2.0 * mu) * tr * alpha;
    if (gamma <= 0.0) return;
    if (!(s32[2] > 0.0f)) {
        // det F <= 0 (or NaN): log of the signed singular value is NaN, as in the
        // shader; the particle is disabled by the update.
#pragma unroll
        for (int i = 0; i < 3; i++) ns[i] = 1.0 + (double)expm1f((float)(strain[i] - dev[i] * (gamma / dn)));
        const double new_det = ns[0] * ns[1] * ns[2];
        state[0] = (float)((double)state[0] * prev_det / new_det);
        state[2] = (float)((double)state[2] + log_near1(prev_det) - log_near1(new_det));
        state[1] = (float)((double)state[1] + gamma);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                f32[i * 3 + j] = (float)((double)u[i * 3 + 0] * ns[0] * (double)v[j * 3 + 0] +
                                         (double)u[i * 3 + 1] * ns[1] * (double)v[j * 3 + 1] +
                                         (double)u[i * 3 + 2] * ns[2] * (double)v[j * 3 + 2]);
        return;
    }
    // Projection onto the cone: F' = U diag(s') V^T needs FP64-accurate singular
    // vectors. Refine V, take the singular values |F v_i| from it and redo the
    // return mapping with them (the decision above used the same values to second
    // order), then F' = sum_i s'_i / |F v_i| (F v_i) v_i^T.
    double fd[9], vd[9], bd[9];
#pragma unroll
    for (int k = 0; k < 9; k++) fd[k] = f32[k];
    refine_right_vectors(fd, v, vd);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) bd[i * 3 + j] = fd[i * 3 + 0] * vd[0 * 3 + j] + fd[i * 3 + 1] * vd[1 * 3 + j] + fd[i * 3 + 2] * vd[2 * 3 + j];
    double sr[3];
#pragma unroll
    for (int j = 0; j < 3; j++) sr[j] = sqrt(bd[0 * 3 + j] * bd[0 * 3 + j] + bd[1 * 3 + j] * bd[1 * 3 + j] + bd[2 * 3 + j] * bd[2 * 3 + j]);
#pragma unroll
    for (int i = 0; i < 3; i++) strain[i] = log_near1(sr[i]) + (double)state[2] / d;
    tr = strain[0] + strain[1] + strain[2];
#pragma unroll
    for (int i = 0; i < 3; i++) dev[i] = strain[i] - tr / d;
    dn = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
    gamma = dn + (d * lambda + 2.0 * mu) / (2.0 * mu) * tr * alpha;
    if (gamma <= 0.0) return;
#pragma unroll
    for (int i = 0; i < 3; i++) ns[i] = 1.0 + (double)expm1f((float)(strain[i] - dev[i] * (gamma / dn)));
    hard = gamma;
    const double det_r = sr[0] * sr[1] * sr[2];
    const double new_det = ns[0] * ns[1] * ns[2];
    state[0] = (float)((double)state[0] * det_r / new_det);
    state[2] = (float)((double)state[2] + log_near1(det_r) - log_near1(new_det));
    state[1] = (float)((double)state[1] + hard);
    double rs[3];
#pragma unroll
    for (int j = 0; j < 3; j++) rs[j] = ns[j] / sr[j];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            f32[i * 3 + j] = (float)(bd[i * 3 + 0] * rs[0] * vd[j * 3 + 0] + bd[i * 3 + 1] * rs[1] * vd[j * 3 + 1] +
                                     bd[i * 3 + 2] * rs[2] * vd[j * 3 + 2]);
}

// ---------------------------------------------------------------------------
// Shapes, poses and collisions (shaders/slosh/rbd, shaders/slosh/collision).
//
// The contact geometry is evaluated without FMA contraction, with the operation
// order of plain IEEE single precision. A contact normal is (projection - node) /
// distance, so for a node near a surface its rounding error grows as 1 / distance:
// contracting some products (the compiler's default) moves such normals by up to
// 1e-4, where the uncontracted evaluation reproduces the CPU's bits.

__device__ inline float3 quat_rotate(float4 q, float3 v) {
#pragma clang fp contract(off)
    float3 t = make_float3(2.0f * (q.y * v.z - q.z * v.y), 2.0f * (q.z * v.x - q.x * v.z), 2.0f * (q.x * v.y - q.y * v.x));
    return make_float3(v.x + q.w * t.x + (q.y * t.z - q.z * t.y), v.y + q.w * t.y + (q.z * t.x - q.x * t.z),
                       v.z + q.w * t.z + (q.x * t.y - q.y * t.x));
}

// Sim3 (quaternion, translation, uniform scale) point transforms.
__device__ inline float3 sim_mul_pt(float4 rot, float4 ts, float3 p) {
#pragma clang fp contract(off)
    float3 r = quat_rotate(rot, make_float3(p.x * ts.w, p.y * ts.w, p.z * ts.w));
    return make_float3(r.x + ts.x, r.y + ts.y, r.z + ts.z);
}
__device__ inline float3 sim_inv_mul_pt(float4 rot, float4 ts, float3 p) {
#pragma clang fp contract(off)
    float3 r = quat_rotate(make_float4(-rot.x, -rot.y, -rot.z, rot.w), make_float3(p.x - ts.x, p.y - ts.y, p.z - ts.z));
    return make_float3(r.x / ts.w, r.y / ts.w, r.z / ts.w);
}

__device__ inline float sgn(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// cuboid.slang:34-63
__device__ inline bool cuboid_project_boundary(float3 he, float3 pt, float3 *out) {
#pragma clang fp contract(off)
    float3 op;
    op.x = pt.x + (fmaxf(-he.x - pt.x, 0.0f) - fmaxf(pt.x - he.x, 0.0f));
    op.y = pt.y + (fmaxf(-he.y - pt.y, 0.0f) - fmaxf(pt.y - he.y, 0.0f));
    op.z = pt.z + (fmaxf(-he.z - pt.z, 0.0f) - fmaxf(pt.z - he.z, 0.0f));
    float sx = sgn(pt.x), sy = sgn(pt.y), sz = sgn(pt.z);
    sx = sx + (fabsf(sx) - 1.0f);
    sy = sy + (fabsf(sy) - 1.0f);
    sz = sz + (fabsf(sz) - 1.0f);
    float dx = he.x - sx * pt.x, dy = he.y - sy * pt.y, dz = he.z - sz * pt.z;
    bool pick_x = dx <= dy && dx <= dz;
    bool pick_y = dy <= dx && dy <= dz;
    float3 ip = pt;
    if (pick_x) ip.x = pt.x + dx * sx;
    else if (pick_y) ip.y = pt.y + dy * sy;
    else ip.z = pt.z + dz * sz;
    bool inside = pt.x == op.x && pt.y == op.y && pt.z == op.z;
    *out = inside ? ip : op;
    return inside;
}

// ball.slang:39-50
__device__ inline bool ball_project_boundary(float radius, float3 pt, float3 *out) {
#pragma clang fp contract(off)
    float dist = sqrtf(pt.x * pt.x + pt.y * pt.y + pt.z * pt.z);
    if (dist != 0.0f) {
        float k = radius / dist;
        *out = make_float3(pt.x * k, pt.y * k, pt.z * k);
    } else {
        *out = make_float3(0.0f, radius, 0.0f);
    }
    return dist <= radius;
}

// segment.slang:10-28: closest point of segment [a, b] to p.
__device__ inline float3 segment_project(float3 a, float3 b, float3 p) {
#pragma clang fp contract(off)
    const float3 ab = make_float3(b.x - a.x, b.y - a.y, b.z - a.z);
    const float3 ap = make_float3(p.x - a.x, p.y - a.y, p.z - a.z);
    const float ab_ap = ab.x * ap.x + ab.y * ap.y + ab.z * ap.z;
    const float sqnab = ab.x * ab.x + ab.y * ab.y + ab.z * ab.z;
    if (ab_ap <= 0.0f) return a;
    if (ab_ap >= sqnab) return b;
    const float u = ab_ap / sqnab;
    return make_float3(a.x + ab.x * u, a.y + ab.y * u, a.z + ab.z * u);
}

// capsule.slang:41-58 and 77-100 (segment a..b, radius r).
__device__ inline bool capsule_project_boundary(float3 a, float3 b, float r, float3 pt, float3 *out) {
#pragma clang fp contract(off)
    const float3 on_axis = segment_project(a, b, pt);
    const float3 d = make_float3(pt.x - on_axis.x, pt.y - on_axis.y, pt.z - on_axis.z);
    const float dist = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    if (dist > 0.0f) {
        const float k = r / dist;
        *out = make_float3(on_axis.x + d.x * k, on_axis.y + d.y * k, on_axis.z + d.z * k);
        return dist <= r;
    }
    // On the axis: the first vector of the orthonormal basis of the axis direction.
    float3 ax = make_float3(b.x - a.x, b.y - a.y, b.z - a.z);
    const float len = sqrtf(ax.x * ax.x + ax.y * ax.y + ax.z * ax.z);
    ax = make_float3(ax.x / len, ax.y / len, ax.z / len);
    const float sign = ax.z >= 0.0f ? 1.0f : -1.0f;
    const float ka = -1.0f / (sign + ax.z);
    const float kb = ax.x * ax.y * ka;
    const float3 dir = make_float3(1.0f + sign * ax.x * ax.x * ka, sign * kb, -sign * ax.x);
    *out = make_float3(on_axis.x + dir.x * r, on_axis.y + dir.y * r, on_axis.z + dir.z * r);
    return true;
}

// |(x, z)| and the unit direction of (x, z) from the y axis ((1, 0) on the axis).
__device__ inline float planar_dir(float3 pt, float *dx, float *dz) {
#pragma clang fp contract(off)
    const float pd = sqrtf(pt.x * pt.x + pt.z * pt.z);
    *dx = pd > 0.0f ? pt.x / pd : 1.0f;
    *dz = pd > 0.0f ? pt.z / pd : 0.0f;
    return pd;
}

// cylinder.slang:68-128 (axis y, half height hh, radius r).
__device__ inline bool cylinder_project_boundary(float hh, float r, float3 pt, float3 *out) {
#pragma clang fp contract(off)
    float dx, dz;
    const float pd = planar_dir(pt, &dx, &dz);
    const float px = dx * r, pz = dz * r;
    if (pt.y >= -hh && pt.y <= hh && pd <= r) {
        const float to_top = hh - pt.y, to_bottom = pt.y - (-hh), to_side = r - pd;
        if (to_top < to_bottom && to_top < to_side) *out = make_float3(pt.x, hh, pt.z);
        else if (to_bottom < to_top && to_bottom < to_side) *out = make_float3(pt.x, -hh, pt.z);
        else *out = make_float3(px, pt.y, pz);
        return true;
    }
    if (pt.y > hh || pt.y < -hh) {
        const float y = pt.y > hh ? hh : -hh;
        *out = pd <= r ? make_float3(pt.x, y, pt.z) : make_float3(px, y, pz);
    } else {
        *out = make_float3(px, pt.y, pz);
    }
    return false;
}

__device__ inline float3 cross3f(float3 a, float3 b) {
#pragma clang fp contract(off)
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// cone.slang:61-120 (axis y, apex at +hh, base disc of radius r at -hh).
__device__ inline bool cone_project_boundary(float hh, float r, float3 pt, float3 *out) {
#pragma clang fp contract(off)
    float dx, dz;
    const float pd = planar_dir(pt, &dx, &dz);
    const float3 on_base = make_float3(pt.x, -hh, pt.z);
    if (pt.y < -hh && pd <= r) {
        *out = on_base;
        return false;
    }
    const float3 rim = make_float3(dx * r, -hh, dz * r);
    const float3 apex = make_float3(0.0f, hh, 0.0f);
    const float3 side = make_float3(rim.x - apex.x, rim.y - apex.y, rim.z - apex.z);
    const float3 proj = segment_project(apex, rim, pt);
    const float3 c1 = cross3f(side, make_float3(pt.x - apex.x, pt.y - apex.y, pt.z - apex.z));
    const float3 c2 = cross3f(side, make_float3(0.0f, -2.0f * hh, 0.0f));
    if (pt.y >= -hh && pt.y <= hh && c1.x * c2.x + c1.y * c2.y + c1.z * c2.z >= 0.0f) {
        const float3 tp = make_float3(proj.x - pt.x, proj.y - pt.y, proj.z - pt.z);
        const float3 tb = make_float3(on_base.x - pt.x, on_base.y - pt.y, on_base.z - pt.z);
        *out = (tp.x * tp.x + tp.y * tp.y + tp.z * tp.z > tb.x * tb.x + tb.y * tb.y + tb.z * tb.z) ? on_base : proj;
        return true;
    }
    *out = proj;
    return false;
}

struct Colliders {
    const float4 *shapes; // 2 per body: a, b
    const float4 *poses;  // 2 per body: rotation, translation_scale
    const uint2 *mats;    // (type, friction bits)
    uint32_t count;
    const float4 *mesh_vtx;   // trimesh collision vertices: BVH AABBs, vertices, pseudo-normals
    const uint32_t *mesh_idx; // trimesh collision indices: BVH topology, triangles
};

// triangle_project_local_point_and_get_location (trimesh.slang:304-377): closest
// point of triangle (a, b, c) to p; feature 0 = vertex (id 0..2 = a, b, c),
// 1 = edge (id 0 = ab, 1 = bc, 2 = ca), 2 = face, 3 = degenerate (the point itself).
__device__ inline float3 triangle_project(float3 a, float3 b, float3 c, float3 p, uint32_t *feat, uint32_t *id) {
#pragma clang fp contract(off)
    const float3 ab = make_float3(b.x - a.x, b.y - a.y, b.z - a.z);
    const float3 ac = make_float3(c.x - a.x, c.y - a.y, c.z - a.z);
    const float3 ap = make_float3(p.x - a.x, p.y - a.y, p.z - a.z);
    auto dot = [](float3 u, float3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; };
    const float ab_ap = dot(ab, ap), ac_ap = dot(ac, ap);
    if (ab_ap <= 0.0f && ac_ap <= 0.0f) {
        *feat = 0u; *id = 0u;
        return a;
    }
    const float3 bp = make_float3(p.x - b.x, p.y - b.y, p.z - b.z);
    const float ab_bp = dot(ab, bp), ac_bp = dot(ac, bp);
    if (ab_bp >= 0.0f && ac_bp <= ab_bp) {
        *feat = 0u; *id = 1u;
        return b;
    }
    const float3 cp = make_float3(p.x - c.x, p.y - c.y, p.z - c.z);
    const float ab_cp = dot(ab, cp), ac_cp = dot(ac, cp);
    if (ac_cp >= 0.0f && ab_cp <= ac_cp) {
        *feat = 0u; *id = 2u;
        return c;
    }
    const float3 bc = make_float3(c.x - b.x, c.y - b.y, c.z - b.z);
    // stable_check_edges_voronoi (trimesh.slang:257-302)
    const float3 n = cross3f(ab, ac);
    const float vc = dot(n, cross3f(ab, ap));
    if (vc < 0.0f && ab_ap >= 0.0f && ab_bp <= 0.0f) {
        const float v = ab_ap / dot(ab, ab);
        *feat = 1u; *id = 0u;
        return make_float3(a.x + ab.x * v, a.y + ab.y * v, a.z + ab.z * v);
    }
    const float vb = -dot(n, cross3f(ac, cp));
    if (vb < 0.0f && ac_ap >= 0.0f && ac_cp <= 0.0f) {
        const float w = ac_ap / dot(ac, ac);
        *feat = 1u; *id = 2u;
        return make_float3(a.x + ac.x * w, a.y + ac.y * w, a.z + ac.z * w);
    }
    const float va = dot(n, cross3f(bc, bp));
    if (va < 0.0f && ac_bp - ab_bp >= 0.0f && ab_cp - ac_cp >= 0.0f) {
        const float w = dot(bc, bp) / dot(bc, bc);
        *feat = 1u; *id = 1u;
        return make_float3(b.x + bc.x * w, b.y + bc.y * w, b.z + bc.z * w);
    }
    if (va + vb + vc != 0.0f) {
        const float denom = 1.0f / (va + vb + vc);
        const float v = vb * denom, w = vc * denom;
        *feat = 2u; *id = 0u;
        return make_float3(a.x + ab.x * v + ac.x * w, a.y + ab.y * v + ac.y * w, a.z + ab.z * v + ac.z * w);
    }
    *feat = 3u; *id = 0u;
    return p;
}

// TriMesh::project_local_point (trimesh.slang:111-154): the closest triangle within
// max_dist by the stackless BVH walk (links point forward, checked on the host, so
// the walk ends after at most `nodes` steps), inside by the pseudo-normal of the
// feature the point projects on. Returns whether a triangle was found.
__device__ inline bool trimesh_project(const Colliders &cs, float4 a, float4 b, float3 pt, float max_dist,
                                       float3 *out, bool *inside) {
#pragma clang fp contract(off)
    const uint32_t vroot = __float_as_uint(a.x), iroot = __float_as_uint(a.y), nodes = __float_as_uint(a.z);
    const uint32_t nvert = __float_as_uint(b.y);
    const uint32_t *topo = cs.mesh_idx + iroot;
    const uint32_t *tri_idx = topo + 3u * nodes;
    const float4 *verts = cs.mesh_vtx + vroot + 2u * nodes;
    float best = max_dist;
    float3 best_p = pt;
    uint32_t best_tri = 0u, best_feat = 3u, best_id = 0u;
    bool found = false;
    uint32_t cur = 0u;
    for (uint32_t steps = 0; cur < nodes && steps < nodes; steps++) {
        const uint32_t entry = topo[3u * cur], exit_ = topo[3u * cur + 1u];
        if (entry == 0xffffffffu) {
            const uint32_t tri = topo[3u * cur + 2u];
            const float4 va = verts[tri_idx[3u * tri]], vb = verts[tri_idx[3u * tri + 1u]],
                         vc = verts[tri_idx[3u * tri + 2u]];
            uint32_t feat, id;
            const float3 p = triangle_project(make_float3(va.x, va.y, va.z), make_float3(vb.x, vb.y, vb.z),
                                              make_float3(vc.x, vc.y, vc.z), pt, &feat, &id);
            const float dx = p.x - pt.x, dy = p.y - pt.y, dz = p.z - pt.z;
            const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
            if (dist <= best) {
                best = dist;
                best_p = p;
                best_tri = tri; best_feat = feat; best_id = id;
                found = true;
            }
            cur = exit_;
        } else {
            const float4 mn = cs.mesh_vtx[vroot + 2u * cur], mx = cs.mesh_vtx[vroot + 2u * cur + 1u];
            const float qx = (pt.x + (fmaxf(mn.x - pt.x, 0.0f) - fmaxf(pt.x - mx.x, 0.0f))) - pt.x;
            const float qy = (pt.y + (fmaxf(mn.y - pt.y, 0.0f) - fmaxf(pt.y - mx.y, 0.0f))) - pt.y;
            const float qz = (pt.z + (fmaxf(mn.z - pt.z, 0.0f) - fmaxf(pt.z - mx.z, 0.0f))) - pt.z;
            cur = sqrtf(qx * qx + qy * qy + qz * qz) < best ? entry : exit_;
        }
    }
    if (!found) return false;
    // TriMesh::pseudo_normal (trimesh.slang:83-109)
    float3 pn;
    const uint32_t *tv = tri_idx + 3u * best_tri;
    if (best_feat == 0u) {
        const float4 q = verts[tv[best_id] + nvert];
        pn = make_float3(q.x, q.y, q.z);
    } else if (best_feat == 1u) {
        const float4 q = verts[2u * nvert + 3u * best_tri + best_id];
        pn = make_float3(q.x, q.y, q.z);
    } else {
        const float4 va = verts[tv[0]], vb = verts[tv[1]], vc = verts[tv[2]];
        pn = cross3f(make_float3(vb.x - va.x, vb.y - va.y, vb.z - va.z),
                     make_float3(vc.x - va.x, vc.y - va.y, vc.z - va.z));
    }
    *inside = pn.x * (pt.x - best_p.x) + pn.y * (pt.y - best_p.y) + pn.z * (pt.z - best_p.z) <= 0.0f;
    *out = best_p;
    return true;
}

struct NodeColl {
    float4 normal_id; // normal.xyz, w = collider id (u32 bits)
    float4 point_valid; // point.xyz, w = valid (u32 bits)
};

// collide(), collide.slang:87-154 (primitive shapes and trimeshes; polylines
// take no part, as in the reference).
// World-space box around the region where collider i can report a contact
// (the shape grown by `reach` along each axis, plus a margin far above the
// rounding of collide()'s transforms). A point outside it is outside the shape
// and farther than `reach` from its projection along some axis, so collide()
// takes nothing from the collider there and skips it. A trimesh's box is its
// BVH root box. Polylines (which collide() skips) and unknown shape types (whose projection is the
// point itself: never a contact) get an empty box. MESH: the launch carries the
// trimesh buffers (otherwise no collider is a trimesh).
template <bool MESH>
__device__ inline void collider_box(const Colliders &cs, uint32_t i, float reach, float4 *lo, float4 *hi) {
    const float4 a = cs.shapes[2 * i];
    const uint32_t ty = __float_as_uint(a.w);
    if (ty > 4u && !(MESH && ty == 6u)) {
        *lo = make_float4(1.0e30f, 1.0e30f, 1.0e30f, 0.0f);
        *hi = make_float4(-1.0e30f, -1.0e30f, -1.0e30f, 0.0f);
        return;
    }
    // The shape's local bounding box (centre, half extents).
    float3 c = make_float3(0.0f, 0.0f, 0.0f), e;
    if (ty == 1u) {
        e = make_float3(a.x, a.y, a.z);
    } else if (ty == 0u) {
        e = make_float3(a.x, a.x, a.x);
    } else if (ty == 2u) {
        const float4 b = cs.shapes[2 * i + 1];
        c = make_float3(0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z));
        e = make_float3(0.5f * fabsf(b.x - a.x) + b.w, 0.5f * fabsf(b.y - a.y) + b.w, 0.5f * fabsf(b.z - a.z) + b.w);
    } else if (MESH && ty == 6u) {
        // The BVH root's box (the first node's AABB).
        const uint32_t vroot = __float_as_uint(a.x);
        const float4 mn = cs.mesh_vtx[vroot], mx = cs.mesh_vtx[vroot + 1u];
        c = make_float3(0.5f * (mn.x + mx.x), 0.5f * (mn.y + mx.y), 0.5f * (mn.z + mx.z));
        e = make_float3(0.5f * (mx.x - mn.x), 0.5f * (mx.y - mn.y), 0.5f * (mx.z - mn.z));
    } else {
        e = make_float3(a.y, a.x, a.y);  // cone / cylinder: radius, half height (axis y), radius
    }
    const float4 rot = cs.poses[2 * i], ts = cs.poses[2 * i + 1];
    // World box of the rotated, scaled local box: centre transformed, half extents |R| e |scale|.
    const float3 cw = sim_mul_pt(rot, ts, c);
    const float3 cx = quat_rotate(rot, make_float3(1.0f, 0.0f, 0.0f));
    const float3 cy = quat_rotate(rot, make_float3(0.0f, 1.0f, 0.0f));
    const float3 cz = quat_rotate(rot, make_float3(0.0f, 0.0f, 1.0f));
    const float sc = fabsf(ts.w);
    const float3 ext = make_float3((fabsf(cx.x) * e.x + fabsf(cy.x) * e.y + fabsf(cz.x) * e.z) * sc,
                                   (fabsf(cx.y) * e.x + fabsf(cy.y) * e.y + fabsf(cz.y) * e.z) * sc,
                                   (fabsf(cx.z) * e.x + fabsf(cy.z) * e.y + fabsf(cz.z) * e.z) * sc);
    const float size = fmaxf(fmaxf(ext.x, ext.y), ext.z) + fmaxf(fmaxf(fabsf(cw.x), fabsf(cw.y)), fabsf(cw.z));
    const float m = reach * 1.01f + 1.0e-4f * size;
    *lo = make_float4(cw.x - ext.x - m, cw.y - ext.y - m, cw.z - ext.z - m, 0.0f);
    *hi = make_float4(cw.x + ext.x + m, cw.y + ext.y + m, cw.z + ext.z + m, 0.0f);
}

// `box`: collider_box of each collider (reach 1.5 h).
template <bool MESH>
__device__ inline NodeColl collide(const Colliders &cs, const float4 *box, float h, float3 point) {
#pragma clang fp contract(off)
    NodeColl c;
    c.normal_id = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u));
    c.point_valid = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u));
    float dist_cap = h * 1.5f;
    float best = 1.0e10f;
    for (uint32_t i = 0; i < cs.count; i++) {
        const float4 lo = box[2 * i], hi = box[2 * i + 1];
        if (point.x < lo.x || point.x > hi.x || point.y < lo.y || point.y > hi.y || point.z < lo.z || point.z > hi.z)
            continue;
        float4 a = cs.shapes[2 * i];
        uint32_t ty = __float_as_uint(a.w);
        if (ty == 5u) continue;
        float4 rot = cs.poses[2 * i], ts = cs.poses[2 * i + 1];
        float3 proj;
        bool inside = false;
        if (MESH && ty == 6u) {
            // collide.slang:120-130: only a triangle within dist_cap counts.
            float3 lproj;
            if (!trimesh_project(cs, a, cs.shapes[2 * i + 1], sim_inv_mul_pt(rot, ts, point), dist_cap, &lproj,
                                 &inside))
                continue;
            proj = sim_mul_pt(rot, ts, lproj);
        } else if (ty <= 4u) {
            float3 lp = sim_inv_mul_pt(rot, ts, point);
            float3 lproj;
            if (ty == 1u) {
                inside = cuboid_project_boundary(make_float3(a.x, a.y, a.z), lp, &lproj);
            } else if (ty == 0u) {
                inside = ball_project_boundary(a.x, lp, &lproj);
            } else if (ty == 2u) {
                const float4 b = cs.shapes[2 * i + 1];
                inside = capsule_project_boundary(make_float3(a.x, a.y, a.z), make_float3(b.x, b.y, b.z), b.w, lp,
                                                  &lproj);
            } else if (ty == 3u) {
                inside = cone_project_boundary(a.x, a.y, lp, &lproj);
            } else {
                inside = cylinder_project_boundary(a.x, a.y, lp, &lproj);
            }
            proj = sim_mul_pt(rot, ts, lproj);
        } else {
            proj = point; // shape.slang:200
        }
        float3 d = make_float3(proj.x - point.x, proj.y - point.y, proj.z - point.z);
        if (inside || (fabsf(d.x) <= dist_cap && fabsf(d.y) <= dist_cap && fabsf(d.z) <= dist_cap)) {
            float sign = inside ? -1.0f : 1.0f;
            float dist = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
            if (sign * dist < best && dist != 0.0f) {
                float k = dist * -sign;
                c.normal_id = make_float4(d.x / k, d.y / k, d.z / k, __uint_as_float(i));
                c.point_valid = make_float4(proj.x, proj.y, proj.z, __uint_as_float(1u));
                best = dist;
            }
        }
    }
    return c;
}

// BoundaryCondition::project_velocity, boundary_condition.slang:31-89
__device__ inline float3 project_velocity(uint2 bc, float3 v, float3 n, float dist, float dt) {
    uint32_t ty = bc.x;
    float friction = __uint_as_float(bc.y);
    if (ty == 0u) {
        if (dist <= 0.0f) return make_float3(0.0f, 0.0f, 0.0f);
        return v;
    }
    if (ty == 1u) {
        if (dist <= 0.0f) {
            float nv = v.x * n.x + v.y * n.y + v.z * n.z;
            return make_float3(v.x - n.x * nv, v.y - n.y * nv, v.z - n.z * nv);
        }
        return v;
    }
    if (ty == 2u || ty == 3u) {
        float nv = v.x * n.x + v.y * n.y + v.z * n.z;
        if (nv < 0.0f) {
            if (dist <= 0.0f) {
                float3 tv = make_float3(v.x - n.x * nv, v.y - n.y * nv, v.z - n.z * nv);
                float tl = sqrtf(tv.x * tv.x + tv.y * tv.y + tv.z * tv.z);
                float3 dir = make_float3(0.0f, 0.0f, 0.0f);
                if (tl > 1.0e-8f) dir = make_float3(tv.x / tl, tv.y / tl, tv.z / tl);
                float m = fmaxf(0.0f, tl + friction * nv);
                return make_float3(dir.x * m, dir.y * m, dir.z * m);
            } else if (-nv * dt > dist) {
                float k = dist / dt + nv;
                return make_float3(v.x - k * n.x, v.y - k * n.y, v.z - k * n.z);
            }
        }
        return v;
    }
    return v;
}

// ---------------------------------------------------------------------------
// Particle state views.

struct PBuf {
    float4 *pos;    // xyz, w = (scene << 1 | enabled) bits
    float4 *vm;     // velocity xyz, mass
    float *C;       // [9][pcap]
    float *F;       // [9][pcap]
    float *vgd;     // [pcap]
    uint32_t *ids;  // original index
    float *dps;     // [3][pcap] Drucker-Prager state (sand only)
};

struct PStatic {
    const float4 *prop;  // by original id: (V0, damping, fixed bits, radius)
    const float4 *model; // by original id: (tag bits, lambda, mu, cfl) of the elastic part
    const float4 *dp0;   // by original id: (h0, h1, h2, h3)
    const float4 *dp1;   // by original id: (lambda, mu, 0, 0) of the plastic part
    const float4 *scene_rec;  // batched, one record per scene: prop, model, dp0, dp1 per scene id
};

// A particle's static records: the single record, its scene's record, or its own
// (by original id).
__device__ inline void static_records(const StepParams &P, const PStatic &st, uint32_t oid, uint32_t scene,
                                      float4 *prop, float4 *model) {
    if (P.uniform) {
        *prop = P.u_prop;
        *model = P.u_model;
    } else if (P.scene_uniform) {
        *prop = st.scene_rec[4 * (size_t)scene];
        *model = st.scene_rec[4 * (size_t)scene + 1];
    } else {
        *prop = st.prop[oid];
        *model = st.model[oid];
    }
}

__device__ inline void plastic_records(const StepParams &P, const PStatic &st, uint32_t oid, uint32_t scene,
                                       float4 *dp0, float4 *dp1) {
    if (P.uniform) {
        *dp0 = P.u_dp0;
        *dp1 = P.u_dp1;
    } else if (P.scene_uniform) {
        *dp0 = st.scene_rec[4 * (size_t)scene + 2];
        *dp1 = st.scene_rec[4 * (size_t)scene + 3];
    } else {
        *dp0 = st.dp0[oid];
        *dp1 = st.dp1[oid];
    }
}

// G2P of one particle from the 6x6x6 LDS halo (g2p.slang:150-265):
// v = sum w v_i, C = (4/h^2) sum w v_i dpt_i^T, vel_grad_det = (4/h^2) sum w v_i . dpt_i.
// With dpt_i = ref + s_i h (s_i in {0,1,2}^3 the node's stencil shift), the sums
// are formed from V = sum w v_i and S_a = sum w s_i,a v_i (a = x, y, z):
// sum w v_i dpt_i,a = V ref_a + h S_a.
__device__ inline void g2p_particle(const float4 *__restrict__ sh, float h, float inv_d, float3 p, bool enabled,
                                    float3 *vel, float *vg, float *vgd_out) {
    float V[3] = {0.0f, 0.0f, 0.0f}, S[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    float ref[3] = {0.0f, 0.0f, 0.0f};
    if (enabled) {
        float w[3][3];
        float pp[3] = {p.x, p.y, p.z};
        uint32_t lc[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            float ac = rintf(div_h(pp[i], h)) - 1.0f;
            ref[i] = ac * h - pp[i];
            k_eval_all(div_h(-ref[i], h), w[i]);
            float ab = floorf(ac / 4.0f) * 4.0f;
            lc[i] = (uint32_t)(ac - ab);
        }
        const uint32_t packed = lc[0] + lc[1] * 6u + lc[2] * 36u;
        // V = sum_n w_n v_n and S[d] = sum_n s_d(n) w_n v_n (s_d = the node's stencil
        // offset 0..2 along d), with the separable weight w_n = wx wy wz summed
        // one axis at a time: rows along x first, then y within each z slab, then z.
        const float wx2 = 2.0f * w[0][2], wy2 = 2.0f * w[1][2], wz2 = 2.0f * w[2][2];
#pragma unroll
        for (int sz = 0; sz < 3; sz++) {
            float bz[3] = {0.0f, 0.0f, 0.0f}, myz[3] = {0.0f, 0.0f, 0.0f}, mxz[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int sy = 0; sy < 3; sy++) {
                const uint32_t base = packed + (uint32_t)sy * 6u + (uint32_t)sz * 36u;
                const float4 c0 = sh[base], c1 = sh[base + 1u], c2 = sh[base + 2u];
                const float r[3] = {w[0][0] * c0.x + w[0][1] * c1.x + w[0][2] * c2.x,
                                    w[0][0] * c0.y + w[0][1] * c1.y + w[0][2] * c2.y,
                                    w[0][0] * c0.z + w[0][1] * c1.z + w[0][2] * c2.z};
                const float mx[3] = {w[0][1] * c1.x + wx2 * c2.x, w[0][1] * c1.y + wx2 * c2.y,
                                     w[0][1] * c1.z + wx2 * c2.z};
                const float wy = w[1][sy];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    bz[c] += wy * r[c];
                    mxz[c] += wy * mx[c];
                    if (sy == 1) myz[c] += wy * r[c];
                    if (sy == 2) myz[c] += wy2 * r[c];
                }
            }
            const float wz = w[2][sz];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                V[c] += wz * bz[c];
                S[0][c] += wz * mxz[c];
                S[1][c] += wz * myz[c];
                if (sz == 1) S[2][c] += wz * bz[c];
                if (sz == 2) S[2][c] += wz2 * bz[c];
            }
        }
    }
    *vel = make_float3(V[0], V[1], V[2]);
    float vgd = 0.0f;
#pragma unroll
    for (int r = 0; r < 3; r++) {
#pragma unroll
        for (int c = 0; c < 3; c++) vg[r * 3 + c] = inv_d * (V[c] * ref[r] + h * S[r][c]);
        vgd += V[r] * ref[r] + h * S[r][r];
    }
    *vgd_out = inv_d * vgd;
}

// particle_update.slang:32-160 for one particle. `aff` holds the velocity
// gradient on entry and the APIC affine matrix on exit. No multiply-add
// contraction here or in the stress helpers: F - I and mu (F F^T - I) are small
// differences of O(1) values, so each rounding step shows in them; unfused
// products and sums round like the shader source's separate operations. MODELS is the set of
// material tags the kernel instance handles (bit t = tag t), so scenes without
// SVD-based models do not carry that code's register footprint.
template <uint32_t MODELS, bool F64>
__device__ inline void update_particle(const StepParams &P, float3 &pos, float3 &v, float &mass, uint32_t &enabled,
                                       float *aff, float *F, float4 prop, float4 model, const PStatic &st,
                                       uint32_t oid, uint32_t scene, float *dps) {
#pragma clang fp contract(off)
    const float dt = P.dt, h = P.h;
    float vl = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    if (vl > P.h_dt) {
        float l2 = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
        v = make_float3(v.x / l2 * h / dt, v.y / l2 * h / dt, v.z / l2 * h / dt);
    }
    float damp = 1.0f + prop.y * dt;
    if (damp != 1.0f) v = make_float3(v.x / damp, v.y / damp, v.z / damp);
    bool fixed = __float_as_uint(prop.z) != 0u;
    if (fixed) v = make_float3(0.0f, 0.0f, 0.0f);
    float3 np = make_float3(pos.x + v.x * dt, pos.y + v.y * dt, pos.z + v.z * dt);
    // F += mul(F, affine * dt)
    float adt[9], fa[9];
#pragma unroll
    for (int k = 0; k < 9; k++) adt[k] = aff[k] * dt;
    mul3(F, adt, fa);
#pragma unroll
    for (int k = 0; k < 9; k++) F[k] = F[k] + fa[k];
    if (fixed) {
#pragma unroll
        for (int k = 0; k < 9; k++) aff[k] = 0.0f;
    }
    // DefaultParticleModel::update, default.slang:51-92
    uint32_t tag = __float_as_uint(model.x);
    float stress[9];
    if ((MODELS & 2u) && tag == 1u) {
        nh_stress(model.y, model.z, F, stress);
    } else if ((MODELS & 1u) && tag == 0u) {
        if (F64) fc_stress_f64(model.y, model.z, F, stress);
        else fc_stress(model.y, model.z, F, stress);
    } else if (((MODELS & 4u) && tag == 2u) || ((MODELS & 8u) && tag == 3u)) {
        float4 hp, pl;
        plastic_records(P, st, oid, scene, &hp, &pl);
        if (F64) dp_project_f64(dps, hp, pl.x, pl.y, F);
        else dp_project(dps, hp, pl.x, pl.y, F);
        if (tag == 3u) nh_stress(model.y, model.z, F, stress);
        else if (F64) fc_stress_f64(model.y, model.z, F, stress);
        else fc_stress(model.y, model.z, F, stress);
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) stress[k] = 0.0f;
    }
    float sk = prop.x * P.inv_d * dt;
    bool ok = np.x == np.x && np.y == np.y && np.z == np.z && det3(F) > 0.0f;
    if (ok) {
        pos = np;
#pragma unroll
        for (int k = 0; k < 9; k++) aff[k] = aff[k] * mass - stress[k] * sk;
    } else {
        enabled = 0u;
        v = make_float3(0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int k = 0; k < 9; k++) {
            F[k] = (k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f;
            aff[k] = 0.0f;
        }
        mass = 0.0f;
    }
}

// ---------------------------------------------------------------------------
// A block's particles ("members").

// In the steady state a block's members are the particles of its range in the
// previous substep's storage that are still in the block ("candidates": the
// range is read whole and filtered by block; prev_stay of them pass) and the
// movers listed for it (MOVER_CAP list slots per block, the rest in the
// overflow list). After a full sort they are the sorted list
// perm[blk_start, blk_start + blk_cnt).
struct Members {
    uint32_t fast;
    const uint32_t *blist;                   // block list (header ids = table slots)
    const uint32_t *nbr;                     // neighbour tables (NBR_PREV: id in the previous table)
    const uint4 *bin;                        // per block: range (start, count) and stayers of its previous-table block
    const uint32_t *mcount, *mlist;          // movers received per block, the first MOVER_CAP of them
    const uint32_t *ovf_slot, *ovf_hid;      // the others
    const uint32_t *perm, *blk_start, *blk_cnt;  // the full sort's result
    const uint32_t *counters;                // CTR_NUM_BLOCKS, CTR_OVERFLOW of this substep
};

struct BlockRange {
    uint32_t cand_lo, cand_n;  // candidates: storage slots [cand_lo, cand_lo + cand_n)
    const uint32_t *list;      // then the members list[0, list_n)
    uint32_t list_n;
    uint32_t total;            // members, overflow movers included
    bool overflow;             // some of the block's movers are in the overflow list
};

// FAST (a template parameter, so that each kernel instance loads only the
// buffers of its mode): the steady state's stayer ranges and mover lists.
template <bool FAST>
__device__ inline BlockRange block_range(const Members &M, uint32_t hid) {
    BlockRange r;
    if constexpr (FAST) {
        const uint4 in = M.bin[hid];
        const uint32_t mc = M.mcount[hid];
        r.cand_lo = in.x;
        r.cand_n = in.y;
        r.list = M.mlist + (size_t)hid * MOVER_CAP;
        r.list_n = mc < MOVER_CAP ? mc : MOVER_CAP;
        r.total = in.z + mc;
        r.overflow = mc > MOVER_CAP;
    } else {
        r.cand_lo = 0u;
        r.cand_n = 0u;
        r.list = M.perm + M.blk_start[hid];
        r.list_n = M.blk_cnt[hid];
        r.total = r.list_n;
        r.overflow = false;
    }
    return r;
}

// Storage slot of entry m of a block's sequence (candidates, then list members).
__device__ inline uint32_t entry_slot(const BlockRange &r, uint32_t m) {
    return m < r.cand_n ? r.cand_lo + m : r.list[m - r.cand_n];
}

// Block key of the particle at x (w = scene << 1 | enabled bits): the key the
// sort files it under (a block outside the key space aliases another block of
// its key space, and is filed, counted and processed under that key).
__device__ inline uint32_t key_at(float h, uint32_t batched, float4 x) {
    int bx, by, bz;
    uint32_t cell;
    assoc_block_cell(h, x.x, x.y, x.z, &bx, &by, &bz, &cell);
    return pack_key(batched, __float_as_uint(x.w) >> 1, bx, by, bz);
}

// ---------------------------------------------------------------------------
// Kernels.

// P2G for one block per workgroup of 3 waves.
// Phase 1 (per chunk of up to P2G_CHUNK members of the block): each thread
// loads one particle (stayers are contiguous in storage) and stages in LDS what
// every node of its stencil shares: mass, the APIC momentum at the stencil's
// first node A0 = m v + C^T (x_0 - x_p), the particle's offset f from that node
// in cells (the kernel weights follow from it in a few operations, cheaper than
// staging them: LDS per particle bounds how many workgroups share a CU), and C.
// The chunk is grouped by associated cell in LDS (a rank per particle from an
// LDS counter per cell, a scan of the 64 counts, a list of staging indices).
// Phase 2: lane = associated cell of the block, wave = z offset kz of the
// stencil; each lane accumulates, over the particles of its cell in the chunk,
// the contributions w_n (A0 + C^T (n h)) to the 9 nodes n = (kx, ky, kz) in 36
// registers.
// Phase 3: each wave sums its registers into its own slab tile (z planes kz to
// kz + 3 of the 6x6x6 tile, in the LDS of the staging), in 9 rounds in which its
// 64 lanes touch 64 distinct nodes (plain read-modify-write; no LDS or global
// atomics on floats), all three waves at once; then each node of the tile is
// the sum of the slabs that cover it, in wave order, written straight to HBM.
// Each cell's summation order follows the order its particles were grouped in.
// Cells hold ~8 particles, so lanes read particles ~8 apart: one padding element
// every 16 spreads 16 consecutive cells' first particles over the 16 LDS bank
// groups of 16-byte elements.
__device__ inline uint32_t p2g_slot(uint32_t q) { return q + (q >> 4); }

// Next substep's hash table, reset here and filled with primary blocks by the
// fused G2P / particle-update kernel (replaces k_reset_table + k_touch_primary
// of the next substep).
struct NextTable {
    uint32_t *keys, *counters;
    uint32_t prime;
};

// Block totals and starts of a steady-state substep, in block-list order, by
// one workgroup of 192 threads: total = stayers of the block's previous range +
// movers received. Runs beside the per-block work of the same launch (nothing
// in it reads them).
__device__ void p2g_block_scan(const Members &M, uint32_t *__restrict__ blk_start, uint32_t *__restrict__ blk_cnt,
                               uint32_t *wsum) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t B = M.counters[CTR_NUM_BLOCKS];
    // 24 entries per thread: up to 4608 blocks in one pass of two dependent loads.
    constexpr uint32_t PER = 24u, TILE = 192u * PER;
    uint32_t carry = 0u;
    for (uint32_t base = 0; base < B; base += TILE) {
        uint32_t hid[PER], c[PER];
#pragma unroll
        for (uint32_t j = 0; j < PER; j++) {
            const uint32_t k = base + tid * PER + j;
            hid[j] = k < B ? M.blist[k] : SLOSH_NONE;
        }
        uint32_t t = 0u;
#pragma unroll
        for (uint32_t j = 0; j < PER; j++) {
            c[j] = hid[j] != SLOSH_NONE ? M.bin[hid[j]].z + M.mcount[hid[j]] : 0u;
            t += c[j];
        }
        uint32_t x = t;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(x, d);
            if ((int)lane >= d) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        const uint32_t w0 = wsum[0], w1 = wsum[1], w2 = wsum[2];
        uint32_t e = carry + (wave > 0 ? w0 : 0u) + (wave > 1 ? w1 : 0u) + x - t;
#pragma unroll
        for (uint32_t j = 0; j < PER; j++) {
            if (hid[j] != SLOSH_NONE) {
                blk_start[hid[j]] = e;
                blk_cnt[hid[j]] = c[j];
            }
            e += c[j];
        }
        carry += w0 + w1 + w2;
        __syncthreads();  // wsum is reused by the next tile
    }
}

// LDS element of node (x, y, z) of a P2G slab tile (z = 0..3), 16-byte elements.
// Within a z plane, x < 4 maps to x + 4 y (0..23) and x >= 4 to x + 4 y + 24
// (28..49), one to one;
// over any 4x4 patch of a plane (the 16 lanes of one merge access, at offsets
// 0..2 in x and y) the element index mod 16, its LDS bank group, takes 16
// different values (the patch's columns differ mod 4, and each column's 4 rows
// cover the residues of that class).
#define P2G_SLAB_LDS 200
__device__ inline uint32_t p2g_tile_lds(uint32_t x, uint32_t y, uint32_t z) {
    return x + 4u * y + (x >= 4u ? 24u : 0u) + 50u * z;
}

// 35,088 bytes at 512 particles per chunk: four workgroups (12 waves) fit in a
// CU's 160 KiB of LDS, as many as the registers allow.
template <int P2G_CHUNK>
struct P2GShared {
    static constexpr int LDS = P2G_CHUNK + P2G_CHUNK / 16;
    union {
        // (m, A0), (f, C0), (C1..C4), (C5..C8) per member, grouped by associated cell
        float4 sp4[4][LDS];
        // after the block's last chunk: each wave's sums over its slab (z planes kz..kz+3)
        float4 slab[3][P2G_SLAB_LDS];
        uint32_t extra[192];  // this block's overflow movers of one round: read by
                              // p2g_chunk before it stages anything
    };
    uint32_t lcnt[64];
    uint32_t wsum[4];
};
static_assert(sizeof(float4) * 3 * P2G_SLAB_LDS <= sizeof(P2GShared<192>::sp4),
              "the slab tiles take no LDS beyond the staging of the smallest chunk");
static_assert(sizeof(P2GShared<512>) <= 163840 / 4, "P2G at 512 per chunk must fit four workgroups per CU");

// Phase 1 and 2 for `kn` entries of the block's sequence; entry t's storage slot
// is slot_of(t). An entry is a member if its associated block is this block
// (candidates that moved to another block are skipped; every other entry is one
// by construction). Every thread of the workgroup calls it. `tot` (wave 0, lane
// = cell) accumulates the block's cell counts.
template <int P2G_CHUNK, typename SlotOf>
__device__ inline void p2g_chunk(const StepParams &P, P2GShared<P2G_CHUNK> &S, uint32_t kn, SlotOf slot_of, const float4 *__restrict__ pos, const float4 *__restrict__ vm,
                                 const float *__restrict__ C, const float *__restrict__ fdt, uint32_t pcap,
                                 uint32_t block_key, float (&acc)[9][4], uint32_t &tot) {
    constexpr int R = (P2G_CHUNK + 191) / 192;
    const uint32_t kz = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const float h = P.h;
    const float kzh = (float)kz * h;
    // The slab's z weight is one piece of the quadratic B-spline of the offset f,
    // the same for the whole wave: wz = a (f - b)^2 + c (no divergent branch).
    const uint32_t kzu = __builtin_amdgcn_readfirstlane(kz);
    const float wza = kzu == 1u ? -1.0f : 0.5f;
    const float wzb = kzu == 0u ? 1.5f : (kzu == 1u ? 1.0f : 0.5f);
    const float wzc = kzu == 1u ? 0.75f : 0.0f;
    __syncthreads();  // the previous chunk's readers are done; the cell counters are clear
    // All of this thread's loads are issued before any is consumed.
    uint32_t pp[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t t = threadIdx.x + 192u * r;
        pp[r] = t < kn ? slot_of(t) : SLOSH_NONE;
    }
    float4 ps[R], vv[R];
    float cc[R][9];
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (pp[r] != SLOSH_NONE) {
            ps[r] = pos[pp[r]];
            vv[r] = vm[pp[r]];
#pragma unroll
            for (int k = 0; k < 9; k++) cc[r][k] = C[(size_t)k * pcap + pp[r]];
        }
    }
    // Per member: the associated cell, a rank within it, and the staged terms.
    uint32_t cl[R], rk[R];
    float a0[R][3], fo[R][3];
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (pp[r] == SLOSH_NONE) continue;
        float mom[3] = {vv[r].x * vv[r].w, vv[r].y * vv[r].w, vv[r].z * vv[r].w};
        if (P.has_force) {
#pragma unroll
            for (int i = 0; i < 3; i++) mom[i] += fdt[(size_t)i * pcap + pp[r]];
        }
        // The associated cell: node 0 of the stencil sits at d0 = x_0 - x_p; its
        // block (floor of the cell / 4) and its place in that block.
        const float xp[3] = {ps[r].x, ps[r].y, ps[r].z};
        float d0[3];
        int blk[3], lc[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const float ac = rintf(div_h(xp[i], h)) - 1.0f;
            d0[i] = ac * h - xp[i];
            fo[r][i] = div_h(-d0[i], h);
            const float fb = floorf(ac / 4.0f);
            blk[i] = (int)fb;
            lc[i] = (int)(ac - fb * 4.0f);
        }
        // A candidate that moved to another block is that block's member.
        if (pack_key(P.batched, __float_as_uint(ps[r].w) >> 1, blk[0], blk[1], blk[2]) != block_key) {
            pp[r] = SLOSH_NONE;
            continue;
        }
        const float *c = cc[r];
#pragma unroll
        for (int i = 0; i < 3; i++) a0[r][i] = mom[i] + (d0[0] * c[i] + d0[1] * c[3 + i] + d0[2] * c[6 + i]);
        cl[r] = (uint32_t)(lc[0] + 4 * lc[1] + 16 * lc[2]);
        rk[r] = atomicAdd(&S.lcnt[cl[r]], 1u);
    }
    __syncthreads();
    // Every wave scans the 64 counts (lane = cell) and writes the same starts.
    const uint32_t cnt = S.lcnt[lane];
    uint32_t x = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if ((int)lane >= d) x += y;
    }
    const uint32_t cs = x - cnt;
    if (kz == 0) tot += cnt;
    // Each member's cell start from the lane of its cell (every wave holds all 64).
    uint32_t cstart[R];
#pragma unroll
    for (int r = 0; r < R; r++) cstart[r] = (uint32_t)__shfl((int)cs, pp[r] != SLOSH_NONE ? (int)cl[r] : 0);
    // Staging in cell order: a cell's members are consecutive.
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (pp[r] == SLOSH_NONE) continue;
        const uint32_t q = p2g_slot(cstart[r] + rk[r]);
        S.sp4[0][q] = make_float4(vv[r].w, a0[r][0], a0[r][1], a0[r][2]);
        S.sp4[1][q] = make_float4(fo[r][0], fo[r][1], fo[r][2], cc[r][0]);
        S.sp4[2][q] = make_float4(cc[r][1], cc[r][2], cc[r][3], cc[r][4]);
        S.sp4[3][q] = make_float4(cc[r][5], cc[r][6], cc[r][7], cc[r][8]);
    }
    __syncthreads();
    for (uint32_t j = cs; j < cs + cnt; j++) {
        const uint32_t q = p2g_slot(j);
        const float4 s0 = S.sp4[0][q], s1 = S.sp4[1][q], s2 = S.sp4[2][q], s3 = S.sp4[3][q];
        const float m = s0.x;
        const float a0s[3] = {s0.y, s0.z, s0.w};
        const float tz = s1.z - wzb;
        const float wz = fmaf(wza * tz, tz, wzc);
        float wx[3], wy[3], wyz[3], az[3];
        const float c[9] = {s1.w, s2.x, s2.y, s2.z, s2.w, s3.x, s3.y, s3.z, s3.w};
        k_eval_all(s1.x, wx);
        k_eval_all(s1.y, wy);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            wyz[i] = wy[i] * wz;
            // A0 + C^T (0, 0, kz h): shared by the 9 nodes of the slab.
            az[i] = a0s[i] + kzh * c[6 + i];
        }
#pragma unroll
        for (int ky = 0; ky < 3; ky++) {
            float ay[3];
#pragma unroll
            for (int i = 0; i < 3; i++) ay[i] = ky == 0 ? az[i] : az[i] + (float)ky * h * c[3 + i];
#pragma unroll
            for (int kx = 0; kx < 3; kx++) {
                const int k = kx + 3 * ky;
                const float wgt = wx[kx] * wyz[ky];
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const float a = kx == 0 ? ay[i] : ay[i] + (float)kx * h * c[i];
                    acc[k][i] += a * wgt;
                }
                acc[k][3] += m * wgt;
            }
        }
    }
    __syncthreads();
    if (kz == 0) S.lcnt[lane] = 0u;
}

// Chunks of up to 384 members fit in 128 VGPRs (4 waves/SIMD; their LDS allows 5
// workgroups per CU); 512 needs ~156 (3 waves/SIMD, as its 35 KB of LDS allow).
template <int P2G_CHUNK, bool FAST>
__global__ __launch_bounds__(192, P2G_CHUNK <= 384 ? 4 : 1) void k_p2g(StepParams P, NextTable nt, const float4 *__restrict__ pos,
                                             const float4 *__restrict__ vm, const float *__restrict__ C,
                                             const float *__restrict__ fdt, uint32_t pcap, Members M,
                                             const uint32_t *__restrict__ keys, float4 *__restrict__ tiles,
                                             uint32_t *__restrict__ cell_cnt, uint32_t *__restrict__ blk_start_out,
                                             uint32_t *__restrict__ blk_cnt_out) {
    __shared__ P2GShared<P2G_CHUNK> S;
    if (nt.prime) {
        for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < P.cap; s += gridDim.x * blockDim.x)
            nt.keys[s] = SLOSH_NONE;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            nt.counters[CTR_NUM_BLOCKS] = 0u;
            nt.counters[CTR_SORTED_ALLOC] = 0u;
            nt.counters[CTR_MOVERS] = 0u;
            nt.counters[CTR_OVERFLOW] = 0u;
        }
    }
    // Steady state: workgroup 0 computes the block totals and starts.
    const uint32_t first = FAST ? 1u : 0u;
    if (FAST && blockIdx.x == 0) {
        p2g_block_scan(M, blk_start_out, blk_cnt_out, S.wsum);
        return;
    }
    const uint32_t B = M.counters[CTR_NUM_BLOCKS];
    const uint32_t n_ovf = M.counters[CTR_OVERFLOW];
    const uint32_t kz = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;

    if (threadIdx.x < 64) S.lcnt[threadIdx.x] = 0u;
    for (uint32_t k = blockIdx.x - first; k < B; k += gridDim.x - first) {
        const uint32_t hid = M.blist[k];
        const BlockRange br = block_range<FAST>(M, hid);
        // A block without particles (a "+1" neighbour only) has an all-zero tile:
        // nothing is written and the grid update skips it.
        if (br.total == 0) {
            if (threadIdx.x < 64) cell_cnt[(size_t)hid * NODES_PER_BLOCK + threadIdx.x] = 0u;
            continue;
        }
        const uint32_t block_key = keys[hid];
        float acc[9][4];
#pragma unroll
        for (int q = 0; q < 9; q++) acc[q][0] = acc[q][1] = acc[q][2] = acc[q][3] = 0.0f;
        uint32_t tot = 0u;
        const uint32_t n_main = br.cand_n + br.list_n;
        for (uint32_t chunk = 0; chunk < n_main; chunk += P2G_CHUNK) {
            const uint32_t kn = n_main - chunk < (uint32_t)P2G_CHUNK ? n_main - chunk : (uint32_t)P2G_CHUNK;
            p2g_chunk<P2G_CHUNK>(P, S, kn, [&](uint32_t t) { return entry_slot(br, chunk + t); }, pos, vm, C, fdt,
                                 pcap, block_key, acc, tot);
        }
        if (br.overflow) {
            // This block's movers beyond its list: picked out of the overflow list,
            // one round of 192 entries at a time.
            for (uint32_t ob = 0; ob < n_ovf; ob += 192u) {
                const uint32_t i = ob + threadIdx.x;
                const bool mine = i < n_ovf && M.ovf_hid[i] == hid;
                const uint64_t mm = __ballot(mine);
                __syncthreads();  // the previous round's readers of `extra` are done
                if (lane == 0) S.wsum[kz] = (uint32_t)__popcll(mm);
                __syncthreads();
                const uint32_t before = (kz > 0 ? S.wsum[0] : 0u) + (kz > 1 ? S.wsum[1] : 0u);
                const uint32_t kn = S.wsum[0] + S.wsum[1] + S.wsum[2];
                if (mine) S.extra[before + (uint32_t)__popcll(mm & lanemask_lt())] = M.ovf_slot[i];
                if (kn > 0)
                    p2g_chunk<P2G_CHUNK>(P, S, kn, [&](uint32_t t) { return S.extra[t]; }, pos, vm, C, fdt, pcap,
                                         block_key, acc, tot);
            }
        }
        // Each wave adds its registers into its own slab tile (sharing the staging's
        // LDS), in 9 rounds in which its 64 lanes touch 64 distinct nodes; then each
        // node is the sum of the slabs that cover it, in wave order.
        __syncthreads();  // the staging's readers are done
        float4 *slab = S.slab[kz];
        for (uint32_t j = lane; j < P2G_SLAB_LDS; j += 64) slab[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
#pragma unroll
        for (int ky = 0; ky < 3; ky++)
#pragma unroll
            for (int kx = 0; kx < 3; kx++) {
                const int q = kx + 3 * ky;
                const uint32_t node = p2g_tile_lds(lx + kx, ly + ky, lz);
                float4 t = slab[node];
                t.x += acc[q][0];
                t.y += acc[q][1];
                t.z += acc[q][2];
                t.w += acc[q][3];
                slab[node] = t;
                // The next round's nodes overlap other lanes' nodes of this
                // round: complete this round's LDS writes first.
                __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < TILE_NODES; j += 192) {
            const uint32_t x = j % 6u, y = (j / 6u) % 6u, z = j / 36u;
            float4 t = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
            for (uint32_t w = 0; w < 3; w++) {
                if (z < w || z > w + 3u) continue;
                const float4 u = S.slab[w][p2g_tile_lds(x, y, z - w)];
                t.x += u.x;
                t.y += u.y;
                t.z += u.z;
                t.w += u.w;
            }
            tiles[(size_t)hid * TILE_NODES + tile_index(x, y, z)] = t;
        }
        if (kz == 0) cell_cnt[(size_t)hid * NODES_PER_BLOCK + lane] = tot;
    }
}

// mode 0: node momentum from tiles; mode 1: nodes already hold momentum (P2G hook).
// do_update 0: only combine tiles (grid update disabled).
template <bool MESH>
__global__ __launch_bounds__(256) void k_grid_update(StepParams P, const float4 *__restrict__ tiles,
                                                     float4 *__restrict__ nodes, const uint32_t *__restrict__ nbr,
                                                     const uint32_t *__restrict__ blist,
                                                     const uint32_t *__restrict__ keys, Colliders cs,
                                                     const NodeColl *__restrict__ coll_prev, NodeColl *__restrict__ coll_cur,
                                                     const uint32_t *__restrict__ counters,
                                                     const uint32_t *__restrict__ blk_cnt, uint32_t mode,
                                                     uint32_t do_update) {
    __shared__ float4 cbox[2 * SLOSH_MAX_COLLIDERS];
    if (threadIdx.x < cs.count) collider_box<MESH>(cs, threadIdx.x, 1.5f * P.h, &cbox[2 * threadIdx.x], &cbox[2 * threadIdx.x + 1]);
    __syncthreads();
    const uint32_t B = counters[CTR_NUM_BLOCKS];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const int x = lane & 3, y = (lane >> 2) & 3, z = lane >> 4;
    const uint32_t waves_per_grid = gridDim.x * 4;
    for (uint32_t k = blockIdx.x * 4 + wave; k < B; k += waves_per_grid) {
        const uint32_t hid = blist[k];
        float4 acc;
        if (mode == 0) {
            // Tiles of blocks without particles were not written (they count as
            // zero). Each tile value is loaded together with its block's particle
            // count and then selected by it, so the tile loads wait only for the
            // neighbour ids, not for the counts (every tile lies inside the
            // allocation, written or not). The part of a "-1" neighbour's tile
            // that lands on this block is one contiguous region of that tile.
            float4 t[8];
            uint32_t c[8];
            t[0] = tiles[(size_t)hid * TILE_NODES + x + 4 * y + 16 * z];
            c[0] = blk_cnt[hid];
#pragma unroll
            for (int d = 1; d < 8; d++) {
                const int ox = ASSOC[d][0], oy = ASSOC[d][1], oz = ASSOC[d][2];
                t[d] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                c[d] = 0u;
                if ((ox == 0 || x <= 1) && (oy == 0 || y <= 1) && (oz == 0 || z <= 1)) {
                    const uint32_t src = nbr[(size_t)hid * NBR_STRIDE + NBR_MINUS + d - 1];
                    if (src != SLOSH_NONE) {
                        t[d] = tiles[(size_t)src * TILE_NODES + tile_index(x + 4 * ox, y + 4 * oy, z + 4 * oz)];
                        c[d] = blk_cnt[src];
                    }
                }
            }
            acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
            for (int d = 0; d < 8; d++) {
                // Bitwise select: a stale tile (possibly NaN) never enters the sum.
                const bool on = c[d] != 0u;
                const float4 v = make_float4(__uint_as_float(on ? __float_as_uint(t[d].x) : 0u),
                                             __uint_as_float(on ? __float_as_uint(t[d].y) : 0u),
                                             __uint_as_float(on ? __float_as_uint(t[d].z) : 0u),
                                             __uint_as_float(on ? __float_as_uint(t[d].w) : 0u));
                if (d == 0) {
                    acc = v;
                } else {
                    acc.x += v.x;
                    acc.y += v.y;
                    acc.z += v.z;
                    acc.w += v.w;
                }
            }
        } else {
            acc = nodes[(size_t)hid * NODES_PER_BLOCK + lane];
        }
        if (!do_update) {
            nodes[(size_t)hid * NODES_PER_BLOCK + lane] = acc;
            continue;
        }
        int bx, by, bz;
        uint32_t scene;
        unpack_key(P.batched, keys[hid], &bx, &by, &bz, &scene);
        const float h = P.h, dt = P.dt;
        float3 cell_pos = make_float3((float)(bx * 4 + x) * h, (float)(by * 4 + y) * h, (float)(bz * 4 + z) * h);
        // update_single_cell, grid_update.slang:154-163
        const float mass = acc.w;
        const float inv_mass = mass > 0.0f ? 1.0f / mass : 0.0f;
        const float lim = P.h_dt;
        float3 v;
        v.x = fminf(fmaxf((acc.x + mass * P.gravity[0] * dt) * inv_mass, -lim), lim);
        v.y = fminf(fmaxf((acc.y + mass * P.gravity[1] * dt) * inv_mass, -lim), lim);
        v.z = fminf(fmaxf((acc.z + mass * P.gravity[2] * dt) * inv_mass, -lim), lim);
        if (cs.count > 0) {
            NodeColl c;
            bool have = true;
            if (P.coll_cache) {
                uint32_t prev = nbr[(size_t)hid * NBR_STRIDE + NBR_PREV];
                if (prev != SLOSH_NONE && coll_prev != nullptr) c = coll_prev[(size_t)prev * NODES_PER_BLOCK + lane];
                else c = collide<MESH>(cs, cbox, h, cell_pos);
                coll_cur[(size_t)hid * NODES_PER_BLOCK + lane] = c;
            } else if (mass > 0.0f) {
                // Static colliders: the cached record equals a fresh collide(); a
                // zero-mass node has v = 0, which every boundary condition keeps.
                c = collide<MESH>(cs, cbox, h, cell_pos);
            } else {
                have = false;
            }
            if (have && __float_as_uint(c.point_valid.w) != 0u) {
                uint2 bc = cs.mats[__float_as_uint(c.normal_id.w)];
                float3 n = make_float3(c.normal_id.x, c.normal_id.y, c.normal_id.z);
                float3 dd = make_float3(cell_pos.x - c.point_valid.x, cell_pos.y - c.point_valid.y,
                                        cell_pos.z - c.point_valid.z);
                float dist = (dd.x * n.x + dd.y * n.y + dd.z * n.z) - h;
                v = project_velocity(bc, v, n, dist, dt);
            }
        }
        nodes[(size_t)hid * NODES_PER_BLOCK + lane] = make_float4(v.x, v.y, v.z, mass);
    }
}

struct Bodies {
    float4 *poses;        // 2 per body
    float4 *vels;         // 2 per body: linear, angular
    const float4 *mprops; // 5 per body: inv_inertia rows 0..2, inv_mass, com (world)
    const float4 *local_com; // 1 per body
    int *impulses;        // 6 per body: linear xyz, angular xyz (fixed point x 1e5)
    uint32_t count;
};

__device__ inline float len3(float3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }

// rigid_impulses.slang:70-121 with body.slang:63-130, for body `id`.
__device__ void integrate_body(const StepParams &P, const Bodies &b, uint32_t id) {
    const float dt = P.dt;
    float3 lin, ang;
    lin.x = (float)b.impulses[6 * id + 0] / 1.0e5f;
    lin.y = (float)b.impulses[6 * id + 1] / 1.0e5f;
    lin.z = (float)b.impulses[6 * id + 2] / 1.0e5f;
    ang.x = (float)b.impulses[6 * id + 3] / 1.0e5f;
    ang.y = (float)b.impulses[6 * id + 4] / 1.0e5f;
    ang.z = (float)b.impulses[6 * id + 5] / 1.0e5f;
#pragma unroll
    for (int k = 0; k < 6; k++) b.impulses[6 * id + k] = 0;
    const float4 ii0 = b.mprops[5 * id + 0], ii1 = b.mprops[5 * id + 1], ii2 = b.mprops[5 * id + 2];
    const float4 im = b.mprops[5 * id + 3];
    float4 vl = b.vels[2 * id], va = b.vels[2 * id + 1];
    float3 nl = make_float3(vl.x + im.x * lin.x, vl.y + im.y * lin.y, vl.z + im.z * lin.z);
    float3 na = make_float3(va.x + (ang.x * ii0.x + ang.y * ii1.x + ang.z * ii2.x),
                            va.y + (ang.x * ii0.y + ang.y * ii1.y + ang.z * ii2.y),
                            va.z + (ang.x * ii0.z + ang.y * ii1.z + ang.z * ii2.z));
    float ln = len3(nl), an = len3(na);
    float lin_limit = 0.1f * P.h / dt;
    if (len3(lin) != 0.0f || len3(ang) != 0.0f) {
        if (ln > lin_limit) nl = make_float3(nl.x * (lin_limit / ln), nl.y * (lin_limit / ln), nl.z * (lin_limit / ln));
        if (an > 1.0f) na = make_float3(na.x * (1.0f / an), na.y * (1.0f / an), na.z * (1.0f / an));
    }
    // integrate_velocity
    float4 rot = b.poses[2 * id], ts = b.poses[2 * id + 1];
    float4 lc = b.local_com[id];
    float3 init_com = sim_mul_pt(rot, ts, make_float3(lc.x, lc.y, lc.z));
    float3 sa = make_float3(na.x * dt, na.y * dt, na.z * dt);
    float angle = len3(sa);
    float4 dq = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    if (angle != 0.0f) {
        float s = sinf(angle * 0.5f) / angle;
        dq = make_float4(sa.x * s, sa.y * s, sa.z * s, cosf(angle * 0.5f));
    }
    float3 rr = quat_rotate(dq, make_float3(ts.x - init_com.x, ts.y - init_com.y, ts.z - init_com.z));
    float4 nq;
    nq.x = dq.w * rot.x + dq.x * rot.w + dq.y * rot.z - dq.z * rot.y;
    nq.y = dq.w * rot.y - dq.x * rot.z + dq.y * rot.w + dq.z * rot.x;
    nq.z = dq.w * rot.z + dq.x * rot.y - dq.y * rot.x + dq.z * rot.w;
    nq.w = dq.w * rot.w - dq.x * rot.x - dq.y * rot.y - dq.z * rot.z;
    float sq = nq.x * nq.x + nq.y * nq.y + nq.z * nq.z + nq.w * nq.w;
    float k = 0.5f * (3.0f - sq);
    b.poses[2 * id] = make_float4(nq.x * k, nq.y * k, nq.z * k, nq.w * k);
    b.poses[2 * id + 1] = make_float4(init_com.x + rr.x * ts.w + nl.x * dt, init_com.y + rr.y * ts.w + nl.y * dt,
                                      init_com.z + rr.z * ts.w + nl.z * dt, ts.w);
    nl.x += P.gravity[0] * (im.x != 0.0f ? 1.0f : 0.0f) * dt;
    nl.y += P.gravity[1] * (im.y != 0.0f ? 1.0f : 0.0f) * dt;
    nl.z += P.gravity[2] * (im.z != 0.0f ? 1.0f : 0.0f) * dt;
    b.vels[2 * id] = make_float4(nl.x, nl.y, nl.z, 0.0f);
    b.vels[2 * id + 1] = make_float4(na.x, na.y, na.z, 0.0f);
}


__global__ void k_rigid_update(StepParams P, Bodies b) {
    if (threadIdx.x < b.count) integrate_body(P, b, threadIdx.x);
}

__device__ inline void load_halo(float4 *__restrict__ sh, const float4 *__restrict__ nodes,
                                 const uint32_t *__restrict__ nbr, uint32_t hid) {
    for (uint32_t j = threadIdx.x; j < TILE_NODES; j += blockDim.x) {
        uint32_t sx = j % 6u, sy = (j / 6u) % 6u, sz = j / 36u;
        uint32_t ox = sx >= 4u, oy = sy >= 4u, oz = sz >= 4u;
        uint32_t oi = ox * 4u + oy * 2u + oz;
        uint32_t src = oi == 0u ? hid : nbr[(size_t)hid * NBR_STRIDE + oi - 1u];
        uint32_t node = (sx - 4u * ox) + 4u * (sy - 4u * oy) + 16u * (sz - 4u * oz);
        sh[j] = src != SLOSH_NONE ? nodes[(size_t)src * NODES_PER_BLOCK + node] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// What the fused G2P kernel prepares for the next substep: its table (keys
// inserted, a new block's mover count cleared), the list of particles that
// leave their block, and every block's stayer count.
struct NextSort {
    uint32_t *mcount;         // next table: movers per block (cleared at insertion)
    uint32_t *mv_key, *mv_slot;  // movers: new block key, output slot
    uint32_t *stay;           // this table: particles of each block that stay in it
    uint32_t *errors;         // this substep's error word (ERR_KEY_RANGE)
};

// F64: the SVD-based steps refined to FP64 accuracy. The bound keeps that
// variant at three waves per SIMD (168 VGPRs).
//
// Output order: a block's members take the output slots [start, start + total)
// of the block in order of arrival of their wavefronts (one LDS atomic per
// wavefront), so a block's particles stay together in storage. The particles
// whose updated position leaves the block are listed, with their new block key,
// for the next substep's sort; the others are counted (the block's stayers).
template <uint32_t MODELS, int NT, bool F64, bool FAST>
__global__ __launch_bounds__(NT, F64 ? (MODELS == 3u ? 4 : 3) : 1) void k_g2p_update(StepParams P, PBuf in, PBuf out, PStatic st, uint32_t pcap,
                                                    float *__restrict__ fdt, Members M,
                                                    const uint32_t *__restrict__ blk_start,
                                                    const float4 *__restrict__ nodes,
                                                    const uint32_t *__restrict__ keys, NextTable nt, NextSort ns,
                                                    Bodies bodies) {
    __shared__ float4 sh[TILE_NODES];
    __shared__ uint32_t n_out, n_stay;
    // integrate_bodies: reads no particle or grid state (the impulses it applies
    // are only written by the host), so it runs alongside the particle update.
    if (blockIdx.x == 0 && threadIdx.x < bodies.count) integrate_body(P, bodies, threadIdx.x);
    const uint32_t B = M.counters[CTR_NUM_BLOCKS];
    const uint32_t n_ovf = M.counters[CTR_OVERFLOW];
    const uint32_t lane = threadIdx.x & 63u;
    BlockInit init;
    init.mcount = ns.mcount;
    init.blk_cnt = nullptr;
    init.cell_cnt = nullptr;
    // A block with stayers inserts its 8 blocks into the next table while the
    // workgroup works on its next block (threads 0-7, key in `pk`): the first
    // probe's load is issued with the next block's halo loads.
    uint32_t pk = SLOSH_NONE;
    for (uint32_t k = blockIdx.x; k < B; k += gridDim.x) {
        uint32_t probe = SLOSH_NONE;
        if (pk != SLOSH_NONE) probe = nt.keys[hash_key(pk) & (P.cap - 1u)];
        const uint32_t hid = M.blist[k];
        const BlockRange br = block_range<FAST>(M, hid);
        const uint32_t cnt = br.total;
        if (cnt == 0) {
            if (threadIdx.x == 0) ns.stay[hid] = 0u;
            if (pk != SLOSH_NONE && probe != pk) hash_insert(nt.keys, P.cap, pk, nt.counters, init);
            pk = SLOSH_NONE;
            continue;
        }
        const uint32_t s0 = blk_start[hid];
        const uint32_t block_key = keys[hid];
        int bx, by, bz;
        uint32_t scene;
        unpack_key(P.batched, block_key, &bx, &by, &bz, &scene);
        bool stays = false;
        load_halo(sh, nodes, M.nbr, hid);
        if (threadIdx.x == 0) {
            n_out = 0u;
            n_stay = 0u;
        }
        __syncthreads();
        if (pk != SLOSH_NONE && probe != pk) hash_insert(nt.keys, P.cap, pk, nt.counters, init);
        pk = SLOSH_NONE;
        // The block's entries (a candidate is a member only if it is still in the
        // block), then, only if some of its movers overflowed their list, the
        // overflow list, whose entries of other blocks are skipped. Per member:
        // G2P, update, output at the next free slot of the block's range (one LDS
        // atomic per wavefront), listed if it leaves the block, its blocks
        // inserted into the next table. Every lane of the workgroup runs the same
        // number of iterations (the wave-level counting and run detection see all
        // lanes).
        const uint32_t n_main = br.cand_n + br.list_n;
        const uint32_t n_all = n_main + (br.overflow ? n_ovf : 0u);
        for (uint32_t jb = 0; jb < n_all; jb += blockDim.x) {
            const uint32_t m = jb + threadIdx.x;
            bool entry = m < n_all;
            uint32_t p = 0u;
            if (m < n_main) {
                p = entry_slot(br, m);
            } else if (entry) {
                entry = M.ovf_hid[m - n_main] == hid;
                p = M.ovf_slot[m - n_main];
            }
            float4 ps = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            // The neo-Hookean-only instance has registers to spare: it issues the
            // deformation gradient and id loads with the position load (nearly
            // every entry is a member), the others after the G2P sums.
            constexpr bool EARLY = MODELS == 2u;
            float F[9];
            uint32_t oid = 0u;
            if (entry) {
                ps = in.pos[p];
                if (EARLY) {
#pragma unroll
                    for (int q = 0; q < 9; q++) F[q] = in.F[(size_t)q * pcap + p];
                    oid = in.ids[p];
                }
            }
            const bool valid = entry && key_at(P.h, P.batched, ps) == block_key;
            const uint64_t vmask = __ballot(valid);
            uint32_t ob = 0u;
            if (lane == 0 && vmask) ob = atomicAdd(&n_out, (uint32_t)__popcll(vmask));
            const uint32_t j = __shfl(ob, 0) + (uint32_t)__popcll(vmask & lanemask_lt());
            const uint32_t slot = s0 + j;
            uint32_t next_key = 0u;
            if (valid) {
                const uint32_t wbits = __float_as_uint(ps.w);
                uint32_t enabled = wbits & 1u;
                float3 pos = make_float3(ps.x, ps.y, ps.z);
                float3 v;
                float aff[9];
                float vgd;
                g2p_particle(sh, P.h, P.inv_d, pos, enabled != 0u, &v, aff, &vgd);
                const uint32_t scene = wbits >> 1;
                // Uniform masses (one for all, or one per scene) skip the 16-byte
                // velocity / mass load (the velocity comes from the grid): only the
                // update disables, and it zeroes the mass.
                float mass = P.mass_uniform ? (enabled ? P.u_mass : 0.0f)
                             : P.scene_mass  ? (enabled ? st.scene_rec[4 * (size_t)scene + 3].z : 0.0f)
                                             : in.vm[p].w;
                if (!EARLY) {
#pragma unroll
                    for (int q = 0; q < 9; q++) F[q] = in.F[(size_t)q * pcap + p];
                    oid = in.ids[p];
                }
                float4 prop, model;
                static_records(P, st, oid, scene, &prop, &model);
                // Only Drucker-Prager particles read or write their plastic state
                // (the other particles' state slots are never read).
                const uint32_t tag = __float_as_uint(model.x);
                const bool plastic = P.has_sand && (tag == 2u || tag == 3u);
                float dps[3] = {0.0f, 0.0f, 0.0f};
                if (plastic) {
#pragma unroll
                    for (int q = 0; q < 3; q++) dps[q] = in.dps[(size_t)q * pcap + p];
                }
                update_particle<MODELS, F64>(P, pos, v, mass, enabled, aff, F, prop, model, st, oid, scene, dps);
                out.pos[slot] = make_float4(pos.x, pos.y, pos.z, __uint_as_float((wbits & ~1u) | enabled));
                out.vm[slot] = make_float4(v.x, v.y, v.z, mass);
#pragma unroll
                for (int q = 0; q < 9; q++) {
                    out.C[(size_t)q * pcap + slot] = aff[q];
                    out.F[(size_t)q * pcap + slot] = F[q];
                }
                out.vgd[slot] = vgd;
                out.ids[slot] = oid;
                if (plastic) {
#pragma unroll
                    for (int q = 0; q < 3; q++) out.dps[(size_t)q * pcap + slot] = dps[q];
                }
                if (P.has_force) {
                    fdt[slot] = 0.0f;
                    fdt[(size_t)pcap + slot] = 0.0f;
                    fdt[2 * (size_t)pcap + slot] = 0.0f;
                }
                next_key = key_at(P.h, P.batched, make_float4(pos.x, pos.y, pos.z, __uint_as_float(wbits)));
            }
            const bool mover = valid && next_key != block_key;
            stays |= valid && !mover;
            const uint64_t sm = __ballot(valid && !mover);
            if (lane == 0 && sm) atomicAdd(&n_stay, (uint32_t)__popcll(sm));
            const uint64_t mv = __ballot(mover);
            if (mv) {
                // Only a particle that leaves its block can leave the key space (a
                // block outside it would alias another one): reported, once per
                // wavefront.
                bool out_of_range = false;
                if (mover) {
                    const float4 np = out.pos[slot];
                    int nx, ny, nz;
                    uint32_t cell;
                    assoc_block_cell(P.h, np.x, np.y, np.z, &nx, &ny, &nz, &cell);
                    out_of_range = !block_in_key_range(P.batched, nx, ny, nz);
                }
                if (lane == 0 && __ballot(out_of_range)) atomicOr(ns.errors, ERR_KEY_RANGE);
                if (nt.prime) {
                    // The movers, listed for the next substep's sort.
                    uint32_t lb = 0u;
                    if (lane == 0) lb = atomicAdd(&nt.counters[CTR_MOVERS], (uint32_t)__popcll(mv));
                    lb = __shfl(lb, 0);
                    if (mover) {
                        const uint32_t e = lb + (uint32_t)__popcll(mv & lanemask_lt());
                        ns.mv_key[e] = next_key;
                        ns.mv_slot[e] = slot;
                    }
                }
            }
            if (nt.prime) {
                // touch_primary_blocks and touch_neighbor_blocks of the next
                // substep (sort.slang:50-92). Particles that stay in this block
                // only flag it (its 8 blocks are inserted once below); the others
                // insert theirs, one round per run of equal keys.
                const uint64_t heads = run_heads(next_key, mover) & mv;
                insert_blocks(nt.keys, P.cap, P.batched, next_key, heads, nt.counters, init);
            }
        }
        // The barrier also ends this block's reads of `sh`, `n_out` and `n_stay`.
        const bool any_stays = __syncthreads_or(stays);
        if (threadIdx.x == 0) ns.stay[hid] = n_stay;
        if (nt.prime && any_stays && threadIdx.x < 8)
            pk = pack_key(P.batched, scene, bx + BLOCK_ASSOC[threadIdx.x][0], by + BLOCK_ASSOC[threadIdx.x][1],
                          bz + BLOCK_ASSOC[threadIdx.x][2]);
    }
    if (pk != SLOSH_NONE) hash_insert(nt.keys, P.cap, pk, nt.counters, init);
}

// WgTimestepBounds (timestep_bound.slang:18-92): per enabled particle, the
// material's P-wave sound-speed bound cfl h / max(|v|, c) (models/utils.slang:
// 119-135; elastic part of sand models, default.slang:94-127) and the velocity
// bound h / |v|, as the u32 fixed point of GpuTimestepBounds::secs_to_int,
// reduced with atomicMin into *out.
__device__ inline uint32_t secs_to_int(float secs) {
    const float x = floorf(secs * 1.0e12f);
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967295.0f) return 0xffffffffu;
    return (uint32_t)x;
}

// The bound of particle i (0xffffffff for a disabled particle).
__device__ inline uint32_t particle_dt_bound(const StepParams &P, const PBuf &buf, const PStatic &st, uint32_t pcap,
                                             uint32_t i) {
#pragma clang fp contract(off)
    const float h = P.h;
    const float4 ps = buf.pos[i];
    if ((__float_as_uint(ps.w) & 1u) == 0u) return 0xffffffffu;
    const float4 vm = buf.vm[i];
    const uint32_t oid = buf.ids[i];
    float4 prop, model;
    static_records(P, st, oid, __float_as_uint(ps.w) >> 1, &prop, &model);
    float F[9];
#pragma unroll
    for (int k = 0; k < 9; k++) F[k] = buf.F[(size_t)k * pcap + i];
    const float speed = sqrtf(vm.x * vm.x + vm.y * vm.y + vm.z * vm.z);
    float dt = 0.0f;
    if (__float_as_uint(model.x) < 4u) {
        const float lambda = model.y, mu = model.z, cfl = model.w;
        const float bulk = lambda + 2.0f * mu / 3.0f;
        const float shear = mu;
        const float density0 = vm.w / prop.x;
        const float curr_density = density0 / fmaxf(det3(F), 1.0e-6f);
        const float sound_speed = sqrtf((bulk + shear * 4.0f / 3.0f) / curr_density);
        dt = cfl * h / fmaxf(speed, sound_speed);
    }
    dt = fminf(dt, h / speed);
    return secs_to_int(dt);
}

__device__ inline uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}

__global__ __launch_bounds__(256) void k_timestep_bound(StepParams P, PBuf buf, PStatic st, uint32_t pcap,
                                                        uint32_t *__restrict__ out) {
    uint32_t best = 0xffffffffu;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P.n; i += gridDim.x * blockDim.x)
        best = min(best, particle_dt_bound(P, buf, st, pcap, i));
    best = wave_min_u32(best);
    if (lane_id() == 0 && best != 0xffffffffu) atomicMin(out, best);
}

// The same bound in two passes with no atomics and nothing to clear first:
// workgroup g stores the minimum over particles [g * TSB_CHUNK, (g + 1) *
// TSB_CHUNK) into partials[g] (0xffffffff if none is enabled), then one
// workgroup stores the minimum of the partials into *out. Integer minima, so
// the result is bit-identical to k_timestep_bound's.
__global__ __launch_bounds__(256) void k_timestep_bound_partials(StepParams P, PBuf buf, PStatic st, uint32_t pcap,
                                                                 uint32_t *__restrict__ partials) {
    __shared__ uint32_t wmin[4];
    uint32_t best = 0xffffffffu;
    const uint32_t base = blockIdx.x * TSB_CHUNK;
    const uint32_t end = min(P.n, base + TSB_CHUNK);
    for (uint32_t i = base + threadIdx.x; i < end; i += 256) best = min(best, particle_dt_bound(P, buf, st, pcap, i));
    best = wave_min_u32(best);
    if (lane_id() == 0) wmin[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = min(min(wmin[0], wmin[1]), min(wmin[2], wmin[3]));
}

__global__ __launch_bounds__(256) void k_min_partials(const uint32_t *__restrict__ partials, uint32_t count,
                                                      uint32_t *__restrict__ out) {
    __shared__ uint32_t wmin[4];
    uint32_t best = 0xffffffffu;
    for (uint32_t i = threadIdx.x; i < count; i += 256) best = min(best, partials[i]);
    best = wave_min_u32(best);
    if (lane_id() == 0) wmin[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) *out = min(min(wmin[0], wmin[1]), min(wmin[2], wmin[3]));
}

// G2P only, in place (hook path): writes velocity, velocity gradient, vel_grad_det.
__global__ __launch_bounds__(256) void k_g2p(StepParams P, PBuf buf, uint32_t pcap, const uint32_t *__restrict__ perm,
                                             const uint32_t *__restrict__ blist,
                                             const uint32_t *__restrict__ blk_start,
                                             const uint32_t *__restrict__ blk_cnt, const uint32_t *__restrict__ nbr,
                                             const float4 *__restrict__ nodes, const uint32_t *__restrict__ counters) {
    __shared__ float4 sh[TILE_NODES];
    const uint32_t B = counters[CTR_NUM_BLOCKS];
    for (uint32_t k = blockIdx.x; k < B; k += gridDim.x) {
        const uint32_t hid = blist[k];
        const uint32_t s0 = blk_start[hid], cnt = blk_cnt[hid];
        if (cnt == 0) continue;
        load_halo(sh, nodes, nbr, hid);
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < cnt; j += blockDim.x) {
            const uint32_t p = perm[s0 + j];
            const float4 ps = buf.pos[p];
            float3 v;
            float g[9];
            float vgd;
            g2p_particle(sh, P.h, P.inv_d, make_float3(ps.x, ps.y, ps.z), (__float_as_uint(ps.w) & 1u) != 0u, &v, g,
                         &vgd);
            float4 o = buf.vm[p];
            buf.vm[p] = make_float4(v.x, v.y, v.z, o.w);
#pragma unroll
            for (int k = 0; k < 9; k++) buf.C[(size_t)k * pcap + p] = g[k];
            buf.vgd[p] = vgd;
        }
        __syncthreads();
    }
}

// Particle update only, in place (hook path), over every particle.
template <uint32_t MODELS, bool F64>
__global__ __launch_bounds__(256, F64 ? 3 : 1) void k_particle_update(StepParams P, PBuf buf, PStatic st, uint32_t pcap,
                                                         float *__restrict__ fdt) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    const float4 ps = buf.pos[p];
    const uint32_t wbits = __float_as_uint(ps.w);
    uint32_t enabled = wbits & 1u;
    float3 pos = make_float3(ps.x, ps.y, ps.z);
    float4 vmv = buf.vm[p];
    float3 v = make_float3(vmv.x, vmv.y, vmv.z);
    float mass = vmv.w;
    float aff[9], F[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        aff[k] = buf.C[(size_t)k * pcap + p];
        F[k] = buf.F[(size_t)k * pcap + p];
    }
    const uint32_t oid = buf.ids[p];
    float dps[3] = {0.0f, 0.0f, 0.0f};
    if (P.has_sand) {
#pragma unroll
        for (int k = 0; k < 3; k++) dps[k] = buf.dps[(size_t)k * pcap + p];
    }
    float4 prop, model;
    static_records(P, st, oid, wbits >> 1, &prop, &model);
    update_particle<MODELS, F64>(P, pos, v, mass, enabled, aff, F, prop, model, st, oid, wbits >> 1, dps);
    buf.pos[p] = make_float4(pos.x, pos.y, pos.z, __uint_as_float((wbits & ~1u) | enabled));
    buf.vm[p] = make_float4(v.x, v.y, v.z, mass);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        buf.C[(size_t)k * pcap + p] = aff[k];
        buf.F[(size_t)k * pcap + p] = F[k];
    }
    if (P.has_sand) {
#pragma unroll
        for (int k = 0; k < 3; k++) buf.dps[(size_t)k * pcap + p] = dps[k];
    }
    if (P.has_force) {
        fdt[p] = 0.0f;
        fdt[(size_t)pcap + p] = 0.0f;
        fdt[2 * (size_t)pcap + p] = 0.0f;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers (declarations in slosh_api.hip).

static inline uint32_t div_up_t(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }
static inline uint32_t min_t(uint32_t a, uint32_t b) { return a < b ? a : b; }

static inline PBuf to_pbuf(const ParticleBuffers &b) {
    PBuf p;
    p.pos = b.pos;
    p.vm = b.vm;
    p.C = b.C;
    p.F = b.F;
    p.vgd = b.vgd;
    p.ids = b.ids;
    p.dps = b.dps;
    return p;
}

// Workgroups of a per-block kernel: one block (per_wg blocks) per workgroup up to
// the tuned persistent width; workgroups loop over the device-side block count.
static uint32_t block_grid(const StepParams &P, uint32_t per_wg) {
    return min_t(div_up_t(P.cap, per_wg), P.block_wgs ? P.block_wgs : 8192u);
}

static Members members_of(const TransferBuffers &t) {
    Members M;
    M.fast = t.fast;
    M.blist = t.blist;
    M.nbr = t.nbr;
    M.bin = t.bin;
    M.mcount = t.mcount;
    M.mlist = t.mlist;
    M.ovf_slot = t.ovf_slot;
    M.ovf_hid = t.ovf_hid;
    M.perm = t.perm;
    M.blk_start = t.blk_start;
    M.blk_cnt = t.blk_cnt;
    M.counters = t.counters;
    return M;
}

hipError_t launch_p2g(const StepParams &P, const ParticleBuffers &pb, uint32_t pcap, const TransferBuffers &t,
                      const NextTableBuffers &next, const PassEvents *ev, hipStream_t stream) {
    NextTable nt;
    nt.keys = next.keys;
    nt.counters = next.counters;
    nt.prime = next.prime;
    auto kern = t.fast ? (P.p2g_chunk == 192u   ? k_p2g<192, true>
                          : P.p2g_chunk == 256u ? k_p2g<256, true>
                          : P.p2g_chunk == 384u ? k_p2g<384, true>
                                                : k_p2g<512, true>)
                       : (P.p2g_chunk == 192u   ? k_p2g<192, false>
                          : P.p2g_chunk == 256u ? k_p2g<256, false>
                          : P.p2g_chunk == 384u ? k_p2g<384, false>
                                                : k_p2g<512, false>);
    // Steady state: one more workgroup (the first) computes the block starts.
    launch_kernel(ev, kern, dim3(block_grid(P, 1) + (t.fast ? 1u : 0u)), dim3(192), 0, stream, P, nt,
                  (const float4 *)pb.pos, (const float4 *)pb.vm, (const float *)pb.C, (const float *)t.fdt, pcap,
                  members_of(t), t.keys, t.tiles, t.cell_cnt, t.blk_start, t.blk_cnt);
    return hipGetLastError();
}

hipError_t launch_grid_update(const StepParams &P, const TransferBuffers &t, const ColliderBuffers &c, uint32_t mode,
                              uint32_t do_update, const PassEvents *ev, hipStream_t stream) {
    Colliders cs;
    cs.shapes = c.shapes;
    cs.poses = c.poses;
    cs.mats = c.mats;
    cs.count = c.count;
    cs.mesh_vtx = c.mesh_vtx;
    cs.mesh_idx = c.mesh_idx;
    // Trimesh projection (a BVH walk) is compiled into its own instance so that
    // scenes without meshes keep the smaller kernel.
    launch_kernel(ev, c.mesh_vtx ? k_grid_update<true> : k_grid_update<false>, dim3(block_grid(P, 4)), dim3(256), 0,
                  stream, P,
                  (const float4 *)t.tiles, t.nodes, t.nbr, t.blist, t.keys, cs, (const NodeColl *)t.coll_prev,
                  (NodeColl *)t.coll_cur, t.counters, (const uint32_t *)t.blk_cnt, mode, do_update);
    return hipGetLastError();
}

hipError_t launch_g2p_update(const StepParams &P, const ParticleBuffers &in, const ParticleBuffers &out,
                             const StaticBuffers &s, uint32_t pcap, const TransferBuffers &t,
                             const NextTableBuffers &next, const BodyBuffers &bb, const PassEvents *ev,
                             hipStream_t stream) {
    NextTable nt;
    nt.keys = next.keys;
    nt.counters = next.counters;
    nt.prime = next.prime;
    NextSort ns;
    ns.mcount = next.mcount;
    ns.mv_key = next.mv_key;
    ns.mv_slot = next.mv_slot;
    ns.stay = t.stay;
    ns.errors = t.errors;
    Bodies b;
    b.poses = bb.poses;
    b.vels = bb.vels;
    b.mprops = bb.mprops;
    b.local_com = bb.local_com;
    b.impulses = bb.impulses;
    b.count = bb.count;
    PStatic st;
    st.prop = s.prop;
    st.model = s.model;
    st.dp0 = s.dp0;
    st.dp1 = s.dp1;
    st.scene_rec = s.scene_rec;
    // The lightest instance that handles every material present: neo-Hookean
    // only (no SVD code), fixed-corotated and neo-Hookean, every material.
    const bool nh_only = (P.tag_mask & ~2u) == 0u;
    const bool elastic_only = (P.tag_mask & ~3u) == 0u;
    const bool narrow = P.g2p_threads == 128u;
    auto pick = [&](auto fast) {
        constexpr bool F = decltype(fast)::value;
        return nh_only ? (narrow ? k_g2p_update<2u, 128, false, F> : k_g2p_update<2u, 256, false, F>)
               : elastic_only
                   ? (P.svd_f64 ? (narrow ? k_g2p_update<3u, 128, true, F> : k_g2p_update<3u, 256, true, F>)
                                : (narrow ? k_g2p_update<3u, 128, false, F> : k_g2p_update<3u, 256, false, F>))
                   : (P.svd_f64 ? (narrow ? k_g2p_update<15u, 128, true, F> : k_g2p_update<15u, 256, true, F>)
                                : (narrow ? k_g2p_update<15u, 128, false, F> : k_g2p_update<15u, 256, false, F>));
    };
    auto kern = t.fast ? pick(std::true_type{}) : pick(std::false_type{});
    launch_kernel(ev, kern, dim3(block_grid(P, 1)), dim3(narrow ? 128 : 256), 0, stream, P, to_pbuf(in),
                  to_pbuf(out), st, pcap, t.fdt, members_of(t), (const uint32_t *)t.blk_start,
                  (const float4 *)t.nodes, t.keys, nt, ns, b);
    return hipGetLastError();
}

hipError_t launch_g2p(const StepParams &P, const ParticleBuffers &b, uint32_t pcap, const TransferBuffers &t,
                      const PassEvents *ev, hipStream_t stream) {
    launch_kernel(ev, k_g2p, dim3(block_grid(P, 1)), dim3(256), 0, stream, P, to_pbuf(b), pcap, t.perm, t.blist,
                  (const uint32_t *)t.blk_start, (const uint32_t *)t.blk_cnt, t.nbr, (const float4 *)t.nodes,
                  t.counters);
    return hipGetLastError();
}

hipError_t launch_particle_update(const StepParams &P, const ParticleBuffers &b, const StaticBuffers &s, uint32_t pcap,
                                  float *fdt, const PassEvents *ev, hipStream_t stream) {
    if (P.n == 0) return hipSuccess;
    PStatic st;
    st.prop = s.prop;
    st.model = s.model;
    st.dp0 = s.dp0;
    st.dp1 = s.dp1;
    st.scene_rec = s.scene_rec;
    auto kern = P.tag_mask == 2u ? k_particle_update<2u, false>
                                 : (P.svd_f64 ? k_particle_update<15u, true> : k_particle_update<15u, false>);
    launch_kernel(ev, kern, dim3(div_up_t(P.n, 256)), dim3(256), 0, stream, P, to_pbuf(b), st, pcap, fdt);
    return hipGetLastError();
}

// For each run of equal scene ids among slots [slot0, slot0 + n): the by-id
// records of the run's first particle into its scene's table entry.
__global__ __launch_bounds__(256) void k_scene_records(const float4 *__restrict__ pos, const uint32_t *__restrict__ ids,
                                                       uint32_t slot0, uint32_t n, PStatic st,
                                                       float4 *__restrict__ scene_rec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t scene = __float_as_uint(pos[slot0 + i].w) >> 1;
    if (i > 0 && (__float_as_uint(pos[slot0 + i - 1].w) >> 1) == scene) return;
    const uint32_t oid = ids[slot0 + i];
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    scene_rec[4 * (size_t)scene] = st.prop[oid];
    scene_rec[4 * (size_t)scene + 1] = st.model[oid];
    scene_rec[4 * (size_t)scene + 2] = st.dp0 ? st.dp0[oid] : zero;
    scene_rec[4 * (size_t)scene + 3] = st.dp1 ? st.dp1[oid] : zero;
}

// Then each scene's mass, the largest particle mass of its uploaded particles
// (the enabled ones share it, the disabled ones have +0), into the unused third
// component of its plastic record (zeroed by k_scene_records). As unsigned bits,
// non-negative floats order as their values. An upload holds each scene's
// particles in one run, so a wave almost always covers one scene: it reduces
// first and sends one atomic (per-particle atomics on one address per scene
// serialize in L2: 11.4M of them on 256 records took ~10 ms).
__global__ __launch_bounds__(256) void k_scene_mass(const float4 *__restrict__ pos, const float4 *__restrict__ vm,
                                                    uint32_t slot0, uint32_t n, float4 *__restrict__ scene_rec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const uint32_t scene = valid ? __float_as_uint(pos[slot0 + i].w) >> 1 : SLOSH_NONE;
    const uint32_t m = valid ? __float_as_uint(vm[slot0 + i].w) : 0u;
    // Lane 0 holds the wave's first index, so it is valid whenever any lane is.
    const uint32_t s0 = (uint32_t)__shfl((int)scene, 0);
    if (__ballot(valid && scene != s0) == 0ull) {
        uint32_t x = m;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x = max(x, (uint32_t)__shfl_xor((int)x, o));
        if (lane_id() == 0 && s0 != SLOSH_NONE) atomicMax((uint32_t *)&scene_rec[4 * (size_t)s0 + 3].z, x);
    } else if (valid) {
        atomicMax((uint32_t *)&scene_rec[4 * (size_t)scene + 3].z, m);
    }
}

hipError_t launch_scene_records(const float4 *pos, const float4 *vm, const uint32_t *ids, uint32_t slot0, uint32_t n,
                                const StaticBuffers &s, float4 *scene_rec, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    PStatic st;
    st.prop = s.prop;
    st.model = s.model;
    st.dp0 = s.dp0;
    st.dp1 = s.dp1;
    st.scene_rec = nullptr;
    hipLaunchKernelGGL(k_scene_records, dim3(div_up_t(n, 256)), dim3(256), 0, stream, pos, ids, slot0, n, st,
                       scene_rec);
    hipLaunchKernelGGL(k_scene_mass, dim3(div_up_t(n, 256)), dim3(256), 0, stream, pos, vm, slot0, n, scene_rec);
    return hipGetLastError();
}

hipError_t launch_timestep_bound(const StepParams &P, const ParticleBuffers &b, const StaticBuffers &s, uint32_t pcap,
                                 uint32_t *out, hipStream_t stream) {
    if (P.n == 0) return hipSuccess;
    PStatic st;
    st.prop = s.prop;
    st.model = s.model;
    st.dp0 = s.dp0;
    st.dp1 = s.dp1;
    st.scene_rec = s.scene_rec;
    hipLaunchKernelGGL(k_timestep_bound, dim3(min_t(div_up_t(P.n, 256), 4096)), dim3(256), 0, stream, P, to_pbuf(b),
                       st, pcap, out);
    return hipGetLastError();
}

hipError_t launch_timestep_bound_partials(const StepParams &P, const ParticleBuffers &b, const StaticBuffers &s,
                                          uint32_t pcap, uint32_t *partials, uint32_t *out, hipStream_t stream) {
    PStatic st;
    st.prop = s.prop;
    st.model = s.model;
    st.dp0 = s.dp0;
    st.dp1 = s.dp1;
    st.scene_rec = s.scene_rec;
    const uint32_t groups = div_up_t(P.n, TSB_CHUNK);
    if (groups > 0)
        hipLaunchKernelGGL(k_timestep_bound_partials, dim3(groups), dim3(256), 0, stream, P, to_pbuf(b), st, pcap,
                           partials);
    hipLaunchKernelGGL(k_min_partials, dim3(1), dim3(256), 0, stream, partials, groups, out);
    return hipGetLastError();
}

hipError_t launch_rigid_update(const StepParams &P, const BodyBuffers &bb, const PassEvents *ev, hipStream_t stream) {
    if (bb.count == 0) return hipSuccess;
    Bodies b;
    b.poses = bb.poses;
    b.vels = bb.vels;
    b.mprops = bb.mprops;
    b.local_com = bb.local_com;
    b.impulses = bb.impulses;
    b.count = bb.count;
    launch_kernel(ev, k_rigid_update, dim3(1), dim3(64), 0, stream, P, b);
    return hipGetLastError();
}
