// mdx_device.h -- device-side math of the Langevin step (gfx950).
//
// Reference citations are relative to the reference's Code/ folder.  The FP64
// forms keep the reference's operation order where it matters for rounding
// (min-image rule, cut-off tests, clamp constants).  The FP32 forms
// reformulate the patch angles with sin = |n x u| and theta = atan2(sin, cos)
// (no 1 - a^2 cancellation) and take every discontinuous decision (cut-offs,
// cone edges, the d1 = 0 zone) in FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdx {

constexpr int MAXT = 8;    // particle types
constexpr int MAXP = 6;    // patches per type
constexpr int MAXEXCL = 8; // exclusions per particle

// ----------------------------------------------------------------- Philox
// Philox4x32-10, Salmon, Moraes, Dror, Shaw (SC'11).
struct U4 {
    uint32_t a, b, c, d;
};

__device__ __forceinline__ U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t lo0 = 0xD2511F53u * ctr.a;
        uint32_t hi0 = __umulhi(0xD2511F53u, ctr.a);
        uint32_t lo1 = 0xCD9E8D57u * ctr.c;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, ctr.c);
        U4 n;
        n.a = hi1 ^ ctr.b ^ k0;
        n.b = lo1;
        n.c = hi0 ^ ctr.d ^ k1;
        n.d = lo0;
        ctr = n;
    }
    return ctr;
}

// uniform on [-sqrt 3, sqrt 3] with unit variance, the transform of
// generate_uniform_random_matrix (DataStructures/matrix2.cpp:1807-1821)
template <class R> __device__ __forceinline__ R unit_uniform(uint32_t u)
{
    R f = ((R)u + (R)0.5) * (R)2.3283064365386963e-10;
    return (R)3.464101615 * f - (R)1.732050808;
}

// ------------------------------------------------------------------ geometry
// cube::distance_vector min-image rule (geometry.h:277-294): shift when
// as^2 > (l/2)^2 by SIGN(l, as).
__device__ __forceinline__ double mimage(double as, double l, double l2s, int periodic)
{
    if (periodic && as * as > l2s)
        as = as - (as >= 0 ? l : -l);
    return as;
}

// Basic/basic.h:58
__device__ __forceinline__ int fasterfloor(double x) { return x > 0 ? (int)x : (int)x - 1; }

// cube::assign_box per axis (geometry.h:513-516)
__device__ __forceinline__ int axis_cell(double x, int n, double l)
{
    int m = fasterfloor((double)n * x / l);
    if (m < 0)
        m = 0;
    else if (m > n - 1)
        m = n - 1;
    return m;
}

// cube::correct_position_and_momentum (geometry.h:403-445) for one
// coordinate: the reference's loop of single box-length shifts (periodic) or
// one reflection (walls), so a coordinate one step outside the box rounds as
// the reference's does.  Bounded: after WRAP_MAX shifts (a coordinate that
// far out cannot come from one step with a finite force) the remainder is
// taken by floor, and a result outside [0, l] (a reflection from more than a
// box length out) is clamped.  Returns false, leaving x and p untouched, for
// a non-finite x or p; the caller then keeps the particle's previous state.
constexpr int WRAP_MAX = 64;
template <class R>
__host__ __device__ __forceinline__ bool wrap1(double &x, R &p, double l, int periodic)
{
    if (!(__builtin_isfinite(x) && __builtin_isfinite((double)p)))
        return false;
    if (periodic) {
        int it = 0;
        if (x < 0) {
            while (x < 0 && it++ < WRAP_MAX)
                x += l;
        } else if (x > l) {
            while (x > l && it++ < WRAP_MAX)
                x -= l;
        }
        if (x < 0 || x > l) {
            x -= l * __builtin_floor(x / l);
            x = x < 0 ? 0.0 : (x > l ? l : x);
        }
    } else {
        if (x < 0) {
            x = -x;
            p = -p;
        } else if (x > l) {
            x = l - (x - l);
            p = -p;
        }
        x = x < 0 ? 0.0 : (x > l ? l : x);
    }
    return true;
}

// ------------------------------------------------------------- orientations
// The mixed-precision engine stores each orientation as a unit quaternion
// (w, x, y, z) in 16 B instead of the 36 B of an FP32 3x3 matrix, and
// rebuilds the lab -> body matrix on load:
//   m = [[1-2(y^2+z^2), 2(xy-wz),     2(xz+wy)    ],
//        [2(xy+wz),     1-2(x^2+z^2), 2(yz-wx)    ],
//        [2(xz-wy),     2(yz+wx),     1-2(x^2+y^2)]]
// A rotated matrix goes back to a quaternion by Shepperd's method (the
// branch on the largest of the trace and the diagonal keeps the square root
// well conditioned) and is normalised, which also keeps the orientation
// orthonormal.  R(q) = R(-q) bit for bit, so the sign the conversion picks
// does not matter to the dynamics.
template <class R> __host__ __device__ __forceinline__ void quat_to_mat(R w, R x, R y, R z, R m[9])
{
    const R xx = x * x, yy = y * y, zz = z * z;
    const R xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y, wz = w * z;
    m[0] = (R)1 - (R)2 * (yy + zz);
    m[1] = (R)2 * (xy - wz);
    m[2] = (R)2 * (xz + wy);
    m[3] = (R)2 * (xy + wz);
    m[4] = (R)1 - (R)2 * (xx + zz);
    m[5] = (R)2 * (yz - wx);
    m[6] = (R)2 * (xz - wy);
    m[7] = (R)2 * (yz + wx);
    m[8] = (R)1 - (R)2 * (xx + yy);
}

__host__ __device__ __forceinline__ float sqrt_(float v) { return __builtin_sqrtf(v); }
__host__ __device__ __forceinline__ double sqrt_(double v) { return __builtin_sqrt(v); }

// a / S: FP64 divides (the conversion of uploaded FP64 matrices, bitwise
// as the host's); FP32 multiplies by the one reciprocal iS = 1 / S
template <class R> __host__ __device__ __forceinline__ R qdiv(R a, R S, R iS)
{
    if constexpr (sizeof(R) == 8)
        return a / S;
    else
        return a * iS;
}

template <class R> __host__ __device__ __forceinline__ void mat_to_quat(const R m[9], R q[4])
{
    const R tr = m[0] + m[4] + m[8];
    R w, x, y, z;
    if (tr > (R)0) {
        const R S = sqrt_(tr + (R)1) * (R)2, iS = (R)1 / S;
        w = (R)0.25 * S;
        x = qdiv(m[7] - m[5], S, iS);
        y = qdiv(m[2] - m[6], S, iS);
        z = qdiv(m[3] - m[1], S, iS);
    } else if (m[0] > m[4] && m[0] > m[8]) {
        const R S = sqrt_((R)1 + m[0] - m[4] - m[8]) * (R)2, iS = (R)1 / S;
        w = qdiv(m[7] - m[5], S, iS);
        x = (R)0.25 * S;
        y = qdiv(m[1] + m[3], S, iS);
        z = qdiv(m[2] + m[6], S, iS);
    } else if (m[4] > m[8]) {
        const R S = sqrt_((R)1 + m[4] - m[0] - m[8]) * (R)2, iS = (R)1 / S;
        w = qdiv(m[2] - m[6], S, iS);
        x = qdiv(m[1] + m[3], S, iS);
        y = (R)0.25 * S;
        z = qdiv(m[5] + m[7], S, iS);
    } else {
        const R S = sqrt_((R)1 + m[8] - m[0] - m[4]) * (R)2, iS = (R)1 / S;
        w = qdiv(m[3] - m[1], S, iS);
        x = qdiv(m[2] + m[6], S, iS);
        y = qdiv(m[5] + m[7], S, iS);
        z = (R)0.25 * S;
    }
    const R inv = (R)1 / sqrt_(w * w + x * x + y * y + z * z);
    q[0] = w * inv;
    q[1] = x * inv;
    q[2] = y * inv;
    q[3] = z * inv;
}

// ------------------------------------------------------------------- pairs
// HSPotential::force (potential.h:112-125)
__device__ __forceinline__ double hs_force(double x, double epsilon, double sigma6,
                                           double sigma12)
{
    double a = x * x * x;
    a = a * a;
    return 4 * epsilon * (12 * sigma12 / (x * (a * a)) - 6 * sigma6 / (x * a));
}
__device__ __forceinline__ float hs_force(float x, float epsilon, float sigma6, float sigma12)
{
    float a = x * x * x;
    a = a * a;
    return 4.f * epsilon * (12.f * sigma12 / (x * (a * a)) - 6.f * sigma6 / (x * a));
}
// WCAPotential::energy r < paramlimit branch with attrepsilon = 0 (potential.h:179-185)
template <class R> __device__ __forceinline__ R wca_energy(R x, R epsilon, R sigma6, R sigma12)
{
    R a = x * x * x;
    a = a * a;
    return (R)4 * epsilon * (sigma12 / (a * a) - sigma6 / a) + epsilon;
}

// derived per-row constants of one KernFrenkelOnePatch2
struct PatchRow {
    double eps, d, thetam, costm, kk, idist; // kk = pi / (2 theta_m), idist = 1.4 d
};

// ---------------------------------------------------- Kern-Frenkel, FP64 form
// KernFrenkelOnePatch2::force_and_torque (potentialtheta.h:517-750) for one
// patch pair whose lab normals n1 (particle "1") and n2 are given, u unit
// from particle 1 to ... (u = (x1 - x2)/r), a1 = -n1.u, a2 = n2.u already
// known to pass the cone test.  Adds the force on 1 to f, the torques to t1,
// t2; returns the energy.
__device__ __forceinline__ double kf2_pair_fp64(const double u[3], double rij, const double n1[3],
                                                const double n2[3], double argthetai,
                                                double argthetaj, const PatchRow &pr,
                                                double f[3], double t1[3], double t2[3])
{
    const double PI = 3.14159265358979323846;
    double thetam = pr.thetam;
    if (argthetai > 1.)
        argthetai = 0.999999;
    if (argthetaj > 1.)
        argthetaj = 0.999999;
    double thetai = acos(argthetai);
    double thetaj = acos(argthetaj);
    double ci = cos(PI * thetai / (2. * thetam));
    double cj = cos(PI * thetaj / (2. * thetam));
    double si = sin(PI * thetai / (2. * thetam));
    double sj = sin(PI * thetaj / (2. * thetam));
    double fang = ci * cj;
    double fac = (rij / pr.d);
    double fac2 = fac * fac;
    double fac4 = fac2 * fac2;
    double fac8 = fac4 * fac4;
    double fac10 = fac2 * fac8;
    double expf_ = exp(-0.5 * fac10);
    double pot = -pr.eps * expf_;
    double potf = -5. * pr.eps * fac8 * (fac / pr.d) * expf_;
    double fx = fang * potf * u[0];
    double fy = fang * potf * u[1];
    double fz = fang * potf * u[2];
    double d1 = fabs(argthetai * argthetai - 1) < 1E-5 ? 0.0 : -1. / sqrt(1 - argthetai * argthetai);
    double d2 = fabs(argthetaj * argthetaj - 1) < 1E-5 ? 0.0 : -1. / sqrt(1 - argthetaj * argthetaj);
    double k = PI / (2 * thetam);
    double temp1 = -pot * (k * si * d1 * cj);
    double temp2 = -pot * (k * sj * d2 * ci);
    fx += (temp1 / rij) * (n1[0] + argthetai * u[0]);
    fy += (temp1 / rij) * (n1[1] + argthetai * u[1]);
    fz += (temp1 / rij) * (n1[2] + argthetai * u[2]);
    fx += -(temp2 / rij) * (n2[0] - argthetaj * u[0]);
    fy += -(temp2 / rij) * (n2[1] - argthetaj * u[1]);
    fz += -(temp2 / rij) * (n2[2] - argthetaj * u[2]);
    double gi = -k * pot * si * d1 * cj;
    double gj = -k * pot * sj * d2 * ci;
    f[0] += fx;
    f[1] += fy;
    f[2] += fz;
    t1[0] += gi * (u[2] * n1[1] - u[1] * n1[2]);
    t1[1] += gi * (-u[2] * n1[0] + u[0] * n1[2]);
    t1[2] += gi * (u[1] * n1[0] - u[0] * n1[1]);
    t2[0] += gj * (-u[2] * n2[1] + u[1] * n2[2]);
    t2[1] += gj * (u[2] * n2[0] - u[0] * n2[2]);
    t2[2] += gj * (-u[1] * n2[0] + u[0] * n2[1]);
    return pot * fang;
}

// ---------------------------------------------------- Kern-Frenkel, FP32 form
// Same potential with theta = atan2(|n x u|, a) and d1 = -1/|n x u|; the
// zones of the reference (d1 = 0 when |a^2 - 1| < 1e-5, clamp of a > 1) are
// decided by the caller in FP64 and passed as zero1/zero2.
__device__ __forceinline__ float kf2_pair_fp32(const float u[3], float rij, const float n1[3],
                                               const float n2[3], float a1, float a2, bool zero1,
                                               bool zero2, float eps, float d, float kk,
                                               float f[3], float t1[3], float t2[3])
{
#pragma clang fp contract(fast)
    // sin(theta) from the cross products (full relative precision near 0)
    float c1x = n1[1] * u[2] - n1[2] * u[1], c1y = n1[2] * u[0] - n1[0] * u[2],
          c1z = n1[0] * u[1] - n1[1] * u[0];
    float c2x = n2[1] * u[2] - n2[2] * u[1], c2y = n2[2] * u[0] - n2[0] * u[2],
          c2z = n2[0] * u[1] - n2[1] * u[0];
    float s1 = sqrtf(c1x * c1x + c1y * c1y + c1z * c1z);
    float s2 = sqrtf(c2x * c2x + c2y * c2y + c2z * c2z);
    float th1 = atan2f(s1, a1);
    float th2 = atan2f(s2, a2);
    float si, ci, sj, cj;
    sincosf(kk * th1, &si, &ci);
    sincosf(kk * th2, &sj, &cj);
    float fang = ci * cj;
    float fac = rij / d;
    float fac2 = fac * fac;
    float fac4 = fac2 * fac2;
    float fac8 = fac4 * fac4;
    float fac10 = fac2 * fac8;
    float ex = expf(-0.5f * fac10);
    float pot = -eps * ex;
    float potf = -5.f * eps * fac8 * (fac / d) * ex;
    float rd1 = zero1 ? 0.f : -1.f / s1;
    float rd2 = zero2 ? 0.f : -1.f / s2;
    float temp1 = -pot * (kk * si * rd1 * cj);
    float temp2 = -pot * (kk * sj * rd2 * ci);
    float ir = 1.f / rij;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        f[k] += fang * potf * u[k] + temp1 * ir * (n1[k] + a1 * u[k]) -
                temp2 * ir * (n2[k] - a2 * u[k]);
    }
    // torque_i = g_i (n1 x u), torque_j = g_j (u x n2)
    float gi = -kk * pot * si * rd1 * cj;
    float gj = -kk * pot * sj * rd2 * ci;
    t1[0] += gi * c1x;
    t1[1] += gi * c1y;
    t1[2] += gi * c1z;
    t2[0] -= gj * c2x;
    t2[1] -= gj * c2y;
    t2[2] -= gj * c2z;
    return pot * fang;
}

// ----------------------------------------------------------------- bending
// BendingPotential::force (potential3.h:43-114): force on member `role`
// (0,1,2) of the triplet with ab = x2 - x1, bc = x3 - x2.
template <class R>
__device__ __forceinline__ void bend_harmonic(const R ab[3], const R bc[3], R k0, R theta0,
                                              int role, R out[3])
{
    R a = ab[0] * bc[0] + ab[1] * bc[1] + ab[2] * bc[2];
    R mabsq = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
    R mbcsq = bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2];
    R mab = sqrt(mabsq), mbc = sqrt(mbcsq);
    R temp1 = a / (mab * mbc);
    if (temp1 > (R)1.)
        temp1 = (R)0.999999999;
    else if (temp1 < (R)-1)
        temp1 = (R)-0.9999999999;
    R theta = acos(temp1);
    R fac1 = sqrt((R)1. - temp1 * temp1);
    R fa = -k0 * (theta - theta0);
    out[0] = out[1] = out[2] = 0;
    if (fabs(theta - theta0) < (R)1E-10)
        return;
    for (int j = 0; j < 3; j++) {
        if (role == 0)
            out[j] = fa * (bc[j] * mabsq - ab[j] * a) / (mabsq * fac1 * mab * mbc);
        else if (role == 1)
            out[j] = fa * ((ab[j] - bc[j]) * mabsq * mbcsq + a * (-(bc[j] * mabsq) + ab[j] * mbcsq)) /
                     (fac1 * mab * mbc * mabsq * mbcsq);
        else
            out[j] = fa * (bc[j] * a - ab[j] * mbcsq) / (mbcsq * fac1 * mab * mbc);
    }
}

// cosine bending U = k (1 - ab.bc / (|ab||bc|))
template <class R>
__device__ __forceinline__ void bend_cosine(const R ab[3], const R bc[3], R k0, int role, R out[3])
{
    R a = ab[0] * bc[0] + ab[1] * bc[1] + ab[2] * bc[2];
    R mabsq = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
    R mbcsq = bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2];
    R inv = (R)1 / sqrt(mabsq * mbcsq);
    R c = a * inv;
    for (int j = 0; j < 3; j++) {
        R f1 = -k0 * (bc[j] * inv - c * ab[j] / mabsq);
        R f3 = k0 * (ab[j] * inv - c * bc[j] / mbcsq);
        out[j] = role == 0 ? f1 : (role == 2 ? f3 : -(f1 + f3));
    }
}

} // namespace mdx
