// mdx_engine.hip -- MI355X (gfx950) engine behind include/mdx.h.
//
// Per step (rotor integrator, Condensate.cpp:352-377 of the reference):
//   kick/drift/rotate/bin  -> exclusive scan of cell counts -> fill cell lists
//   -> order each cell's list -> gather forces + sphere + half kick
// The gather kernel computes every particle's force and torque by looping
// over its 27 neighbour cells (full neighbour list, no Newton-3 scatter, no
// atomics), so results are deterministic.  Inside one mdx_step call the
// gather kernel of step s also performs the first half of step s+1 (second
// half kick, next half kick, drift, rotation, binning) so that the per-step
// state crosses HBM once.  Positions and orientations are double buffered
// because neighbours read them while the owner writes the next ones.
//
// Reference paths in comments are relative to the reference's Code/ folder.
#include "mdx_device.h"
#include "../../include/mdx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace mdx;

// ---------------------------------------------------------------- errors
static thread_local std::string g_err;

static int set_err(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(call)                                                                           \
    do {                                                                                       \
        hipError_t e__ = (call);                                                               \
        if (e__ != hipSuccess)                                                                 \
            return set_err(MDX_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e__), \
                           __FILE__, __LINE__);                                                \
    } while (0)

// ------------------------------------------------------------ device structs
struct DevRep {
    double l[3], l2s[3];
    int nc[3];
    int coff;
    int first, count;
    double Rt, Rr;      // sqrt(2 gamma kT / dt), sqrt(2 gamma_r kT / dt) (LangevinR.cpp:51-52)
    double bead_r;      // sqrt(0.5 kT gamma m dt) (Langevin.cpp:44)
    int prow;           // first patch row of this replica
};

struct StepCtl {
    unsigned long long epoch;     // force evaluations so far (noise counter)
    unsigned long long call_base; // epoch at the start of the current call
    unsigned int status;          // bit0 force guard, bit1 non-finite
    unsigned int pad;
};

enum { M_CONS = 0, M_PRIME = 1, M_STEP = 2, M_FUSED = 3, M_BEAD = 4, M_ENERGY = 5 };

template <class R> struct KArgs {
    int N;
    int rotor;
    int periodic;
    const uint32_t *meta; // replica << 8 | type
    const DevRep *reps;
    const double *X;      // current positions N x 3
    double *Xn;           // next positions
    const R *Q;           // current orientations N x 9
    R *Qn;
    R *p, *L, *F, *T;     // momenta, body angular momenta, carried total force/torque
    double *Fout, *Tout;  // conservative output
    double *obs;          // per particle energy + virial (10)
    int *cell, *rank, *count;
    const int *start, *list;
    // WCA (HSPotential)
    int wca;
    double wca_eps, wca_s6, wca_s12, wca_pl;
    // patches
    int has_patches;
    int np[MAXT];
    int pstart[MAXT];
    int rstart[MAXT * MAXT];
    unsigned long long active_mask; // bit ta * MAXT + tc: type pair with a nonzero row
    const double *pdirs;
    const PatchRow *prows;
    double maxd2;
    double rmax2;
    // topology
    const int *bond_off, *bond_j;
    const double *bond_k, *bond_r0;
    int bond_kind;
    const int *ang_off, *ang_ref; // per particle: angle index * 4 + role
    const int *ang_ijk;
    const double *ang_k, *ang_t0;
    int ang_form;
    const int *exc_off, *exc_j;
    const int *act_prev, *act_next;
    double fa;
    // integrator
    double dt, hdt, c1, m, gamma, gammar, I[3];
    double b_c5c2, b_c5c3q, b_c5c4, b_c2; // bead coefficients without the kT part
    int reortho_every;
    // noise
    uint32_t key0, key1;
    const double *noise; // injected rows (ncomp per particle) or NULL
    StepCtl *ctl;
    double fguard;
};

// ------------------------------------------------------------ small helpers
template <class R> __device__ __forceinline__ void lab_normal(const R q[9], const double *nb, R n[3])
{
    // n = Q^T n_body (potentialtheta.h:559-565)
    n[0] = (R)nb[0] * q[0] + (R)nb[1] * q[3] + (R)nb[2] * q[6];
    n[1] = (R)nb[0] * q[1] + (R)nb[1] * q[4] + (R)nb[2] * q[7];
    n[2] = (R)nb[0] * q[2] + (R)nb[1] * q[5] + (R)nb[2] * q[8];
}

__device__ __forceinline__ void lab_normal_d(const float q[9], const double *nb, double n[3])
{
    n[0] = nb[0] * (double)q[0] + nb[1] * (double)q[3] + nb[2] * (double)q[6];
    n[1] = nb[0] * (double)q[1] + nb[1] * (double)q[4] + nb[2] * (double)q[7];
    n[2] = nb[0] * (double)q[2] + nb[1] * (double)q[5] + nb[2] * (double)q[8];
}

// LangevinNVTR::genfullmat (LangevinR.cpp:392-461)
template <class R>
__device__ __forceinline__ void genfullmat(R dt, const R I[3], const R L[3], R up[9])
{
    R tempx = dt * L[0] / I[0];
    R tempy = dt * L[1] / I[1];
    R tempz = dt * L[2] / I[2];
    R qx = tempx / (R)4., qy = tempy / (R)4., hz = tempz / (R)2.;
    R fac1 = ((R)1 - qx * qx) / ((R)1 + qx * qx);
    R fac2 = tempx / ((R)2. * ((R)1 + qx * qx));
    R fac3 = ((R)1 - qy * qy) / ((R)1 + qy * qy);
    R fac4 = tempy / ((R)2. * ((R)1 + qy * qy));
    R fac5 = ((R)1 - hz * hz) / ((R)1 + hz * hz);
    R fac6 = (tempz) / ((R)1 + hz * hz);
    up[0] = -fac4 * fac4 + fac3 * fac3 * fac5;
    up[1] = fac3 * (fac2 * fac4 * ((R)1 + fac5) - fac1 * fac6);
    up[2] = fac3 * (fac1 * fac4 * ((R)1 + fac5) + fac2 * fac6);
    up[3] = fac3 * (fac2 * fac4 * ((R)1 + fac5) + fac1 * fac6);
    up[4] = fac1 * fac1 * fac5 + fac2 * fac2 * (-fac3 * fac3 + fac4 * fac4 * fac5);
    up[5] = -(fac1 * fac2 * (fac3 * fac3 + fac5 - fac4 * fac4 * fac5)) +
            (fac1 * fac1 + fac2 * fac2) * fac4 * fac6;
    up[6] = -(fac1 * fac3 * fac4 * ((R)1 + fac5)) + fac2 * fac3 * fac6;
    up[7] = fac1 * fac2 * (fac3 * fac3 + fac5 - fac4 * fac4 * fac5) +
            (fac1 * fac1 + fac2 * fac2) * fac4 * fac6;
    up[8] = -(fac2 * fac2 * fac5) + fac1 * fac1 * (fac3 * fac3 - fac4 * fac4 * fac5);
}

// LangevinNVTR::rotate (LangevinR.cpp:595-685): L <- R L, Q <- R^T Q
template <class R> __device__ __forceinline__ void rotate(R dt, const R I[3], R L[3], R q[9])
{
    R rr[9];
    genfullmat(dt, I, L, rr);
    R jx = L[0], jy = L[1], jz = L[2];
    L[0] = jx * rr[0] + jy * rr[1] + jz * rr[2];
    L[1] = jx * rr[3] + jy * rr[4] + jz * rr[5];
    L[2] = jx * rr[6] + jy * rr[7] + jz * rr[8];
    R o[9];
#pragma unroll
    for (int k = 0; k < 9; k++)
        o[k] = q[k];
    q[0] = o[0] * rr[0] + o[3] * rr[3] + o[6] * rr[6];
    q[1] = o[1] * rr[0] + o[4] * rr[3] + o[7] * rr[6];
    q[2] = o[2] * rr[0] + o[5] * rr[3] + o[8] * rr[6];
    q[3] = o[0] * rr[1] + o[3] * rr[4] + o[6] * rr[7];
    q[4] = o[1] * rr[1] + o[4] * rr[4] + o[7] * rr[7];
    q[5] = o[2] * rr[1] + o[5] * rr[4] + o[8] * rr[7];
    q[6] = o[0] * rr[2] + o[3] * rr[5] + o[6] * rr[8];
    q[7] = o[1] * rr[2] + o[4] * rr[5] + o[7] * rr[8];
    q[8] = o[2] * rr[2] + o[5] * rr[5] + o[8] * rr[8];
}

// symmetric re-orthonormalisation in FP64: Q <- 1.5 Q - 0.5 Q Q^T Q (one
// Newton-Schulz step of Q (Q^T Q)^(-1/2); deviations are ~1e-6 in FP32)
template <class R> __device__ __forceinline__ void reortho(R q[9])
{
    double a[9], m[9];
    for (int k = 0; k < 9; k++)
        a[k] = (double)q[k];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
            m[3 * r + c] = a[3 * r] * a[3 * c] + a[3 * r + 1] * a[3 * c + 1] + a[3 * r + 2] * a[3 * c + 2];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double s = m[3 * r] * a[c] + m[3 * r + 1] * a[3 + c] + m[3 * r + 2] * a[6 + c];
            q[3 * r + c] = (R)(1.5 * a[3 * r + c] - 0.5 * s);
        }
}

// cube::correct_position_and_momentum (geometry.h:403-445) for one coordinate
template <class R>
__device__ __forceinline__ void wrap1(double &x, R &p, double l, int periodic)
{
    if (periodic) {
        if (x < 0) {
            while (x < 0)
                x += l;
        } else if (x > l) {
            while (x > l)
                x -= l;
        }
    } else {
        if (x < 0) {
            x = -x;
            p = -p;
        } else if (x > l) {
            x = l - (x - l);
            p = -p;
        }
    }
}

template <class R> __device__ __forceinline__ void flag_force(const KArgs<R> &a, double f)
{
    if (a.fguard > 0 && fabs(f) > a.fguard)
        atomicOr(&a.ctl->status, 1u);
}

template <class R>
__device__ __forceinline__ void load_noise(const KArgs<R> &a, int i, int ncomp, R out[6])
{
    unsigned long long ep = a.ctl->epoch;
    if (a.noise) {
        const double *row = a.noise + (size_t)(ep - a.ctl->call_base) * (size_t)a.N * ncomp +
                            (size_t)i * ncomp;
        for (int k = 0; k < ncomp; k++)
            out[k] = (R)row[k];
    } else {
        U4 c0 = {(uint32_t)i, (uint32_t)ep, (uint32_t)(ep >> 32), 0u};
        U4 r0 = philox4x32_10(c0, a.key0, a.key1);
        out[0] = unit_uniform<R>(r0.a);
        out[1] = unit_uniform<R>(r0.b);
        out[2] = unit_uniform<R>(r0.c);
        if (ncomp > 3) {
            out[3] = unit_uniform<R>(r0.d);
            U4 c1 = {(uint32_t)i, (uint32_t)ep, (uint32_t)(ep >> 32), 1u};
            U4 r1 = philox4x32_10(c1, a.key0, a.key1);
            out[4] = unit_uniform<R>(r1.a);
            out[5] = unit_uniform<R>(r1.b);
        }
    }
}

template <class R>
__device__ __forceinline__ void bin_particle(const KArgs<R> &a, int i, const DevRep &rp,
                                             const double x[3])
{
    int cx = axis_cell(x[0], rp.nc[0], rp.l[0]);
    int cy = axis_cell(x[1], rp.nc[1], rp.l[1]);
    int cz = axis_cell(x[2], rp.nc[2], rp.l[2]);
    int c = rp.coff + cx + rp.nc[0] * (cy + rp.nc[1] * cz);
    a.cell[i] = c;
    a.rank[i] = atomicAdd(&a.count[c], 1);
}

// ------------------------------------------------------------- gather kernel
// One thread per particle.  Accumulates the conservative force (lab) and
// patch torque (lab) of particle i over its neighbour cells and topology,
// then applies the mode's integrator epilogue.
template <class R, int MODE>
__global__ void __launch_bounds__(256) k_gather(KArgs<R> a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N)
        return;
    const uint32_t mi = a.meta[i];
    const int ti = mi & 255;
    const DevRep rp = a.reps[mi >> 8];
    const double xi0 = a.X[3 * i], xi1 = a.X[3 * i + 1], xi2 = a.X[3 * i + 2];
    R Fi[3] = {0, 0, 0}, Ti[3] = {0, 0, 0};
    double epot = 0, vir[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    constexpr bool FP64 = sizeof(R) == 8;

    // own orientation and patch normals
    R qi[9];
    R ni[MAXP][3];
    const int npi = a.has_patches ? a.np[ti] : 0;
    if (a.rotor) {
#pragma unroll
        for (int k = 0; k < 9; k++)
            qi[k] = a.Q[9 * (size_t)i + k];
        for (int ia = 0; ia < npi; ia++)
            lab_normal(qi, a.pdirs + 3 * (a.pstart[ti] + ia), ni[ia]);
    }
    // own exclusions
    int exc[MAXEXCL];
    int nexc = 0;
    if (a.exc_off) {
        int e0 = a.exc_off[i], e1 = a.exc_off[i + 1];
        nexc = e1 - e0;
        for (int k = 0; k < nexc && k < MAXEXCL; k++)
            exc[k] = a.exc_j[e0 + k];
    }

    if (a.wca || npi > 0) {
        const int c = a.cell[i] - rp.coff;
        const int cx = c % rp.nc[0], cy = (c / rp.nc[0]) % rp.nc[1], cz = c / (rp.nc[0] * rp.nc[1]);
        int lo[3], hi[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            lo[k] = rp.nc[k] >= 3 ? -1 : 0;
            hi[k] = rp.nc[k] >= 2 ? 1 : 0;
        }
        for (int oz = lo[2]; oz <= hi[2]; oz++) {
            int z = cz + oz;
            z = z < 0 ? z + rp.nc[2] : (z >= rp.nc[2] ? z - rp.nc[2] : z);
            for (int oy = lo[1]; oy <= hi[1]; oy++) {
                int y = cy + oy;
                y = y < 0 ? y + rp.nc[1] : (y >= rp.nc[1] ? y - rp.nc[1] : y);
                for (int ox = lo[0]; ox <= hi[0]; ox++) {
                    int x = cx + ox;
                    x = x < 0 ? x + rp.nc[0] : (x >= rp.nc[0] ? x - rp.nc[0] : x);
                    const int c2 = rp.coff + x + rp.nc[0] * (y + rp.nc[1] * z);
                    const int s0 = a.start[c2], s1 = a.start[c2 + 1];
                    for (int s = s0; s < s1; s++) {
                        const int j = a.list[s];
                        if (j == i)
                            continue;
                        double d[3];
                        d[0] = mimage(xi0 - a.X[3 * j], rp.l[0], rp.l2s[0], a.periodic);
                        d[1] = mimage(xi1 - a.X[3 * j + 1], rp.l[1], rp.l2s[1], a.periodic);
                        d[2] = mimage(xi2 - a.X[3 * j + 2], rp.l[2], rp.l2s[2], a.periodic);
                        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                        if (!(r2 < a.rmax2))
                            continue;
                        // ---------------- WCA (MD.cpp:1700-1758, potential.h:112-125)
                        if (a.wca) {
                            bool ex = false;
                            for (int k = 0; k < nexc && k < MAXEXCL; k++)
                                ex |= exc[k] == j;
                            const double r = sqrt(r2);
                            if (!ex && r < a.wca_pl) {
                                if constexpr (FP64) {
                                    double f1 = hs_force(r, a.wca_eps, a.wca_s6, a.wca_s12);
                                    flag_force(a, f1);
                                    for (int k = 0; k < 3; k++) {
                                        double fac = f1 * d[k] / r;
                                        Fi[k] += (R)fac;
                                        if (MODE == M_ENERGY)
                                            for (int l = 0; l < 3; l++)
                                                vir[3 * k + l] += 0.5 * d[l] * fac;
                                    }
                                    if (MODE == M_ENERGY)
                                        epot += 0.5 * wca_energy(r, a.wca_eps, a.wca_s6, a.wca_s12);
                                } else {
                                    float rf = (float)r;
                                    float f1 = hs_force(rf, (float)a.wca_eps, (float)a.wca_s6,
                                                        (float)a.wca_s12);
                                    flag_force(a, f1);
                                    float fr = f1 / rf;
                                    for (int k = 0; k < 3; k++) {
                                        float fac = fr * (float)d[k];
                                        Fi[k] += (R)fac;
                                        if (MODE == M_ENERGY)
                                            for (int l = 0; l < 3; l++)
                                                vir[3 * k + l] += 0.5 * d[l] * fac;
                                    }
                                    if (MODE == M_ENERGY)
                                        epot += 0.5 * (double)wca_energy(
                                                          rf, (float)a.wca_eps,
                                                          (float)a.wca_s6, (float)a.wca_s12);
                                }
                            }
                        }
                        // ------------- patches (LangevinR.cpp:911-1047, potentialtheta.h:517-750)
                        if (npi > 0 && r2 < a.maxd2) {
                            const int tj = a.meta[j] & 255;
                            const int npj = a.np[tj];
                            if (npj == 0)
                                continue;
                            // roles: p1 < p2, exchanged when type(p1) > type(p2)
                            int pa = i < j ? i : j, pc = i < j ? j : i;
                            int tpa = pa == i ? ti : tj, tpc = pa == i ? tj : ti;
                            if (tpa > tpc) {
                                int t = pa;
                                pa = pc;
                                pc = t;
                                t = tpa;
                                tpa = tpc;
                                tpc = t;
                            }
                            if (!((a.active_mask >> (tpa * MAXT + tpc)) & 1ull))
                                continue;
                            const bool i_is_a = pa == i;
                            const double rij = sqrt(r2);
                            const double sg = i_is_a ? 1.0 : -1.0;
                            double ud[3] = {sg * d[0] / rij, sg * d[1] / rij, sg * d[2] / rij};
                            R qj[9];
                            for (int k = 0; k < 9; k++)
                                qj[k] = a.Q[9 * (size_t)j + k];
                            R nj[MAXP][3];
                            for (int ib = 0; ib < npj; ib++)
                                lab_normal(qj, a.pdirs + 3 * (a.pstart[tj] + ib), nj[ib]);
                            const int npa = a.np[tpa], npc = a.np[tpc];
                            const int rowbase = rp.prow + a.rstart[tpa * MAXT + tpc];
                            R ua[3] = {(R)ud[0], (R)ud[1], (R)ud[2]};
                            R arg1[MAXP], arg2[MAXP];
                            for (int ia = 0; ia < npa; ia++) {
                                const R *n1 = i_is_a ? ni[ia] : nj[ia];
                                arg1[ia] = -(n1[0] * ua[0] + n1[1] * ua[1] + n1[2] * ua[2]);
                            }
                            for (int ib = 0; ib < npc; ib++) {
                                const R *n2 = i_is_a ? nj[ib] : ni[ib];
                                arg2[ib] = (n2[0] * ua[0] + n2[1] * ua[1] + n2[2] * ua[2]);
                            }
                            for (int ia = 0; ia < npa; ia++) {
                                for (int ib = 0; ib < npc; ib++) {
                                    const PatchRow pr = a.prows[rowbase + ia * npc + ib];
                                    if (rij > pr.idist)
                                        continue;
                                    R f[3] = {0, 0, 0}, t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
                                    const R *n1 = i_is_a ? ni[ia] : nj[ia];
                                    const R *n2 = i_is_a ? nj[ib] : ni[ib];
                                    double e;
                                    if constexpr (FP64) {
                                        if (!(arg1[ia] > pr.costm && arg2[ib] > pr.costm))
                                            continue;
                                        double fd[3] = {0, 0, 0}, t1d[3] = {0, 0, 0},
                                               t2d[3] = {0, 0, 0};
                                        double n1d[3] = {(double)n1[0], (double)n1[1], (double)n1[2]};
                                        double n2d[3] = {(double)n2[0], (double)n2[1], (double)n2[2]};
                                        e = kf2_pair_fp64(ud, rij, n1d, n2d, (double)arg1[ia],
                                                          (double)arg2[ib], pr, fd, t1d, t2d);
                                        for (int k = 0; k < 3; k++) {
                                            f[k] = (R)fd[k];
                                            t1[k] = (R)t1d[k];
                                            t2[k] = (R)t2d[k];
                                        }
                                    } else {
                                        // cone and d1 = 0 decisions in FP64 near their edges
                                        double a1 = (double)arg1[ia], a2 = (double)arg2[ib];
                                        const double tol = 4e-6;
                                        const int tpi = i_is_a ? ti : tj, tpj = i_is_a ? tj : ti;
                                        const R *qa = i_is_a ? qi : qj;
                                        const R *qc = i_is_a ? qj : qi;
                                        if (fabs(a1 - pr.costm) < tol || a1 > 0.9999) {
                                            double nd[3];
                                            lab_normal_d(qa,
                                                         a.pdirs + 3 * (a.pstart[tpi] + ia), nd);
                                            a1 = -(nd[0] * ud[0] + nd[1] * ud[1] + nd[2] * ud[2]);
                                        }
                                        if (fabs(a2 - pr.costm) < tol || a2 > 0.9999) {
                                            double nd[3];
                                            lab_normal_d(qc,
                                                         a.pdirs + 3 * (a.pstart[tpj] + ib), nd);
                                            a2 = (nd[0] * ud[0] + nd[1] * ud[1] + nd[2] * ud[2]);
                                        }
                                        if (!(a1 > pr.costm && a2 > pr.costm))
                                            continue;
                                        bool z1 = fabs(a1 * a1 - 1) < 1E-5;
                                        bool z2 = fabs(a2 * a2 - 1) < 1E-5;
                                        float ff[3] = {0, 0, 0}, t1f[3] = {0, 0, 0},
                                              t2f[3] = {0, 0, 0};
                                        float uf[3] = {(float)ud[0], (float)ud[1], (float)ud[2]};
                                        float n1f[3] = {(float)n1[0], (float)n1[1], (float)n1[2]};
                                        float n2f[3] = {(float)n2[0], (float)n2[1], (float)n2[2]};
                                        e = kf2_pair_fp32(uf, (float)rij, n1f, n2f,
                                                          (float)arg1[ia], (float)arg2[ib], z1,
                                                          z2, (float)pr.eps, (float)pr.d,
                                                          (float)pr.kk, ff, t1f, t2f);
                                        for (int k = 0; k < 3; k++) {
                                            f[k] = (R)ff[k];
                                            t1[k] = (R)t1f[k];
                                            t2[k] = (R)t2f[k];
                                        }
                                    }
                                    if (i_is_a) {
                                        for (int k = 0; k < 3; k++) {
                                            Fi[k] += f[k];
                                            Ti[k] += t1[k];
                                        }
                                    } else {
                                        for (int k = 0; k < 3; k++) {
                                            Fi[k] -= f[k];
                                            Ti[k] += t2[k];
                                        }
                                    }
                                    if (MODE == M_ENERGY) {
                                        epot += 0.5 * e;
                                        for (int k = 0; k < 3; k++)
                                            for (int l = 0; l < 3; l++)
                                                vir[3 * k + l] += 0.5 * d[l] * (double)(sg * f[k]);
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }
    }

    // ---------------------------------------- bonds (MD.cpp:1684-1764, 2037-2111)
    if (a.bond_off) {
        for (int b = a.bond_off[i]; b < a.bond_off[i + 1]; b++) {
            const int j = a.bond_j[b];
            double d[3];
            d[0] = mimage(xi0 - a.X[3 * j], rp.l[0], rp.l2s[0], a.periodic);
            d[1] = mimage(xi1 - a.X[3 * j + 1], rp.l[1], rp.l2s[1], a.periodic);
            d[2] = mimage(xi2 - a.X[3 * j + 2], rp.l[2], rp.l2s[2], a.periodic);
            const double dis = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const double r = sqrt(dis);
            const double k = a.bond_k[b], r0 = a.bond_r0[b];
            double f1;
            if (a.bond_kind == MDX_BOND_FENE)
                f1 = -k * r / (1 - (r * r) / (r0 * r0)); // FENEPotential::force (potential.h:356)
            else
                f1 = -k * (r - r0);
            flag_force(a, f1);
            for (int c = 0; c < 3; c++) {
                double fac = f1 * d[c] / r;
                Fi[c] += (R)fac;
                if (MODE == M_ENERGY)
                    for (int l = 0; l < 3; l++)
                        vir[3 * c + l] += 0.5 * d[l] * fac;
            }
            if (MODE == M_ENERGY) {
                if (a.bond_kind == MDX_BOND_FENE)
                    epot += 0.5 * (-(k / 2.) * r0 * r0 * log(1.0 - r * r / (r0 * r0)));
                else
                    epot += 0.5 * 0.5 * k * (r - r0) * (r - r0);
            }
        }
    }
    // ------------------------------------ angles (MD.cpp:1925-1983, potential3.h)
    if (a.ang_off) {
        for (int q = a.ang_off[i]; q < a.ang_off[i + 1]; q++) {
            const int ref = a.ang_ref[q];
            const int t = ref >> 2, role = ref & 3;
            const int p1 = a.ang_ijk[3 * t], p2 = a.ang_ijk[3 * t + 1], p3 = a.ang_ijk[3 * t + 2];
            double ab[3], bc[3];
            for (int k = 0; k < 3; k++) {
                ab[k] = mimage(a.X[3 * p2 + k] - a.X[3 * p1 + k], rp.l[k], rp.l2s[k], a.periodic);
                bc[k] = mimage(a.X[3 * p3 + k] - a.X[3 * p2 + k], rp.l[k], rp.l2s[k], a.periodic);
            }
            R abr[3] = {(R)ab[0], (R)ab[1], (R)ab[2]}, bcr[3] = {(R)bc[0], (R)bc[1], (R)bc[2]};
            R out[3];
            if (a.ang_form == MDX_ANGLE_HARMONIC)
                bend_harmonic(abr, bcr, (R)a.ang_k[t], (R)a.ang_t0[t], role, out);
            else
                bend_cosine(abr, bcr, (R)a.ang_k[t], role, out);
            for (int k = 0; k < 3; k++)
                Fi[k] += out[k];
            if (MODE == M_ENERGY && role == 1) {
                double c = (ab[0] * bc[0] + ab[1] * bc[1] + ab[2] * bc[2]) /
                           sqrt((ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]) *
                                (bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]));
                c = fmin(1.0, fmax(-1.0, c));
                if (a.ang_form == MDX_ANGLE_HARMONIC)
                    epot += 0.5 * a.ang_k[t] * (acos(c) - a.ang_t0[t]) * (acos(c) - a.ang_t0[t]);
                else
                    epot += a.ang_k[t] * (1 - c);
            }
        }
    }
    // -------------------------------------------------- tangential active force
    if (a.act_prev) {
        int nx = a.act_next[i] < 0 ? i : a.act_next[i];
        int pv = a.act_prev[i] < 0 ? i : a.act_prev[i];
        if (nx != pv) {
            double t[3];
            for (int k = 0; k < 3; k++)
                t[k] = mimage(a.X[3 * nx + k] - a.X[3 * pv + k], rp.l[k], rp.l2s[k], a.periodic);
            double inv = 1.0 / sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
            for (int k = 0; k < 3; k++)
                Fi[k] += (R)(a.fa * t[k] * inv);
        }
    }

    if (!(isfinite((double)Fi[0]) && isfinite((double)Fi[1]) && isfinite((double)Fi[2]) &&
          isfinite((double)Ti[0]) && isfinite((double)Ti[1]) && isfinite((double)Ti[2])))
        atomicOr(&a.ctl->status, 2u);

    // ------------------------------------------------------------ epilogues
    if (MODE == M_CONS) {
        for (int k = 0; k < 3; k++) {
            a.Fout[3 * i + k] = (double)Fi[k];
            a.Tout[3 * i + k] = (double)Ti[k];
        }
        return;
    }
    if (MODE == M_ENERGY) {
        double *o = a.obs + 10 * (size_t)i;
        o[0] = epot;
        for (int k = 0; k < 9; k++)
            o[1 + k] = vir[k];
        return;
    }
    if (MODE == M_BEAD) {
        // LangevinNVT::advance_mom (Langevin.cpp:254-269) then advance_pos
        // (Langevin.cpp:205-223) and binning of the new position
        R nz[6];
        load_noise(a, i, 3, nz);
        const R rr = (R)rp.bead_r;
        const R c5 = (R)1 - (R)(a.gamma * a.dt / 2.);
        const R c4 = (R)a.b_c2 * rr;
        double xn[3] = {xi0, xi1, xi2};
        for (int k = 0; k < 3; k++) {
            R pk = a.p[3 * i + k];
            pk = (R)a.b_c5c2 * pk + (R)a.b_c5c3q * Fi[k] + (c5 * c4 + rr) * nz[k];
            if (FP64)
                xn[k] = xn[k] + a.c1 * (double)pk;
            else
                xn[k] = xn[k] + (double)((R)a.c1 * pk);
            wrap1(xn[k], pk, rp.l[k], a.periodic);
            a.p[3 * i + k] = pk;
            a.Xn[3 * i + k] = xn[k];
        }
        bin_particle(a, i, rp, xn);
        return;
    }

    // rotor: create_forces_and_torques_sphere (LangevinR.cpp:475-578)
    R pl[3] = {a.p[3 * i], a.p[3 * i + 1], a.p[3 * i + 2]};
    R ll[3] = {a.L[3 * i], a.L[3 * i + 1], a.L[3 * i + 2]};
    R nz[6];
    load_noise(a, i, 6, nz);
    {
        const R *q = qi;
        R pbx = q[0] * pl[0] + q[1] * pl[1] + q[2] * pl[2];
        R pby = q[3] * pl[0] + q[4] * pl[1] + q[5] * pl[2];
        R pbz = q[6] * pl[0] + q[7] * pl[1] + q[8] * pl[2];
        const R gm = (R)(-a.gamma / a.m);
        R fx = gm * pbx + (R)rp.Rt * nz[0];
        R fy = gm * pby + (R)rp.Rt * nz[1];
        R fz = gm * pbz + (R)rp.Rt * nz[2];
        R tbx = (R)(-a.gammar / a.I[0]) * ll[0] + (R)rp.Rr * nz[3];
        R tby = (R)(-a.gammar / a.I[1]) * ll[1] + (R)rp.Rr * nz[4];
        R tbz = (R)(-a.gammar / a.I[2]) * ll[2] + (R)rp.Rr * nz[5];
        R Fx = Fi[0] + fx * q[0] + fy * q[3] + fz * q[6];
        R Fy = Fi[1] + fx * q[1] + fy * q[4] + fz * q[7];
        R Fz = Fi[2] + fx * q[2] + fy * q[5] + fz * q[8];
        R Tx = tbx + Ti[0] * q[0] + Ti[1] * q[1] + Ti[2] * q[2];
        R Ty = tby + Ti[0] * q[3] + Ti[1] * q[4] + Ti[2] * q[5];
        R Tz = tbz + Ti[0] * q[6] + Ti[1] * q[7] + Ti[2] * q[8];
        Fi[0] = Fx;
        Fi[1] = Fy;
        Fi[2] = Fz;
        Ti[0] = Tx;
        Ti[1] = Ty;
        Ti[2] = Tz;
    }
    if (MODE == M_PRIME) {
        for (int k = 0; k < 3; k++) {
            a.F[3 * i + k] = Fi[k];
            a.T[3 * i + k] = Ti[k];
        }
        return;
    }
    // advancemom_halfstep (LangevinR.cpp:327-347)
    const R hdt = (R)a.hdt;
    for (int k = 0; k < 3; k++) {
        pl[k] += hdt * Fi[k];
        ll[k] += hdt * Ti[k];
    }
    if (MODE == M_STEP) {
        for (int k = 0; k < 3; k++) {
            a.p[3 * i + k] = pl[k];
            a.L[3 * i + k] = ll[k];
            a.F[3 * i + k] = Fi[k];
            a.T[3 * i + k] = Ti[k];
        }
        return;
    }
    // M_FUSED: first half of the next step with the same F, T
    for (int k = 0; k < 3; k++) {
        pl[k] += hdt * Fi[k];
        ll[k] += hdt * Ti[k];
    }
    double xn[3] = {xi0, xi1, xi2};
    for (int k = 0; k < 3; k++) {
        if (FP64)
            xn[k] = xn[k] + a.c1 * (double)pl[k];
        else
            xn[k] = xn[k] + (double)((R)a.c1 * pl[k]);
        wrap1(xn[k], pl[k], rp.l[k], a.periodic);
    }
    const R Ir[3] = {(R)a.I[0], (R)a.I[1], (R)a.I[2]};
    rotate((R)a.dt, Ir, ll, qi);
    if (!FP64 && a.reortho_every > 0 && ((a.ctl->epoch + 1) % a.reortho_every) == 0)
        reortho(qi);
    for (int k = 0; k < 3; k++) {
        a.p[3 * i + k] = pl[k];
        a.L[3 * i + k] = ll[k];
        a.Xn[3 * i + k] = xn[k];
    }
    for (int k = 0; k < 9; k++)
        a.Qn[9 * (size_t)i + k] = qi[k];
    bin_particle(a, i, rp, xn);
}

// first half of a rotor step from the carried F, T: half kick, drift + wrap
// (Langevin.cpp:205-223), rotate (LangevinR.cpp:687-695), then binning
template <class R> __global__ void __launch_bounds__(256) k_kick_drift(KArgs<R> a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N)
        return;
    constexpr bool FP64 = sizeof(R) == 8;
    const DevRep rp = a.reps[a.meta[i] >> 8];
    const R hdt = (R)a.hdt;
    R pl[3], ll[3], q[9];
    double xn[3];
    for (int k = 0; k < 3; k++) {
        pl[k] = a.p[3 * i + k] + hdt * a.F[3 * i + k];
        ll[k] = a.L[3 * i + k] + hdt * a.T[3 * i + k];
        xn[k] = a.X[3 * i + k];
    }
    for (int k = 0; k < 9; k++)
        q[k] = a.Q[9 * (size_t)i + k];
    for (int k = 0; k < 3; k++) {
        if (FP64)
            xn[k] = xn[k] + a.c1 * (double)pl[k];
        else
            xn[k] = xn[k] + (double)((R)a.c1 * pl[k]);
        wrap1(xn[k], pl[k], rp.l[k], a.periodic);
    }
    const R Ir[3] = {(R)a.I[0], (R)a.I[1], (R)a.I[2]};
    rotate((R)a.dt, Ir, ll, q);
    if (!FP64 && a.reortho_every > 0 && ((a.ctl->epoch + 1) % a.reortho_every) == 0)
        reortho(q);
    for (int k = 0; k < 3; k++) {
        a.p[3 * i + k] = pl[k];
        a.L[3 * i + k] = ll[k];
        a.Xn[3 * i + k] = xn[k];
    }
    for (int k = 0; k < 9; k++)
        a.Qn[9 * (size_t)i + k] = q[k];
    bin_particle(a, i, rp, xn);
}

// binning of the current positions (after an upload)
template <class R> __global__ void __launch_bounds__(256) k_bin(KArgs<R> a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N)
        return;
    const DevRep rp = a.reps[a.meta[i] >> 8];
    double x[3] = {a.X[3 * i], a.X[3 * i + 1], a.X[3 * i + 2]};
    bin_particle(a, i, rp, x);
}

// --------------------------------------------------------------- cell lists
// exclusive scan of count[0..n) into start[0..n], count zeroed after use.
// phase 1: per block (1024 elements) local scan and block total
__global__ void __launch_bounds__(256) k_scan1(const int *count, int *start, int *bsum, int n)
{
    __shared__ int sh[256];
    const int base = blockIdx.x * 1024 + threadIdx.x * 4;
    int v[4], s = 0;
    for (int k = 0; k < 4; k++) {
        v[k] = base + k < n ? count[base + k] : 0;
        s += v[k];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    int run = sh[threadIdx.x] - s;
    for (int k = 0; k < 4; k++) {
        if (base + k < n)
            start[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 255)
        bsum[blockIdx.x] = sh[255];
}

// phase 2: exclusive scan of the block totals (one block)
__global__ void __launch_bounds__(1024) k_scan2(int *bsum, int nb, int *total)
{
    __shared__ int sh[1024];
    const int per = (nb + 1023) / 1024;
    const int b0 = threadIdx.x * per;
    int s = 0;
    for (int k = 0; k < per; k++)
        if (b0 + k < nb)
            s += bsum[b0 + k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    int run = sh[threadIdx.x] - s;
    for (int k = 0; k < per; k++)
        if (b0 + k < nb) {
            int v = bsum[b0 + k];
            bsum[b0 + k] = run;
            run += v;
        }
    if (threadIdx.x == 1023)
        *total = sh[1023];
}

// phase 3: add block offsets, zero the counts, close the array
__global__ void __launch_bounds__(256) k_scan3(int *count, int *start, const int *bsum, int n,
                                               const int *total)
{
    const int base = blockIdx.x * 1024 + threadIdx.x * 4;
    const int off = bsum[blockIdx.x];
    for (int k = 0; k < 4; k++)
        if (base + k < n) {
            start[base + k] += off;
            count[base + k] = 0;
        }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        start[n] = *total;
}

__global__ void __launch_bounds__(256) k_fill(int N, const int *cell, const int *rank,
                                              const int *start, int *list)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N)
        list[start[cell[i]] + rank[i]] = i;
}

// ascending particle order inside each cell (makes the gather order, and so
// the floating-point sums, independent of atomic arrival order)
__global__ void __launch_bounds__(256) k_cellsort(int ncells, const int *start, int *list)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells)
        return;
    int s0 = start[c], s1 = start[c + 1];
    for (int s = s0 + 1; s < s1; s++) {
        int v = list[s];
        int t = s - 1;
        while (t >= s0 && list[t] > v) {
            list[t + 1] = list[t];
            t--;
        }
        list[t + 1] = v;
    }
}

__global__ void k_epoch_inc(StepCtl *ctl)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        ctl->epoch += 1;
}

__global__ void k_set_call_base(StepCtl *ctl)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        ctl->call_base = ctl->epoch;
}

// ---------------------------------------------------- parity entry points
// cube::assign_box with a caller-given n for particles [first, first+count)
__global__ void k_assign_custom(const double *X, int first, int count, double l0, double l1,
                                double l2, int n0, int n1, int n2, int *out)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count)
        return;
    const double *x = X + 3 * (size_t)(first + k);
    int cx = axis_cell(x[0], n0, l0), cy = axis_cell(x[1], n1, l1), cz = axis_cell(x[2], n2, l2);
    out[k] = cx + n0 * (cy + n1 * cz);
}

// cube::distance_less_than (geometry.h:145-164)
__device__ __forceinline__ bool dlt(const double *xi, const double *xj, const double l[3],
                                    int periodic, double cut)
{
    double temp = 0.0;
    const double c2 = cut * cut;
    for (int k = 0; k < 3; k++) {
        double dx = xi[k] - xj[k];
        if (periodic) {
            if (fabs(dx) > l[k] * 0.5)
                dx = dx - (dx >= 0 ? fabs(l[k]) : -fabs(l[k]));
        }
        dx = dx * dx;
        if (dx > c2)
            return false;
        temp += dx;
    }
    return temp < c2;
}

// pass 0: count (write=0) or write sorted partner lists (write=1) of i < j
__global__ void k_pairs(const double *X, int first, int count, double l0, double l1, double l2,
                        int periodic, int n0, int n1, int n2, double cut, const int *cellv,
                        const int *cstart, const int *clist, int *cnt, const long long *off,
                        int *ij, long long cap, int write)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count)
        return;
    const double l[3] = {l0, l1, l2};
    const int nn[3] = {n0, n1, n2};
    const double *xi = X + 3 * (size_t)(first + k);
    const int c = cellv[k];
    const int cc[3] = {c % n0, (c / n0) % n1, c / (n0 * n1)};
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = nn[a] >= 3 ? -1 : 0;
        hi[a] = nn[a] >= 2 ? 1 : 0;
    }
    int found = 0;
    long long o = write ? off[k] : 0;
    for (int oz = lo[2]; oz <= hi[2]; oz++)
        for (int oy = lo[1]; oy <= hi[1]; oy++)
            for (int ox = lo[0]; ox <= hi[0]; ox++) {
                int w[3] = {cc[0] + ox, cc[1] + oy, cc[2] + oz};
                for (int a = 0; a < 3; a++)
                    w[a] = w[a] < 0 ? w[a] + nn[a] : (w[a] >= nn[a] ? w[a] - nn[a] : w[a]);
                int c2 = w[0] + n0 * (w[1] + n1 * w[2]);
                for (int s = cstart[c2]; s < cstart[c2 + 1]; s++) {
                    int j = clist[s];
                    if (j <= k)
                        continue;
                    if (dlt(xi, X + 3 * (size_t)(first + j), l, periodic, cut)) {
                        if (write && o + found < cap) {
                            ij[2 * (o + found)] = k;
                            ij[2 * (o + found) + 1] = j;
                        }
                        found++;
                    }
                }
            }
    if (!write) {
        cnt[k] = found;
        return;
    }
    // order the partners of k ascending
    long long e = o + found < cap ? o + found : cap;
    for (long long s = o + 1; s < e; s++) {
        int v = ij[2 * s + 1];
        long long t = s - 1;
        while (t >= o && ij[2 * t + 1] > v) {
            ij[2 * (t + 1) + 1] = ij[2 * t + 1];
            t--;
        }
        ij[2 * (t + 1) + 1] = v;
    }
}

// ------------------------------------------------------------ reductions
// per-block partial sums (fixed order) of 16 observables over [first, first+count)
template <class R>
__global__ void __launch_bounds__(256) k_obs(int first, int count, const R *p, const R *L,
                                             const double *obs, int has_obs, double m, double I0,
                                             double *partial)
{
    __shared__ double sh[256];
    double acc[16];
    for (int k = 0; k < 16; k++)
        acc[k] = 0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < count; s += gridDim.x * blockDim.x) {
        int i = first + s;
        for (int k = 0; k < 3; k++) {
            double pk = (double)p[3 * i + k], lk = L ? (double)L[3 * i + k] : 0.0;
            acc[k] += (pk * pk) / (2. * m);
            acc[3 + k] += (lk * lk) / (2. * I0);
        }
        if (has_obs)
            for (int k = 0; k < 10; k++)
                acc[6 + k] += obs[10 * (size_t)i + k];
    }
    for (int k = 0; k < 16; k++) {
        sh[threadIdx.x] = acc[k];
        __syncthreads();
        for (int off = 128; off > 0; off >>= 1) {
            if (threadIdx.x < off)
                sh[threadIdx.x] += sh[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            partial[16 * blockIdx.x + k] = sh[0];
        __syncthreads();
    }
}

// precision conversion helpers (device <-> host FP64 layout)
template <class A, class B> __global__ void k_convert(const A *in, B *out, size_t n)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        out[k] = (B)in[k];
}

// ================================================================ host side
namespace {

struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
};

} // namespace

struct mdx_engine {
    // description
    int nrep = 1;
    int N = 0;
    int periodic = 1;
    int prec = MDX_PREC_FP64;
    int integ = MDX_INTEG_ROTOR;
    double dt = 0.005, m = 1, gamma = 1, gammar = 1, I[3] = {0.1, 0.1, 0.1};
    unsigned long long seed = 0;
    int device = 0;
    std::vector<int> nper, first;
    std::vector<double> box, kT;
    hipStream_t stream = nullptr;
    // interactions
    int wca = 0;
    double wca_eps = 1, wca_sigma = 1, wca_pl = 1.12246;
    int ntypes = 0;
    std::vector<int> type, np, pstart, rstart;
    std::vector<double> pdirs;
    std::vector<double> prow_raw; // per replica rows x 3
    int nrows = 0;
    double maxd2 = 2.4;
    unsigned long long active_mask = 0;
    int bond_kind = MDX_BOND_HARMONIC, ang_form = MDX_ANGLE_HARMONIC;
    long long nbonds = 0, nangles = 0, nexcl = 0;
    double fa = 0;
    int has_active = 0;
    double fguard = 1e4;
    int reortho_every = 100;
    // grid
    int ncells = 0;
    std::vector<DevRep> reps;
    bool grid_dirty = true;
    // state flags
    int cur = 0;
    bool binned = false;
    bool primed = false;
    unsigned long long steps_taken = 0;
    size_t bytes_held = 0;
    // device buffers
    Buf X[2], Q[2], P, Lm, Fc, Tc, Fout, Tout, obs;
    Buf meta, dreps, prows, pdirs_d;
    Buf cell, rank, count, start, list, bsum, total;
    Buf bond_off, bond_j, bond_k, bond_r0;
    Buf ang_off, ang_ref, ang_ijk, ang_k, ang_t0;
    Buf exc_off, exc_j, act_prev, act_next;
    Buf noise, ctl, partial, tmpd, tmpi, tmpi2, tmpi3, tmpll;

    size_t rsz() const { return prec == MDX_PREC_FP64 ? 8 : 4; }
};

static int dalloc(mdx_engine *e, Buf &b, size_t bytes)
{
    if (bytes == 0)
        bytes = 8;
    if (b.p && b.bytes >= bytes)
        return MDX_OK;
    if (b.p) {
        HIPCHK(hipFree(b.p));
        e->bytes_held -= b.bytes;
        b.p = nullptr;
    }
    HIPCHK(hipMalloc(&b.p, bytes));
    b.bytes = bytes;
    e->bytes_held += bytes;
    HIPCHK(hipMemsetAsync(b.p, 0, bytes, e->stream));
    return MDX_OK;
}

static void dfree(Buf &b)
{
    if (b.p)
        (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

template <class T> static int upload(mdx_engine *e, Buf &b, const T *h, size_t n)
{
    int rc = dalloc(e, b, n * sizeof(T));
    if (rc)
        return rc;
    if (n)
        HIPCHK(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MDX_OK;
}

static inline unsigned nblk(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// interaction reach: WCA cut-off and the patch reach min(sqrt(maxd2), 1.4 d)
static double reach(const mdx_engine *e)
{
    double r = 0;
    if (e->wca)
        r = std::max(r, e->wca_pl);
    if (e->ntypes > 0 && e->active_mask) {
        double dmax = 0;
        for (size_t k = 0; k < e->prow_raw.size() / 3; k++)
            if (e->prow_raw[3 * k] != 0.0)
                dmax = std::max(dmax, 1.4 * e->prow_raw[3 * k + 1]);
        r = std::max(r, std::min(std::sqrt(e->maxd2), dmax));
    }
    return r;
}

static int build_grid(mdx_engine *e)
{
    if (!e->grid_dirty)
        return MDX_OK;
    double rc = reach(e);
    e->reps.resize(e->nrep);
    int coff = 0;
    for (int r = 0; r < e->nrep; r++) {
        DevRep &d = e->reps[r];
        for (int k = 0; k < 3; k++) {
            d.l[k] = e->box[3 * r + k];
            d.l2s[k] = (d.l[k] / 2.) * (d.l[k] / 2.);
            int n = rc > 0 ? (int)std::floor(d.l[k] / (rc * (1 + 1e-9))) : 1;
            d.nc[k] = std::max(1, std::min(n, 1024));
        }
        d.coff = coff;
        coff += d.nc[0] * d.nc[1] * d.nc[2];
        d.first = e->first[r];
        d.count = e->nper[r];
        double kT = e->kT[r];
        d.Rt = std::sqrt(2 * e->gamma * kT / e->dt);
        d.Rr = std::sqrt(2 * e->gammar * kT / e->dt);
        d.bead_r = std::sqrt(0.5 * kT * (e->gamma) * (e->m) * (e->dt));
        d.prow = r * e->nrows;
    }
    e->ncells = coff;
    int rc2;
    if ((rc2 = upload(e, e->dreps, e->reps.data(), e->reps.size())))
        return rc2;
    if ((rc2 = dalloc(e, e->count, sizeof(int) * (size_t)(e->ncells + 1))))
        return rc2;
    HIPCHK(hipMemsetAsync(e->count.p, 0, sizeof(int) * (size_t)(e->ncells + 1), e->stream));
    if ((rc2 = dalloc(e, e->start, sizeof(int) * (size_t)(e->ncells + 1))))
        return rc2;
    int nb = (e->ncells + 1023) / 1024;
    if ((rc2 = dalloc(e, e->bsum, sizeof(int) * (size_t)std::max(nb, 1))))
        return rc2;
    e->grid_dirty = false;
    e->binned = false;
    e->primed = false;
    return MDX_OK;
}

template <class R> static KArgs<R> make_args(mdx_engine *e)
{
    KArgs<R> a;
    memset(&a, 0, sizeof a);
    a.N = e->N;
    a.rotor = e->integ == MDX_INTEG_ROTOR;
    a.periodic = e->periodic;
    a.meta = (const uint32_t *)e->meta.p;
    a.reps = (const DevRep *)e->dreps.p;
    a.X = (const double *)e->X[e->cur].p;
    a.Xn = (double *)e->X[e->cur ^ 1].p;
    a.Q = (const R *)e->Q[e->cur].p;
    a.Qn = (R *)e->Q[e->cur ^ 1].p;
    a.p = (R *)e->P.p;
    a.L = (R *)e->Lm.p;
    a.F = (R *)e->Fc.p;
    a.T = (R *)e->Tc.p;
    a.Fout = (double *)e->Fout.p;
    a.Tout = (double *)e->Tout.p;
    a.obs = (double *)e->obs.p;
    a.cell = (int *)e->cell.p;
    a.rank = (int *)e->rank.p;
    a.count = (int *)e->count.p;
    a.start = (const int *)e->start.p;
    a.list = (const int *)e->list.p;
    a.wca = e->wca;
    a.wca_eps = e->wca_eps;
    double s3 = e->wca_sigma * e->wca_sigma * e->wca_sigma;
    a.wca_s6 = s3 * s3;
    a.wca_s12 = a.wca_s6 * a.wca_s6;
    a.wca_pl = e->wca_pl;
    a.has_patches = e->ntypes > 0 && e->integ == MDX_INTEG_ROTOR;
    for (int t = 0; t < MAXT; t++) {
        a.np[t] = t < e->ntypes ? e->np[t] : 0;
        a.pstart[t] = t < e->ntypes ? e->pstart[t] : 0;
    }
    for (int k = 0; k < MAXT * MAXT; k++)
        a.rstart[k] = k < (int)e->rstart.size() ? e->rstart[k] : 0;
    a.active_mask = e->active_mask;
    a.pdirs = (const double *)e->pdirs_d.p;
    a.prows = (const PatchRow *)e->prows.p;
    a.maxd2 = e->maxd2;
    double pr = std::max(e->wca ? e->wca_pl * e->wca_pl * (1 + 1e-12) : 0.0,
                         a.has_patches ? e->maxd2 : 0.0);
    a.rmax2 = pr;
    if (e->nbonds) {
        a.bond_off = (const int *)e->bond_off.p;
        a.bond_j = (const int *)e->bond_j.p;
        a.bond_k = (const double *)e->bond_k.p;
        a.bond_r0 = (const double *)e->bond_r0.p;
    }
    a.bond_kind = e->bond_kind;
    if (e->nangles) {
        a.ang_off = (const int *)e->ang_off.p;
        a.ang_ref = (const int *)e->ang_ref.p;
        a.ang_ijk = (const int *)e->ang_ijk.p;
        a.ang_k = (const double *)e->ang_k.p;
        a.ang_t0 = (const double *)e->ang_t0.p;
    }
    a.ang_form = e->ang_form;
    if (e->nexcl) {
        a.exc_off = (const int *)e->exc_off.p;
        a.exc_j = (const int *)e->exc_j.p;
    }
    if (e->has_active) {
        a.act_prev = (const int *)e->act_prev.p;
        a.act_next = (const int *)e->act_next.p;
        a.fa = e->fa;
    }
    a.dt = e->dt;
    a.hdt = e->dt / 2.;
    a.c1 = e->dt / e->m;
    a.m = e->m;
    a.gamma = e->gamma;
    a.gammar = e->gammar;
    for (int k = 0; k < 3; k++)
        a.I[k] = e->I[k];
    // LangevinNVT coefficients (Langevin.cpp:44-52); r depends on kT (per replica)
    double d = (e->gamma * e->dt / 2.), q = (e->dt) / 2.;
    double c2 = (1.0 / (1.0 + (d))), c3 = (1.0 / (1.0 + (d))) * q, c5 = (1 - (d));
    a.b_c2 = c2;
    a.b_c5c2 = c5 * c2;
    a.b_c5c3q = (c5 * (c3) + q);
    a.reortho_every = e->reortho_every;
    a.key0 = (uint32_t)e->seed;
    a.key1 = (uint32_t)(e->seed >> 32);
    a.noise = (const double *)nullptr;
    a.ctl = (StepCtl *)e->ctl.p;
    a.fguard = e->fguard;
    return a;
}

// scan + fill + per-cell ordering of the cell list built by the binning
static int finish_bins(mdx_engine *e)
{
    const int n = e->ncells;
    const int nb = (n + 1023) / 1024;
    hipLaunchKernelGGL(k_scan1, dim3(nb), dim3(256), 0, e->stream, (const int *)e->count.p,
                       (int *)e->start.p, (int *)e->bsum.p, n);
    hipLaunchKernelGGL(k_scan2, dim3(1), dim3(1024), 0, e->stream, (int *)e->bsum.p, nb,
                       (int *)e->total.p);
    hipLaunchKernelGGL(k_scan3, dim3(nb), dim3(256), 0, e->stream, (int *)e->count.p,
                       (int *)e->start.p, (const int *)e->bsum.p, n, (const int *)e->total.p);
    hipLaunchKernelGGL(k_fill, dim3(nblk(e->N)), dim3(256), 0, e->stream, e->N,
                       (const int *)e->cell.p, (const int *)e->rank.p, (const int *)e->start.p,
                       (int *)e->list.p);
    hipLaunchKernelGGL(k_cellsort, dim3(nblk(n)), dim3(256), 0, e->stream, n,
                       (const int *)e->start.p, (int *)e->list.p);
    HIPCHK(hipGetLastError());
    return MDX_OK;
}

template <class R> static int bin_current(mdx_engine *e)
{
    int rc = build_grid(e);
    if (rc)
        return rc;
    KArgs<R> a = make_args<R>(e);
    hipLaunchKernelGGL(k_bin<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, a);
    HIPCHK(hipGetLastError());
    rc = finish_bins(e);
    if (rc)
        return rc;
    e->binned = true;
    return MDX_OK;
}

template <class R, int MODE> static int launch_gather(mdx_engine *e, const double *noise_dev)
{
    KArgs<R> a = make_args<R>(e);
    a.noise = noise_dev;
    hipLaunchKernelGGL((k_gather<R, MODE>), dim3(nblk(e->N)), dim3(256), 0, e->stream, a);
    HIPCHK(hipGetLastError());
    return MDX_OK;
}

static int read_status(mdx_engine *e, unsigned int *st)
{
    StepCtl c;
    HIPCHK(hipMemcpyAsync(&c, e->ctl.p, sizeof c, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    *st = c.status;
    if (c.status) {
        unsigned int z = 0;
        HIPCHK(hipMemcpy((char *)e->ctl.p + offsetof(StepCtl, status), &z, sizeof z,
                         hipMemcpyHostToDevice));
    }
    return MDX_OK;
}

static int status_to_rc(unsigned int st)
{
    if (st & 2u)
        return set_err(MDX_ERR_NAN, "non-finite force or torque");
    if (st & 1u)
        return set_err(MDX_ERR_FORCE, "pair force above the guard (force too big)");
    return MDX_OK;
}

static int noise_upload(mdx_engine *e, const double *h, size_t n, const double **dptr)
{
    *dptr = nullptr;
    if (!h)
        return MDX_OK;
    int rc = dalloc(e, e->noise, n * sizeof(double));
    if (rc)
        return rc;
    HIPCHK(hipMemcpyAsync(e->noise.p, h, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    *dptr = (const double *)e->noise.p;
    return MDX_OK;
}

template <class R> static int do_prime(mdx_engine *e, const double *noise0)
{
    int rc = bin_current<R>(e);
    if (rc)
        return rc;
    const double *nd;
    if ((rc = noise_upload(e, noise0, (size_t)e->N * 6, &nd)))
        return rc;
    hipLaunchKernelGGL(k_set_call_base, dim3(1), dim3(1), 0, e->stream, (StepCtl *)e->ctl.p);
    if ((rc = launch_gather<R, M_PRIME>(e, nd)))
        return rc;
    hipLaunchKernelGGL(k_epoch_inc, dim3(1), dim3(1), 0, e->stream, (StepCtl *)e->ctl.p);
    unsigned int st;
    if ((rc = read_status(e, &st)))
        return rc;
    e->primed = true;
    return status_to_rc(st);
}

template <class R> static int do_step_rotor(mdx_engine *e, int nsteps, const double *noise)
{
    if (!e->primed)
        return set_err(MDX_ERR_STATE, "mdx_step before mdx_prime (the carried F, T are unset)");
    int rc;
    const double *nd;
    if ((rc = noise_upload(e, noise, (size_t)nsteps * e->N * 6, &nd)))
        return rc;
    hipLaunchKernelGGL(k_set_call_base, dim3(1), dim3(1), 0, e->stream, (StepCtl *)e->ctl.p);
    {
        KArgs<R> a = make_args<R>(e);
        hipLaunchKernelGGL(k_kick_drift<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, a);
        HIPCHK(hipGetLastError());
        e->cur ^= 1;
    }
    for (int s = 0; s < nsteps; s++) {
        if ((rc = finish_bins(e)))
            return rc;
        if (s + 1 < nsteps) {
            if ((rc = launch_gather<R, M_FUSED>(e, nd)))
                return rc;
            e->cur ^= 1;
        } else {
            if ((rc = launch_gather<R, M_STEP>(e, nd)))
                return rc;
        }
        hipLaunchKernelGGL(k_epoch_inc, dim3(1), dim3(1), 0, e->stream, (StepCtl *)e->ctl.p);
    }
    e->binned = true;
    e->steps_taken += nsteps;
    unsigned int st;
    if ((rc = read_status(e, &st)))
        return rc;
    return status_to_rc(st);
}

template <class R> static int do_step_bead(mdx_engine *e, int nsteps, const double *noise)
{
    int rc;
    if (!e->binned && (rc = bin_current<R>(e)))
        return rc;
    const double *nd;
    if ((rc = noise_upload(e, noise, (size_t)nsteps * e->N * 3, &nd)))
        return rc;
    hipLaunchKernelGGL(k_set_call_base, dim3(1), dim3(1), 0, e->stream, (StepCtl *)e->ctl.p);
    for (int s = 0; s < nsteps; s++) {
        if ((rc = launch_gather<R, M_BEAD>(e, nd)))
            return rc;
        e->cur ^= 1;
        hipLaunchKernelGGL(k_epoch_inc, dim3(1), dim3(1), 0, e->stream, (StepCtl *)e->ctl.p);
        if ((rc = finish_bins(e)))
            return rc;
    }
    e->binned = true;
    e->steps_taken += nsteps;
    unsigned int st;
    if ((rc = read_status(e, &st)))
        return rc;
    return status_to_rc(st);
}

// ================================================================== C ABI
// (the declarations in mdx.h give these definitions C linkage)

const char *mdx_last_error(void) { return g_err.c_str(); }

int mdx_create(const mdx_desc *d, mdx_engine **out)
{
    if (!d || !out)
        return set_err(MDX_ERR_ARG, "null descriptor or output");
    *out = nullptr;
    if (d->nreplicas < 1 || !d->n_per_replica || !d->box || !d->kT)
        return set_err(MDX_ERR_ARG, "nreplicas >= 1 with counts, boxes and kT required");
    if (d->precision != MDX_PREC_FP64 && d->precision != MDX_PREC_MIXED)
        return set_err(MDX_ERR_ARG, "unknown precision %d", d->precision);
    if (d->integrator != MDX_INTEG_ROTOR && d->integrator != MDX_INTEG_BEAD)
        return set_err(MDX_ERR_ARG, "unknown integrator %d", d->integrator);
    if (!(d->dt > 0) || !(d->m > 0) || !(d->gamma >= 0) || !(d->gammar >= 0))
        return set_err(MDX_ERR_ARG, "dt, m must be > 0 and gamma, gammar >= 0");
    for (int k = 0; k < 3; k++)
        if (!(d->inertia[k] > 0))
            return set_err(MDX_ERR_ARG, "inertia must be > 0");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (d->device < 0 || d->device >= ndev)
        return set_err(MDX_ERR_ARG, "device %d not present (%d devices)", d->device, ndev);
    HIPCHK(hipSetDevice(d->device));
    mdx_engine *e = new mdx_engine();
    e->nrep = d->nreplicas;
    e->periodic = d->periodic ? 1 : 0;
    e->prec = d->precision;
    e->integ = d->integrator;
    e->dt = d->dt;
    e->m = d->m;
    e->gamma = d->gamma;
    e->gammar = d->gammar;
    for (int k = 0; k < 3; k++)
        e->I[k] = d->inertia[k];
    e->seed = d->seed;
    e->device = d->device;
    long long N = 0;
    for (int r = 0; r < e->nrep; r++) {
        if (d->n_per_replica[r] < 1) {
            delete e;
            return set_err(MDX_ERR_ARG, "replica %d has no particles", r);
        }
        for (int k = 0; k < 3; k++)
            if (!(d->box[3 * r + k] > 0)) {
                delete e;
                return set_err(MDX_ERR_ARG, "replica %d box edge %d must be > 0", r, k);
            }
        e->first.push_back((int)N);
        e->nper.push_back(d->n_per_replica[r]);
        N += d->n_per_replica[r];
        e->kT.push_back(d->kT[r]);
        for (int k = 0; k < 3; k++)
            e->box.push_back(d->box[3 * r + k]);
    }
    if (N > (1ll << 30) || e->nrep > (1 << 23)) {
        delete e;
        return set_err(MDX_ERR_ARG, "too many particles or replicas");
    }
    e->N = (int)N;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return set_err(MDX_ERR_HIP, "hipStreamCreate failed");
    }
    size_t rs = e->rsz();
    int rc = 0;
    rc |= dalloc(e, e->X[0], 24 * N);
    rc |= dalloc(e, e->X[1], 24 * N);
    rc |= dalloc(e, e->Q[0], 9 * rs * N);
    rc |= dalloc(e, e->Q[1], 9 * rs * N);
    rc |= dalloc(e, e->P, 3 * rs * N);
    rc |= dalloc(e, e->Lm, 3 * rs * N);
    rc |= dalloc(e, e->Fc, 3 * rs * N);
    rc |= dalloc(e, e->Tc, 3 * rs * N);
    rc |= dalloc(e, e->cell, 4 * N);
    rc |= dalloc(e, e->rank, 4 * N);
    rc |= dalloc(e, e->list, 4 * N);
    rc |= dalloc(e, e->meta, 4 * N);
    rc |= dalloc(e, e->ctl, sizeof(StepCtl));
    rc |= dalloc(e, e->total, 16);
    if (rc) {
        mdx_destroy(e);
        return set_err(MDX_ERR_HIP, "device allocation failed (%s)", g_err.c_str());
    }
    std::vector<uint32_t> meta(N);
    for (int r = 0; r < e->nrep; r++)
        for (int k = 0; k < e->nper[r]; k++)
            meta[e->first[r] + k] = ((uint32_t)r << 8);
    if ((rc = upload(e, e->meta, meta.data(), meta.size()))) {
        mdx_destroy(e);
        return rc;
    }
    *out = e;
    return MDX_OK;
}

void mdx_destroy(mdx_engine *e)
{
    if (!e)
        return;
    (void)hipSetDevice(e->device);
    if (e->stream)
        (void)hipStreamSynchronize(e->stream);
    Buf *bufs[] = {&e->X[0],   &e->X[1],    &e->Q[0],    &e->Q[1],   &e->P,       &e->Lm,
                   &e->Fc,     &e->Tc,      &e->Fout,    &e->Tout,   &e->obs,     &e->meta,
                   &e->dreps,  &e->prows,   &e->pdirs_d, &e->cell,   &e->rank,    &e->count,
                   &e->start,  &e->list,    &e->bsum,    &e->total,  &e->bond_off, &e->bond_j,
                   &e->bond_k, &e->bond_r0, &e->ang_off, &e->ang_ref, &e->ang_ijk, &e->ang_k,
                   &e->ang_t0, &e->exc_off, &e->exc_j,   &e->act_prev, &e->act_next, &e->noise,
                   &e->ctl,    &e->partial, &e->tmpd,    &e->tmpi,   &e->tmpi2,   &e->tmpi3,
                   &e->tmpll};
    for (Buf *b : bufs)
        dfree(*b);
    if (e->stream)
        (void)hipStreamDestroy(e->stream);
    delete e;
}

template <class R> static int upload_real(mdx_engine *e, Buf &dst, const double *h, size_t n)
{
    if (sizeof(R) == 8)
        return upload(e, dst, h, n);
    int rc = upload(e, e->tmpd, h, n);
    if (rc)
        return rc;
    hipLaunchKernelGGL((k_convert<double, R>), dim3(nblk((long long)n)), dim3(256), 0, e->stream,
                       (const double *)e->tmpd.p, (R *)dst.p, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    return MDX_OK;
}

template <class R> static int download_real(mdx_engine *e, const Buf &src, double *h, size_t n)
{
    if (sizeof(R) == 8) {
        HIPCHK(hipMemcpyAsync(h, src.p, n * 8, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        return MDX_OK;
    }
    int rc = dalloc(e, e->tmpd, n * 8);
    if (rc)
        return rc;
    hipLaunchKernelGGL((k_convert<R, double>), dim3(nblk((long long)n)), dim3(256), 0, e->stream,
                       (const R *)src.p, (double *)e->tmpd.p, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h, e->tmpd.p, n * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MDX_OK;
}

template <class R>
static int upload_state_t(mdx_engine *e, const double *x, const double *p, const double *Q,
                          const double *L)
{
    size_t N = e->N;
    int rc;
    if ((rc = upload(e, e->X[e->cur], x, 3 * N)))
        return rc;
    std::vector<double> z;
    if (!p || !L) {
        z.assign(3 * N, 0.0);
    }
    if ((rc = upload_real<R>(e, e->P, p ? p : z.data(), 3 * N)))
        return rc;
    if ((rc = upload_real<R>(e, e->Lm, L ? L : z.data(), 3 * N)))
        return rc;
    std::vector<double> eye;
    if (!Q) {
        eye.assign(9 * N, 0.0);
        for (size_t i = 0; i < N; i++)
            eye[9 * i] = eye[9 * i + 4] = eye[9 * i + 8] = 1.0;
    }
    if ((rc = upload_real<R>(e, e->Q[e->cur], Q ? Q : eye.data(), 9 * N)))
        return rc;
    return MDX_OK;
}

int mdx_upload_state(mdx_engine *e, const double *x, const double *p, const double *Q,
                     const double *L)
{
    if (!e || !x)
        return set_err(MDX_ERR_ARG, "engine and positions required");
    HIPCHK(hipSetDevice(e->device));
    for (long long k = 0; k < 3ll * e->N; k++)
        if (!std::isfinite(x[k]))
            return set_err(MDX_ERR_ARG, "non-finite position at %lld", k / 3);
    int rc = e->prec == MDX_PREC_FP64 ? upload_state_t<double>(e, x, p, Q, L)
                                      : upload_state_t<float>(e, x, p, Q, L);
    e->binned = false;
    e->primed = false;
    return rc;
}

int mdx_download_state(mdx_engine *e, double *x, double *p, double *Q, double *L)
{
    if (!e)
        return set_err(MDX_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(e->device));
    size_t N = e->N;
    int rc = MDX_OK;
    if (x) {
        HIPCHK(hipMemcpyAsync(x, e->X[e->cur].p, 24 * N, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
    }
    bool f64 = e->prec == MDX_PREC_FP64;
    if (p && (rc = f64 ? download_real<double>(e, e->P, p, 3 * N) : download_real<float>(e, e->P, p, 3 * N)))
        return rc;
    if (Q && (rc = f64 ? download_real<double>(e, e->Q[e->cur], Q, 9 * N)
                       : download_real<float>(e, e->Q[e->cur], Q, 9 * N)))
        return rc;
    if (L && (rc = f64 ? download_real<double>(e, e->Lm, L, 3 * N) : download_real<float>(e, e->Lm, L, 3 * N)))
        return rc;
    return MDX_OK;
}

int mdx_download_forces(mdx_engine *e, double *F, double *T)
{
    if (!e)
        return set_err(MDX_ERR_ARG, "null engine");
    if (!e->primed)
        return set_err(MDX_ERR_STATE, "no carried forces (prime or step first)");
    HIPCHK(hipSetDevice(e->device));
    size_t N = e->N;
    bool f64 = e->prec == MDX_PREC_FP64;
    int rc;
    if (F && (rc = f64 ? download_real<double>(e, e->Fc, F, 3 * N) : download_real<float>(e, e->Fc, F, 3 * N)))
        return rc;
    if (T && (rc = f64 ? download_real<double>(e, e->Tc, T, 3 * N) : download_real<float>(e, e->Tc, T, 3 * N)))
        return rc;
    return MDX_OK;
}

int mdx_set_pair_wca(mdx_engine *e, double eps, double sigma, double rcut)
{
    if (!e || !(sigma > 0) || !std::isfinite(eps))
        return set_err(MDX_ERR_ARG, "sigma must be > 0");
    e->wca = 1;
    e->wca_eps = eps;
    e->wca_sigma = sigma;
    e->wca_pl = rcut > 0 ? rcut : 1.12246 * sigma;
    e->grid_dirty = true;
    return MDX_OK;
}

static int patch_rows_upload(mdx_engine *e)
{
    std::vector<PatchRow> rows(e->prow_raw.size() / 3);
    const double PI = 3.14159265358979323846;
    for (size_t k = 0; k < rows.size(); k++) {
        PatchRow &r = rows[k];
        r.eps = e->prow_raw[3 * k];
        r.d = e->prow_raw[3 * k + 1];
        r.thetam = e->prow_raw[3 * k + 2];
        r.costm = std::cos(r.thetam);
        r.kk = PI / (2 * r.thetam);
        r.idist = 1.4 * r.d;
    }
    // a type pair is active when any of its rows in any replica has eps != 0
    e->active_mask = 0;
    for (int ta = 0; ta < e->ntypes; ta++)
        for (int tc = ta; tc < e->ntypes; tc++) {
            int n = e->np[ta] * e->np[tc];
            bool on = false;
            for (int r = 0; r < e->nrep && !on; r++)
                for (int k = 0; k < n; k++)
                    on |= e->prow_raw[3 * ((size_t)r * e->nrows + e->rstart[ta * MAXT + tc] + k)] != 0.0;
            if (on && n > 0)
                e->active_mask |= 1ull << (ta * MAXT + tc);
        }
    e->grid_dirty = true;
    return upload(e, e->prows, rows.data(), rows.size());
}

int mdx_set_patches(mdx_engine *e, int ntypes, const int *type_of_particle,
                    const int *npatch_per_type, const double *dirs, const double *params,
                    double maxd2)
{
    if (!e || ntypes < 1 || ntypes > MAXT || !type_of_particle || !npatch_per_type || !params)
        return set_err(MDX_ERR_ARG, "1..%d types with types, patch counts and params required", MAXT);
    if (!(maxd2 > 0))
        return set_err(MDX_ERR_ARG, "maxd2 must be > 0");
    HIPCHK(hipSetDevice(e->device));
    int ptot = 0;
    for (int t = 0; t < ntypes; t++)
        ptot += npatch_per_type[t] > 0 ? npatch_per_type[t] : 0;
    if (ptot > 0 && !dirs)
        return set_err(MDX_ERR_ARG, "patch directions required");
    e->ntypes = ntypes;
    e->np.assign(npatch_per_type, npatch_per_type + ntypes);
    e->pstart.assign(ntypes, 0);
    int tot = 0;
    for (int t = 0; t < ntypes; t++) {
        if (e->np[t] < 0 || e->np[t] > MAXP)
            return set_err(MDX_ERR_ARG, "type %d: 0..%d patches supported", t, MAXP);
        e->pstart[t] = tot;
        tot += e->np[t];
    }
    e->rstart.assign(MAXT * MAXT, 0);
    int hj = 0;
    for (int i = 0; i < ntypes; i++)
        for (int j = i; j < ntypes; j++) {
            e->rstart[i * MAXT + j] = hj;
            hj += e->np[i] * e->np[j];
        }
    e->nrows = hj;
    for (int k = 0; k < tot; k++) {
        double n2 = dirs[3 * k] * dirs[3 * k] + dirs[3 * k + 1] * dirs[3 * k + 1] +
                    dirs[3 * k + 2] * dirs[3 * k + 2];
        // KernFrenkelOnePatch2 constructor check (potentialtheta.h:415-422)
        if (std::fabs(n2 - 1) > 1E-5)
            return set_err(MDX_ERR_ARG, "patch %d is not a unit vector", k);
    }
    for (int k = 0; k < hj; k++)
        if (!(params[3 * k + 1] > 0) || !(params[3 * k + 2] > 0))
            return set_err(MDX_ERR_ARG, "patch row %d: d and theta_m must be > 0", k);
    e->pdirs.assign(dirs, dirs + 3 * tot);
    e->prow_raw.clear();
    for (int r = 0; r < e->nrep; r++)
        e->prow_raw.insert(e->prow_raw.end(), params, params + 3 * hj);
    e->maxd2 = maxd2;
    e->type.assign(type_of_particle, type_of_particle + e->N);
    std::vector<uint32_t> meta(e->N);
    for (int r = 0; r < e->nrep; r++)
        for (int k = 0; k < e->nper[r]; k++) {
            int i = e->first[r] + k;
            int t = e->type[i];
            if (t < 0 || t >= ntypes)
                return set_err(MDX_ERR_ARG, "particle %d has type %d outside 0..%d", i, t, ntypes - 1);
            meta[i] = ((uint32_t)r << 8) | (uint32_t)t;
        }
    int rc;
    if ((rc = upload(e, e->meta, meta.data(), meta.size())))
        return rc;
    if ((rc = upload(e, e->pdirs_d, e->pdirs.data(), std::max<size_t>(e->pdirs.size(), 3))))
        return rc;
    return patch_rows_upload(e);
}

int mdx_set_replica_patch_params(mdx_engine *e, int replica, const double *params)
{
    if (!e || replica < 0 || replica >= e->nrep || !params || e->ntypes == 0)
        return set_err(MDX_ERR_ARG, "set patches first; replica out of range");
    for (int k = 0; k < e->nrows; k++)
        if (!(params[3 * k + 1] > 0) || !(params[3 * k + 2] > 0))
            return set_err(MDX_ERR_ARG, "patch row %d: d and theta_m must be > 0", k);
    std::copy(params, params + 3 * e->nrows, e->prow_raw.begin() + (size_t)replica * 3 * e->nrows);
    HIPCHK(hipSetDevice(e->device));
    return patch_rows_upload(e);
}

// CSR of per-particle entries from (particle, value) lists
static void csr(int N, const std::vector<std::pair<int, int>> &ent, std::vector<int> &off,
                std::vector<int> &val)
{
    off.assign(N + 1, 0);
    for (auto &p : ent)
        off[p.first + 1]++;
    for (int i = 0; i < N; i++)
        off[i + 1] += off[i];
    val.assign(ent.size(), 0);
    std::vector<int> cur(off.begin(), off.end() - 1);
    for (auto &p : ent)
        val[cur[p.first]++] = p.second;
}

static int check_same_replica(mdx_engine *e, const int *idx, int n, long long row)
{
    int r0 = -1;
    for (int k = 0; k < n; k++) {
        int i = idx[k];
        if (i < 0 || i >= e->N)
            return set_err(MDX_ERR_ARG, "entry %lld: particle %d out of range", row, i);
        int r = (int)(std::upper_bound(e->first.begin(), e->first.end(), i) - e->first.begin()) - 1;
        if (r0 >= 0 && r != r0)
            return set_err(MDX_ERR_ARG, "entry %lld couples different replicas", row);
        r0 = r;
    }
    return MDX_OK;
}

int mdx_set_bonds(mdx_engine *e, long long nbonds, const int *ij, const double *k,
                  const double *r0, int kind)
{
    if (!e || nbonds < 0 || (nbonds > 0 && (!ij || !k || !r0)))
        return set_err(MDX_ERR_ARG, "bond arrays required");
    if (kind != MDX_BOND_HARMONIC && kind != MDX_BOND_FENE)
        return set_err(MDX_ERR_ARG, "unknown bond kind %d", kind);
    HIPCHK(hipSetDevice(e->device));
    std::vector<std::pair<int, int>> ent;
    ent.reserve(2 * nbonds);
    for (long long b = 0; b < nbonds; b++) {
        int rc = check_same_replica(e, ij + 2 * b, 2, b);
        if (rc)
            return rc;
        if (ij[2 * b] == ij[2 * b + 1])
            return set_err(MDX_ERR_ARG, "bond %lld joins a particle to itself", b);
        ent.push_back({ij[2 * b], (int)b});
        ent.push_back({ij[2 * b + 1], (int)b});
    }
    std::vector<int> off, val;
    csr(e->N, ent, off, val);
    // per-entry partner, k, r0 in CSR order
    std::vector<int> bj(val.size());
    std::vector<double> bk(val.size()), br(val.size());
    for (int i = 0; i < e->N; i++)
        for (int q = off[i]; q < off[i + 1]; q++) {
            long long b = val[q];
            bj[q] = ij[2 * b] == i ? ij[2 * b + 1] : ij[2 * b];
            bk[q] = k[b];
            br[q] = r0[b];
        }
    int rc;
    if ((rc = upload(e, e->bond_off, off.data(), off.size())) ||
        (rc = upload(e, e->bond_j, bj.data(), bj.size())) ||
        (rc = upload(e, e->bond_k, bk.data(), bk.size())) ||
        (rc = upload(e, e->bond_r0, br.data(), br.size())))
        return rc;
    e->nbonds = nbonds;
    e->bond_kind = kind;
    return MDX_OK;
}

int mdx_set_angles(mdx_engine *e, long long nangles, const int *ijk, const double *k,
                   const double *theta0, int form)
{
    if (!e || nangles < 0 || (nangles > 0 && (!ijk || !k || !theta0)))
        return set_err(MDX_ERR_ARG, "angle arrays required");
    if (form != MDX_ANGLE_HARMONIC && form != MDX_ANGLE_COSINE)
        return set_err(MDX_ERR_ARG, "unknown angle form %d", form);
    if (nangles > (1ll << 29))
        return set_err(MDX_ERR_ARG, "too many angles");
    HIPCHK(hipSetDevice(e->device));
    std::vector<std::pair<int, int>> ent;
    ent.reserve(3 * nangles);
    for (long long t = 0; t < nangles; t++) {
        int rc = check_same_replica(e, ijk + 3 * t, 3, t);
        if (rc)
            return rc;
        for (int role = 0; role < 3; role++)
            ent.push_back({ijk[3 * t + role], (int)(t * 4 + role)});
    }
    std::vector<int> off, val;
    csr(e->N, ent, off, val);
    int rc;
    if ((rc = upload(e, e->ang_off, off.data(), off.size())) ||
        (rc = upload(e, e->ang_ref, val.data(), val.size())) ||
        (rc = upload(e, e->ang_ijk, ijk, (size_t)3 * nangles)) ||
        (rc = upload(e, e->ang_k, k, (size_t)nangles)) ||
        (rc = upload(e, e->ang_t0, theta0, (size_t)nangles)))
        return rc;
    e->nangles = nangles;
    e->ang_form = form;
    return MDX_OK;
}

int mdx_set_exclusions(mdx_engine *e, long long npairs, const int *ij)
{
    if (!e || npairs < 0 || (npairs > 0 && !ij))
        return set_err(MDX_ERR_ARG, "exclusion pairs required");
    HIPCHK(hipSetDevice(e->device));
    std::vector<std::pair<int, int>> ent;
    for (long long q = 0; q < npairs; q++) {
        int rc = check_same_replica(e, ij + 2 * q, 2, q);
        if (rc)
            return rc;
        ent.push_back({ij[2 * q], ij[2 * q + 1]});
        ent.push_back({ij[2 * q + 1], ij[2 * q]});
    }
    std::vector<int> off, val;
    csr(e->N, ent, off, val);
    for (int i = 0; i < e->N; i++)
        if (off[i + 1] - off[i] > MAXEXCL)
            return set_err(MDX_ERR_ARG, "particle %d has more than %d exclusions", i, MAXEXCL);
    int rc;
    if ((rc = upload(e, e->exc_off, off.data(), off.size())) ||
        (rc = upload(e, e->exc_j, val.data(), std::max<size_t>(val.size(), 1))))
        return rc;
    e->nexcl = npairs;
    return MDX_OK;
}

int mdx_set_active(mdx_engine *e, double fa, const int *prev, const int *next)
{
    if (!e || !prev || !next)
        return set_err(MDX_ERR_ARG, "prev and next arrays required");
    for (int i = 0; i < e->N; i++)
        if (prev[i] >= e->N || next[i] >= e->N)
            return set_err(MDX_ERR_ARG, "active neighbour of %d out of range", i);
    HIPCHK(hipSetDevice(e->device));
    int rc;
    if ((rc = upload(e, e->act_prev, prev, e->N)) || (rc = upload(e, e->act_next, next, e->N)))
        return rc;
    e->fa = fa;
    e->has_active = 1;
    return MDX_OK;
}

int mdx_set_force_guard(mdx_engine *e, double fmax)
{
    if (!e || fmax < 0)
        return set_err(MDX_ERR_ARG, "guard must be >= 0");
    e->fguard = fmax;
    return MDX_OK;
}

int mdx_cell_assign(mdx_engine *e, int replica, const int n[3], int *cell_of)
{
    if (!e || replica < 0 || replica >= e->nrep || !n || !cell_of)
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (n[0] < 1 || n[1] < 1 || n[2] < 1)
        return set_err(MDX_ERR_ARG, "n must be >= 1");
    HIPCHK(hipSetDevice(e->device));
    int cnt = e->nper[replica];
    int rc = dalloc(e, e->tmpi, 4 * (size_t)cnt);
    if (rc)
        return rc;
    const double *b = &e->box[3 * replica];
    hipLaunchKernelGGL(k_assign_custom, dim3(nblk(cnt)), dim3(256), 0, e->stream,
                       (const double *)e->X[e->cur].p, e->first[replica], cnt, b[0], b[1], b[2],
                       n[0], n[1], n[2], (int *)e->tmpi.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(cell_of, e->tmpi.p, 4 * (size_t)cnt, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MDX_OK;
}

int mdx_neighbor_pairs(mdx_engine *e, int replica, const int n[3], double cut, int *ij,
                       long long cap, long long *npairs)
{
    if (!e || replica < 0 || replica >= e->nrep || !n || !npairs || cap < 0 || (cap > 0 && !ij))
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (n[0] < 1 || n[1] < 1 || n[2] < 1 || !(cut > 0))
        return set_err(MDX_ERR_ARG, "n >= 1 and cut > 0 required");
    const double *b = &e->box[3 * replica];
    HIPCHK(hipSetDevice(e->device));
    long long nc = (long long)n[0] * n[1] * n[2];
    if (nc > (1ll << 28))
        return set_err(MDX_ERR_ARG, "too many cells");
    int cnt = e->nper[replica];
    int rc;
    if ((rc = dalloc(e, e->tmpi, 4 * (size_t)cnt)) || (rc = dalloc(e, e->tmpi2, 4 * (size_t)cnt)) ||
        (rc = dalloc(e, e->tmpi3, 4 * (size_t)(nc + 1))) ||
        (rc = dalloc(e, e->tmpll, 8 * (size_t)(cnt + 1))))
        return rc;
    // host-side counting sort of the cell ids (parity path, not the hot loop)
    const double *X = (const double *)e->X[e->cur].p;
    hipLaunchKernelGGL(k_assign_custom, dim3(nblk(cnt)), dim3(256), 0, e->stream, X,
                       e->first[replica], cnt, b[0], b[1], b[2], n[0], n[1], n[2],
                       (int *)e->tmpi.p);
    HIPCHK(hipGetLastError());
    std::vector<int> cellv(cnt), cstart(nc + 1, 0), clist(cnt);
    HIPCHK(hipMemcpyAsync(cellv.data(), e->tmpi.p, 4 * (size_t)cnt, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int k = 0; k < cnt; k++)
        cstart[cellv[k] + 1]++;
    for (long long c = 0; c < nc; c++)
        cstart[c + 1] += cstart[c];
    {
        std::vector<int> fill(cstart.begin(), cstart.end() - 1);
        for (int k = 0; k < cnt; k++)
            clist[fill[cellv[k]]++] = k;
    }
    if ((rc = upload(e, e->tmpi3, cstart.data(), cstart.size())) ||
        (rc = upload(e, e->tmpi2, clist.data(), clist.size())))
        return rc;
    if ((rc = dalloc(e, e->partial, 4 * (size_t)cnt)))
        return rc;
    hipLaunchKernelGGL(k_pairs, dim3(nblk(cnt)), dim3(256), 0, e->stream, X, e->first[replica],
                       cnt, b[0], b[1], b[2], e->periodic, n[0], n[1], n[2], cut,
                       (const int *)e->tmpi.p, (const int *)e->tmpi3.p, (const int *)e->tmpi2.p,
                       (int *)e->partial.p, (const long long *)nullptr, (int *)nullptr, 0ll, 0);
    HIPCHK(hipGetLastError());
    std::vector<int> pc(cnt);
    HIPCHK(hipMemcpyAsync(pc.data(), e->partial.p, 4 * (size_t)cnt, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    std::vector<long long> off(cnt + 1, 0);
    for (int k = 0; k < cnt; k++)
        off[k + 1] = off[k] + pc[k];
    *npairs = off[cnt];
    if (cap > 0) {
        if ((rc = upload(e, e->tmpll, off.data(), off.size())))
            return rc;
        Buf out;
        if ((rc = dalloc(e, out, 8 * (size_t)std::min(cap, off[cnt]) + 8)))
            return rc;
        hipLaunchKernelGGL(k_pairs, dim3(nblk(cnt)), dim3(256), 0, e->stream, X,
                           e->first[replica], cnt, b[0], b[1], b[2], e->periodic, n[0], n[1],
                           n[2], cut, (const int *)e->tmpi.p, (const int *)e->tmpi3.p,
                           (const int *)e->tmpi2.p, (int *)nullptr, (const long long *)e->tmpll.p,
                           (int *)out.p, std::min(cap, off[cnt]), 1);
        hipError_t le = hipGetLastError();
        if (le == hipSuccess)
            le = hipMemcpyAsync(ij, out.p, 8 * (size_t)std::min(cap, off[cnt]),
                                hipMemcpyDeviceToHost, e->stream);
        if (le == hipSuccess)
            le = hipStreamSynchronize(e->stream);
        e->bytes_held -= out.bytes;
        dfree(out);
        if (le != hipSuccess)
            return set_err(MDX_ERR_HIP, "pair listing failed: %s", hipGetErrorString(le));
    }
    if (off[cnt] > cap)
        return set_err(MDX_ERR_CAPACITY, "%lld pairs, capacity %lld", off[cnt], cap);
    return MDX_OK;
}

template <class R> static int compute_forces_t(mdx_engine *e, double *F, double *T)
{
    int rc;
    if ((rc = bin_current<R>(e)))
        return rc;
    if ((rc = dalloc(e, e->Fout, 24 * (size_t)e->N)) || (rc = dalloc(e, e->Tout, 24 * (size_t)e->N)))
        return rc;
    if ((rc = launch_gather<R, M_CONS>(e, nullptr)))
        return rc;
    HIPCHK(hipMemcpyAsync(F, e->Fout.p, 24 * (size_t)e->N, hipMemcpyDeviceToHost, e->stream));
    if (T)
        HIPCHK(hipMemcpyAsync(T, e->Tout.p, 24 * (size_t)e->N, hipMemcpyDeviceToHost, e->stream));
    unsigned int st;
    if ((rc = read_status(e, &st)))
        return rc;
    return status_to_rc(st & 2u);
}

int mdx_compute_forces(mdx_engine *e, double *F, double *T)
{
    if (!e || !F)
        return set_err(MDX_ERR_ARG, "engine and F required");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? compute_forces_t<double>(e, F, T)
                                    : compute_forces_t<float>(e, F, T);
}

int mdx_prime(mdx_engine *e, const double *noise0)
{
    if (!e)
        return set_err(MDX_ERR_ARG, "null engine");
    if (e->integ != MDX_INTEG_ROTOR)
        return set_err(MDX_ERR_STATE, "prime applies to the rotor integrator");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? do_prime<double>(e, noise0) : do_prime<float>(e, noise0);
}

int mdx_step(mdx_engine *e, int nsteps, const double *noise)
{
    if (!e || nsteps < 0)
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (nsteps == 0)
        return MDX_OK;
    HIPCHK(hipSetDevice(e->device));
    if (e->integ == MDX_INTEG_ROTOR)
        return e->prec == MDX_PREC_FP64 ? do_step_rotor<double>(e, nsteps, noise)
                                        : do_step_rotor<float>(e, nsteps, noise);
    return e->prec == MDX_PREC_FP64 ? do_step_bead<double>(e, nsteps, noise)
                                    : do_step_bead<float>(e, nsteps, noise);
}

template <class R> static int observables_t(mdx_engine *e, int replica, double *out)
{
    int rc;
    if (!e->binned && (rc = bin_current<R>(e)))
        return rc;
    if ((rc = dalloc(e, e->obs, 80 * (size_t)e->N)))
        return rc;
    if ((rc = launch_gather<R, M_ENERGY>(e, nullptr)))
        return rc;
    int first = replica < 0 ? 0 : e->first[replica];
    int cnt = replica < 0 ? e->N : e->nper[replica];
    const int nb = std::min(1024u, nblk(cnt));
    if ((rc = dalloc(e, e->partial, 16 * 8 * (size_t)nb)))
        return rc;
    hipLaunchKernelGGL(k_obs<R>, dim3(nb), dim3(256), 0, e->stream, first, cnt,
                       (const R *)e->P.p, e->integ == MDX_INTEG_ROTOR ? (const R *)e->Lm.p : nullptr,
                       (const double *)e->obs.p, 1, e->m, e->I[0], (double *)e->partial.p);
    HIPCHK(hipGetLastError());
    std::vector<double> part(16 * nb);
    HIPCHK(hipMemcpyAsync(part.data(), e->partial.p, 8 * part.size(), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int k = 0; k < 16; k++) {
        double s = 0;
        for (int b = 0; b < nb; b++)
            s += part[16 * b + k];
        out[k] = s;
    }
    for (int k = 0; k < 6; k++)
        out[k] /= (double)cnt;
    return MDX_OK;
}

int mdx_observables(mdx_engine *e, int replica, double *out)
{
    if (!e || !out || replica < -1 || replica >= e->nrep)
        return set_err(MDX_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? observables_t<double>(e, replica, out)
                                    : observables_t<float>(e, replica, out);
}

int mdx_info(mdx_engine *e, double *out)
{
    if (!e || !out)
        return set_err(MDX_ERR_ARG, "bad arguments");
    int rc = build_grid(e);
    if (rc)
        return rc;
    out[0] = e->N;
    out[1] = e->ncells;
    out[2] = e->reps[0].l[0] / e->reps[0].nc[0];
    out[3] = (double)e->bytes_held;
    out[4] = (double)e->steps_taken;
    return MDX_OK;
}

template <class R>
static int profile_t(mdx_engine *e, int nsteps, int cap, const char **names, double *ms, int *nk)
{
    // kernels of one fused rotor/bead step, each bracketed by events
    static const char *kn[] = {"gather", "scan", "fill", "cellsort", "kick_drift"};
    const int K = 5;
    double acc[K] = {0, 0, 0, 0, 0};
    int cntk[K] = {0, 0, 0, 0, 0};
    hipEvent_t ev[8];
    for (int k = 0; k < 8; k++)
        HIPCHK(hipEventCreate(&ev[k]));
    int rc;
    if (e->integ == MDX_INTEG_ROTOR && !e->primed && (rc = do_prime<R>(e, nullptr)))
        return rc;
    if (!e->binned && (rc = bin_current<R>(e)))
        return rc;
    hipLaunchKernelGGL(k_set_call_base, dim3(1), dim3(1), 0, e->stream, (StepCtl *)e->ctl.p);
    if (e->integ == MDX_INTEG_ROTOR) {
        KArgs<R> a = make_args<R>(e);
        HIPCHK(hipEventRecord(ev[0], e->stream));
        hipLaunchKernelGGL(k_kick_drift<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, a);
        HIPCHK(hipEventRecord(ev[1], e->stream));
        HIPCHK(hipEventSynchronize(ev[1]));
        float t;
        HIPCHK(hipEventElapsedTime(&t, ev[0], ev[1]));
        acc[4] += t;
        cntk[4]++;
        e->cur ^= 1;
    }
    const int n = e->ncells;
    const int nb = (n + 1023) / 1024;
    for (int s = 0; s < nsteps; s++) {
        HIPCHK(hipEventRecord(ev[0], e->stream));
        hipLaunchKernelGGL(k_scan1, dim3(nb), dim3(256), 0, e->stream, (const int *)e->count.p,
                           (int *)e->start.p, (int *)e->bsum.p, n);
        hipLaunchKernelGGL(k_scan2, dim3(1), dim3(1024), 0, e->stream, (int *)e->bsum.p, nb,
                           (int *)e->total.p);
        hipLaunchKernelGGL(k_scan3, dim3(nb), dim3(256), 0, e->stream, (int *)e->count.p,
                           (int *)e->start.p, (const int *)e->bsum.p, n, (const int *)e->total.p);
        HIPCHK(hipEventRecord(ev[1], e->stream));
        hipLaunchKernelGGL(k_fill, dim3(nblk(e->N)), dim3(256), 0, e->stream, e->N,
                           (const int *)e->cell.p, (const int *)e->rank.p, (const int *)e->start.p,
                           (int *)e->list.p);
        HIPCHK(hipEventRecord(ev[2], e->stream));
        hipLaunchKernelGGL(k_cellsort, dim3(nblk(n)), dim3(256), 0, e->stream, n,
                           (const int *)e->start.p, (int *)e->list.p);
        HIPCHK(hipEventRecord(ev[3], e->stream));
        if (e->integ == MDX_INTEG_ROTOR) {
            if ((rc = launch_gather<R, M_FUSED>(e, nullptr)))
                return rc;
        } else {
            if ((rc = launch_gather<R, M_BEAD>(e, nullptr)))
                return rc;
        }
        HIPCHK(hipEventRecord(ev[4], e->stream));
        e->cur ^= 1;
        hipLaunchKernelGGL(k_epoch_inc, dim3(1), dim3(1), 0, e->stream, (StepCtl *)e->ctl.p);
        HIPCHK(hipEventSynchronize(ev[4]));
        float t;
        HIPCHK(hipEventElapsedTime(&t, ev[0], ev[1]));
        acc[1] += t;
        HIPCHK(hipEventElapsedTime(&t, ev[1], ev[2]));
        acc[2] += t;
        HIPCHK(hipEventElapsedTime(&t, ev[2], ev[3]));
        acc[3] += t;
        HIPCHK(hipEventElapsedTime(&t, ev[3], ev[4]));
        acc[0] += t;
        cntk[0]++;
        cntk[1]++;
        cntk[2]++;
        cntk[3]++;
    }
    // leave the state at a step boundary: bin, then a plain (non-fused) step
    if ((rc = finish_bins(e)))
        return rc;
    if (e->integ == MDX_INTEG_ROTOR) {
        if ((rc = launch_gather<R, M_STEP>(e, nullptr)))
            return rc;
        hipLaunchKernelGGL(k_epoch_inc, dim3(1), dim3(1), 0, e->stream, (StepCtl *)e->ctl.p);
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    e->steps_taken += nsteps + (e->integ == MDX_INTEG_ROTOR ? 1 : 0);
    for (int k = 0; k < 8; k++)
        (void)hipEventDestroy(ev[k]);
    int w = 0;
    for (int k = 0; k < K && w < cap; k++) {
        if (!cntk[k])
            continue;
        names[w] = kn[k];
        ms[w] = acc[k] / cntk[k];
        w++;
    }
    *nk = w;
    unsigned int st;
    if ((rc = read_status(e, &st)))
        return rc;
    return status_to_rc(st);
}

int mdx_profile_step(mdx_engine *e, int nsteps, int cap, const char **names, double *ms, int *nk)
{
    if (!e || nsteps < 1 || !names || !ms || !nk)
        return set_err(MDX_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? profile_t<double>(e, nsteps, cap, names, ms, nk)
                                    : profile_t<float>(e, nsteps, cap, names, ms, nk);
}
