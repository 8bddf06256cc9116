// mdx_kernels.h -- gfx950 kernels of the Langevin step.
//
// Storage order.  Particles live in "slots".  Every kernel that produces the
// next state is launched with one thread per slot of the *cell list* order:
// thread s takes the particle in old slot list[s], and writes its next state
// to slot s of the other buffer.  The state is therefore re-sorted into cell
// order every step without an extra pass, so that the threads of a wavefront
// handle particles of the same or adjacent cells (shared neighbour cells,
// coalesced gathers).  Each particle carries its original index (id): noise
// counters, the p1 < p2 roles of the patch loop, topology and the host
// interface all use it.
//
// Reference paths in comments are relative to the reference's Code/ folder.
#pragma once

#include "mdx_device.h"
#include "../../include/mdx.h"

namespace mdx {

// one position record: FP64 position + original index + (replica << 8 | type)
struct __align__(16) PosRec {
    double x, y, z;
    uint32_t id;
    uint32_t meta;
};

struct DevRep {
    double l[3], l2s[3];
    int nc[3];
    int coff;
    int first, count;
    double Rt, Rr;   // sqrt(2 gamma kT / dt), sqrt(2 gamma_r kT / dt) (LangevinR.cpp:51-52)
    double bead_r;   // sqrt(0.5 kT gamma m dt) (Langevin.cpp:44)
    int prow;        // first patch row of this replica
};

struct StepCtl {
    unsigned long long epoch;     // force evaluations so far (noise counter)
    unsigned long long call_base; // epoch at the start of the current call
    unsigned int status;          // bit0 force guard, bit1 non-finite, bit2 candidate capacity,
                                  // bit3 binding capacity
    unsigned int pad;
    // neighbour list: rebuild flag, displacement bound since the last build
    unsigned int rebuild; // 1: the next step rebuilds the list before its walk
    unsigned int nbuilds; // list builds so far
    double D;             // sum over steps of the max drift since the last build
    // a walk queued more than pcap patch candidates for some position (the
    // overflow kernel's gate; cleared by the integrator kernel after it)
    unsigned int ovf;
    // the first force evaluation (epoch) that raised a status bit; ~0 while
    // the status is clear.  From then on no moving kernel advances a particle.
    unsigned long long fail_epoch;
    // 1: a list build has written nbr since the compacted walk last laid its
    // entries out per block (the next compacted walk does so); cleared by the
    // kernel that closes the list step (k_decide, k_step_end)
    unsigned int fresh;
};

enum { M_CONS = 0, M_PRIME = 1, M_STEP = 2, M_FUSED = 3, M_BEAD = 4, M_ENERGY = 5 };

template <class R> struct KArgs {
    int N;
    int rotor;
    int periodic;
    const DevRep *reps;
    // state: current (read) and next (written) buffers
    const PosRec *X;
    PosRec *Xn;
    const R *Q;
    R *Qn;
    const R *p, *L, *F, *T;
    R *pn, *Ln, *Fn, *Tn;
    const int *slot; // slot of each original index
    int *slotn;
    // cell lists
    const int *cell; // cell of each current slot
    int *celln;      // cell of each next slot
    int *rank, *count;
    const int *start, *list;
    // outputs by original index
    double *Fout, *Tout, *obs;
    // split walk -> finish: WCA force and patch candidates per cell-list position
    R *Fw;
    int *ncand, *cand; // cand[q * N + s], q < pcap
    // patch candidates held per position (PMAXG, or kmax in list steps,
    // where a position's candidates are among its list entries);
    // ovf_fatal: no overflow kernel follows the walk -- a position with more
    // candidates than that (its list overflowed too) raises status bit 2
    int pcap, ovf_fatal;
    R *Tp;             // patch torque per position (compacted patch kernel)
    int nrep; // replicas (1: the replica record is loaded uniformly)
    // neighbour list (lmode): slot s walks nbr[k * N + s], k < nn[s] (<= kmax),
    // else the 27 cells of its cell at the last build, bcell[s].  nbrc: the
    // same entries per block of 256 slots, back to back in slot order from
    // nbrc[first slot * kmax] (slots over kmax contribute none), laid out by
    // the first compacted walk after each build
    int lmode;
    const int *nbr, *nbrc;
    int *nbrw, *nbrcw, *nn, *bcell;
    int kmax;
    // list radii squared: interaction reach + skin; WCA reach + skin for the
    // pairs that cannot interact through patches
    double rl2, rl2w, skin;
    // drift maxima of the moving kernels' blocks (float bits, atomic max into
    // slot block % NDSLOT); k_decide reduces and clears them
    unsigned int *dslot;
    // WCA (HSPotential)
    int wca;
    double wca_eps, wca_s6, wca_s12, wca_pl, wca_pl2, wca_pl2lo;
    // patches
    int has_patches;
    int np[MAXT];
    int pstart[MAXT];
    int rstart[MAXT * MAXT];
    unsigned long long active_mask; // bit ta * MAXT + tc: type pair with a nonzero row
    const double *pdirs;
    const float *pdirs_f; // the same directions in FP32 (the own-patch screen)
    const PatchRow *prows;
    double maxd2, rmax2, costm_screen;
    double costm_screen_f; // the own-patch screen's threshold (FP32 margin)
    // topology (indexed by original index)
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
    double gm_m, gr_I[3]; // -gamma / m, -gammar / I[k]
    double b_c5c2, b_c5c3q, b_c2;
    // noise
    uint32_t key0, key1;
    const double *noise; // injected rows (ncomp per particle, original order) or NULL
    StepCtl *ctl;
    double fguard;
    // one-bond-per-patch binding: the bound-pair force and torque (lab) by
    // slot, added by the integrator where the position's binding candidate
    // count (bncand) is non-zero (null: no binding)
    const double *bF, *bT;
    const int *bncand;
};

// ------------------------------------------------------- orientation store
// FP64: the 3 x 3 lab -> body matrix (9 doubles per particle, the reference's
// layout, LangevinR.h:67-68).  Mixed: a unit quaternion, 4 floats (16 B)
// per particle, the matrix rebuilt on load (quat_to_mat).
template <class R> struct Orient {
    static constexpr int W = 9;
};
template <> struct Orient<float> {
    static constexpr int W = 4;
};

template <class R> __device__ __forceinline__ void load_orient(const R *Q, size_t i, R m[9])
{
    if constexpr (sizeof(R) == 8) {
#pragma unroll
        for (int k = 0; k < 9; k++)
            m[k] = Q[9 * i + k];
    } else {
        const float4 q = reinterpret_cast<const float4 *>(Q)[i];
        quat_to_mat<float>(q.x, q.y, q.z, q.w, m);
    }
}

// store the rotated matrix m (mixed: as its normalised quaternion)
template <class R> __device__ __forceinline__ void store_orient(R *Q, size_t i, const R m[9])
{
    if constexpr (sizeof(R) == 8) {
#pragma unroll
        for (int k = 0; k < 9; k++)
            Q[9 * i + k] = m[k];
    } else {
        float q[4];
        mat_to_quat<float>(m, q);
        reinterpret_cast<float4 *>(Q)[i] = make_float4(q[0], q[1], q[2], q[3]);
    }
}

// an orientation that did not rotate, copied as stored (no round trip)
template <class R> __device__ __forceinline__ void copy_orient(R *Qn, size_t s, const R *Q, size_t i)
{
    if constexpr (sizeof(R) == 8) {
#pragma unroll
        for (int k = 0; k < 9; k++)
            Qn[9 * s + k] = Q[9 * i + k];
    } else {
        reinterpret_cast<float4 *>(Qn)[s] = reinterpret_cast<const float4 *>(Q)[i];
    }
}

// ------------------------------------------------------------ small helpers
// n = Q^T n_body in FP64 with the reference's operation order
// (potentialtheta.h:559-565); Q promoted from R
template <class R> __device__ __forceinline__ void lab_normal_exact(const R *q, const double *nb,
                                                                    double n[3])
{
    n[0] = nb[0] * (double)q[0] + nb[1] * (double)q[3] + nb[2] * (double)q[6];
    n[1] = nb[0] * (double)q[1] + nb[1] * (double)q[4] + nb[2] * (double)q[7];
    n[2] = nb[0] * (double)q[2] + nb[1] * (double)q[5] + nb[2] * (double)q[8];
}

// LangevinNVTR::genfullmat (LangevinR.cpp:392-461)
// FP64 with the reference's divisions; FP32 with one hardware reciprocal
// (v_rcp_f32, 1 ulp) per axis in place of its nine divisions
template <class R>
__device__ __forceinline__ void genfullmat(R dt, const R I[3], const R L[3], R up[9])
{
    R fac1, fac2, fac3, fac4, fac5, fac6;
    if constexpr (sizeof(R) == 8) {
        R tempx = dt * L[0] / I[0];
        R tempy = dt * L[1] / I[1];
        R tempz = dt * L[2] / I[2];
        R qx = tempx / (R)4., qy = tempy / (R)4., hz = tempz / (R)2.;
        fac1 = ((R)1 - qx * qx) / ((R)1 + qx * qx);
        fac2 = tempx / ((R)2. * ((R)1 + qx * qx));
        fac3 = ((R)1 - qy * qy) / ((R)1 + qy * qy);
        fac4 = tempy / ((R)2. * ((R)1 + qy * qy));
        fac5 = ((R)1 - hz * hz) / ((R)1 + hz * hz);
        fac6 = (tempz) / ((R)1 + hz * hz);
    } else {
        const float tempx = dt * L[0] * __builtin_amdgcn_rcpf(I[0]);
        const float tempy = dt * L[1] * __builtin_amdgcn_rcpf(I[1]);
        const float tempz = dt * L[2] * __builtin_amdgcn_rcpf(I[2]);
        const float qx = tempx * 0.25f, qy = tempy * 0.25f, hz = tempz * 0.5f;
        const float ix = __builtin_amdgcn_rcpf(1.f + qx * qx);
        const float iy = __builtin_amdgcn_rcpf(1.f + qy * qy);
        const float iz = __builtin_amdgcn_rcpf(1.f + hz * hz);
        fac1 = (1.f - qx * qx) * ix;
        fac2 = tempx * 0.5f * ix;
        fac3 = (1.f - qy * qy) * iy;
        fac4 = tempy * 0.5f * iy;
        fac5 = (1.f - hz * hz) * iz;
        fac6 = tempz * iz;
    }
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

// status bits: 1 a pair or bond term above the force guard, 2 a non-finite
// force, torque or updated state; the first epoch that raised one is kept
template <class R> __device__ __forceinline__ void raise_status(const KArgs<R> &a, unsigned bit)
{
    atomicOr(&a.ctl->status, bit);
    atomicMin(&a.ctl->fail_epoch, a.ctl->epoch);
}

// |f| of one term above the guard (the reference prints "force too big" and
// pauses: MD.cpp:1711-1749)
template <class R> __device__ __forceinline__ bool flag_force(const KArgs<R> &a, double f)
{
    if (a.fguard > 0 && fabs(f) > a.fguard) {
        raise_status(a, 1u);
        return true;
    }
    return false;
}

// a status bit is raised (by this kernel so far or an earlier one): the run
// stops, and moving kernels keep every particle's state from here on
template <class R> __device__ __forceinline__ bool stopped(const KArgs<R> &a)
{
    return __hip_atomic_load(&a.ctl->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}

// noise of particle (original index) o for the current force evaluation
template <class R>
__device__ __forceinline__ void load_noise(const KArgs<R> &a, uint32_t o, int ncomp, R out[6])
{
    unsigned long long ep = a.ctl->epoch;
    if (a.noise) {
        const double *row = a.noise + (size_t)(ep - a.ctl->call_base) * (size_t)a.N * ncomp +
                            (size_t)o * ncomp;
        for (int k = 0; k < ncomp; k++)
            out[k] = (R)row[k];
    } else {
        U4 c0 = {o, (uint32_t)ep, (uint32_t)(ep >> 32), 0u};
        U4 r0 = philox4x32_10(c0, a.key0, a.key1);
        out[0] = unit_uniform<R>(r0.a);
        out[1] = unit_uniform<R>(r0.b);
        out[2] = unit_uniform<R>(r0.c);
        if (ncomp > 3) {
            out[3] = unit_uniform<R>(r0.d);
            U4 c1 = {o, (uint32_t)ep, (uint32_t)(ep >> 32), 1u};
            U4 r1 = philox4x32_10(c1, a.key0, a.key1);
            out[4] = unit_uniform<R>(r1.a);
            out[5] = unit_uniform<R>(r1.b);
        }
    }
}

// cube::assign_box with the engine's grid; count the particle into its cell
template <class R>
__device__ __forceinline__ void bin_particle(const KArgs<R> &a, int s, const DevRep &rp,
                                             const double x[3], int *cellout)
{
    int cx = axis_cell(x[0], rp.nc[0], rp.l[0]);
    int cy = axis_cell(x[1], rp.nc[1], rp.l[1]);
    int cz = axis_cell(x[2], rp.nc[2], rp.l[2]);
    int c = rp.coff + cx + rp.nc[0] * (cy + rp.nc[1] * cz);
    cellout[s] = c;
    a.rank[s] = atomicAdd(&a.count[c], 1);
}

template <class R>
__device__ __forceinline__ void min_image(const KArgs<R> &a, const DevRep &rp, double x0,
                                          double x1, double x2, const PosRec &b, double d[3])
{
    d[0] = mimage(x0 - b.x, rp.l[0], rp.l2s[0], a.periodic);
    d[1] = mimage(x1 - b.y, rp.l[1], rp.l2s[1], a.periodic);
    d[2] = mimage(x2 - b.z, rp.l[2], rp.l2s[2], a.periodic);
}

// The raw difference x_i - x_j is the minimum image for every pair the list
// (or the cells of its build) can hold when the box has walls, or when x_i
// lies farther than two cells plus the skin from every face: such a pair was
// in adjacent cells at the build (under two cells apart per axis) and each
// side has moved less than half the skin since.  (min_image would return the
// same value; this skips its tests.)
template <class R>
__device__ __forceinline__ bool face_clear(const KArgs<R> &a, const DevRep &rp, const PosRec &x)
{
    if (!a.periodic)
        return true;
    const double xs[3] = {x.x, x.y, x.z};
    bool ok = true;
    for (int k = 0; k < 3; k++) {
        const double m = 2.0 * rp.l[k] / rp.nc[k] + a.skin;
        ok = ok && xs[k] > m && xs[k] < rp.l[k] - m;
    }
    return ok;
}

// ------------------------------------------------------------- gather kernel
// the walk screens patch candidates (cone screens) before handing them over
#ifndef MDX_WALK_SCREEN
#define MDX_WALK_SCREEN 1
#endif
// neighbour-list entries loaded together per round of the list walk
#ifndef MDX_LBATCH
#define MDX_LBATCH 4
#endif
constexpr int PMAX = 8;   // deferred patch candidates per particle held in LDS
constexpr int PMAXG = 16; // patch candidates per particle handed from walk to finish
enum { PH_ALL = 0, PH_WALK = 1, PH_FINISH = 2, PH_INTEGRATE = 3, PH_INTEGRATE_F = 4 };

// Workgroups go to the 8 XCDs round-robin; this gives each XCD one contiguous
// range of blocks (of cell-ordered slots), so the neighbours a block gathers
// are mostly in its own XCD's L2.  A bijection on [0, G).
__device__ __forceinline__ int xcd_block(int b, int G)
{
    const int q = G >> 3, r = G & 7, x = b & 7;
    return x * q + min(x, r) + (b >> 3);
}

// the replica record of a particle; one replica: a uniform load
template <class R> __device__ __forceinline__ DevRep rep_rec(const KArgs<R> &a, uint32_t meta)
{
    return a.nrep == 1 ? a.reps[0] : a.reps[meta >> 8];
}

// neighbour slots of slot i: the 9 (y, z) rows of the 27-cell stencil; the x
// neighbours of a row are contiguous in the cell list (one range, or two at
// the periodic seam; fewer cells when a side has fewer than 3)
template <class R, class Fn>
__device__ __forceinline__ void for_each_in_cells(const KArgs<R> &a, const DevRep &rp, int cabs, int i,
                                                  Fn &&fn)
{
    const int c = cabs - rp.coff;
    const int n0 = rp.nc[0], n1 = rp.nc[1], n2 = rp.nc[2];
    const int cx = c % n0, cy = (c / n0) % n1, cz = c / (n0 * n1);
    const int ylo = n1 >= 3 ? -1 : 0, yhi = n1 >= 2 ? 1 : 0;
    const int zlo = n2 >= 3 ? -1 : 0, zhi = n2 >= 2 ? 1 : 0;
    for (int oz = zlo; oz <= zhi; oz++) {
        int z = cz + oz;
        z = z < 0 ? z + n2 : (z >= n2 ? z - n2 : z);
        for (int oy = ylo; oy <= yhi; oy++) {
            int y = cy + oy;
            y = y < 0 ? y + n1 : (y >= n1 ? y - n1 : y);
            const int rb = rp.coff + n0 * (y + n1 * z);
            int r0, r1, r2 = 0, r3 = 0;
            if (n0 >= 3) {
                if (cx == 0) {
                    r0 = a.start[rb];
                    r1 = a.start[rb + 2];
                    r2 = a.start[rb + n0 - 1];
                    r3 = a.start[rb + n0];
                } else if (cx == n0 - 1) {
                    r0 = a.start[rb + n0 - 2];
                    r1 = a.start[rb + n0];
                    r2 = a.start[rb];
                    r3 = a.start[rb + 1];
                } else {
                    r0 = a.start[rb + cx - 1];
                    r1 = a.start[rb + cx + 2];
                }
            } else {
                r0 = a.start[rb];
                r1 = a.start[rb + n0];
            }
            for (int k = r0; k < r1; k++) {
                const int j = a.list[k];
                if (j != i)
                    fn(a.X[j], j);
            }
            for (int k = r2; k < r3; k++) {
                const int j = a.list[k];
                if (j != i)
                    fn(a.X[j], j);
            }
        }
    }
}

template <class R, class Fn>
__device__ __forceinline__ void for_each_neighbor(const KArgs<R> &a, const DevRep &rp, int i, Fn &&fn)
{
    for_each_in_cells(a, rp, a.cell[i], i, fn);
}

// nn[i]: entry count, and LIST_EXC when the list holds a pair without WCA
// (an exclusion kept for its patch interaction); lists without that flag hold
// no excluded pair, so their walk needs no exclusion test
constexpr int LIST_EXC = 1 << 30;
template <class R> __device__ __forceinline__ int list_count(const KArgs<R> &a, int i)
{
    return a.nn[i] & (LIST_EXC - 1);
}
// exclusions must be tested for slot i's pairs in this walk
template <class R> __device__ __forceinline__ bool needs_exclusions(const KArgs<R> &a, int i)
{
    if (!a.exc_off)
        return false;
    if (!a.lmode)
        return true;
    const int v = a.nn[i];
    return (v & LIST_EXC) || (v & (LIST_EXC - 1)) > a.kmax;
}

// neighbours of slot i from the list (order of the build's cell walk), or,
// for a particle with more than kmax listed, the 27 cells of the build.  Both
// are supersets of the pairs within reach while the displacement bound holds.
template <class R, class Fn>
__device__ __forceinline__ void for_each_listed(const KArgs<R> &a, const DevRep &rp, int i, Fn &&fn)
{
    // the first MDX_LBATCH entries are loaded together with the count (the
    // list has at least that many slots per particle; entries beyond the
    // count are stale slot numbers of an earlier build, never used)
    int j0[MDX_LBATCH];
#pragma unroll
    for (int u = 0; u < MDX_LBATCH; u++)
        j0[u] = a.nbr[(size_t)u * a.N + i];
    const int n = list_count(a, i);
    if (n <= a.kmax) {
        // MDX_LBATCH entries per round: their index and position loads overlap
        for (int k0 = 0; k0 < n; k0 += MDX_LBATCH) {
            int j[MDX_LBATCH];
            PosRec x[MDX_LBATCH];
#pragma unroll
            for (int u = 0; u < MDX_LBATCH; u++)
                j[u] = k0 + u < n ? (k0 == 0 ? j0[u] : a.nbr[(size_t)(k0 + u) * a.N + i]) : i;
#pragma unroll
            for (int u = 0; u < MDX_LBATCH; u++)
                x[u] = a.X[j[u]];
#pragma unroll
            for (int u = 0; u < MDX_LBATCH; u++)
                if (k0 + u < n)
                    fn(x[u], j[u]);
        }
    } else {
        for_each_in_cells(a, rp, a.bcell[i], i, fn);
    }
}

template <class R, class Fn>
__device__ __forceinline__ void for_each_near(const KArgs<R> &a, const DevRep &rp, int i, Fn &&fn)
{
    if (a.lmode)
        for_each_listed(a, rp, i, fn);
    else
        for_each_neighbor(a, rp, i, fn);
}

// block max of the drift into one of NDSLOT slots (atomic max on the float
// bits, which order like the values for drifts >= 0; a non-finite drift
// counts as infinite, so it forces a rebuild); k_decide reduces the slots
constexpr int NDSLOT = 64;
template <class R> __device__ __forceinline__ void drift_bound(const KArgs<R> &a, float d)
{
    __shared__ float wmax[4];
    for (int o = 32; o > 0; o >>= 1)
        d = fmaxf(d, __shfl_xor(d, o));
    if ((threadIdx.x & 63) == 0)
        wmax[threadIdx.x >> 6] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = 0.f;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++)
            m = fmaxf(m, wmax[w]);
        if (!(m <= 3.0e38f))
            m = __int_as_float(0x7f800000);
        atomicMax(&a.dslot[blockIdx.x & (NDSLOT - 1)], __float_as_uint(m));
    }
}

template <class R> struct Acc {
    R F[3], T[3];
    double epot, vir[9];
    bool bad; // a term of this particle was above the force guard
};

// roles of a patch pair: "a" = lower original index, exchanged when its type
// is higher (GeneralPatch rows exist for ta <= tc; the potential is symmetric)
__device__ __forceinline__ bool patch_roles(uint32_t oi, int ti, uint32_t oj, int tj, int &ta,
                                            int &tc)
{
    bool i_is_a = oi < oj;
    ta = i_is_a ? ti : tj;
    tc = i_is_a ? tj : ti;
    if (ta > tc) {
        int t = ta;
        ta = tc;
        tc = t;
        i_is_a = !i_is_a;
    }
    return i_is_a;
}

// all patch pairs between particle i and neighbour slot j
// (LangevinR.cpp:923-1044, potentialtheta.h:517-750)
template <class R, int MODE>
__device__ __forceinline__ void patch_pair(const KArgs<R> &a, const DevRep &rp, const PosRec &xi,
                                           int ti, const R *qi, int j, Acc<R> &acc)
{
    constexpr bool FP64 = sizeof(R) == 8;
    const PosRec xj = a.X[j];
    double d[3];
    min_image(a, rp, xi.x, xi.y, xi.z, xj, d);
    const double r2s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double rij = sqrt(r2s);
    const int tj = xj.meta & 255;
    int ta, tc;
    const bool i_is_a = patch_roles(xi.id, ti, xj.id, tj, ta, tc);
    const double sg = i_is_a ? 1.0 : -1.0;
    const double u[3] = {sg * d[0] / rij, sg * d[1] / rij, sg * d[2] / rij};
    R qj[9];
    load_orient(a.Q, (size_t)j, qj);
    const R *qa = i_is_a ? qi : qj;
    const R *qc = i_is_a ? qj : qi;
    // body-frame components of u screen the cone tests; the decisions are
    // then taken on the reference's own FP64 expressions
    const R ur[3] = {(R)u[0], (R)u[1], (R)u[2]};
    const R va[3] = {qa[0] * ur[0] + qa[1] * ur[1] + qa[2] * ur[2],
                     qa[3] * ur[0] + qa[4] * ur[1] + qa[5] * ur[2],
                     qa[6] * ur[0] + qa[7] * ur[1] + qa[8] * ur[2]};
    const R vc[3] = {qc[0] * ur[0] + qc[1] * ur[1] + qc[2] * ur[2],
                     qc[3] * ur[0] + qc[4] * ur[1] + qc[5] * ur[2],
                     qc[6] * ur[0] + qc[7] * ur[1] + qc[8] * ur[2]};
    const int npa = a.np[ta], npc = a.np[tc];
    const double *da = a.pdirs + 3 * a.pstart[ta];
    const double *dc = a.pdirs + 3 * a.pstart[tc];
    const int rowbase = rp.prow + a.rstart[ta * MAXT + tc];
    const R scr = (R)a.costm_screen;
    for (int ia = 0; ia < npa; ia++) {
        const R s1 = -((R)da[3 * ia] * va[0] + (R)da[3 * ia + 1] * va[1] + (R)da[3 * ia + 2] * va[2]);
        if (!(s1 > scr))
            continue;
        double n1[3];
        lab_normal_exact(qa, da + 3 * ia, n1);
        const double a1 = -(n1[0] * u[0] + n1[1] * u[1] + n1[2] * u[2]);
        for (int ib = 0; ib < npc; ib++) {
            const R s2 = (R)dc[3 * ib] * vc[0] + (R)dc[3 * ib + 1] * vc[1] + (R)dc[3 * ib + 2] * vc[2];
            if (!(s2 > scr))
                continue;
            const PatchRow pr = a.prows[rowbase + ia * npc + ib];
            if (rij > pr.idist || !(a1 > pr.costm))
                continue;
            double n2[3];
            lab_normal_exact(qc, dc + 3 * ib, n2);
            const double a2 = (n2[0] * u[0] + n2[1] * u[1] + n2[2] * u[2]);
            if (!(a2 > pr.costm))
                continue;
            R f[3] = {0, 0, 0}, t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
            double e;
            if constexpr (FP64) {
                e = kf2_pair_fp64(u, rij, n1, n2, a1, a2, pr, f, t1, t2);
            } else {
                const float uf[3] = {(float)u[0], (float)u[1], (float)u[2]};
                const float n1f[3] = {(float)n1[0], (float)n1[1], (float)n1[2]};
                const float n2f[3] = {(float)n2[0], (float)n2[1], (float)n2[2]};
                e = kf2_pair_fp32(uf, (float)rij, n1f, n2f, (float)a1, (float)a2,
                                  fabs(a1 * a1 - 1) < 1E-5, fabs(a2 * a2 - 1) < 1E-5,
                                  (float)pr.eps, (float)pr.d, (float)pr.kk, f, t1, t2);
            }
            if (i_is_a) {
                for (int q = 0; q < 3; q++) {
                    acc.F[q] += f[q];
                    acc.T[q] += t1[q];
                }
            } else {
                for (int q = 0; q < 3; q++) {
                    acc.F[q] -= f[q];
                    acc.T[q] += t2[q];
                }
            }
            if (MODE == M_ENERGY) {
                acc.epot += 0.5 * e;
                for (int q = 0; q < 3; q++)
                    for (int l = 0; l < 3; l++)
                        acc.vir[3 * q + l] += 0.5 * d[l] * sg * (double)f[q];
            }
        }
    }
}

// WCA term of one pair (MD.cpp:1700-1758, potential.h:112-125) into acc:
// d = x_i - x_j minimum-imaged, r2s = |d|^2, exclusions [exc0, exc1) of i
template <class R, int MODE>
__device__ __forceinline__ void wca_term(const KArgs<R> &a, const double d[3], double r2s,
                                         uint32_t idj, int exc0, int exc1, Acc<R> &acc)
{
    constexpr bool FP64 = sizeof(R) == 8;
    // ---------------- WCA (MD.cpp:1700-1758, potential.h:112-125)
    // (r^2 against the squared cut-off with a margin first; the
    // decision itself is the reference's sqrt(r^2) < cut-off)
    if (a.wca && r2s < a.wca_pl2) {
        // FP64: the reference's sqrt(r^2) < cut-off; mixed: the same
        // decision taken on r^2 outside a 1e-12 band around the
        // squared cut-off (where the two cannot differ), sqrt inside it
        double rij = 0;
        bool inside;
        if constexpr (FP64) {
            rij = sqrt(r2s);
            inside = rij < a.wca_pl;
        } else {
            inside = r2s < a.wca_pl2lo || sqrt(r2s) < a.wca_pl;
        }
        if (inside) {
            bool ex = false;
            for (int e = exc0; e < exc1; e++)
                ex |= (uint32_t)a.exc_j[e] == idj;
            if (!ex) {
                if constexpr (FP64) {
                    const double f1 = hs_force(rij, a.wca_eps, a.wca_s6, a.wca_s12);
                    acc.bad |= flag_force(a, f1);
                    for (int q = 0; q < 3; q++) {
                        const double fac = f1 * d[q] / rij;
                        acc.F[q] += fac;
                        if (MODE == M_ENERGY)
                            for (int l = 0; l < 3; l++)
                                acc.vir[3 * q + l] += 0.5 * d[l] * fac;
                    }
                    if (MODE == M_ENERGY)
                        acc.epot += 0.5 * wca_energy(rij, a.wca_eps, a.wca_s6, a.wca_s12);
                } else {
                    // 4 eps (12 s12 / r^13 - 6 s6 / r^7) from 1/r
                    const float r2f = (float)r2s;
                    const float ir = rsqrtf(r2f);
                    const float rf = r2f * ir;
                    const float ir2 = ir * ir;
                    const float ir6 = ir2 * ir2 * ir2;
                    const float f1 = 4.f * (float)a.wca_eps * ir * ir6 *
                                     (12.f * (float)a.wca_s12 * ir6 - 6.f * (float)a.wca_s6);
                    acc.bad |= flag_force(a, f1);
                    const float fr = f1 * ir;
                    for (int q = 0; q < 3; q++) {
                        const float fac = fr * (float)d[q];
                        acc.F[q] += fac;
                        if (MODE == M_ENERGY)
                            for (int l = 0; l < 3; l++)
                                acc.vir[3 * q + l] += 0.5 * d[l] * fac;
                    }
                    if (MODE == M_ENERGY)
                        acc.epot += 0.5 * (double)wca_energy(rf, (float)a.wca_eps,
                                                             (float)a.wca_s6,
                                                             (float)a.wca_s12);
                }
            }
        }
    }
}

// The cone screen of patch_pair on the particle's own patches alone (no
// load of the partner), in FP32 for both precisions: false when none of its
// patches passes, in which case patch_pair adds exactly nothing for the pair.
// Its error (~1e-7 with the orientation rounded to FP32 and a reciprocal
// square root) is far inside the screen's margin below cos(theta_m) (1e-5),
// so a pair it rejects fails patch_pair's exact FP64 cone test as well.  The
// walk uses it to hand over only candidates that can bind; every walk form
// and the overflow path use this same screen, so they agree on the candidates.
template <class R>
__device__ __forceinline__ bool own_screen(const KArgs<R> &a, uint32_t oi, int ti, const float *qi,
                                           const PosRec &xj, const double d[3], double r2s)
{
    const int tj = xj.meta & 255;
    int ta, tc;
    const bool i_is_a = patch_roles(oi, ti, xj.id, tj, ta, tc);
    // u = (d / r), sign folded into the test below
    const float ri = rsqrtf((float)r2s);
    const float ur[3] = {(float)d[0] * ri, (float)d[1] * ri, (float)d[2] * ri};
    const float v[3] = {qi[0] * ur[0] + qi[1] * ur[1] + qi[2] * ur[2],
                        qi[3] * ur[0] + qi[4] * ur[1] + qi[5] * ur[2],
                        qi[6] * ur[0] + qi[7] * ur[1] + qi[8] * ur[2]};
    const int t = i_is_a ? ta : tc;
    const int np = a.np[t];
    const float *dd = a.pdirs_f + 3 * a.pstart[t];
    const float scr = (float)a.costm_screen_f;
    bool pass = false;
    for (int k = 0; k < np; k++) {
        const float sd = dd[3 * k] * v[0] + dd[3 * k + 1] * v[1] + dd[3 * k + 2] * v[2];
        // u = +d/r for role a, whose test is -(n . u); u = -d/r for role c,
        // whose test is n . u: both are -(n . d/r)
        pass |= -sd > scr;
    }
    return pass;
}

// the moved position of slot s into the next buffer (the whole record: a
// write of x, y, z alone measured no faster)
template <class R>
__device__ __forceinline__ void store_position(const KArgs<R> &a, int s, const PosRec &xi,
                                               const double xn[3])
{
    PosRec o = xi;
    o.x = xn[0];
    o.y = xn[1];
    o.z = xn[2];
    a.Xn[s] = o;
}

// Force (lab) and patch torque (lab) of one particle, then the mode's
// integrator epilogue.  Full neighbour list: every pair is evaluated from
// both sides, each side keeps its own share; no atomics.  The neighbour loop
// evaluates WCA and queues patch candidates (pairs within max_distance_check
// whose types carry patch rows); the patch pairs are evaluated after the loop
// (fewer live registers in the loop).
// PH_ALL: walk, patches and integrator in one kernel; PH_WALK: the walk
// only (WCA force and patch candidates to global memory); PH_FINISH: the
// patches from those candidates, topology and the integrator.  The split
// keeps the register-heavy patch code out of the neighbour walk.
// PH_INTEGRATE_F: as PH_INTEGRATE with the force given in fin and no patch
// torque (a particle without patch candidates, integrated by the walk)
template <class R, int MODE, int PH, class Trav>
__device__ __forceinline__ void process_particle(const KArgs<R> &a, int s, int *plist, Trav &&trav,
                                                 float &drift, const R *fin = nullptr)
{
    constexpr bool FP64 = sizeof(R) == 8;
    const int i = a.lmode ? s : a.list[s];
    const PosRec xi = a.X[i];
    const uint32_t oi = xi.id;
    const int ti = xi.meta & 255;
    const DevRep rp = rep_rec(a, xi.meta);
    Acc<R> acc;
    acc.F[0] = acc.F[1] = acc.F[2] = 0;
    acc.T[0] = acc.T[1] = acc.T[2] = 0;
    acc.epot = 0;
    acc.bad = false;
    for (int k = 0; k < 9; k++)
        acc.vir[k] = 0;

    R qi[9];
    float qs[9]; // qi in FP32, for the own-patch screen
    // the bead integrator carries no orientations: no patch code in its kernels
    constexpr bool INTG = PH == PH_INTEGRATE || PH == PH_INTEGRATE_F;
    const int npi = (MODE != M_BEAD && !INTG && a.has_patches) ? a.np[ti] : 0;
    if (MODE != M_BEAD && a.rotor) {
        load_orient(a.Q, (size_t)i, qi);
#pragma unroll
        for (int k = 0; k < 9; k++)
            qs[k] = (float)qi[k];
    }
    int exc0 = 0, exc1 = 0;
    if (needs_exclusions(a, i)) {
        exc0 = a.exc_off[oi];
        exc1 = a.exc_off[oi + 1];
    }

    // candidate filter of the patch loop: within max_distance_check
    // (LangevinR.cpp:938) and a type pair with patch rows
    auto is_patch_candidate = [&](const PosRec &xj, double r2s) -> bool {
        if (!(npi > 0 && r2s < a.maxd2))
            return false;
        const int tj = xj.meta & 255;
        if (a.np[tj] == 0)
            return false;
        int ta, tc;
        patch_roles(oi, ti, xj.id, tj, ta, tc);
        return (a.active_mask >> (ta * MAXT + tc)) & 1ull;
    };
    // the candidates the split form hands over: those that pass the screen
    auto is_listed_candidate = [&](const PosRec &xj, int js, const double d[3], double r2s) -> bool {
        (void)js;
        return is_patch_candidate(xj, r2s) && own_screen(a, oi, ti, qs, xj, d, r2s);
    };
    if (PH == PH_FINISH) {
        for (int k = 0; k < 3; k++)
            acc.F[k] = a.Fw[3 * (size_t)s + k];
        const int npc = a.ncand[s];
        const int nq = npc < a.pcap ? npc : a.pcap;
        for (int q = 0; q < nq; q++)
            patch_pair<R, MODE>(a, rp, xi, ti, qi, a.cand[(size_t)q * a.N + s], acc);
        if (npc > a.pcap) {
            int seen = 0;
            trav(i, rp, [&](const PosRec &xj, int j) {
                double d[3];
                min_image(a, rp, xi.x, xi.y, xi.z, xj, d);
                const double r2s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                const int js = j >= 0 ? j : a.slot[xj.id];
                if (!(r2s < a.rmax2) ||
                    !(MDX_WALK_SCREEN ? is_listed_candidate(xj, js, d, r2s) : is_patch_candidate(xj, r2s)))
                    return;
                if (seen++ >= a.pcap)
                    patch_pair<R, MODE>(a, rp, xi, ti, qi, js, acc);
            });
        }
    }
    if (PH == PH_INTEGRATE) {
        // force with patches and patch torque from the compacted patch kernel
        // (which leaves positions without candidates untouched)
        const bool pt = a.has_patches && a.ncand[s] > 0;
        for (int k = 0; k < 3; k++) {
            acc.F[k] = a.Fw[3 * (size_t)s + k];
            acc.T[k] = pt ? a.Tp[3 * (size_t)s + k] : (R)0;
        }
        // the bound pairs' terms after the pair forces (the order of
        // Condensate.cpp:471-473: F = WCA, then the binding routine adds)
        if (a.bF && a.bncand[s] > 0)
            for (int k = 0; k < 3; k++) {
                acc.F[k] += (R)a.bF[3 * (size_t)i + k];
                acc.T[k] += (R)a.bT[3 * (size_t)i + k];
            }
    }
    if (PH == PH_INTEGRATE_F) {
        for (int k = 0; k < 3; k++) {
            acc.F[k] = fin[k];
            acc.T[k] = (R)0;
        }
    }
    if (PH != PH_FINISH && !INTG && (a.wca || npi > 0)) {
        int npc = 0;
        trav(i, rp, [&](const PosRec &xj, int j) {
            double d[3];
            min_image(a, rp, xi.x, xi.y, xi.z, xj, d);
            const double r2s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (!(r2s < a.rmax2))
                return;
            wca_term<R, MODE>(a, d, r2s, xj.id, exc0, exc1, acc);
            if (PH == PH_WALK) {
                const int js = j >= 0 ? j : a.slot[xj.id];
                if (MDX_WALK_SCREEN ? is_listed_candidate(xj, js, d, r2s) : is_patch_candidate(xj, r2s)) {
                    if (npc < a.pcap)
                        a.cand[(size_t)npc * a.N + s] = js;
                    npc++;
                }
            } else if (is_patch_candidate(xj, r2s)) {
                if (npc < PMAX)
                    plist[npc * 256 + threadIdx.x] = j >= 0 ? j : a.slot[xj.id];
                npc++;
            }
        });
        if (PH == PH_WALK) {
            for (int k = 0; k < 3; k++)
                a.Fw[3 * (size_t)s + k] = acc.F[k];
            a.ncand[s] = npc;
            if (npc > a.pcap)
                atomicOr(&a.ctl->ovf, 1u);
            return;
        }
        // -------- patches (LangevinR.cpp:911-1047), in neighbour-loop order
        const int nq = npc < PMAX ? npc : PMAX;
        for (int q = 0; q < nq; q++)
            patch_pair<R, MODE>(a, rp, xi, ti, qi, plist[q * 256 + threadIdx.x], acc);
        if (npc > PMAX) {
            // dense neighbourhoods: a second traversal takes the rest in order
            int seen = 0;
            trav(i, rp, [&](const PosRec &xj, int j) {
                double d[3];
                min_image(a, rp, xi.x, xi.y, xi.z, xj, d);
                const double r2s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (!(r2s < a.rmax2) || !is_patch_candidate(xj, r2s))
                    return;
                if (seen++ >= PMAX)
                    patch_pair<R, MODE>(a, rp, xi, ti, qi, j >= 0 ? j : a.slot[xj.id], acc);
            });
        }
    }
    if (PH == PH_WALK) {
        for (int k = 0; k < 3; k++)
            a.Fw[3 * (size_t)s + k] = 0;
        a.ncand[s] = 0;
        return;
    }
    R *Fi = acc.F, *Ti = acc.T;
    double &epot = acc.epot;
    double *vir = acc.vir;

    // ---------------------------------------- bonds (MD.cpp:1684-1764, 2037-2111)
    if (a.bond_off) {
        for (int b = a.bond_off[oi]; b < a.bond_off[oi + 1]; b++) {
            const PosRec xj = a.X[a.slot[a.bond_j[b]]];
            double d[3];
            min_image(a, rp, xi.x, xi.y, xi.z, xj, d);
            const double dis = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const double r = sqrt(dis);
            const double k = a.bond_k[b], r0 = a.bond_r0[b];
            double f1;
            if (a.bond_kind == MDX_BOND_FENE)
                f1 = -k * r / (1 - (r * r) / (r0 * r0)); // FENEPotential::force (potential.h:356)
            else
                f1 = -k * (r - r0);
            acc.bad |= flag_force(a, f1);
            for (int c = 0; c < 3; c++) {
                const double fac = f1 * d[c] / r;
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
        for (int q = a.ang_off[oi]; q < a.ang_off[oi + 1]; q++) {
            const int ref = a.ang_ref[q];
            const int t = ref >> 2, role = ref & 3;
            const PosRec x1 = a.X[a.slot[a.ang_ijk[3 * t]]];
            const PosRec x2 = a.X[a.slot[a.ang_ijk[3 * t + 1]]];
            const PosRec x3 = a.X[a.slot[a.ang_ijk[3 * t + 2]]];
            double ab[3], bc[3];
            min_image(a, rp, x2.x, x2.y, x2.z, x1, ab);
            min_image(a, rp, x3.x, x3.y, x3.z, x2, bc);
            R abr[3] = {(R)ab[0], (R)ab[1], (R)ab[2]}, bcr[3] = {(R)bc[0], (R)bc[1], (R)bc[2]};
            R out[3];
            if (a.ang_form == MDX_ANGLE_HARMONIC)
                bend_harmonic(abr, bcr, (R)a.ang_k[t], (R)a.ang_t0[t], role, out);
            else
                bend_cosine(abr, bcr, (R)a.ang_k[t], role, out);
            for (int k = 0; k < 3; k++)
                Fi[k] += out[k];
            if (MODE == M_ENERGY && role == 1) {
                // the triplet's energy (BendingPotential::energy,
                // potential3.h:32-41, or k (1 - cos theta)) and virial
                // sum_i F_i x_i = -F1 ab + F3 bc, once, by the middle member
                const double ac = ab[0] * bc[0] + ab[1] * bc[1] + ab[2] * bc[2];
                const double absq = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
                const double bcsq = bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2];
                if (a.ang_form == MDX_ANGLE_HARMONIC) {
                    const double c = fmin(1.0, fmax(-1.0, ac / (sqrt(absq) * sqrt(bcsq))));
                    epot += 0.5 * a.ang_k[t] * (acos(c) - a.ang_t0[t]) * (acos(c) - a.ang_t0[t]);
                } else {
                    epot += a.ang_k[t] * (1 - ac / sqrt(absq * bcsq));
                }
                R f1[3], f3[3];
                if (a.ang_form == MDX_ANGLE_HARMONIC) {
                    bend_harmonic(abr, bcr, (R)a.ang_k[t], (R)a.ang_t0[t], 0, f1);
                    bend_harmonic(abr, bcr, (R)a.ang_k[t], (R)a.ang_t0[t], 2, f3);
                } else {
                    bend_cosine(abr, bcr, (R)a.ang_k[t], 0, f1);
                    bend_cosine(abr, bcr, (R)a.ang_k[t], 2, f3);
                }
                for (int q2 = 0; q2 < 3; q2++)
                    for (int l = 0; l < 3; l++)
                        vir[3 * q2 + l] += -(double)f1[q2] * ab[l] + (double)f3[q2] * bc[l];
            }
        }
    }
    // -------------------------------------------------- tangential active force
    if (a.act_prev) {
        const int nx = a.act_next[oi] < 0 ? (int)oi : a.act_next[oi];
        const int pv = a.act_prev[oi] < 0 ? (int)oi : a.act_prev[oi];
        if (nx != pv) {
            const PosRec xa = a.X[a.slot[nx]], xb = a.X[a.slot[pv]];
            double t[3];
            min_image(a, rp, xa.x, xa.y, xa.z, xb, t);
            const double inv = 1.0 / sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
            for (int k = 0; k < 3; k++)
                Fi[k] += (R)(a.fa * t[k] * inv);
        }
    }

    const bool finite_ft = isfinite((double)Fi[0]) && isfinite((double)Fi[1]) &&
                           isfinite((double)Fi[2]) && isfinite((double)Ti[0]) &&
                           isfinite((double)Ti[1]) && isfinite((double)Ti[2]);
    if (!finite_ft)
        raise_status(a, 2u);

    // ------------------------------------------------------------ epilogues
    if (MODE == M_CONS) {
        for (int k = 0; k < 3; k++) {
            a.Fout[3 * (size_t)oi + k] = (double)Fi[k];
            a.Tout[3 * (size_t)oi + k] = (double)Ti[k];
        }
        return;
    }
    if (MODE == M_ENERGY) {
        double *o = a.obs + 10 * (size_t)oi;
        o[0] = epot;
        for (int k = 0; k < 9; k++)
            o[1 + k] = vir[k];
        return;
    }
    // The moving epilogues keep a particle's state (and carry no force) when
    // its force or torque is non-finite, one of its terms is above the guard,
    // its updated state is non-finite, or the run has stopped (a status bit
    // raised by any particle so far); the inputs are still intact in the
    // current buffers, so the kept state is re-read from there.
    const bool hold = !finite_ft || acc.bad || stopped(a);
    // list mode keeps the slots: both slot maps already hold them
    if (!a.lmode)
        a.slotn[oi] = s;
    if (MODE == M_BEAD) {
        // LangevinNVT::advance_mom (Langevin.cpp:254-269), advance_pos
        // (Langevin.cpp:205-223), binning of the new position
        R nz[6];
        load_noise(a, oi, 3, nz);
        const R rr = (R)rp.bead_r;
        const R c5 = (R)1 - (R)(a.gamma * a.dt / 2.);
        const R c4 = (R)a.b_c2 * rr;
        double xn[3] = {xi.x, xi.y, xi.z};
        R pk[3];
        double dd = 0;
        bool ok = !hold;
        for (int k = 0; k < 3; k++) {
            pk[k] = a.p[3 * (size_t)i + k];
            pk[k] = (R)a.b_c5c2 * pk[k] + (R)a.b_c5c3q * Fi[k] + (c5 * c4 + rr) * nz[k];
            double step;
            if constexpr (FP64)
                step = a.c1 * pk[k];
            else
                step = (double)((R)a.c1 * pk[k]);
            xn[k] = xn[k] + step;
            dd += step * step;
            ok = ok && wrap1(xn[k], pk[k], rp.l[k], a.periodic);
        }
        if (!ok) {
            if (!hold)
                raise_status(a, 2u);
            xn[0] = xi.x;
            xn[1] = xi.y;
            xn[2] = xi.z;
            for (int k = 0; k < 3; k++)
                pk[k] = a.p[3 * (size_t)i + k];
            dd = 0;
        }
        for (int k = 0; k < 3; k++)
            a.pn[3 * (size_t)s + k] = pk[k];
        store_position(a, s, xi, xn);
        if (a.lmode)
            drift = (float)sqrt(dd);
        else
            bin_particle(a, s, rp, xn, a.celln);
        return;
    }

    // rotor: create_forces_and_torques_sphere (LangevinR.cpp:475-578)
    R pl[3] = {a.p[3 * (size_t)i], a.p[3 * (size_t)i + 1], a.p[3 * (size_t)i + 2]};
    R ll[3] = {a.L[3 * (size_t)i], a.L[3 * (size_t)i + 1], a.L[3 * (size_t)i + 2]};
    R nz[6];
    load_noise(a, oi, 6, nz);
    {
        const R *q = qi;
        const R pbx = q[0] * pl[0] + q[1] * pl[1] + q[2] * pl[2];
        const R pby = q[3] * pl[0] + q[4] * pl[1] + q[5] * pl[2];
        const R pbz = q[6] * pl[0] + q[7] * pl[1] + q[8] * pl[2];
        const R gm = (R)a.gm_m;
        const R fx = gm * pbx + (R)rp.Rt * nz[0];
        const R fy = gm * pby + (R)rp.Rt * nz[1];
        const R fz = gm * pbz + (R)rp.Rt * nz[2];
        const R tbx = (R)a.gr_I[0] * ll[0] + (R)rp.Rr * nz[3];
        const R tby = (R)a.gr_I[1] * ll[1] + (R)rp.Rr * nz[4];
        const R tbz = (R)a.gr_I[2] * ll[2] + (R)rp.Rr * nz[5];
        const R Fx = Fi[0] + fx * q[0] + fy * q[3] + fz * q[6];
        const R Fy = Fi[1] + fx * q[1] + fy * q[4] + fz * q[7];
        const R Fz = Fi[2] + fx * q[2] + fy * q[5] + fz * q[8];
        const R Tx = tbx + Ti[0] * q[0] + Ti[1] * q[1] + Ti[2] * q[2];
        const R Ty = tby + Ti[0] * q[3] + Ti[1] * q[4] + Ti[2] * q[5];
        const R Tz = tbz + Ti[0] * q[6] + Ti[1] * q[7] + Ti[2] * q[8];
        Fi[0] = Fx;
        Fi[1] = Fy;
        Fi[2] = Fz;
        Ti[0] = Tx;
        Ti[1] = Ty;
        Ti[2] = Tz;
    }
    const R hdt = (R)a.hdt;
    if (MODE == M_STEP || MODE == M_FUSED)
        for (int k = 0; k < 3; k++) {
            // advancemom_halfstep (LangevinR.cpp:327-347)
            pl[k] += hdt * Fi[k];
            ll[k] += hdt * Ti[k];
        }
    if (MODE == M_PRIME || MODE == M_STEP) {
        bool ok = !hold;
        for (int k = 0; k < 3; k++)
            ok = ok && isfinite((double)pl[k]) && isfinite((double)ll[k]) &&
                 isfinite((double)Fi[k]) && isfinite((double)Ti[k]);
        if (!ok) {
            if (!hold)
                raise_status(a, 2u);
            for (int k = 0; k < 3; k++) {
                pl[k] = a.p[3 * (size_t)i + k];
                ll[k] = a.L[3 * (size_t)i + k];
                Fi[k] = Ti[k] = (R)0;
            }
        }
        a.Xn[s] = xi;
        copy_orient(a.Qn, (size_t)s, a.Q, (size_t)i);
        for (int k = 0; k < 3; k++) {
            a.pn[3 * (size_t)s + k] = pl[k];
            a.Ln[3 * (size_t)s + k] = ll[k];
            a.Fn[3 * (size_t)s + k] = Fi[k];
            a.Tn[3 * (size_t)s + k] = Ti[k];
        }
        return;
    }
    // M_FUSED: the next step's half kick with the same F, T, drift, rotation
    for (int k = 0; k < 3; k++) {
        pl[k] += hdt * Fi[k];
        ll[k] += hdt * Ti[k];
    }
    double xn[3] = {xi.x, xi.y, xi.z};
    double dd = 0;
    bool ok = !hold;
    for (int k = 0; k < 3; k++) {
        double step;
        if constexpr (FP64)
            step = a.c1 * pl[k];
        else
            step = (double)((R)a.c1 * pl[k]);
        xn[k] = xn[k] + step;
        dd += step * step;
        ok = ok && wrap1(xn[k], pl[k], rp.l[k], a.periodic);
    }
    const R Ir[3] = {(R)a.I[0], (R)a.I[1], (R)a.I[2]};
    rotate((R)a.dt, Ir, ll, qi);
    for (int k = 0; k < 3; k++)
        ok = ok && isfinite((double)ll[k]);
    for (int k = 0; k < 9; k++)
        ok = ok && isfinite((double)qi[k]);
    if (!ok) {
        if (!hold)
            raise_status(a, 2u);
        xn[0] = xi.x;
        xn[1] = xi.y;
        xn[2] = xi.z;
        for (int k = 0; k < 3; k++) {
            pl[k] = a.p[3 * (size_t)i + k];
            ll[k] = a.L[3 * (size_t)i + k];
        }
        dd = 0;
    }
    for (int k = 0; k < 3; k++) {
        a.pn[3 * (size_t)s + k] = pl[k];
        a.Ln[3 * (size_t)s + k] = ll[k];
    }
    if (ok)
        store_orient(a.Qn, (size_t)s, qi);
    else
        copy_orient(a.Qn, (size_t)s, a.Q, (size_t)i);
    store_position(a, s, xi, xn);
    if (a.lmode)
        drift = (float)sqrt(dd);
    else
        bin_particle(a, s, rp, xn, a.celln);
}

// the first kernel after a list build clears the request and the bound
template <class R> __device__ __forceinline__ void list_built(const KArgs<R> &a)
{
    if (a.lmode && blockIdx.x == 0 && threadIdx.x == 0 && a.ctl->rebuild) {
        a.ctl->rebuild = 0u;
        a.ctl->D = 0.0;
    }
}

// one thread per cell-list position (list mode: per slot)
template <class R, int MODE> __global__ void __launch_bounds__(256) k_gather(KArgs<R> a)
{
    __shared__ int plist[PMAX * 256];
    list_built(a);
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    float drift = 0.f;
    if (s < a.N)
        process_particle<R, MODE, PH_ALL>(
            a, s, plist, [&](int i, const DevRep &rp, auto &&fn) { for_each_near(a, rp, i, fn); },
            drift);
    if ((MODE == M_FUSED || MODE == M_BEAD) && a.lmode)
        drift_bound(a, drift);
}

// split form: the neighbour walk ...
template <class R, int MODE> __global__ void __launch_bounds__(256) k_walk(KArgs<R> a)
{
    list_built(a);
    const int s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (s >= a.N)
        return;
    float drift = 0.f;
    process_particle<R, MODE, PH_WALK>(
        a, s, nullptr, [&](int i, const DevRep &rp, auto &&fn) { for_each_near(a, rp, i, fn); },
        drift);
}

// List-mode walk, compacted: the list entries of a block's 256 positions are
// laid out back to back and evaluated one per thread (every lane busy, instead
// of each lane looping over its own list with the wave waiting for its longest
// list); each position then adds its entries' WCA forces and appends its
// patch candidates in list order -- the same sums in the same order as the
// one-lane walk.  Positions with more than kmax entries walk the build's
// cells in their own lane.
// INT: positions without patch candidates are integrated here (their force
// is complete), the others are left to the patch kernel
// FP32 without the integrator: at most 80 VGPRs, so 6 waves per SIMD (what
// its LDS allows) hide more of the gathers' latency (a few bytes of spill,
// 13 % faster on the 1M condensate); the FP64 form is held to 4 waves by its
// LDS and the fused form needs the integrator's registers
template <class R, int MODE, bool INT>
__global__ void __launch_bounds__(256, (sizeof(R) == 4 && !INT) ? 6 : 1) k_walkc(KArgs<R> a)
{
    __shared__ int off[257];
    __shared__ PosRec sx[256];
    __shared__ float sq[256 * 9]; // orientations in FP32 for the own-patch screen
    __shared__ int sexc[256 * 2];
    __shared__ unsigned char spl[256]; // face_clear of each position
    __shared__ R pf[256 * 3];
    __shared__ int pj[256];
    __shared__ int wsum[4];
    list_built(a);
    const int tid = threadIdx.x;
    const int base = xcd_block(blockIdx.x, gridDim.x) * 256;
    const int s = base + tid;
    const bool live = s < a.N;
    const int n = live ? list_count(a, s) : 0;
    const bool fb = n > a.kmax;
    const int cnt = fb ? 0 : n;
    PosRec xi;
    if (live) {
        xi = a.X[s];
        sx[tid] = xi;
        if (a.has_patches && a.np[xi.meta & 255] > 0) {
            R qm[9];
            load_orient(a.Q, (size_t)s, qm);
#pragma unroll
            for (int k = 0; k < 9; k++)
                sq[9 * tid + k] = (float)qm[k];
        }
        const bool ne = needs_exclusions(a, s);
        sexc[2 * tid] = ne ? a.exc_off[xi.id] : 0;
        sexc[2 * tid + 1] = ne ? a.exc_off[xi.id + 1] : 0;
        spl[tid] = face_clear(a, rep_rec(a, xi.meta), xi);
    }
    // exclusive scan of the entry counts
    int v = cnt;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o);
        if ((tid & 63) >= o)
            v += t;
    }
    if ((tid & 63) == 63)
        wsum[tid >> 6] = v;
    __syncthreads();
    int wb = 0;
    for (int w = 0; w < (tid >> 6); w++)
        wb += wsum[w];
    const int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    off[tid] = wb + v - cnt;
    if (tid == 255)
        off[256] = total;
    __syncthreads();

    // WCA of a pair into acc; true for a patch candidate that passes the
    // own-patch screen (the one-lane walk's rule)
    auto visit = [&](const PosRec &xo, const float *qo, int e0, int e1, bool pl,
                     const PosRec &xj, Acc<R> &acc) -> bool {
        const DevRep rp = rep_rec(a, xo.meta);
        double d[3];
        if (pl) {
            d[0] = xo.x - xj.x;
            d[1] = xo.y - xj.y;
            d[2] = xo.z - xj.z;
        } else {
            min_image(a, rp, xo.x, xo.y, xo.z, xj, d);
        }
        const double r2s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (!(r2s < a.rmax2))
            return false;
        wca_term<R, MODE>(a, d, r2s, xj.id, e0, e1, acc);
        const int to = xo.meta & 255;
        if (!(a.has_patches && a.np[to] > 0 && r2s < a.maxd2))
            return false;
        const int tj = xj.meta & 255;
        if (a.np[tj] == 0)
            return false;
        int ta, tc;
        patch_roles(xo.id, to, xj.id, tj, ta, tc);
        if (!((a.active_mask >> (ta * MAXT + tc)) & 1ull))
            return false;
        return !MDX_WALK_SCREEN || own_screen(a, xo.id, to, qo, xj, d, r2s);
    };

    R F[3] = {0, 0, 0};
    int npc = 0;
    auto owner_of = [&](int e) -> int {
        int lo = 0, hi = 255;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (off[mid] <= e)
                lo = mid;
            else
                hi = mid - 1;
        }
        return lo;
    };
    const int *lst = a.nbrc + (size_t)base * a.kmax;
    // a list built since the last compacted walk: this block's entries are
    // first laid out back to back (and read back below by this block)
    if (a.ctl->fresh) {
        int *seg = a.nbrcw + (size_t)base * a.kmax;
        for (int e = tid; e < total; e += 256) {
            const int o = owner_of(e);
            seg[e] = a.nbr[(size_t)(e - off[o]) * a.N + base + o];
        }
        __threadfence_block();
        __syncthreads();
    }
    // software pipeline: round r evaluates the record loaded in round r - 1
    // and loads the record of round r + 1 and the list entry of round r + 2
    int j1 = tid < total ? lst[tid] : 0;
    int j2 = tid + 256 < total ? lst[tid + 256] : 0;
    PosRec x1;
    if (tid < total)
        x1 = a.X[j1];
    for (int r0 = 0; r0 < total; r0 += 256) {
        const int e = r0 + tid;
        const int j = j1;
        const PosRec xj = x1;
        j1 = j2;
        if (e + 256 < total)
            x1 = a.X[j1];
        if (e + 512 < total)
            j2 = lst[e + 512];
        if (e < total) {
            const int o = owner_of(e);
            const PosRec xo = sx[o];
            Acc<R> pa;
            pa.F[0] = pa.F[1] = pa.F[2] = 0;
            pa.bad = false;
            const bool c = visit(xo, sq + 9 * o, sexc[2 * o], sexc[2 * o + 1], spl[o], xj, pa);
            for (int k = 0; k < 3; k++)
                pf[3 * tid + k] = pa.F[k];
            pj[tid] = c ? j : -1;
        }
        __syncthreads();
        const int lo = max(off[tid], r0), hi = min(off[tid] + cnt, r0 + 256);
        for (int e2 = lo; e2 < hi; e2++) {
            for (int k = 0; k < 3; k++)
                F[k] += pf[3 * (e2 - r0) + k];
            const int jc = pj[e2 - r0];
            if (jc >= 0) {
                if (npc < a.pcap)
                    a.cand[(size_t)npc * a.N + s] = jc;
                npc++;
            }
        }
        __syncthreads();
    }
    float drift = 0.f;
    if (live) {
        if (fb) {
            const DevRep rp = rep_rec(a, xi.meta);
            Acc<R> pa;
            pa.F[0] = pa.F[1] = pa.F[2] = 0;
            pa.bad = false;
            for_each_in_cells(a, rp, a.bcell[s], s, [&](const PosRec &xj, int j) {
                if (visit(xi, sq + 9 * tid, sexc[2 * tid], sexc[2 * tid + 1], spl[tid], xj, pa)) {
                    if (npc < a.pcap)
                        a.cand[(size_t)npc * a.N + s] = j;
                    npc++;
                }
            });
            for (int k = 0; k < 3; k++)
                F[k] = pa.F[k];
        }
        a.ncand[s] = npc;
        if (npc > a.pcap) {
            if (a.ovf_fatal)
                raise_status(a, 4u);
            else
                atomicOr(&a.ctl->ovf, 1u);
        }
        if (!(INT && npc == 0))
            for (int k = 0; k < 3; k++)
                a.Fw[3 * (size_t)s + k] = F[k];
    }
    if (INT) {
        // the positions without candidates, compacted to the first lanes
        // (slot in pj, force in pf), are integrated one per lane
        const bool here = live && npc == 0;
        const unsigned long long m = __ballot(here);
        const int lane = tid & 63;
        const int rank = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0)
            wsum[tid >> 6] = __popcll(m);
        __syncthreads();
        int at = rank;
        for (int w = 0; w < (tid >> 6); w++)
            at += wsum[w];
        const int nint = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (here) {
            pj[at] = s;
            for (int k = 0; k < 3; k++)
                pf[3 * at + k] = F[k];
        }
        __syncthreads();
        if (tid < nint) {
            const R fin[3] = {pf[3 * tid], pf[3 * tid + 1], pf[3 * tid + 2]};
            process_particle<R, MODE, PH_INTEGRATE_F>(
                a, pj[tid], nullptr,
                [&](int i, const DevRep &rp, auto &&fn) { for_each_near(a, rp, i, fn); }, drift,
                fin);
        }
        if (MODE == M_FUSED || MODE == M_BEAD)
            drift_bound(a, drift);
    }
}

// ... then the patch pairs, topology and the integrator
template <class R, int MODE> __global__ void __launch_bounds__(256) k_finish(KArgs<R> a)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    float drift = 0.f;
    if (s < a.N)
        process_particle<R, MODE, PH_FINISH>(
            a, s, nullptr, [&](int i, const DevRep &rp, auto &&fn) { for_each_near(a, rp, i, fn); },
            drift);
    if ((MODE == M_FUSED || MODE == M_BEAD) && a.lmode)
        drift_bound(a, drift);
}

// the patch candidates of slot so beyond the pcap handed over by the walk:
// the neighbours are walked again in walk order and the terms after the
// first pcap candidates added to Fw / Tp (k_patch has added the others)
template <class R, int MODE>
__device__ __forceinline__ void patch_overflow(const KArgs<R> &a, int so)
{
    const int i = a.lmode ? so : a.list[so];
    const PosRec xi = a.X[i];
    const int ti = xi.meta & 255;
    const DevRep rp = rep_rec(a, xi.meta);
    R qi[9];
    float qs[9]; // qi in FP32, for the own-patch screen
    load_orient(a.Q, (size_t)i, qi);
    for (int k2 = 0; k2 < 9; k2++)
        qs[k2] = (float)qi[k2];
    Acc<R> acc;
    acc.bad = false;
    for (int k2 = 0; k2 < 3; k2++) {
        acc.F[k2] = a.Fw[3 * (size_t)so + k2];
        acc.T[k2] = a.Tp[3 * (size_t)so + k2];
    }
    const int npi = a.np[ti];
    int seen = 0;
    for_each_near(a, rp, i, [&](const PosRec &xj, int j) {
        double d[3];
        min_image(a, rp, xi.x, xi.y, xi.z, xj, d);
        const double r2s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (!(r2s < a.rmax2) || !(npi > 0 && r2s < a.maxd2))
            return;
        const int tj = xj.meta & 255;
        if (a.np[tj] == 0)
            return;
        int ta, tc;
        patch_roles(xi.id, ti, xj.id, tj, ta, tc);
        if (!((a.active_mask >> (ta * MAXT + tc)) & 1ull))
            return;
        if (MDX_WALK_SCREEN && !own_screen(a, xi.id, ti, qs, xj, d, r2s))
            return;
        const int js = j >= 0 ? j : a.slot[xj.id];
        if (seen++ >= a.pcap)
            patch_pair<R, MODE>(a, rp, xi, ti, qi, js, acc);
    });
    for (int k2 = 0; k2 < 3; k2++) {
        a.Fw[3 * (size_t)so + k2] = acc.F[k2];
        a.Tp[3 * (size_t)so + k2] = acc.T[k2];
    }
}

// split form, compacted: the patch candidates of a block of pblk positions
// are laid out back to back and evaluated one per thread (no idle lanes on
// positions without candidates; large blocks keep the whole
// grid resident).  The thread holding a position's first candidate of a
// round then adds that round's terms in candidate order to the position's
// force (the walk force in Fw) and patch torque (Tp): the same order as the
// one-kernel form, and bitwise equal to it while at most one patch pair of a
// particle pair passes the cone tests (true for cos(theta_m) >= 0.5 with
// these patch sets).  Positions without candidates are not touched (the
// integrator reads ncand).
constexpr int PBLK = 2048; // largest block of positions (pblk: 256 .. PBLK, a multiple of 256)
// INT: the positions with candidates are then integrated here (the walk
// integrated the others)
template <class R, int MODE, bool INT>
__global__ void __launch_bounds__(256) k_patch(KArgs<R> a, int pblk)
{
    const int PER = pblk / 256;
    __shared__ int off[PBLK + 1];
    __shared__ R ent[256 * 6];
    __shared__ int wsum[4];
    const int tid = threadIdx.x;
    const int base = xcd_block(blockIdx.x, gridDim.x) * pblk;
    // exclusive scan of min(ncand, pcap) in position order (position
    // k * 256 + tid of the block, k < PER)
    int carry = 0;
    for (int k = 0; k < PER; k++) {
        const int p = base + k * 256 + tid;
        const int nc = p < a.N ? a.ncand[p] : 0;
        const int nq = nc < a.pcap ? nc : a.pcap;
        int v = nq;
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(v, o);
            if ((tid & 63) >= o)
                v += t;
        }
        if ((tid & 63) == 63)
            wsum[tid >> 6] = v;
        __syncthreads();
        int wb = 0;
        for (int w = 0; w < (tid >> 6); w++)
            wb += wsum[w];
        off[k * 256 + tid] = carry + wb + v - nq;
        carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (tid == 0)
        off[pblk] = carry;
    __syncthreads();
    const int total = carry;
    for (int r0 = 0; r0 < total; r0 += 256) {
        const int e = r0 + tid;
        int o = 0;
        if (e < total) {
            // owner: the last position whose offset is <= e
            int lo = 0, hi = pblk - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (off[mid] <= e)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            o = lo;
            const int so = base + o;
            const int q = e - off[o];
            const int i = a.lmode ? so : a.list[so];
            const PosRec xo = a.X[i];
            const DevRep rp = rep_rec(a, xo.meta);
            R qo[9];
            load_orient(a.Q, (size_t)i, qo);
            Acc<R> ea;
            for (int k = 0; k < 3; k++)
                ea.F[k] = ea.T[k] = 0;
            ea.bad = false;
            patch_pair<R, MODE>(a, rp, xo, xo.meta & 255, qo, a.cand[(size_t)q * a.N + so], ea);
            for (int k = 0; k < 3; k++) {
                ent[6 * tid + k] = ea.F[k];
                ent[6 * tid + 3 + k] = ea.T[k];
            }
        }
        __syncthreads();
        if (e < total) {
            const int first = max(off[o], r0);
            if (e == first) {
                const int last = min(off[o + 1], r0 + 256);
                const size_t so = (size_t)base + o;
                R F[3], T[3];
                for (int k = 0; k < 3; k++) {
                    F[k] = a.Fw[3 * so + k];
                    T[k] = first == off[o] ? (R)0 : a.Tp[3 * so + k];
                }
                for (int e2 = first; e2 < last; e2++)
                    for (int k = 0; k < 3; k++) {
                        F[k] += ent[6 * (e2 - r0) + k];
                        T[k] += ent[6 * (e2 - r0) + 3 + k];
                    }
                for (int k = 0; k < 3; k++) {
                    a.Fw[3 * so + k] = F[k];
                    a.Tp[3 * so + k] = T[k];
                }
            }
        }
        __threadfence_block();
        __syncthreads();
    }
    // candidates beyond the handed-over ones (rare): inline in the fused form,
    // else k_patch_ovf after this kernel (keeps this kernel at 3 waves per SIMD)
    if (INT)
        for (int k = 0; k < PER; k++) {
            const int so = base + k * 256 + tid;
            if (so < a.N && a.ncand[so] > a.pcap)
                patch_overflow<R, MODE>(a, so);
        }
    if (INT) {
        // the positions with candidates, compacted into off[], integrated one
        // per lane
        __threadfence_block();
        __syncthreads();
        int cnt = 0;
        for (int k = 0; k < PER; k++) {
            const int so = base + k * 256 + tid;
            const bool here = so < a.N && a.ncand[so] > 0;
            const unsigned long long m = __ballot(here);
            const int lane = tid & 63;
            const int rank = __popcll(m & ((1ull << lane) - 1ull));
            if (lane == 0)
                wsum[tid >> 6] = __popcll(m);
            __syncthreads();
            int wb = 0;
            for (int w = 0; w < (tid >> 6); w++)
                wb += wsum[w];
            if (here)
                off[cnt + wb + rank] = so;
            cnt += wsum[0] + wsum[1] + wsum[2] + wsum[3];
            __syncthreads();
        }
        float drift = 0.f;
        for (int t = tid; t < cnt; t += 256) {
            float d = 0.f;
            process_particle<R, MODE, PH_INTEGRATE>(
                a, off[t], nullptr,
                [&](int i, const DevRep &rp, auto &&fn) { for_each_near(a, rp, i, fn); }, d);
            drift = fmaxf(drift, d);
        }
        if (MODE == M_FUSED || MODE == M_BEAD)
            drift_bound(a, drift);
    }
}

// the candidates beyond pcap of the separate form's positions (after k_patch;
// grid-stride over a small grid; almost every step has none, and then the
// walk left the gate closed)
template <class R, int MODE> __global__ void __launch_bounds__(256) k_patch_ovf(KArgs<R> a)
{
    if (!a.ctl->ovf)
        return;
    for (int so = blockIdx.x * blockDim.x + threadIdx.x; so < a.N; so += gridDim.x * blockDim.x)
        if (a.ncand[so] > a.pcap)
            patch_overflow<R, MODE>(a, so);
}

// ... and the topology terms and the integrator
template <class R, int MODE> __global__ void __launch_bounds__(256) k_integrate(KArgs<R> a)
{
    const int s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        a.ctl->ovf = 0u; // the overflow kernel (before this one) has read it
    float drift = 0.f;
    if (s < a.N)
        process_particle<R, MODE, PH_INTEGRATE>(
            a, s, nullptr, [&](int i, const DevRep &rp, auto &&fn) { for_each_near(a, rp, i, fn); },
            drift);
    if ((MODE == M_FUSED || MODE == M_BEAD) && a.lmode)
        drift_bound(a, drift);
}

// neighbour list of every particle from the cell list of the current
// positions: the slots within sqrt(rl2) in the order of a 27-cell walk
// (z rows -1, 0, +1 outer, y rows -1, 0, +1 inner, then the cells x - 1,
// x, x + 1 of the row), written to nbr[k * N + slot] with the count in nn;
// the compacted walk lays the entries out per block afterwards (fresh).
// gated: only when a rebuild was requested.
// One thread per cell-list position: a block's particles then lie in a few
// consecutive x-rows of cells (the list is in cell order).  The block first
// finds, for each z offset (one wave each), the distinct neighbour rows of
// its rows and loads their cell offsets into LDS; then, one z offset at a
// time, it stages those rows' particles (position record and slot, read with
// coalesced loads from the list) and walks them there.  Each staged row
// holds its last cell, the row, then its first cell again, so that the three
// cells about any x -- across the periodic seam too -- are one contiguous
// run.  The walk first marks the candidates within the list radius in a bit
// mask (a distance per candidate, no other work), then takes the marked ones
// in order.  A block whose particles span more than NB_W rows or two
// replicas, a grid with fewer than 3 cells on a side, and a z offset whose
// rows hold more than NB_CAP particles or NB_ROWCAP cell offsets walk global
// memory instead.  NB_CAP: 800 (38 KB of LDS, 4 blocks per CU) for sparse
// systems, 1152 (50 KB, 3 blocks) where an x-row of cells holds many
// particles (the host picks it from the mean particles per row).
constexpr int NB_W = 16;        // own x-rows of a block walked from LDS
constexpr int NB_CAP_S = 800;   // staged particles per z offset (sparse)
constexpr int NB_CAP_D = 1152;  // (dense)
constexpr int NB_ROWCAP = 1152; // staged cell offsets per z offset

// the list ranges [r0, r1) and [r2, r3) of the cells x - 1, x, x + 1 about
// x-cell cx, in that order, in a row whose cell offsets are off[0..n0]
// (across the seam: two ranges; the whole row when it has fewer than 3 cells)
template <class T>
__device__ __forceinline__ void row_ranges(const T *off, int n0, int cx, int &r0, int &r1, int &r2,
                                           int &r3)
{
    r2 = r3 = 0;
    if (n0 >= 3) {
        if (cx == 0) {
            r0 = off[n0 - 1];
            r1 = off[n0];
            r2 = off[0];
            r3 = off[2];
        } else if (cx == n0 - 1) {
            r0 = off[n0 - 2];
            r1 = off[n0];
            r2 = off[0];
            r3 = off[1];
        } else {
            r0 = off[cx - 1];
            r1 = off[cx + 2];
        }
    } else {
        r0 = off[0];
        r1 = off[n0];
    }
}

template <class R, int NB_CAP>
__global__ void __launch_bounds__(256) k_nbuild(KArgs<R> a, const int *cells, int gated)
{
    if (gated && !a.ctl->rebuild)
        return;
    __shared__ PosRec lpos[NB_CAP];
    __shared__ int lslot[NB_CAP];
    // per z offset p: n0 + 1 cell offsets per staged row, relative to the
    // row's first list position hk0
    __shared__ unsigned short rowoff[3][NB_ROWCAP];
    __shared__ int crow[3][3 * NB_W];     // first cell of each candidate row
    __shared__ int hrow[3][3 * NB_W];     // its staged row
    __shared__ int hcell[3][3 * NB_W];    // first cell of each staged row
    __shared__ int hk0[3][3 * NB_W];      // list position of its first particle
    __shared__ int hoff[3][3 * NB_W + 1]; // LDS index of its first staged particle
    __shared__ int pinfo[3][3]; // staged rows; staged particles (> NB_CAP: global); a row over NB_CAP
    __shared__ int ends[4];     // replica and cell of the first and last position
    __shared__ int binfo[3];    // replica (-1: global walk), first own row, own rows
    const int tid = threadIdx.x;
    if (blockIdx.x == 0 && tid == 0) {
        atomicAdd(&a.ctl->nbuilds, 1u);
        a.ctl->fresh = 1u;
    }
    const int p0 = xcd_block(blockIdx.x, gridDim.x) * blockDim.x;
    const int pos = p0 + tid;
    const int plast = min(p0 + (int)blockDim.x, a.N) - 1;
    const bool live = pos <= plast;
    int s = 0, c = 0;
    PosRec xi;
    if (live) {
        s = a.list[pos];
        xi = a.X[s];
        c = cells[s];
        a.bcell[s] = c;
        if (pos == p0) {
            ends[0] = (int)(xi.meta >> 8);
            ends[1] = c;
        }
        if (pos == plast) {
            ends[2] = (int)(xi.meta >> 8);
            ends[3] = c;
        }
    }
    __syncthreads();
    if (tid == 0) {
        // the first and last positions bound the block's cells (cell order)
        int use = -1, rf = 0, w = 0;
        if (ends[0] == ends[2]) {
            const DevRep &rp = a.reps[ends[0]];
            const int n0 = rp.nc[0];
            rf = (ends[1] - rp.coff) / n0;
            w = (ends[3] - rp.coff) / n0 - rf + 1;
            if (n0 >= 3 && rp.nc[1] >= 3 && rp.nc[2] >= 3 && w <= NB_W)
                use = ends[0];
        }
        binfo[0] = use;
        binfo[1] = rf;
        binfo[2] = w;
    }
    __syncthreads();

    int n = 0;
    bool kept_excluded = false;
    int exc0 = 0, exc1 = 0, ti = 0;
    DevRep rp;
    if (live) {
        rp = rep_rec(a, xi.meta);
        if (a.exc_off) {
            exc0 = a.exc_off[xi.id];
            exc1 = a.exc_off[xi.id + 1];
        }
        ti = xi.meta & 255;
    }
    // the pair's types carry an active patch row
    auto patch_active = [&](const PosRec &xj) -> bool {
        const int tj = xj.meta & 255;
        if (!(a.has_patches && a.np[ti] > 0 && a.np[tj] > 0))
            return false;
        int ta, tc;
        patch_roles(xi.id, ti, xj.id, tj, ta, tc);
        return (a.active_mask >> (ta * MAXT + tc)) & 1ull;
    };
    // plain: no candidate can be a periodic image (the cell is off the box
    // faces, at least 5 cells per side, or walls), so the minimum image is the
    // plain difference (the same value min_image returns)
    bool plain = false;
    auto dist2 = [&](const PosRec &xj) -> double {
        double d[3];
        if (plain) {
            d[0] = xi.x - xj.x;
            d[1] = xi.y - xj.y;
            d[2] = xi.z - xj.z;
        } else {
            min_image(a, rp, xi.x, xi.y, xi.z, xj, d);
        }
        return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    };
    // a candidate within the list radius (r2 < rl2): listed unless beyond
    // the WCA reach without patch terms, or an exclusion without them
    auto take = [&](const PosRec &xj, int j, double r2) {
        if (!(r2 < a.rl2w) && !patch_active(xj))
            return;
        bool ex = false;
        for (int e = exc0; e < exc1; e++)
            ex |= (uint32_t)a.exc_j[e] == xj.id;
        if (ex) {
            if (!patch_active(xj))
                return;
            kept_excluded = true;
        }
        if (n < a.kmax)
            a.nbrw[(size_t)n * a.N + s] = j;
        n++;
    };
    auto consider = [&](const PosRec &xj, int j) {
        const double r2 = dist2(xj);
        if (r2 < a.rl2)
            take(xj, j, r2);
    };
    auto global_range = [&](int r0, int r1) {
        for (int k = r0; k < r1; k++) {
            const int j = a.list[k];
            if (j != s)
                consider(a.X[j], j);
        }
    };

    if (binfo[0] < 0) {
        // the 27 cells from global memory, in the same order
        if (live) {
            const int cl = c - rp.coff;
            const int n0 = rp.nc[0], n1 = rp.nc[1], n2 = rp.nc[2];
            const int cx = cl % n0, cy = (cl / n0) % n1, cz = cl / (n0 * n1);
            const int ylo = n1 >= 3 ? -1 : 0, yhi = n1 >= 2 ? 1 : 0;
            const int zlo = n2 >= 3 ? -1 : 0, zhi = n2 >= 2 ? 1 : 0;
            for (int oz = zlo; oz <= zhi; oz++) {
                int z = cz + oz;
                z = z < 0 ? z + n2 : (z >= n2 ? z - n2 : z);
                for (int oy = ylo; oy <= yhi; oy++) {
                    int y = cy + oy;
                    y = y < 0 ? y + n1 : (y >= n1 ? y - n1 : y);
                    int r0, r1, r2, r3;
                    row_ranges(a.start + rp.coff + n0 * (y + n1 * z), n0, cx, r0, r1, r2, r3);
                    global_range(r0, r1);
                    global_range(r2, r3);
                }
            }
        }
    } else {
        const DevRep &rb = a.reps[binfo[0]];
        const int n0 = rb.nc[0], n1 = rb.nc[1], n2 = rb.nc[2];
        const int rf = binfo[1], w = binfo[2];
        const int cc = c - rb.coff;
        const int cx = cc % n0, ro = cc / n0 - rf;
        const int cy = (cc / n0) % n1, cz = cc / (n0 * n1);
        plain = !a.periodic || (n0 >= 5 && n1 >= 5 && n2 >= 5 && cx >= 1 && cx <= n0 - 2 &&
                                cy >= 1 && cy <= n1 - 2 && cz >= 1 && cz <= n2 - 2);
        const int wv = tid >> 6, lane = tid & 63;
        // wave p (z offset p - 1): the candidate rows, own row rf + m / 3
        // moved by m % 3 - 1 in y and by p - 1 in z; the distinct ones are
        // staged in order of first appearance
        if (wv < 3) {
            const bool act = lane < 3 * w;
            int rid = -1;
            if (act) {
                const int r = rf + lane / 3;
                int yy = r % n1 + lane % 3 - 1, zz = r / n1 + wv - 1;
                yy = yy < 0 ? yy + n1 : (yy >= n1 ? yy - n1 : yy);
                zz = zz < 0 ? zz + n2 : (zz >= n2 ? zz - n2 : zz);
                rid = rb.coff + n0 * (yy + n1 * zz);
                crow[wv][lane] = rid;
            }
            // the first lane holding the same row
            int first = lane;
            for (int m = 3 * w - 1; m >= 0; m--) {
                const int o = __shfl(rid, m);
                if (m < lane && o == rid)
                    first = m;
            }
            const bool uniq = act && first == lane;
            const unsigned long long bm = __ballot(uniq);
            const int h = __popcll(bm & ((1ull << lane) - 1ull));
            if (uniq)
                hcell[wv][h] = rid;
            const int hf = __shfl(h, first);
            if (act)
                hrow[wv][lane] = hf;
            if (lane == 0) {
                pinfo[wv][0] = __popcll(bm);
                pinfo[wv][2] = 0;
            }
        }
        __syncthreads();
        // the cell offsets of every staged row of the three z offsets
        for (int p = 0; p < 3; p++) {
            const int nh = pinfo[p][0];
            if (nh * (n0 + 1) <= NB_ROWCAP)
                for (int t = tid; t < nh * (n0 + 1); t += blockDim.x) {
                    const int h = t / (n0 + 1);
                    const int x = t - h * (n0 + 1);
                    const int k0 = a.start[hcell[p][h]];
                    const int r = a.start[hcell[p][h] + x] - k0;
                    rowoff[p][t] = (unsigned short)r;
                    if (x == 0)
                        hk0[p][h] = k0;
                    if (r > NB_CAP)
                        pinfo[p][2] = 1; // (then the 16-bit offsets may not hold it)
                }
        }
        __syncthreads();
        if (wv < 3) {
            // LDS index of each staged row's first particle: exclusive scan
            // of the staged rows' sizes (the row plus its last and first
            // cells once more)
            const int nh = pinfo[wv][0];
            const bool fit = nh * (n0 + 1) <= NB_ROWCAP && !pinfo[wv][2];
            int cnt = 0;
            if (fit && lane < nh) {
                const unsigned short *off = rowoff[wv] + lane * (n0 + 1);
                cnt = off[n0] + (off[n0] - off[n0 - 1]) + off[1];
            }
            int v = cnt;
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(v, o);
                if (lane >= o)
                    v += t;
            }
            if (lane <= nh)
                hoff[wv][lane] = v - cnt;
            if (lane == 63)
                pinfo[wv][1] = fit ? v : NB_CAP + 1;
        }
        __syncthreads();
        for (int p = 0; p < 3; p++) {
            const int nh = pinfo[p][0];
            const int total = pinfo[p][1];
            const bool staged = total <= NB_CAP;
            if (staged)
                for (int f = tid; f < total; f += blockDim.x) {
                    int lo = 0, hi = nh - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (hoff[p][mid] <= f)
                            lo = mid;
                        else
                            hi = mid - 1;
                    }
                    // the staged row: last cell, the row, first cell
                    const unsigned short *off = rowoff[p] + lo * (n0 + 1);
                    const int cl = off[n0] - off[n0 - 1], cr = off[n0];
                    int o = f - hoff[p][lo];
                    o = o < cl ? off[n0 - 1] + o : (o < cl + cr ? o - cl : o - cl - cr);
                    const int j = a.list[hk0[p][lo] + o];
                    lslot[f] = j;
                    lpos[f] = a.X[j];
                }
            __syncthreads();
            if (live && staged) {
                // the thread's candidates of this z offset, in walk order:
                // three LDS runs, one per row (cells cx - 1 .. cx + 1 start at
                // staged cell cx, the last cell standing first)
                int qs[3], cum[3], tot = 0;
                for (int dy = 0; dy < 3; dy++) {
                    const int h = hrow[p][3 * ro + dy];
                    const unsigned short *off = rowoff[p] + h * (n0 + 1);
                    const int cl = off[n0] - off[n0 - 1];
                    const int b0 = cx == 0 ? 0 : cl + off[cx - 1];
                    const int b1 = cx + 3 == n0 + 2 ? cl + off[n0] + off[1] : cl + off[cx + 2];
                    qs[dy] = hoff[p][h] + b0;
                    tot += b1 - b0;
                    cum[dy] = tot;
                }
                // marks, 64 candidates at a time, then the marked ones in order
                for (int t0 = 0; t0 < tot; t0 += 64) {
                    const int te = min(tot, t0 + 64);
                    unsigned long long hit = 0ull;
                    for (int t = t0; t < te; t++) {
                        const int q = t < cum[0] ? qs[0] + t
                                                 : (t < cum[1] ? qs[1] + t - cum[0] : qs[2] + t - cum[1]);
                        const PosRec &xj = lpos[q];
                        if (xj.id != xi.id && dist2(xj) < a.rl2)
                            hit |= 1ull << (t - t0);
                    }
                    while (hit) {
                        const int t = t0 + __builtin_ctzll(hit);
                        hit &= hit - 1ull;
                        const int q = t < cum[0] ? qs[0] + t
                                                 : (t < cum[1] ? qs[1] + t - cum[0] : qs[2] + t - cum[1]);
                        take(lpos[q], lslot[q], dist2(lpos[q]));
                    }
                }
            } else if (live) {
                for (int dy = 0; dy < 3; dy++) {
                    int r0, r1, r2, r3;
                    row_ranges(a.start + crow[p][3 * ro + dy], n0, cx, r0, r1, r2, r3);
                    global_range(r0, r1);
                    global_range(r2, r3);
                }
            }
            __syncthreads();
        }
    }
    if (live)
        a.nn[s] = n | (kept_excluded ? LIST_EXC : 0);
}

// ------------------------------------------------------- LDS-staged tiles
// A block owns 256 consecutive cell-list positions.  When they lie in one
// replica and one z plane, in x-rows ya..yb, the block stages the 3 (nr + 2)
// halo rows (z-1..z+1 x ya-1..yb+1) -- positions, slots and per-cell offsets --
// into LDS and walks the neighbour cells there; otherwise (replica or plane
// boundary, small grids, over-full tiles) it walks global memory.
constexpr int TILE_CAP = 1536;   // staged particles
constexpr int TILE_ROWCAP = 4096; // staged cell offsets
constexpr int TILE_OWNROWS = 12; // own x-rows per tile
constexpr int TILE_HROWS = 3 * (TILE_OWNROWS + 2);

template <class R, int MODE> __global__ void __launch_bounds__(256) k_tile(KArgs<R> a)
{
    __shared__ int plist[PMAX * 256];
    __shared__ PosRec lpos[TILE_CAP];
    __shared__ int rowoff[TILE_ROWCAP];
    __shared__ int hbase[TILE_HROWS + 1];
    __shared__ int hcell[TILE_HROWS];
    __shared__ int tinfo[8]; // use_lds, ya, nr, z, rep, nh
    const int s0 = blockIdx.x * blockDim.x;
    const int s1 = min(a.N, s0 + (int)blockDim.x);
    if (threadIdx.x == 0) {
        int use = 0;
        const PosRec x0 = a.X[a.list[s0]], x1 = a.X[a.list[s1 - 1]];
        const int rep = x0.meta >> 8;
        if ((x1.meta >> 8) == (uint32_t)rep) {
            const DevRep &rp = a.reps[rep];
            const int n0 = rp.nc[0], n1 = rp.nc[1], n2 = rp.nc[2];
            const int r0 = (a.cell[a.list[s0]] - rp.coff) / n0;
            const int r1 = (a.cell[a.list[s1 - 1]] - rp.coff) / n0;
            const int ya = r0 % n1, z = r0 / n1, yb = r1 % n1;
            const int nr = r1 - r0 + 1;
            if (n0 >= 3 && n1 >= 3 && n2 >= 3 && r1 / n1 == z && nr <= TILE_OWNROWS &&
                nr + 2 <= n1 && 3 * (nr + 2) * (n0 + 1) <= TILE_ROWCAP) {
                use = 1;
                tinfo[1] = ya;
                tinfo[2] = nr;
                tinfo[3] = z;
                tinfo[4] = rep;
                tinfo[5] = 3 * (nr + 2);
            }
            (void)yb;
        }
        tinfo[0] = use;
    }
    __syncthreads();
    if (tinfo[0]) {
        const DevRep &rp = a.reps[tinfo[4]];
        const int n0 = rp.nc[0], n1 = rp.nc[1], n2 = rp.nc[2];
        const int ya = tinfo[1], nr = tinfo[2], z = tinfo[3], nh = tinfo[5];
        // halo row h = (dz + 1) (nr + 2) + (y - ya + 1), first cell of its x-row
        if ((int)threadIdx.x < nh) {
            const int h = threadIdx.x;
            const int dz = h / (nr + 2) - 1, yy = ya - 1 + h % (nr + 2);
            const int zz = z + dz < 0 ? z + dz + n2 : (z + dz >= n2 ? z + dz - n2 : z + dz);
            const int y2 = yy < 0 ? yy + n1 : (yy >= n1 ? yy - n1 : yy);
            const int c0 = rp.coff + n0 * (y2 + n1 * zz);
            hcell[h] = c0;
            hbase[h + 1] = a.start[c0 + n0] - a.start[c0];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            hbase[0] = 0;
            for (int h = 0; h < nh; h++)
                hbase[h + 1] += hbase[h];
            if (hbase[nh] > TILE_CAP)
                tinfo[0] = 0;
        }
        __syncthreads();
    }
    if (tinfo[0]) {
        const DevRep &rp = a.reps[tinfo[4]];
        const int n0 = rp.nc[0];
        const int nh = tinfo[5];
        const int total = hbase[nh];
        // per-cell offsets of the staged rows (LDS index of each cell's first particle)
        for (int t = threadIdx.x; t < nh * (n0 + 1); t += blockDim.x) {
            const int h = t / (n0 + 1), x = t % (n0 + 1);
            rowoff[t] = hbase[h] + a.start[hcell[h] + x] - a.start[hcell[h]];
        }
        // staged particles, flattened over the rows
        for (int f = threadIdx.x; f < total; f += blockDim.x) {
            int lo = 0, hi = nh - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (hbase[mid] <= f)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const int k = a.start[hcell[lo]] + (f - hbase[lo]);
            lpos[f] = a.X[a.list[k]];
        }
        __syncthreads();
        const int s = s0 + threadIdx.x;
        if (s >= s1)
            return;
        const int ya = tinfo[1], nr = tinfo[2];
        float drift_unused = 0.f;
        process_particle<R, MODE, PH_ALL>(a, s, plist, [&](int i, const DevRep &rq, auto &&fn) {
            const uint32_t self = a.X[i].id;
            const int c = a.cell[i] - rq.coff;
            const int cx = c % n0;
            const int cy = (c / n0) % rq.nc[1];
            for (int dz = 0; dz < 3; dz++) {
                for (int dy = 0; dy < 3; dy++) {
                    const int h = dz * (nr + 2) + (cy - ya + dy);
                    const int *ro = rowoff + h * (n0 + 1);
                    int q0, q1, q2 = 0, q3 = 0;
                    if (cx == 0) {
                        q0 = ro[0];
                        q1 = ro[2];
                        q2 = ro[n0 - 1];
                        q3 = ro[n0];
                    } else if (cx == n0 - 1) {
                        q0 = ro[n0 - 2];
                        q1 = ro[n0];
                        q2 = ro[0];
                        q3 = ro[1];
                    } else {
                        q0 = ro[cx - 1];
                        q1 = ro[cx + 2];
                    }
                    // slots are recovered from the id only for patch candidates
                    for (int q = q0; q < q1; q++)
                        if (lpos[q].id != self)
                            fn(lpos[q], -1);
                    for (int q = q2; q < q3; q++)
                        if (lpos[q].id != self)
                            fn(lpos[q], -1);
                }
            }
        }, drift_unused);
        return;
    }
    const int s = s0 + threadIdx.x;
    if (s >= s1)
        return;
    float drift_unused = 0.f;
    process_particle<R, MODE, PH_ALL>(a, s, plist, [&](int i, const DevRep &rp, auto &&fn) {
        for_each_neighbor(a, rp, i, fn);
    }, drift_unused);
}

// first half of a rotor step from the carried F, T (no reordering): half kick,
// drift + wrap (Langevin.cpp:205-223), rotate (LangevinR.cpp:687-695), binning
template <class R> __device__ __forceinline__ float kick_drift_one(const KArgs<R> &a, int i)
{
    constexpr bool FP64 = sizeof(R) == 8;
    PosRec xr = a.X[i];
    const DevRep rp = rep_rec(a, xr.meta);
    const R hdt = (R)a.hdt;
    R pl[3], ll[3], q[9];
    double xn[3] = {xr.x, xr.y, xr.z};
    for (int k = 0; k < 3; k++) {
        pl[k] = a.p[3 * (size_t)i + k] + hdt * a.F[3 * (size_t)i + k];
        ll[k] = a.L[3 * (size_t)i + k] + hdt * a.T[3 * (size_t)i + k];
    }
    load_orient(a.Q, (size_t)i, q);
    double dd = 0;
    // a stopped run (status raised) keeps every particle's state, and a
    // particle whose updated state is non-finite keeps its own
    const bool hold = stopped(a);
    bool ok = !hold;
    for (int k = 0; k < 3; k++) {
        double step;
        if constexpr (FP64)
            step = a.c1 * pl[k];
        else
            step = (double)((R)a.c1 * pl[k]);
        xn[k] = xn[k] + step;
        dd += step * step;
        ok = ok && wrap1(xn[k], pl[k], rp.l[k], a.periodic);
    }
    const R Ir[3] = {(R)a.I[0], (R)a.I[1], (R)a.I[2]};
    rotate((R)a.dt, Ir, ll, q);
    for (int k = 0; k < 3; k++)
        ok = ok && isfinite((double)ll[k]);
    for (int k = 0; k < 9; k++)
        ok = ok && isfinite((double)q[k]);
    if (!ok) {
        if (!hold)
            raise_status(a, 2u);
        xn[0] = xr.x;
        xn[1] = xr.y;
        xn[2] = xr.z;
        for (int k = 0; k < 3; k++) {
            pl[k] = a.p[3 * (size_t)i + k];
            ll[k] = a.L[3 * (size_t)i + k];
        }
        dd = 0;
    }
    for (int k = 0; k < 3; k++) {
        a.pn[3 * (size_t)i + k] = pl[k];
        a.Ln[3 * (size_t)i + k] = ll[k];
    }
    if (ok)
        store_orient(a.Qn, (size_t)i, q);
    else
        copy_orient(a.Qn, (size_t)i, a.Q, (size_t)i);
    xr.x = xn[0];
    xr.y = xn[1];
    xr.z = xn[2];
    a.Xn[i] = xr;
    if (!a.lmode) {
        a.slotn[xr.id] = i;
        bin_particle(a, i, rp, xn, a.celln);
    }
    return (float)sqrt(dd);
}

template <class R> __global__ void __launch_bounds__(256) k_kick_drift(KArgs<R> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float drift = 0.f;
    if (i < a.N)
        drift = kick_drift_one(a, i);
    if (a.lmode)
        drift_bound(a, drift);
}

// binning of the current positions in place (after an upload or a reorder)
template <class R> __global__ void __launch_bounds__(256) k_bin(KArgs<R> a, const StepCtl *gate)
{
    if (gate && !gate->rebuild)
        return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N)
        return;
    const PosRec xr = a.X[i];
    const DevRep rp = rep_rec(a, xr.meta);
    const double x[3] = {xr.x, xr.y, xr.z};
    bin_particle(a, i, rp, x, a.celln);
}

// --------------------------------------------------------------- cell lists
// exclusive scan of count[0..n) into start[0..n], counts zeroed after use.
// phase 1: per block b (1024 elements) local scan and block total
__device__ __forceinline__ void scan1_block(const int *count, int *start, int *bsum, int n, int b)
{
    __shared__ int sh[256];
    const int base = b * 1024 + threadIdx.x * 4;
    int v[4], s = 0;
    for (int k = 0; k < 4; k++) {
        v[k] = base + k < n ? count[base + k] : 0;
        s += v[k];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int t = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
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
        bsum[b] = sh[255];
}

// phase 3: add block offsets, zero the counts, close the array
// (bsum holds the block totals of phase 1; block b of nb sums those before it)
__device__ __forceinline__ void scan3_block(int *count, int *start, const int *bsum, int n, int b,
                                            int nb)
{
    __shared__ int sh[256];
    int part = 0;
    for (int q = threadIdx.x; q < b; q += 256)
        part += bsum[q];
    sh[threadIdx.x] = part;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    const int base = b * 1024 + threadIdx.x * 4;
    const int off = sh[0];
    for (int k = 0; k < 4; k++)
        if (base + k < n) {
            start[base + k] += off;
            count[base + k] = 0;
        }
    if (b == nb - 1 && threadIdx.x == 0)
        start[n] = off + bsum[b];
}

// gate: when non-null, each kernel runs only if a list rebuild was requested
__global__ void __launch_bounds__(256) k_scan1(const int *count, int *start, int *bsum, int n,
                                               const StepCtl *gate)
{
    if (gate && !gate->rebuild)
        return;
    scan1_block(count, start, bsum, n, blockIdx.x);
}

__global__ void __launch_bounds__(256) k_scan3(int *count, int *start, const int *bsum, int n,
                                               const StepCtl *gate)
{
    if (gate && !gate->rebuild)
        return;
    scan3_block(count, start, bsum, n, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void fill_one(const int *cell, const int *rank, const int *start,
                                         int *list, int i)
{
    list[start[cell[i]] + rank[i]] = i;
}

__global__ void __launch_bounds__(256) k_fill(int N, const int *cell, const int *rank,
                                              const int *start, int *list, const StepCtl *gate)
{
    if (gate && !gate->rebuild)
        return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N)
        fill_one(cell, rank, start, list, i);
}

// ascending original index inside each cell: makes the gather order, and so
// the floating-point sums, independent of atomic arrival order and of the
// slot numbering (a state uploaded in original order and the same state held
// in cell order walk their cells, and build their lists, in the same order).
// The particle that arrived first in its cell (rank 0, unique per non-empty
// cell and not changed by the sort) sorts the cell, so empty cells cost
// nothing.
__device__ __forceinline__ void cellsort_one(const int *cell, const int *rank, const int *start,
                                             int *list, const PosRec *X, int i)
{
    if (rank[i] != 0)
        return;
    const int c = cell[i];
    const int s0 = start[c], s1 = start[c + 1];
    for (int s = s0 + 1; s < s1; s++) {
        const int v = list[s];
        const uint32_t idv = X[v].id;
        int t = s - 1;
        while (t >= s0 && X[list[t]].id > idv) {
            list[t + 1] = list[t];
            t--;
        }
        list[t + 1] = v;
    }
}

__global__ void __launch_bounds__(256) k_cellsort(int N, const int *cell, const int *rank,
                                                  const int *start, int *list, const PosRec *X,
                                                  const StepCtl *gate)
{
    if (gate && !gate->rebuild)
        return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N)
        cellsort_one(cell, rank, start, list, X, i);
}

__global__ void k_epoch_inc(StepCtl *ctl)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        ctl->epoch += 1;
}

// the last kernel of a list step that moves nothing (M_STEP): the step's
// force evaluation is counted, and the compacted walk has laid out any new list
__global__ void k_step_end(StepCtl *ctl)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->epoch += 1;
        ctl->fresh = 0u;
    }
}

// a list was just built at the current positions: no rebuild pending, bound 0
__global__ void k_list_reset(StepCtl *ctl)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->rebuild = 0u;
        ctl->D = 0.0;
    }
}

// the grid max of the block drifts of the kernel before is added to the
// displacement bound D; a list rebuild is requested once 2 D reaches the skin
// (a pair within reach now was within reach + skin at the build).
// epoch_inc: also count the step's force evaluation; step_end: the list
// step's kernels are done (the compacted walk has laid out any new list)
__global__ void __launch_bounds__(64) k_decide(StepCtl *ctl, unsigned int *dslot, double skin,
                                               int epoch_inc, int step_end)
{
    // one wave: each lane takes and clears one slot
    float m = __uint_as_float(dslot[threadIdx.x]);
    dslot[threadIdx.x] = 0u;
    for (int o = 32; o > 0; o >>= 1)
        m = fmaxf(m, __shfl_xor(m, o));
    if (threadIdx.x == 0) {
        // margin for the rounding of the drift and of the FP64 position update
        const double D = ctl->D + (double)m * (1 + 1e-6) + 1e-9;
        ctl->D = D;
        if (2 * D >= skin)
            ctl->rebuild = 1u;
        if (epoch_inc)
            ctl->epoch += 1;
        if (step_end)
            ctl->fresh = 0u;
    }
}


__global__ void k_set_call_base(StepCtl *ctl)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        ctl->call_base = ctl->epoch;
}

// ---------------------------------------------------- parity entry points
// positions of the particles with original index in [first, first+count),
// in original order (x only; for the parity paths)
__global__ void k_positions_by_id(const PosRec *X, int N, int first, int count, double *out)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N)
        return;
    PosRec r = X[s];
    int o = (int)r.id - first;
    if (o < 0 || o >= count)
        return;
    out[3 * (size_t)o] = r.x;
    out[3 * (size_t)o + 1] = r.y;
    out[3 * (size_t)o + 2] = r.z;
}

// cube::assign_box with a caller-given n (original order)
__global__ void k_assign_custom(const double *X, int count, double l0, double l1, double l2,
                                int n0, int n1, int n2, int *out)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count)
        return;
    const double *x = X + 3 * (size_t)k;
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

// pairs i < j (original order) of the reference's stencil and predicate:
// count (write = 0) or write each particle's partners in ascending order
__global__ void k_pairs(const double *X, int count, double l0, double l1, double l2, int periodic,
                        int n0, int n1, int n2, double cut, const int *cellv, const int *cstart,
                        const int *clist, int *cnt, const long long *off, int *ij, long long cap,
                        int write)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count)
        return;
    const double l[3] = {l0, l1, l2};
    const int nn[3] = {n0, n1, n2};
    const double *xi = X + 3 * (size_t)k;
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
                    if (dlt(xi, X + 3 * (size_t)j, l, periodic, cut)) {
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

// distance_graph (Analysis/AnalyzeGrowth.cpp:136-177): connected components
// of the pair graph, each particle labelled with the smallest index of its
// component -- the fixed point of the reference's ConnectedComponentsParallel_Old
// (Bindings/GraphAlgorithms.cpp:672-810), which starts from ascending labels
// and only moves a root to a smaller adjacent label.  Here a label always
// names a particle of the same component and never increases, so the order of
// the concurrent updates does not change the fixed point; the host repeats
// hook + jump until neither changes a label.
__global__ void k_cc_init(int *lab, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        lab[i] = i;
}

// each kept pair (both indices outside [lo, hi] when lo <= hi): the larger
// of the two labels is moved to the smaller
__global__ void k_cc_hook(const int *ij, long long np, int lo, int hi, int *lab, int *changed)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np)
        return;
    const int u = ij[2 * k], v = ij[2 * k + 1];
    if (lo <= hi && !((u < lo || u > hi) && (v < lo || v > hi)))
        return;
    const int lu = lab[u], lv = lab[v];
    if (lu == lv)
        return;
    const int m = min(lu, lv), M = max(lu, lv);
    if (atomicMin(&lab[M], m) > m)
        atomicOr(changed, 1);
}

// each label to the root of its chain (labels decrease along a chain, so the
// walk ends within n reads)
__global__ void k_cc_jump(int *lab, int n, int *changed)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int l0 = lab[i];
    int x = l0;
    for (int t = 0; t < n; t++) {
        const int y = lab[x];
        if (y >= x)
            break;
        x = y;
    }
    if (x != l0) {
        lab[i] = x;
        atomicOr(changed, 1);
    }
}

// calcg1 (Analysis/PairCorrelation.cpp:9-155): every unordered pair of
// stencil-adjacent cells once (j > k), the distance of cube::distance
// (geometry.h:107-123), bin floor(dis / dx) counted when below gs.  Counts are
// integers: the atomic order does not change them.
constexpr int GMAXB = 4096;
__global__ void __launch_bounds__(256) k_gcount(const double *X, int count, double l0, double l1,
                                                double l2, int periodic, int n0, int n1, int n2,
                                                const int *cellv, const int *cstart,
                                                const int *clist, double dxb, int gs,
                                                unsigned long long *hist)
{
    __shared__ unsigned int sh[GMAXB];
    for (int b = threadIdx.x; b < gs; b += blockDim.x)
        sh[b] = 0u;
    __syncthreads();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) {
        const double l[3] = {l0, l1, l2};
        const int nn[3] = {n0, n1, n2};
        const double *xi = X + 3 * (size_t)k;
        const int c = cellv[k];
        const int cc[3] = {c % n0, (c / n0) % n1, c / (n0 * n1)};
        int lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = nn[a] >= 3 ? -1 : 0;
            hi[a] = nn[a] >= 2 ? 1 : 0;
        }
        for (int oz = lo[2]; oz <= hi[2]; oz++)
            for (int oy = lo[1]; oy <= hi[1]; oy++)
                for (int ox = lo[0]; ox <= hi[0]; ox++) {
                    int w[3] = {cc[0] + ox, cc[1] + oy, cc[2] + oz};
                    for (int a = 0; a < 3; a++)
                        w[a] = w[a] < 0 ? w[a] + nn[a] : (w[a] >= nn[a] ? w[a] - nn[a] : w[a]);
                    const int c2 = w[0] + n0 * (w[1] + n1 * w[2]);
                    for (int s = cstart[c2]; s < cstart[c2 + 1]; s++) {
                        const int j = clist[s];
                        if (j <= k)
                            continue;
                        const double *xj = X + 3 * (size_t)j;
                        double temp = 0.0;
                        for (int a = 0; a < 3; a++) {
                            double d = xi[a] - xj[a];
                            if (periodic && fabs(d) > l[a] * 0.5)
                                d = d - (d >= 0 ? fabs(l[a]) : -fabs(l[a]));
                            temp += d * d;
                        }
                        const int in = (int)floor(sqrt(temp) / dxb);
                        if (in < gs)
                            atomicAdd(&sh[in], 1u);
                    }
                }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < gs; b += blockDim.x)
        if (sh[b])
            atomicAdd(&hist[b], (unsigned long long)sh[b]);
}

// patch bonds per particle: partners with a patch pair whose Kern-Frenkel
// term is non-zero (eps != 0, r <= 1.4 d and both cone tests passing,
// potentialtheta.h:517-750), decided with patch_pair's FP64 expressions
template <class R>
__global__ void __launch_bounds__(256) k_patch_bonds(KArgs<R> a, int *bonds)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.N)
        return;
    const int i = a.list[s];
    const PosRec xi = a.X[i];
    const int ti = xi.meta & 255;
    const DevRep rp = rep_rec(a, xi.meta);
    int nb = 0;
    if (a.has_patches && a.np[ti] > 0) {
        R qi[9];
        load_orient(a.Q, (size_t)i, qi);
        for_each_neighbor(a, rp, i, [&](const PosRec &xj, int j) {
            double d[3];
            min_image(a, rp, xi.x, xi.y, xi.z, xj, d);
            const double r2s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (!(r2s < a.maxd2))
                return;
            const int tj = xj.meta & 255;
            if (a.np[tj] == 0)
                return;
            int ta, tc;
            const bool i_is_a = patch_roles(xi.id, ti, xj.id, tj, ta, tc);
            if (!((a.active_mask >> (ta * MAXT + tc)) & 1ull))
                return;
            const double rij = sqrt(r2s);
            const double sg = i_is_a ? 1.0 : -1.0;
            const double u[3] = {sg * d[0] / rij, sg * d[1] / rij, sg * d[2] / rij};
            R qj[9];
            load_orient(a.Q, (size_t)j, qj);
            const R *qa = i_is_a ? qi : qj;
            const R *qc = i_is_a ? qj : qi;
            const int npa = a.np[ta], npc = a.np[tc];
            const double *da = a.pdirs + 3 * a.pstart[ta];
            const double *dc = a.pdirs + 3 * a.pstart[tc];
            const int rowbase = rp.prow + a.rstart[ta * MAXT + tc];
            bool bonded = false;
            for (int ia = 0; ia < npa && !bonded; ia++) {
                double n1[3];
                lab_normal_exact(qa, da + 3 * ia, n1);
                const double a1 = -(n1[0] * u[0] + n1[1] * u[1] + n1[2] * u[2]);
                for (int ib = 0; ib < npc && !bonded; ib++) {
                    const PatchRow pr = a.prows[rowbase + ia * npc + ib];
                    if (pr.eps == 0.0 || rij > pr.idist || !(a1 > pr.costm))
                        continue;
                    double n2[3];
                    lab_normal_exact(qc, dc + 3 * ib, n2);
                    const double a2 = (n2[0] * u[0] + n2[1] * u[1] + n2[2] * u[2]);
                    bonded = a2 > pr.costm;
                }
            }
            nb += bonded ? 1 : 0;
        });
    }
    bonds[xi.id] = nb;
}

// ------------------------------------------------------------ reductions
// per-block partial sums (fixed order) of 16 observables: slot range
// ------------------------------------------------ single-stage entry points
// One reference call each, on the current buffer in place (slot order); host
// arrays are staged by original index.

// LangevinNVTR::advancemom_halfstep (LangevinR.cpp:327-347); L may be null
template <class R>
__global__ void __launch_bounds__(256) k_stage_halfkick(int N, const PosRec *X, R *p, R *L,
                                                        const double *F, const double *T, double hdt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const size_t o = X[i].id;
    for (int k = 0; k < 3; k++) {
        p[3 * (size_t)i + k] += (R)hdt * (R)F[3 * o + k];
        if (L)
            L[3 * (size_t)i + k] += (R)hdt * (R)T[3 * o + k];
    }
}

// LangevinNVT::advance_pos (Langevin.cpp:205-223) + cube::correct_position_and_momentum
// (geometry.h:403-445)
template <class R>
__global__ void __launch_bounds__(256) k_stage_advance_pos(int N, PosRec *X, R *p, const DevRep *reps,
                                                           double c1, int periodic, StepCtl *ctl)
{
    constexpr bool FP64 = sizeof(R) == 8;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    PosRec xr = X[i];
    const DevRep rp = reps[xr.meta >> 8];
    double x[3] = {xr.x, xr.y, xr.z};
    R pk[3];
    bool ok = true;
    for (int k = 0; k < 3; k++) {
        pk[k] = p[3 * (size_t)i + k];
        if constexpr (FP64)
            x[k] = x[k] + c1 * pk[k];
        else
            x[k] = x[k] + (double)((R)c1 * pk[k]);
        ok = ok && wrap1(x[k], pk[k], rp.l[k], periodic);
    }
    if (!ok) {
        // non-finite: the particle keeps its position and momentum
        atomicOr(&ctl->status, 2u);
        return;
    }
    for (int k = 0; k < 3; k++)
        p[3 * (size_t)i + k] = pk[k];
    xr.x = x[0];
    xr.y = x[1];
    xr.z = x[2];
    X[i] = xr;
}

// LangevinNVTR::rotate (LangevinR.cpp:595-695)
template <class R>
__global__ void __launch_bounds__(256) k_stage_rotate(int N, R *L, R *Q, double dt, double I0,
                                                      double I1, double I2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    R l[3], q[9];
    for (int k = 0; k < 3; k++)
        l[k] = L[3 * (size_t)i + k];
    load_orient(Q, (size_t)i, q);
    const R Ir[3] = {(R)I0, (R)I1, (R)I2};
    rotate((R)dt, Ir, l, q);
    for (int k = 0; k < 3; k++)
        L[3 * (size_t)i + k] = l[k];
    store_orient(Q, (size_t)i, q);
}

// LangevinNVTR::create_forces_and_torques_sphere (LangevinR.cpp:475-578):
// F (lab) in/out, T lab in -> body out, RT N x 6; all by original index
template <class R>
__global__ void __launch_bounds__(256) k_stage_sphere(int N, const PosRec *X, const R *p, const R *L,
                                                      const R *Q, const DevRep *reps, double *F,
                                                      double *T, const double *RT, double gm,
                                                      double gr0, double gr1, double gr2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const PosRec xr = X[i];
    const size_t o = xr.id;
    const DevRep rp = reps[xr.meta >> 8];
    R q[9];
    load_orient(Q, (size_t)i, q);
    const R pl[3] = {p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]};
    const R ll[3] = {L[3 * (size_t)i], L[3 * (size_t)i + 1], L[3 * (size_t)i + 2]};
    const R pbx = q[0] * pl[0] + q[1] * pl[1] + q[2] * pl[2];
    const R pby = q[3] * pl[0] + q[4] * pl[1] + q[5] * pl[2];
    const R pbz = q[6] * pl[0] + q[7] * pl[1] + q[8] * pl[2];
    const double *nz = RT + 6 * o;
    const R fx = (R)gm * pbx + (R)rp.Rt * (R)nz[0];
    const R fy = (R)gm * pby + (R)rp.Rt * (R)nz[1];
    const R fz = (R)gm * pbz + (R)rp.Rt * (R)nz[2];
    const R tbx = (R)gr0 * ll[0] + (R)rp.Rr * (R)nz[3];
    const R tby = (R)gr1 * ll[1] + (R)rp.Rr * (R)nz[4];
    const R tbz = (R)gr2 * ll[2] + (R)rp.Rr * (R)nz[5];
    const R F0 = (R)F[3 * o], F1 = (R)F[3 * o + 1], F2 = (R)F[3 * o + 2];
    const R T0 = (R)T[3 * o], T1 = (R)T[3 * o + 1], T2 = (R)T[3 * o + 2];
    F[3 * o] = (double)(F0 + fx * q[0] + fy * q[3] + fz * q[6]);
    F[3 * o + 1] = (double)(F1 + fx * q[1] + fy * q[4] + fz * q[7]);
    F[3 * o + 2] = (double)(F2 + fx * q[2] + fy * q[5] + fz * q[8]);
    T[3 * o] = (double)(tbx + T0 * q[0] + T1 * q[1] + T2 * q[2]);
    T[3 * o + 1] = (double)(tby + T0 * q[3] + T1 * q[4] + T2 * q[5]);
    T[3 * o + 2] = (double)(tbz + T0 * q[6] + T1 * q[7] + T2 * q[8]);
}

// LangevinNVT::advance_mom (Langevin.cpp:254-269), coefficients Langevin.cpp:29-53
// (r per replica through its kT)
template <class R>
__global__ void __launch_bounds__(256) k_stage_advance_mom(int N, const PosRec *X, R *p,
                                                           const DevRep *reps, const double *F,
                                                           const double *Rn, double c5c2,
                                                           double c5c3q, double c2, double c5)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const PosRec xr = X[i];
    const size_t o = xr.id;
    const DevRep rp = reps[xr.meta >> 8];
    const R rr = (R)rp.bead_r;
    const R c4 = (R)c2 * rr;
    for (int k = 0; k < 3; k++)
        p[3 * (size_t)i + k] = (R)c5c2 * p[3 * (size_t)i + k] + (R)c5c3q * (R)F[3 * o + k] +
                               ((R)c5 * c4 + rr) * (R)Rn[3 * o + k];
}

// [first, first+count) for p, L; the same range of original indices for obs
template <class R>
__global__ void __launch_bounds__(256) k_obs(int first, int count, const R *p, const R *L,
                                             const double *obs, double m, double I0,
                                             double *partial)
{
    __shared__ double sh[256];
    double acc[16];
    for (int k = 0; k < 16; k++)
        acc[k] = 0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < count; s += gridDim.x * blockDim.x) {
        int i = first + s;
        for (int k = 0; k < 3; k++) {
            double pk = (double)p[3 * (size_t)i + k], lk = L ? (double)L[3 * (size_t)i + k] : 0.0;
            acc[k] += (pk * pk) / (2. * m);
            acc[3 + k] += (lk * lk) / (2. * I0);
        }
        for (int k = 0; k < 10; k++)
            acc[6 + k] += obs[10 * (size_t)i + k];
    }
    for (int k = 0; k < 16; k++) {
        sh[threadIdx.x] = acc[k];
        __syncthreads();
        for (int off = 128; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off)
                sh[threadIdx.x] += sh[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            partial[16 * blockIdx.x + k] = sh[0];
        __syncthreads();
    }
}

// MD::calculatestress (MD.cpp:2381-2424) with the engine's WCA as the MD's
// interaction (HSPotential::force2mdx, potential.h:126-139): per particle the
// half of its pair terms f(r^2) sqrt(r^2) d_k d_j, d = x_i - x_j
// minimum-imaged as distance3v (geometry.h:250-276); every pair within the
// cut-off is visited from both sides (the 27 cells of the current cell list)
template <class R>
__device__ __forceinline__ void stress_terms(const KArgs<R> &a, int i, const PosRec &xi, double st[9])
{
    const DevRep rp = rep_rec(a, xi.meta);
    const double pl2 = a.wca_pl * a.wca_pl;
    for_each_neighbor(a, rp, i, [&](const PosRec &xj, int j) {
        (void)j;
        double d[3];
        min_image(a, rp, xi.x, xi.y, xi.z, xj, d);
        const double dis = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (!(dis < pl2))
            return;
        const double x3 = dis * dis * dis;
        const double f1 = 4 * a.wca_eps * (12 * a.wca_s12 / (dis * (x3 * x3)) - 6 * a.wca_s6 / (dis * x3));
        for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++)
                st[3 * k + l] += 0.5 * (f1 * sqrt(dis) * d[k] * d[l]);
    });
}

template <class R>
__global__ void __launch_bounds__(256) k_stress(KArgs<R> a, double *out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.N)
        return;
    const int i = a.list[s];
    const PosRec xi = a.X[i];
    double st[9];
    for (int k = 0; k < 9; k++)
        st[k] = 0;
    if (a.wca)
        stress_terms(a, i, xi, st);
    for (int k = 0; k < 9; k++)
        out[9 * (size_t)xi.id + k] = st[k];
}

// the same stress terms summed per block (grid-stride over the cell list,
// particles with original index in [first, first + count)), block sums in
// a fixed order into partial[9 * block + k]
template <class R>
__global__ void __launch_bounds__(256) k_stress_partial(KArgs<R> a, int first, int count,
                                                        double *partial)
{
    __shared__ double sh[256];
    double st[9];
    for (int k = 0; k < 9; k++)
        st[k] = 0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < a.N; s += gridDim.x * blockDim.x) {
        const int i = a.list[s];
        const PosRec xi = a.X[i];
        if (a.wca && (int)xi.id >= first && (int)xi.id < first + count)
            stress_terms(a, i, xi, st);
    }
    for (int k = 0; k < 9; k++) {
        sh[threadIdx.x] = st[k];
        __syncthreads();
        for (int off = 128; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off)
                sh[threadIdx.x] += sh[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            partial[9 * (size_t)blockIdx.x + k] = sh[0];
        __syncthreads();
    }
}

// out[k] = sum over b < nb of partial[9 b + k], in block order (one wave)
__global__ void __launch_bounds__(64) k_sum_partials9(const double *partial, int nb, double *out)
{
    const int k = threadIdx.x;
    if (k >= 9)
        return;
    double v = 0;
    for (int b = 0; b < nb; b++)
        v += partial[9 * (size_t)b + k];
    out[k] = v;
}

// per-block sums (fixed order) of width-value rows first..first+count of in
template <int W>
__global__ void __launch_bounds__(256) k_sum_rows(int first, int count, const double *in,
                                                  double *partial)
{
    __shared__ double sh[256];
    double acc[W];
    for (int k = 0; k < W; k++)
        acc[k] = 0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < count; s += gridDim.x * blockDim.x)
        for (int k = 0; k < W; k++)
            acc[k] += in[(size_t)W * (first + s) + k];
    for (int k = 0; k < W; k++) {
        sh[threadIdx.x] = acc[k];
        __syncthreads();
        for (int off = 128; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off)
                sh[threadIdx.x] += sh[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            partial[W * blockIdx.x + k] = sh[0];
        __syncthreads();
    }
}

// precision conversion (device <-> host FP64 layout)
template <class A, class B> __global__ void k_convert(const A *in, B *out, size_t n)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        out[k] = (B)in[k];
}

// a host array by original index (width values per particle, FP64) into the
// slot order of the current buffer
template <class R>
__global__ void k_by_id_to_slots(const PosRec *X, int N, const double *in, R *out, int width)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N)
        return;
    const size_t o = X[s].id;
    for (int k = 0; k < width; k++)
        out[(size_t)width * s + k] = (R)in[width * o + k];
}

// host matrices (FP64, 9 per particle, slot order) into the mixed engine's
// quaternions; Shepperd's method in FP64, rounded once
__global__ void k_mats_to_quats(const double *in, int N, float *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    double m[9], q[4];
    for (int k = 0; k < 9; k++)
        m[k] = in[9 * (size_t)i + k];
    mat_to_quat<double>(m, q);
    reinterpret_cast<float4 *>(out)[i] = make_float4((float)q[0], (float)q[1], (float)q[2], (float)q[3]);
}

// the mixed engine's quaternions as FP64 rotation matrices (9 per particle,
// slot order), normalised in FP64 first (the stored FP32 norm is 1 only to
// FP32 rounding)
__global__ void k_quats_to_mats(const float *in, int N, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const float4 q = reinterpret_cast<const float4 *>(in)[i];
    const double w = q.x, x = q.y, y = q.z, z = q.w;
    const double inv = 1.0 / sqrt(w * w + x * x + y * y + z * z);
    double m[9];
    quat_to_mat<double>(w * inv, x * inv, y * inv, z * inv, m);
    for (int k = 0; k < 9; k++)
        out[9 * (size_t)i + k] = m[k];
}

// new (replica << 8 | type) words for the current slots from an array by
// original index
__global__ void k_set_meta(PosRec *X, int N, const uint32_t *meta_by_id)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < N)
        X[s].meta = meta_by_id[X[s].id];
}

} // namespace mdx
