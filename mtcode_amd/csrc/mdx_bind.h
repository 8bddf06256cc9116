// One-bond-per-patch binding: the update of the bound state and the forces
// of the bound pairs, LangevinNVTR::calculate_forces_and_torques3D_onlyone_nonlets
// (Rotational/LangevinRforce.cpp:2874-4136) with BindingModelSingle
// (Bindings/BindingModelSingle.cpp:1-119).
//
// Patches are numbered per particle in original-index order (poff[i] is the
// number of the first patch of particle i, its patches follow in type order),
// the numbering of GeneralPatch::which_patch (GeneralPatch.cpp:292-307) for
// particles sorted by type.  Four kernels per call (after a clear of the
// per-patch candidate counts):
//   k_bind_screen  one thread per particle: the neighbours within the largest
//                  interaction distance whose patches pass an FP32 cone
//                  screen on both sides (a superset of the candidates) --
//                  in the step the pair walk collects them instead (its own
//                  side's screen);
//   k_bind_edges   one thread per particle with screened neighbours: its
//                  candidate patch pairs (both cones, a distance bound of 1,
//                  0.7 or 0.5 x the interaction distance by bound state, a
//                  score) decided in FP64, each appended to the row of the
//                  thread's own patch -- no atomics, both particles of a pair
//                  compute the same FP64 values -- and each row sorted by
//                  partner;
//   k_bind_decide  one thread per patch: a search of its component of the
//                  candidate graph that stops at the first smaller patch, so
//                  only the component's smallest patch goes on; it breaks
//                  its members' bonds to non-candidates and decides the
//                  component (1 patch: unbound; 2 and 3: the model's doublet
//                  and triplet rules; more: the highest-scored edge removed
//                  until no part exceeds 3, each part decided in order of its
//                  smallest patch) with Philox draws keyed by the smallest
//                  patch, the call's draw counter and the decision's index;
//   k_bind_forces  one thread per particle: the Kern-Frenkel force and
//                  torques of its bound patches on itself.
// Every cut-off, cone and bound-state decision is taken in FP64 on the
// reference's own expressions.
#pragma once

#include "mdx_kernels.h"

namespace mdx {

constexpr int BIND_DEG = 8;  // candidate partners kept per patch
constexpr int BIND_CMAX = 64; // patches per component

struct BindArgs {
    const int *poff;  // first patch of each original index (N + 1)
    const int *owner; // original index of each patch
    int *isbound, *boundto;
    int *adj_n;       // candidates per patch
    int *adj_p;       // partner patches, BIND_DEG per patch, ascending
    double *adj_s;    // their scores
    int nh;
    double mk2;       // squared largest interaction distance (max_check)
    double cmin;      // smallest cos(theta_m) of any row (the cone screen)
    double on_rate, off_rate;
    uint32_t key0, key1;
    unsigned long long draw;
    int draw_epoch; // 1: the draw counter is the force-evaluation count (steps)
    StepCtl *ctl;
    // 0: BindingModelSingle(on_rate, off_rate); 1: the species-table form of
    // BindingModelBinary / BindingModelTernary: species[patch] in 1..3,
    // dtab[pair * 4 + 2 * before + k] (k 0: rate to unbound, 1: to bound) by
    // sorted species pair 11,12,13,22,23,33, ttab[triple * 16 + 4 * from + to]
    // (from: bound 12, 23, 13, none; to: 12, 23, 13) and tr4[triple * 4 + from]
    // (to none) by sorted species triple 111,112,113,122,123,133,222,223,233,333
    int table;
    const int *species;
    const double *dtab, *ttab, *tr4;
};

// status bit 3: a patch with more than BIND_DEG candidates, or a component
// of more than BIND_CMAX patches
__device__ __forceinline__ void bind_status(const BindArgs &b, unsigned bit)
{
    atomicOr(&b.ctl->status, bit);
    atomicMin(&b.ctl->fail_epoch, b.ctl->epoch);
}

// 53-bit uniform in [0, 1) of Philox4x32-10 at {root, draw lo, draw hi, k}
__device__ __forceinline__ double bind_uniform(const BindArgs &b, int root, int k)
{
    const unsigned long long draw = b.draw_epoch ? b.ctl->epoch : b.draw;
    const U4 o = philox4x32_10(
        U4{(uint32_t)root, (uint32_t)draw, (uint32_t)(draw >> 32), (uint32_t)k}, b.key0, b.key1);
    return ((double)(o.a >> 5) * 67108864.0 + (double)(o.b >> 6)) * (1.0 / 9007199254740992.0);
}

// the patches of a particle (orientation q, type t) whose cone value
// sign * (n . u) passes the FP32 screen: n . u = nb . (Q u), against the
// smallest cos(theta_m) of any row less 1e-5 (well above the FP32 error of
// these unit-vector products), so a patch inside its own row's cone always
// passes; sign -1 for the particle in role "a" (its value is -(n . u))
template <class R>
__device__ __forceinline__ uint32_t bind_cone_mask(const KArgs<R> &a, const BindArgs &b,
                                                   const R *q, int t, const double u[3], float sign)
{
    const float uf[3] = {(float)u[0], (float)u[1], (float)u[2]};
    float v[3];
    for (int r = 0; r < 3; r++)
        v[r] = sign * ((float)q[3 * r] * uf[0] + (float)q[3 * r + 1] * uf[1] +
                       (float)q[3 * r + 2] * uf[2]);
    const float *f = a.pdirs_f + 3 * a.pstart[t];
    const float scr_f = (float)(b.cmin - 1e-5);
    uint32_t m = 0;
    for (int k = 0; k < a.np[t]; k++)
        if (f[3 * k] * v[0] + f[3 * k + 1] * v[1] + f[3 * k + 2] * v[2] > scr_f)
            m |= 1u << k;
    return m;
}

// the neighbours of slot i within the binding radius whose patches pass
// the cone screen on both sides (bind_cone_mask): fn(xj, j) for each
template <class R, class Fn>
__device__ __forceinline__ void bind_screened(const KArgs<R> &a, const BindArgs &b, const DevRep &rp,
                                              int i, const PosRec &xi, int ti, const R *qi,
                                              Fn &&fn)
{
    auto visit = [&](const PosRec &xj, int j) {
        double d[3];
        min_image(a, rp, xi.x, xi.y, xi.z, xj, d);
        const double r2s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (!(r2s < b.mk2))
            return;
        const int tj = xj.meta & 255;
        if (a.np[tj] == 0)
            return;
        int ta, tc;
        const bool i_is_a = patch_roles(xi.id, ti, xj.id, tj, ta, tc);
        const double dis = sqrt(r2s);
        const double sg = i_is_a ? 1.0 : -1.0;
        const double u[3] = {sg * d[0] / dis, sg * d[1] / dis, sg * d[2] / dis};
        R qj[9];
        load_orient(a.Q, (size_t)j, qj);
        if (bind_cone_mask(a, b, i_is_a ? qi : qj, ta, u, -1.0f) &&
            bind_cone_mask(a, b, i_is_a ? qj : qi, tc, u, 1.0f))
            fn(xj, j);
    };
    // the neighbour list in list steps (its loads batched), else the 27 cells
    for_each_near(a, rp, i, visit);
}

constexpr int BIND_SCR = 8; // screened neighbours held per particle

// the candidate counts of every patch cleared (coalesced), before the edges
__global__ void __launch_bounds__(256) k_bind_clear(int *adj_n, int nh)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nh)
        adj_n[p] = 0;
}

// pass 1 (one call on a frozen state): the screened neighbours of each
// cell-list position s into scr[q * N + s], q < BIND_SCR, their count in
// nscr[s] (beyond BIND_SCR the edge kernel screens the neighbours again)
template <class R>
__global__ void __launch_bounds__(256) k_bind_screen(KArgs<R> a, BindArgs b, int *scr, int *nscr)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.N)
        return;
    const int i = a.lmode ? s : a.list[s];
    const PosRec xi = a.X[i];
    const int ti = xi.meta & 255;
    const int npi = a.np[ti];
    int cnt = 0;
    if (npi > 0) {
        const DevRep rp = rep_rec(a, xi.meta);
        R qi[9];
        load_orient(a.Q, (size_t)i, qi);
        bind_screened(a, b, rp, i, xi, ti, qi, [&](const PosRec &, int j) {
            if (cnt < BIND_SCR)
                scr[(size_t)cnt * a.N + s] = j;
            cnt++;
        });
    }
    nscr[s] = cnt;
}

// pass 2, candidate patch pairs (LangevinRforce.cpp:2914-3075): for each
// screened neighbour the FP64 cone values of the patches, decided on the
// row's own threshold, the interaction distance and then the bound-state
// distance (which never exceeds it) -- the reference's decision -- with the
// score; each candidate appended to the row of the thread's own patch (no
// atomics: both particles of a pair compute the same FP64 values), rows
// sorted by partner.  Per-patch counts are packed 4 bits each.
// scr, nscr: the screened neighbours (slots) of each position and their
// count, cap of them held (the screen kernel's, or the candidates the pair
// walk collected in steps, per cell-list position); beyond cap the thread
// screens its neighbours again itself
template <class R>
__device__ __forceinline__ void bind_edges_one(const KArgs<R> &a, const BindArgs &b, const int *scr,
                                               int cnt, int cap, int s)
{
    static_assert(MAXP * 4 <= 32 && BIND_DEG < 16, "packed candidate counts");
    const int i = a.lmode ? s : a.list[s];
    const PosRec xi = a.X[i];
    const int ti = xi.meta & 255;
    const int npi = a.np[ti];
    const int base = b.poff[xi.id];
    uint32_t degs = 0; // candidates of own patch k in bits 4k..4k+3
    const DevRep rp = rep_rec(a, xi.meta);
    R qi[9];
    load_orient(a.Q, (size_t)i, qi);
    bool over = false;
    auto exact = [&](const PosRec &xj, int j) {
        double d[3];
        min_image(a, rp, xi.x, xi.y, xi.z, xj, d);
        const double r2s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const int tj = xj.meta & 255;
        int ta, tc;
        const bool i_is_a = patch_roles(xi.id, ti, xj.id, tj, ta, tc);
        const double dis = sqrt(r2s);
        const double sg = i_is_a ? 1.0 : -1.0;
        const double u[3] = {sg * d[0] / dis, sg * d[1] / dis, sg * d[2] / dis};
        R qj[9];
        load_orient(a.Q, (size_t)j, qj);
        const R *qa = i_is_a ? qi : qj;
        const R *qc = i_is_a ? qj : qi;
        const int npa = a.np[ta], npc = a.np[tc];
        // the FP32 screen of both sides first (as in bind_screened)
        const uint32_t ma = bind_cone_mask(a, b, qa, ta, u, -1.0f);
        if (!ma)
            return;
        const uint32_t mc = bind_cone_mask(a, b, qc, tc, u, 1.0f);
        if (!mc)
            return;
        const double *da = a.pdirs + 3 * a.pstart[ta];
        const double *dc = a.pdirs + 3 * a.pstart[tc];
        const int rowbase = rp.prow + a.rstart[ta * MAXT + tc];
        const int pj = b.poff[xj.id];
        for (uint32_t m1 = ma; m1; m1 &= m1 - 1) {
            const int ia = __builtin_ctz(m1);
            double n1[3];
            lab_normal_exact(qa, da + 3 * ia, n1);
            const double argthetai = -(n1[0] * u[0] + n1[1] * u[1] + n1[2] * u[2]);
            if (!(argthetai > b.cmin))
                continue;
            for (uint32_t m2 = mc; m2; m2 &= m2 - 1) {
                const int ic = __builtin_ctz(m2);
                double n2[3];
                lab_normal_exact(qc, dc + 3 * ic, n2);
                const double argthetaj = (n2[0] * u[0] + n2[1] * u[1] + n2[2] * u[2]);
                const PatchRow &pr = a.prows[rowbase + ia * npc + ic];
                const double ctm = pr.costm, disp = pr.idist;
                if (!(argthetai > ctm && argthetaj > ctm && dis < disp))
                    continue;
                const int wp1 = (i_is_a ? base : pj) + ia, wp2 = (i_is_a ? pj : base) + ic;
                const bool cond1 = b.boundto[wp1] == wp2 && b.boundto[wp2] == wp1;
                const bool b1 = b.isbound[wp1] != 0, b2 = b.isbound[wp2] != 0;
                double disp2;
                if (b1 && b2 && cond1)
                    disp2 = disp;
                else if (b1 && b2 && !cond1)
                    disp2 = 0.5 * disp;
                else if (!b1 != !b2)
                    disp2 = 0.7 * disp;
                else
                    disp2 = disp;
                if (!(dis < disp2))
                    continue;
                const double scr_v =
                    (1 - (argthetai - ctm)) + (1 - (argthetaj - ctm)) + 2 * (dis / disp2);
                const int own = i_is_a ? ia : ic;
                const int partner = i_is_a ? wp2 : wp1;
                const int dg = (degs >> (4 * own)) & 15u;
                if (dg < BIND_DEG) {
                    const size_t at = (size_t)(base + own) * BIND_DEG + dg;
                    b.adj_p[at] = partner;
                    b.adj_s[at] = scr_v;
                    degs += 1u << (4 * own);
                } else {
                    over = true;
                }
            }
        }
    };
    if (cnt <= cap) {
        for (int q = 0; q < cnt; q++) {
            const int j = scr[(size_t)q * a.N + s];
            exact(a.X[j], j);
        }
    } else {
        bind_screened(a, b, rp, i, xi, ti, qi, exact);
    }
    if (over)
        bind_status(b, 8u);
    for (int k = 0; k < npi; k++) {
        const int dg = (degs >> (4 * k)) & 15u;
        const size_t r0 = (size_t)(base + k) * BIND_DEG;
        for (int q = 1; q < dg; q++) // insertion sort by partner
            for (int t = q; t > 0 && b.adj_p[r0 + t - 1] > b.adj_p[r0 + t]; t--) {
                const int tp = b.adj_p[r0 + t];
                b.adj_p[r0 + t] = b.adj_p[r0 + t - 1];
                b.adj_p[r0 + t - 1] = tp;
                const double ts = b.adj_s[r0 + t];
                b.adj_s[r0 + t] = b.adj_s[r0 + t - 1];
                b.adj_s[r0 + t - 1] = ts;
            }
        if (dg)
            b.adj_n[base + k] = dg;
    }
}

// the positions of a block with screened neighbours are compacted to its
// leading lanes first (a ballot and a prefix over the waves), so that the
// few that have any run together
template <class R>
__global__ void __launch_bounds__(256) k_bind_edges(KArgs<R> a, BindArgs b, const int *scr,
                                                    const int *nscr, int cap)
{
    __shared__ int off[256], num[256];
    __shared__ int wsum[4];
    const int tid = threadIdx.x;
    const int s = blockIdx.x * 256 + tid;
    const int c = s < a.N ? nscr[s] : 0;
    const unsigned long long m = __ballot(c > 0);
    const int lane = tid & 63;
    const int rank = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0)
        wsum[tid >> 6] = __popcll(m);
    __syncthreads();
    int wb = 0;
    for (int w = 0; w < (tid >> 6); w++)
        wb += wsum[w];
    if (c > 0) {
        off[wb + rank] = s;
        num[wb + rank] = c;
    }
    __syncthreads();
    const int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    for (int t = tid; t < total; t += 256)
        bind_edges_one(a, b, scr, num[t], cap, off[t]);
}

__device__ __forceinline__ void bind_pair(const BindArgs &b, int i1, int i2)
{
    b.boundto[i1] = i2;
    b.boundto[i2] = i1;
    b.isbound[i1] = 1;
    b.isbound[i2] = 1;
}

__device__ __forceinline__ int bind_pair_code(int a, int c) // sorted species, 1..3
{
    return a == 1 ? c - 1 : a == 2 ? c + 1 : 5;
}

__device__ __forceinline__ int bind_triple_code(int a, int c, int e) // sorted
{
    if (a == 1)
        return c == 1 ? e - 1 : c == 2 ? (e == 2 ? 3 : 4) : 5;
    if (a == 2)
        return c == 2 ? (e == 2 ? 6 : 7) : 8;
    return 9;
}

// BindingModelSingle::doublet (BindingModelSingle.cpp:10-25), or the
// species-table form (BindingModelTernary.cpp:523-558, BindingModelBinary.cpp:
// 97-166), on i1 < i2 (LangevinRforce.cpp:3262-3302)
__device__ __forceinline__ void bind_decide2(const BindArgs &b, int i1, int i2, double ran)
{
    bool bind;
    if (b.table) {
        const bool before =
            b.boundto[i1] == i2 && b.boundto[i2] == i1 && b.isbound[i1] && b.isbound[i2];
        const int s1 = b.species[i1], s2 = b.species[i2];
        const double *r = b.dtab + 4 * bind_pair_code(s1 < s2 ? s1 : s2, s1 < s2 ? s2 : s1);
        const int bb = before ? 1 : 0;
        const double r1 = r[bb * 2 + 0];
        const double r2 = r[bb * 2 + 1];
        const double totalr = r1 + r2;
        const double rr = totalr * ran;
        bind = !(rr < r1);
    } else {
        const double totalr = b.off_rate + b.on_rate;
        const double rr = totalr * ran;
        bind = !(rr < b.off_rate);
    }
    if (bind) {
        bind_pair(b, i1, i2);
    } else {
        b.isbound[i1] = 0;
        b.isbound[i2] = 0;
    }
}

// sort_triplet_and_save_permutation (basic.h:166-245): the sorted species
// and the permutation code (p1 * 10 + p2)
__device__ __forceinline__ int bind_sort3(int t1, int t2, int t3, int &a, int &c, int &e)
{
    if (t1 <= t2) {
        if (t2 <= t3) {
            a = t1; c = t2; e = t3;
            return 12;
        }
        if (t1 <= t3) {
            a = t1; c = t3; e = t2;
            return 13;
        }
        a = t3; c = t1; e = t2;
        return 31;
    }
    if (t1 <= t3) {
        a = t2; c = t1; e = t3;
        return 21;
    }
    if (t2 <= t3) {
        a = t2; c = t3; e = t1;
        return 23;
    }
    a = t3; c = t2; e = t1;
    return 32;
}

// save_to_type (basic.h:373-419): {12, 23, 13} into the sorted frame
__device__ __forceinline__ void bind_to_type(const bool *x, int o, bool *y)
{
    switch (o) {
    case 12: y[0] = x[0]; y[1] = x[1]; y[2] = x[2]; break;
    case 13: y[0] = x[2]; y[1] = x[1]; y[2] = x[0]; break;
    case 21: y[0] = x[0]; y[1] = x[2]; y[2] = x[1]; break;
    case 23: y[0] = x[1]; y[2] = x[0]; y[1] = x[2]; break;
    case 31: y[0] = x[1]; y[1] = x[2]; y[2] = x[0]; break;
    default: y[1] = x[0]; y[0] = x[1]; y[2] = x[2]; break;
    }
}

// inverse_save_to_type (basic.h:421-468)
__device__ __forceinline__ void bind_from_type(const bool *x, int o, bool *y)
{
    switch (o) {
    case 12: y[0] = x[0]; y[1] = x[1]; y[2] = x[2]; break;
    case 13: y[2] = x[0]; y[1] = x[1]; y[0] = x[2]; break;
    case 21: y[0] = x[0]; y[2] = x[1]; y[1] = x[2]; break;
    case 23: y[1] = x[0]; y[0] = x[2]; y[2] = x[1]; break;
    case 31: y[2] = x[0]; y[0] = x[1]; y[1] = x[2]; break;
    default: y[1] = x[0]; y[0] = x[1]; y[2] = x[2]; break;
    }
}

// BindingModelSingle::triplet (BindingModelSingle.cpp:40-119), or the
// species-table form (BindingModelTernary.cpp:791-870; the binary model's
// is its identity-permutation case), on i1 < i2 < i3 with connectivity c
// (LangevinRforce.cpp:3303-3497)
__device__ __forceinline__ void bind_decide3(const BindArgs &b, int i1, int i2, int i3, bool c12,
                                             bool c23, bool c13, double ran)
{
    const int *bt = b.boundto, *ib = b.isbound;
    const bool b12 = bt[i1] == i2 && bt[i2] == i1 && ib[i1] && ib[i2];
    const bool b23 = bt[i2] == i3 && bt[i3] == i2 && ib[i2] && ib[i3];
    const bool b13 = bt[i1] == i3 && bt[i3] == i1 && ib[i1] && ib[i3];
    if (b.table) {
        int sa, sc, se;
        const int o = bind_sort3(b.species[i1], b.species[i2], b.species[i3], sa, sc, se);
        const bool bb[3] = {b12, b23, b13}, cc[3] = {c12, c23, c13};
        bool tb[3], tc[3];
        bind_to_type(bb, o, tb);
        bind_to_type(cc, o, tc);
        const int f = tb[0] ? 0 : tb[1] ? 1 : tb[2] ? 2 : 3;
        const int code = bind_triple_code(sa, sc, se);
        const double *r = b.ttab + 16 * code;
        const double r1 = tc[0] * r[f * 4 + 0];
        const double r2 = tc[1] * r[f * 4 + 1];
        const double r3 = tc[2] * r[f * 4 + 2];
        const double r4 = b.tr4[4 * code + f];
        const double totalr = r1 + r2 + r3 + r4;
        const double rr = totalr * ran;
        bool ta[3] = {false, false, false}, a[3];
        if (rr < r1)
            ta[0] = true;
        else if (rr >= r1 && rr < r1 + r2)
            ta[1] = true;
        else if (rr >= r1 + r2 && rr < r1 + r2 + r3)
            ta[2] = true;
        bind_from_type(ta, o, a);
        if (a[0]) {
            bind_pair(b, i1, i2);
            b.isbound[i3] = 0;
        } else if (a[1]) {
            bind_pair(b, i2, i3);
            b.isbound[i1] = 0;
        } else if (a[2]) {
            bind_pair(b, i1, i3);
            b.isbound[i2] = 0;
        } else {
            b.isbound[i1] = 0;
            b.isbound[i2] = 0;
            b.isbound[i3] = 0;
        }
        return;
    }
    const double on = b.on_rate, off = b.off_rate;
    double r1, r2, r3, r4;
    if (b12) {
        r1 = c12 * on;
        r2 = c23 * ((off / 3.) + 0.001);
        r3 = c13 * ((off / 3.) + 0.001);
        r4 = (off / 3.) + 0.001;
    } else if (b23) {
        r1 = c12 * ((off / 3.) + 0.001);
        r2 = c23 * (on);
        r3 = c13 * ((off / 3.) + 0.001);
        r4 = ((off / 3.) + 0.001);
    } else if (b13) {
        r1 = c12 * ((off / 3.) + 0.001);
        r2 = c23 * ((off / 3.) + 0.001);
        r3 = c13 * on;
        r4 = (off / 3.) + 0.001;
    } else {
        r1 = c12 * on / 3.;
        r2 = c23 * on / 3.;
        r3 = c13 * on / 3.;
        r4 = off;
    }
    const double totalr = r1 + r2 + r3 + r4;
    const double rr = totalr * ran;
    if (rr < r1) {
        bind_pair(b, i1, i2);
        b.isbound[i3] = 0;
    } else if (rr >= r1 && rr < r1 + r2) {
        bind_pair(b, i2, i3);
        b.isbound[i1] = 0;
    } else if (rr >= r1 + r2 && rr < r1 + r2 + r3) {
        bind_pair(b, i1, i3);
        b.isbound[i2] = 0;
    } else {
        b.isbound[i1] = 0;
        b.isbound[i2] = 0;
        b.isbound[i3] = 0;
    }
}

// connectivity of a three-patch component from the first one's degree and
// first neighbour and the second one's degree; false if malformed
__device__ __forceinline__ bool bind_conn3(int nb1, int first1, int nb2, int i2, int i3, bool &c12,
                                           bool &c23, bool &c13)
{
    if (nb1 == 1) {
        if (first1 == i2) {
            c12 = true;
            c13 = false;
            c23 = true;
        } else if (first1 == i3) {
            c13 = true;
            c12 = false;
            c23 = true;
        } else {
            return false;
        }
    } else if (nb1 == 2) {
        c12 = true;
        c13 = true;
        c23 = nb2 != 1;
    } else {
        return false;
    }
    return true;
}

__device__ __forceinline__ int uf_find(unsigned char *par, int x)
{
    while (par[x] != x)
        x = par[x] = par[par[x]];
    return x;
}

// a component of more than three patches (LangevinRforce.cpp:3531-3870):
// mem (ascending, s of them) are its patches, root the smallest
__device__ void bind_large(const BindArgs &b, const int *mem, int s, int root)
{
    constexpr int EMAX = BIND_CMAX * BIND_DEG / 2;
    unsigned char ea[EMAX], eb[EMAX];
    double es[EMAX];
    int ne = 0;
    // the component's edges in local indices (RecapitulateEdges,
    // LangevinRforceUtil.cpp:508-575): members in order, each one's partners
    // in order, each unordered edge once
    for (int j = 0; j < s; j++) {
        const int g = mem[j];
        const int dg = b.adj_n[g];
        for (int k = 0; k < dg; k++) {
            const int pg = b.adj_p[(size_t)g * BIND_DEG + k];
            int lo = 0, hi = s - 1, f2 = -1;
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                if (mem[mid] == pg) {
                    f2 = mid;
                    break;
                }
                if (mem[mid] < pg)
                    lo = mid + 1;
                else
                    hi = mid - 1;
            }
            if (f2 < 0) // (every partner of a member is a member)
                continue;
            const int x = j < f2 ? j : f2, y = j < f2 ? f2 : j;
            bool already = false;
            for (int q = 0; q < ne; q++)
                if (ea[q] == x && eb[q] == y) {
                    already = true;
                    break;
                }
            if (!already && ne < EMAX) {
                ea[ne] = (unsigned char)x;
                eb[ne] = (unsigned char)y;
                es[ne] = b.adj_s[(size_t)g * BIND_DEG + k];
                ne++;
            }
        }
    }
    unsigned char par[BIND_CMAX];
    for (;;) {
        // drop the first highest-scored edge (std::max_element), once before
        // the first test and then while a part exceeds 3 patches
        int best = 0;
        for (int q = 1; q < ne; q++)
            if (es[best] < es[q])
                best = q;
        for (int q = best; q + 1 < ne; q++) {
            ea[q] = ea[q + 1];
            eb[q] = eb[q + 1];
            es[q] = es[q + 1];
        }
        ne--;
        for (int j = 0; j < s; j++)
            par[j] = (unsigned char)j;
        for (int q = 0; q < ne; q++) {
            const int r1 = uf_find(par, ea[q]), r2 = uf_find(par, eb[q]);
            if (r1 < r2)
                par[r2] = (unsigned char)r1;
            else if (r2 < r1)
                par[r1] = (unsigned char)r2;
        }
        unsigned char cnt[BIND_CMAX];
        for (int j = 0; j < s; j++)
            cnt[j] = 0;
        int big = 0;
        for (int j = 0; j < s; j++) {
            const int r = uf_find(par, j);
            cnt[r]++;
            big = cnt[r] > big ? cnt[r] : big;
        }
        if (big <= 3)
            break;
    }
    // local degrees and first neighbours of what is left (PairHistogram)
    unsigned char lens[BIND_CMAX], first_nb[BIND_CMAX];
    for (int j = 0; j < s; j++)
        lens[j] = 0;
    for (int q = 0; q < ne; q++) {
        if (lens[ea[q]]++ == 0)
            first_nb[ea[q]] = eb[q];
        if (lens[eb[q]]++ == 0)
            first_nb[eb[q]] = ea[q];
    }
    // bonds to patches that are no longer neighbours break
    for (int k1 = 0; k1 < s; k1++) {
        const int g = mem[k1];
        if (b.isbound[g]) {
            bool outside = true;
            for (int q = 0; q < ne; q++)
                if ((ea[q] == k1 && mem[eb[q]] == b.boundto[g]) ||
                    (eb[q] == k1 && mem[ea[q]] == b.boundto[g])) {
                    outside = false;
                    break;
                }
            if (outside)
                b.isbound[g] = 0;
        }
    }
    // the parts in order of their smallest local index
    int it = 0;
    for (int j = 0; j < s; j++) {
        if (uf_find(par, j) != j)
            continue;
        int m[3], z = 0;
        for (int q = j; q < s && z < 3; q++)
            if (uf_find(par, q) == j)
                m[z++] = q;
        if (z == 1) {
            b.isbound[mem[m[0]]] = 0;
        } else if (z == 2) {
            bind_decide2(b, mem[m[0]], mem[m[1]], bind_uniform(b, root, 1 + it));
            it++;
        } else {
            bool c12 = false, c23 = false, c13 = false;
            const int ii1 = m[0], ii2 = m[1], ii3 = m[2];
            if (!bind_conn3(lens[ii1], lens[ii1] > 0 ? first_nb[ii1] : -1, lens[ii2], ii2, ii3,
                            c12, c23, c13)) {
                bind_status(b, 8u);
                return;
            }
            bind_decide3(b, mem[ii1], mem[ii2], mem[ii3], c12, c23, c13,
                         bind_uniform(b, root, 1 + it));
            it++;
        }
    }
}

// bound-state decisions per component (LangevinRforce.cpp:3098-3870)
__global__ void __launch_bounds__(256) k_bind_decide(BindArgs b)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= b.nh)
        return;
    if (b.adj_n[p] == 0) { // a component of one patch
        b.isbound[p] = 0;
        return;
    }
    int mem[BIND_CMAX];
    int cnt = 1, head = 0;
    mem[0] = p;
    while (head < cnt) {
        const int u = mem[head++];
        const int dg = b.adj_n[u];
        for (int k = 0; k < dg; k++) {
            const int v = b.adj_p[(size_t)u * BIND_DEG + k];
            if (v < p)
                return; // not the component's smallest patch
            bool seen = false;
            for (int q = 0; q < cnt; q++)
                if (mem[q] == v) {
                    seen = true;
                    break;
                }
            if (seen)
                continue;
            if (cnt == BIND_CMAX) {
                bind_status(b, 8u);
                return;
            }
            mem[cnt++] = v;
        }
    }
    for (int q = 1; q < cnt; q++) // ascending
        for (int t = q; t > 0 && mem[t - 1] > mem[t]; t--) {
            const int tp = mem[t];
            mem[t] = mem[t - 1];
            mem[t - 1] = tp;
        }
    // bonds whose partner is not a candidate break (LangevinRforce.cpp:3098-3125)
    for (int q = 0; q < cnt; q++) {
        const int g = mem[q];
        if (b.isbound[g]) {
            bool outside = true;
            const int dg = b.adj_n[g];
            for (int k = 0; k < dg; k++)
                if (b.adj_p[(size_t)g * BIND_DEG + k] == b.boundto[g]) {
                    outside = false;
                    break;
                }
            if (outside)
                b.isbound[g] = 0;
        }
    }
    if (cnt == 2) {
        bind_decide2(b, mem[0], mem[1], bind_uniform(b, p, 0));
    } else if (cnt == 3) {
        bool c12 = false, c23 = false, c13 = false;
        const int i1 = mem[0], i2 = mem[1], i3 = mem[2];
        if (!bind_conn3(b.adj_n[i1], b.adj_p[(size_t)i1 * BIND_DEG], b.adj_n[i2], i2, i3, c12, c23,
                        c13)) {
            bind_status(b, 8u);
            return;
        }
        bind_decide3(b, i1, i2, i3, c12, c23, c13, bind_uniform(b, p, 0));
    } else {
        bind_large(b, mem, cnt, p);
    }
}

// forces and torques of the bound pairs (LangevinRforce.cpp:4030-4095,
// KernFrenkelOnePatch2::force_and_torque, potentialtheta.h:517-750), each
// particle its own share, written to Fb, Tb (3 per slot, or per original
// index with by_id; lab frame)
// ncand (per cell-list position, nullable): a particle without candidate
// neighbours has no bound patch after the decisions (an edge is seen from
// both of its particles): by slot (the step's form) nothing is written for
// it -- the integrator reads the terms only where ncand > 0 -- and by
// original index zeros
template <class R>
__device__ __forceinline__ void bind_forces_one(const KArgs<R> &a, const BindArgs &b, double *Fb,
                                                double *Tb, int by_id, int s)
{
    constexpr bool FP64 = sizeof(R) == 8;
    const int i = a.lmode ? s : a.list[s];
    const PosRec xi = a.X[i];
    const int ti = xi.meta & 255;
    const int npi = a.np[ti];
    double F[3] = {0, 0, 0}, T[3] = {0, 0, 0};
    if (npi > 0) {
        const DevRep rp = rep_rec(a, xi.meta);
        const int base = b.poff[xi.id];
        R qi[9];
        load_orient(a.Q, (size_t)i, qi);
        for (int k = 0; k < npi; k++) {
            if (!b.isbound[base + k])
                continue;
            const int wq = b.boundto[base + k];
            const int idj = b.owner[wq];
            const int j = a.slot[idj];
            const PosRec xj = a.X[j];
            const int tj = xj.meta & 255;
            double d[3];
            min_image(a, rp, xi.x, xi.y, xi.z, xj, d);
            const double rij = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            int ta, tc;
            const bool i_is_a = patch_roles(xi.id, ti, xj.id, tj, ta, tc);
            const double sg = i_is_a ? 1.0 : -1.0;
            const double u[3] = {sg * d[0] / rij, sg * d[1] / rij, sg * d[2] / rij};
            R qj[9];
            load_orient(a.Q, (size_t)j, qj);
            const R *qa = i_is_a ? qi : qj;
            const R *qc = i_is_a ? qj : qi;
            const int ia = i_is_a ? k : wq - b.poff[idj];
            const int ic = i_is_a ? wq - b.poff[idj] : k;
            const int npc = a.np[tc];
            const PatchRow pr = a.prows[rp.prow + a.rstart[ta * MAXT + tc] + ia * npc + ic];
            if (rij > pr.idist)
                continue;
            double n1[3], n2[3];
            lab_normal_exact(qa, a.pdirs + 3 * (a.pstart[ta] + ia), n1);
            lab_normal_exact(qc, a.pdirs + 3 * (a.pstart[tc] + ic), n2);
            const double a1 = -(n1[0] * u[0] + n1[1] * u[1] + n1[2] * u[2]);
            const double a2 = (n2[0] * u[0] + n2[1] * u[1] + n2[2] * u[2]);
            if (!(a1 > pr.costm && a2 > pr.costm))
                continue;
            R f[3] = {0, 0, 0}, t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
            if constexpr (FP64) {
                kf2_pair_fp64(u, rij, n1, n2, a1, a2, pr, f, t1, t2);
            } else {
                const float uf[3] = {(float)u[0], (float)u[1], (float)u[2]};
                const float n1f[3] = {(float)n1[0], (float)n1[1], (float)n1[2]};
                const float n2f[3] = {(float)n2[0], (float)n2[1], (float)n2[2]};
                kf2_pair_fp32(uf, (float)rij, n1f, n2f, (float)a1, (float)a2,
                              fabs(a1 * a1 - 1) < 1E-5, fabs(a2 * a2 - 1) < 1E-5, (float)pr.eps,
                              (float)pr.d, (float)pr.kk, f, t1, t2);
            }
            for (int q = 0; q < 3; q++) {
                F[q] += i_is_a ? (double)f[q] : -(double)f[q];
                T[q] += i_is_a ? (double)t1[q] : (double)t2[q];
            }
        }
    }
    const size_t o = 3 * (size_t)(by_id ? (int)xi.id : i);
    for (int q = 0; q < 3; q++) {
        Fb[o + q] = F[q];
        Tb[o + q] = T[q];
    }
}

template <class R>
__global__ void __launch_bounds__(256) k_bind_forces(KArgs<R> a, BindArgs b, double *Fb, double *Tb,
                                                     int by_id, const int *ncand)
{
    const int tid = threadIdx.x;
    const int s = blockIdx.x * 256 + tid;
    if (by_id || !ncand) {
        if (s < a.N)
            bind_forces_one(a, b, Fb, Tb, by_id, s);
        return;
    }
    // the positions with candidates compacted to the block's leading lanes
    __shared__ int off[256];
    __shared__ int wsum[4];
    const bool here = s < a.N && ncand[s] > 0;
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
        off[wb + rank] = s;
    __syncthreads();
    const int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    for (int t = tid; t < total; t += 256)
        bind_forces_one(a, b, Fb, Tb, 0, off[t]);
}

// LangevinNVTR::calculate_patch_list (Rotational/LangevinR.cpp:1702-1853) on
// a caller's pair list (original indices), one thread per pair: a pair
// closer than maxdis (max_check + 1) sets bit ia * np(t2) + ib of mask[k] for
// each patch ia of p1 and ib of p2 whose cone arguments satisfy
// (a_i - maxang) + (a_j - maxang) > -0.4, with p1 < p2 (particles numbered
// by type, so t1 <= t2, checked by the host).  FP64 in the reference's
// expression order; orientations as the engine holds them.
// bound: LangevinNVTR::calculate_bound_pairs (LangevinR.cpp:1854-1998)
// instead: closer than maxdis = max_check, and both cone arguments above the
// row's cos(theta_m) (computed on the host, as the reference's cos).
template <class R>
__global__ void __launch_bounds__(256) k_patch_list(KArgs<R> a, const int *pairs, long long npairs,
                                                    double maxdis, double maxang, int bound,
                                                    unsigned long long *mask)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs)
        return;
    int p1 = pairs[2 * k], p2 = pairs[2 * k + 1];
    if (p2 < p1) {
        const int t = p1;
        p1 = p2;
        p2 = t;
    }
    const int s1 = a.slot[p1], s2 = a.slot[p2];
    const PosRec x1 = a.X[s1], x2 = a.X[s2];
    const DevRep rp = rep_rec(a, x1.meta);
    double un[3];
    min_image(a, rp, x1.x, x1.y, x1.z, x2, un);
    double dis = un[0] * un[0] + un[1] * un[1] + un[2] * un[2];
    dis = sqrt(dis);
    unsigned long long m = 0ull;
    if (dis < maxdis) {
        for (int c = 0; c < 3; c++)
            un[c] /= dis;
        R q1[9], q2[9];
        load_orient(a.Q, (size_t)s1, q1);
        load_orient(a.Q, (size_t)s2, q2);
        const int t1 = x1.meta & 255, t2 = x2.meta & 255;
        const int n1 = a.np[t1], n2 = a.np[t2];
        const int rowbase = rp.prow + a.rstart[t1 * MAXT + t2];
        for (int ia = 0; ia < n1; ia++) {
            const double *b1 = a.pdirs + 3 * (a.pstart[t1] + ia);
            const double nx1 = b1[0] * (double)q1[0] + b1[1] * (double)q1[3] + b1[2] * (double)q1[6];
            const double ny1 = b1[0] * (double)q1[1] + b1[1] * (double)q1[4] + b1[2] * (double)q1[7];
            const double nz1 = b1[0] * (double)q1[2] + b1[1] * (double)q1[5] + b1[2] * (double)q1[8];
            const double ai = -(nx1 * un[0] + ny1 * un[1] + nz1 * un[2]);
            for (int ib = 0; ib < n2; ib++) {
                const double *b2 = a.pdirs + 3 * (a.pstart[t2] + ib);
                const double nx2 =
                    b2[0] * (double)q2[0] + b2[1] * (double)q2[3] + b2[2] * (double)q2[6];
                const double ny2 =
                    b2[0] * (double)q2[1] + b2[1] * (double)q2[4] + b2[2] * (double)q2[7];
                const double nz2 =
                    b2[0] * (double)q2[2] + b2[1] * (double)q2[5] + b2[2] * (double)q2[8];
                const double aj = nx2 * un[0] + ny2 * un[1] + nz2 * un[2];
                bool pass;
                if (bound) {
                    const double ctm = a.prows[rowbase + ia * n2 + ib].costm;
                    pass = ai > ctm && aj > ctm;
                } else {
                    pass = (ai - maxang) + (aj - maxang) > -0.4;
                }
                if (pass)
                    m |= 1ull << (ia * n2 + ib);
            }
        }
    }
    mask[k] = m;
}

} // namespace mdx
