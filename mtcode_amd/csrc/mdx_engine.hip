// mdx_engine.hip -- MI355X (gfx950) engine behind include/mdx.h (host side).
//
// Per step (rotor integrator, Condensate.cpp:352-377 of the reference):
//   exclusive scan of cell counts -> fill cell lists -> order each cell's list
//   -> gather forces + sphere + half kick (+ the next step's half kick, drift,
//      rotation and binning, fused)
// The gather kernel (mdx_kernels.h) computes every particle's force and torque
// by looping over its 27 neighbour cells (full neighbour list, no Newton-3
// scatter, no atomics), so results are deterministic, and re-sorts the state
// into cell order as it writes the next step.
//
// Reference paths in comments are relative to the reference's Code/ folder.
#include "mdx_kernels.h"
#include "mdx_bind.h"
#include "../../include/mdx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
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

// ================================================================ host side
namespace {

struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    std::vector<char> sig; // launch arguments it was captured with
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
    std::vector<uint32_t> meta; // by original index
    hipStream_t stream = nullptr;
    // interactions
    int wca = 0;
    double wca_eps = 1, wca_sigma = 1, wca_pl = 1.12246;
    int ntypes = 0;
    std::vector<int> np, pstart, rstart;
    std::vector<double> pdirs;
    std::vector<double> prow_raw; // per replica rows x 3
    int nrows = 0;
    double maxd2 = 2.4;
    double costm_min = 1.0;
    unsigned long long active_mask = 0;
    int bond_kind = MDX_BOND_HARMONIC, ang_form = MDX_ANGLE_HARMONIC;
    long long nbonds = 0, nangles = 0, nexcl = 0;
    double fa = 0;
    int has_active = 0;
    double fguard = 1e4;
    int use_tiles = 0; // LDS-staged neighbour tiles (MDX_TILES=1 enables)
    // positions per block of the compacted patch kernel (256 .. PBLK, a
    // multiple of 256; 0: chosen from N) (MDX_PBLK)
    int pblk = 0;
    // LDS staging of the list build: -1 chosen from the particles per x-row,
    // 0 the smaller, 1 the larger (MDX_NBDENSE)
    int nb_dense = -1;
    // walk / patch / integrator kernels (1) or one kernel per step (0); by
    // default the one-kernel form up to SPLIT_MIN_N particles, where launches
    // dominate the step (MDX_SPLIT=0/1 forces either)
    int use_split = 1;
    // neighbour list: skin > 0 enables it for mdx_step (MDX_SKIN), kmax slots per
    // particle (MDX_KMAX), state re-sorted into cell order every resort_every
    // steps (MDX_RESORT)
    double skin = 0.4;
    int kmax = 64;
    int resort_every = 200;
    bool list_valid = false;
    bool counted = false; // cell counts of the current positions are held (not yet scanned)
    int since_resort = 0;
    // grid
    int ncells = 0;
    std::vector<DevRep> reps;
    bool grid_dirty = true;
    // state
    int cur = 0;
    bool binned = false;
    bool primed = false;
    unsigned long long steps_taken = 0;
    size_t bytes_held = 0;
    // device buffers (double-buffered state in slot order)
    Buf X[2], Q[2], P[2], Lm[2], Fc[2], Tc[2], slot[2], cell[2];
    Buf Fout, Tout, obs;
    Buf dreps, prows, pdirs_d, pdirs_f;
    Buf rank, count, start, list, bsum;
    Buf bond_off, bond_j, bond_k, bond_r0;
    Buf ang_off, ang_ref, ang_ijk, ang_k, ang_t0;
    Buf exc_off, exc_j, act_prev, act_next;
    Buf noise, ctl, partial, tmpd, tmpi, tmpi2, tmpi3, tmpll, tmpx;
    Buf Fw, ncand, cand, Tp; // split walk -> (patch ->) finish
    int use_compact = 1;     // compacted patch kernel + integrate kernel (MDX_COMPACT=0: one finish kernel)
    int use_walkc = 1;       // compacted list walk (MDX_WALKC=0: one lane per position)
    int use_graph = 1;       // list steps replayed from captured graphs (MDX_GRAPH=0: eager)
    // list steps integrate in the walk / patch kernels (MDX_FUSE=1), in a
    // kernel of their own (0), or the former when at most a quarter of the
    // particles carry patches (2, the default): integrating in the walk costs
    // the walk occupancy, which pays only where few particles go on to the
    // patch kernel
    int use_fuse = 2;
    int n_patchy = 0; // particles whose type carries patches
    bool list_eager = false; // an eager list step ran since the last list build
    GraphSlot graphs[2];     // per current state buffer
    Buf nbr, nbrc, nn, bcell; // neighbour list (column-major, per-block compact copy)
    Buf dslot;           // drift maxima slots (NDSLOT), cleared by k_decide
    hipEvent_t prof_mid = nullptr;  // recorded between walk and finish while profiling
    hipEvent_t prof_mid2 = nullptr; // recorded between patch and integrate while profiling
    bool last_split = false;       // the last gather launch used the split form
    // one-bond-per-patch binding (mdx_set_binding): patch numbering, bound
    // state, candidate rows, the bound-pair forces by original index
    int bind_on = 0;
    double bind_on_rate = 0, bind_off_rate = 0, bind_mk = 0, bind_cmin = 1;
    int nh = 0;
    Buf bpoff, bowner, bisbound, bboundto, badj_n, badj_p, badj_s, bF, bT, bscr, bnscr;
    Buf plpairs, plmask; // mdx_patch_list: the caller's pairs, a bit mask per pair
    int bind_table = 0; // 1: the species-table model (mdx_set_binding_tables)
    Buf bspecies, bdtab, bttab, btr4;

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

template <class T> static int download(mdx_engine *e, const Buf &b, T *h, size_t n)
{
    if (n)
        HIPCHK(hipMemcpyAsync(h, b.p, n * sizeof(T), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MDX_OK;
}

static inline unsigned nblk(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

static int scratch_setup(int dev); // device address of the no-allocation scratch

// interaction reach: WCA cut-off and the patch reach min(sqrt(maxd2), 1.4 d)
static double reach(const mdx_engine *e)
{
    double r = 0;
    if (e->wca)
        r = std::max(r, e->wca_pl);
    if (e->ntypes > 0 && e->active_mask && e->integ == MDX_INTEG_ROTOR) {
        double dmax = 0;
        for (size_t k = 0; k < e->prow_raw.size() / 3; k++)
            if (e->prow_raw[3 * k] != 0.0)
                dmax = std::max(dmax, 1.4 * e->prow_raw[3 * k + 1]);
        r = std::max(r, std::min(std::sqrt(e->maxd2), dmax));
    }
    // binding candidates lie within the largest interaction distance of any
    // row (max_check), whatever max_distance_check
    if (e->bind_on)
        r = std::max(r, e->bind_mk);
    return r;
}

static int build_grid(mdx_engine *e)
{
    if (!e->grid_dirty)
        return MDX_OK;
    // with the neighbour list the cells hold the list reach (interaction + skin)
    const double r0 = reach(e);
    const double rc = r0 > 0 && e->skin > 0 ? r0 + e->skin : r0;
    e->reps.resize(e->nrep);
    long long coff = 0;
    for (int r = 0; r < e->nrep; r++) {
        DevRep &d = e->reps[r];
        for (int k = 0; k < 3; k++) {
            d.l[k] = e->box[3 * r + k];
            d.l2s[k] = (d.l[k] / 2.) * (d.l[k] / 2.);
            // cells no smaller than the interaction reach (margin for rounding)
            int n = rc > 0 ? (int)std::floor(d.l[k] / (rc * (1 + 1e-9))) : 1;
            d.nc[k] = std::max(1, std::min(n, 1024));
        }
        d.coff = (int)coff;
        coff += (long long)d.nc[0] * d.nc[1] * d.nc[2];
        d.first = e->first[r];
        d.count = e->nper[r];
        const double kT = e->kT[r];
        d.Rt = std::sqrt(2 * e->gamma * kT / e->dt);
        d.Rr = std::sqrt(2 * e->gammar * kT / e->dt);
        d.bead_r = std::sqrt(0.5 * kT * (e->gamma) * (e->m) * (e->dt));
        d.prow = r * e->nrows;
    }
    if (coff > (1ll << 30))
        return set_err(MDX_ERR_ARG, "cell grid too large (%lld cells)", coff);
    e->ncells = (int)coff;
    int rc2;
    if ((rc2 = upload(e, e->dreps, e->reps.data(), e->reps.size())))
        return rc2;
    if ((rc2 = dalloc(e, e->count, sizeof(int) * (size_t)(e->ncells + 1))))
        return rc2;
    HIPCHK(hipMemsetAsync(e->count.p, 0, sizeof(int) * (size_t)(e->ncells + 1), e->stream));
    if ((rc2 = dalloc(e, e->start, sizeof(int) * (size_t)(e->ncells + 1))))
        return rc2;
    const int nb = (e->ncells + 1023) / 1024;
    if (nb > 1024 * 1024)
        return set_err(MDX_ERR_ARG, "cell grid too large for the scan");
    if ((rc2 = dalloc(e, e->bsum, sizeof(int) * (size_t)std::max(nb, 1))))
        return rc2;
    e->grid_dirty = false;
    e->binned = false;
    e->primed = false;
    e->counted = false;
    e->list_valid = false;
    return MDX_OK;
}

template <class R> static KArgs<R> make_args(mdx_engine *e)
{
    KArgs<R> a;
    memset(&a, 0, sizeof a);
    const int c = e->cur, n = e->cur ^ 1;
    a.N = e->N;
    a.rotor = e->integ == MDX_INTEG_ROTOR;
    a.periodic = e->periodic;
    a.reps = (const DevRep *)e->dreps.p;
    a.X = (const PosRec *)e->X[c].p;
    a.Xn = (PosRec *)e->X[n].p;
    a.Q = (const R *)e->Q[c].p;
    a.Qn = (R *)e->Q[n].p;
    a.p = (const R *)e->P[c].p;
    a.pn = (R *)e->P[n].p;
    a.L = (const R *)e->Lm[c].p;
    a.Ln = (R *)e->Lm[n].p;
    a.F = (const R *)e->Fc[c].p;
    a.Fn = (R *)e->Fc[n].p;
    a.T = (const R *)e->Tc[c].p;
    a.Tn = (R *)e->Tc[n].p;
    a.slot = (const int *)e->slot[c].p;
    a.slotn = (int *)e->slot[n].p;
    a.cell = (const int *)e->cell[c].p;
    a.celln = (int *)e->cell[n].p;
    a.rank = (int *)e->rank.p;
    a.count = (int *)e->count.p;
    a.start = (const int *)e->start.p;
    a.list = (const int *)e->list.p;
    a.pcap = PMAXG;
    a.Fout = (double *)e->Fout.p;
    a.Tout = (double *)e->Tout.p;
    a.obs = (double *)e->obs.p;
    a.Fw = (R *)e->Fw.p;
    a.ncand = (int *)e->ncand.p;
    a.cand = (int *)e->cand.p;
    a.wca = e->wca;
    a.wca_eps = e->wca_eps;
    const double s3 = e->wca_sigma * e->wca_sigma * e->wca_sigma;
    a.wca_s6 = s3 * s3;
    a.wca_s12 = a.wca_s6 * a.wca_s6;
    a.wca_pl = e->wca_pl;
    a.wca_pl2 = e->wca_pl * e->wca_pl * (1 + 1e-12);
    a.nrep = e->nrep;
    // with binding the pair walk carries no Kern-Frenkel terms: the bound
    // pairs' forces come from the binding kernels (Condensate.cpp:468-473)
    a.has_patches =
        e->ntypes > 0 && e->integ == MDX_INTEG_ROTOR && e->active_mask != 0 && !e->bind_on;
    for (int t = 0; t < MAXT; t++) {
        a.np[t] = t < e->ntypes ? e->np[t] : 0;
        a.pstart[t] = t < e->ntypes ? e->pstart[t] : 0;
    }
    for (int k = 0; k < MAXT * MAXT; k++)
        a.rstart[k] = k < (int)e->rstart.size() ? e->rstart[k] : 0;
    a.active_mask = e->active_mask;
    a.pdirs = (const double *)e->pdirs_d.p;
    // the own-patch screen reads the directions in FP32 (both precisions)
    a.pdirs_f = (const float *)e->pdirs_f.p;
    a.wca_pl2lo = e->wca_pl * e->wca_pl * (1 - 1e-12);
    a.prows = (const PatchRow *)e->prows.p;
    a.maxd2 = e->maxd2;
    a.rmax2 = std::max(e->wca ? e->wca_pl * e->wca_pl * (1 + 1e-12) : 0.0,
                       a.has_patches ? e->maxd2 : 0.0);
    // the screen of the cone test runs in R precision: FP32 needs a margin
    a.costm_screen = e->costm_min - (e->prec == MDX_PREC_FP64 ? 1e-12 : 1e-5);
    a.costm_screen_f = e->costm_min - 1e-5;
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
    // the friction factors of the sphere step, divided here once (the same
    // IEEE quotients the device would form)
    a.gm_m = -e->gamma / e->m;
    for (int k = 0; k < 3; k++)
        a.gr_I[k] = -e->gammar / e->I[k];
    // LangevinNVT coefficients (Langevin.cpp:44-52); r depends on kT (per replica)
    const double d = (e->gamma * e->dt / 2.), q = (e->dt) / 2.;
    const double c2 = (1.0 / (1.0 + (d))), c3 = (1.0 / (1.0 + (d))) * q, c5 = (1 - (d));
    a.b_c2 = c2;
    a.b_c5c2 = c5 * c2;
    a.b_c5c3q = (c5 * (c3) + q);
    a.key0 = (uint32_t)e->seed;
    a.key1 = (uint32_t)(e->seed >> 32);
    a.noise = nullptr;
    a.ctl = (StepCtl *)e->ctl.p;
    a.fguard = e->fguard;
    a.lmode = 0;
    a.nbr = (const int *)e->nbr.p;
    a.nbrw = (int *)e->nbr.p;
    a.nbrc = (const int *)e->nbrc.p;
    a.nbrcw = (int *)e->nbrc.p;
    a.nn = (int *)e->nn.p;
    a.bcell = (int *)e->bcell.p;
    a.kmax = e->kmax;
    const double rl = reach(e) + e->skin;
    a.rl2 = rl * rl;
    // (the decisions on the WCA cut-off carry a 1e-12 relative margin); one
    // radius when every type pair is patch-active (no per-pair test needed)
    bool all_active = a.has_patches;
    for (int ta = 0; ta < e->ntypes && all_active; ta++)
        for (int tc = ta; tc < e->ntypes; tc++)
            all_active &= (e->active_mask >> (ta * MAXT + tc)) & 1ull;
    const double rw = e->wca ? e->wca_pl * (1 + 1e-12) + e->skin : 0.0;
    a.rl2w = all_active ? a.rl2 : (e->wca ? std::min(rw * rw, a.rl2) : 0.0);
    a.bF = a.bT = nullptr; // (set by the binding step's launch)
    a.bncand = nullptr;
    if (e->bind_on)
        a.rl2w = a.rl2; // every type pair can bind: one list radius (max_check + skin)
    a.skin = e->skin;
    a.dslot = (unsigned int *)e->dslot.p;
    return a;
}

// scan + fill + per-cell ordering of the cell list built by the binning of
// the current slots (cells: the binning's cell of each slot; gate: run only
// when a list rebuild was requested)
static int finish_bins(mdx_engine *e, const StepCtl *gate = nullptr, const int *cells = nullptr)
{
    const int n = e->ncells;
    const int nb = (n + 1023) / 1024;
    if (!cells)
        cells = (const int *)e->cell[e->cur].p;
    hipLaunchKernelGGL(k_scan1, dim3(nb), dim3(256), 0, e->stream, (const int *)e->count.p,
                       (int *)e->start.p, (int *)e->bsum.p, n, gate);
    hipLaunchKernelGGL(k_scan3, dim3(nb), dim3(256), 0, e->stream, (int *)e->count.p,
                       (int *)e->start.p, (const int *)e->bsum.p, n, gate);
    hipLaunchKernelGGL(k_fill, dim3(nblk(e->N)), dim3(256), 0, e->stream, e->N, cells,
                       (const int *)e->rank.p, (const int *)e->start.p, (int *)e->list.p, gate);
    hipLaunchKernelGGL(k_cellsort, dim3(nblk(e->N)), dim3(256), 0, e->stream, e->N, cells,
                       (const int *)e->rank.p, (const int *)e->start.p, (int *)e->list.p,
                       (const PosRec *)e->X[e->cur].p, gate);
    HIPCHK(hipGetLastError());
    return MDX_OK;
}

// cell counts of the current slots into cells (gate as above)
template <class R> static int count_current(mdx_engine *e, int *cells, const StepCtl *gate)
{
    KArgs<R> a = make_args<R>(e);
    a.celln = cells;
    hipLaunchKernelGGL(k_bin<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, a, gate);
    HIPCHK(hipGetLastError());
    return MDX_OK;
}

// cell list of the current slots (no reordering)
template <class R> static int bin_current(mdx_engine *e)
{
    int rc = build_grid(e);
    if (rc)
        return rc;
    if (!e->counted && (rc = count_current<R>(e, (int *)e->cell[e->cur].p, nullptr)))
        return rc;
    if ((rc = finish_bins(e)))
        return rc;
    e->counted = false;
    e->binned = true;
    // the cell list no longer describes the positions of the neighbour-list build
    e->list_valid = false;
    return MDX_OK;
}

template <class R, int MODE>
static int launch_bind_step(mdx_engine *e, const double *noise_dev, int lmode);

template <class R, int MODE>
static int launch_gather(mdx_engine *e, const double *noise_dev, int lmode = 0)
{
    if (e->bind_on && (MODE == M_PRIME || MODE == M_STEP || MODE == M_FUSED))
        return launch_bind_step<R, MODE>(e, noise_dev, lmode);
    KArgs<R> a = make_args<R>(e);
    a.noise = noise_dev;
    a.lmode = lmode;
    // split form: patch systems, and list steps of the bead integrator (walk,
    // then the integrator kernel; no patch kernel)
    const bool bead_list = MODE == M_BEAD && lmode && e->use_walkc;
    const bool split = e->use_split && ((a.has_patches && (MODE == M_PRIME || MODE == M_STEP ||
                                                           MODE == M_FUSED || MODE == M_CONS)) ||
                                        bead_list);
    if (split) {
        int rc;
        // list steps of the compacted walk hold every candidate of a position
        // with at most kmax list entries (its candidates are among them)
        const bool list_walk = lmode && e->use_walkc;
        a.pcap = list_walk ? std::max(PMAXG, e->kmax) : PMAXG;
        if ((rc = dalloc(e, e->Fw, 3 * e->rsz() * (size_t)e->N)) ||
            (rc = dalloc(e, e->ncand, 4 * (size_t)e->N)) ||
            (rc = dalloc(e, e->cand, 4 * (size_t)a.pcap * e->N)))
            return rc;
        if (e->use_compact && (rc = dalloc(e, e->Tp, 3 * e->rsz() * (size_t)e->N)))
            return rc;
        a.Fw = (R *)e->Fw.p;
        a.ncand = (int *)e->ncand.p;
        a.cand = (int *)e->cand.p;
        a.Tp = (R *)e->Tp.p;
        // positions per patch block: as many as keep about 512 blocks
        int pblk = 256;
        while (pblk < PBLK && (long long)e->N / (2 * pblk) >= 512)
            pblk *= 2;
        if (e->pblk > 0)
            pblk = e->pblk;
        // fused form (list steps): the walk integrates the particles without
        // patch candidates, the patch kernel the others
        const bool fuse_ok = e->use_fuse == 1 ||
                             (e->use_fuse == 2 && (!a.has_patches || 4LL * e->n_patchy <= e->N));
        const bool fuse = lmode && e->use_walkc && fuse_ok && e->use_compact &&
                          (MODE == M_FUSED || MODE == M_STEP || MODE == M_BEAD);
        if (fuse) {
            hipLaunchKernelGGL((k_walkc<R, MODE, true>), dim3(nblk(e->N)), dim3(256), 0, e->stream,
                               a);
            if (e->prof_mid)
                HIPCHK(hipEventRecord(e->prof_mid, e->stream));
            e->last_split = true;
            if (!bead_list) {
                hipLaunchKernelGGL((k_patch<R, MODE, true>), dim3(nblk(e->N, pblk)), dim3(256), 0,
                                   e->stream, a, pblk);
            }
            if (e->prof_mid2)
                HIPCHK(hipEventRecord(e->prof_mid2, e->stream));
            HIPCHK(hipGetLastError());
            return MDX_OK;
        }
        if (lmode && e->use_walkc) {
            // no overflow kernel after the compacted form's patch kernel: a
            // position with more candidates than kmax stops the run
            KArgs<R> w = a;
            w.ovf_fatal = e->use_compact ? 1 : 0;
            hipLaunchKernelGGL((k_walkc<R, MODE, false>), dim3(nblk(e->N)), dim3(256), 0, e->stream,
                               w);
        } else
            hipLaunchKernelGGL((k_walk<R, MODE>), dim3(nblk(e->N)), dim3(256), 0, e->stream, a);
        if (e->prof_mid)
            HIPCHK(hipEventRecord(e->prof_mid, e->stream));
        e->last_split = true;
        if (bead_list) {
            if (e->prof_mid2)
                HIPCHK(hipEventRecord(e->prof_mid2, e->stream));
            hipLaunchKernelGGL((k_integrate<R, MODE>), dim3(nblk(e->N)), dim3(256), 0, e->stream,
                               a);
        } else if (e->use_compact) {
            hipLaunchKernelGGL((k_patch<R, MODE, false>), dim3(nblk(e->N, pblk)), dim3(256), 0,
                               e->stream, a, pblk);
            // (list steps hold every candidate: no overflow kernel)
            if (!list_walk)
                hipLaunchKernelGGL((k_patch_ovf<R, MODE>), dim3(std::min(nblk(e->N), 512u)),
                                   dim3(256), 0, e->stream, a);
            if (e->prof_mid2)
                HIPCHK(hipEventRecord(e->prof_mid2, e->stream));
            hipLaunchKernelGGL((k_integrate<R, MODE>), dim3(nblk(e->N)), dim3(256), 0, e->stream,
                               a);
        } else {
            if (e->prof_mid2)
                HIPCHK(hipEventRecord(e->prof_mid2, e->stream));
            hipLaunchKernelGGL((k_finish<R, MODE>), dim3(nblk(e->N)), dim3(256), 0, e->stream,
                               a);
        }
        HIPCHK(hipGetLastError());
        return MDX_OK;
    }
    e->last_split = false;
    // (k_tile: no list mode, nothing to close)
    if (e->use_tiles && !lmode)
        hipLaunchKernelGGL((k_tile<R, MODE>), dim3(nblk(e->N)), dim3(256), 0, e->stream, a);
    else
        hipLaunchKernelGGL((k_gather<R, MODE>), dim3(nblk(e->N)), dim3(256), 0, e->stream,
                           a);
    HIPCHK(hipGetLastError());
    return MDX_OK;
}

// status of the calls since the last read (and clear it): st = status bits,
// *step = the force evaluation that first raised one, counted from the start
// of the current call (0: the carried-force evaluation of mdx_prime, or the
// first step's kernels; k: step k)
static int read_status(mdx_engine *e, unsigned int *st, long long *step = nullptr)
{
    StepCtl c;
    HIPCHK(hipMemcpyAsync(&c, e->ctl.p, sizeof c, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    *st = c.status;
    if (step)
        *step = c.status ? (long long)(c.fail_epoch - c.call_base) : -1;
    if (c.status) {
        unsigned int z = 0;
        unsigned long long none = ~0ull;
        HIPCHK(hipMemcpyAsync((char *)e->ctl.p + offsetof(StepCtl, status), &z, sizeof z,
                              hipMemcpyHostToDevice, e->stream));
        HIPCHK(hipMemcpyAsync((char *)e->ctl.p + offsetof(StepCtl, fail_epoch), &none,
                              sizeof none, hipMemcpyHostToDevice, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
    }
    return MDX_OK;
}

// step >= 1: the step of the call whose force evaluation raised the status
// (the steps after it moved no particle); 0: mdx_prime's evaluation
static int status_to_rc(unsigned int st, long long step = -1)
{
    char where[96] = "";
    if (step == 0)
        snprintf(where, sizeof where, " in the carried-force evaluation");
    else if (step > 0)
        snprintf(where, sizeof where, " at step %lld of the call (no particle moved after it)", step);
    if (st & 8u)
        return set_err(MDX_ERR_CAPACITY,
                       "a patch has more than %d binding candidates or a candidate component more "
                       "than %d patches%s", BIND_DEG, BIND_CMAX, where);
    if (st & 4u)
        return set_err(MDX_ERR_CAPACITY,
                       "a particle has more patch candidates than the neighbour list holds (kmax): "
                       "raise kmax with mdx_set_neighbor_list%s", where);
    if (st & 2u)
        return set_err(MDX_ERR_NAN, "non-finite force, torque or updated state%s", where);
    if (st & 1u)
        return set_err(MDX_ERR_FORCE, "pair force above the guard (force too big)%s", where);
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

static void launch_small(mdx_engine *e, void (*k)(StepCtl *))
{
    hipLaunchKernelGGL(k, dim3(1), dim3(1), 0, e->stream, (StepCtl *)e->ctl.p);
}

template <class R> static int do_prime(mdx_engine *e, const double *noise0)
{
    int rc = bin_current<R>(e);
    if (rc)
        return rc;
    const double *nd;
    if ((rc = noise_upload(e, noise0, (size_t)e->N * 6, &nd)))
        return rc;
    launch_small(e, k_set_call_base);
    if ((rc = launch_gather<R, M_PRIME>(e, nd)))
        return rc;
    e->cur ^= 1;
    launch_small(e, k_epoch_inc);
    e->binned = false; // the state moved to cell order; its list is stale
    e->counted = false;
    e->list_valid = false;
    unsigned int st;
    long long at;
    if ((rc = read_status(e, &st, &at)))
        return rc;
    e->primed = true;
    return status_to_rc(st, at);
}

static int nbr_alloc(mdx_engine *e)
{
    int rc;
    if ((rc = dalloc(e, e->nbr, 4 * (size_t)e->kmax * e->N)) ||
        (rc = dalloc(e, e->nbrc, 4 * (size_t)e->kmax * nblk(e->N) * 256)) ||
        (rc = dalloc(e, e->nn, 4 * (size_t)e->N)) || (rc = dalloc(e, e->bcell, 4 * (size_t)e->N)) ||
        (rc = dalloc(e, e->dslot, 4 * (size_t)NDSLOT)))
        return rc;
    return MDX_OK;
}

// after a kernel that moved the particles in list mode (its blocks wrote
// drift slots); epoch_inc: also count the step's force evaluation;
// step_end: the list step's kernels are done (clears fresh)
static int decide(mdx_engine *e, int epoch_inc, int step_end)
{
    hipLaunchKernelGGL(k_decide, dim3(1), dim3(NDSLOT), 0, e->stream, (StepCtl *)e->ctl.p,
                       (unsigned int *)e->dslot.p, e->skin, epoch_inc, step_end);
    HIPCHK(hipGetLastError());
    return MDX_OK;
}

// neighbour list of the current slots from the cell list just built
template <class R> static int build_list(mdx_engine *e, const int *cells, int gated)
{
    KArgs<R> a = make_args<R>(e);
    a.lmode = 1;
    // the larger LDS staging where an x-row of cells holds many particles
    // (about 3 + 256 / row rows of particles staged per z offset of a block)
    long long rows = 0;
    for (const DevRep &r : e->reps)
        rows += (long long)r.nc[1] * r.nc[2];
    const double per_row = rows > 0 ? (double)e->N / (double)rows : 0.0;
    const bool dense =
        e->nb_dense >= 0
            ? e->nb_dense != 0
            : per_row > 0 && (std::ceil(256.0 / per_row) + 3.0) * per_row > 0.9 * NB_CAP_S;
    if (dense)
        hipLaunchKernelGGL((k_nbuild<R, NB_CAP_D>), dim3(nblk(e->N)), dim3(256), 0, e->stream, a,
                           cells, gated);
    else
        hipLaunchKernelGGL((k_nbuild<R, NB_CAP_S>), dim3(nblk(e->N)), dim3(256), 0, e->stream, a,
                           cells, gated);
    HIPCHK(hipGetLastError());
    return MDX_OK;
}

// optional per-step events: ev[5 s + 0] step start, + 1 after the cell / list
// work, + 2 after the walk, + 3 after the patch kernel (split form), + 4 step end
constexpr int NEV = 5;
static inline void mark(mdx_engine *e, hipEvent_t *ev, int s, int k)
{
    if (ev)
        (void)hipEventRecord(ev[NEV * s + k], e->stream);
}

// the kernels of one list step (gated list upkeep, walk / patch / integrate,
// displacement decision); no host state changes, so it can be captured
template <class R, int MODE>
static int launch_list_step(mdx_engine *e, const double *nd, hipEvent_t *ev, int s)
{
    int rc;
    const StepCtl *gate = (const StepCtl *)e->ctl.p;
    if ((rc = count_current<R>(e, (int *)e->bcell.p, gate)))
        return rc;
    if ((rc = finish_bins(e, gate, (const int *)e->bcell.p)))
        return rc;
    if ((rc = build_list<R>(e, (const int *)e->bcell.p, 1)))
        return rc;
    mark(e, ev, s, 1);
    e->prof_mid = ev ? ev[NEV * s + 2] : nullptr;
    e->prof_mid2 = ev ? ev[NEV * s + 3] : nullptr;
    rc = launch_gather<R, MODE>(e, nd, 1);
    e->prof_mid = e->prof_mid2 = nullptr;
    if (rc)
        return rc;
    if (MODE != M_STEP) {
        if ((rc = decide(e, 1, 1)))
            return rc;
    } else {
        launch_small(e, k_step_end);
    }
    return MDX_OK;
}

// a list step from a graph captured per state buffer (the kernels' arguments
// differ only in which buffer is current); recaptured when the arguments
// change.  The graph holds kernel nodes only (no memset or copy nodes).
template <class R, int MODE> static int list_step_graph(mdx_engine *e)
{
    KArgs<R> a = make_args<R>(e);
    a.lmode = 1;
    // the arguments, the grid and the launch-form switches (which kernels
    // the step launches)
    const int form[7] = {e->ncells,      e->N,         e->use_split, e->use_compact,
                         e->use_walkc,   e->use_fuse,  e->use_tiles};
    const double rates[3] = {e->bind_on_rate, e->bind_off_rate, (double)e->bind_table};
    std::vector<char> sig(sizeof a + sizeof form + sizeof rates);
    memcpy(sig.data(), &a, sizeof a);
    memcpy(sig.data() + sizeof a, form, sizeof form);
    memcpy(sig.data() + sizeof a + sizeof form, rates, sizeof rates);
    GraphSlot &g = e->graphs[e->cur];
    if (!g.exec || g.sig != sig) {
        if (g.exec) {
            HIPCHK(hipGraphExecDestroy(g.exec));
            g.exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
        int rc = launch_list_step<R, MODE>(e, nullptr, nullptr, 0);
        hipError_t ce = hipStreamEndCapture(e->stream, &graph);
        if (rc)
            return rc;
        if (ce != hipSuccess)
            return set_err(MDX_ERR_HIP, "stream capture failed: %s", hipGetErrorString(ce));
        hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
            g.exec = nullptr;
            return set_err(MDX_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ie));
        }
        g.sig = sig;
    }
    HIPCHK(hipGraphLaunch(g.exec, e->stream));
    return MDX_OK;
}

// one step of the loop.  MODE: M_FUSED / M_STEP (rotor), M_BEAD.
// Without the list (skin 0) every step scans the counts its predecessor took
// and walks the 27 cells; the gather re-sorts the state into cell order.
// With the list, a re-sort step does the same and then builds the list at
// the new positions; the steps in between walk the list in place, and a
// rebuild runs (gated kernels) whenever the displacement bound asks for it.
template <class R, int MODE>
static int enqueue_one(mdx_engine *e, const double *nd, hipEvent_t *ev, int s, bool last)
{
    int rc;
    const bool lm = e->skin > 0;
    mark(e, ev, s, 0);
    if (!lm) {
        if (!e->counted && (rc = count_current<R>(e, (int *)e->cell[e->cur].p, nullptr)))
            return rc;
        if ((rc = finish_bins(e)))
            return rc;
        mark(e, ev, s, 1);
        e->prof_mid = ev ? ev[NEV * s + 2] : nullptr;
        e->prof_mid2 = ev ? ev[NEV * s + 3] : nullptr;
        rc = launch_gather<R, MODE>(e, nd, 0);
        e->prof_mid = e->prof_mid2 = nullptr;
        if (rc)
            return rc;
        e->cur ^= 1;
        launch_small(e, k_epoch_inc);
        e->counted = MODE != M_STEP; // FUSED and BEAD bin the new positions
        e->binned = false;
    } else if (!e->list_valid || e->since_resort >= e->resort_every) {
        if ((rc = nbr_alloc(e)))
            return rc;
        if (!e->counted && (rc = count_current<R>(e, (int *)e->cell[e->cur].p, nullptr)))
            return rc;
        if ((rc = finish_bins(e)))
            return rc;
        mark(e, ev, s, 1);
        e->prof_mid = ev ? ev[NEV * s + 2] : nullptr;
        e->prof_mid2 = ev ? ev[NEV * s + 3] : nullptr;
        rc = launch_gather<R, MODE>(e, nd, 0);
        e->prof_mid = e->prof_mid2 = nullptr;
        if (rc)
            return rc;
        e->cur ^= 1;
        launch_small(e, k_epoch_inc);
        if (MODE != M_STEP) {
            // the gather binned the new positions: cell list, then the list
            if ((rc = finish_bins(e)))
                return rc;
            if ((rc = build_list<R>(e, (const int *)e->cell[e->cur].p, 0)))
                return rc;
            launch_small(e, k_list_reset);
            e->list_eager = false;
            // list steps keep the slots and do not rewrite the slot map
            HIPCHK(hipMemcpyAsync(e->slot[e->cur ^ 1].p, e->slot[e->cur].p, 4 * (size_t)e->N,
                                  hipMemcpyDeviceToDevice, e->stream));
            e->list_valid = true;
            e->since_resort = 0;
            e->binned = MODE == M_BEAD;
        } else {
            e->list_valid = false; // re-sorted, list not rebuilt
        }
        e->counted = false;
    } else {
        // replayed from a captured graph once an eager list step has run
        // with the same launch arguments (engine noise, not profiling)
        const bool graph = e->use_graph && !ev && !nd && MODE != M_STEP && e->list_eager;
        if (graph)
            rc = list_step_graph<R, MODE>(e);
        else
            rc = launch_list_step<R, MODE>(e, nd, ev, s);
        if (rc)
            return rc;
        e->list_eager = true;
        e->cur ^= 1;
        e->since_resort++;
        e->binned = false;
        e->counted = false;
    }
    if (ev && !e->last_split) {
        mark(e, ev, s, 2);
        mark(e, ev, s, 3);
    }
    mark(e, ev, s, 4);
    (void)last;
    return MDX_OK;
}

// enqueue nsteps rotor steps (requires primed F, T); ev: see mark()
template <class R> static int enqueue_rotor(mdx_engine *e, int nsteps, const double *nd,
                                            hipEvent_t *ev = nullptr, hipEvent_t *kev = nullptr)
{
    int rc;
    // the first kernel reads the replica table and bins into the cell grid
    if ((rc = build_grid(e)))
        return rc;
    launch_small(e, k_set_call_base);
    {
        KArgs<R> a = make_args<R>(e);
        // with a live list the drift counts toward its displacement bound;
        // otherwise the new positions are binned for the next step
        a.lmode = e->skin > 0 && e->list_valid;
        if (kev)
            HIPCHK(hipEventRecord(kev[0], e->stream));
        hipLaunchKernelGGL(k_kick_drift<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, a);
        HIPCHK(hipGetLastError());
        if (a.lmode && (rc = decide(e, 0, 0)))
            return rc;
        if (kev)
            HIPCHK(hipEventRecord(kev[1], e->stream));
        e->cur ^= 1;
        e->counted = !a.lmode;
    }
    for (int s = 0; s < nsteps; s++) {
        const bool last = s + 1 == nsteps;
        rc = last ? enqueue_one<R, M_STEP>(e, nd, ev, s, last)
                  : enqueue_one<R, M_FUSED>(e, nd, ev, s, last);
        if (rc)
            return rc;
    }
    e->binned = false;
    e->steps_taken += nsteps;
    return MDX_OK;
}

template <class R> static int do_step_rotor(mdx_engine *e, int nsteps, const double *noise)
{
    int rc;
    // a change of the interactions or of the list skin since mdx_prime
    // rebuilds the grid and drops the carried forces computed with the old ones
    if ((rc = build_grid(e)))
        return rc;
    if (!e->primed)
        return set_err(MDX_ERR_STATE, "mdx_step before mdx_prime or mdx_upload_forces (the carried "
                                      "F, T are unset, or the interactions changed since)");
    const double *nd;
    if ((rc = noise_upload(e, noise, (size_t)nsteps * e->N * 6, &nd)))
        return rc;
    if ((rc = enqueue_rotor<R>(e, nsteps, nd)))
        return rc;
    unsigned int st;
    long long at;
    if ((rc = read_status(e, &st, &at)))
        return rc;
    // step k of the call evaluates its forces at epoch call_base + k - 1
    return status_to_rc(st, at >= 0 ? at + 1 : at);
}

template <class R> static int enqueue_bead(mdx_engine *e, int nsteps, const double *nd,
                                           hipEvent_t *ev = nullptr)
{
    int rc;
    if ((rc = build_grid(e)))
        return rc;
    launch_small(e, k_set_call_base);
    for (int s = 0; s < nsteps; s++)
        if ((rc = enqueue_one<R, M_BEAD>(e, nd, ev, s, s + 1 == nsteps)))
            return rc;
    e->steps_taken += nsteps;
    return MDX_OK;
}

template <class R> static int do_step_bead(mdx_engine *e, int nsteps, const double *noise)
{
    int rc;
    const double *nd;
    if ((rc = noise_upload(e, noise, (size_t)nsteps * e->N * 3, &nd)))
        return rc;
    if ((rc = enqueue_bead<R>(e, nsteps, nd)))
        return rc;
    unsigned int st;
    long long at;
    if ((rc = read_status(e, &st, &at)))
        return rc;
    return status_to_rc(st, at >= 0 ? at + 1 : at);
}

// ids of the current slots (host copy)
static int slot_ids(mdx_engine *e, std::vector<PosRec> &rec)
{
    rec.resize(e->N);
    return download(e, e->X[e->cur], rec.data(), rec.size());
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

// slot-ordered device array (width values per particle) to a host array in
// original order
template <class R>
static int download_by_id(mdx_engine *e, const Buf &src, const std::vector<PosRec> &rec, int width,
                          double *h)
{
    const size_t n = (size_t)e->N * width;
    std::vector<double> tmp(n);
    if (sizeof(R) == 8) {
        int rc = download(e, src, tmp.data(), n);
        if (rc)
            return rc;
    } else {
        int rc = dalloc(e, e->tmpd, n * 8);
        if (rc)
            return rc;
        hipLaunchKernelGGL((k_convert<R, double>), dim3(nblk((long long)n)), dim3(256), 0,
                           e->stream, (const R *)src.p, (double *)e->tmpd.p, n);
        HIPCHK(hipGetLastError());
        if ((rc = download(e, e->tmpd, tmp.data(), n)))
            return rc;
    }
    for (int s = 0; s < e->N; s++)
        std::memcpy(h + (size_t)rec[s].id * width, tmp.data() + (size_t)s * width, 8 * width);
    return MDX_OK;
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
    if (d->nreplicas > (1 << 23))
        return set_err(MDX_ERR_ARG, "too many replicas");
    if (d->precision != MDX_PREC_FP64 && d->precision != MDX_PREC_MIXED)
        return set_err(MDX_ERR_ARG, "unknown precision %d", d->precision);
    if (d->integrator != MDX_INTEG_ROTOR && d->integrator != MDX_INTEG_BEAD)
        return set_err(MDX_ERR_ARG, "unknown integrator %d", d->integrator);
    if (!(d->dt > 0) || !(d->m > 0) || !(d->gamma >= 0) || !(d->gammar >= 0))
        return set_err(MDX_ERR_ARG, "dt, m must be > 0 and gamma, gammar >= 0");
    for (int k = 0; k < 3; k++)
        if (!(d->inertia[k] > 0))
            return set_err(MDX_ERR_ARG, "inertia must be > 0");
    long long N = 0;
    for (int r = 0; r < d->nreplicas; r++) {
        if (d->n_per_replica[r] < 1)
            return set_err(MDX_ERR_ARG, "replica %d has no particles", r);
        for (int k = 0; k < 3; k++)
            if (!(d->box[3 * r + k] > 0))
                return set_err(MDX_ERR_ARG, "replica %d box edge %d must be > 0", r, k);
        if (!(d->kT[r] >= 0))
            return set_err(MDX_ERR_ARG, "replica %d kT must be >= 0", r);
        N += d->n_per_replica[r];
    }
    if (N > (1ll << 30))
        return set_err(MDX_ERR_ARG, "too many particles");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (d->device < 0 || d->device >= ndev)
        return set_err(MDX_ERR_ARG, "device %d not present (%d devices)", d->device, ndev);
    HIPCHK(hipSetDevice(d->device));
    if (int src = scratch_setup(d->device))
        return src;
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
    e->N = (int)N;
    // measured on tetrahedral systems (FP64 / mixed): the one-kernel form is
    // 5-14 % faster at 4,096-65,536 particles and 36 % slower at 262,144
    constexpr long long SPLIT_MIN_N = 100000;
    e->use_split = N > SPLIT_MIN_N;
    if (const char *t = std::getenv("MDX_TILES"))
        e->use_tiles = std::atoi(t) != 0;
    if (const char *t = std::getenv("MDX_PBLK"))
        e->pblk = std::min(PBLK, std::max(0, std::atoi(t) / 256 * 256));
    if (const char *t = std::getenv("MDX_NBDENSE"))
        e->nb_dense = std::atoi(t) != 0;
    if (const char *t = std::getenv("MDX_SPLIT"))
        e->use_split = std::atoi(t) != 0;
    if (const char *t = std::getenv("MDX_FUSE"))
        e->use_fuse = std::atoi(t);
    if (const char *t = std::getenv("MDX_GRAPH"))
        e->use_graph = std::atoi(t) != 0;
    if (const char *t = std::getenv("MDX_WALKC"))
        e->use_walkc = std::atoi(t) != 0;
    if (const char *t = std::getenv("MDX_COMPACT"))
        e->use_compact = std::atoi(t) != 0;
    if (const char *t = std::getenv("MDX_SKIN"))
        e->skin = std::max(0.0, std::atof(t));
    if (const char *t = std::getenv("MDX_KMAX"))
        e->kmax = std::max(MDX_LBATCH, std::atoi(t));
    if (const char *t = std::getenv("MDX_RESORT"))
        e->resort_every = std::max(1, std::atoi(t));
    long long f = 0;
    for (int r = 0; r < e->nrep; r++) {
        e->first.push_back((int)f);
        e->nper.push_back(d->n_per_replica[r]);
        f += d->n_per_replica[r];
        e->kT.push_back(d->kT[r]);
        for (int k = 0; k < 3; k++)
            e->box.push_back(d->box[3 * r + k]);
        for (int k = 0; k < d->n_per_replica[r]; k++)
            e->meta.push_back((uint32_t)r << 8);
    }
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return set_err(MDX_ERR_HIP, "hipStreamCreate failed");
    }
    const size_t rs = e->rsz();
    int rc = 0;
    for (int b = 0; b < 2 && !rc; b++) {
        rc |= dalloc(e, e->X[b], sizeof(PosRec) * N);
        rc |= dalloc(e, e->P[b], 3 * rs * N);
        rc |= dalloc(e, e->slot[b], 4 * N);
        rc |= dalloc(e, e->cell[b], 4 * N);
        if (e->integ == MDX_INTEG_ROTOR) {
            rc |= dalloc(e, e->Q[b], (size_t)(rs == 8 ? 72 : 16) * N);
            rc |= dalloc(e, e->Lm[b], 3 * rs * N);
            rc |= dalloc(e, e->Fc[b], 3 * rs * N);
            rc |= dalloc(e, e->Tc[b], 3 * rs * N);
        }
    }
    rc |= dalloc(e, e->rank, 4 * N);
    rc |= dalloc(e, e->list, 4 * N);
    rc |= dalloc(e, e->ctl, sizeof(StepCtl));
    if (!rc) {
        StepCtl c0;
        memset(&c0, 0, sizeof c0);
        c0.fail_epoch = ~0ull;
        rc |= upload(e, e->ctl, &c0, 1);
    }
    if (rc) {
        std::string msg = g_err;
        mdx_destroy(e);
        return set_err(MDX_ERR_HIP, "device allocation failed (%s)", msg.c_str());
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
    Buf *bufs[] = {&e->Fout,    &e->Tout,    &e->obs,    &e->dreps,  &e->prows,   &e->pdirs_d,
                   &e->pdirs_f,
                   &e->rank,    &e->count,   &e->start,  &e->list,   &e->bsum,
                   &e->bond_off, &e->bond_j, &e->bond_k, &e->bond_r0, &e->ang_off, &e->ang_ref,
                   &e->ang_ijk, &e->ang_k,   &e->ang_t0, &e->exc_off, &e->exc_j,  &e->act_prev,
                   &e->act_next, &e->noise,  &e->ctl,    &e->partial, &e->tmpd,   &e->tmpi,
                   &e->tmpi2,   &e->tmpi3,   &e->tmpll,  &e->tmpx,    &e->Fw,
                   &e->ncand,   &e->cand,    &e->nbr,    &e->nbrc,   &e->nn,     &e->bcell,
                   &e->dslot,   &e->Tp,      &e->bpoff,  &e->bowner, &e->bisbound, &e->bboundto,
                   &e->badj_n,  &e->badj_p,  &e->badj_s, &e->bF,     &e->bT,      &e->bscr,
                   &e->bnscr,   &e->bspecies, &e->bdtab, &e->bttab,  &e->btr4,
                   &e->plpairs, &e->plmask};
    for (Buf *b : bufs)
        dfree(*b);
    for (GraphSlot &g : e->graphs)
        if (g.exec)
            (void)hipGraphExecDestroy(g.exec);
    for (int b = 0; b < 2; b++) {
        dfree(e->X[b]);
        dfree(e->Q[b]);
        dfree(e->P[b]);
        dfree(e->Lm[b]);
        dfree(e->Fc[b]);
        dfree(e->Tc[b]);
        dfree(e->slot[b]);
        dfree(e->cell[b]);
    }
    if (e->stream)
        (void)hipStreamDestroy(e->stream);
    delete e;
}

template <class R>
static int upload_state_t(mdx_engine *e, const double *x, const double *p, const double *Q,
                          const double *L)
{
    const size_t N = e->N;
    e->cur = 0;
    std::vector<PosRec> rec(N);
    std::vector<int> slot(N);
    for (size_t i = 0; i < N; i++) {
        rec[i].x = x[3 * i];
        rec[i].y = x[3 * i + 1];
        rec[i].z = x[3 * i + 2];
        rec[i].id = (uint32_t)i;
        rec[i].meta = e->meta[i];
        slot[i] = (int)i;
    }
    int rc;
    if ((rc = upload(e, e->X[0], rec.data(), N)) || (rc = upload(e, e->slot[0], slot.data(), N)))
        return rc;
    std::vector<double> z;
    if (!p || (!L && e->integ == MDX_INTEG_ROTOR))
        z.assign(3 * N, 0.0);
    if ((rc = upload_real<R>(e, e->P[0], p ? p : z.data(), 3 * N)))
        return rc;
    if (e->integ == MDX_INTEG_ROTOR) {
        if ((rc = upload_real<R>(e, e->Lm[0], L ? L : z.data(), 3 * N)))
            return rc;
        std::vector<double> eye;
        if (!Q) {
            eye.assign(9 * N, 0.0);
            for (size_t i = 0; i < N; i++)
                eye[9 * i] = eye[9 * i + 4] = eye[9 * i + 8] = 1.0;
        }
        const double *Qh = Q ? Q : eye.data();
        if (sizeof(R) == 8) {
            if ((rc = upload(e, e->Q[0], Qh, 9 * N)))
                return rc;
        } else {
            // the mixed engine stores unit quaternions
            if ((rc = upload(e, e->tmpd, Qh, 9 * N)))
                return rc;
            hipLaunchKernelGGL(k_mats_to_quats, dim3(nblk((long long)N)), dim3(256), 0, e->stream,
                               (const double *)e->tmpd.p, (int)N, (float *)e->Q[0].p);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(e->stream));
        }
    }
    return MDX_OK;
}

int mdx_upload_state(mdx_engine *e, const double *x, const double *p, const double *Q,
                     const double *L)
{
    if (!e || !x)
        return set_err(MDX_ERR_ARG, "engine and positions required");
    HIPCHK(hipSetDevice(e->device));
    for (long long k = 0; k < 3ll * e->N; k++)
        if (!std::isfinite(x[k]) || (p && !std::isfinite(p[k])) || (L && !std::isfinite(L[k])))
            return set_err(MDX_ERR_ARG, "non-finite position or momentum at particle %lld", k / 3);
    if (Q)
        for (long long k = 0; k < 9ll * e->N; k++)
            if (!std::isfinite(Q[k]))
                return set_err(MDX_ERR_ARG, "non-finite orientation at particle %lld", k / 9);
    int rc = e->prec == MDX_PREC_FP64 ? upload_state_t<double>(e, x, p, Q, L)
                                      : upload_state_t<float>(e, x, p, Q, L);
    e->binned = false;
    e->primed = false;
    e->counted = false;
    e->list_valid = false;
    return rc;
}

int mdx_download_state(mdx_engine *e, double *x, double *p, double *Q, double *L)
{
    if (!e)
        return set_err(MDX_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(e->device));
    std::vector<PosRec> rec;
    int rc = slot_ids(e, rec);
    if (rc)
        return rc;
    if (x)
        for (int s = 0; s < e->N; s++) {
            double *o = x + 3 * (size_t)rec[s].id;
            o[0] = rec[s].x;
            o[1] = rec[s].y;
            o[2] = rec[s].z;
        }
    const bool f64 = e->prec == MDX_PREC_FP64;
    const bool rot = e->integ == MDX_INTEG_ROTOR;
    const int c = e->cur;
    if (p && (rc = f64 ? download_by_id<double>(e, e->P[c], rec, 3, p)
                       : download_by_id<float>(e, e->P[c], rec, 3, p)))
        return rc;
    if (Q) {
        if (rot) {
            if (f64) {
                if ((rc = download_by_id<double>(e, e->Q[c], rec, 9, Q)))
                    return rc;
            } else {
                // quaternions as FP64 matrices, then by original index
                if ((rc = dalloc(e, e->tmpx, 72 * (size_t)e->N)))
                    return rc;
                hipLaunchKernelGGL(k_quats_to_mats, dim3(nblk(e->N)), dim3(256), 0, e->stream,
                                   (const float *)e->Q[c].p, e->N, (double *)e->tmpx.p);
                HIPCHK(hipGetLastError());
                if ((rc = download_by_id<double>(e, e->tmpx, rec, 9, Q)))
                    return rc;
            }
        } else {
            for (int i = 0; i < e->N; i++)
                for (int k = 0; k < 9; k++)
                    Q[9 * (size_t)i + k] = (k % 4 == 0) ? 1.0 : 0.0;
        }
    }
    if (L) {
        if (rot) {
            if ((rc = f64 ? download_by_id<double>(e, e->Lm[c], rec, 3, L)
                          : download_by_id<float>(e, e->Lm[c], rec, 3, L)))
                return rc;
        } else {
            std::fill(L, L + 3 * (size_t)e->N, 0.0);
        }
    }
    return MDX_OK;
}

int mdx_download_forces(mdx_engine *e, double *F, double *T)
{
    if (!e)
        return set_err(MDX_ERR_ARG, "null engine");
    if (!e->primed || e->integ != MDX_INTEG_ROTOR)
        return set_err(MDX_ERR_STATE, "no carried forces (rotor engine: prime or step first)");
    HIPCHK(hipSetDevice(e->device));
    std::vector<PosRec> rec;
    int rc = slot_ids(e, rec);
    if (rc)
        return rc;
    const bool f64 = e->prec == MDX_PREC_FP64;
    const int c = e->cur;
    if (F && (rc = f64 ? download_by_id<double>(e, e->Fc[c], rec, 3, F)
                       : download_by_id<float>(e, e->Fc[c], rec, 3, F)))
        return rc;
    if (T && (rc = f64 ? download_by_id<double>(e, e->Tc[c], rec, 3, T)
                       : download_by_id<float>(e, e->Tc[c], rec, 3, T)))
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
    e->costm_min = 1.0;
    for (size_t k = 0; k < rows.size(); k++) {
        PatchRow &r = rows[k];
        r.eps = e->prow_raw[3 * k];
        r.d = e->prow_raw[3 * k + 1];
        r.thetam = e->prow_raw[3 * k + 2];
        r.costm = std::cos(r.thetam);
        r.kk = PI / (2 * r.thetam);
        r.idist = 1.4 * r.d;
        e->costm_min = std::min(e->costm_min, r.costm);
    }
    // a type pair is active when any of its rows in any replica has eps != 0
    e->active_mask = 0;
    for (int ta = 0; ta < e->ntypes; ta++)
        for (int tc = ta; tc < e->ntypes; tc++) {
            const int n = e->np[ta] * e->np[tc];
            bool on = false;
            for (int r = 0; r < e->nrep && !on; r++)
                for (int k = 0; k < n; k++)
                    on |= e->prow_raw[3 * ((size_t)r * e->nrows + e->rstart[ta * MAXT + tc] + k)] != 0.0;
            if (on && n > 0)
                e->active_mask |= 1ull << (ta * MAXT + tc);
        }
    e->grid_dirty = true;
    return upload(e, e->prows, rows.data(), std::max<size_t>(rows.size(), 1));
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
    int tot = 0;
    for (int t = 0; t < ntypes; t++) {
        if (npatch_per_type[t] < 0 || npatch_per_type[t] > MAXP)
            return set_err(MDX_ERR_ARG, "type %d: 0..%d patches supported", t, MAXP);
        tot += npatch_per_type[t];
    }
    if (tot > 0 && !dirs)
        return set_err(MDX_ERR_ARG, "patch directions required");
    for (int k = 0; k < tot; k++) {
        const double n2 = dirs[3 * k] * dirs[3 * k] + dirs[3 * k + 1] * dirs[3 * k + 1] +
                          dirs[3 * k + 2] * dirs[3 * k + 2];
        // KernFrenkelOnePatch2 constructor check (potentialtheta.h:415-422)
        if (std::fabs(n2 - 1) > 1E-5)
            return set_err(MDX_ERR_ARG, "patch %d is not a unit vector", k);
    }
    for (int i = 0; i < e->N; i++)
        if (type_of_particle[i] < 0 || type_of_particle[i] >= ntypes)
            return set_err(MDX_ERR_ARG, "particle %d has type %d outside 0..%d", i,
                           type_of_particle[i], ntypes - 1);
    int hj = 0;
    std::vector<int> rstart(MAXT * MAXT, 0);
    for (int i = 0; i < ntypes; i++)
        for (int j = i; j < ntypes; j++) {
            rstart[i * MAXT + j] = hj;
            hj += npatch_per_type[i] * npatch_per_type[j];
        }
    for (int k = 0; k < hj; k++)
        if (!(params[3 * k + 1] > 0) || !(params[3 * k + 2] > 0))
            return set_err(MDX_ERR_ARG, "patch row %d: d and theta_m must be > 0", k);
    e->ntypes = ntypes;
    e->bind_on = 0; // the patch numbering changes: mdx_set_binding again
    e->np.assign(npatch_per_type, npatch_per_type + ntypes);
    e->pstart.assign(ntypes, 0);
    for (int t = 1; t < ntypes; t++)
        e->pstart[t] = e->pstart[t - 1] + e->np[t - 1];
    e->rstart = rstart;
    e->nrows = hj;
    e->pdirs.assign(dirs ? dirs : params, dirs ? dirs + 3 * tot : params);
    e->prow_raw.clear();
    for (int r = 0; r < e->nrep; r++)
        e->prow_raw.insert(e->prow_raw.end(), params, params + 3 * hj);
    e->maxd2 = maxd2;
    e->n_patchy = 0;
    for (int i = 0; i < e->N; i++)
        e->n_patchy += npatch_per_type[type_of_particle[i]] > 0;
    for (int r = 0; r < e->nrep; r++)
        for (int k = 0; k < e->nper[r]; k++) {
            const int i = e->first[r] + k;
            e->meta[i] = ((uint32_t)r << 8) | (uint32_t)type_of_particle[i];
        }
    int rc;
    if ((rc = upload(e, e->tmpi, e->meta.data(), e->meta.size())))
        return rc;
    hipLaunchKernelGGL(k_set_meta, dim3(nblk(e->N)), dim3(256), 0, e->stream,
                       (PosRec *)e->X[e->cur].p, e->N, (const uint32_t *)e->tmpi.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    std::vector<double> pd = e->pdirs;
    pd.resize(std::max<size_t>(pd.size(), 3), 0.0);
    std::vector<float> pf(pd.begin(), pd.end());
    if ((rc = upload(e, e->pdirs_f, pf.data(), pf.size())))
        return rc;
    if ((rc = upload(e, e->pdirs_d, pd.data(), pd.size())))
        return rc;
    return patch_rows_upload(e);
}

static void bind_limits(const mdx_engine *e, double &mk, double &cmin);

int mdx_set_replica_patch_params(mdx_engine *e, int replica, const double *params)
{
    if (!e || replica < 0 || replica >= e->nrep || !params || e->ntypes == 0)
        return set_err(MDX_ERR_ARG, "set patches first; replica out of range");
    for (int k = 0; k < e->nrows; k++)
        if (!(params[3 * k + 1] > 0) || !(params[3 * k + 2] > 0))
            return set_err(MDX_ERR_ARG, "patch row %d: d and theta_m must be > 0", k);
    std::copy(params, params + 3 * e->nrows, e->prow_raw.begin() + (size_t)replica * 3 * e->nrows);
    HIPCHK(hipSetDevice(e->device));
    if (e->bind_on) {
        const double mk0 = e->bind_mk;
        bind_limits(e, e->bind_mk, e->bind_cmin);
        if (e->bind_mk != mk0) {
            e->grid_dirty = true;
            e->binned = false;
            e->list_valid = false;
        }
    }
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
        const int i = idx[k];
        if (i < 0 || i >= e->N)
            return set_err(MDX_ERR_ARG, "entry %lld: particle %d out of range", row, i);
        const int r = (int)(std::upper_bound(e->first.begin(), e->first.end(), i) - e->first.begin()) - 1;
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
    if (nbonds > (1ll << 29))
        return set_err(MDX_ERR_ARG, "too many bonds");
    HIPCHK(hipSetDevice(e->device));
    std::vector<std::pair<int, int>> ent;
    ent.reserve(2 * nbonds);
    for (long long b = 0; b < nbonds; b++) {
        int rc = check_same_replica(e, ij + 2 * b, 2, b);
        if (rc)
            return rc;
        if (ij[2 * b] == ij[2 * b + 1])
            return set_err(MDX_ERR_ARG, "bond %lld joins a particle to itself", b);
        if (kind == MDX_BOND_FENE && !(r0[b] > 0))
            return set_err(MDX_ERR_ARG, "bond %lld: FENE R0 must be > 0", b);
        ent.push_back({ij[2 * b], (int)b});
        ent.push_back({ij[2 * b + 1], (int)b});
    }
    std::vector<int> off, val;
    csr(e->N, ent, off, val);
    std::vector<int> bj(val.size());
    std::vector<double> bk(val.size()), br(val.size());
    for (int i = 0; i < e->N; i++)
        for (int q = off[i]; q < off[i + 1]; q++) {
            const long long b = val[q];
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

int mdx_clear_interactions(mdx_engine *e, int terms)
{
    if (!e || terms < 0)
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (terms & MDX_TERM_WCA)
        e->wca = 0;
    if (terms & MDX_TERM_PATCHES) {
        e->ntypes = 0;
        e->active_mask = 0;
        e->bind_on = 0; // (binding needs the patches: mdx_set_binding again)
    }
    if (terms & MDX_TERM_BONDS)
        e->nbonds = 0;
    if (terms & MDX_TERM_ANGLES)
        e->nangles = 0;
    if (terms & MDX_TERM_EXCLUSIONS)
        e->nexcl = 0;
    if (terms & MDX_TERM_ACTIVE)
        e->has_active = 0;
    e->grid_dirty = true;
    return MDX_OK;
}

int mdx_set_force_guard(mdx_engine *e, double fmax)
{
    if (!e || !(fmax >= 0))
        return set_err(MDX_ERR_ARG, "guard must be >= 0");
    e->fguard = fmax;
    return MDX_OK;
}

// positions of replica r in original order into e->tmpx (count x 3)
static int replica_positions(mdx_engine *e, int replica)
{
    const int cnt = e->nper[replica];
    int rc = dalloc(e, e->tmpx, 24 * (size_t)cnt);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_positions_by_id, dim3(nblk(e->N)), dim3(256), 0, e->stream,
                       (const PosRec *)e->X[e->cur].p, e->N, e->first[replica], cnt,
                       (double *)e->tmpx.p);
    HIPCHK(hipGetLastError());
    return MDX_OK;
}

int mdx_cell_assign(mdx_engine *e, int replica, const int n[3], int *cell_of)
{
    if (!e || replica < 0 || replica >= e->nrep || !n || !cell_of)
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (n[0] < 1 || n[1] < 1 || n[2] < 1)
        return set_err(MDX_ERR_ARG, "n must be >= 1");
    HIPCHK(hipSetDevice(e->device));
    const int cnt = e->nper[replica];
    int rc;
    if ((rc = replica_positions(e, replica)) || (rc = dalloc(e, e->tmpi, 4 * (size_t)cnt)))
        return rc;
    const double *b = &e->box[3 * replica];
    hipLaunchKernelGGL(k_assign_custom, dim3(nblk(cnt)), dim3(256), 0, e->stream,
                       (const double *)e->tmpx.p, cnt, b[0], b[1], b[2], n[0], n[1], n[2],
                       (int *)e->tmpi.p);
    HIPCHK(hipGetLastError());
    return download(e, e->tmpi, cell_of, (size_t)cnt);
}

// positions of replica r in original order (tmpx), their cells on an
// n[0] x n[1] x n[2] grid (tmpi) and that grid's cell list (tmpi3: starts,
// tmpi2: members in original order); a parity / analysis path, not the step
static int custom_cells(mdx_engine *e, int replica, const int n[3])
{
    const long long nc = (long long)n[0] * n[1] * n[2];
    if (nc > (1ll << 28))
        return set_err(MDX_ERR_ARG, "too many cells");
    const int cnt = e->nper[replica];
    const double *b = &e->box[3 * replica];
    int rc;
    if ((rc = replica_positions(e, replica)) || (rc = dalloc(e, e->tmpi, 4 * (size_t)cnt)) ||
        (rc = dalloc(e, e->tmpi2, 4 * (size_t)cnt)) || (rc = dalloc(e, e->tmpi3, 4 * (size_t)(nc + 1))))
        return rc;
    hipLaunchKernelGGL(k_assign_custom, dim3(nblk(cnt)), dim3(256), 0, e->stream,
                       (const double *)e->tmpx.p, cnt, b[0], b[1], b[2], n[0], n[1], n[2],
                       (int *)e->tmpi.p);
    HIPCHK(hipGetLastError());
    std::vector<int> cellv(cnt), cstart(nc + 1, 0), clist(cnt);
    if ((rc = download(e, e->tmpi, cellv.data(), (size_t)cnt)))
        return rc;
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
    return MDX_OK;
}

int mdx_pair_histogram(mdx_engine *e, int replica, double max, double dx, long long *counts,
                       int cap, int *nbins)
{
    if (!e || replica < 0 || replica >= e->nrep || !nbins || cap < 0 || (cap > 0 && !counts))
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (!(max > 0) || !(dx > 0))
        return set_err(MDX_ERR_ARG, "max and dx must be > 0");
    const double gsd = std::floor(max / dx);
    if (gsd > GMAXB)
        return set_err(MDX_ERR_ARG, "at most %d bins", GMAXB);
    const int gs = (int)gsd;
    HIPCHK(hipSetDevice(e->device));
    // calcg1: floor(l / max) cells per side (PairCorrelation.cpp:14)
    const double *b = &e->box[3 * replica];
    int n[3];
    for (int k = 0; k < 3; k++) {
        n[k] = (int)std::floor(b[k] / max);
        if (n[k] < 1)
            return set_err(MDX_ERR_ARG, "max larger than the box");
    }
    int rc;
    if ((rc = custom_cells(e, replica, n)) || (rc = dalloc(e, e->tmpll, 8 * (size_t)std::max(gs, 1))))
        return rc;
    HIPCHK(hipMemsetAsync(e->tmpll.p, 0, 8 * (size_t)std::max(gs, 1), e->stream));
    const int cnt = e->nper[replica];
    hipLaunchKernelGGL(k_gcount, dim3(nblk(cnt)), dim3(256), 0, e->stream,
                       (const double *)e->tmpx.p, cnt, b[0], b[1], b[2], e->periodic, n[0], n[1],
                       n[2], (const int *)e->tmpi.p, (const int *)e->tmpi3.p,
                       (const int *)e->tmpi2.p, dx, gs, (unsigned long long *)e->tmpll.p);
    HIPCHK(hipGetLastError());
    *nbins = gs;
    std::vector<long long> h((size_t)std::max(gs, 1));
    if ((rc = download(e, e->tmpll, h.data(), h.size())))
        return rc;
    for (int k = 0; k < std::min(gs, cap); k++)
        counts[k] = h[(size_t)k];
    if (cap < gs)
        return set_err(MDX_ERR_CAPACITY, "%d bins, room for %d", gs, cap);
    return MDX_OK;
}

template <class R> static int patch_bonds_t(mdx_engine *e, int *bonds)
{
    int rc;
    if (!e->binned && (rc = bin_current<R>(e)))
        return rc;
    if ((rc = dalloc(e, e->tmpi, 4 * (size_t)e->N)))
        return rc;
    KArgs<R> a = make_args<R>(e);
    hipLaunchKernelGGL(k_patch_bonds<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, a,
                       (int *)e->tmpi.p);
    HIPCHK(hipGetLastError());
    return download(e, e->tmpi, bonds, (size_t)e->N);
}

int mdx_patch_bonds(mdx_engine *e, int *bonds)
{
    if (!e || !bonds)
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (e->integ != MDX_INTEG_ROTOR)
        return set_err(MDX_ERR_STATE, "patch bonds need the rotor integrator");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? patch_bonds_t<double>(e, bonds)
                                    : patch_bonds_t<float>(e, bonds);
}

// the pairs of MD::calculatepairs as a set (k_pairs: count pass, offsets,
// write pass) into the device buffer out, at most cap of them; *npairs = all
static int pair_list(mdx_engine *e, int replica, const int n[3], double cut, Buf &out,
                     long long cap, long long *npairs)
{
    const int cnt = e->nper[replica];
    const double *b = &e->box[3 * replica];
    int rc;
    if ((rc = custom_cells(e, replica, n)) || (rc = dalloc(e, e->tmpll, 8 * (size_t)(cnt + 1))) ||
        (rc = dalloc(e, e->partial, 4 * (size_t)cnt)))
        return rc;
    const double *X = (const double *)e->tmpx.p;
    hipLaunchKernelGGL(k_pairs, dim3(nblk(cnt)), dim3(256), 0, e->stream, X, cnt, b[0], b[1],
                       b[2], e->periodic, n[0], n[1], n[2], cut, (const int *)e->tmpi.p,
                       (const int *)e->tmpi3.p, (const int *)e->tmpi2.p, (int *)e->partial.p,
                       (const long long *)nullptr, (int *)nullptr, 0ll, 0);
    HIPCHK(hipGetLastError());
    std::vector<int> pc(cnt);
    if ((rc = download(e, e->partial, pc.data(), (size_t)cnt)))
        return rc;
    std::vector<long long> off(cnt + 1, 0);
    for (int k = 0; k < cnt; k++)
        off[k + 1] = off[k] + pc[k];
    *npairs = off[cnt];
    const long long w = std::min(cap, off[cnt]);
    if (w > 0) {
        if ((rc = upload(e, e->tmpll, off.data(), off.size())) ||
            (rc = dalloc(e, out, 8 * (size_t)w)))
            return rc;
        hipLaunchKernelGGL(k_pairs, dim3(nblk(cnt)), dim3(256), 0, e->stream, X, cnt, b[0],
                           b[1], b[2], e->periodic, n[0], n[1], n[2], cut, (const int *)e->tmpi.p,
                           (const int *)e->tmpi3.p, (const int *)e->tmpi2.p, (int *)nullptr,
                           (const long long *)e->tmpll.p, (int *)out.p, w, 1);
        HIPCHK(hipGetLastError());
    }
    return MDX_OK;
}

static void release(mdx_engine *e, Buf &b)
{
    e->bytes_held -= b.bytes;
    dfree(b);
}

int mdx_neighbor_pairs(mdx_engine *e, int replica, const int n[3], double cut, int *ij,
                       long long cap, long long *npairs)
{
    if (!e || replica < 0 || replica >= e->nrep || !n || !npairs || cap < 0 || (cap > 0 && !ij))
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (n[0] < 1 || n[1] < 1 || n[2] < 1 || !(cut > 0))
        return set_err(MDX_ERR_ARG, "n >= 1 and cut > 0 required");
    HIPCHK(hipSetDevice(e->device));
    Buf out;
    int rc = pair_list(e, replica, n, cut, out, cap, npairs);
    const long long w = std::min(cap, *npairs);
    if (rc == MDX_OK && w > 0) {
        hipError_t le = hipMemcpyAsync(ij, out.p, 8 * (size_t)w, hipMemcpyDeviceToHost, e->stream);
        if (le == hipSuccess)
            le = hipStreamSynchronize(e->stream);
        if (le != hipSuccess)
            rc = set_err(MDX_ERR_HIP, "pair listing failed: %s", hipGetErrorString(le));
    }
    release(e, out);
    if (rc)
        return rc;
    if (*npairs > cap)
        return set_err(MDX_ERR_CAPACITY, "%lld pairs, capacity %lld", *npairs, cap);
    return MDX_OK;
}

int mdx_clusters(mdx_engine *e, int replica, const int n[3], double cut, int ex_lo, int ex_hi,
                 int *labels, int *nclusters)
{
    if (!e || replica < 0 || replica >= e->nrep || !n || !labels)
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (n[0] < 1 || n[1] < 1 || n[2] < 1 || !(cut > 0))
        return set_err(MDX_ERR_ARG, "n >= 1 and cut > 0 required");
    HIPCHK(hipSetDevice(e->device));
    const int cnt = e->nper[replica];
    Buf pairs, lab, flag;
    long long np = 0;
    int rc = pair_list(e, replica, n, cut, pairs, (long long)1 << 40, &np);
    if (!rc && !(rc = dalloc(e, lab, 4 * (size_t)cnt)) && !(rc = dalloc(e, flag, 4))) {
        hipLaunchKernelGGL(k_cc_init, dim3(nblk(cnt)), dim3(256), 0, e->stream, (int *)lab.p, cnt);
        // every round changes at least one label and labels only decrease:
        // the rounds end; the cap only guards against a broken invariant
        const int max_rounds = 100000;
        int round = 0;
        for (; round < max_rounds; round++) {
            hipError_t le = hipMemsetAsync(flag.p, 0, 4, e->stream);
            if (le != hipSuccess) {
                rc = set_err(MDX_ERR_HIP, "%s", hipGetErrorString(le));
                break;
            }
            if (np > 0)
                hipLaunchKernelGGL(k_cc_hook, dim3(nblk(np)), dim3(256), 0, e->stream,
                                   (const int *)pairs.p, np, ex_lo, ex_hi, (int *)lab.p,
                                   (int *)flag.p);
            hipLaunchKernelGGL(k_cc_jump, dim3(nblk(cnt)), dim3(256), 0, e->stream,
                               (int *)lab.p, cnt, (int *)flag.p);
            int changed = 0;
            if ((le = hipGetLastError()) != hipSuccess ||
                (le = hipMemcpyAsync(&changed, flag.p, 4, hipMemcpyDeviceToHost, e->stream)) !=
                    hipSuccess ||
                (le = hipStreamSynchronize(e->stream)) != hipSuccess) {
                rc = set_err(MDX_ERR_HIP, "connected components: %s", hipGetErrorString(le));
                break;
            }
            if (!changed)
                break;
        }
        if (!rc && round == max_rounds)
            rc = set_err(MDX_ERR_STATE, "connected components did not settle");
        if (!rc)
            rc = download(e, lab, labels, (size_t)cnt);
    }
    release(e, pairs);
    release(e, lab);
    release(e, flag);
    if (rc)
        return rc;
    int nc = 0;
    for (int i = 0; i < cnt; i++)
        nc += labels[i] == i;
    if (nclusters)
        *nclusters = nc;
    return MDX_OK;
}

template <class R> static BindArgs bind_args(mdx_engine *e, unsigned long long draw);

template <class R> static int compute_forces_t(mdx_engine *e, double *F, double *T)
{
    int rc;
    if (!e->binned && (rc = bin_current<R>(e)))
        return rc;
    if ((rc = dalloc(e, e->Fout, 24 * (size_t)e->N)) || (rc = dalloc(e, e->Tout, 24 * (size_t)e->N)))
        return rc;
    if ((rc = launch_gather<R, M_CONS>(e, nullptr)))
        return rc;
    if ((rc = download(e, e->Fout, F, 3 * (size_t)e->N)))
        return rc;
    if (T && (rc = download(e, e->Tout, T, 3 * (size_t)e->N)))
        return rc;
    if (e->bind_on) {
        // with binding: the bound pairs' terms of the current bound state
        // (no update), as the step adds them
        hipLaunchKernelGGL(k_bind_forces<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream,
                           make_args<R>(e), bind_args<R>(e, 0), (double *)e->bF.p,
                           (double *)e->bT.p, 1, (const int *)nullptr);
        HIPCHK(hipGetLastError());
        std::vector<double> fb(3 * (size_t)e->N), tb(3 * (size_t)e->N);
        if ((rc = download(e, e->bF, fb.data(), fb.size())) ||
            (rc = download(e, e->bT, tb.data(), tb.size())))
            return rc;
        for (size_t k = 0; k < fb.size(); k++) {
            F[k] += fb[k];
            if (T)
                T[k] += tb[k];
        }
    }
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

// ------------------------------------------------------------ binding

// ComboPatch::max_check (combopatch.h:15, GeneralPatch.cpp:93-94): 1.0
// raised to the largest 1.4 d of any row (potentialtheta.h:417); and the
// smallest cos(theta_m) of any row (the binding kernel's cone screen)
static void bind_limits(const mdx_engine *e, double &mk, double &cmin)
{
    mk = 1.0;
    cmin = 1.0;
    for (size_t k = 0; k < e->prow_raw.size() / 3; k++) {
        mk = std::max(mk, 1.4 * e->prow_raw[3 * k + 1]);
        cmin = std::min(cmin, std::cos(e->prow_raw[3 * k + 2]));
    }
}

int mdx_set_binding(mdx_engine *e, double on_rate, double off_rate, int *npatches)
{
    if (!e)
        return set_err(MDX_ERR_ARG, "null engine");
    if (e->integ != MDX_INTEG_ROTOR || e->ntypes == 0)
        return set_err(MDX_ERR_STATE, "binding needs the rotor integrator and patches");
    if (!(on_rate >= 0) || !(off_rate >= 0) || !(on_rate + off_rate > 0))
        return set_err(MDX_ERR_ARG, "rates must be >= 0 with a positive sum");
    HIPCHK(hipSetDevice(e->device));
    std::vector<int> poff(e->N + 1, 0);
    for (int i = 0; i < e->N; i++)
        poff[i + 1] = poff[i] + e->np[e->meta[i] & 255];
    const int nh = poff[e->N];
    if (nh == 0)
        return set_err(MDX_ERR_STATE, "no particle carries patches");
    std::vector<int> owner(nh);
    for (int i = 0; i < e->N; i++)
        for (int k = poff[i]; k < poff[i + 1]; k++)
            owner[k] = i;
    double mk, cmin;
    bind_limits(e, mk, cmin);
    int rc;
    if ((rc = upload(e, e->bpoff, poff.data(), poff.size())) ||
        (rc = upload(e, e->bowner, owner.data(), owner.size())) ||
        (rc = dalloc(e, e->bisbound, 4 * (size_t)nh)) ||
        (rc = dalloc(e, e->bboundto, 4 * (size_t)nh)) ||
        (rc = dalloc(e, e->badj_n, 4 * (size_t)nh)) ||
        (rc = dalloc(e, e->badj_p, 4 * (size_t)nh * BIND_DEG)) ||
        (rc = dalloc(e, e->badj_s, 8 * (size_t)nh * BIND_DEG)) ||
        (rc = dalloc(e, e->bF, 24 * (size_t)e->N)) || (rc = dalloc(e, e->bT, 24 * (size_t)e->N)) ||
        (rc = dalloc(e, e->bscr, 4 * (size_t)BIND_SCR * e->N)) ||
        (rc = dalloc(e, e->bnscr, 4 * (size_t)e->N)))
        return rc;
    HIPCHK(hipMemsetAsync(e->bisbound.p, 0, 4 * (size_t)nh, e->stream));
    HIPCHK(hipMemsetAsync(e->bboundto.p, 0, 4 * (size_t)nh, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (!e->bind_on || e->bind_mk != mk) {
        e->grid_dirty = true;
        e->binned = false;
        e->list_valid = false;
    }
    e->bind_on = 1;
    e->bind_table = 0;
    e->bind_on_rate = on_rate;
    e->bind_off_rate = off_rate;
    e->bind_mk = mk;
    e->bind_cmin = cmin;
    e->nh = nh;
    if (npatches)
        *npatches = nh;
    return MDX_OK;
}

int mdx_set_binding_tables(mdx_engine *e, const int *species, const double *dtab,
                           const double *ttab, const double *tr4)
{
    if (!e || !species || !dtab || !ttab || !tr4)
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (!e->bind_on)
        return set_err(MDX_ERR_STATE, "call mdx_set_binding first");
    for (int k = 0; k < e->nh; k++)
        if (species[k] < 1 || species[k] > 3)
            return set_err(MDX_ERR_ARG, "species[%d] = %d outside 1..3", k, species[k]);
    for (int k = 0; k < 24; k++)
        if (!(dtab[k] >= 0))
            return set_err(MDX_ERR_ARG, "doublet rates must be >= 0");
    for (int k = 0; k < 160; k++)
        if (!(ttab[k] >= 0))
            return set_err(MDX_ERR_ARG, "triplet rates must be >= 0");
    for (int k = 0; k < 40; k++)
        if (!(tr4[k] >= 0))
            return set_err(MDX_ERR_ARG, "triplet rates must be >= 0");
    HIPCHK(hipSetDevice(e->device));
    int rc;
    if ((rc = upload(e, e->bspecies, species, (size_t)e->nh)) ||
        (rc = upload(e, e->bdtab, dtab, 24)) || (rc = upload(e, e->bttab, ttab, 160)) ||
        (rc = upload(e, e->btr4, tr4, 40)))
        return rc;
    e->bind_table = 1;
    return MDX_OK;
}

int mdx_set_bindings(mdx_engine *e, const int *isbound, const int *boundto)
{
    if (!e || !isbound || !boundto)
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (!e->bind_on)
        return set_err(MDX_ERR_STATE, "call mdx_set_binding first");
    for (int k = 0; k < e->nh; k++)
        if (boundto[k] < 0 || boundto[k] >= e->nh)
            return set_err(MDX_ERR_ARG, "boundto[%d] = %d outside 0..%d", k, boundto[k], e->nh - 1);
    HIPCHK(hipSetDevice(e->device));
    std::vector<int> ib(isbound, isbound + e->nh);
    for (int &v : ib)
        v = v != 0;
    int rc;
    if ((rc = upload(e, e->bisbound, ib.data(), ib.size())) ||
        (rc = upload(e, e->bboundto, boundto, (size_t)e->nh)))
        return rc;
    return MDX_OK;
}

int mdx_get_bindings(mdx_engine *e, int *isbound, int *boundto)
{
    if (!e || !isbound || !boundto)
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (!e->bind_on)
        return set_err(MDX_ERR_STATE, "call mdx_set_binding first");
    HIPCHK(hipSetDevice(e->device));
    int rc;
    if ((rc = download(e, e->bisbound, isbound, (size_t)e->nh)) ||
        (rc = download(e, e->bboundto, boundto, (size_t)e->nh)))
        return rc;
    return MDX_OK;
}

template <class R> static BindArgs bind_args(mdx_engine *e, unsigned long long draw)
{
    BindArgs b;
    b.poff = (const int *)e->bpoff.p;
    b.owner = (const int *)e->bowner.p;
    b.isbound = (int *)e->bisbound.p;
    b.boundto = (int *)e->bboundto.p;
    b.adj_n = (int *)e->badj_n.p;
    b.adj_p = (int *)e->badj_p.p;
    b.adj_s = (double *)e->badj_s.p;
    b.nh = e->nh;
    b.mk2 = e->bind_mk * e->bind_mk;
    b.cmin = e->bind_cmin;
    b.on_rate = e->bind_on_rate;
    b.off_rate = e->bind_off_rate;
    // the noise key with its second word changed: binding draws and the
    // Langevin noise are different streams
    b.key0 = (uint32_t)e->seed;
    b.key1 = (uint32_t)(e->seed >> 32) ^ 0xB1D0B1D0u;
    b.draw = draw;
    b.ctl = (StepCtl *)e->ctl.p;
    b.table = e->bind_table;
    b.species = (const int *)e->bspecies.p;
    b.dtab = (const double *)e->bdtab.p;
    b.ttab = (const double *)e->bttab.p;
    b.tr4 = (const double *)e->btr4.p;
    return b;
}

// the four binding kernels on the current state (the cell list built, or
// the neighbour list in list steps: a.lmode); draw_epoch: the draw counter
// is the force-evaluation count held on the device
// (cand, ncand, cap: the candidates the pair walk collected; null: the
// screen kernel collects them)
template <class R>
static int launch_bind(mdx_engine *e, const KArgs<R> &a, unsigned long long draw, int draw_epoch,
                       const int *cand = nullptr, const int *ncand = nullptr, int cap = 0)
{
    BindArgs b = bind_args<R>(e, draw);
    b.draw_epoch = draw_epoch;
    hipLaunchKernelGGL(k_bind_clear, dim3(nblk(e->nh)), dim3(256), 0, e->stream, b.adj_n, e->nh);
    if (!cand) {
        hipLaunchKernelGGL(k_bind_screen<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, a, b,
                           (int *)e->bscr.p, (int *)e->bnscr.p);
        cand = (const int *)e->bscr.p;
        ncand = (const int *)e->bnscr.p;
        cap = BIND_SCR;
    }
    hipLaunchKernelGGL(k_bind_edges<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, a, b, cand,
                       ncand, cap);
    hipLaunchKernelGGL(k_bind_decide, dim3(nblk(e->nh)), dim3(256), 0, e->stream, b);
    hipLaunchKernelGGL(k_bind_forces<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, a, b,
                       (double *)e->bF.p, (double *)e->bT.p, draw_epoch ? 0 : 1, ncand);
    HIPCHK(hipGetLastError());
    return MDX_OK;
}

template <class R> static int bind_forces_t(mdx_engine *e, unsigned long long draw, double *F,
                                            double *T)
{
    int rc;
    if (!e->binned && (rc = bin_current<R>(e)))
        return rc;
    if ((rc = launch_bind<R>(e, make_args<R>(e), draw, 0)))
        return rc;
    unsigned int st;
    if ((rc = read_status(e, &st)))
        return rc;
    if ((rc = status_to_rc(st)))
        return rc;
    if (F && (rc = download(e, e->bF, F, 3 * (size_t)e->N)))
        return rc;
    if (T && (rc = download(e, e->bT, T, 3 * (size_t)e->N)))
        return rc;
    return MDX_OK;
}

// a force evaluation of the step with binding (the loop of Condensate::
// run_singlebond, Condensate.cpp:464-478): the pair walk (WCA, no patch
// terms), the binding kernels at the same positions, then the integrator
// kernel, which adds the bound pairs' terms (never the fused forms)
template <class R, int MODE>
static int launch_bind_step(mdx_engine *e, const double *noise_dev, int lmode)
{
    KArgs<R> a = make_args<R>(e);
    a.noise = noise_dev;
    a.lmode = lmode;
    const bool list_walk = lmode && e->use_walkc;
    a.pcap = list_walk ? std::max(PMAXG, e->kmax) : PMAXG;
    int rc;
    if ((rc = dalloc(e, e->Fw, 3 * e->rsz() * (size_t)e->N)) ||
        (rc = dalloc(e, e->ncand, 4 * (size_t)e->N)) ||
        (rc = dalloc(e, e->cand, 4 * (size_t)a.pcap * e->N)) ||
        (rc = dalloc(e, e->Tp, 3 * e->rsz() * (size_t)e->N)))
        return rc;
    a.Fw = (R *)e->Fw.p;
    a.ncand = (int *)e->ncand.p;
    a.cand = (int *)e->cand.p;
    a.Tp = (R *)e->Tp.p;
    // the walk computes WCA and collects the binding candidates: neighbours
    // within max_check of any patch-carrying type pair whose own patches
    // pass an FP32 cone screen at the smallest cos(theta_m) less 1e-5 (the
    // binding kernel decides them exactly); no Kern-Frenkel terms
    KArgs<R> w = a;
    w.has_patches = 1;
    w.maxd2 = e->bind_mk * e->bind_mk;
    w.rmax2 = std::max(a.rmax2, w.maxd2);
    w.costm_screen_f = e->bind_cmin - 1e-5;
    w.active_mask = 0;
    for (int ta = 0; ta < e->ntypes; ta++)
        for (int tc = ta; tc < e->ntypes; tc++)
            if (e->np[ta] > 0 && e->np[tc] > 0)
                w.active_mask |= 1ull << (ta * MAXT + tc);
    w.ovf_fatal = 1;
    if (list_walk)
        hipLaunchKernelGGL((k_walkc<R, MODE, false>), dim3(nblk(e->N)), dim3(256), 0, e->stream,
                           w);
    else
        hipLaunchKernelGGL((k_walk<R, MODE>), dim3(nblk(e->N)), dim3(256), 0, e->stream, w);
    if (e->prof_mid)
        HIPCHK(hipEventRecord(e->prof_mid, e->stream));
    e->last_split = true;
    if ((rc = launch_bind<R>(e, a, 0, 1, (const int *)e->cand.p, (const int *)e->ncand.p,
                             a.pcap)))
        return rc;
    if (e->prof_mid2)
        HIPCHK(hipEventRecord(e->prof_mid2, e->stream));
    // the integrator adds the bound pairs' terms where the walk found
    // binding candidates
    a.bF = (const double *)e->bF.p;
    a.bT = (const double *)e->bT.p;
    a.bncand = (const int *)e->ncand.p;
    hipLaunchKernelGGL((k_integrate<R, MODE>), dim3(nblk(e->N)), dim3(256), 0, e->stream, a);
    HIPCHK(hipGetLastError());
    return MDX_OK;
}

int mdx_bind_forces(mdx_engine *e, unsigned long long draw, double *F, double *T)
{
    if (!e)
        return set_err(MDX_ERR_ARG, "null engine");
    if (!e->bind_on)
        return set_err(MDX_ERR_STATE, "call mdx_set_binding first");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? bind_forces_t<double>(e, draw, F, T)
                                    : bind_forces_t<float>(e, draw, F, T);
}

// calculate_patch_list (bound 0) or calculate_bound_pairs (bound 1) on the
// device (k_patch_list), the masks expanded into rows (p1, p2, potn, patch of
// p1, patch of p2) in pair order, ascending potn
template <class R>
static int patch_list_t(mdx_engine *e, const int *pairs, long long npairs, int *out,
                        long long cap, long long *count, int bound)
{
    int rc;
    if (!e->binned && (rc = bin_current<R>(e)))
        return rc;
    if (npairs == 0)
        return MDX_OK;
    if ((rc = upload(e, e->plpairs, pairs, 2 * (size_t)npairs)) ||
        (rc = dalloc(e, e->plmask, 8 * (size_t)npairs)))
        return rc;
    double mk, cmin;
    bind_limits(e, mk, cmin);
    // ComboPatch::max_ang starts at 0.6 (combopatch.h:15); GeneralPatch
    // lowers it to the smallest cos(theta_m) (GeneralPatch.cpp:98-99)
    const double maxang = std::min(0.6, cmin);
    const KArgs<R> a = make_args<R>(e);
    const unsigned int nb = (unsigned int)((npairs + 255) / 256);
    hipLaunchKernelGGL(k_patch_list<R>, dim3(nb), dim3(256), 0, e->stream, a,
                       (const int *)e->plpairs.p, npairs, bound ? mk : mk + 1.0, maxang, bound,
                       (unsigned long long *)e->plmask.p);
    HIPCHK(hipGetLastError());
    std::vector<unsigned long long> m((size_t)npairs);
    if ((rc = download(e, e->plmask, m.data(), m.size())))
        return rc;
    std::vector<int> poff((size_t)e->N + 1, 0);
    for (int i = 0; i < e->N; i++)
        poff[i + 1] = poff[i] + e->np[e->meta[i] & 255];
    long long n = 0;
    for (long long k = 0; k < npairs; k++) {
        if (!m[k])
            continue;
        const int p1 = std::min(pairs[2 * k], pairs[2 * k + 1]);
        const int p2 = std::max(pairs[2 * k], pairs[2 * k + 1]);
        const int t1 = e->meta[p1] & 255, t2 = e->meta[p2] & 255;
        for (int bit = 0; bit < 64; bit++) {
            if (!((m[k] >> bit) & 1ull))
                continue;
            if (n < cap) {
                int *r = out + 5 * n;
                r[0] = p1;
                r[1] = p2;
                r[2] = a.rstart[t1 * MAXT + t2] + bit;
                r[3] = poff[p1] + bit / e->np[t2];
                r[4] = poff[p2] + bit % e->np[t2];
            }
            n++;
        }
    }
    *count = n;
    if (n > cap)
        return set_err(MDX_ERR_CAPACITY, "%s: %lld rows, room for %lld",
                       bound ? "bound pairs" : "patch list", n, cap);
    return MDX_OK;
}

static int patch_pairs(mdx_engine *e, const int *pairs, long long npairs, int *out,
                       long long cap, long long *count, int bound)
{
    if (!e || !count || npairs < 0 || cap < 0 || (npairs > 0 && !pairs) || (cap > 0 && !out))
        return set_err(MDX_ERR_ARG, "bad arguments");
    *count = 0;
    if (e->integ != MDX_INTEG_ROTOR || e->ntypes == 0)
        return set_err(MDX_ERR_STATE, "the patch list needs the rotor integrator and patches");
    if (e->nrep != 1)
        return set_err(MDX_ERR_STATE, "the patch list is for a single system");
    for (int i = 1; i < e->N; i++)
        if ((e->meta[i] & 255) < (e->meta[i - 1] & 255))
            return set_err(MDX_ERR_STATE,
                           "the patch list needs particles numbered by type (type ranges)");
    for (long long k = 0; k < 2 * npairs; k++)
        if (pairs[k] < 0 || pairs[k] >= e->N)
            return set_err(MDX_ERR_ARG, "pairs[%lld] = %d outside 0..%d", k, pairs[k], e->N - 1);
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64
               ? patch_list_t<double>(e, pairs, npairs, out, cap, count, bound)
               : patch_list_t<float>(e, pairs, npairs, out, cap, count, bound);
}

int mdx_patch_list(mdx_engine *e, const int *pairs, long long npairs, int *out, long long cap,
                   long long *count)
{
    return patch_pairs(e, pairs, npairs, out, cap, count, 0);
}

int mdx_bound_pairs(mdx_engine *e, const int *pairs, long long npairs, int *out, long long cap,
                    long long *count)
{
    return patch_pairs(e, pairs, npairs, out, cap, count, 1);
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
    const int first = replica < 0 ? 0 : e->first[replica];
    const int cnt = replica < 0 ? e->N : e->nper[replica];
    const int nb = (int)std::min(1024u, nblk(cnt));
    if ((rc = dalloc(e, e->partial, 16 * 8 * (size_t)nb)))
        return rc;
    hipLaunchKernelGGL(k_obs<R>, dim3(nb), dim3(256), 0, e->stream, first, cnt,
                       (const R *)e->P[e->cur].p,
                       e->integ == MDX_INTEG_ROTOR ? (const R *)e->Lm[e->cur].p : nullptr,
                       (const double *)e->obs.p, e->m, e->I[0], (double *)e->partial.p);
    HIPCHK(hipGetLastError());
    std::vector<double> part(16 * (size_t)nb);
    if ((rc = download(e, e->partial, part.data(), part.size())))
        return rc;
    for (int k = 0; k < 16; k++) {
        double s = 0;
        for (int b = 0; b < nb; b++)
            s += part[16 * (size_t)b + k];
        out[k] = s;
    }
    for (int k = 0; k < 6; k++)
        out[k] /= (double)cnt;
    unsigned int st;
    if ((rc = read_status(e, &st)))
        return rc;
    return status_to_rc(st & 2u);
}

template <class R> static int stress_t(mdx_engine *e, int replica, double *out)
{
    int rc;
    if (!e->binned && (rc = bin_current<R>(e)))
        return rc;
    if ((rc = dalloc(e, e->obs, 80 * (size_t)e->N)))
        return rc;
    KArgs<R> a = make_args<R>(e);
    hipLaunchKernelGGL(k_stress<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, a,
                       (double *)e->obs.p);
    const int first = replica < 0 ? 0 : e->first[replica];
    const int cnt = replica < 0 ? e->N : e->nper[replica];
    const int nb = (int)std::min(1024u, nblk(cnt));
    if ((rc = dalloc(e, e->partial, 9 * 8 * (size_t)nb)))
        return rc;
    hipLaunchKernelGGL(k_sum_rows<9>, dim3(nb), dim3(256), 0, e->stream, first, cnt,
                       (const double *)e->obs.p, (double *)e->partial.p);
    HIPCHK(hipGetLastError());
    std::vector<double> part(9 * (size_t)nb);
    if ((rc = download(e, e->partial, part.data(), part.size())))
        return rc;
    for (int k = 0; k < 9; k++) {
        double v = 0;
        for (int b = 0; b < nb; b++)
            v += part[9 * (size_t)b + k];
        out[k] = v;
    }
    return MDX_OK;
}

// ------------------------------------------------ no-allocation stress
// Scratch of mdx_stress_noalloc: per-block partial sums of the stress (9
// doubles per block of a grid of at most NOALLOC_BLOCKS blocks) and the final
// 9 sums after them.  One copy per GPU, held until the process ends; host
// code uses the address hipGetSymbolAddress gives for the current device
// (g_scratch_ptr), never the array's name.
constexpr int NOALLOC_BLOCKS = 2048;
static __device__ __align__(256) double scratch[9 * (NOALLOC_BLOCKS + 1)];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[MAX_DEVICES];

// the device address of scratch on device dev (current); once per device
static int scratch_setup(int dev)
{
    if (dev < 0 || dev >= MAX_DEVICES)
        return set_err(MDX_ERR_ARG, "device %d beyond the %d the library tracks", dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev])
        return MDX_OK;
    void *p = nullptr;
    hipError_t he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (he != hipSuccess || !p)
        return set_err(MDX_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                       hipGetErrorString(he));
    g_scratch_ptr[dev] = (double *)p;
    return MDX_OK;
}

template <class R> static int stress_noalloc_t(mdx_engine *e, int replica, double *out)
{
    // the cell grid and its buffers must exist already: building them
    // allocates
    if (e->grid_dirty || !e->count.p || !e->start.p || !e->bsum.p)
        return set_err(MDX_ERR_STATE, "mdx_stress_noalloc needs the cell grid built (call "
                                      "mdx_prime, mdx_step, mdx_stress or mdx_observables first)");
    double *scr = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        scr = g_scratch_ptr[e->device];
    }
    if (!scr)
        return set_err(MDX_ERR_STATE, "scratch not set up on device %d", e->device);
    int rc;
    if (!e->binned && (rc = bin_current<R>(e)))
        return rc;
    KArgs<R> a = make_args<R>(e);
    const int first = replica < 0 ? 0 : e->first[replica];
    const int cnt = replica < 0 ? e->N : e->nper[replica];
    const int nb = (int)std::min<unsigned>(NOALLOC_BLOCKS, nblk(e->N));
    hipLaunchKernelGGL(k_stress_partial<R>, dim3(nb), dim3(256), 0, e->stream, a, first, cnt, scr);
    hipLaunchKernelGGL(k_sum_partials9, dim3(1), dim3(64), 0, e->stream, (const double *)scr, nb,
                       scr + 9 * (size_t)NOALLOC_BLOCKS);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, scr + 9 * (size_t)NOALLOC_BLOCKS, 9 * sizeof(double),
                          hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MDX_OK;
}

int mdx_stress_noalloc(mdx_engine *e, int replica, double *out)
{
    if (!e || !out || replica < -1 || replica >= e->nrep)
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (!e->wca)
        return set_err(MDX_ERR_STATE, "calculatestress needs the pair interaction (mdx_set_pair_wca)");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? stress_noalloc_t<double>(e, replica, out)
                                    : stress_noalloc_t<float>(e, replica, out);
}

int mdx_stress(mdx_engine *e, int replica, double *out)
{
    if (!e || !out || replica < -1 || replica >= e->nrep)
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (!e->wca)
        return set_err(MDX_ERR_STATE, "calculatestress needs the pair interaction (mdx_set_pair_wca)");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? stress_t<double>(e, replica, out)
                                    : stress_t<float>(e, replica, out);
}

int mdx_observables(mdx_engine *e, int replica, double *out)
{
    if (!e || !out || replica < -1 || replica >= e->nrep)
        return set_err(MDX_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? observables_t<double>(e, replica, out)
                                    : observables_t<float>(e, replica, out);
}

// ------------------------------------------------------ noise counter, resume
int mdx_get_noise_counter(mdx_engine *e, unsigned long long *counter)
{
    if (!e || !counter)
        return set_err(MDX_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(e->device));
    StepCtl c;
    HIPCHK(hipMemcpyAsync(&c, e->ctl.p, sizeof c, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    *counter = c.epoch;
    return MDX_OK;
}

int mdx_set_noise_counter(mdx_engine *e, unsigned long long counter)
{
    if (!e)
        return set_err(MDX_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipMemcpyAsync((char *)e->ctl.p + offsetof(StepCtl, epoch), &counter, sizeof counter,
                          hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MDX_OK;
}

template <class R> static int upload_forces_t(mdx_engine *e, const double *F, const double *T)
{
    const size_t n3 = 3 * (size_t)e->N;
    int rc;
    // the step that follows bins the particles: the cell grid (replica table,
    // cell counts) must exist, as mdx_prime would have built it
    if ((rc = build_grid(e)) || (rc = upload(e, e->tmpd, F, n3)) || (rc = upload(e, e->tmpx, T, n3)))
        return rc;
    const int c = e->cur;
    hipLaunchKernelGGL((k_by_id_to_slots<R>), dim3(nblk(e->N)), dim3(256), 0, e->stream,
                       (const PosRec *)e->X[c].p, e->N, (const double *)e->tmpd.p, (R *)e->Fc[c].p,
                       3);
    hipLaunchKernelGGL((k_by_id_to_slots<R>), dim3(nblk(e->N)), dim3(256), 0, e->stream,
                       (const PosRec *)e->X[c].p, e->N, (const double *)e->tmpx.p, (R *)e->Tc[c].p,
                       3);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    e->primed = true;
    return MDX_OK;
}

int mdx_upload_forces(mdx_engine *e, const double *F, const double *T)
{
    if (!e || !F || !T)
        return set_err(MDX_ERR_ARG, "engine, F and T required");
    if (e->integ != MDX_INTEG_ROTOR)
        return set_err(MDX_ERR_STATE, "carried forces belong to the rotor integrator");
    for (long long k = 0; k < 3ll * e->N; k++)
        if (!std::isfinite(F[k]) || !std::isfinite(T[k]))
            return set_err(MDX_ERR_ARG, "non-finite carried force or torque at particle %lld",
                           k / 3);
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? upload_forces_t<double>(e, F, T)
                                    : upload_forces_t<float>(e, F, T);
}

// -------------------------------------------------- the step's neighbour list
// The list mdx_step walks, built at the current positions by the step's own
// kernels (binning, scan, fill, per-cell order, k_nbuild), returned as the
// set of unordered pairs (original indices, i < j, ascending).  The list is
// full (each pair from both sides): both sides are checked to agree.
template <class R>
static int list_pairs_t(mdx_engine *e, int *ij, long long cap, long long *npairs, double *radii)
{
    if (!(e->skin > 0))
        return set_err(MDX_ERR_STATE, "no neighbour list (skin 0)");
    int rc;
    if ((rc = build_grid(e)) || (rc = nbr_alloc(e)))
        return rc;
    if (!e->counted && (rc = count_current<R>(e, (int *)e->cell[e->cur].p, nullptr)))
        return rc;
    if ((rc = finish_bins(e)) || (rc = build_list<R>(e, (const int *)e->cell[e->cur].p, 0)))
        return rc;
    e->counted = false;
    e->binned = true;
    // the step rebuilds its list (and its displacement bound) at its next step
    e->list_valid = false;
    const int N = e->N;
    std::vector<int> nn(N), nb((size_t)e->kmax * N);
    std::vector<PosRec> rec;
    if ((rc = download(e, e->nn, nn.data(), nn.size())) ||
        (rc = download(e, e->nbr, nb.data(), nb.size())) || (rc = slot_ids(e, rec)))
        return rc;
    std::vector<long long> fwd, bwd;
    for (int s = 0; s < N; s++) {
        const int n = nn[s] & (LIST_EXC - 1);
        if (n > e->kmax)
            return set_err(MDX_ERR_STATE,
                           "particle %u has %d list entries, more than kmax = %d (it walks its "
                           "cells instead)", rec[s].id, n, e->kmax);
        for (int k = 0; k < n; k++) {
            const long long a = rec[s].id, b = rec[nb[(size_t)k * N + s]].id;
            (a < b ? fwd : bwd).push_back(a < b ? (a << 32) | b : (b << 32) | a);
        }
    }
    std::sort(fwd.begin(), fwd.end());
    std::sort(bwd.begin(), bwd.end());
    if (fwd != bwd)
        return set_err(MDX_ERR_STATE, "neighbour list not symmetric (%zu vs %zu entries)",
                       fwd.size(), bwd.size());
    *npairs = (long long)fwd.size();
    for (long long k = 0; k < std::min(cap, *npairs); k++) {
        ij[2 * k] = (int)(fwd[k] >> 32);
        ij[2 * k + 1] = (int)(fwd[k] & 0xffffffffll);
    }
    if (radii) {
        // the build compares r^2 with the squares of these (make_args)
        KArgs<R> a = make_args<R>(e);
        const double rl = reach(e) + e->skin;
        const double rw = e->wca ? e->wca_pl * (1 + 1e-12) + e->skin : 0.0;
        radii[0] = rl;
        radii[1] = a.rl2w == a.rl2 ? rl : rw;
    }
    if (*npairs > cap)
        return set_err(MDX_ERR_CAPACITY, "%lld pairs, capacity %lld", *npairs, cap);
    return MDX_OK;
}

int mdx_list_pairs(mdx_engine *e, int *ij, long long cap, long long *npairs, double *radii)
{
    if (!e || !npairs || cap < 0 || (cap > 0 && !ij))
        return set_err(MDX_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? list_pairs_t<double>(e, ij, cap, npairs, radii)
                                    : list_pairs_t<float>(e, ij, cap, npairs, radii);
}

int mdx_info(mdx_engine *e, double *out)
{
    if (!e || !out)
        return set_err(MDX_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(e->device));
    int rc = build_grid(e);
    if (rc)
        return rc;
    out[0] = e->N;
    out[1] = e->ncells;
    out[2] = e->reps[0].l[0] / e->reps[0].nc[0];
    out[3] = (double)e->bytes_held;
    out[4] = (double)e->steps_taken;
    StepCtl c;
    HIPCHK(hipMemcpyAsync(&c, e->ctl.p, sizeof c, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    out[5] = (double)c.nbuilds;
    out[6] = e->skin;
    out[7] = e->kmax;
    return MDX_OK;
}

// ------------------------------------------------ single-stage entry points
// Each changes the engine's state in place (slot order kept); afterwards the
// carried forces of mdx_step are stale (mdx_prime again) and, if positions
// moved, so are the cell and neighbour lists.
static void stage_done(mdx_engine *e, bool moved)
{
    e->primed = false;
    if (moved) {
        e->binned = false;
        e->counted = false;
        e->list_valid = false;
    }
}

static int stage_upload(mdx_engine *e, Buf &b, const double *h, size_t n)
{
    int rc = dalloc(e, b, n * sizeof(double));
    if (rc)
        return rc;
    HIPCHK(hipMemcpyAsync(b.p, h, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    return MDX_OK;
}

template <class R> static int stage_halfkick_t(mdx_engine *e, const double *F, const double *T)
{
    int rc;
    const size_t n3 = 3 * (size_t)e->N;
    if ((rc = stage_upload(e, e->tmpd, F, n3)))
        return rc;
    const bool rot = e->integ == MDX_INTEG_ROTOR;
    if (rot && (rc = stage_upload(e, e->tmpx, T, n3)))
        return rc;
    hipLaunchKernelGGL(k_stage_halfkick<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, e->N,
                       (const PosRec *)e->X[e->cur].p, (R *)e->P[e->cur].p,
                       rot ? (R *)e->Lm[e->cur].p : nullptr, (const double *)e->tmpd.p,
                       rot ? (const double *)e->tmpx.p : nullptr, e->dt / 2.);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    stage_done(e, false);
    return MDX_OK;
}

int mdx_stage_halfkick(mdx_engine *e, const double *F, const double *T)
{
    if (!e || !F || (e->integ == MDX_INTEG_ROTOR && !T))
        return set_err(MDX_ERR_ARG, "engine, F (and T for rotors) required");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? stage_halfkick_t<double>(e, F, T)
                                    : stage_halfkick_t<float>(e, F, T);
}

template <class R> static int stage_advance_pos_t(mdx_engine *e)
{
    int rc = build_grid(e);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_stage_advance_pos<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, e->N,
                       (PosRec *)e->X[e->cur].p, (R *)e->P[e->cur].p, (const DevRep *)e->dreps.p,
                       e->dt / e->m, e->periodic, (StepCtl *)e->ctl.p);
    HIPCHK(hipGetLastError());
    stage_done(e, true);
    unsigned int st;
    if ((rc = read_status(e, &st)))
        return rc;
    return status_to_rc(st);
}

int mdx_stage_advance_pos(mdx_engine *e)
{
    if (!e)
        return set_err(MDX_ERR_ARG, "null engine");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? stage_advance_pos_t<double>(e) : stage_advance_pos_t<float>(e);
}

template <class R> static int stage_rotate_t(mdx_engine *e)
{
    hipLaunchKernelGGL(k_stage_rotate<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, e->N,
                       (R *)e->Lm[e->cur].p, (R *)e->Q[e->cur].p, e->dt, e->I[0], e->I[1], e->I[2]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    stage_done(e, false);
    return MDX_OK;
}

int mdx_stage_rotate(mdx_engine *e)
{
    if (!e)
        return set_err(MDX_ERR_ARG, "null engine");
    if (e->integ != MDX_INTEG_ROTOR)
        return set_err(MDX_ERR_STATE, "rotate applies to the rotor integrator");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? stage_rotate_t<double>(e) : stage_rotate_t<float>(e);
}

template <class R> static int stage_sphere_t(mdx_engine *e, double *F, double *T, const double *RT)
{
    int rc = build_grid(e);
    if (rc)
        return rc;
    const size_t n3 = 3 * (size_t)e->N;
    if ((rc = stage_upload(e, e->tmpd, F, n3)) || (rc = stage_upload(e, e->tmpx, T, n3)) ||
        (rc = stage_upload(e, e->noise, RT, 2 * n3)))
        return rc;
    hipLaunchKernelGGL(k_stage_sphere<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, e->N,
                       (const PosRec *)e->X[e->cur].p, (const R *)e->P[e->cur].p,
                       (const R *)e->Lm[e->cur].p, (const R *)e->Q[e->cur].p,
                       (const DevRep *)e->dreps.p, (double *)e->tmpd.p, (double *)e->tmpx.p,
                       (const double *)e->noise.p, -e->gamma / e->m, -e->gammar / e->I[0],
                       -e->gammar / e->I[1], -e->gammar / e->I[2]);
    HIPCHK(hipGetLastError());
    if ((rc = download(e, e->tmpd, F, n3)) || (rc = download(e, e->tmpx, T, n3)))
        return rc;
    stage_done(e, false);
    return MDX_OK;
}

int mdx_stage_sphere(mdx_engine *e, double *F, double *T, const double *RT)
{
    if (!e || !F || !T || !RT)
        return set_err(MDX_ERR_ARG, "engine, F, T and RT required");
    if (e->integ != MDX_INTEG_ROTOR)
        return set_err(MDX_ERR_STATE, "the sphere update applies to the rotor integrator");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? stage_sphere_t<double>(e, F, T, RT)
                                    : stage_sphere_t<float>(e, F, T, RT);
}

template <class R> static int stage_advance_mom_t(mdx_engine *e, const double *F, const double *Rn)
{
    int rc = build_grid(e);
    if (rc)
        return rc;
    const size_t n3 = 3 * (size_t)e->N;
    if ((rc = stage_upload(e, e->tmpd, F, n3)) || (rc = stage_upload(e, e->tmpx, Rn, n3)))
        return rc;
    const double d = (e->gamma * e->dt / 2.), q = (e->dt) / 2.;
    const double c2 = (1.0 / (1.0 + (d))), c3 = (1.0 / (1.0 + (d))) * q, c5 = (1 - (d));
    hipLaunchKernelGGL(k_stage_advance_mom<R>, dim3(nblk(e->N)), dim3(256), 0, e->stream, e->N,
                       (const PosRec *)e->X[e->cur].p, (R *)e->P[e->cur].p,
                       (const DevRep *)e->dreps.p, (const double *)e->tmpd.p,
                       (const double *)e->tmpx.p, c5 * c2, (c5 * (c3) + q), c2, c5);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    stage_done(e, false);
    return MDX_OK;
}

int mdx_stage_advance_mom(mdx_engine *e, const double *F, const double *R)
{
    if (!e || !F || !R)
        return set_err(MDX_ERR_ARG, "engine, F and R required");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? stage_advance_mom_t<double>(e, F, R)
                                    : stage_advance_mom_t<float>(e, F, R);
}

// Autotuner: time trial steps of each candidate (skin x walk form) on this
// engine's own system and keep the fastest.  The trial steps are real steps
// (the state advances); rotors are primed again after each change.
template <class R> static int autotune_t(mdx_engine *e, int trial, int cap, double *report, int *ncfg)
{
    static const double skins[] = {0.25, 0.4, 0.6, 0.9};
    static const int walks[] = {1, 0};
    hipEvent_t t0, t1;
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    int k = 0, best = -1, rc = MDX_OK;
    double best_ms = 1e300;
    const bool rotor = e->integ == MDX_INTEG_ROTOR;
    // warm-up at the current settings first: a system started from rest is
    // still heating towards kT for about 1 / gamma, and its slow particles
    // would make the first candidates' list rebuilds look rare
    if (rotor && !e->primed && (rc = do_prime<R>(e, nullptr)))
        return rc;
    if ((rc = rotor ? enqueue_rotor<R>(e, 4 * trial, nullptr) : enqueue_bead<R>(e, 4 * trial, nullptr)))
        return rc;
    for (double sk : skins)
        for (int wv : walks) {
            if (sk != e->skin)
                e->grid_dirty = true;
            e->skin = sk;
            e->use_walkc = wv;
            e->list_valid = false;
            if (rotor && (rc = do_prime<R>(e, nullptr)))
                break;
            // one call to settle the list and the graphs, then the timed call
            if ((rc = rotor ? enqueue_rotor<R>(e, trial, nullptr) : enqueue_bead<R>(e, trial, nullptr)))
                break;
            HIPCHK(hipEventRecord(t0, e->stream));
            if ((rc = rotor ? enqueue_rotor<R>(e, trial, nullptr) : enqueue_bead<R>(e, trial, nullptr)))
                break;
            HIPCHK(hipEventRecord(t1, e->stream));
            HIPCHK(hipEventSynchronize(t1));
            float ms;
            HIPCHK(hipEventElapsedTime(&ms, t0, t1));
            const double per = ms / trial;
            if (k < cap) {
                report[3 * k] = sk;
                report[3 * k + 1] = wv;
                report[3 * k + 2] = per;
            }
            if (per < best_ms) {
                best_ms = per;
                best = k;
            }
            k++;
        }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc)
        return rc;
    const double bsk = skins[best / 2];
    if (bsk != e->skin)
        e->grid_dirty = true;
    e->skin = bsk;
    e->use_walkc = walks[best % 2];
    e->list_valid = false;
    if (rotor && (rc = do_prime<R>(e, nullptr)))
        return rc;
    *ncfg = k;
    unsigned int st;
    if ((rc = read_status(e, &st)))
        return rc;
    return status_to_rc(st);
}

int mdx_autotune(mdx_engine *e, int trial_steps, int cap, double *report, int *nconfigs)
{
    if (!e || trial_steps < 1 || cap < 0 || (cap > 0 && !report) || !nconfigs)
        return set_err(MDX_ERR_ARG, "bad arguments");
    if (e->integ == MDX_INTEG_ROTOR && !e->primed)
        return set_err(MDX_ERR_STATE, "mdx_autotune before mdx_prime");
    HIPCHK(hipSetDevice(e->device));
    return e->prec == MDX_PREC_FP64 ? autotune_t<double>(e, trial_steps, cap, report, nconfigs)
                                    : autotune_t<float>(e, trial_steps, cap, report, nconfigs);
}

int mdx_set_neighbor_list(mdx_engine *e, double skin, int kmax, int resort_every)
{
    if (!e || !(skin >= 0) || kmax < MDX_LBATCH || resort_every < 1)
        return set_err(MDX_ERR_ARG, "skin >= 0, kmax >= %d and resort_every >= 1 required",
                       MDX_LBATCH);
    HIPCHK(hipSetDevice(e->device));
    if (skin != e->skin)
        e->grid_dirty = true; // cells hold the interaction reach plus the skin
    e->skin = skin;
    e->kmax = kmax;
    e->resort_every = resort_every;
    e->list_valid = false;
    return MDX_OK;
}

// per-phase device times of nsteps steps launched eagerly with HIP events on
// the engine's stream (the same launches as mdx_step); the state advances by
// nsteps steps
template <class R>
static int profile_t(mdx_engine *e, int nsteps, int cap, const char **names, double *ms, int *nk)
{
    int rc;
    if (e->integ == MDX_INTEG_ROTOR && !e->primed && (rc = do_prime<R>(e, nullptr)))
        return rc;
    std::vector<hipEvent_t> ev(NEV * (size_t)nsteps + 2);
    for (auto &x : ev)
        HIPCHK(hipEventCreate(&x));
    if (e->integ == MDX_INTEG_ROTOR)
        rc = enqueue_rotor<R>(e, nsteps, nullptr, ev.data() + 2, ev.data());
    else
        rc = enqueue_bead<R>(e, nsteps, nullptr, ev.data() + 2);
    if (rc)
        return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    // bins: cell/list work; walk; patch; finish (integrator, or patch +
    // integrator in one kernel); gather: walk..end; step: all
    static const char *kn[] = {"bins", "walk", "patch", "finish", "gather", "step",
                               "kick_drift", "bins_min", "bins_max"};
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 1e30, 0};
    const hipEvent_t *se = ev.data() + 2;
    for (int s = 0; s < nsteps; s++) {
        float t[4], tall;
        for (int k = 0; k < 4; k++)
            HIPCHK(hipEventElapsedTime(&t[k], se[NEV * s + k], se[NEV * s + k + 1]));
        HIPCHK(hipEventElapsedTime(&tall, se[NEV * s], se[NEV * s + 4]));
        acc[0] += t[0];
        acc[1] += t[1];
        acc[2] += t[2];
        acc[3] += t[3];
        acc[4] += t[1] + t[2] + t[3];
        acc[5] += tall;
        acc[7] = std::min(acc[7], (double)t[0] * nsteps);
        acc[8] = std::max(acc[8], (double)t[0] * nsteps);
    }
    if (e->integ == MDX_INTEG_ROTOR) {
        float t;
        HIPCHK(hipEventElapsedTime(&t, ev[0], ev[1]));
        acc[6] = t * nsteps; // reported once, not averaged
    }
    for (auto &x : ev)
        (void)hipEventDestroy(x);
    const bool compact = e->last_split && e->use_compact;
    int w = 0;
    for (int k = 0; k < 9 && w < cap; k++) {
        if (k == 6 && e->integ != MDX_INTEG_ROTOR)
            continue;
        if (k == 2 && !compact)
            continue; // no separate patch kernel
        if ((k == 1 || k == 3) && !e->last_split)
            continue; // one gather kernel
        names[w] = kn[k];
        ms[w++] = acc[k] / nsteps;
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
