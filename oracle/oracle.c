/*
 * oracle.c -- FP64 CPU restatement of mtcode's per-timestep Langevin step.
 *
 * TEST INFRASTRUCTURE ONLY (a result check).  Nothing in the product links or
 * calls this file.  Reference citations are relative to the reference's Code/
 * directory.  See oracle.h for what pins each function.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SQR(x) ((x) * (x))
/* Basic/basic.h:54 */
#define SIGN(a, b) ((b) >= 0 ? fabs(a) : -fabs(a))

/* Basic/basic.h:32 -- the reference's pi is a long double, which makes every
 * patch angle expression that involves it long-double arithmetic. */
static const long double PI_L = 3.1415926535897932384626433832795028841971693993751L;

/* ------------------------------------------------------------------ geometry */

int orc_fasterfloor(double x) { return x > 0 ? (int)x : (int)x - 1; }

int orc_assign_box(const orc_box *b, const double *x3, const int n[3])
{
    /* geometry.h:508-521; dim = [1, n, n^2] (MD.cpp:1071-1078) */
    int cs[3];
    for (int k = 0; k < 3; k++) {
        int m = orc_fasterfloor(n[k] * (x3[k]) / b->l[k]);
        if (m < 0)
            m = 0;
        else if (m > n[k] - 1)
            m = n[k] - 1;
        cs[k] = m;
    }
    return cs[0] + n[0] * cs[1] + n[0] * n[1] * cs[2];
}

void orc_assign_all(const orc_box *b, int N, const double *x, const int n[3], int *cell)
{
    for (int i = 0; i < N; i++)
        cell[i] = orc_assign_box(b, x + 3 * i, n);
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

int orc_boxes_relationships(const int n[3], int *out)
{
    /* geometry.h:523-609: offsets (t%3-1, (t/3)%3-1, (t/9)%3-1), periodic
     * wrap per axis, then each row sorted and deduplicated; the column count
     * is that of row 0 (geometry.h:584-592). */
    int nc = n[0] * n[1] * n[2];
    int ncols = -1;
    for (int c = 0; c < nc; c++) {
        int cx = c % n[0], cy = (c / n[0]) % n[1], cz = c / (n[0] * n[1]);
        int row[27];
        for (int t = 0; t < 27; t++) {
            int o[3] = {t % 3 - 1, (t / 3) % 3 - 1, (t / 9) % 3 - 1};
            int cc[3] = {cx + o[0], cy + o[1], cz + o[2]};
            for (int k = 0; k < 3; k++) {
                if (cc[k] < 0)
                    cc[k] += n[k];
                else if (cc[k] > n[k] - 1)
                    cc[k] -= n[k];
            }
            row[t] = cc[0] + n[0] * cc[1] + n[0] * n[1] * cc[2];
        }
        qsort(row, 27, sizeof(int), cmp_int);
        int u = 0;
        for (int t = 0; t < 27; t++)
            if (t == 0 || row[t] != row[u - 1])
                row[u++] = row[t];
        if (ncols < 0)
            ncols = u;
        for (int t = 0; t < ncols; t++)
            out[(long long)c * ncols + t] = t < u ? row[t] : 0;
    }
    return ncols;
}

int orc_distance_less_than(const orc_box *b, const double *xi, const double *xj, double cut)
{
    /* geometry.h:145-164 */
    double temp = 0.0;
    for (int k = 0; k < 3; k++) {
        double dx = xi[k] - xj[k];
        if (b->periodic) {
            if (fabs(dx) > b->l[k] * 0.5)
                dx = dx - SIGN(b->l[k], dx);
        }
        dx = SQR(dx);
        if (dx > SQR(cut))
            return 0;
        temp += dx;
    }
    return temp < SQR(cut);
}

void orc_distance_vector(const orc_box *b, const double *xi, const double *xj, double un[3],
                         double *r2)
{
    /* geometry.h:277-294; l2s = SQR(l/2.) (geometry.h:56) */
    double temp = 0.0;
    for (int k = 0; k < 3; k++) {
        double as = xi[k] - xj[k];
        if (b->periodic) {
            double l2s = SQR(b->l[k] / 2.);
            if (SQR(as) > l2s)
                as = as - SIGN(b->l[k], as);
        }
        temp += SQR(as);
        un[k] = as;
    }
    *r2 = temp;
}

long long orc_pairs_reference(const orc_box *b, int N, const double *x, const int n[3],
                              double cut, int *ij, long long cap)
{
    /* MD::calculatepairs (MD.cpp:1035-1098): particles pushed into their cell
     * in index order; precalculatepairs (MD.cpp:575-713): for each cell c1
     * and each stencil column, same cell -> i<j, higher cell -> all. */
    int nc = n[0] * n[1] * n[2];
    int *cell = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
    int *cnt = (int *)calloc((size_t)nc + 1, sizeof(int));
    int *start = (int *)calloc((size_t)nc + 1, sizeof(int));
    int *lst = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
    int *rel = (int *)malloc(sizeof(int) * (size_t)nc * 27);
    orc_assign_all(b, N, x, n, cell);
    for (int i = 0; i < N; i++)
        cnt[cell[i]]++;
    for (int c = 0; c < nc; c++)
        start[c + 1] = start[c] + cnt[c];
    memset(cnt, 0, sizeof(int) * (size_t)nc);
    for (int i = 0; i < N; i++)
        lst[start[cell[i]] + cnt[cell[i]]++] = i;
    int ncols = orc_boxes_relationships(n, rel);
    long long np = 0;
    for (int c1 = 0; c1 < nc; c1++) {
        for (int c2 = 0; c2 < ncols; c2++) {
            int box1 = c1, box2 = rel[(long long)c1 * ncols + c2];
            int b1s = start[box1 + 1] - start[box1];
            int b2s = start[box2 + 1] - start[box2];
            if (box1 == box2) {
                for (int i = 0; i < b1s; i++)
                    for (int j = i + 1; j < b2s; j++) {
                        int p1 = lst[start[box1] + i], p2 = lst[start[box2] + j];
                        if (orc_distance_less_than(b, x + 3 * p1, x + 3 * p2, cut)) {
                            if (np < cap) {
                                ij[2 * np] = p1;
                                ij[2 * np + 1] = p2;
                            }
                            np++;
                        }
                    }
            } else if (box2 > box1) {
                for (int i = 0; i < b1s; i++)
                    for (int j = 0; j < b2s; j++) {
                        int p1 = lst[start[box1] + i], p2 = lst[start[box2] + j];
                        if (orc_distance_less_than(b, x + 3 * p1, x + 3 * p2, cut)) {
                            if (np < cap) {
                                ij[2 * np] = p1;
                                ij[2 * np + 1] = p2;
                            }
                            np++;
                        }
                    }
            }
        }
    }
    free(cell);
    free(cnt);
    free(start);
    free(lst);
    free(rel);
    return np;
}

/* ---------------------------------------------------------------- potentials */

double orc_hs_force(double x, double epsilon, double sigma, double rcut)
{
    /* HSPotential::force (potential.h:112-125); paramlimit = 1.12246*sigma
     * (potential.h:84) unless an explicit cut-off is given. */
    double paramlimit = rcut > 0 ? rcut : 1.12246 * sigma;
    double s3 = sigma * sigma * sigma;
    double sigma6 = SQR(s3), sigma12 = SQR(sigma6);
    if (x < paramlimit) {
        double a = x * x * x;
        a = SQR(a);
        return 4 * epsilon * (12 * sigma12 / (x * SQR(a)) - 6 * sigma6 / (x * a));
    }
    return 0;
}

double orc_wca_energy(double x, double epsilon, double sigma, double rcut)
{
    /* WCAPotential::energy, r < paramlimit branch (potential.h:179-185), with
     * attrepsilon = 0 */
    double paramlimit = rcut > 0 ? rcut : 1.12246 * sigma;
    double s3 = sigma * sigma * sigma;
    double sigma6 = SQR(s3), sigma12 = SQR(sigma6);
    if (x < paramlimit) {
        double a = x * x * x;
        a = SQR(a);
        return 4 * epsilon * (sigma12 / SQR(a) - sigma6 / a) + epsilon;
    }
    return 0;
}

void orc_pair_forces(const orc_box *b, int N, const double *x, const int *ij, long long np,
                     int kind, double p0, double p1, double p2, const double *kk,
                     const double *r0, double *F)
{
    (void)N;
    /* MD::calculateforces (MD.cpp:1684-1764); calculateforcesharmonic
     * (MD.cpp:2037-2111) for per-pair rest lengths. */
    for (long long i = 0; i < np; ++i) {
        int a = ij[2 * i], c = ij[2 * i + 1];
        double un[3], dis;
        orc_distance_vector(b, x + 3 * a, x + 3 * c, un, &dis);
        double r = sqrt(dis);
        double f1;
        if (kind == 0) {
            f1 = orc_hs_force(r, p0, p1, p2);
        } else if (kind == 1) {
            double k = kk ? kk[i] : p0;
            double rr = r0 ? r0[i] : p1;
            f1 = -k * (r - rr);
        } else {
            /* FENEPotential::force (potential.h:356) */
            f1 = -p0 * r / (1 - SQR(r) / SQR(p1));
        }
        for (int j = 0; j < 3; ++j) {
            double fac = f1 * un[j] / sqrt(dis);
            F[3 * a + j] += fac;
            F[3 * c + j] += -fac;
        }
    }
}

static double dot3(const double *u, const double *v)
{
    /* scalar() (DataStructures/vector1.cpp:500) sums in index order */
    double s = 0;
    for (int k = 0; k < 3; k++)
        s += u[k] * v[k];
    return s;
}

void orc_angle_forces(const orc_box *b, int N, const double *x, const int *ijk, long long nt,
                      int form, const double *kk, const double *th0, double *F)
{
    (void)N;
    for (long long t = 0; t < nt; ++t) {
        int p1 = ijk[3 * t], p2 = ijk[3 * t + 1], p3 = ijk[3 * t + 2];
        double ab[3], bc[3], d1, d2;
        /* MD.cpp:1947-1948: un1 = x2 - x1, un2 = x3 - x2 */
        orc_distance_vector(b, x + 3 * p2, x + 3 * p1, ab, &d1);
        orc_distance_vector(b, x + 3 * p3, x + 3 * p2, bc, &d2);
        double f1[3] = {0, 0, 0}, f2[3] = {0, 0, 0}, f3[3] = {0, 0, 0};
        double k0 = kk[t];
        if (form == 0) {
            /* BendingPotential::force (potential3.h:43-114) */
            double theta0 = th0[t];
            double a = dot3(ab, bc);
            double mabsq = dot3(ab, ab);
            double mbcsq = dot3(bc, bc);
            double mab = sqrt(mabsq), mbc = sqrt(mbcsq);
            double temp1 = a / (mab * mbc);
            if (temp1 > 1.)
                temp1 = 0.999999999;
            else if (temp1 < -1)
                temp1 = -0.9999999999;
            double theta = acos(temp1);
            double fac1 = sqrt(1. - SQR(temp1));
            double fa = -k0 * (theta - theta0);
            if (!(fabs(theta - theta0) < 1E-10)) {
                for (int j = 0; j < 3; j++)
                    f1[j] = fa * (bc[j] * mabsq - ab[j] * a) / (mabsq * fac1 * mab * mbc);
                for (int j = 0; j < 3; j++)
                    f2[j] = fa * ((ab[j] - bc[j]) * mabsq * mbcsq +
                                  a * (-(bc[j] * mabsq) + ab[j] * mbcsq)) /
                            (fac1 * mab * mbc * mabsq * mbcsq);
                for (int j = 0; j < 3; j++)
                    f3[j] = fa * (bc[j] * a - ab[j] * mbcsq) / (mbcsq * fac1 * mab * mbc);
            }
        } else {
            /* cosine bending U = k (1 - c), c = ab.bc / (|ab||bc|);
             * F1 = -k dc/dab, F3 = k dc/dbc, F2 = -(F1+F3) */
            double a = dot3(ab, bc);
            double mabsq = dot3(ab, ab), mbcsq = dot3(bc, bc);
            double inv = 1.0 / sqrt(mabsq * mbcsq);
            double c = a * inv;
            for (int j = 0; j < 3; j++) {
                double dcab = bc[j] * inv - c * ab[j] / mabsq;
                double dcbc = ab[j] * inv - c * bc[j] / mbcsq;
                f1[j] = -k0 * dcab;
                f3[j] = k0 * dcbc;
                f2[j] = -(f1[j] + f3[j]);
            }
        }
        for (int k = 0; k < 3; k++) {
            F[3 * p1 + k] += f1[k];
            F[3 * p2 + k] += f2[k];
            F[3 * p3 + k] += f3[k];
        }
    }
}

void orc_active_forces(const orc_box *b, int N, const double *x, const int *prev,
                       const int *next, double fa, double *F)
{
    for (int i = 0; i < N; i++) {
        int a = next[i] < 0 ? i : next[i];
        int c = prev[i] < 0 ? i : prev[i];
        if (a == c)
            continue;
        double t[3], r2;
        orc_distance_vector(b, x + 3 * a, x + 3 * c, t, &r2);
        double inv = 1.0 / sqrt(r2);
        for (int k = 0; k < 3; k++)
            F[3 * i + k] += fa * t[k] * inv;
    }
}

/* patch normal in the lab frame: n = Q^T n_body (potentialtheta.h:559-565) */
static void lab_patch(const double *q, const double nb[3], double n[3])
{
    n[0] = nb[0] * q[0] + nb[1] * q[3] + nb[2] * q[6];
    n[1] = nb[0] * q[1] + nb[1] * q[4] + nb[2] * q[7];
    n[2] = nb[0] * q[2] + nb[1] * q[5] + nb[2] * q[8];
}

void orc_kf2_force_torque(const double *un, double rij, const double *Q1, const double *Q2,
                          const double nb1[3], const double nb2[3], double att, double dis,
                          double thetam, double *o)
{
    /* KernFrenkelOnePatch2::force_and_torque (potentialtheta.h:517-750) */
    for (int k = 0; k < 9; k++)
        o[k] = 0.0;
    double interaction_distance = 1.4 * dis;
    if (rij > interaction_distance)
        return;
    double n1[3], n2[3];
    lab_patch(Q1, nb1, n1);
    lab_patch(Q2, nb2, n2);
    double nx1 = n1[0], ny1 = n1[1], nz1 = n1[2];
    double nx2 = n2[0], ny2 = n2[1], nz2 = n2[2];
    double argthetai = -(nx1 * un[0] + ny1 * un[1] + nz1 * un[2]);
    double argthetaj = (nx2 * un[0] + ny2 * un[1] + nz2 * un[2]);
    if (!(argthetai > cos(thetam) && argthetaj > cos(thetam)))
        return;
    if (argthetai > 1.)
        argthetai = 0.999999;
    if (argthetaj > 1.)
        argthetaj = 0.999999;
    double thetai = acos(argthetai);
    double thetaj = acos(argthetaj);
    double f = (double)(cosl(PI_L * thetai / (2. * thetam)) * cosl(PI_L * thetaj / (2. * thetam)));
    double fac = (rij / dis);
    double fac2 = SQR(fac);
    double fac4 = SQR(fac2);
    double fac8 = SQR(fac4);
    double fac10 = fac2 * fac8;
    double expf_ = exp(-0.5 * fac10);
    double pot = -att * expf_;
    double potf = -5. * att * fac8 * (fac / dis) * expf_;
    double fx = f * potf * un[0];
    double fy = f * potf * un[1];
    double fz = f * potf * un[2];
    double d1, d2;
    if (fabs(SQR(argthetai) - 1) < 1E-5)
        d1 = 0.0;
    else
        d1 = -1. / sqrt(1 - SQR(argthetai));
    if (fabs(SQR(argthetaj) - 1) < 1E-5)
        d2 = 0.0;
    else
        d2 = -1. / sqrt(1 - SQR(argthetaj));
    double temp1 = (double)(-pot * ((PI_L / (2 * thetam)) * sinl(PI_L * thetai / (2. * thetam)) *
                                    d1 * cosl((PI_L * thetaj) / (2 * thetam))));
    double temp2 = (double)(-pot * ((PI_L / (2 * thetam)) * sinl(PI_L * thetaj / (2. * thetam)) *
                                    d2 * cosl((PI_L * thetai) / (2 * thetam))));
    fx += (temp1 / rij) * (nx1 + (argthetai)*un[0]);
    fy += (temp1 / rij) * (ny1 + (argthetai)*un[1]);
    fz += (temp1 / rij) * (nz1 + (argthetai)*un[2]);
    fx += -(temp2 / rij) * (nx2 - (argthetaj)*un[0]);
    fy += -(temp2 / rij) * (ny2 - (argthetaj)*un[1]);
    fz += -(temp2 / rij) * (nz2 - (argthetaj)*un[2]);
    long double gi = -(PI_L / (2. * thetam)) * pot * sinl(PI_L * thetai / (2 * thetam)) * d1 *
                     cosl(PI_L * thetaj / (2 * thetam));
    long double gj = -(PI_L / (2. * thetam)) * pot * sinl(PI_L * thetaj / (2 * thetam)) * d2 *
                     cosl(PI_L * thetai / (2 * thetam));
    o[0] = fx;
    o[1] = fy;
    o[2] = fz;
    o[3] = (double)(gi * (un[2] * ny1 - un[1] * nz1));
    o[4] = (double)(gi * (-un[2] * nx1 + un[0] * nz1));
    o[5] = (double)(gi * (un[1] * nx1 - un[0] * ny1));
    o[6] = (double)(gj * (-un[2] * ny2 + un[1] * nz2));
    o[7] = (double)(gj * (un[2] * nx2 - un[0] * nz2));
    o[8] = (double)(gj * (-un[1] * nx2 + un[0] * ny2));
}

double orc_kf2_energy(const double *un, double rij, const double *Q1, const double *Q2,
                      const double nb1[3], const double nb2[3], double att, double dis,
                      double thetam)
{
    /* KernFrenkelOnePatch2::energy (potentialtheta.h:448-515) */
    if (rij > 1.4 * dis)
        return 0.;
    double n1[3], n2[3];
    lab_patch(Q1, nb1, n1);
    lab_patch(Q2, nb2, n2);
    double argthetai = -(n1[0] * un[0] + n1[1] * un[1] + n1[2] * un[2]);
    double argthetaj = (n2[0] * un[0] + n2[1] * un[1] + n2[2] * un[2]);
    if (argthetai > cos(thetam) && argthetaj > cos(thetam)) {
        if (argthetai > 1.)
            argthetai = 0.999999;
        if (argthetaj > 1.)
            argthetaj = 0.999999;
        double thetai = acos(argthetai);
        double thetaj = acos(argthetaj);
        double f =
            (double)(cosl(PI_L * thetai / (2. * thetam)) * cosl(PI_L * thetaj / (2. * thetam)));
        double fac = (rij / dis);
        double fac2 = SQR(fac);
        double fac4 = SQR(fac2);
        double fac8 = SQR(fac4);
        double fac10 = fac2 * fac8;
        double expf_ = exp(-0.5 * fac10);
        return -att * expf_ * f;
    }
    return 0.0;
}

/* patch layout of GeneralPatch (GeneralPatch.cpp:38-116): first patch row of
 * each type, and the first parameter row of each type pair ti<=tj. */
static void patch_layout(const orc_patchset *ps, int pstart[ORC_MAXTYPES],
                         int rstart[ORC_MAXTYPES][ORC_MAXTYPES])
{
    int s = 0;
    for (int t = 0; t < ps->ntypes; t++) {
        pstart[t] = s;
        s += ps->npatch[t];
    }
    int hj = 0;
    for (int i = 0; i < ps->ntypes; i++)
        for (int j = i; j < ps->ntypes; j++) {
            rstart[i][j] = hj;
            hj += ps->npatch[i] * ps->npatch[j];
        }
}

static void patch_pair(const orc_box *b, const double *x, const double *Q, const orc_patchset *ps,
                       int pstart[ORC_MAXTYPES], int rstart[ORC_MAXTYPES][ORC_MAXTYPES],
                       double maxd2, int p1, int p2, double *F, double *T, double *energy,
                       double *W)
{
    /* LangevinR.cpp:923-1044 for one listed pair */
    if (p2 < p1) {
        int tp1 = p1;
        p1 = p2;
        p2 = tp1;
    }
    double un[3], dis;
    orc_distance_vector(b, x + 3 * p1, x + 3 * p2, un, &dis);
    if (!(dis < maxd2))
        return;
    dis = sqrt(dis);
    for (int k = 0; k < 3; k++)
        un[k] /= dis;
    /* GeneralPatch::UpdateIteratorSafe (GeneralPatch.cpp:275-282) needs
     * type(p1) <= type(p2); with a descending type order the pair is
     * evaluated with its roles exchanged (the potential is symmetric). */
    int a = p1, c = p2;
    if (ps->type[p1] > ps->type[p2]) {
        a = p2;
        c = p1;
        for (int k = 0; k < 3; k++)
            un[k] = -un[k];
    }
    int k1 = ps->type[a], k2 = ps->type[c];
    int r0 = rstart[k1][k2];
    for (int ia = 0; ia < ps->npatch[k1]; ia++)
        for (int ib = 0; ib < ps->npatch[k2]; ib++) {
            const double *prm = ps->params + 3 * (r0 + ia * ps->npatch[k2] + ib);
            const double *nb1 = ps->dirs + 3 * (pstart[k1] + ia);
            const double *nb2 = ps->dirs + 3 * (pstart[k2] + ib);
            if (energy) {
                *energy += orc_kf2_energy(un, dis, Q + 9 * a, Q + 9 * c, nb1, nb2, prm[0], prm[1],
                                          prm[2]);
                if (!W)
                    continue;
            }
            double o[9];
            orc_kf2_force_torque(un, dis, Q + 9 * a, Q + 9 * c, nb1, nb2, prm[0], prm[1], prm[2],
                                 o);
            if (W) {
                /* force on a times x_a - x_c = un * dis */
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++)
                        W[3 * k + l] += o[k] * (un[l] * dis);
                continue;
            }
            F[3 * a + 0] += o[0];
            F[3 * a + 1] += o[1];
            F[3 * a + 2] += o[2];
            F[3 * c + 0] += -o[0];
            F[3 * c + 1] += -o[1];
            F[3 * c + 2] += -o[2];
            T[3 * a + 0] += o[3];
            T[3 * a + 1] += o[4];
            T[3 * a + 2] += o[5];
            T[3 * c + 0] += o[6];
            T[3 * c + 1] += o[7];
            T[3 * c + 2] += o[8];
        }
}

void orc_patch_forces(const orc_box *b, int N, const double *x, const double *Q, const int *ij,
                      long long np, const orc_patchset *ps, double maxd2, double *F, double *T)
{
    (void)N;
    int pstart[ORC_MAXTYPES], rstart[ORC_MAXTYPES][ORC_MAXTYPES];
    patch_layout(ps, pstart, rstart);
    for (long long i = 0; i < np; ++i)
        patch_pair(b, x, Q, ps, pstart, rstart, maxd2, ij[2 * i], ij[2 * i + 1], F, T, NULL, NULL);
}

double orc_patch_energy(const orc_box *b, int N, const double *x, const double *Q, const int *ij,
                        long long np, const orc_patchset *ps, double maxd2)
{
    (void)N;
    int pstart[ORC_MAXTYPES], rstart[ORC_MAXTYPES][ORC_MAXTYPES];
    patch_layout(ps, pstart, rstart);
    double e = 0;
    for (long long i = 0; i < np; ++i)
        patch_pair(b, x, Q, ps, pstart, rstart, maxd2, ij[2 * i], ij[2 * i + 1], NULL, NULL, &e,
                   NULL);
    return e;
}

/* --------------------------------------------------------------- integrators */

void orc_sphere(const orc_params *P, int N, const double *p, const double *L, const double *Q,
                double *F, double *T, const double *RT)
{
    /* create_forces_and_torques_sphere (LangevinR.cpp:475-578); Rt, Rr from
     * LangevinR.cpp:51-52 */
    double Rt = sqrt(2 * P->gamma * P->kT / P->dt);
    double Rr = sqrt(2 * P->gammar * P->kT / P->dt);
    for (int i = 0; i < N; i++) {
        const double *q = Q + 9 * i;
        double plx = p[3 * i], ply = p[3 * i + 1], plz = p[3 * i + 2];
        double llx = L[3 * i], lly = L[3 * i + 1], llz = L[3 * i + 2];
        double pbx = q[0] * plx + q[1] * ply + q[2] * plz;
        double pby = q[3] * plx + q[4] * ply + q[5] * plz;
        double pbz = q[6] * plx + q[7] * ply + q[8] * plz;
        double fbrfx = (-P->gamma / P->m) * pbx;
        double fbrfy = (-P->gamma / P->m) * pby;
        double fbrfz = (-P->gamma / P->m) * pbz;
        double tbrfx = (-P->gammar / P->I[0]) * llx;
        double tbrfy = (-P->gammar / P->I[1]) * lly;
        double tbrfz = (-P->gammar / P->I[2]) * llz;
        const double *R = RT + 6 * i;
        double fbrrx = Rt * R[0], fbrry = Rt * R[1], fbrrz = Rt * R[2];
        double tbrrx = Rr * R[3], tbrry = Rr * R[4], tbrrz = Rr * R[5];
        double fx = F[3 * i] + (fbrfx + fbrrx) * q[0] + (fbrfy + fbrry) * q[3] +
                    (fbrfz + fbrrz) * q[6];
        double fy = F[3 * i + 1] + (fbrfx + fbrrx) * q[1] + (fbrfy + fbrry) * q[4] +
                    (fbrfz + fbrrz) * q[7];
        double fz = F[3 * i + 2] + (fbrfx + fbrrx) * q[2] + (fbrfy + fbrry) * q[5] +
                    (fbrfz + fbrrz) * q[8];
        double t0 = T[3 * i], t1 = T[3 * i + 1], t2 = T[3 * i + 2];
        double tbx = tbrfx + tbrrx + t0 * q[0] + t1 * q[1] + t2 * q[2];
        double tby = tbrfy + tbrry + t0 * q[3] + t1 * q[4] + t2 * q[5];
        double tbz = tbrfz + tbrrz + t0 * q[6] + t1 * q[7] + t2 * q[8];
        F[3 * i] = fx;
        F[3 * i + 1] = fy;
        F[3 * i + 2] = fz;
        T[3 * i] = tbx;
        T[3 * i + 1] = tby;
        T[3 * i + 2] = tbz;
    }
}

void orc_halfkick(const orc_params *P, int N, double *p, double *L, const double *F,
                  const double *T)
{
    /* LangevinR.cpp:327-347 */
    for (int i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) {
            p[3 * i + k] += (P->dt / 2.) * F[3 * i + k];
            L[3 * i + k] += (P->dt / 2.) * T[3 * i + k];
        }
}

void orc_advance_pos(const orc_params *P, const orc_box *b, int N, double *x, double *p)
{
    /* Langevin.cpp:205-223 then geometry.h:403-445 */
    double c1 = (P->dt / P->m);
    for (int i = 0; i < N; i++)
        for (int k = 0; k < 3; k++)
            x[3 * i + k] = x[3 * i + k] + c1 * p[3 * i + k];
    for (int k = 0; k < 3; k++) {
        double l = b->l[k];
        for (int j = 0; j < N; j++) {
            double *r = x + 3 * j + k;
            if (b->periodic) {
                double temp = *r;
                if (temp < 0) {
                    while (temp < 0)
                        temp += l;
                    *r = temp;
                } else if (temp > l) {
                    while (temp > l)
                        temp -= l;
                    *r = temp;
                }
            } else {
                if (*r < 0) {
                    *r = -*r;
                    p[3 * j + k] = -p[3 * j + k];
                } else if (*r > l) {
                    *r = l - (*r - l);
                    p[3 * j + k] = -p[3 * j + k];
                }
            }
        }
    }
}

void orc_genfullmat(const orc_params *P, const double *L3, double up[9])
{
    /* LangevinR.cpp:392-461: symmetric split Rx(h/2) Ry(h/2) Rz(h) Ry(h/2)
     * Rx(h/2) written through Cayley factors */
    double ix = P->I[0], iy = P->I[1], iz = P->I[2], dt = P->dt;
    double jx = L3[0], jy = L3[1], jz = L3[2];
    double tempx = dt * jx / ix;
    double tempy = dt * jy / iy;
    double tempz = dt * jz / iz;
    double fac1 = (1 - SQR(tempx / 4.)) / (1 + SQR(tempx / 4.));
    double fac2 = tempx / (2. * (1 + SQR(tempx / 4.)));
    double fac3 = (1 - SQR(tempy / 4.)) / (1 + SQR(tempy / 4.));
    double fac4 = tempy / (2. * (1 + SQR(tempy / 4.)));
    double fac5 = (1 - SQR(tempz / 2.)) / (1 + SQR(tempz / 2.));
    double fac6 = (tempz) / (1 + SQR(tempz / 2.));
    up[0] = -SQR(fac4) + SQR(fac3) * fac5;
    up[1] = fac3 * (fac2 * fac4 * (1 + fac5) - fac1 * fac6);
    up[2] = fac3 * (fac1 * fac4 * (1 + fac5) + fac2 * fac6);
    up[3] = fac3 * (fac2 * fac4 * (1 + fac5) + fac1 * fac6);
    up[4] = SQR(fac1) * fac5 + SQR(fac2) * (-SQR(fac3) + SQR(fac4) * fac5);
    up[5] = -(fac1 * fac2 * (SQR(fac3) + fac5 - SQR(fac4) * fac5)) +
            (SQR(fac1) + SQR(fac2)) * fac4 * fac6;
    up[6] = -(fac1 * fac3 * fac4 * (1 + fac5)) + fac2 * fac3 * fac6;
    up[7] = fac1 * fac2 * (SQR(fac3) + fac5 - SQR(fac4) * fac5) +
            (SQR(fac1) + SQR(fac2)) * fac4 * fac6;
    up[8] = -(SQR(fac2) * fac5) + SQR(fac1) * (SQR(fac3) - SQR(fac4) * fac5);
}

void orc_rotate(const orc_params *P, int N, double *L, double *Q)
{
    /* LangevinR.cpp:595-685: L <- R L, Q <- R^T Q */
    for (int i = 0; i < N; i++) {
        double rr[9];
        orc_genfullmat(P, L + 3 * i, rr);
        double jx = L[3 * i], jy = L[3 * i + 1], jz = L[3 * i + 2];
        double j0 = jx * rr[0] + jy * rr[1] + jz * rr[2];
        double j1 = jx * rr[3] + jy * rr[4] + jz * rr[5];
        double j2 = jx * rr[6] + jy * rr[7] + jz * rr[8];
        double *o = Q + 9 * i;
        double q0 = o[0] * rr[0] + o[3] * rr[3] + o[6] * rr[6];
        double q1 = o[1] * rr[0] + o[4] * rr[3] + o[7] * rr[6];
        double q2 = o[2] * rr[0] + o[5] * rr[3] + o[8] * rr[6];
        double q3 = o[0] * rr[1] + o[3] * rr[4] + o[6] * rr[7];
        double q4 = o[1] * rr[1] + o[4] * rr[4] + o[7] * rr[7];
        double q5 = o[2] * rr[1] + o[5] * rr[4] + o[8] * rr[7];
        double q6 = o[0] * rr[2] + o[3] * rr[5] + o[6] * rr[8];
        double q7 = o[1] * rr[2] + o[4] * rr[5] + o[7] * rr[8];
        double q8 = o[2] * rr[2] + o[5] * rr[5] + o[8] * rr[8];
        L[3 * i] = j0;
        L[3 * i + 1] = j1;
        L[3 * i + 2] = j2;
        o[0] = q0;
        o[1] = q1;
        o[2] = q2;
        o[3] = q3;
        o[4] = q4;
        o[5] = q5;
        o[6] = q6;
        o[7] = q7;
        o[8] = q8;
    }
}

void orc_advance_mom(const orc_params *P, int N, double *p, const double *F, const double *R)
{
    /* Langevin.cpp:254-269 with the coefficients of Langevin.cpp:29-53 */
    double d = (P->gamma * P->dt / 2.);
    double q = (P->dt) / 2.;
    double r = sqrt(0.5 * P->kT * (P->gamma) * (P->m) * (P->dt));
    double c2 = (1.0 / (1.0 + (d)));
    double c3 = (1.0 / (1.0 + (d))) * q;
    double c4 = (1.0 / (1.0 + (d))) * r;
    double c5 = (1 - (d));
    for (int i = 0; i < N; i++)
        for (int k = 0; k < 3; k++)
            p[3 * i + k] =
                c5 * c2 * p[3 * i + k] + (c5 * (c3) + q) * F[3 * i + k] + (c5 * (c4) + r) * R[3 * i + k];
}

int orc_calcg1(const orc_box *b, int N, const double *x, double max, double dx,
               long long *counts, int cap)
{
    /* PairCorrelation.cpp:140-150 binning; geometry.h:107-123 distance */
    int gs = (int)floor(max / dx);
    for (int k = 0; k < gs && k < cap; k++)
        counts[k] = 0;
    for (int i = 0; i < N; i++)
        for (int j = i + 1; j < N; j++) {
            double temp = 0.0;
            for (int k = 0; k < 3; k++) {
                double d = x[3 * i + k] - x[3 * j + k];
                if (b->periodic && fabs(d) > b->l[k] * 0.5)
                    d = d - (d >= 0 ? fabs(b->l[k]) : -fabs(b->l[k]));
                temp += d * d;
            }
            double dis = sqrt(temp);
            int in = (int)floor(dis / dx);
            if (in < gs && in < cap)
                counts[in] += 1;
        }
    return gs;
}

void orc_measured_temperature(const orc_params *P, int N, const double *p, const double *L,
                              double out[6])
{
    /* LangevinR.cpp:170-193 */
    double mo[3] = {0, 0, 0}, an[3] = {0, 0, 0};
    for (int i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) {
            mo[k] += (p[3 * i + k] * p[3 * i + k]) / (2. * P->m);
            an[k] += (L[3 * i + k] * L[3 * i + k]) / (2. * P->I[0]);
        }
    for (int k = 0; k < 3; k++) {
        out[k] = mo[k] / (double)N;
        out[3 + k] = an[k] / (double)N;
    }
}

/* --------------------------------------------------------------- noise / RNG */

static uint32_t mulhilo32(uint32_t a, uint32_t b, uint32_t *hi)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    return (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    /* Salmon et al. SC'11: multipliers 0xD2511F53, 0xCD9E8D57; Weyl key
     * increments 0x9E3779B9, 0xBB67AE85; 10 rounds, key bumped between rounds */
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t hi0, hi1;
        uint32_t lo0 = mulhilo32(0xD2511F53u, c0, &hi0);
        uint32_t lo1 = mulhilo32(0xCD9E8D57u, c2, &hi1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

static uint64_t splitmix64_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

double orc_noise_uniform(uint64_t seed, uint64_t idx)
{
    uint64_t z = splitmix64_mix(seed + (idx + 1) * 0x9e3779b97f4a7c15ULL);
    double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    return 3.464101615 * u - 1.732050808;
}

void orc_noise_fill(uint64_t seed, uint64_t first, long long count, double *out)
{
    for (long long k = 0; k < count; k++)
        out[k] = orc_noise_uniform(seed, first + (uint64_t)k);
}

/* ------------------------------------------------------- step compositions */

static int cmp_ll(const void *a, const void *b)
{
    long long x = *(const long long *)a, y = *(const long long *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* the exclusion pairs as sorted keys min * 2^32 + max, for a binary search */
static long long *exclusion_keys(const orc_interactions *it)
{
    long long *k = (long long *)malloc(sizeof(long long) * (size_t)(it->nexcl > 0 ? it->nexcl : 1));
    if (!k)
        return NULL;
    for (long long q = 0; q < it->nexcl; q++) {
        long long e0 = it->excl[2 * q], e1 = it->excl[2 * q + 1];
        k[q] = e0 < e1 ? (e0 << 32) | e1 : (e1 << 32) | e0;
    }
    qsort(k, (size_t)it->nexcl, sizeof(long long), cmp_ll);
    return k;
}

static int is_excluded(const long long *keys, long long n, int a, int c)
{
    long long key = a < c ? ((long long)a << 32) | c : ((long long)c << 32) | a;
    long long lo = 0, hi = n;
    while (lo < hi) {
        long long mid = (lo + hi) / 2;
        if (keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < n && keys[lo] == key;
}

typedef struct {
    int *ij;
    long long np, cap;
    int *wca_ij; /* pairs minus exclusions */
    long long nwca;
} pairbuf;

static int build_pairs(const orc_box *b, int N, const double *x, const orc_interactions *it,
                       pairbuf *pb)
{
    long long np = orc_pairs_reference(b, N, x, it->cells_per_side, it->list_cut, pb->ij, pb->cap);
    if (np > pb->cap) {
        free(pb->ij);
        free(pb->wca_ij);
        pb->cap = np + np / 4 + 16;
        pb->ij = (int *)malloc(sizeof(int) * 2 * pb->cap);
        pb->wca_ij = (int *)malloc(sizeof(int) * 2 * pb->cap);
        if (!pb->ij || !pb->wca_ij)
            return -1;
        np = orc_pairs_reference(b, N, x, it->cells_per_side, it->list_cut, pb->ij, pb->cap);
    }
    pb->np = np;
    pb->nwca = 0;
    if (it->nexcl > 0) {
        long long *keys = exclusion_keys(it);
        if (!keys)
            return -1;
        for (long long k = 0; k < np; k++) {
            int a = pb->ij[2 * k], c = pb->ij[2 * k + 1];
            if (is_excluded(keys, it->nexcl, a, c))
                continue;
            pb->wca_ij[2 * pb->nwca] = a;
            pb->wca_ij[2 * pb->nwca + 1] = c;
            pb->nwca++;
        }
        free(keys);
    }
    return 0;
}

static void conservative(const orc_box *b, int N, const double *x, const double *Q,
                         const orc_interactions *it, const pairbuf *pb, double *F, double *T)
{
    /* Condensate.cpp:358-362: F = calculateforces(pairs, HS), T = 0, then the
     * patch loop adds into F and T.  Bonded / active terms (no reference in
     * the rotor loop) are added between the two. */
    memset(F, 0, sizeof(double) * 3 * (size_t)N);
    if (T)
        memset(T, 0, sizeof(double) * 3 * (size_t)N);
    if (it->wca_on) {
        const int *wij = it->nexcl > 0 ? pb->wca_ij : pb->ij;
        long long nw = it->nexcl > 0 ? pb->nwca : pb->np;
        orc_pair_forces(b, N, x, wij, nw, 0, it->wca_eps, it->wca_sigma, it->wca_rcut, NULL, NULL,
                        F);
    }
    if (it->nbonds > 0)
        orc_pair_forces(b, N, x, it->bonds, it->nbonds, it->bond_kind,
                        it->bond_k ? it->bond_k[0] : 0, it->bond_r0 ? it->bond_r0[0] : 1, 0,
                        it->bond_kind == 1 ? it->bond_k : NULL,
                        it->bond_kind == 1 ? it->bond_r0 : NULL, F);
    if (it->nangles > 0)
        orc_angle_forces(b, N, x, it->angles, it->nangles, it->angle_form, it->angle_k,
                         it->angle_theta0, F);
    if (it->fa != 0.0 && it->prev && it->next)
        orc_active_forces(b, N, x, it->prev, it->next, it->fa, F);
    if (it->patches && T)
        orc_patch_forces(b, N, x, Q, pb->ij, pb->np, it->patches, it->maxd2, F, T);
}

int orc_compute_forces(const orc_box *b, int N, const double *x, const double *Q,
                       const orc_interactions *it, double *F, double *T)
{
    pairbuf pb = {NULL, 0, 0, NULL, 0};
    pb.cap = (long long)N * 64 + 64;
    pb.ij = (int *)malloc(sizeof(int) * 2 * pb.cap);
    pb.wca_ij = (int *)malloc(sizeof(int) * 2 * pb.cap);
    if (!pb.ij || !pb.wca_ij || build_pairs(b, N, x, it, &pb))
        return -1;
    conservative(b, N, x, Q, it, &pb, F, T);
    free(pb.ij);
    free(pb.wca_ij);
    return 0;
}

void orc_stress(const orc_box *b, int N, const double *x, const int *ij, long long np,
                double eps, double sigma, double rcut, double out[9])
{
    (void)N;
    /* MD::calculatestress (MD.cpp:2381-2424) with HSPotential::force2mdx
     * (potential.h:126-139) as the MD's interaction: d from distance3v
     * (geometry.h:250-276, x_p1 - x_p2 minimum-imaged, not normalised),
     * stress(k, j) += force2mdx(r^2) * sqrt(r^2) * d_k * d_j */
    double paramlimit = rcut > 0 ? rcut : 1.12246 * sigma;
    double s3 = sigma * sigma * sigma;
    double sigma6 = SQR(s3), sigma12 = SQR(sigma6);
    for (int k = 0; k < 9; k++)
        out[k] = 0;
    for (long long i = 0; i < np; i++) {
        const double *xa = x + 3 * ij[2 * i], *xc = x + 3 * ij[2 * i + 1];
        double d[3] = {xa[0] - xc[0], xa[1] - xc[1], xa[2] - xc[2]};
        if (b->periodic)
            for (int k = 0; k < 3; k++)
                if (SQR(d[k]) > SQR(b->l[k] / 2.))
                    d[k] = d[k] - (d[k] >= 0 ? fabs(b->l[k]) : -fabs(b->l[k]));
        double dis = SQR(d[0]) + SQR(d[1]) + SQR(d[2]);
        double f1 = 0;
        if (dis < SQR(paramlimit)) {
            double a = dis * dis * dis;
            f1 = 4 * eps * (12 * sigma12 / (dis * SQR(a)) - 6 * sigma6 / (dis * a));
        }
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++)
                out[3 * k + j] += f1 * sqrt(dis) * d[k] * d[j];
    }
}

/* pair term force on a (f1 along d = x_a - x_c) into W and U */
static void ev_pair(double f1, double e, const double d[3], double r, double *U, double W[9])
{
    *U += e;
    for (int k = 0; k < 3; k++) {
        double fk = f1 * d[k] / r;
        for (int l = 0; l < 3; l++)
            W[3 * k + l] += fk * d[l];
    }
}

int orc_energy_virial(const orc_box *b, int N, const double *x, const double *Q,
                      const orc_interactions *it, double *U, double W[9])
{
    pairbuf pb = {NULL, 0, 0, NULL, 0};
    pb.cap = (long long)N * 64 + 64;
    pb.ij = (int *)malloc(sizeof(int) * 2 * pb.cap);
    pb.wca_ij = (int *)malloc(sizeof(int) * 2 * pb.cap);
    if (!pb.ij || !pb.wca_ij || build_pairs(b, N, x, it, &pb))
        return -1;
    *U = 0;
    for (int k = 0; k < 9; k++)
        W[k] = 0;
    if (it->wca_on) {
        /* HSPotential::force and WCAPotential::energy (potential.h:112-125,
         * 179-185) on the listed pairs minus exclusions */
        const int *wij = it->nexcl > 0 ? pb.wca_ij : pb.ij;
        long long nw = it->nexcl > 0 ? pb.nwca : pb.np;
        for (long long i = 0; i < nw; i++) {
            double d[3], dis;
            orc_distance_vector(b, x + 3 * wij[2 * i], x + 3 * wij[2 * i + 1], d, &dis);
            double r = sqrt(dis);
            ev_pair(orc_hs_force(r, it->wca_eps, it->wca_sigma, it->wca_rcut),
                    orc_wca_energy(r, it->wca_eps, it->wca_sigma, it->wca_rcut), d, r, U, W);
        }
    }
    for (long long i = 0; i < it->nbonds; i++) {
        /* HarmonicPotential / FENEPotential energy and force (potential.h:265-297, 345-376) */
        double d[3], dis;
        orc_distance_vector(b, x + 3 * it->bonds[2 * i], x + 3 * it->bonds[2 * i + 1], d, &dis);
        double r = sqrt(dis), k = it->bond_k[i], r0 = it->bond_r0[i], f1, e;
        if (it->bond_kind == 2) {
            f1 = -k * r / (1 - SQR(r) / SQR(r0));
            e = r < r0 ? -(k / 2.) * SQR(r0) * log(1.0 - SQR(r) / SQR(r0)) : 1.E10 * SQR(r);
        } else {
            f1 = -k * (r - r0);
            e = 0.5 * k * SQR(r - r0);
        }
        ev_pair(f1, e, d, r, U, W);
    }
    for (long long t = 0; t < it->nangles; t++) {
        /* BendingPotential::energy (potential3.h:32-41) or k (1 - cos theta);
         * the triplet's virial sum_i F_i x_i = -F1 ab + F3 bc */
        int p1 = it->angles[3 * t], p2 = it->angles[3 * t + 1], p3 = it->angles[3 * t + 2];
        double ab[3], bc[3], d1, d2, F[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        orc_distance_vector(b, x + 3 * p2, x + 3 * p1, ab, &d1);
        orc_distance_vector(b, x + 3 * p3, x + 3 * p2, bc, &d2);
        double xs[9] = {0, 0, 0, ab[0], ab[1], ab[2], ab[0] + bc[0], ab[1] + bc[1], ab[2] + bc[2]};
        int tri[3] = {0, 1, 2};
        double kk = it->angle_k[t], th0 = it->angle_theta0[t];
        orc_box open = {{1e300, 1e300, 1e300}, 0};
        orc_angle_forces(&open, 3, xs, tri, 1, it->angle_form, &kk, &th0, F);
        double a = dot3(ab, bc), mab = sqrt(dot3(ab, ab)), mbc = sqrt(dot3(bc, bc));
        if (it->angle_form == 0)
            *U += 0.5 * kk * SQR(acos(a / (mab * mbc)) - th0);
        else
            *U += kk * (1 - a / sqrt(dot3(ab, ab) * dot3(bc, bc)));
        for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++)
                W[3 * k + l] += -F[k] * ab[l] + F[6 + k] * bc[l];
    }
    if (it->patches) {
        int pstart[ORC_MAXTYPES], rstart[ORC_MAXTYPES][ORC_MAXTYPES];
        patch_layout(it->patches, pstart, rstart);
        for (long long i = 0; i < pb.np; ++i)
            patch_pair(b, x, Q, it->patches, pstart, rstart, it->maxd2, pb.ij[2 * i],
                       pb.ij[2 * i + 1], NULL, NULL, U, W);
    }
    free(pb.ij);
    free(pb.wca_ij);
    return 0;
}

int orc_run_rotors(const orc_params *P, const orc_box *b, int N, double *x, double *p, double *Q,
                   double *L, const orc_interactions *it, int nsteps, const double *noise,
                   int snap_every, double *snaps)
{
    /* Condensate::run (Condensate.cpp:288-377) */
    pairbuf pb = {NULL, 0, 0, NULL, 0};
    pb.cap = (long long)N * 64 + 64;
    pb.ij = (int *)malloc(sizeof(int) * 2 * pb.cap);
    pb.wca_ij = (int *)malloc(sizeof(int) * 2 * pb.cap);
    double *F = (double *)malloc(sizeof(double) * 3 * (size_t)N);
    double *T = (double *)malloc(sizeof(double) * 3 * (size_t)N);
    if (!pb.ij || !pb.wca_ij || !F || !T || build_pairs(b, N, x, it, &pb))
        return -1;
    conservative(b, N, x, Q, it, &pb, F, T);
    orc_sphere(P, N, p, L, Q, F, T, noise);
    int rebuild = it->rebuild_every > 0 ? it->rebuild_every : 1;
    for (int i = 0; i < nsteps; i++) {
        if (i > 0 && i % rebuild == 0)
            if (build_pairs(b, N, x, it, &pb))
                return -1;
        orc_halfkick(P, N, p, L, F, T);
        orc_advance_pos(P, b, N, x, p);
        orc_rotate(P, N, L, Q);
        conservative(b, N, x, Q, it, &pb, F, T);
        orc_sphere(P, N, p, L, Q, F, T, noise + (size_t)(i + 1) * N * 6);
        orc_halfkick(P, N, p, L, F, T);
        if (snap_every > 0 && (i + 1) % snap_every == 0)
            memcpy(snaps + (size_t)((i + 1) / snap_every - 1) * N * 3, x,
                   sizeof(double) * 3 * (size_t)N);
    }
    free(pb.ij);
    free(pb.wca_ij);
    free(F);
    free(T);
    return 0;
}

int orc_run_beads(const orc_params *P, const orc_box *b, int N, double *x, double *p,
                  const orc_interactions *it, int nsteps, const double *noise, int snap_every,
                  double *snaps)
{
    /* mainActivePolymer.cpp:381-426: F = F1(bonds) + F2(pairs), then
     * advance_mom(F, R), advance_pos() */
    pairbuf pb = {NULL, 0, 0, NULL, 0};
    pb.cap = (long long)N * 64 + 64;
    pb.ij = (int *)malloc(sizeof(int) * 2 * pb.cap);
    pb.wca_ij = (int *)malloc(sizeof(int) * 2 * pb.cap);
    double *F = (double *)malloc(sizeof(double) * 3 * (size_t)N);
    if (!pb.ij || !pb.wca_ij || !F || build_pairs(b, N, x, it, &pb))
        return -1;
    int rebuild = it->rebuild_every > 0 ? it->rebuild_every : 1;
    for (int i = 0; i < nsteps; i++) {
        if (i > 0 && i % rebuild == 0)
            if (build_pairs(b, N, x, it, &pb))
                return -1;
        conservative(b, N, x, NULL, it, &pb, F, NULL);
        orc_advance_mom(P, N, p, F, noise + (size_t)i * N * 3);
        orc_advance_pos(P, b, N, x, p);
        if (snap_every > 0 && (i + 1) % snap_every == 0)
            memcpy(snaps + (size_t)((i + 1) / snap_every - 1) * N * 3, x,
                   sizeof(double) * 3 * (size_t)N);
    }
    free(pb.ij);
    free(pb.wca_ij);
    free(F);
    return 0;
}

/* -------------------------------------------------------- distance_graph */

static int uf_root(int *lab, int x)
{
    while (lab[x] != x) {
        lab[x] = lab[lab[x]];
        x = lab[x];
    }
    return x;
}

long long orc_dgraph_labels(const orc_box *b, int N, const double *x, const int n[3], double cut,
                            int lo, int hi, int *labels)
{
    long long np = orc_pairs_reference(b, N, x, n, cut, NULL, 0);
    int *ij = (int *)malloc(sizeof(int) * 2 * (size_t)(np > 0 ? np : 1));
    if (!ij)
        return -1;
    orc_pairs_reference(b, N, x, n, cut, ij, np);
    for (int i = 0; i < N; i++)
        labels[i] = i;
    for (long long k = 0; k < np; k++) {
        const int p1 = ij[2 * k], p2 = ij[2 * k + 1];
        if (lo <= hi && !((p1 < lo || p1 > hi) && (p2 < lo || p2 > hi)))
            continue;
        const int r1 = uf_root(labels, p1), r2 = uf_root(labels, p2);
        if (r1 < r2)
            labels[r2] = r1;
        else if (r2 < r1)
            labels[r1] = r2;
    }
    for (int i = 0; i < N; i++)
        labels[i] = uf_root(labels, i);
    free(ij);
    return np;
}

/* ------------------------------------------------- one-bond-per-patch binding */

typedef struct {
    int a, b;
    double s;
} bind_edge;

/* 53-bit uniform in [0, 1) from two Philox words */
static double bind_u53(const uint32_t o[4])
{
    return ((double)(o[0] >> 5) * 67108864.0 + (double)(o[1] >> 6)) * (1.0 / 9007199254740992.0);
}

double orc_bind_uniform(const uint32_t key[2], uint64_t draw, int root, int k)
{
    const uint32_t ctr[4] = {(uint32_t)root, (uint32_t)draw, (uint32_t)(draw >> 32), (uint32_t)k};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    return bind_u53(o);
}

static int pair_code(int a, int b) /* sorted species a <= b, 1..3 */
{
    static const int c[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    return c[a - 1][b - 1];
}

static int triple_code(int a, int b, int c) /* sorted a <= b <= c */
{
    if (a == 1)
        return b == 1 ? (c == 1 ? 0 : c == 2 ? 1 : 2) : b == 2 ? (c == 2 ? 3 : 4) : 5;
    if (a == 2)
        return b == 2 ? (c == 2 ? 6 : 7) : 8;
    return 9;
}

/* BindingModelSingle::doublet (Bindings/BindingModelSingle.cpp:10-25), or
 * the species-table form (BindingModelTernary.cpp:523-558) */
static int bind_doublet(const orc_bind_model *m, int before, int i1, int i2, double ran)
{
    if (m->table) {
        int s1 = m->species[i1], s2 = m->species[i2];
        const double *r = m->dtab + 4 * pair_code(s1 < s2 ? s1 : s2, s1 < s2 ? s2 : s1);
        const int b = before ? 1 : 0;
        double r1 = r[b * 2 + 0];
        double r2 = r[b * 2 + 1];
        double totalr = r1 + r2;
        double rr = totalr * ran;
        return !(rr < r1);
    }
    double totalr = m->off_rate + m->on_rate; /* tr, BindingModelSingle.cpp:5-8 */
    double rr = totalr * ran;
    return !(rr < m->off_rate);
}

/* sort_triplet_and_save_permutation (basic.h:166-245) */
static void sort3_perm(int t1, int t2, int t3, int *i, int *p)
{
    if (t1 <= t2) {
        if (t2 <= t3) {
            i[0] = t1; i[1] = t2; i[2] = t3; p[0] = 1; p[1] = 2; p[2] = 3;
        } else if (t1 <= t3) {
            i[0] = t1; i[1] = t3; i[2] = t2; p[0] = 1; p[1] = 3; p[2] = 2;
        } else {
            i[0] = t3; i[1] = t1; i[2] = t2; p[0] = 3; p[1] = 1; p[2] = 2;
        }
    } else {
        if (t1 <= t3) {
            i[0] = t2; i[1] = t1; i[2] = t3; p[0] = 2; p[1] = 1; p[2] = 3;
        } else if (t2 <= t3) {
            i[0] = t2; i[1] = t3; i[2] = t1; p[0] = 2; p[1] = 3; p[2] = 1;
        } else {
            i[0] = t3; i[1] = t2; i[2] = t1; p[0] = 3; p[1] = 2; p[2] = 1;
        }
    }
}

/* save_to_type (basic.h:373-419): b = {12, 23, 13} into the sorted frame */
static void to_type(const int *b, const int *o, int *a)
{
    if (o[0] == 1) {
        if (o[1] == 2) { a[0] = b[0]; a[1] = b[1]; a[2] = b[2]; }
        else { a[0] = b[2]; a[1] = b[1]; a[2] = b[0]; }
    } else if (o[0] == 2) {
        if (o[1] == 1) { a[0] = b[0]; a[1] = b[2]; a[2] = b[1]; }
        else { a[0] = b[1]; a[2] = b[0]; a[1] = b[2]; }
    } else {
        if (o[1] == 1) { a[0] = b[1]; a[1] = b[2]; a[2] = b[0]; }
        else { a[1] = b[0]; a[0] = b[1]; a[2] = b[2]; }
    }
}

/* inverse_save_to_type (basic.h:421-468) */
static void from_type(const int *b, const int *o, int *a)
{
    if (o[0] == 1) {
        if (o[1] == 2) { a[0] = b[0]; a[1] = b[1]; a[2] = b[2]; }
        else { a[2] = b[0]; a[1] = b[1]; a[0] = b[2]; }
    } else if (o[0] == 2) {
        if (o[1] == 1) { a[0] = b[0]; a[2] = b[1]; a[1] = b[2]; }
        else { a[1] = b[0]; a[0] = b[2]; a[2] = b[1]; }
    } else {
        if (o[1] == 1) { a[2] = b[0]; a[0] = b[1]; a[1] = b[2]; }
        else { a[1] = b[0]; a[0] = b[1]; a[2] = b[2]; }
    }
}

/* the species-table triplet (BindingModelTernary.cpp:791-870; the binary
 * model's is its identity-permutation case, BindingModelBinary.cpp:236-470) */
static int bind_triplet_table(const orc_bind_model *m, int i1, int i2, int i3, int b12, int b23,
                              int b13, int c12, int c23, int c13, double ran)
{
    int sp[3], o[3];
    sort3_perm(m->species[i1], m->species[i2], m->species[i3], sp, o);
    const int bb[3] = {b12, b23, b13}, cc[3] = {c12, c23, c13};
    int tb[3], tc[3];
    to_type(bb, o, tb);
    to_type(cc, o, tc);
    int f;
    if (tb[0])
        f = 0;
    else if (tb[1])
        f = 1;
    else if (tb[2])
        f = 2;
    else
        f = 3;
    const int code = triple_code(sp[0], sp[1], sp[2]);
    const double *r = m->ttab + 16 * code;
    double r1 = tc[0] * r[f * 4 + 0];
    double r2 = tc[1] * r[f * 4 + 1];
    double r3 = tc[2] * r[f * 4 + 2];
    double r4 = m->tr4[4 * code + f];
    double totalr = r1 + r2 + r3 + r4;
    double rr = totalr * ran;
    int ta[3] = {0, 0, 0}, a[3];
    if (rr < r1)
        ta[0] = 1;
    else if (rr >= r1 && rr < r1 + r2)
        ta[1] = 1;
    else if (rr >= r1 + r2 && rr < r1 + r2 + r3)
        ta[2] = 1;
    from_type(ta, o, a);
    return a[0] ? 1 : a[1] ? 2 : a[2] ? 3 : 0;
}

/* BindingModelSingle::triplet (BindingModelSingle.cpp:40-119): which pair
 * ends bound (1: 12, 2: 23, 3: 13, 0: none) */
static int bind_triplet(const orc_bind_model *m, int b12, int b23, int b13, int c12, int c23,
                        int c13, double ran)
{
    const double on = m->on_rate, off = m->off_rate;
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
    double totalr = r1 + r2 + r3 + r4;
    double rr = totalr * ran;
    if (rr < r1)
        return 1;
    if (rr >= r1 && rr < r1 + r2)
        return 2;
    if (rr >= r1 + r2 && rr < r1 + r2 + r3)
        return 3;
    return 0;
}

typedef struct {
    int *isbound, *boundto;
    int *pairs; /* bound pairs in decision order */
    long long npairs;
} bind_out;

static void bind_set_pair(bind_out *o, int i1, int i2)
{
    o->boundto[i1] = i2;
    o->boundto[i2] = i1;
    o->isbound[i1] = 1;
    o->isbound[i2] = 1;
    o->pairs[2 * o->npairs] = i1;
    o->pairs[2 * o->npairs + 1] = i2;
    o->npairs++;
}

static void sort3(int t1, int t2, int t3, int *i1, int *i2, int *i3)
{
    int v[3] = {t1, t2, t3};
    for (int a = 0; a < 2; a++)
        for (int c = 0; c < 2 - a; c++)
            if (v[c] > v[c + 1]) {
                int t = v[c];
                v[c] = v[c + 1];
                v[c + 1] = t;
            }
    *i1 = v[0];
    *i2 = v[1];
    *i3 = v[2];
}

/* a doublet i1 < i2 (LangevinRforce.cpp:3262-3302) */
static void bind_decide2(const orc_bind_model *m, bind_out *o, int i1, int i2, double ran)
{
    const int before = o->boundto[i1] == i2 && o->boundto[i2] == i1 && o->isbound[i1] &&
                       o->isbound[i2];
    if (bind_doublet(m, before, i1, i2, ran))
        bind_set_pair(o, i1, i2);
    else {
        o->isbound[i1] = 0;
        o->isbound[i2] = 0;
    }
}

/* a triplet i1 < i2 < i3 with connectivity c (LangevinRforce.cpp:3303-3497) */
static void bind_decide3(const orc_bind_model *m, bind_out *o, int i1, int i2, int i3, int c12,
                         int c23, int c13, double ran)
{
    const int *bt = o->boundto, *ib = o->isbound;
    int b12 = bt[i1] == i2 && bt[i2] == i1 && ib[i1] && ib[i2];
    int b23 = bt[i2] == i3 && bt[i3] == i2 && ib[i2] && ib[i3];
    int b13 = bt[i1] == i3 && bt[i3] == i1 && ib[i1] && ib[i3];
    int w = m->table ? bind_triplet_table(m, i1, i2, i3, b12, b23, b13, c12, c23, c13, ran)
                     : bind_triplet(m, b12, b23, b13, c12, c23, c13, ran);
    if (w == 1) {
        bind_set_pair(o, i1, i2);
        o->isbound[i3] = 0;
    } else if (w == 2) {
        bind_set_pair(o, i2, i3);
        o->isbound[i1] = 0;
    } else if (w == 3) {
        bind_set_pair(o, i1, i3);
        o->isbound[i2] = 0;
    } else {
        o->isbound[i1] = 0;
        o->isbound[i2] = 0;
        o->isbound[i3] = 0;
    }
}

/* connectivity of a three-node component from node 1's degree and
 * neighbours (LangevinRforce.cpp:3333-3402 / 3768-3819); -1 if malformed */
static int bind_conn3(int nb1, int first1, int nb2, int i2, int i3, int *c12, int *c23, int *c13)
{
    if (nb1 == 1) {
        if (first1 == i2) {
            *c12 = 1;
            *c13 = 0;
            *c23 = 1;
        } else if (first1 == i3) {
            *c13 = 1;
            *c12 = 0;
            *c23 = 1;
        } else
            return -1;
    } else if (nb1 == 2) {
        *c12 = 1;
        *c13 = 1;
        *c23 = nb2 == 1 ? 0 : 1;
    } else
        return -1;
    return 0;
}

/* DFUtilstore / ConnectedComponents (Bindings/GraphAlgorithms.cpp:467-531) */
static void bind_dfs(int i, const int *adj, const int *lens, int depth, char *seen, int *ind,
                     int *iter)
{
    seen[i] = 1;
    ind[(*iter)++] = i;
    for (int j = 0; j < lens[i]; j++) {
        int p = adj[i * depth + j];
        if (!seen[p])
            bind_dfs(p, adj, lens, depth, seen, ind, iter);
    }
}

/* PairHistogram of local edges (GraphAlgorithms.cpp:225-240) */
static void bind_hist(const bind_edge *E, int ne, int n, int depth, int *adj, int *lens)
{
    for (int i = 0; i < n; i++)
        lens[i] = 0;
    for (int k = 0; k < ne; k++) {
        adj[E[k].a * depth + lens[E[k].a]++] = E[k].b;
        adj[E[k].b * depth + lens[E[k].b]++] = E[k].a;
    }
}

/* component boundaries dem[0..nc] and the member order ind; returns nc */
static int bind_components(const bind_edge *E, int ne, int n, int *adj, int *lens, char *seen,
                           int *ind, int *dem)
{
    bind_hist(E, ne, n, n, adj, lens);
    for (int i = 0; i < n; i++)
        seen[i] = 0;
    int iter = 0, nc = 0;
    dem[0] = 0;
    for (int i = 0; i < n; i++)
        if (!seen[i]) {
            bind_dfs(i, adj, lens, n, seen, ind, &iter);
            dem[++nc] = iter;
        }
    return nc;
}

/* std::max_element with mdpairwdCompareScore: the first highest score */
static void bind_drop_max(bind_edge *E, int *ne)
{
    int best = 0;
    for (int k = 1; k < *ne; k++)
        if (E[best].s < E[k].s)
            best = k;
    for (int k = best; k + 1 < *ne; k++)
        E[k] = E[k + 1];
    (*ne)--;
}

static int bind_local(const int *mem, int s, int g)
{
    int lo = 0, hi = s - 1;
    while (lo <= hi) {
        int mid = (lo + hi) / 2;
        if (mem[mid] == g)
            return mid;
        if (mem[mid] < g)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return -1;
}

int orc_bind_forces(const orc_box *b, int N, const double *x, const double *Q,
                    const orc_patchset *ps, const orc_bind_model *m, int *isbound, int *boundto,
                    double *F, double *T, long long *stats)
{
    int pstart[ORC_MAXTYPES], rstart[ORC_MAXTYPES][ORC_MAXTYPES];
    patch_layout(ps, pstart, rstart);
    int *poff = (int *)malloc(sizeof(int) * (size_t)(N + 1));
    poff[0] = 0;
    for (int i = 0; i < N; i++)
        poff[i + 1] = poff[i] + ps->npatch[ps->type[i]];
    const int nh = poff[N];
    int *owner = (int *)malloc(sizeof(int) * (size_t)(nh + 1));
    for (int i = 0; i < N; i++)
        for (int k = poff[i]; k < poff[i + 1]; k++)
            owner[k] = i;
    /* ComboPatch::max_check (combopatch.h:15, GeneralPatch.cpp:93-94):
     * 1.0, raised to the largest interaction_distance = 1.4 d
     * (potentialtheta.h:417) */
    double mk = 1.0;
    int nrows = 0;
    for (int i = 0; i < ps->ntypes; i++)
        for (int j = i; j < ps->ntypes; j++)
            nrows += ps->npatch[i] * ps->npatch[j];
    for (int r = 0; r < nrows; r++)
        if (1.4 * ps->params[3 * r + 1] > mk)
            mk = 1.4 * ps->params[3 * r + 1];

    /* candidate edges (LangevinRforce.cpp:2914-3075), pairs p1 < p2 */
    long long ecap = 1024, ne = 0;
    bind_edge *edges = (bind_edge *)malloc(sizeof(bind_edge) * (size_t)ecap);
    for (int p1 = 0; p1 < N; p1++)
        for (int p2 = p1 + 1; p2 < N; p2++) {
            double un[3], dis;
            orc_distance_vector(b, x + 3 * p1, x + 3 * p2, un, &dis);
            if (!(dis < mk * mk))
                continue;
            dis = sqrt(dis);
            for (int k = 0; k < 3; k++)
                un[k] /= dis;
            int a = p1, c = p2;
            if (ps->type[p1] > ps->type[p2]) { /* roles exchanged, as in patch_pair */
                a = p2;
                c = p1;
                for (int k = 0; k < 3; k++)
                    un[k] = -un[k];
            }
            const int k1 = ps->type[a], k2 = ps->type[c];
            const int r0 = rstart[k1][k2];
            for (int ia = 0; ia < ps->npatch[k1]; ia++)
                for (int ic = 0; ic < ps->npatch[k2]; ic++) {
                    const double *prm = ps->params + 3 * (r0 + ia * ps->npatch[k2] + ic);
                    double n1[3], n2[3];
                    lab_patch(Q + 9 * a, ps->dirs + 3 * (pstart[k1] + ia), n1);
                    lab_patch(Q + 9 * c, ps->dirs + 3 * (pstart[k2] + ic), n2);
                    double argthetai = -(n1[0] * un[0] + n1[1] * un[1] + n1[2] * un[2]);
                    double argthetaj = (n2[0] * un[0] + n2[1] * un[1] + n2[2] * un[2]);
                    double disp = 1.4 * prm[1], thetam = prm[2];
                    double ctm = cos(thetam);
                    const int wp1 = poff[a] + ia, wp2 = poff[c] + ic;
                    int cond1 = boundto[wp1] == wp2 && boundto[wp2] == wp1;
                    int b1 = isbound[wp1] != 0, b2 = isbound[wp2] != 0;
                    double disp2;
                    if (b1 && b2 && cond1)
                        disp2 = disp;
                    else if (b1 && b2 && !cond1)
                        disp2 = 0.5 * disp;
                    else if (!b1 != !b2)
                        disp2 = 0.7 * disp;
                    else
                        disp2 = disp;
                    if (argthetai > ctm && argthetaj > ctm && dis < disp2) {
                        double scr1 = 1 - (argthetai - cos(thetam));
                        double scr2 = 1 - (argthetaj - cos(thetam));
                        double scr3 = 2 * (dis / disp2);
                        if (ne == ecap) {
                            ecap *= 2;
                            edges = (bind_edge *)realloc(edges, sizeof(bind_edge) * (size_t)ecap);
                        }
                        edges[ne].a = wp1;
                        edges[ne].b = wp2;
                        edges[ne].s = scr1 + scr2 + scr3;
                        ne++;
                    }
                }
        }

    /* per-patch adjacency (PairHistogramExtended, GraphAlgorithms.cpp:264-351),
     * each patch's partners in ascending order */
    int *deg = (int *)calloc((size_t)nh + 1, sizeof(int));
    int *aoff = (int *)malloc(sizeof(int) * (size_t)(nh + 1));
    for (long long k = 0; k < ne; k++) {
        deg[edges[k].a]++;
        deg[edges[k].b]++;
    }
    aoff[0] = 0;
    for (int i = 0; i < nh; i++)
        aoff[i + 1] = aoff[i] + deg[i];
    int *apart = (int *)malloc(sizeof(int) * (size_t)(aoff[nh] + 1));
    double *ascr = (double *)malloc(sizeof(double) * (size_t)(aoff[nh] + 1));
    int *fill = (int *)calloc((size_t)nh + 1, sizeof(int));
    for (long long k = 0; k < ne; k++) {
        int u = edges[k].a, v = edges[k].b;
        apart[aoff[u] + fill[u]] = v;
        ascr[aoff[u] + fill[u]++] = edges[k].s;
        apart[aoff[v] + fill[v]] = u;
        ascr[aoff[v] + fill[v]++] = edges[k].s;
    }
    for (int i = 0; i < nh; i++) /* insertion sort by partner */
        for (int k = aoff[i] + 1; k < aoff[i + 1]; k++)
            for (int q = k; q > aoff[i] && apart[q - 1] > apart[q]; q--) {
                int tp = apart[q];
                apart[q] = apart[q - 1];
                apart[q - 1] = tp;
                double ts = ascr[q];
                ascr[q] = ascr[q - 1];
                ascr[q - 1] = ts;
            }

    /* bonds whose partner is no longer a candidate break
     * (LangevinRforce.cpp:3098-3125) */
    for (int i = 0; i < nh; i++)
        if (isbound[i]) {
            int outside = 1;
            for (int k = aoff[i]; k < aoff[i + 1]; k++)
                if (apart[k] == boundto[i]) {
                    outside = 0;
                    break;
                }
            if (outside)
                isbound[i] = 0;
        }

    /* components labelled by their smallest patch (ConnectedComponentsParallel_Old,
     * GraphAlgorithms.cpp:672-760, converges to the component minimum), sizes
     * and members in ascending order (SingleHistogram, GraphAlgorithms.cpp:2-49) */
    int *lab = (int *)malloc(sizeof(int) * (size_t)(nh + 1));
    for (int i = 0; i < nh; i++)
        lab[i] = i;
    for (long long k = 0; k < ne; k++) {
        int r1 = uf_root(lab, edges[k].a), r2 = uf_root(lab, edges[k].b);
        if (r1 < r2)
            lab[r2] = r1;
        else if (r2 < r1)
            lab[r1] = r2;
    }
    int *ccs = (int *)calloc((size_t)nh + 1, sizeof(int));
    for (int i = 0; i < nh; i++) {
        lab[i] = uf_root(lab, i);
        ccs[lab[i]]++;
    }
    int *moff = (int *)malloc(sizeof(int) * (size_t)(nh + 1));
    moff[0] = 0;
    for (int i = 0; i < nh; i++)
        moff[i + 1] = moff[i] + ccs[i];
    int *mem = (int *)malloc(sizeof(int) * (size_t)(nh + 1));
    for (int i = 0; i < nh; i++)
        fill[i] = 0;
    for (int i = 0; i < nh; i++)
        mem[moff[lab[i]] + fill[lab[i]]++] = i;

    /* random numbers: the reference's rand() draws (Generate_Random_Numbers_Needed,
     * GraphAlgorithms.cpp:431-438, then generate_random_matrix,
     * matrix2.cpp:1788-1805, for the large clusters) or Philox per cluster */
    int smax = 0, nlarge = 0, nmulti = 0;
    for (int i = 0; i < nh; i++) {
        if (ccs[i] > smax)
            smax = ccs[i];
        if (ccs[i] > 1)
            nmulti++;
        if (ccs[i] > 3)
            nlarge++;
    }
    const int ncols = smax / 2 + 1;
    double *rsmall = NULL, *rlarge = NULL;
    if (m->rng == 0) {
        srand(m->seed);
        rsmall = (double *)calloc((size_t)nh + 1, sizeof(double));
        for (int i = 0; i < nh; i++)
            if (ccs[i] > 1)
                rsmall[i] = (double)(rand()) / ((double)(RAND_MAX));
        rlarge = (double *)malloc(sizeof(double) * ((size_t)nlarge * ncols + 1));
        for (long long k = 0; k < (long long)nlarge * ncols; k++)
            rlarge[k] = (((double)rand() / (RAND_MAX)));
    }

    bind_out o = {isbound, boundto, (int *)malloc(sizeof(int) * (size_t)(nh + 2)), 0};
    /* clusters of 1, 2 and 3 patches (LangevinRforce.cpp:3245-3512) */
    for (int r = 0; r < nh; r++) {
        const int s = ccs[r];
        if (s == 0 || s > 3) {
            continue;
        }
        const int *mm = mem + moff[r];
        if (s == 1) {
            isbound[mm[0]] = 0;
            continue;
        }
        double ran = m->rng == 0 ? rsmall[r] : orc_bind_uniform(m->key, m->draw, r, 0);
        if (s == 2) {
            bind_decide2(m, &o, mm[0], mm[1], ran);
        } else {
            int i1 = mm[0], i2 = mm[1], i3 = mm[2], c12 = 0, c23 = 0, c13 = 0;
            if (bind_conn3(deg[i1], deg[i1] > 0 ? apart[aoff[i1]] : -1, deg[i2], i2, i3, &c12,
                           &c23, &c13) < 0)
                return -2;
            bind_decide3(m, &o, i1, i2, i3, c12, c23, c13, ran);
        }
    }
    /* larger clusters (LangevinRforce.cpp:3531-3870): drop the highest-scored
     * edge until no component exceeds 3 patches, then decide each component */
    int allc = 0;
    for (int r = 0; r < nh; r++) {
        const int s = ccs[r];
        if (s <= 3)
            continue;
        const int *mm = mem + moff[r];
        int emax = 0;
        for (int j = 0; j < s; j++)
            emax += deg[mm[j]];
        bind_edge *E = (bind_edge *)malloc(sizeof(bind_edge) * (size_t)(emax + 1));
        int nE = 0;
        for (int j = 0; j < s; j++) /* RecapitulateEdges (LangevinRforceUtil.cpp:508-575) */
            for (int k = aoff[mm[j]]; k < aoff[mm[j] + 1]; k++) {
                int f1 = j, f2 = bind_local(mm, s, apart[k]);
                bind_edge e1;
                e1.a = f1 < f2 ? f1 : f2;
                e1.b = f1 < f2 ? f2 : f1;
                e1.s = ascr[k];
                int already = 0;
                for (int q = 0; q < nE; q++)
                    if (E[q].a == e1.a && E[q].b == e1.b) {
                        already = 1;
                        break;
                    }
                if (!already)
                    E[nE++] = e1;
            }
        int *adj = (int *)malloc(sizeof(int) * (size_t)s * s), *lens = (int *)malloc(sizeof(int) * s);
        int *ind = (int *)malloc(sizeof(int) * s), *dem = (int *)malloc(sizeof(int) * (s + 1));
        char *seen = (char *)malloc((size_t)s);
        bind_drop_max(E, &nE);
        int nc = bind_components(E, nE, s, adj, lens, seen, ind, dem);
        for (;;) {
            int big = 0;
            for (int j = 0; j < nc; j++)
                if (dem[j + 1] - dem[j] > big)
                    big = dem[j + 1] - dem[j];
            if (big <= 3)
                break;
            bind_drop_max(E, &nE);
            nc = bind_components(E, nE, s, adj, lens, seen, ind, dem);
        }
        bind_hist(E, nE, s, s, adj, lens); /* PairHistogram -> boindicesx, lensx */
        for (int k1 = 0; k1 < s; k1++) {
            const int g = mm[k1];
            if (isbound[g]) {
                int outside = 1;
                for (int k = 0; k < lens[k1]; k++)
                    if (mm[adj[k1 * s + k]] == boundto[g]) {
                        outside = 0;
                        break;
                    }
                if (outside)
                    isbound[g] = 0;
            }
        }
        int it = 0;
        for (int j = 0; j < nc; j++) {
            const int z = dem[j + 1] - dem[j];
            if (z == 1) {
                isbound[mm[ind[dem[j]]]] = 0;
            } else if (z == 2) {
                int n1 = ind[dem[j]], n2 = ind[dem[j] + 1];
                int i1 = mm[n1 < n2 ? n1 : n2], i2 = mm[n1 < n2 ? n2 : n1];
                double ran = m->rng == 0 ? rlarge[(size_t)allc * ncols + it]
                                         : orc_bind_uniform(m->key, m->draw, r, 1 + it);
                it++;
                bind_decide2(m, &o, i1, i2, ran);
            } else {
                int ii1, ii2, ii3, c12 = 0, c23 = 0, c13 = 0;
                sort3(ind[dem[j]], ind[dem[j] + 1], ind[dem[j] + 2], &ii1, &ii2, &ii3);
                if (bind_conn3(lens[ii1], lens[ii1] > 0 ? adj[ii1 * s] : -1, lens[ii2], ii2, ii3,
                               &c12, &c23, &c13) < 0)
                    return -3;
                double ran = m->rng == 0 ? rlarge[(size_t)allc * ncols + it]
                                         : orc_bind_uniform(m->key, m->draw, r, 1 + it);
                it++;
                bind_decide3(m, &o, mm[ii1], mm[ii2], mm[ii3], c12, c23, c13, ran);
            }
        }
        free(E);
        free(adj);
        free(lens);
        free(ind);
        free(dem);
        free(seen);
        allc++;
    }

    /* forces and torques of the bound pairs (LangevinRforce.cpp:4030-4095) */
    for (long long q = 0; q < o.npairs; q++) {
        const int w1 = o.pairs[2 * q], w2 = o.pairs[2 * q + 1];
        int p1 = owner[w1], p2 = owner[w2];
        if (p2 < p1) {
            int t = p1;
            p1 = p2;
            p2 = t;
        }
        double un[3], dis;
        orc_distance_vector(b, x + 3 * p1, x + 3 * p2, un, &dis);
        dis = sqrt(dis);
        for (int k = 0; k < 3; k++)
            un[k] /= dis;
        int a = p1, c = p2;
        if (ps->type[p1] > ps->type[p2]) {
            a = p2;
            c = p1;
            for (int k = 0; k < 3; k++)
                un[k] = -un[k];
        }
        const int wa = owner[w1] == a ? w1 : w2, wc = owner[w1] == a ? w2 : w1;
        const int k1 = ps->type[a], k2 = ps->type[c];
        const int ia = wa - poff[a], ic = wc - poff[c];
        const double *prm = ps->params + 3 * (rstart[k1][k2] + ia * ps->npatch[k2] + ic);
        double f[9];
        orc_kf2_force_torque(un, dis, Q + 9 * a, Q + 9 * c, ps->dirs + 3 * (pstart[k1] + ia),
                             ps->dirs + 3 * (pstart[k2] + ic), prm[0], prm[1], prm[2], f);
        F[3 * a + 0] += f[0];
        F[3 * a + 1] += f[1];
        F[3 * a + 2] += f[2];
        F[3 * c + 0] += -f[0];
        F[3 * c + 1] += -f[1];
        F[3 * c + 2] += -f[2];
        T[3 * a + 0] += f[3];
        T[3 * a + 1] += f[4];
        T[3 * a + 2] += f[5];
        T[3 * c + 0] += f[6];
        T[3 * c + 1] += f[7];
        T[3 * c + 2] += f[8];
    }
    if (stats) {
        stats[0] = ne;
        stats[1] = nmulti;
        stats[2] = nlarge;
        stats[3] = smax;
        stats[4] = o.npairs;
    }
    free(poff);
    free(owner);
    free(edges);
    free(deg);
    free(aoff);
    free(apart);
    free(ascr);
    free(fill);
    free(lab);
    free(ccs);
    free(moff);
    free(mem);
    free(rsmall);
    free(rlarge);
    free(o.pairs);
    return 0;
}

/* the loop of calculate_patch_list (bound 0) and calculate_bound_pairs
 * (bound 1, LangevinR.cpp:1854-1998: both cone arguments above the row's
 * cos(theta_m)) */
static long long patch_rows(const orc_box *b, int N, const double *x, const double *Q,
                            const int *ij, long long np, const orc_patchset *ps, double maxdis,
                            double maxang, int bound, int *out, long long cap)
{
    int pstart[ORC_MAXTYPES], rstart[ORC_MAXTYPES][ORC_MAXTYPES];
    patch_layout(ps, pstart, rstart);
    int *poff = (int *)malloc(sizeof(int) * (size_t)(N + 1));
    poff[0] = 0;
    for (int i = 0; i < N; i++)
        poff[i + 1] = poff[i] + ps->npatch[ps->type[i]];
    long long n = 0;
    for (long long k = 0; k < np; k++) {
        int p1 = ij[2 * k], p2 = ij[2 * k + 1];
        if (p2 < p1) {
            int t = p1;
            p1 = p2;
            p2 = t;
        }
        double un[3], dis;
        orc_distance_vector(b, x + 3 * p1, x + 3 * p2, un, &dis);
        dis = sqrt(dis);
        if (!(dis < maxdis))
            continue;
        for (int c = 0; c < 3; c++)
            un[c] /= dis;
        const int k1 = ps->type[p1], k2 = ps->type[p2];
        if (k1 > k2) {
            free(poff);
            return -1;
        }
        for (int ia = 0; ia < ps->npatch[k1]; ia++)
            for (int ib = 0; ib < ps->npatch[k2]; ib++) {
                double n1[3], n2[3];
                lab_patch(Q + 9 * p1, ps->dirs + 3 * (pstart[k1] + ia), n1);
                lab_patch(Q + 9 * p2, ps->dirs + 3 * (pstart[k2] + ib), n2);
                double ai = -(n1[0] * un[0] + n1[1] * un[1] + n1[2] * un[2]);
                double aj = (n2[0] * un[0] + n2[1] * un[1] + n2[2] * un[2]);
                const int row = rstart[k1][k2] + ia * ps->npatch[k2] + ib;
                int pass;
                if (bound)
                    pass = ai > cos(ps->params[3 * row + 2]) && aj > cos(ps->params[3 * row + 2]);
                else
                    pass = (ai - maxang) + (aj - maxang) > -0.4;
                if (pass) {
                    if (n < cap) {
                        int *r = out + 5 * n;
                        r[0] = p1;
                        r[1] = p2;
                        r[2] = row;
                        r[3] = poff[p1] + ia;
                        r[4] = poff[p2] + ib;
                    }
                    n++;
                }
            }
    }
    free(poff);
    return n;
}

long long orc_patch_list(const orc_box *b, int N, const double *x, const double *Q,
                         const int *ij, long long np, const orc_patchset *ps, double maxdis,
                         double maxang, int *out, long long cap)
{
    return patch_rows(b, N, x, Q, ij, np, ps, maxdis, maxang, 0, out, cap);
}

long long orc_bound_pairs(const orc_box *b, int N, const double *x, const double *Q,
                          const int *ij, long long np, const orc_patchset *ps, double maxdis,
                          int *out, long long cap)
{
    return patch_rows(b, N, x, Q, ij, np, ps, maxdis, 0.0, 1, out, cap);
}
