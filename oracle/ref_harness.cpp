/*
 * ref_harness.cpp -- drives the reference's own C++ hot-path classes on
 * files, to make golden fixtures and to time the reference's CPU path.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference
 * sources where they lie (the reference headers include their .cpp files);
 * nothing from the reference is copied here, this file only calls it.  The
 * binary goes to oracle/_ref/ and is never linked into the product.
 *
 * Usage: mtref <command> <job-dir>
 *   job-dir/job.txt holds "key value" lines; arrays are raw little-endian
 *   files job-dir/<name>.bin (float64 or int32 as documented per command).
 * Commands:
 *   cells      x(N,3) -> cells(N) via cube::assign_box, boxes(tb,ncols) via
 *              cube::generate_boxes_relationships
 *   pairs      x -> pairs(P,2) via MD::calculatepairs
 *   forces     x,Q -> F,T via MD::calculateforces(pairs, HSPotential) then
 *              LangevinNVTR::calculate_forces_and_torques3D(pairs, ComboPatch)
 *   bonded     x, bonds, angles -> Fb, Fa via MD::calculateforces(bonds, pot)
 *              and MD::calculateforces_threebody(triplets, BendingPotential)
 *   traj_rotor x,p,Q,L -> final state + snapshots, Condensate::run step order
 *              with the injected noise of orc_noise_uniform
 *   traj_bead  x,p -> final state, mainActivePolymer step order, injected noise
 *   time_rotor x,Q -> wall-clock of the Condensate::run loop (OpenMP, rand())
 *   long_rotor x,Q -> snapshots of a long Condensate::run loop (rand() noise)
 *   calcg1    x    -> calcg1 pair-distance histogram (Analysis/PairCorrelation.cpp)
 *   csv       x    -> the matrix written by the reference's CSV operator
 *   temperature p, L -> LangevinNVTR::measured_temperature
 *   pairpot   r    -> WCAPotential::energy, HSPotential::force and force2mdx
 *   stress    x    -> MD::calculatestress with an HSPotential interaction
 *   bind      x, Q, isbound, boundto -> one call of the one-bond-per-patch
 *              routine (calculate_forces_and_torques3D_onlyone_nonlets) with
 *              BindingModelSingle, BindingModelBinary or BindingModelTernary
 */
#include <stdlib.h>
#include <iostream>
#include <fstream>
#include <stdio.h>
#include <stdarg.h>
#include <vector>
#include <algorithm>
#include <stdexcept>
#include <sstream>
#include <string>
#include <iomanip>
#include <map>
#include <sys/stat.h>
#include <random>
#include <parallel/algorithm>
#include <chrono>
#include <unistd.h>
#include <execinfo.h>
#include <signal.h>
#if defined(_OPENMP)
#include <omp.h>
#else
typedef int omp_int_t;
inline omp_int_t omp_get_thread_num() { return 0; }
inline omp_int_t omp_get_max_threads() { return 1; }
inline omp_int_t omp_get_num_threads() { return 1; }
#endif
using namespace std;
using namespace std::chrono;

#include "Basic/basic.h"
#include "DataStructures/vector1.h"
#include "DataStructures/matrix2.h"
#include "DataStructures/matrix2.cpp"
#include "MDBase/MD.h"
#include "MDBase/Langevin.h"
#include "MDBase/Rotational/LangevinR.h"

#include "MDBase/Analysis/PairCorrelation.cpp"

#include "oracle.h"

static string g_dir;
static map<string, string> g_job;

static void die(const string &m)
{
    fprintf(stderr, "mtref: %s\n", m.c_str());
    exit(2);
}

static void load_job()
{
    ifstream f((g_dir + "/job.txt").c_str());
    if (!f)
        die("missing job.txt");
    string k, v;
    while (f >> k >> v)
        g_job[k] = v;
}

static double jd(const char *k, double def = NAN)
{
    auto it = g_job.find(k);
    if (it == g_job.end()) {
        if (std::isnan(def))
            die(string("missing key ") + k);
        return def;
    }
    return atof(it->second.c_str());
}
static long long ji(const char *k, long long def = -1234567)
{
    auto it = g_job.find(k);
    if (it == g_job.end()) {
        if (def == -1234567)
            die(string("missing key ") + k);
        return def;
    }
    return atoll(it->second.c_str());
}

template <class T> static vector<T> rd(const string &name, size_t count)
{
    vector<T> v(count);
    FILE *f = fopen((g_dir + "/" + name + ".bin").c_str(), "rb");
    if (!f)
        die("missing " + name);
    if (count && fread(v.data(), sizeof(T), count, f) != count)
        die("short " + name);
    fclose(f);
    return v;
}

template <class T> static void wr(const string &name, const T *p, size_t count)
{
    FILE *f = fopen((g_dir + "/" + name + ".bin").c_str(), "wb");
    if (!f)
        die("cannot write " + name);
    if (count)
        fwrite(p, sizeof(T), count, f);
    fclose(f);
}

static matrix<double> to_mat(const vector<double> &v, int n, int c)
{
    matrix<double> m(n, c);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < c; j++)
            m(i, j) = v[(size_t)i * c + j];
    return m;
}

static vector<double> from_mat(matrix<double> &m)
{
    int n = m.getnrows(), c = m.getncols();
    vector<double> v((size_t)n * c);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < c; j++)
            v[(size_t)i * c + j] = m(i, j);
    return v;
}

static cube make_cube()
{
    bool per = ji("periodic") != 0;
    vector1<bool> pb(3, per);
    return cube(jd("L"), pb, 3);
}

/* ComboPatch from the job: patch_kind tetra (TetrahedralPatch), bivalent
 * (BivalentPatch) or general (GeneralPatch with ntypes, npatch_<t>,
 * numpertype_<t> cumulative, params.bin rows x 3, dirs.bin) */
static ComboPatch *make_patches()
{
    string kind = g_job.count("patch_kind") ? g_job["patch_kind"] : "none";
    if (kind == "tetra")
        return new TetrahedralPatch(jd("eps_p"), jd("d"), jd("thetam"));
    if (kind == "bivalent")
        return new BivalentPatch(jd("eps_p"), jd("d"), jd("thetam"));
    if (kind == "general") {
        int nt = (int)ji("ntypes");
        vector1<int> np(nt), npt(nt);
        int nrows = 0, ndirs = 0;
        for (int t = 0; t < nt; t++) {
            char k1[64], k2[64];
            snprintf(k1, sizeof k1, "npatch_%d", t);
            snprintf(k2, sizeof k2, "numpertype_%d", t);
            np[t] = (int)ji(k1);
            npt[t] = (int)ji(k2);
            ndirs += np[t];
        }
        for (int i = 0; i < nt; i++)
            for (int j = i; j < nt; j++)
                nrows += np[i] * np[j];
        matrix<double> params = to_mat(rd<double>("params", (size_t)nrows * 3), nrows, 3);
        matrix<double> dirs = to_mat(rd<double>("dirs", (size_t)ndirs * 3), ndirs, 3);
        return new GeneralPatch(np, npt, params, dirs);
    }
    return NULL;
}

static int cmd_cells()
{
    int N = (int)ji("N"), n = (int)ji("n");
    cube geo = make_cube();
    matrix<double> x = to_mat(rd<double>("x", (size_t)N * 3), N, 3);
    vector1<int> dim(3);
    dim[0] = 1;
    dim[1] = n;
    dim[2] = n * n;
    vector<int> cells(N);
    for (int i = 0; i < N; i++)
        cells[i] = geo.assign_box(x, i, dim, n);
    wr("cells", cells.data(), cells.size());
    int tb;
    matrix<int> boxes = geo.generate_boxes_relationships(n, tb);
    vector<int> bx((size_t)boxes.getnrows() * boxes.getncols());
    for (int i = 0; i < boxes.getnrows(); i++)
        for (int j = 0; j < boxes.getncols(); j++)
            bx[(size_t)i * boxes.getncols() + j] = boxes(i, j);
    wr("boxes", bx.data(), bx.size());
    int shape[2] = {boxes.getnrows(), boxes.getncols()};
    wr("boxes_shape", shape, 2);
    return 0;
}

static matrix<int> *make_pairs(LangevinNVT &b, cube &geo, int n, double cut)
{
    int tb;
    matrix<int> boxes = geo.generate_boxes_relationships(n, tb);
    return b.calculatepairs(boxes, cut);
}

static void write_pairs(matrix<int> *pairs)
{
    vector<int> pv((size_t)pairs->getnrows() * 2);
    for (int i = 0; i < pairs->getnrows(); i++) {
        pv[2 * (size_t)i] = (*pairs)(i, 0);
        pv[2 * (size_t)i + 1] = (*pairs)(i, 1);
    }
    wr("pairs", pv.data(), pv.size());
}

static int cmd_pairs()
{
    int N = (int)ji("N");
    cube geo = make_cube();
    LangevinNVT b(geo);
    matrix<double> x = to_mat(rd<double>("x", (size_t)N * 3), N, 3);
    b.setdat(x);
    matrix<int> *pairs = make_pairs(b, geo, (int)ji("n"), jd("cut"));
    write_pairs(pairs);
    delete pairs;
    return 0;
}

static int cmd_forces()
{
    int N = (int)ji("N");
    cube geo = make_cube();
    LangevinNVTR b(geo);
    matrix<double> x = to_mat(rd<double>("x", (size_t)N * 3), N, 3);
    matrix<double> Q = to_mat(rd<double>("Q", (size_t)N * 9), N, 9);
    matrix<double> p0(N, 3), L0(N, 3);
    b.initialize(x, p0, Q, L0);
    b.setmaxdistance(jd("maxd2", 2.4));
    matrix<int> *pairs = make_pairs(b, geo, (int)ji("n"), jd("cut"));
    write_pairs(pairs);
    HSPotential wsa(jd("wca_eps"), jd("wca_sigma"));
    matrix<double> F = b.calculateforces(*pairs, wsa);
    vector<double> Fw = from_mat(F);
    wr("F_wca", Fw.data(), Fw.size());
    matrix<double> T(N, 3);
    ComboPatch *pots = make_patches();
    double epot = 0.0;
    if (pots) {
        b.calculate_forces_and_torques3D(*pairs, *pots, F, T);
        /* patch energy with the force loop's own pair filter */
        double maxd2 = jd("maxd2", 2.4);
        for (int i = 0; i < pairs->getnrows(); i++) {
            int p1 = (*pairs)(i, 0), p2 = (*pairs)(i, 1);
            if (p2 < p1)
                std::swap(p1, p2);
            vector1<double> un(3);
            double dis;
            matrix<double> xx = b.getdat();
            geo.distance_vector(xx, p1, p2, un, dis);
            if (!(dis < maxd2))
                continue;
            dis = sqrt(dis);
            un /= dis;
            int **q = new int *;
            if (pots->safe) {
                pots->UpdateIterator(p1, p2);
                *q = *pots->p;
            } else
                pots->UpdateIteratorSafe(p1, p2, q);
            for (int tp = 1; tp < (*q)[0] + 1; tp++)
                epot += pots->potential_bundle[(*q)[tp]]->energy(un, dis, b.getorientation(), p1, p2);
            delete q;
        }
    }
    vector<double> Fv = from_mat(F), Tv = from_mat(T);
    wr("F", Fv.data(), Fv.size());
    wr("T", Tv.data(), Tv.size());
    wr("epot", &epot, 1);
    delete pairs;
    return 0;
}

static int cmd_bonded()
{
    int N = (int)ji("N");
    cube geo = make_cube();
    LangevinNVT b(geo);
    matrix<double> x = to_mat(rd<double>("x", (size_t)N * 3), N, 3);
    b.setdat(x);
    int nb = (int)ji("nbonds"), na = (int)ji("nangles");
    vector<int> bl = rd<int>("bonds", (size_t)nb * 2);
    vector<int> al = rd<int>("angles", (size_t)na * 3);
    matrix<int> bonds(nb, 2), tri(na, 3);
    for (int i = 0; i < nb; i++) {
        bonds(i, 0) = bl[2 * i];
        bonds(i, 1) = bl[2 * i + 1];
    }
    for (int i = 0; i < na; i++)
        for (int k = 0; k < 3; k++)
            tri(i, k) = al[3 * i + k];
    string kind = g_job["bond_kind"];
    matrix<double> Fb;
    if (kind == "harmonic") {
        HarmonicPotential hp(jd("bond_k"), jd("bond_r0"));
        Fb = b.calculateforces(bonds, hp);
    } else if (kind == "fene") {
        FENEPotential fp(jd("bond_k"), jd("bond_r0"));
        Fb = b.calculateforces(bonds, fp);
    } else {
        /* per-bond rest lengths: calculateforcesharmonic */
        matrix<double> r0 = to_mat(rd<double>("bond_r0s", (size_t)nb), nb, 1);
        Fb = b.calculateforcesharmonic(bonds, r0, jd("bond_k"));
    }
    BendingPotential bp(jd("angle_k"), jd("angle_theta0"));
    matrix<double> Fa = b.calculateforces_threebody(tri, bp);
    vector<double> v1 = from_mat(Fb), v2 = from_mat(Fa);
    wr("Fb", v1.data(), v1.size());
    wr("Fa", v2.data(), v2.size());
    return 0;
}

static void set_rotor_params(LangevinNVTR &b)
{
    /* Condensate::Condensate (Condensate.cpp:70-79) setter order */
    b.setdt(jd("dt"));
    b.setgamma(jd("gamma"));
    b.setgammar(jd("gammar"));
    b.setkT(jd("kT"));
    b.setmaxdistance(jd("maxd2", 2.4));
}

static int cmd_traj_rotor()
{
    int N = (int)ji("N"), nsteps = (int)ji("nsteps"), n = (int)ji("n");
    int rebuild = (int)ji("rebuild", 20), snap = (int)ji("snap_every", 0);
    uint64_t seed = (uint64_t)ji("noise_seed");
    cube geo = make_cube();
    LangevinNVTR b(geo);
    matrix<double> x = to_mat(rd<double>("x", (size_t)N * 3), N, 3);
    matrix<double> p = to_mat(rd<double>("p", (size_t)N * 3), N, 3);
    matrix<double> Q = to_mat(rd<double>("Q", (size_t)N * 9), N, 9);
    matrix<double> L = to_mat(rd<double>("L", (size_t)N * 3), N, 3);
    b.initialize(x, p, Q, L);
    set_rotor_params(b);
    ComboPatch *pots = make_patches();
    double cut = jd("cut");
    int tb;
    matrix<int> boxes = geo.generate_boxes_relationships(n, tb);
    matrix<int> *pairs = b.calculatepairs(boxes, cut);
    HSPotential wsa(jd("wca_eps"), jd("wca_sigma"));
    matrix<double> F(N, 3), T(N, 3), RT(N, 6);
    vector<double> snaps;
    /* Condensate.cpp:318-331 */
    F = b.calculateforces(*pairs, wsa);
    if (pots)
        b.calculate_forces_and_torques3D(*pairs, *pots, F, T);
    vector<double> nz((size_t)N * 6);
    orc_noise_fill(seed, 0, (long long)N * 6, nz.data());
    RT = to_mat(nz, N, 6);
    b.create_forces_and_torques_sphere(F, T, RT);
    /* Condensate.cpp:335-377 */
    for (int i = 0; i < nsteps; i++) {
        if (i > 0 && i % rebuild == 0) {
            delete pairs;
            pairs = b.calculatepairs(boxes, cut);
        }
        b.advancemom_halfstep(F, T);
        b.advance_pos();
        b.rotate();
        F = b.calculateforces(*pairs, wsa);
        T.reset_parallel(0.0);
        if (pots)
            b.calculate_forces_and_torques3D(*pairs, *pots, F, T);
        orc_noise_fill(seed, (uint64_t)(i + 1) * N * 6, (long long)N * 6, nz.data());
        RT = to_mat(nz, N, 6);
        b.create_forces_and_torques_sphere(F, T, RT);
        b.advancemom_halfstep(F, T);
        if (snap > 0 && (i + 1) % snap == 0) {
            matrix<double> xm = b.getdat();
            vector<double> xs = from_mat(xm);
            snaps.insert(snaps.end(), xs.begin(), xs.end());
        }
    }
    matrix<double> xf = b.getdat(), pf = b.getmom();
    vector<double> v1 = from_mat(xf), v2 = from_mat(pf), v3 = from_mat(b.getorientation()),
                   v4 = from_mat(b.getangmom());
    wr("x_out", v1.data(), v1.size());
    wr("p_out", v2.data(), v2.size());
    wr("Q_out", v3.data(), v3.size());
    wr("L_out", v4.data(), v4.size());
    wr("snaps", snaps.data(), snaps.size());
    delete pairs;
    return 0;
}

static int cmd_traj_bead()
{
    int N = (int)ji("N"), nsteps = (int)ji("nsteps"), n = (int)ji("n");
    int rebuild = (int)ji("rebuild", 20), snap = (int)ji("snap_every", 0);
    uint64_t seed = (uint64_t)ji("noise_seed");
    cube geo = make_cube();
    LangevinNVT b(geo);
    matrix<double> x = to_mat(rd<double>("x", (size_t)N * 3), N, 3);
    matrix<double> p = to_mat(rd<double>("p", (size_t)N * 3), N, 3);
    b.setdat(x);
    b.setmom(p);
    b.setgamma(jd("gamma"));
    b.setdt(jd("dt"));
    b.setkT(jd("kT"));
    int nb = (int)ji("nbonds");
    vector<int> bl = rd<int>("bonds", (size_t)nb * 2);
    matrix<int> bonds(nb, 2);
    for (int i = 0; i < nb; i++) {
        bonds(i, 0) = bl[2 * i];
        bonds(i, 1) = bl[2 * i + 1];
    }
    HarmonicPotential hp(jd("bond_k"), jd("bond_r0"));
    HSPotential wsa(jd("wca_eps"), jd("wca_sigma"));
    double cut = jd("cut");
    int tb;
    matrix<int> boxes = geo.generate_boxes_relationships(n, tb);
    matrix<int> *pairs = b.calculatepairs(boxes, cut);
    vector<double> nz((size_t)N * 3), snaps;
    /* mainActivePolymer.cpp:381-426 */
    for (int i = 0; i < nsteps; i++) {
        if (i > 0 && i % rebuild == 0) {
            delete pairs;
            pairs = b.calculatepairs(boxes, cut);
        }
        matrix<double> F1(b.calculateforces(bonds, hp));
        matrix<double> F2(b.calculateforces(*pairs, wsa));
        matrix<double> F = F1;
        F += F2;
        orc_noise_fill(seed, (uint64_t)i * N * 3, (long long)N * 3, nz.data());
        matrix<double> R = to_mat(nz, N, 3);
        b.advance_mom(F, R);
        b.advance_pos();
        if (snap > 0 && (i + 1) % snap == 0) {
            matrix<double> xs = b.getdat();
            vector<double> v = from_mat(xs);
            snaps.insert(snaps.end(), v.begin(), v.end());
        }
    }
    matrix<double> xf = b.getdat(), pf = b.getmom();
    vector<double> v1 = from_mat(xf), v2 = from_mat(pf);
    wr("x_out", v1.data(), v1.size());
    wr("p_out", v2.data(), v2.size());
    wr("snaps", snaps.data(), snaps.size());
    delete pairs;
    return 0;
}

static int cmd_time_rotor()
{
    /* The Condensate::run loop (Condensate.cpp:288-377) exactly as the
     * reference runs it: rand() noise, 3.5 Verlet list rebuilt every 20 steps,
     * OpenMP inside each call.  Prints one line: seconds per step. */
    int N = (int)ji("N"), nsteps = (int)ji("nsteps"), n = (int)ji("n");
    cube geo = make_cube();
    LangevinNVTR b(geo);
    matrix<double> x = to_mat(rd<double>("x", (size_t)N * 3), N, 3);
    matrix<double> Q = to_mat(rd<double>("Q", (size_t)N * 9), N, 9);
    matrix<double> p(N, 3), L(N, 3);
    b.initialize(x, p, Q, L);
    set_rotor_params(b);
    ComboPatch *pots = make_patches();
    double cut = jd("cut");
    srand(1);
    auto t0 = high_resolution_clock::now();
    int tb;
    matrix<int> boxes = geo.generate_boxes_relationships(n, tb);
    matrix<int> *pairs = b.calculatepairs(boxes, cut);
    HSPotential wsa(jd("wca_eps"), jd("wca_sigma"));
    matrix<double> F(N, 3), T(N, 3), RT(N, 6);
    F = b.calculateforces(*pairs, wsa);
    b.calculate_forces_and_torques3D(*pairs, *pots, F, T);
    generate_uniform_random_matrix(RT);
    b.create_forces_and_torques_sphere(F, T, RT);
    auto t1 = high_resolution_clock::now();
    for (int i = 0; i < nsteps; i++) {
        vector1<double> meas(6);
        b.measured_temperature(meas);
        if (i > 0 && i % 20 == 0) {
            delete pairs;
            pairs = b.calculatepairs(boxes, cut);
        }
        b.advancemom_halfstep(F, T);
        b.advance_pos();
        b.rotate();
        F = b.calculateforces(*pairs, wsa);
        T.reset_parallel(0.0);
        b.calculate_forces_and_torques3D(*pairs, *pots, F, T);
        generate_uniform_random_matrix(RT);
        b.create_forces_and_torques_sphere(F, T, RT);
        b.advancemom_halfstep(F, T);
    }
    auto t2 = high_resolution_clock::now();
    double setup = duration_cast<duration<double>>(t1 - t0).count();
    double loop = duration_cast<duration<double>>(t2 - t1).count();
    printf("{\"setup_s\": %.6f, \"loop_s\": %.6f, \"steps\": %d, \"N\": %d, \"threads\": %d, "
           "\"pairs\": %d}\n",
           setup, loop, nsteps, N, omp_get_max_threads(), pairs->getnrows());
    delete pairs;
    return 0;
}

static int cmd_long_rotor()
{
    /* The Condensate::run loop (Condensate.cpp:288-377) with the reference's
     * own rand() noise (srand(seed)) for nsteps steps; x, p, Q, L appended to
     * snap_*.bin after every `every` steps. */
    int N = (int)ji("N"), nsteps = (int)ji("nsteps"), n = (int)ji("n"), every = (int)ji("every");
    cube geo = make_cube();
    LangevinNVTR b(geo);
    matrix<double> x = to_mat(rd<double>("x", (size_t)N * 3), N, 3);
    matrix<double> Q = to_mat(rd<double>("Q", (size_t)N * 9), N, 9);
    matrix<double> p(N, 3), L(N, 3);
    b.initialize(x, p, Q, L);
    set_rotor_params(b);
    ComboPatch *pots = make_patches();
    double cut = jd("cut");
    srand((unsigned)ji("seed"));
    int tb;
    matrix<int> boxes = geo.generate_boxes_relationships(n, tb);
    matrix<int> *pairs = b.calculatepairs(boxes, cut);
    HSPotential wsa(jd("wca_eps"), jd("wca_sigma"));
    matrix<double> F(N, 3), T(N, 3), RT(N, 6);
    F = b.calculateforces(*pairs, wsa);
    b.calculate_forces_and_torques3D(*pairs, *pots, F, T);
    generate_uniform_random_matrix(RT);
    b.create_forces_and_torques_sphere(F, T, RT);
    FILE *fx = fopen((g_dir + "/snap_x.bin").c_str(), "wb");
    FILE *fp = fopen((g_dir + "/snap_p.bin").c_str(), "wb");
    FILE *fq = fopen((g_dir + "/snap_Q.bin").c_str(), "wb");
    FILE *fl = fopen((g_dir + "/snap_L.bin").c_str(), "wb");
    if (!fx || !fp || !fq || !fl)
        die("cannot open snapshot files");
    for (int i = 0; i < nsteps; i++) {
        if (i > 0 && i % 20 == 0) {
            delete pairs;
            pairs = b.calculatepairs(boxes, cut);
        }
        b.advancemom_halfstep(F, T);
        b.advance_pos();
        b.rotate();
        F = b.calculateforces(*pairs, wsa);
        T.reset_parallel(0.0);
        b.calculate_forces_and_torques3D(*pairs, *pots, F, T);
        generate_uniform_random_matrix(RT);
        b.create_forces_and_torques_sphere(F, T, RT);
        b.advancemom_halfstep(F, T);
        if ((i + 1) % every == 0) {
            matrix<double> xs = b.getdat(), ps = b.getmom();
            matrix<double> &qs = b.getorientation(), &ls = b.getangmom();
            for (int k = 0; k < N; k++) {
                for (int c = 0; c < 3; c++) {
                    double v = xs(k, c);
                    fwrite(&v, 8, 1, fx);
                    v = ps(k, c);
                    fwrite(&v, 8, 1, fp);
                    v = ls(k, c);
                    fwrite(&v, 8, 1, fl);
                }
                for (int c = 0; c < 9; c++) {
                    double v = qs(k, c);
                    fwrite(&v, 8, 1, fq);
                }
            }
        }
    }
    fclose(fx);
    fclose(fp);
    fclose(fq);
    fclose(fl);
    delete pairs;
    return 0;
}

static int cmd_calcg1()
{
    /* calcg1 (Analysis/PairCorrelation.cpp:9-155): pair-distance histogram */
    int N = (int)ji("N");
    cube geo = make_cube();
    matrix<double> x = to_mat(rd<double>("x", (size_t)N * 3), N, 3);
    vector1<int> g = calcg1(x, geo, jd("max"), jd("dx"));
    vector<long long> out(g.getsize());
    for (int i = 0; i < g.getsize(); i++)
        out[i] = g[i];
    wr("counts", out.data(), out.size());
    return 0;
}

static int cmd_dgraph()
{
    /* distance_graph (Analysis/AnalyzeGrowth.cpp:136-177): the pairs of
       calculatepairs_parallel within the binding distance, kept when both
       indices lie outside [ex_lo, ex_hi] (the excludea predicate,
       AnalyzeGrowth.cpp:123-134; ex_lo > ex_hi keeps all, index_test), then
       ConnectedComponentsParallel_Old (Bindings/GraphAlgorithms.cpp:672-810)
       from ascending labels */
    int N = (int)ji("N");
    cube geo = make_cube();
    LangevinNVT b(geo);
    matrix<double> x = to_mat(rd<double>("x", (size_t)N * 3), N, 3);
    b.setdat(x);
    int tb;
    matrix<int> boxes = geo.generate_boxes_relationships((int)ji("n"), tb);
    matrix<int> *pairs = b.calculatepairs_parallel(boxes, jd("cut"));
    const int lo = (int)ji("ex_lo"), hi = (int)ji("ex_hi");
    vector<mdpair> edgelist;
    for (int i = 0; i < pairs->getNsafe(); i++) {
        int p1 = (*pairs)(i, 0), p2 = (*pairs)(i, 1);
        bool keep = true;
        if (lo <= hi)
            keep = (p1 < lo || p1 > hi) && (p2 < lo || p2 > hi);
        if (keep)
            edgelist.push_back(mdpair(p1, p2));
    }
    string sg = "a";
    vector1<int> idx(N, sg);
    ConnectedComponentsParallel_Old(edgelist, idx);
    vector<int> out(N);
    for (int i = 0; i < N; i++)
        out[i] = idx[i];
    wr("labels", out.data(), out.size());
    delete pairs;
    return 0;
}

static int cmd_csv()
{
    /* the reference's matrix CSV output (operator<<=, matrix2.cpp:898-912) */
    int N = (int)ji("N"), M = (int)ji("M");
    matrix<double> x = to_mat(rd<double>("x", (size_t)N * M), N, M);
    ofstream f((g_dir + "/ref.csv").c_str());
    f <<= x;
    return 0;
}

static int cmd_temperature()
{
    /* LangevinNVTR::measured_temperature(vector1<double>&) (LangevinR.cpp:170-193)
       on given momenta and angular momenta */
    int N = (int)ji("N");
    cube geo = make_cube();
    LangevinNVTR b(geo);
    matrix<double> x(N, 3);
    matrix<double> p = to_mat(rd<double>("p", (size_t)N * 3), N, 3);
    matrix<double> L = to_mat(rd<double>("L", (size_t)N * 3), N, 3);
    matrix<double> Q(N, 9);
    for (int i = 0; i < N; i++)
        Q(i, 0) = Q(i, 4) = Q(i, 8) = 1.0;
    b.initialize(x, p, Q, L);
    if (g_job.count("m"))
        b.setm(jd("m"));
    vector1<double> out(6);
    b.measured_temperature(out);
    vector<double> v(6);
    for (int k = 0; k < 6; k++)
        v[k] = out[k];
    wr("temperature", v.data(), v.size());
    return 0;
}

static int cmd_pairpot()
{
    /* WCAPotential::energy and HSPotential::force / force2mdx at given r
       (potential.h:79-200) */
    int n = (int)ji("n");
    vector<double> r = rd<double>("r", (size_t)n);
    WCAPotential w(jd("eps"), jd("sigma"), jd("attr", 0.0));
    HSPotential h(jd("eps"), jd("sigma"));
    vector<double> e(n), f(n), f2(n);
    for (int i = 0; i < n; i++) {
        e[i] = w.energy(r[i]);
        f[i] = h.force(r[i]);
        f2[i] = h.force2mdx(r[i] * r[i]);
    }
    wr("energy", e.data(), e.size());
    wr("force", f.data(), f.size());
    wr("force2mdx", f2.data(), f2.size());
    return 0;
}

static int cmd_stress()
{
    /* MD::calculatestress (MD.cpp:2381-2424) with an HSPotential as the MD's
       interaction (MD::setinteractions), over the pairs of calculatepairs */
    int N = (int)ji("N");
    cube geo = make_cube();
    LangevinNVT b(geo);
    matrix<double> x = to_mat(rd<double>("x", (size_t)N * 3), N, 3);
    b.setdat(x);
    HSPotential h(jd("wca_eps"), jd("wca_sigma"));
    b.setinteractions(h);
    matrix<int> *pairs = make_pairs(b, geo, (int)ji("n"), jd("cut"));
    matrix<double> st = b.calculatestress(*pairs);
    vector<double> v = from_mat(st);
    wr("stress", v.data(), v.size());
    write_pairs(pairs);
    delete pairs;
    return 0;
}

/* bind: x, Q, isbound(nh), boundto(nh) -> F, T, isbound, boundto after one
 * LangevinNVTR::calculate_forces_and_torques3D_onlyone_nonlets with
 * BindingModelSingle(on_rate, off_rate) (F, T start at zero: the binding
 * forces alone), srand(seed) first; the patch list from
 * calculate_patch_list over MD::calculatepairs at `cut` (Condensate.cpp:452-463).
 * Run with one OpenMP thread, so that the routine's lists are in its serial
 * order. */
static int cmd_bind()
{
    int N = (int)ji("N");
    cube geo = make_cube();
    LangevinNVTR b(geo);
    matrix<double> x = to_mat(rd<double>("x", (size_t)N * 3), N, 3);
    matrix<double> Q = to_mat(rd<double>("Q", (size_t)N * 9), N, 9);
    matrix<double> p0(N, 3), L0(N, 3);
    b.initialize(x, p0, Q, L0);
    ComboPatch *pots = make_patches();
    if (!pots)
        die("bind needs patches");
    int nh = pots->get_total_patches(N);
    vector<int> ib = rd<int>("isbound", (size_t)nh), bt = rd<int>("boundto", (size_t)nh);
    BinaryBindStore bbs;
    vector1<bool> isbound(nh);
    vector1<int> boundto(nh);
    for (int i = 0; i < nh; i++) {
        isbound[i] = ib[i] != 0;
        boundto[i] = bt[i];
    }
    bbs.isbound = isbound;
    bbs.boundto = boundto;
    matrix<int> *pairs = make_pairs(b, geo, (int)ji("n"), jd("cut"));
    vector<patchint> opairs = b.calculate_patch_list(*pairs, *pots);
    vector<int> runs_diff = adjacency(opairs);
    matrix<double> F(N, 3), T(N, 3);
    /* the model: single (on_rate, off_rate), binary (div, setup_equilibrium)
     * or ternary (SortingFunctionNonUniform div1 div2 np, setup t0..t14); for
     * the table models the species of each patch and the rate tables are
     * written out in the oracle's layout (pairs 11,12,13,22,23,33; triples
     * 111,112,113,122,123,133,222,223,233,333; tr4 = the rate to none each
     * model reads) */
    string model = g_job.count("model") ? g_job["model"] : "single";
    vector<int> species(nh, 0);
    vector<double> dtab(6 * 4, 0.0), ttab(10 * 16, 0.0), tr4(10 * 4, 0.0);
    auto put_d = [&](int code, vector1<double> &r) {
        for (int k = 0; k < 4; k++)
            dtab[4 * code + k] = r[k];
    };
    auto put_t = [&](int code, vector1<double> &r, vector1<double> &r4src) {
        for (int k = 0; k < 16; k++)
            ttab[16 * code + k] = r[k];
        for (int f = 0; f < 4; f++)
            tr4[4 * code + f] = r4src[4 * f + 3];
    };
    srand((unsigned int)ji("seed"));
    if (model == "binary") {
        const int div = (int)ji("div");
        BindingModelBinary bm(div);
        bm.setup_equilibrium();
        for (int i = 0; i < nh; i++)
            species[i] = i < div ? 1 : 2;
        put_d(0, bm.doubr11);
        put_d(1, bm.doubr12);
        put_d(3, bm.doubr22);
        put_t(0, bm.tripr111, bm.tripr111);
        put_t(1, bm.tripr112, bm.tripr112);
        put_t(3, bm.tripr122, bm.tripr111);
        put_t(6, bm.tripr222, bm.tripr111);
        b.calculate_forces_and_torques3D_onlyone_nonlets(opairs, runs_diff, *pots, bbs, bm, F, T);
    } else if (model == "ternary") {
        SortingFunctionNonUniform sf;
        sf.div1 = (int)ji("div1");
        sf.div2 = (int)ji("div2");
        sf.np = (int)ji("np");
        BindingModelTernary<SortingFunctionNonUniform> bm(sf);
        double t[15];
        for (int k = 0; k < 15; k++) {
            char key[8];
            snprintf(key, sizeof key, "t%d", k);
            t[k] = jd(key);
        }
        bm.setup(t[0], t[1], t[2], t[3], t[4], t[5], t[6], t[7], t[8], t[9], t[10], t[11], t[12],
                 t[13], t[14]);
        for (int i = 0; i < nh; i++)
            species[i] = sf(i);
        put_d(0, bm.doubr11);
        put_d(1, bm.doubr12);
        put_d(2, bm.doubr13);
        put_d(3, bm.doubr22);
        put_d(4, bm.doubr23);
        put_d(5, bm.doubr33);
        vector1<double> *tt[10] = {&bm.tripr111, &bm.tripr112, &bm.tripr113, &bm.tripr122,
                                   &bm.tripr123, &bm.tripr133, &bm.tripr222, &bm.tripr223,
                                   &bm.tripr233, &bm.tripr333};
        for (int c = 0; c < 10; c++)
            put_t(c, *tt[c], *tt[c]);
        b.calculate_forces_and_torques3D_onlyone_nonlets(opairs, runs_diff, *pots, bbs, bm, F, T);
    } else {
        BindingModelSingle bm(jd("on_rate"), jd("off_rate"));
        b.calculate_forces_and_torques3D_onlyone_nonlets(opairs, runs_diff, *pots, bbs, bm, F, T);
    }
    wr("species", species.data(), species.size());
    wr("dtab", dtab.data(), dtab.size());
    wr("ttab", ttab.data(), ttab.size());
    wr("tr4", tr4.data(), tr4.size());
    for (int i = 0; i < nh; i++) {
        ib[i] = bbs.isbound[i] ? 1 : 0;
        bt[i] = bbs.boundto[i];
    }
    vector<double> Fv = from_mat(F), Tv = from_mat(T);
    wr("F", Fv.data(), Fv.size());
    wr("T", Tv.data(), Tv.size());
    wr("isbound_out", ib.data(), ib.size());
    wr("boundto_out", bt.data(), bt.size());
    int np = (int)opairs.size();
    wr("npatchpairs", &np, 1);
    delete pairs;
    return 0;
}

/* LangevinNVTR::calculate_patch_list (LangevinR.cpp:1702-1853) on the pairs
 * of calculatepairs(cut), and adjacency (LangevinR.cpp:2000-2036): the pairs,
 * the patch list (p1, p2, potn, patch1, patch2 per row) and the run starts;
 * calculate_bound_pairs (LangevinR.cpp:1854-1998) on the same pairs */
static int cmd_plist()
{
    int N = (int)ji("N");
    cube geo = make_cube();
    LangevinNVTR b(geo);
    matrix<double> x = to_mat(rd<double>("x", (size_t)N * 3), N, 3);
    matrix<double> Q = to_mat(rd<double>("Q", (size_t)N * 9), N, 9);
    matrix<double> p0(N, 3), L0(N, 3);
    b.initialize(x, p0, Q, L0);
    ComboPatch *pots = make_patches();
    if (!pots)
        die("plist needs patches");
    matrix<int> *pairs = make_pairs(b, geo, (int)ji("n"), jd("cut"));
    write_pairs(pairs);
    vector<patchint> opairs = b.calculate_patch_list(*pairs, *pots);
    vector<int> rows(5 * opairs.size());
    for (size_t k = 0; k < opairs.size(); k++) {
        int a1, a2, pn, w1, w2;
        opairs[k].get_all(a1, a2, pn, w1, w2);
        rows[5 * k] = a1;
        rows[5 * k + 1] = a2;
        rows[5 * k + 2] = pn;
        rows[5 * k + 3] = w1;
        rows[5 * k + 4] = w2;
    }
    wr("plist", rows.data(), rows.size());
    vector<int> runs = adjacency(opairs);
    wr("adjacency", runs.data(), runs.size());
    vector<patchint> bpairs = b.calculate_bound_pairs(*pairs, *pots);
    vector<int> brows(5 * bpairs.size());
    for (size_t k = 0; k < bpairs.size(); k++) {
        int a1, a2, pn, w1, w2;
        bpairs[k].get_all(a1, a2, pn, w1, w2);
        brows[5 * k] = a1;
        brows[5 * k + 1] = a2;
        brows[5 * k + 2] = pn;
        brows[5 * k + 3] = w1;
        brows[5 * k + 4] = w2;
    }
    wr("blist", brows.data(), brows.size());
    double lim[2] = {pots->max_check, pots->max_ang};
    wr("limits", lim, 2);
    delete pairs;
    return 0;
}

int main(int argc, char **argv)
{
    if (argc != 3)
        die("usage: mtref <command> <job-dir>");
    string cmd = argv[1];
    g_dir = argv[2];
    load_job();
    if (cmd == "cells")
        return cmd_cells();
    if (cmd == "pairs")
        return cmd_pairs();
    if (cmd == "forces")
        return cmd_forces();
    if (cmd == "bonded")
        return cmd_bonded();
    if (cmd == "traj_rotor")
        return cmd_traj_rotor();
    if (cmd == "traj_bead")
        return cmd_traj_bead();
    if (cmd == "time_rotor")
        return cmd_time_rotor();
    if (cmd == "long_rotor")
        return cmd_long_rotor();
    if (cmd == "calcg1")
        return cmd_calcg1();
    if (cmd == "csv")
        return cmd_csv();
    if (cmd == "dgraph")
        return cmd_dgraph();
    if (cmd == "temperature")
        return cmd_temperature();
    if (cmd == "pairpot")
        return cmd_pairpot();
    if (cmd == "stress")
        return cmd_stress();
    if (cmd == "bind")
        return cmd_bind();
    if (cmd == "plist")
        return cmd_plist();
    die("unknown command " + cmd);
    return 2;
}
