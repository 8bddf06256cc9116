// dropin_driver.cpp -- drives the C++ host classes (mtcode_amd/host/mtcode_amd.hpp)
// the way the reference's drivers drive MDBase, for tests/test_dropin.py:
//
//   rotor: the loop of Condensate::run (Condensate.cpp:318-377) with WCA and
//          TetrahedralPatch, either call by call (mode "stages": calculatepairs,
//          calculateforces, calculate_forces_and_torques3D,
//          create_forces_and_torques_sphere, advancemom_halfstep, advance_pos,
//          rotate) or as device-resident steps (mode "run")
//   bead:  the bead loop (forces, advance_mom, advance_pos;
//          mainActivePolymer.cpp:390-425) with WCA, harmonic bonds and cosine
//          bending, call by call or through run()
//
// usage: dropin_driver <dir>
//   dir/job.txt  key value lines: integ mode N L nsteps dt gamma gammar kT
//                precision eps_p d thetam nbonds nangles kbond r0 kappa
//   dir/x.bin (N x 3), Q.bin (N x 9, rotor), noise.bin, bonds.bin, angles.bin
//   writes dir/out_x.bin, out_p.bin (+ out_Q.bin, out_L.bin for rotors)
//   integ "stencil" with key n: writes the stencil table out_stencil.bin (no GPU)
//   integ "csv" with N, M: outfunc / importcsv of x.bin (no GPU)
//   rotor mode "output": Condensate::run(nsteps, every) with its CSV files
//   rotor mode "bindout": Condensate::run_singlebond(nsteps, every) with its
//          CSV files (pos, orientation, bindings, connectedcomponents)
//   rotor mode "bind" / "bindrun": Condensate::run_singlebond's loop
//          (BinaryBindStore, BindingModelSingle(on_rate, off_rate),
//          calculate_forces_and_torques3D_onlyone_nonlets) call by call or
//          on the device; also writes out_isbound.bin, out_boundto.bin
//   integ "dgraph" with N, L, n, cut, ex_lo, ex_hi: distance_graph labels of
//          x.bin to out_labels.bin
#include "mtcode_amd.hpp"

#include <cstdio>
#include <fstream>
#include <unistd.h>
#include <map>
#include <string>
#include <vector>

using namespace mtamd;

static std::string g_dir;
static std::map<std::string, std::string> g_job;

template <class T> static std::vector<T> rd(const std::string &name, size_t n)
{
    std::vector<T> v(n);
    std::ifstream f(g_dir + "/" + name, std::ios::binary);
    if (!f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(T))))
        throw mdx_error(MDX_ERR_ARG, "cannot read " + name);
    return v;
}

template <class T> static void wr(const std::string &name, const T *p, size_t n)
{
    std::ofstream f(g_dir + "/" + name, std::ios::binary);
    f.write(reinterpret_cast<const char *>(p), (std::streamsize)(n * sizeof(T)));
}

static double jd(const std::string &k) { return std::stod(g_job.at(k)); }
static int ji(const std::string &k) { return std::stoi(g_job.at(k)); }

template <class T> static matrix<T> to_mat(const std::vector<T> &v, int n, int m)
{
    matrix<T> a(n, m);
    std::copy(v.begin(), v.end(), a.getdata());
    return a;
}

static matrix<double> noise_row(const std::vector<double> &nz, int row, int N, int w)
{
    matrix<double> R(N, w);
    std::copy(nz.begin() + (size_t)row * N * w, nz.begin() + (size_t)(row + 1) * N * w, R.getdata());
    return R;
}

static int rotor_main()
{
    const int N = ji("N"), nsteps = ji("nsteps");
    vector1<bool> pb(3, true);
    cube geo(jd("L"), pb, 3);
    LangevinNVTR b(geo);
    b.setprecision(ji("precision"));
    b.setdt(jd("dt"));
    b.setgamma(jd("gamma"));
    b.setgammar(jd("gammar"));
    b.setkT(jd("kT"));
    matrix<double> x = to_mat(rd<double>("x.bin", (size_t)N * 3), N, 3);
    matrix<double> Q = to_mat(rd<double>("Q.bin", (size_t)N * 9), N, 9);
    matrix<double> p(N, 3), L(N, 3);
    b.initialize(x, p, Q, L);
    HSPotential wsa(1.0, 1.0);
    TetrahedralPatch pots(jd("eps_p"), jd("d"), jd("thetam"));
    const std::vector<double> nz = rd<double>("noise.bin", (size_t)(nsteps + 1) * N * 6);
    const std::string mode = g_job.at("mode");
    if (mode == "bindout") {
        // Condensate::run_singlebond(runtime, every, strbase) with its files
        BindingModelSingle bm(jd("on_rate"), jd("off_rate"));
        const int nh = pots.get_total_patches(N);
        BinaryBindStore bbs;
        bbs.isbound = vector1<bool>(nh);
        bbs.boundto = vector1<int>(nh);
        StepTerms t;
        t.wca = &wsa;
        t.patches = &pots;
        if (chdir(g_dir.c_str()) != 0)
            throw mdx_error(MDX_ERR_ARG, "cannot enter " + g_dir);
        b.run_singlebond(nsteps, ji("every"), "_test", t, bbs, bm);
        g_dir = ".";
    } else if (mode == "bind" || mode == "bindrun") {
        // Condensate::run_singlebond (Condensate.cpp:422-478) from an unbound
        // store: call by call, or as device-resident steps; model single
        // (on_rate, off_rate), binary (div, setup_equilibrium) or ternary
        // (SortingFunctionNonUniform div1 div2 np, tables from dtab.bin /
        // ttab.bin in the pair / triple order 11,12,13,22,23,33 and
        // 111,...,333)
        const std::string model = g_job.count("model") ? g_job.at("model") : "single";
        BindingModelSingle single(jd("on_rate"), jd("off_rate"));
        BindingModelBinary binary(g_job.count("div") ? ji("div") : 0);
        binary.setup_equilibrium();
        SortingFunctionNonUniform sf{g_job.count("div1") ? ji("div1") : 0,
                                     g_job.count("div2") ? ji("div2") : 0,
                                     g_job.count("np") ? ji("np") : 1};
        BindingModelTernary<SortingFunctionNonUniform> ternary(sf);
        if (model == "ternary") {
            const std::vector<double> dt = rd<double>("dtab.bin", 24), tt = rd<double>("ttab.bin", 160);
            vector1<double> *d[6] = {&ternary.doubr11, &ternary.doubr12, &ternary.doubr13,
                                     &ternary.doubr22, &ternary.doubr23, &ternary.doubr33};
            vector1<double> *t[10] = {&ternary.tripr111, &ternary.tripr112, &ternary.tripr113,
                                      &ternary.tripr122, &ternary.tripr123, &ternary.tripr133,
                                      &ternary.tripr222, &ternary.tripr223, &ternary.tripr233,
                                      &ternary.tripr333};
            for (int c = 0; c < 6; c++)
                for (int k = 0; k < 4; k++)
                    (*d[c])[k] = dt[(size_t)(4 * c + k)];
            for (int c = 0; c < 10; c++)
                for (int k = 0; k < 16; k++)
                    (*t[c])[k] = tt[(size_t)(16 * c + k)];
        }
        const AbstractBindingModel &bm = model == "binary"    ? (const AbstractBindingModel &)binary
                                         : model == "ternary" ? (const AbstractBindingModel &)ternary
                                                              : (const AbstractBindingModel &)single;
        const int nh = pots.get_total_patches(N);
        BinaryBindStore bbs;
        bbs.isbound = vector1<bool>(nh);
        bbs.boundto = vector1<int>(nh);
        if (mode == "bindrun") {
            StepTerms t;
            t.wca = &wsa;
            t.patches = &pots;
            b.run_singlebond(nsteps, t, bbs, bm, nz.data());
        } else {
            int tb;
            const int n = std::max(1, (int)(geo.l / 4.0));
            matrix<int> boxes = geo.generate_boxes_relationships(n, tb);
            matrix<int> *pairs = b.calculatepairs(boxes, 3.5);
            // the patch list and its runs, as the reference's loop builds them
            // (Condensate.cpp:474-479, every 10 steps); the engine's routine
            // takes its candidates from its own pair walk
            std::vector<patchint> opairs = b.calculate_patch_list(*pairs, pots);
            std::vector<int> runs_diff = adjacency(opairs);
            matrix<double> F = b.calculateforces(*pairs, wsa);
            matrix<double> T(N, 3);
            b.calculate_forces_and_torques3D_onlyone_nonlets(opairs, runs_diff, pots, bbs, bm, F, T);
            matrix<double> RT = noise_row(nz, 0, N, 6);
            b.create_forces_and_torques_sphere(F, T, RT);
            for (int i = 0; i < nsteps; i++) {
                if (i > 0 && i % 20 == 0) {
                    delete pairs;
                    pairs = b.calculatepairs(boxes, 3.5);
                }
                if (i > 0 && i % 10 == 0) {
                    opairs = b.calculate_patch_list(*pairs, pots);
                    runs_diff = adjacency(opairs);
                }
                b.advancemom_halfstep(F, T);
                b.advance_pos();
                b.rotate();
                F = b.calculateforces(*pairs, wsa);
                T.reset(0.0);
                b.calculate_forces_and_torques3D_onlyone_nonlets(opairs, runs_diff, pots, bbs, bm,
                                                                 F, T);
                RT = noise_row(nz, i + 1, N, 6);
                b.create_forces_and_torques_sphere(F, T, RT);
                b.advancemom_halfstep(F, T);
            }
            delete pairs;
        }
        std::vector<int> ib(nh), bt(nh);
        for (int k = 0; k < nh; k++) {
            ib[k] = bbs.isbound[k] ? 1 : 0;
            bt[k] = bbs.boundto[k];
        }
        wr("out_isbound.bin", ib.data(), (size_t)nh);
        wr("out_boundto.bin", bt.data(), (size_t)nh);
    } else if (mode == "plist") {
        // calculate_patch_list, adjacency and calculate_bound_pairs
        // (LangevinR.cpp:1702-2036) on the pairs of pairs.bin
        const int np = ji("npairs");
        matrix<int> pairs = to_mat(rd<int>("pairs.bin", (size_t)np * 2), np, 2);
        const std::vector<patchint> pl = b.calculate_patch_list(pairs, pots);
        const std::vector<int> runs = adjacency(pl);
        std::vector<int> rows(5 * pl.size() + 1);
        for (size_t k = 0; k < pl.size(); k++) {
            int a1, a2, pn, w1, w2;
            pl[k].get_all(a1, a2, pn, w1, w2);
            const int r[5] = {a1, a2, pn, w1, w2};
            std::copy(r, r + 5, rows.begin() + 5 * k);
        }
        wr("out_plist.bin", rows.data(), 5 * pl.size());
        wr("out_adjacency.bin", runs.data(), runs.size());
        const std::vector<patchint> bp = b.calculate_bound_pairs(pairs, pots);
        std::vector<int> brows(5 * bp.size() + 1);
        for (size_t k = 0; k < bp.size(); k++) {
            int a1, a2, pn, w1, w2;
            bp[k].get_all(a1, a2, pn, w1, w2);
            const int r[5] = {a1, a2, pn, w1, w2};
            std::copy(r, r + 5, brows.begin() + 5 * k);
        }
        wr("out_blist.bin", brows.data(), 5 * bp.size());
    } else if (mode == "run") {
        StepTerms t;
        t.wca = &wsa;
        t.patches = &pots;
        b.run(nsteps, t, nz.data());
    } else if (mode == "output") {
        // Condensate::run(runtime, every, strbase) with its CSV output, from dir
        StepTerms t;
        t.wca = &wsa;
        t.patches = &pots;
        if (chdir(g_dir.c_str()) != 0)
            throw mdx_error(MDX_ERR_ARG, "cannot enter " + g_dir);
        b.run(nsteps, ji("every"), "_test", t);
        g_dir = ".";
    } else {
        // Condensate::run, call by call
        int tb;
        const int n = std::max(1, (int)(geo.l / 4.0));
        matrix<int> boxes = geo.generate_boxes_relationships(n, tb);
        matrix<int> *pairs = b.calculatepairs(boxes, 3.5);
        matrix<double> F = b.calculateforces(*pairs, wsa);
        matrix<double> T(N, 3);
        b.calculate_forces_and_torques3D(*pairs, pots, F, T);
        matrix<double> RT = noise_row(nz, 0, N, 6);
        b.create_forces_and_torques_sphere(F, T, RT);
        for (int i = 0; i < nsteps; i++) {
            b.advancemom_halfstep(F, T);
            b.advance_pos();
            b.rotate();
            if (i > 0 && i % 20 == 0) {
                delete pairs;
                pairs = b.calculatepairs(boxes, 3.5);
            }
            F = b.calculateforces(*pairs, wsa);
            T.reset(0.0);
            b.calculate_forces_and_torques3D(*pairs, pots, F, T);
            RT = noise_row(nz, i + 1, N, 6);
            b.create_forces_and_torques_sphere(F, T, RT);
            b.advancemom_halfstep(F, T);
        }
        delete pairs;
    }
    matrix<double> xo = b.getdat(), po = b.getmom();
    wr("out_x.bin", xo.getdata(), (size_t)N * 3);
    wr("out_p.bin", po.getdata(), (size_t)N * 3);
    wr("out_Q.bin", b.getorientation().getdata(), (size_t)N * 9);
    wr("out_L.bin", b.getangmom().getdata(), (size_t)N * 3);
    return 0;
}

static int bead_main()
{
    const int N = ji("N"), nsteps = ji("nsteps");
    vector1<bool> pb(3, true);
    cube geo(jd("L"), pb, 3);
    LangevinNVT b(geo);
    b.setprecision(ji("precision"));
    b.setdt(jd("dt"));
    b.setgamma(jd("gamma"));
    b.setkT(jd("kT"));
    matrix<double> x = to_mat(rd<double>("x.bin", (size_t)N * 3), N, 3);
    b.setdat(x);
    const int nb = ji("nbonds"), na = ji("nangles");
    matrix<int> bonds = to_mat(rd<int>("bonds.bin", (size_t)nb * 2), nb, 2);
    matrix<int> angles = to_mat(rd<int>("angles.bin", (size_t)na * 3), na, 3);
    HSPotential wsa(1.0, 1.0);
    HarmonicPotential harm(jd("kbond"), jd("r0"));
    CosineBendingPotential bend(jd("kappa"));
    const std::vector<double> nz = rd<double>("noise.bin", (size_t)nsteps * N * 3);
    if (g_job.at("mode") == "run") {
        StepTerms t;
        t.wca = &wsa;
        t.bonds = &bonds;
        t.harmonic = &harm;
        t.triplets = &angles;
        t.cosine = &bend;
        b.run(nsteps, t, nz.data());
    } else {
        int tb;
        const int n = std::max(1, (int)(geo.l / 4.0));
        matrix<int> boxes = geo.generate_boxes_relationships(n, tb);
        matrix<int> *pairs = b.calculatepairs(boxes, 3.5);
        for (int i = 0; i < nsteps; i++) {
            if (i > 0 && i % 20 == 0) {
                delete pairs;
                pairs = b.calculatepairs(boxes, 3.5);
            }
            matrix<double> F = b.calculateforces(*pairs, wsa);
            F += b.calculateforces(bonds, harm);
            F += b.calculateforces_threebody(angles, bend);
            matrix<double> R = noise_row(nz, i, N, 3);
            b.advance_mom(F, R);
            b.advance_pos();
        }
        delete pairs;
    }
    matrix<double> xo = b.getdat(), po = b.getmom();
    wr("out_x.bin", xo.getdata(), (size_t)N * 3);
    wr("out_p.bin", po.getdata(), (size_t)N * 3);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: dropin_driver <dir>\n");
        return 2;
    }
    g_dir = argv[1];
    std::ifstream f(g_dir + "/job.txt");
    std::string k, v;
    while (f >> k >> v)
        g_job[k] = v;
    try {
        if (g_job.at("integ") == "csv") {
            // outfunc of a matrix read from x.bin (N x M), and importcsv of it (no GPU)
            const int N = ji("N"), M = ji("M");
            matrix<double> x = to_mat(rd<double>("x.bin", (size_t)N * M), N, M);
            outfunc(x, g_dir + "/out");
            bool err = false;
            matrix<double> y = importcsv(g_dir + "/out.csv", 0.0, err);
            if (err || y.getNsafe() != N || y.getncols() != M)
                throw mdx_error(MDX_ERR_ARG, "importcsv failed");
            wr("out_import.bin", y.getdata(), (size_t)N * M);
            return 0;
        }
        if (g_job.at("integ") == "dgraph") {
            // distance_graph(pos, boxes, geo, cut, i1) with index_test or excludea
            const int N = ji("N");
            vector1<bool> pb(3, true);
            cube geo(jd("L"), pb, 3);
            matrix<double> x = to_mat(rd<double>("x.bin", (size_t)N * 3), N, 3);
            int tb;
            matrix<int> boxes = geo.generate_boxes_relationships(ji("n"), tb);
            index_test all;
            excludea ex(ji("ex_lo"), ji("ex_hi"));
            index_test *i1 = ji("ex_lo") <= ji("ex_hi") ? &ex : &all;
            vector1<int> lab = distance_graph(x, boxes, geo, jd("cut"), i1);
            std::vector<int> out(N);
            for (int i = 0; i < N; i++)
                out[i] = lab[i];
            wr("out_labels.bin", out.data(), (size_t)N);
            return 0;
        }
        if (g_job.at("integ") == "stencil") {
            // cube::generate_boxes_relationships only (no GPU)
            vector1<bool> pb(3, true);
            cube geo(1.0, pb, 3);
            int tb;
            matrix<int> st = geo.generate_boxes_relationships(ji("n"), tb);
            wr("out_stencil.bin", st.getdata(), (size_t)st.getNsafe() * tb);
            return 0;
        }
        return g_job.at("integ") == "rotor" ? rotor_main() : bead_main();
    } catch (const std::exception &e) {
        fprintf(stderr, "dropin_driver: %s\n", e.what());
        return 1;
    }
}
