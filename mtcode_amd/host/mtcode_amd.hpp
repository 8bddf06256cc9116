// mtcode_amd.hpp -- C++ host classes over the engine's C ABI (include/mdx.h)
// with the API of the reference's system, potential and integrator classes,
// so a driver written against MDBase (MD, LangevinNVT, LangevinNVTR, cube,
// HSPotential, ComboPatch, ...) runs its step on the GPU.
//
// Header-only; needs only mdx.h (no HIP headers).  Link with -lmdx.
//
// State lives on the device between calls; the classes keep host mirrors of
// the reference's matrices and copy in either direction only when a getter or
// a setter needs it.  Each force or stage call corresponds to one reference
// call (the stage entry points of the ABI); run() is the fast path: the whole
// loop of Condensate::run (Condensate.cpp:318-377) or of the bead drivers
// (mainActivePolymer.cpp:381-426) as device-resident steps.
//
// Differences from the reference, by design:
//  - errors throw mtamd::mdx_error instead of error() printing, waiting on
//    stdin and exiting (Basic/basic.h:61-71);
//  - the pair list handed to the force calls is accepted for API
//    compatibility: the engine evaluates every pair within the interaction
//    reach, which equals the reference's result whenever the list holds all
//    pairs within the cut-off (as calculatepairs() with cut >= reach builds);
//  - matrix / vector1 are minimal row-major stand-ins holding what the
//    reference's DataStructures/matrix2.h and vector1.h hold.
//
// Reference paths are relative to the reference's Code/ folder.
#ifndef MTCODE_AMD_HPP
#define MTCODE_AMD_HPP

#include "mdx.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace mtamd {

class mdx_error : public std::runtime_error {
  public:
    int code;
    mdx_error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc)
{
    if (rc != MDX_OK)
        throw mdx_error(rc, mdx_last_error());
}

// ------------------------------------------------------------ containers
template <class T> class vector1 {
    // bytes for bool (std::vector<bool> hands out proxies, not references)
    using S = typename std::conditional<std::is_same<T, bool>::value, unsigned char, T>::type;
    std::vector<S> v_;

  public:
    vector1() {}
    explicit vector1(int n, T x = T()) : v_((size_t)n, (S)x) {}
    int getsize() const { return (int)v_.size(); }
    S &operator[](int i) { return v_[(size_t)i]; }
    const S &operator[](int i) const { return v_[(size_t)i]; }
};

// row-major n x m (DataStructures/matrix2.h:24-37)
template <class T> class matrix {
    int n_ = 0, m_ = 0;
    std::vector<T> v_;

  public:
    matrix() {}
    matrix(int n, int m, T x = T()) : n_(n), m_(m), v_((size_t)n * m, x) {}
    int getNsafe() const { return n_; }
    int getnrows() const { return n_; }
    int getncols() const { return m_; }
    T &operator()(int i, int j) { return v_[(size_t)i * m_ + j]; }
    const T &operator()(int i, int j) const { return v_[(size_t)i * m_ + j]; }
    T *getdata() { return v_.data(); }
    const T *getdata() const { return v_.data(); }
    void reset(T x) { std::fill(v_.begin(), v_.end(), x); }
    matrix &operator+=(const matrix &o)
    {
        if (o.n_ != n_ || o.m_ != m_)
            throw mdx_error(MDX_ERR_ARG, "matrix sizes differ in +=");
        for (size_t k = 0; k < v_.size(); k++)
            v_[k] += o.v_[k];
        return *this;
    }
};

// ------------------------------------------------------------ CSV files
// the reference's matrix output (operator<<=, DataStructures/matrix2.cpp:898-912):
// rows of comma-separated values in the stream's default format
template <class T> std::ostream &write_csv(std::ostream &s, const matrix<T> &a)
{
    const int nr = a.getNsafe(), nc = a.getncols();
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < nc; ++j) {
            if (j == nc - 1)
                s << a(i, j) << std::endl;
            else
                s << a(i, j) << ",";
        }
    return s;
}

// the vector1 CSV operator (DataStructures/vector1.cpp:682-693): precision 10,
// comma-separated on one line, no newline; bools as 0 / 1
template <class T> std::ostream &write_csv(std::ostream &s, const vector1<T> &v)
{
    s.precision(10);
    const int n = v.getsize();
    for (int i = 0; i < n; ++i) {
        if (std::is_same<T, bool>::value)
            s << (v[i] ? 1 : 0);
        else
            s << v[i];
        if (i != n - 1)
            s << ",";
    }
    return s;
}

// outfunc (matrix2.cpp:1557-1565): writes <s>.csv
template <class T> void outfunc(const matrix<T> &a, std::string s)
{
    s = s + ".csv";
    std::ofstream f(s.c_str());
    if (!f.is_open())
        throw mdx_error(MDX_ERR_ARG, "failed to open file " + s);
    write_csv(f, a);
}

// importcsv (matrix2.cpp:1596-1640): rows of comma-separated numbers (atof),
// short rows padded with zeros to the longest; error set when the file is
// missing
template <class T> matrix<T> importcsv(std::string filename, T, bool &error)
{
    std::ifstream data(filename.c_str());
    error = false;
    if (!data.is_open()) {
        error = true;
        return matrix<T>();
    }
    std::string line;
    std::vector<std::vector<T>> a;
    while (std::getline(data, line)) {
        std::stringstream ls(line);
        std::string cell;
        std::vector<T> b;
        while (std::getline(ls, cell, ','))
            b.push_back((T)std::atof(cell.c_str()));
        a.push_back(b);
    }
    size_t nz = 0;
    for (auto &r : a)
        nz = std::max(nz, r.size());
    matrix<T> q((int)a.size(), (int)nz);
    for (size_t i = 0; i < a.size(); i++)
        for (size_t j = 0; j < a[i].size(); j++)
            q((int)i, (int)j) = a[i][j];
    return q;
}

// ------------------------------------------------------------ geometry
// cube (MDBase/geometry.h:43-66): edge l, periodic flag per dimension
struct cube {
    int dimension;
    double l;
    vector1<bool> pb;
    bool periodic;
    double l2s;

    cube(double ll, vector1<bool> pbb, int dim) : dimension(dim), l(ll), pb(pbb)
    {
        if (pbb.getsize() != dim)
            throw mdx_error(MDX_ERR_ARG, "boundary conditions must have the system's dimension");
        periodic = pbb[0];
        l2s = (ll / 2.) * (ll / 2.);
    }

    // generate_boxes_relationships (geometry.h:523-609): the 3^3 stencil of
    // each of num^3 cells (index i + num j + num^2 k), periodic, sorted and
    // without repeats; tb receives the row length
    matrix<int> generate_boxes_relationships(int num, int &tb) const
    {
        if (num < 1)
            throw mdx_error(MDX_ERR_ARG, "cells per side must be >= 1");
        const int nc = num * num * num;
        std::vector<std::vector<int>> rows((size_t)nc);
        size_t width = 0;
        for (int k = 0; k < num; k++)
            for (int j = 0; j < num; j++)
                for (int i = 0; i < num; i++) {
                    std::vector<int> r;
                    for (int c = -1; c <= 1; c++)
                        for (int b = -1; b <= 1; b++)
                            for (int a = -1; a <= 1; a++) {
                                const int ii = ((i + a) % num + num) % num;
                                const int jj = ((j + b) % num + num) % num;
                                const int kk = ((k + c) % num + num) % num;
                                r.push_back(ii + num * jj + num * num * kk);
                            }
                    std::sort(r.begin(), r.end());
                    r.erase(std::unique(r.begin(), r.end()), r.end());
                    width = std::max(width, r.size());
                    rows[(size_t)i + num * j + (size_t)num * num * k] = r;
                }
        matrix<int> out(nc, (int)width);
        for (int c = 0; c < nc; c++)
            for (size_t w = 0; w < rows[(size_t)c].size(); w++)
                out(c, (int)w) = rows[(size_t)c][w];
        tb = (int)width;
        return out;
    }
};

// ------------------------------------------------------------ potentials
// HSPotential (Potentials/potential.h:79-125): WCA with the literal 1.12246
// sigma cut-off
struct HSPotential {
    double epsilon, sigma;
    HSPotential(double a, double b) : epsilon(a), sigma(b) {}
};
// HarmonicPotential (potential.h:265-297): 0.5 k (r - x0)^2
struct HarmonicPotential {
    double k, x0;
    HarmonicPotential(double a, double b) : k(a), x0(b) {}
};
// FENEPotential (potential.h:345-376)
struct FENEPotential {
    double kbond, R_0;
    FENEPotential(double a, double b) : kbond(a), R_0(b) {}
};
// BendingPotential (Potentials/potential3.h:23-130): 0.5 k (theta - theta0)^2
struct BendingPotential {
    double k0, theta0;
    BendingPotential(double a, double b) : k0(a), theta0(b) {}
};
// k (1 - cos theta) on consecutive bonds (no reference counterpart)
struct CosineBendingPotential {
    double k;
    explicit CosineBendingPotential(double a) : k(a) {}
};

// ComboPatch (Potentials/ComboPatch/combopatch.h:8-29): patch directions per
// type (body frame), KernFrenkelOnePatch2 rows (eps, d, theta_m) in
// GeneralPatch order, and the particle index ranges of the types
struct ComboPatch {
    std::vector<int> npatch;   // per type
    std::vector<double> dirs;  // sum(npatch) x 3, type-major
    std::vector<double> rows;  // rows x 3
    std::vector<int> num_per_type; // cumulative end index of each type; empty: all type 0
    virtual ~ComboPatch() {}
    int ntypes() const { return (int)npatch.size(); }
    // GeneralPatch::get_total_patches (GeneralPatch.cpp:284-291)
    int get_total_patches(int N) const
    {
        const std::vector<int> t = types(N);
        int nh = 0;
        for (int i = 0; i < N; i++)
            nh += npatch[(size_t)t[(size_t)i]];
        return nh;
    }
    std::vector<int> types(int N) const
    {
        std::vector<int> t((size_t)N, 0);
        if (num_per_type.empty())
            return t;
        int ty = 0;
        for (int i = 0; i < N; i++) {
            while (ty < ntypes() && i >= num_per_type[(size_t)ty])
                ty++;
            if (ty >= ntypes())
                throw mdx_error(MDX_ERR_ARG, "particle index beyond the last type range");
            t[(size_t)i] = ty;
        }
        return t;
    }
};

// TetrahedralPatch(strength, distance, angle) (TetrahedralPatch.cpp:4-82):
// one type, 4 tetrahedral patches, 16 equal rows
struct TetrahedralPatch : ComboPatch {
    TetrahedralPatch(double strr, double disss, double angg)
    {
        npatch = {4};
        const double a = std::sqrt(8.0 / 9.0), b = std::sqrt(2.0 / 9.0), c = std::sqrt(2.0 / 3.0);
        dirs = {a, 0.0, -1.0 / 3.0, -b, c, -1.0 / 3.0, -b, -c, -1.0 / 3.0, 0.0, 0.0, 1.0};
        for (int k = 0; k < 16; k++)
            rows.insert(rows.end(), {strr, disss, angg});
    }
};

// BivalentPatch(strength, distance, angle) (BivalentPatch.h:6-12): one type,
// patches along -x and +x, 4 equal rows
struct BivalentPatch : ComboPatch {
    BivalentPatch(double strr, double disss, double angg)
    {
        npatch = {2};
        dirs = {-1.0, 0.0, 0.0, 1.0, 0.0, 0.0};
        for (int k = 0; k < 4; k++)
            rows.insert(rows.end(), {strr, disss, angg});
    }
};

// GeneralPatch(patches per type, cumulative particle count per type, rows,
// directions) (GeneralPatch.cpp:7-121)
struct GeneralPatch : ComboPatch {
    GeneralPatch(vector1<int> no_patches_per_type, vector1<int> num_per_typee,
                 matrix<double> &params, matrix<double> &orient, bool = true)
    {
        const int nt = no_patches_per_type.getsize();
        if (num_per_typee.getsize() != nt)
            throw mdx_error(MDX_ERR_ARG, "one particle count per type required");
        int nd = 0, nr = 0;
        for (int t = 0; t < nt; t++) {
            npatch.push_back(no_patches_per_type[t]);
            num_per_type.push_back(num_per_typee[t]);
            nd += no_patches_per_type[t];
        }
        for (int i = 0; i < nt; i++)
            for (int j = i; j < nt; j++)
                nr += no_patches_per_type[i] * no_patches_per_type[j];
        if (orient.getNsafe() != nd || orient.getncols() != 3 || params.getNsafe() != nr ||
            params.getncols() < 3)
            throw mdx_error(MDX_ERR_ARG, "GeneralPatch: direction or parameter table size");
        dirs.assign(orient.getdata(), orient.getdata() + 3 * (size_t)nd);
        for (int r = 0; r < nr; r++)
            rows.insert(rows.end(), {params(r, 0), params(r, 1), params(r, 2)});
    }
};

// ------------------------------------------------------------ binding
// BinaryBindStore (Bindings/BinaryBindStore.h:7-11): per patch, bound or not
// and the partner patch (kept while unbound)
// one patch pair of the patch list (Rotational/LangevinR.h:20-57): the
// particles (index1 < index2), their patches and the potential's index
struct patchint {
    unsigned int particle_index1;
    unsigned int particle_index2;
    unsigned int patch_index1;
    unsigned int patch_index2;
    unsigned int potn;
    void get_all(int &a1, int &a2, int &p, int &pi1, int &pi2) const
    {
        a1 = (int)particle_index1;
        a2 = (int)particle_index2;
        p = (int)potn;
        pi1 = (int)patch_index1;
        pi2 = (int)patch_index2;
    }
    void get_particle(int &a1, int &a2) const
    {
        a1 = (int)particle_index1;
        a2 = (int)particle_index2;
    }
    void get_patch(int &p, int &pi1, int &pi2) const
    {
        p = (int)potn;
        pi1 = (int)patch_index1;
        pi2 = (int)patch_index2;
    }
    friend bool check_same_particle(const patchint &a, const patchint &b)
    {
        return a.particle_index1 == b.particle_index1 && a.particle_index2 == b.particle_index2;
    }
};

// adjacency (Rotational/LangevinR.cpp:2000-2036): the positions in a patch
// list where the particle pair differs from the entry before
inline std::vector<int> adjacency(const std::vector<patchint> &c)
{
    std::vector<int> out;
    for (size_t i = 1; i < c.size(); i++)
        if (!check_same_particle(c[i], c[i - 1]))
            out.push_back((int)i);
    return out;
}

struct BinaryBindStore {
    vector1<bool> isbound;
    vector1<int> boundto;
};

// AbstractBindingModel (Bindings/BindingModeAbstract.h:6-20): what a model
// hands the engine (mdx_set_binding, and for the species-table models
// mdx_set_binding_tables)
struct AbstractBindingModel {
    virtual ~AbstractBindingModel() {}
    virtual void apply(mdx_engine *e, int *nh) const = 0;
};

// BindingModelSingle(on_rate, off_rate) (Bindings/BindingModelSingle.cpp:1-8)
struct BindingModelSingle : AbstractBindingModel {
    double on_rate, off_rate;
    BindingModelSingle(double onn, double offf) : on_rate(onn), off_rate(offf) {}
    void apply(mdx_engine *e, int *nh) const override
    {
        check(mdx_set_binding(e, on_rate, off_rate, nh));
    }
};

// the species-table form shared by the binary and ternary models: the
// species of each patch (1..3), the doublet rates of each sorted species
// pair (11,12,13,22,23,33) and the triplet rates of each sorted triple
// (111,112,113,122,123,133,222,223,233,333), with the "to none" rate each
// triple reads
inline void apply_tables(mdx_engine *e, int *nh, const std::vector<int> &species,
                         const vector1<double> *const d[6], const vector1<double> *const t[10],
                         const vector1<double> *const r4[10])
{
    check(mdx_set_binding(e, 1.0, 0.0, nh));
    if ((int)species.size() != *nh)
        throw mdx_error(MDX_ERR_ARG, "species per patch");
    std::vector<double> dtab(24, 0.0), ttab(160, 0.0), tr4(40, 0.0);
    for (int c = 0; c < 6; c++)
        if (d[c])
            for (int k = 0; k < 4; k++)
                dtab[(size_t)(4 * c + k)] = (*d[c])[k];
    for (int c = 0; c < 10; c++)
        if (t[c]) {
            for (int k = 0; k < 16; k++)
                ttab[(size_t)(16 * c + k)] = (*t[c])[k];
            for (int f = 0; f < 4; f++)
                tr4[(size_t)(4 * c + f)] = (*r4[c])[4 * f + 3];
        }
    check(mdx_set_binding_tables(e, species.data(), dtab.data(), ttab.data(), tr4.data()));
}

// BindingModelBinary(div) (Bindings/BindingModelBinary.h:6-41): species 1
// below patch div, 2 from it; setup_equilibrium() sets the model's own
// tables (BindingModelBinary.cpp:8-96).  As in the reference, its 122 and
// 222 triplets read their rate to none from the 111 table.
struct BindingModelBinary : AbstractBindingModel {
    vector1<double> doubr11, doubr12, doubr22, tripr111, tripr112, tripr122, tripr222;
    int div;
    explicit BindingModelBinary(int divv)
        : doubr11(4), doubr12(4), doubr22(4), tripr111(16), tripr112(16), tripr122(16),
          tripr222(16), div(divv)
    {
    }
    void setup_equilibrium()
    {
        const double d[4] = {0.0, 1.0, 0.0, 1.0};
        // from (bound 12, 23, 13, none) x to (12, 23, 13, none)
        const double t[16] = {0.998, 0.001, 0.001, 0.000, 0.001, 0.998, 0.001, 0.000,
                              0.001, 0.001, 0.998, 0.000, 0.33,  0.33,  0.33,  0.01};
        for (int k = 0; k < 4; k++)
            doubr11[k] = doubr12[k] = doubr22[k] = d[k];
        for (int k = 0; k < 16; k++)
            tripr111[k] = tripr112[k] = tripr122[k] = tripr222[k] = t[k];
    }
    void apply(mdx_engine *e, int *nh) const override
    {
        int n = 0;
        check(mdx_set_binding(e, 1.0, 0.0, &n));
        std::vector<int> sp((size_t)n);
        for (int i = 0; i < n; i++)
            sp[(size_t)i] = i < div ? 1 : 2;
        const vector1<double> *d[6] = {&doubr11, &doubr12, nullptr, &doubr22, nullptr, nullptr};
        const vector1<double> *t[10] = {&tripr111, &tripr112, nullptr, &tripr122, nullptr,
                                        nullptr, &tripr222, nullptr, nullptr, nullptr};
        const vector1<double> *r4[10] = {&tripr111, &tripr112, nullptr, &tripr111, nullptr,
                                         nullptr, &tripr111, nullptr, nullptr, nullptr};
        apply_tables(e, nh, sp, d, t, r4);
    }
};

// SortingFunctionUniform / SortingFunctionNonUniform
// (Bindings/BindingModelTernary.h:7-37): the species of a patch index
struct SortingFunctionUniform {
    int div1, div2;
    int operator()(const int &i) const { return i < div1 ? 1 : i < div2 ? 2 : 3; }
};
struct SortingFunctionNonUniform {
    int div1, div2, np;
    int operator()(const int &i) const
    {
        if (i < div1)
            return 1;
        if (i < div2)
            return 2;
        return (i - div2) % np == np - 1 ? 3 : 1;
    }
};

// BindingModelTernary<Q>(sorter) (Bindings/BindingModelTernary.h:38-101):
// species by the sorting function; the caller fills the rate tables (the
// reference's setup(...) builds them from its parameters)
template <typename Q> struct BindingModelTernary : AbstractBindingModel {
    vector1<double> doubr11, doubr12, doubr22, doubr13, doubr23, doubr33;
    vector1<double> tripr111, tripr112, tripr113, tripr122, tripr123, tripr133, tripr222,
        tripr223, tripr233, tripr333;
    Q func;
    explicit BindingModelTernary(Q &f)
        : doubr11(4), doubr12(4), doubr22(4), doubr13(4), doubr23(4), doubr33(4), tripr111(16),
          tripr112(16), tripr113(16), tripr122(16), tripr123(16), tripr133(16), tripr222(16),
          tripr223(16), tripr233(16), tripr333(16), func(f)
    {
    }
    void apply(mdx_engine *e, int *nh) const override
    {
        int n = 0;
        check(mdx_set_binding(e, 1.0, 0.0, &n));
        std::vector<int> sp((size_t)n);
        for (int i = 0; i < n; i++)
            sp[(size_t)i] = func(i);
        const vector1<double> *d[6] = {&doubr11, &doubr12, &doubr13, &doubr22, &doubr23, &doubr33};
        const vector1<double> *t[10] = {&tripr111, &tripr112, &tripr113, &tripr122, &tripr123,
                                        &tripr133, &tripr222, &tripr223, &tripr233, &tripr333};
        apply_tables(e, nh, sp, d, t, t);
    }
};

// ------------------------------------------------------------ integrators
// Interaction terms of a whole step (run()).
struct StepTerms {
    const HSPotential *wca = nullptr;
    const ComboPatch *patches = nullptr;
    const matrix<int> *bonds = nullptr;          // B x 2
    const HarmonicPotential *harmonic = nullptr; // one k, r0 for all bonds, or
    const FENEPotential *fene = nullptr;
    const matrix<int> *triplets = nullptr;       // A x 3
    const BendingPotential *bending = nullptr;
    const CosineBendingPotential *cosine = nullptr;
    const matrix<int> *exclusions = nullptr;     // pairs without WCA
};

// LangevinNVT (MDBase/Langevin.h, Langevin.cpp): translational Langevin
// dynamics of point particles in a cube
class LangevinNVT {
  protected:
    cube geo_;
    double kT_ = 1, gamma_ = 1, dt_ = 0.005, m_ = 1;
    double gammar_ = 1, im_[3] = {0.1, 0.1, 0.1}, maxd2_ = 2.4;
    int integ_;
    int prec_ = MDX_PREC_FP64;
    unsigned long long seed_ = 1;
    int device_ = 0;
    matrix<double> dat_, mom_, orient_, angmom_;
    mdx_engine *e_ = nullptr;
    bool host_stale_ = false; // the device holds a newer state
    bool dev_stale_ = true;   // the host mirrors hold a newer state

    bool rotor() const { return integ_ == MDX_INTEG_ROTOR; }
    int N() const { return dat_.getNsafe(); }

    void pull()
    {
        if (!host_stale_ || !e_)
            return;
        check(mdx_download_state(e_, dat_.getdata(), mom_.getdata(),
                                 rotor() ? orient_.getdata() : nullptr,
                                 rotor() ? angmom_.getdata() : nullptr));
        host_stale_ = false;
    }
    // parameters held by the engine descriptor: a change rebuilds the engine
    void drop_engine()
    {
        if (!e_)
            return;
        pull();
        mdx_destroy(e_);
        e_ = nullptr;
        dev_stale_ = true;
    }
    void touch_state()
    {
        pull();
        dev_stale_ = true;
    }
    mdx_engine *engine()
    {
        if (N() == 0)
            throw mdx_error(MDX_ERR_STATE, "no particles (setdat first)");
        if (!e_) {
            const int n = N();
            const double box[3] = {geo_.l, geo_.l, geo_.l};
            mdx_desc d = {};
            d.nreplicas = 1;
            d.n_per_replica = &n;
            d.box = box;
            d.periodic = geo_.periodic ? 1 : 0;
            d.kT = &kT_;
            d.dt = dt_;
            d.m = m_;
            d.gamma = gamma_;
            d.gammar = gammar_;
            for (int k = 0; k < 3; k++)
                d.inertia[k] = im_[k];
            d.precision = prec_;
            d.integrator = integ_;
            d.seed = seed_;
            d.device = device_;
            check(mdx_create(&d, &e_));
            dev_stale_ = true;
        }
        if (dev_stale_) {
            check(mdx_upload_state(e_, dat_.getdata(), mom_.getdata(),
                                   rotor() ? orient_.getdata() : nullptr,
                                   rotor() ? angmom_.getdata() : nullptr));
            dev_stale_ = false;
            host_stale_ = false;
        }
        return e_;
    }
    void set_param(double &field, double v)
    {
        if (field != v) {
            drop_engine();
            field = v;
        }
    }
    void apply_terms(const StepTerms &t)
    {
        mdx_engine *e = engine();
        check(mdx_clear_interactions(e, MDX_TERM_ALL));
        const int n = N();
        if (t.wca)
            check(mdx_set_pair_wca(e, t.wca->epsilon, t.wca->sigma, 0.0));
        if (t.patches) {
            const std::vector<int> ty = t.patches->types(n);
            check(mdx_set_patches(e, t.patches->ntypes(), ty.data(), t.patches->npatch.data(),
                                  t.patches->dirs.data(), t.patches->rows.data(), maxd2_));
        }
        if (t.bonds && (t.harmonic || t.fene)) {
            const long long nb = t.bonds->getNsafe();
            std::vector<double> k((size_t)nb, t.harmonic ? t.harmonic->k : t.fene->kbond);
            std::vector<double> r((size_t)nb, t.harmonic ? t.harmonic->x0 : t.fene->R_0);
            check(mdx_set_bonds(e, nb, t.bonds->getdata(), k.data(), r.data(),
                                t.harmonic ? MDX_BOND_HARMONIC : MDX_BOND_FENE));
        }
        if (t.triplets && (t.bending || t.cosine)) {
            const long long na = t.triplets->getNsafe();
            std::vector<double> k((size_t)na, t.bending ? t.bending->k0 : t.cosine->k);
            std::vector<double> th((size_t)na, t.bending ? t.bending->theta0 : 0.0);
            check(mdx_set_angles(e, na, t.triplets->getdata(), k.data(), th.data(),
                                 t.bending ? MDX_ANGLE_HARMONIC : MDX_ANGLE_COSINE));
        }
        if (t.exclusions)
            check(mdx_set_exclusions(e, t.exclusions->getNsafe(), t.exclusions->getdata()));
    }
    // conservative forces (and patch torques) of the given terms, lab frame
    void forces(const StepTerms &t, matrix<double> &F, matrix<double> *T)
    {
        apply_terms(t);
        F = matrix<double>(N(), 3);
        matrix<double> Tt(N(), 3);
        check(mdx_compute_forces(e_, F.getdata(), rotor() ? Tt.getdata() : nullptr));
        if (T)
            *T = Tt;
    }

  public:
    explicit LangevinNVT(cube &a) : geo_(a), integ_(MDX_INTEG_BEAD) {}
    LangevinNVT(const LangevinNVT &) = delete;
    LangevinNVT &operator=(const LangevinNVT &) = delete;
    virtual ~LangevinNVT()
    {
        if (e_)
            mdx_destroy(e_);
    }

    // ---- MD (MDBase/MD.h)
    virtual void setdat(const matrix<double> &x)
    {
        if (x.getncols() != 3)
            throw mdx_error(MDX_ERR_ARG, "positions must be N x 3");
        pull();
        if (x.getNsafe() != N())
            drop_engine();
        if (mom_.getNsafe() != x.getNsafe())
            mom_ = matrix<double>(x.getNsafe(), 3);
        dat_ = x;
        dev_stale_ = true;
    }
    matrix<double> getdat()
    {
        pull();
        return dat_;
    }
    int getN() const { return N(); }
    cube getgeo() const { return geo_; }

    // ---- LangevinNVT (MDBase/Langevin.h)
    virtual void setgamma(double g) { set_param(gamma_, g); }
    virtual void setdt(double t) { set_param(dt_, t); }
    virtual void setkT(double t) { set_param(kT_, t); }
    void setm(double v) { set_param(m_, v); }
    void setmom(matrix<double> &p)
    {
        if (p.getNsafe() != N() || p.getncols() != 3)
            throw mdx_error(MDX_ERR_ARG, "momenta must be N x 3");
        touch_state();
        mom_ = p;
    }
    matrix<double> getmom()
    {
        pull();
        return mom_;
    }
    double getkT() const { return kT_; }
    double getm() const { return m_; }
    double getgamma() const { return gamma_; }
    double getdt() const { return dt_; }

    // ---- engine options (no reference counterpart)
    void setprecision(int p) { set_param_i(prec_, p); }
    void setdevice(int d) { set_param_i(device_, d); }
    void setseed(unsigned long long s)
    {
        if (s != seed_) {
            drop_engine();
            seed_ = s;
        }
    }
    mdx_engine *handle() { return engine(); }

    // ---- pair search (MD::calculatepairs, MD.cpp:1035-1098): every i < j
    // with distance below cut (cube::distance_less_than, geometry.h:145-164),
    // cell side from the stencil table's row count; caller deletes
    matrix<int> *calculatepairs(matrix<int> &boxes, double cut)
    {
        int n = (int)std::lround(std::cbrt((double)boxes.getNsafe()));
        if (n < 1 || n * n * n != boxes.getNsafe())
            throw mdx_error(MDX_ERR_ARG, "stencil table must have n^3 rows");
        const int nn[3] = {n, n, n};
        mdx_engine *e = engine();
        long long np = 0;
        int rc = mdx_neighbor_pairs(e, 0, nn, cut, nullptr, 0, &np);
        if (rc != MDX_OK && rc != MDX_ERR_CAPACITY)
            check(rc);
        matrix<int> *P = new matrix<int>((int)np, 2);
        rc = mdx_neighbor_pairs(e, 0, nn, cut, P->getdata(), np, &np);
        if (rc != MDX_OK) {
            delete P;
            check(rc);
        }
        return P;
    }

    // ---- analysis: distance_graph (Analysis/AnalyzeGrowth.cpp:136-177) on
    // the resident positions: the pairs within binding_distance on the
    // stencil table's cells, without those touching indices ex_lo..ex_hi when
    // ex_lo <= ex_hi; label = smallest index of the connected component
    vector1<int> distance_graph(matrix<int> &boxes, double binding_distance, int ex_lo = 0,
                                int ex_hi = -1)
    {
        int n = (int)std::lround(std::cbrt((double)boxes.getNsafe()));
        if (n < 1 || n * n * n != boxes.getNsafe())
            throw mdx_error(MDX_ERR_ARG, "stencil table must have n^3 rows");
        const int nn[3] = {n, n, n};
        std::vector<int> lab((size_t)N());
        check(mdx_clusters(engine(), 0, nn, binding_distance, ex_lo, ex_hi, lab.data(), nullptr));
        vector1<int> out(N());
        for (int i = 0; i < N(); i++)
            out[i] = lab[(size_t)i];
        return out;
    }

    // ---- forces (MD::calculateforces, MD.cpp:1684-1764, and friends); the
    // pair/bond lists are the terms' support (see the header comment)
    matrix<double> calculateforces(matrix<int> &, HSPotential &p)
    {
        StepTerms t;
        t.wca = &p;
        matrix<double> F;
        forces(t, F, nullptr);
        return F;
    }
    matrix<double> calculateforces(matrix<int> &bonds, HarmonicPotential &p)
    {
        StepTerms t;
        t.bonds = &bonds;
        t.harmonic = &p;
        matrix<double> F;
        forces(t, F, nullptr);
        return F;
    }
    matrix<double> calculateforces(matrix<int> &bonds, FENEPotential &p)
    {
        StepTerms t;
        t.bonds = &bonds;
        t.fene = &p;
        matrix<double> F;
        forces(t, F, nullptr);
        return F;
    }
    // per-bond rest lengths, one k (MD.cpp:2037-2111)
    matrix<double> calculateforcesharmonic(matrix<int> &bonds, matrix<double> &bondlengths, double k)
    {
        mdx_engine *e = engine();
        check(mdx_clear_interactions(e, MDX_TERM_ALL));
        const long long nb = bonds.getNsafe();
        std::vector<double> kk((size_t)nb, k), r0((size_t)nb);
        for (long long b = 0; b < nb; b++)
            r0[(size_t)b] = bondlengths((int)b, 0);
        check(mdx_set_bonds(e, nb, bonds.getdata(), kk.data(), r0.data(), MDX_BOND_HARMONIC));
        matrix<double> F(N(), 3), T(N(), 3);
        check(mdx_compute_forces(e, F.getdata(), rotor() ? T.getdata() : nullptr));
        return F;
    }
    matrix<double> calculateforces_threebody(matrix<int> &triplets, BendingPotential &p)
    {
        StepTerms t;
        t.triplets = &triplets;
        t.bending = &p;
        matrix<double> F;
        forces(t, F, nullptr);
        return F;
    }
    matrix<double> calculateforces_threebody(matrix<int> &triplets, CosineBendingPotential &p)
    {
        StepTerms t;
        t.triplets = &triplets;
        t.cosine = &p;
        matrix<double> F;
        forces(t, F, nullptr);
        return F;
    }

    // ---- stages (Langevin.cpp:205-269)
    void advance_pos()
    {
        check(mdx_stage_advance_pos(engine()));
        host_stale_ = true;
    }
    virtual void advance_mom(matrix<double> &F, matrix<double> &R)
    {
        check(mdx_stage_advance_mom(engine(), F.getdata(), R.getdata()));
        host_stale_ = true;
    }

    // ---- whole steps on the device (bead loop, mainActivePolymer.cpp:390-425):
    // noise null (engine Philox) or nsteps x N x 3
    void run(int nsteps, const StepTerms &t, const double *noise = nullptr)
    {
        if (rotor())
            throw mdx_error(MDX_ERR_STATE, "use LangevinNVTR::run for rotors");
        apply_terms(t);
        check(mdx_step(e_, nsteps, noise));
        host_stale_ = true;
    }

  protected:
    void set_param_i(int &field, int v)
    {
        if (field != v) {
            drop_engine();
            field = v;
        }
    }
};

// LangevinNVTR (MDBase/Rotational/LangevinR.h, LangevinR.cpp): rotors with
// orientation (lab -> body) and body-frame angular momentum
class LangevinNVTR : public LangevinNVT {
    unsigned long long bind_draw_ = 0; // draws of the call-by-call binding routine
    // the binding model and the bound state into the engine (after the
    // patches are set)
    void bind_push(const AbstractBindingModel &bm, const BinaryBindStore &bo)
    {
        int nh = 0;
        bm.apply(e_, &nh);
        if (bo.isbound.getsize() != nh || bo.boundto.getsize() != nh)
            throw mdx_error(MDX_ERR_ARG, "BinaryBindStore size is not the number of patches");
        std::vector<int> ib((size_t)nh), bt((size_t)nh);
        for (int k = 0; k < nh; k++) {
            ib[(size_t)k] = bo.isbound[k] ? 1 : 0;
            bt[(size_t)k] = bo.boundto[k];
        }
        check(mdx_set_bindings(e_, ib.data(), bt.data()));
    }
    void bind_pull(BinaryBindStore &bo)
    {
        const int nh = bo.isbound.getsize();
        std::vector<int> ib((size_t)nh), bt((size_t)nh);
        check(mdx_get_bindings(e_, ib.data(), bt.data()));
        for (int k = 0; k < nh; k++) {
            bo.isbound[k] = ib[(size_t)k] != 0;
            bo.boundto[k] = bt[(size_t)k];
        }
    }

  public:
    explicit LangevinNVTR(cube &a) : LangevinNVT(a) { integ_ = MDX_INTEG_ROTOR; }

    void setdat(const matrix<double> &x) override
    {
        LangevinNVT::setdat(x);
        if (orient_.getNsafe() != N()) {
            orient_ = matrix<double>(N(), 9);
            for (int i = 0; i < N(); i++)
                orient_(i, 0) = orient_(i, 4) = orient_(i, 8) = 1.0;
        }
        if (angmom_.getNsafe() != N())
            angmom_ = matrix<double>(N(), 3);
    }
    // initialize(x, p, Q, L) (LangevinR.cpp:308-325)
    void initialize(matrix<double> &x, matrix<double> &p, matrix<double> &Q, matrix<double> &L)
    {
        setdat(x);
        setmom(p);
        setorientation(Q);
        setangularmomenta(L);
    }
    void setgammar(double b) { set_param(gammar_, b); }
    void setmaxdistance(double distance) { maxd2_ = distance; }
    void setIM(const vector1<double> &im)
    {
        if (im.getsize() < 9)
            throw mdx_error(MDX_ERR_ARG, "moment of inertia tensor has 9 entries");
        set_param(im_[0], im[0]);
        set_param(im_[1], im[4]);
        set_param(im_[2], im[8]);
    }
    void setorientation(matrix<double> &ort)
    {
        if (ort.getNsafe() != N() || ort.getncols() != 9)
            throw mdx_error(MDX_ERR_ARG, "set orient dimensions must match in MD");
        touch_state();
        orient_ = ort;
    }
    void setangularmomenta(matrix<double> &amom)
    {
        if (amom.getNsafe() != N() || amom.getncols() != 3)
            throw mdx_error(MDX_ERR_ARG, "angular momenta must be N x 3");
        touch_state();
        angmom_ = amom;
    }
    matrix<double> &getorientation()
    {
        pull();
        return orient_;
    }
    matrix<double> &getangmom()
    {
        pull();
        return angmom_;
    }

    // calculate_forces_and_torques3D(pairs, ComboPatch&, F, T)
    // (LangevinR.cpp:911-1047): adds the patch forces and lab torques
    void calculate_forces_and_torques3D(matrix<int> &, ComboPatch &pots, matrix<double> &F,
                                        matrix<double> &T)
    {
        StepTerms t;
        t.patches = &pots;
        matrix<double> Fp, Tp;
        forces(t, Fp, &Tp);
        F += Fp;
        T += Tp;
    }
    // calculate_forces_and_torques3D_onlyone_nonlets(pairs, divs, ComboPatch&,
    // BinaryBindStore&, AbstractBindingModel&, F, T)
    // (Rotational/LangevinRforce.cpp:2874-4136): updates the bound state and
    // adds the bound pairs' forces and lab torques.  The candidate pairs come
    // from the engine's own neighbour search (the patch-list arguments keep
    // the reference's signature and are not read); the decisions draw the
    // engine's Philox numbers, a new draw per call.
    template <class PL, class DV>
    void calculate_forces_and_torques3D_onlyone_nonlets(PL &, DV &, ComboPatch &iny,
                                                        BinaryBindStore &bo,
                                                        const AbstractBindingModel &bm,
                                                        matrix<double> &F, matrix<double> &T)
    {
        StepTerms t;
        t.patches = &iny;
        apply_terms(t);
        bind_push(bm, bo);
        matrix<double> Fb(N(), 3), Tb(N(), 3);
        check(mdx_bind_forces(e_, bind_draw_++, Fb.getdata(), Tb.getdata()));
        bind_pull(bo);
        F += Fb;
        T += Tb;
    }
    // calculate_patch_list(pairs, ComboPatch&) (LangevinR.cpp:1702-1853):
    // the patch pairs of the listed particle pairs whose cones may bring them
    // into binding range, computed on the device (mdx_patch_list), in pair
    // order and ascending potential; particles numbered by type
    std::vector<patchint> calculate_patch_list(matrix<int> &pairs, ComboPatch &iny)
    {
        return patch_rows(pairs, iny, mdx_patch_list);
    }
    // calculate_bound_pairs(pairs, ComboPatch&) (LangevinR.cpp:1854-1998):
    // the patch pairs within max_check whose cones both pass their
    // potential's cos(theta_m), on the device (mdx_bound_pairs)
    std::vector<patchint> calculate_bound_pairs(matrix<int> &pairs, ComboPatch &iny)
    {
        return patch_rows(pairs, iny, mdx_bound_pairs);
    }

  private:
    std::vector<patchint> patch_rows(matrix<int> &pairs, ComboPatch &iny,
                                     int (*fn)(mdx_engine *, const int *, long long, int *,
                                               long long, long long *))
    {
        StepTerms t;
        t.patches = &iny;
        apply_terms(t);
        const long long np = pairs.getNsafe();
        if (np > 0 && pairs.getncols() < 2)
            throw mdx_error(MDX_ERR_ARG, "pairs must have two columns");
        std::vector<int> ij((size_t)(2 * np));
        for (long long k = 0; k < np; k++) {
            ij[(size_t)(2 * k)] = pairs((int)k, 0);
            ij[(size_t)(2 * k + 1)] = pairs((int)k, 1);
        }
        long long cap = 8 * np, n = 0;
        std::vector<int> rows((size_t)(5 * std::max(1LL, cap)));
        int rc = fn(e_, ij.data(), np, rows.data(), cap, &n);
        if (rc == MDX_ERR_CAPACITY && n > cap) {
            cap = n;
            rows.assign((size_t)(5 * cap), 0);
            rc = fn(e_, ij.data(), np, rows.data(), cap, &n);
        }
        check(rc);
        std::vector<patchint> out((size_t)n);
        for (long long k = 0; k < n; k++) {
            const int *r = rows.data() + 5 * k;
            out[(size_t)k] = {(unsigned int)r[0], (unsigned int)r[1], (unsigned int)r[3],
                              (unsigned int)r[4], (unsigned int)r[2]};
        }
        return out;
    }

  public:
    // the loop of Condensate::run_singlebond (Condensate.cpp:464-478) on the
    // device: the set-up evaluation and nsteps steps of WCA + the bound pairs'
    // terms (t: the WCA and the patches), the bound state carried in bo.
    // noise null (engine Philox) or (nsteps + 1) x N x 6, row 0 for the set-up.
    void run_singlebond(int nsteps, const StepTerms &t, BinaryBindStore &bo,
                        const AbstractBindingModel &bm, const double *noise = nullptr)
    {
        if (!t.patches)
            throw mdx_error(MDX_ERR_ARG, "run_singlebond needs the patches");
        apply_terms(t);
        bind_push(bm, bo);
        check(mdx_prime(e_, noise));
        if (nsteps > 0)
            check(mdx_step(e_, nsteps, noise ? noise + 6 * (size_t)N() : nullptr));
        host_stale_ = true;
        bind_pull(bo);
    }
    // Condensate::run_singlebond(runtime, every, strbase) (Condensate.cpp:
    // 422-618) with its output: after step i with i % every == 0 and
    // i > int(0.9 runtime), pos<strbase>_i=<k>.csv, orientation<...>.csv (at
    // precision 10), bindings<...>.csv (isbound, a newline, boundto) and
    // connectedcomponents<...>.csv (distance_graph at 1.25 on
    // floor(l / (3 size_mol)) cells per side), k = i / every zero-padded to
    // the digits of ceil(runtime / every); the steps on the device with
    // Philox noise
    void run_singlebond(int runtime, int every, const std::string &strbase, const StepTerms &t,
                        BinaryBindStore &bo, const AbstractBindingModel &bm, double size_mol = 1.0)
    {
        if (every < 1)
            throw mdx_error(MDX_ERR_ARG, "every must be >= 1");
        if (!t.patches)
            throw mdx_error(MDX_ERR_ARG, "run_singlebond needs the patches");
        int tf = (int)std::ceil((double)runtime / (double)every);
        int digits = 0;
        do {
            ++digits;
            tf /= 10;
        } while (tf);
        int ccc;
        matrix<int> boxes = geo_.generate_boxes_relationships(
            std::max(1, (int)std::floor(geo_.l / (3. * size_mol))), ccc);
        apply_terms(t);
        bind_push(bm, bo);
        check(mdx_prime(e_, nullptr));
        int done = 0; // steps taken
        for (int i = 0; i < runtime; i++) {
            if (!(i % every == 0 && i > int(0.9 * runtime)))
                continue;
            check(mdx_step(e_, i + 1 - done, nullptr));
            done = i + 1;
            host_stale_ = true;
            bind_pull(bo);
            std::stringstream ss;
            ss << std::setw(digits) << std::setfill('0') << (i / every);
            const std::string k = strbase + "_i=" + ss.str() + ".csv";
            std::ofstream fp(("pos" + k).c_str()), fo(("orientation" + k).c_str()),
                fb(("bindings" + k).c_str()), fc(("connectedcomponents" + k).c_str());
            if (!fp.is_open() || !fo.is_open() || !fb.is_open() || !fc.is_open())
                throw mdx_error(MDX_ERR_ARG, "cannot open output files");
            write_csv(fp, getdat());
            fo << std::setprecision(10);
            write_csv(fo, getorientation());
            write_csv(fb, bo.isbound);
            fb << "\n";
            write_csv(fb, bo.boundto);
            write_csv(fc, distance_graph(boxes, 1.25));
        }
        if (runtime > done) {
            check(mdx_step(e_, runtime - done, nullptr));
            host_stale_ = true;
            bind_pull(bo);
        }
    }
    // create_forces_and_torques_sphere(F, T, RT) (LangevinR.cpp:475-578)
    void create_forces_and_torques_sphere(matrix<double> &F, matrix<double> &T, matrix<double> &RT)
    {
        check(mdx_stage_sphere(engine(), F.getdata(), T.getdata(), RT.getdata()));
    }
    // advancemom_halfstep(F, T) (LangevinR.cpp:327-347)
    void advancemom_halfstep(matrix<double> &F, matrix<double> &T)
    {
        check(mdx_stage_halfkick(engine(), F.getdata(), T.getdata()));
        host_stale_ = true;
    }
    // rotate() (LangevinR.cpp:595-695)
    void rotate()
    {
        check(mdx_stage_rotate(engine()));
        host_stale_ = true;
    }
    // measured_temperature(out) (LangevinR.cpp:170-193): <p_k^2/2m>, <L_k^2/2I>
    void measured_temperature(vector1<double> &out)
    {
        double o[16];
        check(mdx_observables(engine(), -1, o));
        out = vector1<double>(6);
        for (int k = 0; k < 6; k++)
            out[k] = o[k];
    }

    // ---- whole steps on the device: the loop of Condensate::run
    // (Condensate.cpp:318-377), set-up force evaluation included.  noise
    // null (engine Philox) or (nsteps + 1) x N x 6, row 0 for the set-up.
    void run(int nsteps, const StepTerms &t, const double *noise = nullptr)
    {
        apply_terms(t);
        check(mdx_prime(e_, noise));
        if (nsteps > 0)
            check(mdx_step(e_, nsteps, noise ? noise + 6 * (size_t)N() : nullptr));
        host_stale_ = true;
    }

    // Condensate::run(runtime, every, strbase) (Condensate.cpp:288-419) with
    // its output: after step i with i % every == 0, pos<strbase>_i=<k>.csv and
    // orientation<strbase>_i=<k>.csv (k = i / every, zero-padded to the digits
    // of ceil(runtime / every)), in the reference's CSV format
    void run(int runtime, int every, const std::string &strbase, const StepTerms &t)
    {
        if (every < 1)
            throw mdx_error(MDX_ERR_ARG, "every must be >= 1");
        int tf = (int)std::ceil((double)runtime / (double)every);
        int digits = 0;
        do {
            ++digits;
            tf /= 10;
        } while (tf);
        apply_terms(t);
        check(mdx_prime(e_, nullptr));
        int done = 0; // steps taken
        for (int i = 0; i < runtime; i += every) {
            // advance to the end of step i
            check(mdx_step(e_, i + 1 - done, nullptr));
            done = i + 1;
            host_stale_ = true;
            std::stringstream ss;
            ss << std::setw(digits) << std::setfill('0') << (i / every);
            std::ofstream fp(("pos" + strbase + "_i=" + ss.str() + ".csv").c_str());
            std::ofstream fo(("orientation" + strbase + "_i=" + ss.str() + ".csv").c_str());
            if (!fp.is_open() || !fo.is_open())
                throw mdx_error(MDX_ERR_ARG, "cannot open output files");
            write_csv(fp, getdat());
            write_csv(fo, getorientation());
        }
        if (runtime > done) {
            check(mdx_step(e_, runtime - done, nullptr));
            host_stale_ = true;
        }
    }
};

// index_test / excludea (Analysis/AnalyzeGrowth.cpp:115-134): which pairs
// the distance graph keeps -- all, or those with both indices outside
// [N1, N2]
struct index_test {
    virtual ~index_test() {}
    virtual void excluded_range(int &lo, int &hi) const
    {
        lo = 0;
        hi = -1;
    }
};
struct excludea : public index_test {
    int N1, N2;
    excludea(int N11, int N22) : N1(N11), N2(N22) {}
    void excluded_range(int &lo, int &hi) const override
    {
        lo = N1;
        hi = N2;
    }
};

// distance_graph(pos, boxes, geo, binding_distance, i1)
// (Analysis/AnalyzeGrowth.cpp:136-177): connected components of the pairs
// within binding_distance, computed on the GPU
inline vector1<int> distance_graph(matrix<double> &pos, matrix<int> &boxes, cube &geo,
                                   double binding_distance, index_test *i1)
{
    LangevinNVT A(geo);
    A.setdat(pos);
    int lo = 0, hi = -1;
    if (i1)
        i1->excluded_range(lo, hi);
    return A.distance_graph(boxes, binding_distance, lo, hi);
}

} // namespace mtamd

#endif
