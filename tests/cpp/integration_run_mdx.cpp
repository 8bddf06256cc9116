// The reference-side binding of INTEGRATION.md, compiled against the
// reference's own classes (LangevinNVTR, cube, HSPotential, matrix from
// Code/) and linked to libmdx.so: run_mdx hands the class's state to the
// engine, steps it there and copies it back.  A subclass stands in for the
// one-line declaration a maintainer would add to LangevinR.h (the members it
// reads are protected).  Built by tests/test_integration_compile.py where the
// reference sources are present; main() constructs the system and calls
// run_mdx, which on a host without a GPU must come back with the engine's
// error text rather than crash.
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

#include "mdx.h"

class LangevinNVTR_mdx : public LangevinNVTR {
public:
    LangevinNVTR_mdx(cube &c) : LangevinNVTR(c) {}

    // INTEGRATION.md's run_mdx (status returned instead of calling error()
    // so that a host without a GPU reports it)
    int run_mdx(int nsteps, HSPotential &wca, const std::vector<int> &type_of,
                const std::vector<int> &npatch, const std::vector<double> &dirs,
                const std::vector<double> &rows, std::string &why)
    {
        const int n = this->getN();
        const double box[3] = {geo.l, geo.l, geo.l};
        mdx_desc d = {};
        d.nreplicas = 1;
        d.n_per_replica = &n;
        d.box = box;
        d.periodic = geo.periodic;
        d.kT = &kT;
        d.dt = dt;
        d.m = m;
        d.gamma = gamma;
        d.gammar = gammar;
        d.inertia[0] = im[0];
        d.inertia[1] = im[4];
        d.inertia[2] = im[8];
        d.precision = MDX_PREC_MIXED;
        d.integrator = MDX_INTEG_ROTOR;
        d.seed = 1;
        d.device = 0;
        mdx_engine *e = nullptr;
        int rc = mdx_create(&d, &e);
        if (rc != MDX_OK) {
            why = mdx_last_error();
            return rc;
        }
        rc = mdx_upload_state(e, dat->getdata(), mom->getdata(), orient->getdata(),
                              angmom->getdata());
        if (rc == MDX_OK)
            rc = mdx_set_pair_wca(e, wca.epsilon, wca.sigma, 0.0);
        if (rc == MDX_OK)
            rc = mdx_set_patches(e, (int)npatch.size(), type_of.data(), npatch.data(), dirs.data(),
                                 rows.data(), max_distance_check);
        if (rc == MDX_OK)
            rc = mdx_prime(e, nullptr);
        if (rc == MDX_OK)
            rc = mdx_step(e, nsteps, nullptr);
        if (rc == MDX_OK)
            rc = mdx_download_state(e, dat->getdata(), mom->getdata(), orient->getdata(),
                                    angmom->getdata());
        if (rc != MDX_OK)
            why = mdx_last_error();
        mdx_destroy(e);
        return rc;
    }
};

int main()
{
    vector1<bool> pb(3, true);
    cube c(10.0, pb, 3);
    LangevinNVTR_mdx sys(c);
    const int n = 8;
    matrix<double> x(n, 3);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++)
            x(i, k) = 1.0 + 2.0 * ((i >> k) & 1) + 0.5 * k;
    sys.setkT(1.0);
    sys.setdt(0.005);
    sys.setm(1.0);
    sys.setgamma(1.0);
    sys.setgammar(1.0 / 3.0);
    sys.initialize(x); // positions, zero momenta, identity orientations
    HSPotential wca(1.0, 1.0);
    std::vector<int> type_of(n, 0), npatch{1};
    std::vector<double> dirs{1.0, 0.0, 0.0}, rows{6.0, 1.2, 0.31756};
    std::string why;
    const int rc = sys.run_mdx(5, wca, type_of, npatch, dirs, rows, why);
    std::printf("run_mdx rc %d%s%s\n", rc, why.empty() ? "" : ": ", why.c_str());
    return 0;
}
