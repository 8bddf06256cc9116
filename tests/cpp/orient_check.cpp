// Host-side check of the mixed engine's orientation store (quat_to_mat /
// mat_to_quat in mtcode_amd/csrc/mdx_device.h) on random rotations, all four
// branches of Shepperd's method included:
//   - matrix -> quaternion -> matrix reproduces the matrix (FP64: 1e-14,
//     FP32: 2e-6) and the quaternion has unit norm;
//   - an FP32 quaternion written out as an FP64 matrix and read back in
//     (FP64 Shepperd, rounded once: the download / upload round trip of a
//     resumed run) is the same quaternion up to sign within 1e-6 (its norm
//     differs from 1 by FP32 rounding, which the matrix does not keep), and
//     the FP32 matrix of q and of -q is the same bit for bit.
// Prints a summary and "ok"; exit status 1 on a failure.
#include "mdx_device.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

using mdx::mat_to_quat;
using mdx::quat_to_mat;

static void random_rotation(std::mt19937_64 &g, double m[9])
{
    // from a random unit quaternion (uniform on SO(3))
    std::normal_distribution<double> n(0.0, 1.0);
    double q[4], s = 0;
    for (double &v : q) {
        v = n(g);
        s += v * v;
    }
    s = 1.0 / std::sqrt(s);
    quat_to_mat<double>(q[0] * s, q[1] * s, q[2] * s, q[3] * s, m);
}

int main()
{
    std::mt19937_64 g(7);
    int fails = 0, branch[4] = {0, 0, 0, 0};
    double e64 = 0, e32 = 0;
    for (int t = 0; t < 200000; t++) {
        double m[9], q[4], m2[9];
        random_rotation(g, m);
        const double tr = m[0] + m[4] + m[8];
        branch[tr > 0 ? 0 : (m[0] > m[4] && m[0] > m[8] ? 1 : (m[4] > m[8] ? 2 : 3))]++;
        mat_to_quat<double>(m, q);
        quat_to_mat<double>(q[0], q[1], q[2], q[3], m2);
        for (int k = 0; k < 9; k++)
            e64 = std::fmax(e64, std::fabs(m2[k] - m[k]));
        float mf[9], qf[4], mf2[9];
        for (int k = 0; k < 9; k++)
            mf[k] = (float)m[k];
        mat_to_quat<float>(mf, qf);
        quat_to_mat<float>(qf[0], qf[1], qf[2], qf[3], mf2);
        for (int k = 0; k < 9; k++)
            e32 = std::fmax(e32, std::fabs((double)mf2[k] - m[k]));
        const double nrm = (double)qf[0] * qf[0] + (double)qf[1] * qf[1] + (double)qf[2] * qf[2] +
                           (double)qf[3] * qf[3];
        if (std::fabs(nrm - 1) > 1e-6)
            fails++;
        // round trip of a stored FP32 quaternion through an FP64 matrix
        double md[9], qd[4];
        quat_to_mat<double>(qf[0], qf[1], qf[2], qf[3], md);
        mat_to_quat<double>(md, qd);
        float back[4] = {(float)qd[0], (float)qd[1], (float)qd[2], (float)qd[3]};
        const double sg = (double)back[0] * qf[0] + (double)back[1] * qf[1] + (double)back[2] * qf[2] +
                          (double)back[3] * qf[3] < 0 ? -1.0 : 1.0;
        double dq = 0;
        for (int k = 0; k < 4; k++)
            dq = std::fmax(dq, std::fabs(sg * back[k] - qf[k]));
        float ma[9], mb[9];
        quat_to_mat<float>(qf[0], qf[1], qf[2], qf[3], ma);
        quat_to_mat<float>(-qf[0], -qf[1], -qf[2], -qf[3], mb);
        if (dq > 1e-6 || std::memcmp(ma, mb, sizeof ma)) {
            if (fails < 5)
                std::printf("round trip differs: %.9g %.9g %.9g %.9g -> %.9g %.9g %.9g %.9g\n", qf[0],
                            qf[1], qf[2], qf[3], back[0], back[1], back[2], back[3]);
            fails++;
        }
    }
    std::printf("branches %d %d %d %d  max |dm| fp64 %.3g fp32 %.3g\n", branch[0], branch[1],
                branch[2], branch[3], e64, e32);
    if (e64 > 1e-14 || e32 > 2e-6)
        fails++;
    std::printf(fails ? "FAILED %d\n" : "ok\n", fails);
    return fails ? 1 : 0;
}
