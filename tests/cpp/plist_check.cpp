// Host check of the drop-in's patch-list types (mtcode_amd/host/mtcode_amd.hpp):
// patchint and adjacency (the reference's Rotational/LangevinR.h:20-57 and
// LangevinR.cpp:2000-2036) on a patch list read from a file of int32 rows
// (p1, p2, potn, patch1, patch2).  Writes the run starts to the output file
// as int32, checks that get_all / get_particle / get_patch return the fields
// they were given, and prints "ok".  No device call is made.
//   plist_check <rows.bin> <adjacency_out.bin>
#include "mtcode_amd.hpp"

#include <cstdio>
#include <fstream>
#include <vector>

using mtamd::patchint;

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::printf("usage: plist_check <rows.bin> <adjacency_out.bin>\n");
        return 2;
    }
    std::ifstream in(argv[1], std::ios::binary | std::ios::ate);
    const std::streamsize bytes = in.tellg();
    in.seekg(0);
    std::vector<int> v((size_t)bytes / sizeof(int));
    in.read(reinterpret_cast<char *>(v.data()), bytes);
    const size_t m = v.size() / 5;
    std::vector<patchint> pl(m);
    for (size_t k = 0; k < m; k++)
        pl[k] = {(unsigned int)v[5 * k], (unsigned int)v[5 * k + 1], (unsigned int)v[5 * k + 3],
                 (unsigned int)v[5 * k + 4], (unsigned int)v[5 * k + 2]};
    int bad = 0;
    for (size_t k = 0; k < m; k++) {
        int a1, a2, p, w1, w2, b1, b2, q, x1, x2;
        pl[k].get_all(a1, a2, p, w1, w2);
        pl[k].get_particle(b1, b2);
        pl[k].get_patch(q, x1, x2);
        const int *r = &v[5 * k];
        if (a1 != r[0] || a2 != r[1] || p != r[2] || w1 != r[3] || w2 != r[4] || b1 != r[0] ||
            b2 != r[1] || q != r[2] || x1 != r[3] || x2 != r[4])
            bad++;
        if (k > 0 && check_same_particle(pl[k], pl[k - 1]) !=
                         (r[0] == r[-5] && r[1] == r[-4]))
            bad++;
    }
    const std::vector<int> runs = mtamd::adjacency(pl);
    std::ofstream out(argv[2], std::ios::binary);
    out.write(reinterpret_cast<const char *>(runs.data()),
              (std::streamsize)(runs.size() * sizeof(int)));
    std::printf("rows %zu runs %zu mismatches %d\n", m, runs.size(), bad);
    if (bad)
        return 1;
    std::printf("ok\n");
    return 0;
}
