// Host check of the cell-record layout (voreen_amd/csrc/cell_layout.h), the
// header the march kernels and the record builder include:
//   - cell_index maps the cells of a volume to distinct records below
//     cell_record_count, and cell_of_record inverts it (the builder writes
//     record i with the voxels of cell_of_record(i));
//   - records the index never reaches are padding cells beyond the volume;
//   - cell_neighbours(own, i) equals cell_index(lo -+ e_i) for every cell
//     with both neighbours inside the volume, in 32- and 64-bit offsets.
// Volumes of random sizes, brick-boundary crossings included.  Prints
// "ok <neighbour checks>" or the first failure.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "cell_layout.h"

using namespace hipdvr;

template <typename off_t>
static bool check_neighbours(const uint32_t lo[3], const uint32_t d[3], uint32_t nbx, uint32_t nby, long& n) {
    const off_t own = cell_index<off_t>(lo, nbx, nby);
    for (int i = 0; i < 3; ++i) {
        if (lo[i] < 1 || lo[i] + 1 >= d[i]) continue;
        off_t lower, upper;
        cell_neighbours<off_t>(own, i, (off_t)cell_brick_stride(d, i), lower, upper);
        uint32_t lm[3] = {lo[0], lo[1], lo[2]}, lp[3] = {lo[0], lo[1], lo[2]};
        lm[i] -= 1;
        lp[i] += 1;
        if (lower != cell_index<off_t>(lm, nbx, nby) || upper != cell_index<off_t>(lp, nbx, nby)) {
            printf("neighbour axis %d of %u %u %u (dims %u %u %u)\n", i, lo[0], lo[1], lo[2], d[0], d[1], d[2]);
            return false;
        }
        ++n;
    }
    return true;
}

int main(int argc, char** argv) {
    const int volumes = argc > 1 ? atoi(argv[1]) : 60;
    srand(12345);
    long n = 0;
    for (int t = 0; t < volumes; ++t) {
        uint32_t d[3];
        for (int i = 0; i < 3; ++i) d[i] = 1 + (uint32_t)(rand() % (t % 3 == 0 ? 140 : 45));
        const uint32_t nbx = cell_bricks(d[0]), nby = cell_bricks(d[1]);
        const uint64_t count = cell_record_count(d);
        std::vector<char> seen(count, 0);
        for (uint32_t z = 0; z < d[2]; ++z)
            for (uint32_t y = 0; y < d[1]; ++y)
                for (uint32_t x = 0; x < d[0]; ++x) {
                    const uint32_t lo[3] = {x, y, z};
                    const uint64_t r = cell_index<uint64_t>(lo, nbx, nby);
                    if (r >= count || seen[r] || cell_index<uint32_t>(lo, nbx, nby) != (uint32_t)r) {
                        printf("index of %u %u %u (dims %u %u %u)\n", x, y, z, d[0], d[1], d[2]);
                        return 1;
                    }
                    seen[r] = 1;
                    uint32_t c[3];
                    cell_of_record(r, d, c);
                    if (c[0] != x || c[1] != y || c[2] != z) {
                        printf("inverse of record %llu\n", (unsigned long long)r);
                        return 1;
                    }
                    if (!check_neighbours<uint32_t>(lo, d, nbx, nby, n) ||
                        !check_neighbours<uint64_t>(lo, d, nbx, nby, n))
                        return 1;
                }
        for (uint64_t r = 0; r < count; ++r) {
            if (seen[r]) continue;
            uint32_t c[3];
            cell_of_record(r, d, c);
            if (c[0] < d[0] && c[1] < d[1] && c[2] < d[2]) {
                printf("record %llu of an in-volume cell is never indexed\n", (unsigned long long)r);
                return 1;
            }
        }
    }
    printf("ok %ld\n", n);
    return 0;
}
