// Host build of the device nearest-probe search (orf_sss.h, nearest_probe)
// checked against its definition by brute force: the smallest (d^2, index)
// over all probes with d^2 < cell^2.  Probes on a unit sphere (2 cell sizes,
// coordinates near 0, +300 and -2000 so the cell-coordinate rounding margin is
// exercised), 50 duplicated probes (distance ties broken by index), queries on
// probes (d = 0) and near the surface, and a 16-slot table where most cells
// share a slot.  Exit status 0 when every query agrees.
#include "orf_sss.h"

#include <algorithm>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>

using namespace orf;

static int brute_nearest(const std::vector<float4>& pos, float cell, f3 y)
{
    float best = cell * cell;
    int b = -1;
    for (int j = 0; j < (int)pos.size(); ++j) {
        f3 e = mk3(pos[j].x, pos[j].y, pos[j].z) - y;
        float d2 = dot(e, e);
        if (d2 < best || (d2 == best && b >= 0 && j < b)) {
            best = d2;
            b = j;
        }
    }
    return b;
}

int main()
{
    std::mt19937 g(1);
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    long bad = 0, total = 0, found = 0;
    for (int trial = 0; trial < 6; ++trial) {
        const int n = 4000;
        const float cell = trial < 3 ? 0.05f : 0.013f;
        const float off = trial % 3 == 0 ? 0.0f : (trial % 3 == 1 ? 300.0f : -2000.0f);
        const unsigned slots = trial == 5 ? 16u : 8192u;
        std::vector<float4> p(n);
        for (auto& q : p) {
            f3 v = normalize(mk3(U(g), U(g), U(g)));
            q = make_float4(off + v.x, v.y, v.z, 0.0f);
        }
        for (int i = 0; i < 50; ++i) p[n - 1 - i] = p[i * 7];
        // the hashed grid as the probe set builds it: probes sorted by slot
        const float inv = 1.0f / cell;
        std::vector<unsigned> key(n);
        for (int i = 0; i < n; ++i)
            key[i] = probe_hash(probe_cell(p[i].x, inv), probe_cell(p[i].y, inv), probe_cell(p[i].z, inv), slots - 1);
        std::vector<int> order(n);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
        std::vector<float4> sorted(n);
        std::vector<uint2> cells(slots, make_uint2(0u, 0u));
        for (int i = 0; i < n; ++i) sorted[i] = p[order[i]];
        for (int i = 0; i < n; ++i) {
            unsigned k = key[order[i]];
            if (cells[k].y == 0) cells[k].x = (unsigned)i;
            cells[k].y++;
        }
        ProbesGPU pr{};
        pr.pos = sorted.data();
        pr.cells = cells.data();
        pr.table_mask = slots - 1;
        pr.cell = cell;
        pr.count = n;
        for (int q = 0; q < 20000; ++q) {
            f3 y;
            if (q % 4 == 0) {
                const float4& t = sorted[g() % n];
                y = mk3(t.x, t.y, t.z);
            } else {
                f3 v = normalize(mk3(U(g), U(g), U(g))) * (1.0f + 0.02f * U(g));
                y = mk3(off + v.x, v.y, v.z);
            }
            int a = nearest_probe(pr, y), b = brute_nearest(sorted, cell, y);
            ++total;
            found += a >= 0;
            if (a != b) {
                if (bad < 5) std::printf("trial %d query %d: search %d, brute force %d\n", trial, q, a, b);
                ++bad;
            }
        }
    }
    std::printf("nearest_probe: %ld of %ld queries differ from brute force (%ld found a probe)\n", bad, total, found);
    return bad != 0 || found == 0 || found == total;
}
