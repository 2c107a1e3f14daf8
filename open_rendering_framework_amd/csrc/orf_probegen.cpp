// orf_probegen.cpp -- surface probes of a translucent object and its scattering
// coefficients (host).
//
// Probes: a Poisson-disk point set on the object's triangles by dart throwing
// (Cook 1986): candidates are drawn by area (triangle by a cumulative area table,
// then a uniform point with sqrt(xi1) barycentrics as CU/sample_camera.cu:75-83
// draws them) from PCG32(seed, 11); a candidate is kept when no kept probe lies
// within r_min = 0.7 sqrt(A / N) (checked through a hashed grid of cell r_min);
// drawing stops at N probes or 64 N candidates.  The normal is the interpolated
// shading normal.  The probes are then sorted by the slot of their cell in the
// device's hashed grid (cell = 2.5 sqrt(A / N)) so a slot is one contiguous run.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "orf_common.h"
#include "orf_probes_host.h"
#include "orf_sss.h"

namespace orf {

namespace {

struct Pcg {
    uint64_t state = 0, inc = 0;
    Pcg(uint64_t initstate, uint64_t initseq)
    {
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    uint32_t next()
    {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    }
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

unsigned pow2_at_least(size_t n)
{
    unsigned p = 1;
    while (p < n) p <<= 1;
    return p;
}

}  // namespace

ProbeSetHost generate_probes(const float* V, const int32_t* I, const int4* shade, const float4* normals,
                             int64_t ntris, int material, int count, uint32_t seed)
{
    ProbeSetHost ps;
    std::vector<int64_t> tris;
    std::vector<double> cdf;
    double area = 0.0;
    auto vert = [&](int32_t k) { return mk3(V[3 * k], V[3 * k + 1], V[3 * k + 2]); };
    for (int64_t t = 0; t < ntris; ++t) {
        if (shade[t].w != material) continue;
        f3 a = vert(I[3 * t]), b = vert(I[3 * t + 1]), c = vert(I[3 * t + 2]);
        f3 x = cross(b - a, c - a);
        double ar = 0.5 * std::sqrt((double)x.x * x.x + (double)x.y * x.y + (double)x.z * x.z);
        if (!(ar > 0.0)) continue;
        area += ar;
        tris.push_back(t);
        cdf.push_back(area);
    }
    ps.area = area;
    if (tris.empty() || count <= 0) return ps;
    double spacing = std::sqrt(area / count);
    float rmin = (float)(0.7 * spacing);
    float rmin2 = rmin * rmin;
    unsigned dart_slots = pow2_at_least((size_t)count * 2);
    std::vector<int> head(dart_slots, -1), next;
    next.reserve(count);
    float inv_rmin = 1.0f / rmin;
    Pcg rng(seed, 11);
    std::vector<f3> pos, nrm;
    pos.reserve(count);
    nrm.reserve(count);
    int64_t max_tries = (int64_t)count * 64;
    for (int64_t tries = 0; tries < max_tries && (int)pos.size() < count; ++tries) {
        double target = (double)rng.uniform() * area;
        size_t k = (size_t)(std::upper_bound(cdf.begin(), cdf.end(), target) - cdf.begin());
        if (k >= tris.size()) k = tris.size() - 1;
        int64_t t = tris[k];
        float xi1 = std::sqrt(rng.uniform()), xi2 = rng.uniform();
        float u = 1.0f - xi1, v = (1.0f - xi2) * xi1, w = xi1 * xi2;
        f3 a = vert(I[3 * t]), b = vert(I[3 * t + 1]), c = vert(I[3 * t + 2]);
        f3 p = a * u + b * v + c * w;
        int cx = probe_cell(p.x, inv_rmin), cy = probe_cell(p.y, inv_rmin), cz = probe_cell(p.z, inv_rmin);
        bool ok = true;
        for (int dz = -1; dz <= 1 && ok; ++dz)
            for (int dy = -1; dy <= 1 && ok; ++dy)
                for (int dx = -1; dx <= 1 && ok; ++dx)
                    for (int j = head[probe_hash(cx + dx, cy + dy, cz + dz, dart_slots - 1)]; j >= 0; j = next[j]) {
                        f3 e = pos[j] - p;
                        if (dot(e, e) < rmin2) {
                            ok = false;
                            break;
                        }
                    }
        if (!ok) continue;
        f3 n;
        int4 s = shade[t];
        if (s.x >= 0 && s.y >= 0 && s.z >= 0) {
            float4 n0 = normals[s.x], n1 = normals[s.y], n2 = normals[s.z];
            n = normalize(mk3(n0.x, n0.y, n0.z) * u + mk3(n1.x, n1.y, n1.z) * v + mk3(n2.x, n2.y, n2.z) * w);
        } else {
            n = normalize(cross(b - a, c - a));
        }
        unsigned slot = probe_hash(cx, cy, cz, dart_slots - 1);
        next.push_back(head[slot]);
        head[slot] = (int)pos.size();
        pos.push_back(p);
        nrm.push_back(n);
    }
    // hashed grid of the device lookup
    ps.cell = (float)(2.5 * spacing);
    float inv_cell = 1.0f / ps.cell;
    unsigned slots = pow2_at_least(pos.size() * 2);
    ps.table_mask = slots - 1;
    std::vector<unsigned> key(pos.size());
    for (size_t i = 0; i < pos.size(); ++i)
        key[i] = probe_hash(probe_cell(pos[i].x, inv_cell), probe_cell(pos[i].y, inv_cell),
                            probe_cell(pos[i].z, inv_cell), ps.table_mask);
    std::vector<int> order(pos.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
    ps.cells.assign(slots, make_uint2(0u, 0u));
    ps.pos.resize(pos.size());
    ps.nrm.resize(pos.size());
    for (size_t i = 0; i < order.size(); ++i) {
        int j = order[i];
        ps.pos[i] = make_float4(pos[j].x, pos[j].y, pos[j].z, 0.0f);
        ps.nrm[i] = make_float4(nrm[j].x, nrm[j].y, nrm[j].z, 0.0f);
        uint2& c = ps.cells[key[j]];
        if (c.y == 0) c.x = (unsigned)i;
        c.y++;
    }
    ps.rmin = rmin;
    return ps;
}

// RoughTranslucentMaterial::computeCoefficients (ORF/RoughTranslucentMaterial.cpp:366-395)
// with an outside ior of 1.
ScatteringGPU scattering_coefficients(const SssParamsHost& h)
{
    if (h.has_user) {  // ScatteringMaterialProperties given whole (setUserData)
        ScatteringGPU u = h.user;
        u.dipole_model = h.dipole_model;
        return u;
    }
    return sss_coefficients(h.absorption * h.scale, h.scattering * h.scale, h.meancosine, h.ior, h.dipole_model);
}

}  // namespace orf
