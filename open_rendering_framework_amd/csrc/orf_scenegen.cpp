// orf_scenegen.cpp -- procedural workloads (host): displaced icosphere and sky/sun env map.
//
// These generate the benchmark scenes; they have no counterpart in the
// reference (which loads OBJ and HDR files) and are defined here:
//   icosphere : icosahedron subdivided `level` times (10*4^L + 2 vertices,
//               20*4^L triangles, outward winding), each vertex at radius
//               1 + amplitude * noise(frequency * p) for its unit direction p;
//               area-weighted vertex normals.
//   noise     : 3D value noise, lattice values tea<4>(hash(i,j,k), seed) mapped
//               to [-1,1], smoothstep-weighted trilinear interpolation.
//   sky       : lat-long RGBA FP32, row j <-> theta = (1 - (j+0.5)/h) pi,
//               column i <-> phi = (2 (i+0.5)/w - 1) pi, direction
//               (sin th sin ph, cos th, sin th cos ph) (the env convention of
//               ORF/Envmap.h:57-98); sky gradient above the horizon, dim ground
//               below, plus a Gaussian sun lobe exp(-gamma^2 / 2 sigma^2) * peak.
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "orf_common.h"
#include "orf_rng.h"

namespace orf {

static double lattice(int64_t i, int64_t j, int64_t k, uint32_t seed)
{
    uint32_t h = (uint32_t)i * 0x8da6b343u ^ (uint32_t)j * 0xd8163841u ^ (uint32_t)k * 0xcb1ab31fu;
    uint32_t v = tea<4>(h, seed);
    return (double)(v >> 8) * (2.0 / 16777216.0) - 1.0;
}

static double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

double value_noise3(double x, double y, double z, uint32_t seed)
{
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int64_t i = (int64_t)fx, j = (int64_t)fy, k = (int64_t)fz;
    double u = smooth(x - fx), v = smooth(y - fy), w = smooth(z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(i + a, j + b, k + d, seed);
    double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * u;
    double x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * u;
    double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * u;
    double x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * u;
    double y0 = x00 + (x10 - x00) * v;
    double y1 = x01 + (x11 - x01) * v;
    return y0 + (y1 - y0) * w;
}

struct Icosphere {
    std::vector<double> p;  // unit directions, 3 per vertex
    std::vector<int32_t> f;
};

static Icosphere make_icosphere(int level)
{
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    const double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                                {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const int faces[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
                              {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                              {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
                              {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    Icosphere s;
    for (auto& b : base) {
        double l = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        s.p.insert(s.p.end(), {b[0] / l, b[1] / l, b[2] / l});
    }
    for (auto& fc : faces) s.f.insert(s.f.end(), {fc[0], fc[1], fc[2]});
    for (int L = 0; L < level; ++L) {
        std::unordered_map<uint64_t, int32_t> mid;
        mid.reserve(s.f.size());
        std::vector<int32_t> nf;
        nf.reserve(s.f.size() * 4);
        auto midpoint = [&](int32_t a, int32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | (uint32_t)b) : ((uint64_t)b << 32 | (uint32_t)a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            double m[3];
            for (int k = 0; k < 3; ++k) m[k] = 0.5 * (s.p[3 * a + k] + s.p[3 * b + k]);
            double l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
            int32_t id = (int32_t)(s.p.size() / 3);
            s.p.insert(s.p.end(), {m[0] / l, m[1] / l, m[2] / l});
            mid.emplace(key, id);
            return id;
        };
        for (size_t i = 0; i < s.f.size(); i += 3) {
            int32_t a = s.f[i], b = s.f[i + 1], c = s.f[i + 2];
            int32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            nf.insert(nf.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
        }
        s.f.swap(nf);
    }
    return s;
}

}  // namespace orf

extern "C" orf_status orf_gen_icosphere(int level, uint32_t noise_seed, float amplitude, float frequency,
                                        float* verts, float* normals, int64_t* nverts, int32_t* tris,
                                        int64_t* ntris)
{
    ORF_API_BEGIN
    ORF_REQUIRE(level >= 0 && level <= 10, "orf_gen_icosphere: level must be in [0, 10]");
    ORF_REQUIRE(nverts && ntris, "orf_gen_icosphere: null size pointer");
    int64_t nv = 10 * ((int64_t)1 << (2 * level)) + 2;
    int64_t nt = 20 * ((int64_t)1 << (2 * level));
    if (!verts && !tris && !normals) {
        *nverts = nv;
        *ntris = nt;
        return ORF_SUCCESS;
    }
    ORF_REQUIRE(*nverts >= nv && *ntris >= nt, "orf_gen_icosphere: output arrays too small");
    ORF_REQUIRE(verts && tris && normals, "orf_gen_icosphere: null output array");
    orf::Icosphere s = orf::make_icosphere(level);
    ORF_REQUIRE((int64_t)s.p.size() / 3 == nv && (int64_t)s.f.size() / 3 == nt,
                "orf_gen_icosphere: internal size mismatch");
    std::vector<double> pos(s.p.size());
    for (int64_t v = 0; v < nv; ++v) {
        const double* d = &s.p[3 * v];
        double r = 1.0;
        if (amplitude != 0.0f)
            r += (double)amplitude *
                 orf::value_noise3(frequency * d[0], frequency * d[1], frequency * d[2], noise_seed);
        for (int k = 0; k < 3; ++k) pos[3 * v + k] = r * d[k];
    }
    std::vector<double> acc(s.p.size(), 0.0);
    for (int64_t t = 0; t < nt; ++t) {
        int32_t a = s.f[3 * t], b = s.f[3 * t + 1], c = s.f[3 * t + 2];
        double e1[3], e2[3], n[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = pos[3 * b + k] - pos[3 * a + k];
            e2[k] = pos[3 * c + k] - pos[3 * a + k];
        }
        n[0] = e1[1] * e2[2] - e1[2] * e2[1];
        n[1] = e1[2] * e2[0] - e1[0] * e2[2];
        n[2] = e1[0] * e2[1] - e1[1] * e2[0];
        // outward winding: the face normal points away from the centre
        double ctr = 0.0;
        for (int k = 0; k < 3; ++k) ctr += n[k] * (pos[3 * a + k] + pos[3 * b + k] + pos[3 * c + k]);
        if (ctr < 0.0) {
            std::swap(b, c);
            for (double& x : n) x = -x;
        }
        tris[3 * t] = a;
        tris[3 * t + 1] = b;
        tris[3 * t + 2] = c;
        for (int32_t v : {a, b, c})
            for (int k = 0; k < 3; ++k) acc[3 * v + k] += n[k];
    }
    for (int64_t v = 0; v < nv; ++v) {
        double* n = &acc[3 * v];
        double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; ++k) {
            verts[3 * v + k] = (float)pos[3 * v + k];
            normals[3 * v + k] = (float)(n[k] / l);
        }
    }
    *nverts = nv;
    *ntris = nt;
    ORF_API_END
}

extern "C" orf_status orf_gen_sky_envmap(int w, int h, const float sun_dir[3], float sun_sigma_deg,
                                         float sun_peak, float* rgba)
{
    ORF_API_BEGIN
    ORF_REQUIRE(w > 0 && h > 0 && rgba && sun_dir, "orf_gen_sky_envmap: bad arguments");
    double s[3] = {sun_dir[0], sun_dir[1], sun_dir[2]};
    double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    ORF_REQUIRE(sl > 0.0, "orf_gen_sky_envmap: zero sun direction");
    for (double& x : s) x /= sl;
    const double pi = 3.14159265358979323846;
    const double sigma = (double)sun_sigma_deg * pi / 180.0;
    const double horizon[3] = {0.80, 0.76, 0.72}, zenith[3] = {0.12, 0.24, 0.54}, ground[3] = {0.10, 0.09, 0.08};
    for (int j = 0; j < h; ++j) {
        double th = (1.0 - (j + 0.5) / h) * pi;
        for (int i = 0; i < w; ++i) {
            double ph = (2.0 * (i + 0.5) / w - 1.0) * pi;
            double d[3] = {std::sin(th) * std::sin(ph), std::cos(th), std::sin(th) * std::cos(ph)};
            double c[3];
            if (d[1] >= 0.0) {
                double t = std::sqrt(d[1]);
                for (int k = 0; k < 3; ++k) c[k] = horizon[k] * (1.0 - t) + zenith[k] * t;
            } else {
                for (int k = 0; k < 3; ++k) c[k] = ground[k] * (1.0 + 0.5 * d[1]);
            }
            double cg = d[0] * s[0] + d[1] * s[1] + d[2] * s[2];
            double g = std::acos(std::fmax(-1.0, std::fmin(1.0, cg)));
            double sun = (double)sun_peak * std::exp(-g * g / (2.0 * sigma * sigma));
            float* o = rgba + 4 * ((int64_t)j * w + i);
            for (int k = 0; k < 3; ++k) o[k] = (float)(c[k] + sun);
            o[3] = 1.0f;
        }
    }
    ORF_API_END
}
