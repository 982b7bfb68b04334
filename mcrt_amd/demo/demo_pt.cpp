// demo_pt: the reference's path-tracing demo -- loadSceneCornellBox and
// DemoSimple::run (src/examples/pt/demo_pt.cpp:83-203, :657-720) -- on the
// GPU path, written against the C++ facade include/mcrt.hpp.
//
//   demo_pt [--width W] [--height H] [--spp N] [--depth D] [--rr-start K]
//           [--rr-prob P] [--seed S] [--obj FILE] [--albedo A]
//           [--icosphere SUB] [--envmap FILE] [--out PREFIX] [--raw FILE]
//
// Scene: the Cornell box at the reference's x0.1 scale (grey 0.75 floor,
// back wall and ceiling, red and green side walls) with its 2-triangle
// ceiling light (emission 15), plus either the OBJ file (one diffuse
// material of albedo A, Scene::loadModel) or a centred icosphere of
// subdivision SUB built here; --envmap adds an environment map
// (Scene::setEnvmap, as demo_photonmap.cpp:289 does).  Writes PREFIX.hdr (linear film) and
// PREFIX.ppm (after tonemapGamma(2.2)) like the reference's test_pt.hdr /
// test_pt.ppm, and with --raw the float film (H x W x 3, row 0 = bottom).
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "mcrt.hpp"

namespace
{
using mcrt::Vec3;

struct Args
{
    int width = 256, height = 256, spp = 16, depth = 8, rr_start = 3, icosphere = 4;
    float rr_prob = 0.8f, albedo = 0.8f;
    unsigned long long seed = 1;
    std::string obj, out = "test_pt", raw, envmap;
};

bool parse(int argc, char** argv, Args& a)
{
    for (int i = 1; i < argc; i++)
    {
        const std::string k = argv[i];
        if (i + 1 >= argc) return false;
        const char* v = argv[++i];
        if (k == "--width") a.width = std::atoi(v);
        else if (k == "--height") a.height = std::atoi(v);
        else if (k == "--spp") a.spp = std::atoi(v);
        else if (k == "--depth") a.depth = std::atoi(v);
        else if (k == "--rr-start") a.rr_start = std::atoi(v);
        else if (k == "--rr-prob") a.rr_prob = (float)std::atof(v);
        else if (k == "--seed") a.seed = std::strtoull(v, nullptr, 10);
        else if (k == "--obj") a.obj = v;
        else if (k == "--albedo") a.albedo = (float)std::atof(v);
        else if (k == "--icosphere") a.icosphere = std::atoi(v);
        else if (k == "--out") a.out = v;
        else if (k == "--raw") a.raw = v;
        else if (k == "--envmap") a.envmap = v;
        else return false;
    }
    return true;
}

// One quad of the box as two triangles (p0 p1 p2, p0 p2 p3), scaled by 0.1.
void add_quad(mcrt::Scene& s, const float q[4][3], int mat)
{
    const float k = 0.1f;
    Vec3 p[4];
    for (int i = 0; i < 4; i++) p[i] = Vec3(q[i][0] * k, q[i][1] * k, q[i][2] * k);
    s.add(p[0], p[1], p[2], mat);
    s.add(p[0], p[2], p[3], mat);
}

// Cornell box of the reference demo: light first, then red, green, floor,
// back wall, ceiling.
void cornell_box(mcrt::Scene& s)
{
    const int light = s.addMaterial(mcrt::Material(Vec3(15, 15, 15), 25.0f, mcrt::LGHT));
    const int grey = s.addMaterial(mcrt::Material(Vec3(0.75f, 0.75f, 0.75f), 25.0f, mcrt::DIFF));
    const int red = s.addMaterial(mcrt::Material(Vec3(0.75f, 0.25f, 0.25f), 25.0f, mcrt::DIFF));
    const int green = s.addMaterial(mcrt::Material(Vec3(0.25f, 0.75f, 0.25f), 25.0f, mcrt::DIFF));
    const float ly = 548.79999f - 0.5f;
    const float L[4][3] = {{343, ly, 227}, {343, ly, 332}, {213, ly, 332}, {213, ly, 227}};
    const float R[4][3] = {{556, 0, 0}, {556, 0, 559.20001f}, {556, 548.79999f, 559.20001f}, {556, 548.79999f, 0}};
    const float G[4][3] = {{0, 0, 559.20001f}, {0, 0, 0}, {0, 548.79999f, 0}, {0, 548.79999f, 559.20001f}};
    const float F[4][3] = {{556, 0, 0}, {0, 0, 0}, {0, 0, 559.20001f}, {556, 0, 559.20001f}};
    const float B[4][3] = {{556, 0, 559.20001f}, {0, 0, 559.20001f}, {0, 548.79999f, 559.20001f},
                           {556, 548.79999f, 559.20001f}};
    const float C[4][3] = {{556, 548.79999f, 0}, {556, 548.79999f, 559.20001f}, {0, 548.79999f, 559.20001f},
                           {0, 548.79999f, 0}};
    add_quad(s, L, light);
    add_quad(s, R, red);
    add_quad(s, G, green);
    add_quad(s, F, grey);
    add_quad(s, B, grey);
    add_quad(s, C, grey);
}

// Centred icosphere (radius 12) with exact sphere normals at the corners.
void icosphere(mcrt::Scene& s, int sub, float albedo)
{
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<std::array<double, 3>> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                                            {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    auto unit = [](std::array<double, 3> p) {
        const double n = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        return std::array<double, 3>{p[0] / n, p[1] / n, p[2] / n};
    };
    for (auto& p : v) p = unit(p);
    std::vector<std::array<int, 3>> f = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
                                         {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                         {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
                                         {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (int it = 0; it < sub; it++)
    {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            const std::pair<int, int> key(std::min(a, b), std::max(a, b));
            auto i = mid.find(key);
            if (i != mid.end()) return i->second;
            v.push_back(unit({v[a][0] + v[b][0], v[a][1] + v[b][1], v[a][2] + v[b][2]}));
            return mid[key] = (int)v.size() - 1;
        };
        std::vector<std::array<int, 3>> nf;
        for (const auto& q : f)
        {
            const int ab = midpoint(q[0], q[1]), bc = midpoint(q[1], q[2]), ca = midpoint(q[2], q[0]);
            nf.push_back({q[0], ab, ca});
            nf.push_back({q[1], bc, ab});
            nf.push_back({q[2], ca, bc});
            nf.push_back({ab, bc, ca});
        }
        f.swap(nf);
    }
    const int mat = s.addMaterial(mcrt::Material(Vec3(albedo, albedo, albedo), 25.0f, mcrt::DIFF));
    const double c[3] = {27.8, 27.439999, 27.96}, r = 12.0;
    std::vector<float> tri(9 * f.size()), nrm(9 * f.size());
    std::vector<int32_t> m(f.size(), mat);
    for (size_t i = 0; i < f.size(); i++)
        for (int k = 0; k < 3; k++)
            for (int d = 0; d < 3; d++)
            {
                tri[9 * i + 3 * k + d] = (float)(c[d] + r * v[f[i][k]][d]);
                nrm[9 * i + 3 * k + d] = (float)v[f[i][k]][d];
            }
    s.add(tri.data(), nrm.data(), m.data(), f.size());
}
}  // namespace

int main(int argc, char** argv)
{
    Args a;
    if (!parse(argc, argv, a))
    {
        std::fprintf(stderr, "usage: %s [--width W] [--height H] [--spp N] [--depth D] [--rr-start K] "
                             "[--rr-prob P] [--seed S] [--obj FILE] [--albedo A] [--icosphere SUB] "
                             "[--envmap FILE] [--out PREFIX] [--raw FILE]\n", argv[0]);
        return 2;
    }
    try
    {
        mcrt::Context ctx(0);
        std::printf("device: %s\n", ctx.deviceName().c_str());
        mcrt::Scene scene(ctx);
        scene.setCamera(Vec3(278, 273, -800) * 0.1f, Vec3(278, 273, -799) * 0.1f, a.width, a.height, 39.3077f);
        cornell_box(scene);
        if (!a.obj.empty())
        {
            const int mat = scene.addMaterial(mcrt::Material(Vec3(a.albedo, a.albedo, a.albedo), 25.0f, mcrt::DIFF));
            const size_t n = scene.loadModel(a.obj, nullptr, mat);
            std::printf("loaded %zu faces from %s\n", n, a.obj.c_str());
        }
        else
            icosphere(scene, a.icosphere, a.albedo);
        if (!a.envmap.empty())
        {
            scene.setEnvmap(a.envmap);
            std::printf("loaded envmap %s\n", a.envmap.c_str());
        }
        const mcrt_build_stats st = scene.finalize();
        std::printf("init complete: %zu triangles, SBVH %.3f s, %lld refs\n", scene.numTrangles(),
                    st.build_seconds, (long long)st.refs);
        mcrt::RenderParams p;
        p.spp = a.spp;
        p.max_depth = a.depth;
        p.rr_start = a.rr_start;
        p.rr_prob = a.rr_prob;
        p.seed = a.seed;
        mcrt::FrameBuffer fb;
        const auto t0 = std::chrono::steady_clock::now();
        scene.render(p, fb);
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("took < %f > second (%.1f Msamples/s)\n", dt, (double)a.width * a.height * a.spp / dt / 1e6);
        if (!a.raw.empty())
        {
            FILE* f = std::fopen(a.raw.c_str(), "wb");
            if (!f || std::fwrite(fb.data(), sizeof(float), (size_t)a.width * a.height * 3, f) !=
                          (size_t)a.width * a.height * 3)
            {
                std::fprintf(stderr, "cannot write %s\n", a.raw.c_str());
                if (f) std::fclose(f);
                return 1;
            }
            std::fclose(f);
        }
        fb.dumpHDR(a.out + ".hdr");
        fb.tonemapGamma(2.2f);
        fb.dumpPPM(a.out + ".ppm");
    }
    catch (const mcrt::Error& e)
    {
        std::fprintf(stderr, "mcrt error %d: %s\n", e.code(), e.what());
        return 1;
    }
    return 0;
}
