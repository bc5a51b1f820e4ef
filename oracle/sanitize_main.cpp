// sanitize_main.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A host-only program built with the sanitizers (oracle/build_sanitized.py):
// it runs the oracle's whole pipeline on synthetic pairs under several
// parameter sets, and the library's host Delaunay step (the worker pool, the
// split merge over threads, the namespaced Triangle copies) on point sets of
// every kind at every split depth, two meshes at once, checking each mesh
// against the oracle's serial Triangle.  Exit status 0 when every check
// passed; the sanitizers report (and abort on) what they find.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "elas_hip.h"
#include "elas_triangle.h"

extern "C" {
int eo_process(const uint8_t* L, const uint8_t* R, int W, int H, const elas_params* p, float* D1, float* D2,
               void* dump);
int eo_triangulate(const float* sp, int n, int32_t* tri, int cap);
void elas_params_default(elas_params* p);
void elas_params_declared(elas_params* p);
}

namespace {

// textured left image, right image shifted by a few planar disparities
void make_pair(int W, int H, int dmax, uint32_t seed, std::vector<uint8_t>& l, std::vector<uint8_t>& r) {
    std::mt19937 rng(seed);
    std::vector<float> n((size_t)W * H);
    for (auto& x : n) x = (float)(rng() % 256);
    l.assign((size_t)W * H, 0);
    r.assign((size_t)W * H, 0);
    for (int v = 0; v < H; v++)
        for (int u = 0; u < W; u++) {
            float s = 0;
            int c = 0;
            for (int dv = -1; dv <= 1; dv++)
                for (int du = -1; du <= 1; du++) {
                    const int uu = std::min(std::max(u + du, 0), W - 1), vv = std::min(std::max(v + dv, 0), H - 1);
                    s += n[(size_t)vv * W + uu];
                    c++;
                }
            l[(size_t)v * W + u] = (uint8_t)(s / c);
        }
    for (int v = 0; v < H; v++)
        for (int u = 0; u < W; u++) {
            const int d = std::min(dmax - 1, 4 + (u * 3 + v) / (W / 2 + 1) + (u > W / 2 ? 7 : 0));
            const int us = u + d;
            r[(size_t)v * W + u] = us < W ? l[(size_t)v * W + us] : (uint8_t)(rng() % 256);
        }
}

int check_meshes() {
    std::mt19937 rng(7);
    std::vector<std::vector<float>> sets;
    for (int k = 0; k < 6; k++) {  // lattices with holes (support points of a left and a right image)
        std::vector<float> p;
        const int cw = 30 + 17 * k, ch = 20 + 9 * k;
        for (int v = 0; v < ch; v++)
            for (int u = 0; u < cw; u++)
                if (rng() % 100 < 55) {
                    const float d = (float)(rng() % 12);
                    p.push_back((float)(u * 5) - (k % 2 ? d : 0.f));
                    p.push_back((float)(v * 5));
                }
        sets.push_back(p);
    }
    {
        std::vector<float> p;  // duplicates, shuffled
        for (int i = 0; i < 600; i++) {
            const float x = (float)(rng() % 60), y = (float)(rng() % 60);
            for (int r = 0; r < 1 + (int)(rng() % 3); r++) {
                p.push_back(x);
                p.push_back(y);
            }
        }
        sets.push_back(p);
    }
    {
        std::vector<float> p;  // non-integer coordinates
        for (int i = 0; i < 900; i++) {
            p.push_back((float)(rng() % 30000) * 0.01f);
            p.push_back((float)(rng() % 20000) * 0.01f);
        }
        sets.push_back(p);
    }
    {
        std::vector<float> p;  // collinear
        for (int i = 0; i < 50; i++) {
            p.push_back(5.f * i);
            p.push_back(0.f);
        }
        sets.push_back(p);
    }
    int bad = 0;
    for (int depth : {0, 1, 2, 3, 5, -1}) {
        elas_tri_set_split(depth);
        for (size_t a = 0; a < sets.size(); a++) {
            const size_t b = (a + 1) % sets.size();
            std::vector<int32_t> ta, tb;
            ElasTriJob jobs[2];
            jobs[0].pts = sets[a].data();
            jobs[0].n = (int)sets[a].size() / 2;
            jobs[0].out = &ta;
            jobs[1].pts = sets[b].data();
            jobs[1].n = (int)sets[b].size() / 2;
            jobs[1].out = &tb;
            elas_triangulate_jobs(jobs, 2, ELAS_TRI_INSTANCES);
            for (int j = 0; j < 2; j++) {
                const std::vector<float>& s = j ? sets[b] : sets[a];
                const std::vector<int32_t>& t = j ? tb : ta;
                const int n = (int)s.size() / 2, cap = 2 * n + 8;
                std::vector<int32_t> ref(3 * (size_t)cap);
                const int nr = eo_triangulate(s.data(), n, ref.data(), cap);
                if (nr != jobs[j].result || (nr > 0 && memcmp(ref.data(), t.data(), 12 * (size_t)nr) != 0)) {
                    fprintf(stderr, "mesh mismatch: depth %d set %zu (%d vs %d triangles)\n", depth, j ? b : a,
                            jobs[j].result, nr);
                    bad++;
                }
            }
        }
    }
    elas_tri_set_split(-1);
    return bad;
}

int check_pipeline() {
    int bad = 0;
    struct Case {
        int W, H, dmax;
        int preset;  // 0 effective, 1 effective + corners, 2 declared initialisers (post-processing on)
    } cases[] = {{160, 120, 31, 0}, {200, 90, 63, 1}, {97, 131, 15, 2}, {240, 180, 63, 2}};
    for (const Case& c : cases) {
        elas_params p;
        if (c.preset == 2) elas_params_declared(&p);
        else elas_params_default(&p);
        p.disp_max = c.dmax;
        p.add_corners = c.preset >= 1;
        std::vector<uint8_t> l, r;
        make_pair(c.W, c.H, c.dmax, 11u + (uint32_t)c.W, l, r);
        std::vector<float> D1((size_t)c.W * c.H), D2((size_t)c.W * c.H);
        const int st = eo_process(l.data(), r.data(), c.W, c.H, &p, D1.data(), D2.data(), nullptr);
        size_t valid = 0;
        for (float x : D1) valid += x >= 0;
        if (st != 0 || valid < (size_t)c.W * c.H / 4) {
            fprintf(stderr, "pipeline %dx%d preset %d: status %d, %zu valid\n", c.W, c.H, c.preset, st, valid);
            bad++;
        }
    }
    return bad;
}

}  // namespace

// defaults of the C ABI, restated for this host-only build (the library's
// own definitions live in elas_capi.cpp, which needs the HIP runtime)
extern "C" void elas_params_default(elas_params* p) {
    memset(p, 0, sizeof(*p));
    p->disp_max = 255; p->support_threshold = 0.85f; p->support_texture = 10; p->candidate_stepsize = 5;
    p->incon_window_size = 5; p->incon_threshold = 5; p->incon_min_support = 5; p->grid_size = 20;
    p->beta = 0.02f; p->gamma = 3.0f; p->sigma = 1.0f; p->sradius = 2.0f; p->lr_threshold = 2;
    p->speckle_sim_threshold = 1.0f;
}
extern "C" void elas_params_declared(elas_params* p) {
    elas_params_default(p);
    p->support_threshold = 0.95f; p->add_corners = true; p->gamma = 5.0f; p->sradius = 3.0f;
    p->speckle_size = 200; p->ipol_gap_width = 5000; p->filter_median = true;
}

int main() {
    const int bm = check_meshes();
    const int bp = check_pipeline();
    printf("sanitized run: %d mesh mismatches, %d pipeline failures\n", bm, bp);
    return bm || bp ? 1 : 0;
}
