// batch_asan_driver.cpp - the batch autotuner's host path under
// AddressSanitizer: random scenes, a batch, three autotune rounds over the
// 12 graph-replayed launch shapes that bench.py sweeps (each candidate drops,
// captures and instantiates the frame graph), frames, teardown.  Built by
// build_batch_asan.sh with the library's host code instrumented (GPU code
// is not); exits non-zero on any library error, and the sanitizer aborts on
// an invalid host access.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/lambent.h"

#define CHECK(x)                                                                   \
    do {                                                                           \
        lmb_status s_ = (x);                                                       \
        if (s_ != LMB_OK) {                                                        \
            std::printf("FAIL %s -> %d: %s\n", #x, (int)s_, lmb_last_error());     \
            std::fflush(stdout);                                                   \
            return 1;                                                              \
        }                                                                          \
    } while (0)

static void perspective(float fovy_deg, float aspect, float n, float fa, float m[16]) {
    std::memset(m, 0, 64);
    const float f = 1.0f / std::tan(fovy_deg * 3.14159265f / 180.0f / 2.0f);
    m[0] = f / aspect;
    m[5] = f;
    m[10] = (fa + n) / (n - fa);
    m[11] = -1.0f;
    m[14] = (2.0f * fa * n) / (n - fa);
}

static void look_to(const float e[3], const float d[3], float m[16]) {
    float f[3] = {d[0], d[1], d[2]};
    float l = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (float& x : f) x /= l;
    const float up[3] = {0, 1, 0};
    float s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    l = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (float& x : s) x /= l;
    float u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    float c[16] = {s[0], u[0], -f[0], 0, s[1], u[1], -f[1], 0, s[2], u[2], -f[2], 0,
                   -(e[0] * s[0] + e[1] * s[1] + e[2] * s[2]),
                   -(e[0] * u[0] + e[1] * u[1] + e[2] * u[2]),
                   (e[0] * f[0] + e[1] * f[1] + e[2] * f[2]), 1};
    std::memcpy(m, c, 64);
}

int main(int argc, char** argv) {
    const int nscenes = argc > 1 ? std::atoi(argv[1]) : 16;
    std::setvbuf(stdout, nullptr, _IOLBF, 0);
    std::mt19937 rng(4);
    lmb_textures tex;
    CHECK(lmb_textures_create(&tex));
    const uint8_t white[4] = {255, 255, 255, 255};
    uint32_t idx = 0;
    CHECK(lmb_textures_add(tex, LMB_TEXTURE_DIFFUSE, 1, 1, white, &idx));
    std::vector<uint8_t> sky(6 * 4, 255);
    CHECK(lmb_textures_set_background(tex, 1, sky.data()));
    lmb_material mats[2];
    const float uv[2] = {0.5f, 0.5f};
    const uint16_t rc[2] = {0, 0};
    uint32_t scale = 0;
    CHECK(lmb_emission_scale_pack(0.0f, 1.0f, &scale));
    CHECK(lmb_material_new(uv, uv, uv, rc, 0, LMB_NONE, LMB_NONE, scale, nullptr, 0, &mats[0]));
    CHECK(lmb_material_new(uv, uv, uv, rc, 0, LMB_NONE, LMB_NONE, scale, nullptr, 1, &mats[1]));
    std::vector<lmb_scene> scenes;
    std::vector<lmb_batch_item> items;
    std::uniform_real_distribution<float> U(-10.f, 10.f), L(0.05f, 0.3f), D(-1.f, 1.f);
    std::uniform_real_distribution<double> logn(std::log(1000.0), std::log(100000.0));
    const uint32_t res[4] = {128, 256, 384, 512};
    for (int i = 0; i < nscenes; ++i) {
        const uint64_t n = (uint64_t)std::exp(logn(rng));
        std::vector<float> v(12 * n);
        std::vector<uint32_t> m(n);
        for (uint64_t t = 0; t < n; ++t) {
            const float c[3] = {U(rng), U(rng), U(rng)};
            for (int k = 0; k < 3; ++k) {
                for (int a = 0; a < 3; ++a) v[12 * t + 4 * k + a] = c[a] + L(rng) * D(rng);
                v[12 * t + 4 * k + 3] = 0.f;
            }
            m[t] = (uint32_t)(t % 3 == 0);
        }
        lmb_scene_desc d{};
        d.vertices = v.data();
        d.num_vertices = 3 * n;
        d.material_idx = m.data();
        d.materials = mats;
        d.num_materials = 2;
        d.textures = tex;
        lmb_scene s;
        CHECK(lmb_scene_create(&d, &s));
        scenes.push_back(s);
        lmb_batch_item it{};
        it.scene = s;
        it.width = it.height = res[rng() % 4];
        float proj[16], view[16];
        perspective(45.f, 1.f, 0.1f, 200.f, proj);
        const float eye[3] = {0, 0, -25}, dir[3] = {0, 0, 1};
        look_to(eye, dir, view);
        CHECK(lmb_camera_from_proj_view(proj, view, &it.camera));
        items.push_back(it);
        std::printf("scene %d: %llu triangles, %u px\n", i, (unsigned long long)n, it.width);
    }
    lmb_tracer_opts o;
    lmb_tracer_opts_default(&o);
    o.samples = 4;
    o.bounces = 6;
    o.rr_start = 6;
    o.guiding = LMB_GUIDING_OFF;
    o.restir = LMB_RESTIR_DISABLED;
    lmb_batch_config cfg;
    lmb_batch_config_default(&cfg);
    lmb_batch b;
    CHECK(lmb_batch_create(items.data(), (uint32_t)items.size(), &o, &cfg, &b));
    std::printf("batch created\n");
    std::vector<lmb_batch_config> sweep;
    const uint32_t tiles[2][2] = {{8, 8}, {64, 1}};
    for (auto& t : tiles)
        for (uint32_t k : {4u, 8u, 16u})
            for (uint32_t pp : {0u, 1u << 22}) {
                lmb_batch_config c = cfg;
                c.tile_w = t[0];
                c.tile_h = t[1];
                c.blocks_per_cu = k;
                c.refill = 16;
                c.pass_pixels = pp;
                c.use_graph = 1;
                sweep.push_back(c);
            }
    for (int round = 0; round < 3; ++round) {
        lmb_batch_config best;
        std::vector<double> ms(sweep.size());
        CHECK(lmb_batch_autotune(b, sweep.data(), (uint32_t)sweep.size(), 1, &best, ms.data()));
        std::printf("autotune %d: best tile %ux%u blocks %u pass %u\n", round, best.tile_w,
                    best.tile_h, best.blocks_per_cu, best.pass_pixels);
        for (int f = 0; f < 3; ++f) CHECK(lmb_batch_trace(b, 1 + f, nullptr));
        CHECK(lmb_sync(nullptr));
    }
    std::vector<float> img(512 * 512 * 4);
    CHECK(lmb_batch_read_output(b, 0, img.data()));
    CHECK(lmb_batch_destroy(b));
    for (lmb_scene s : scenes) CHECK(lmb_scene_destroy(s));
    CHECK(lmb_textures_destroy(tex));
    std::printf("driver done\n");
    return 0;
}
