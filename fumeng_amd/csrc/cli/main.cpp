/*
 * fumeng_gpu <scene.json>: the CMD_EXE front end (src/cmd/src/main.cpp:13-35)
 * over the C ABI.  Loads the scene with the reference loader semantics,
 * renders it on GPU 0 (wavefront form) and writes <output>.png,
 * <output>_albedo.png and <output>_normal.png (render_target.h:26-31).
 *
 * Output step (renderer.h:124-142): the frame is rendered in ten steps of a
 * tenth of the samples; after each of the first nine the running means are
 * written as progressive snapshots, after the last the scene's post-process
 * list (tone mapping, sRGB conversions) is applied to the colour and the
 * final images are written.  The bytes come from fm_post_process
 * (RealToUInt8, image.cpp:97-110).  The reference's snapshots hold finished
 * pixels only (its thread pool fills the image pixel by pixel); these hold
 * every pixel at a tenth, two tenths, ... of the samples.  <output>.pfm keeps
 * the float colour film before post-processing.  An "oidn_denoise" entry is
 * skipped with a warning (no OIDN here).
 * Returns 0 on success and -1 on any error, like the reference.
 */
#include "../../../include/fm_gpu.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

static uint32_t crc_table[256];
static void crc_init() {
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        crc_table[n] = c;
    }
}
static uint32_t crc32(const unsigned char* b, size_t n, uint32_t c = 0xFFFFFFFFu) {
    for (size_t i = 0; i < n; i++) c = crc_table[(c ^ b[i]) & 0xFF] ^ (c >> 8);
    return c;
}
static void be32(std::vector<unsigned char>& v, uint32_t x) {
    v.push_back((unsigned char)(x >> 24)); v.push_back((unsigned char)(x >> 16));
    v.push_back((unsigned char)(x >> 8)); v.push_back((unsigned char)x);
}
static void chunk(FILE* f, const char* type, const std::vector<unsigned char>& data) {
    std::vector<unsigned char> b;
    be32(b, (uint32_t)data.size());
    b.insert(b.end(), type, type + 4);
    b.insert(b.end(), data.begin(), data.end());
    const uint32_t c = crc32(b.data() + 4, b.size() - 4) ^ 0xFFFFFFFFu;
    be32(b, c);
    fwrite(b.data(), 1, b.size(), f);
}
/* 8-bit RGB PNG with stored (uncompressed) deflate blocks */
static bool write_png(const std::string& path, int w, int h, const unsigned char* rgb8) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return false;
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    fwrite(sig, 1, 8, f);
    std::vector<unsigned char> ihdr;
    be32(ihdr, (uint32_t)w); be32(ihdr, (uint32_t)h);
    ihdr.push_back(8); ihdr.push_back(2); ihdr.push_back(0); ihdr.push_back(0); ihdr.push_back(0);
    chunk(f, "IHDR", ihdr);
    std::vector<unsigned char> raw;
    raw.reserve((size_t)h * (1 + 3 * (size_t)w));
    for (int y = 0; y < h; y++) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb8 + (size_t)y * 3 * w, rgb8 + (size_t)(y + 1) * 3 * w);
    }
    std::vector<unsigned char> z;
    z.push_back(0x78); z.push_back(0x01);
    size_t pos = 0;
    do {
        const size_t n = std::min<size_t>(65535, raw.size() - pos);
        z.push_back(pos + n == raw.size() ? 1 : 0);
        z.push_back((unsigned char)(n & 0xFF)); z.push_back((unsigned char)(n >> 8));
        z.push_back((unsigned char)(~n & 0xFF)); z.push_back((unsigned char)((~n >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + (long)pos, raw.begin() + (long)(pos + n));
        pos += n;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (unsigned char c : raw) { a = (a + c) % 65521u; b = (b + a) % 65521u; }
    be32(z, (b << 16) | a);
    chunk(f, "IDAT", z);
    chunk(f, "IEND", {});
    fclose(f);
    return true;
}
static bool write_pfm(const std::string& path, int w, int h, const float* rgb) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return false;
    fprintf(f, "PF\n%d %d\n-1.0\n", w, h);
    for (int y = h - 1; y >= 0; y--) fwrite(rgb + (size_t)y * 3 * w, sizeof(float), (size_t)3 * w, f);
    fclose(f);
    return true;
}

int main(int argc, char** argv) {
    crc_init();
    if (argc < 2) { fprintf(stderr, "Please input the scene file.\n"); return -1; }
    const std::string scene = argv[1];
    if (scene.find(".json") == std::string::npos) { fprintf(stderr, "Please specify the correct scene file.\n"); return -1; }
    /* optional, after the scene file: --bvh-builder host|device, --bvh-width 2|4
     * (traversal tree of the scene; the image does not depend on them) */
    int builder = FM_BUILDER_HOST, width = 2;
    for (int i = 2; i + 1 < argc; i += 2) {
        const std::string k = argv[i], v = argv[i + 1];
        if (k == "--bvh-builder" && (v == "host" || v == "device")) builder = v == "device" ? FM_BUILDER_DEVICE : FM_BUILDER_HOST;
        else if (k == "--bvh-width" && (v == "2" || v == "4")) width = v == "4" ? 4 : 2;
        else { fprintf(stderr, "unknown option %s %s\n", k.c_str(), v.c_str()); return -1; }
    }
    fm_ctx* ctx = nullptr;
    if (fm_init(0, &ctx) != FM_OK) { fprintf(stderr, "fm_init: %s\n", fm_last_error(nullptr)); return -1; }
    if (fm_set_bvh_builder(ctx, builder) != FM_OK || fm_set_bvh_width(ctx, width) != FM_OK) {
        fprintf(stderr, "options: %s\n", fm_last_error(ctx));
        fm_shutdown(ctx);
        return -1;
    }
    fm_scene* s = nullptr;
    if (fm_scene_load(ctx, scene.c_str(), &s) != FM_OK) { fprintf(stderr, "load: %s\n", fm_last_error(ctx)); fm_shutdown(ctx); return -1; }
    fm_scene_info info;
    fm_scene_get_info(s, &info);
    const int w = info.renderer.width, h = info.renderer.height;
    std::vector<float> color((size_t)w * h * 3), albedo(color.size()), normal(color.size()), post(color.size());
    std::vector<unsigned char> c8(color.size()), a8(color.size()), n8(color.size());
    std::vector<fm_post_op> ops(FM_MAX_POST_OPS);
    int32_t n_ops = 0;
    fm_scene_post_ops(s, ops.data(), FM_MAX_POST_OPS, &n_ops);
    ops.resize((size_t)n_ops);
    for (size_t k = 0; k < ops.size();)
        if (ops[k].type == FM_POST_OIDN_DENOISE) {
            fprintf(stderr, "warning: oidn_denoise is not available, skipped\n");
            ops.erase(ops.begin() + (long)k);
        } else {
            k++;
        }
    const std::string out = info.renderer.output;
    auto write_all = [&](const std::vector<fm_post_op>& post_ops) {
        const int64_t n = (int64_t)w * h;
        return fm_post_process(ctx, post_ops.data(), (int32_t)post_ops.size(), color.data(), n, post.data(), c8.data()) == FM_OK &&
               fm_post_process(ctx, nullptr, 0, albedo.data(), n, nullptr, a8.data()) == FM_OK &&
               fm_post_process(ctx, nullptr, 0, normal.data(), n, nullptr, n8.data()) == FM_OK &&
               write_png(out + ".png", w, h, c8.data()) && write_png(out + "_albedo.png", w, h, a8.data()) &&
               write_png(out + "_normal.png", w, h, n8.data());
    };
    fm_render_params p;
    memset(&p, 0, sizeof p);
    p.rr_depth = -1;
    p.mode = FM_MODE_WAVEFRONT;
    p.seed = 1;
    fm_film film;
    memset(&film, 0, sizeof film);
    film.color = color.data(); film.albedo = albedo.data(); film.normal = normal.data();
    const int spp = info.renderer.spp > 0 ? info.renderer.spp : 1;
    double ms = 0.0;
    int64_t paths = 0, rays = 0;
    int done = 0;
    bool ok = true;
    for (int step = 1; step <= 10 && ok; step++) {
        const int target = (int)((int64_t)spp * step / 10);
        if (target <= done) continue;
        p.spp = target - done;
        p.sample_offset = done;
        p.accumulate = done > 0 ? 1 : 0;
        fm_render_stats st;
        if (fm_render(ctx, s, &p, &film, &st) != FM_OK) {
            fprintf(stderr, "render: %s\n", fm_last_error(ctx));
            fm_scene_destroy(s);
            fm_shutdown(ctx);
            return -1;
        }
        ms += st.kernel_ms; paths += st.paths; rays += st.rays;
        done = target;
        printf("Current progress is : %.2f\n", (double)done / spp);
        ok = write_all(done < spp ? std::vector<fm_post_op>() : ops);
    }
    ok = ok && write_pfm(out + ".pfm", w, h, color.data());
    if (!ok) fprintf(stderr, "output: %s\n", fm_last_error(ctx));
    printf("rendered %dx%d at %d spp in %.1f ms (%.2f Mpaths/s, %.2f Mrays/s)\n", w, h, spp, ms,
           paths / (ms * 1e3), rays / (ms * 1e3));
    fm_scene_destroy(s);
    fm_shutdown(ctx);
    return ok ? 0 : -1;
}
