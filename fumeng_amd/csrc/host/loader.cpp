/*
 * Scene JSON -> fm_scene_desc with the reference SceneLoader's semantics.
 *
 * Follows (paths relative to the reference's src/engine/):
 *   include/engine/create/scene_loader.h:42-302  parse order, defaults, errors
 *   include/engine/core/config.h:9-102            LoadValue: a missing key keeps
 *        the default; a value of the wrong JSON type logs and keeps it; vec3
 *        takes 1 (broadcast) or >= 3 numbers
 *   src/core/config.cpp:6-40                      TextureDesc constant / image
 *   include/engine/core/{camera,renderer,material,geometry,texture}.h  defaults
 *   src/geometry/triangle.cpp:104-110,312-329     OBJ faces in file order
 *   src/image/image.cpp:127-148                   8-bit vs .hdr images
 * File decoders are this project's own: OBJ (v/vt/vn/f), Radiance RGBE .hdr
 * (flat and RLE scanlines), TGA (types 2/10, 24/32 bit) and binary PPM.
 */
#include "loader.h"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

namespace fmh {

static void log_warn(const std::string& m) { fprintf(stderr, "[fumeng] warning: %s\n", m.c_str()); }

/* ---- LoadValue semantics (config.h:35-102) -------------------------------- */
static void load_int(const Json& j, const char* k, int32_t& v) {
    const Json* x = j.find(k);
    if (!x) return;
    if (x->is_number()) v = x->type == Json::INT ? (int32_t)x->i : (int32_t)x->d;
    else log_warn(std::string("Failed when parsing ") + k);
}
static void load_real(const Json& j, const char* k, float& v) {
    const Json* x = j.find(k);
    if (!x) return;
    if (x->is_number()) v = (float)x->num();
    else log_warn(std::string("Failed when parsing ") + k);
}
static void load_bool(const Json& j, const char* k, int32_t& v) {
    const Json* x = j.find(k);
    if (!x) return;
    if (x->type == Json::BOOL) v = x->b ? 1 : 0;
    else log_warn(std::string("Failed when parsing ") + k);
}
static void load_str(const Json& j, const char* k, std::string& v) {
    const Json* x = j.find(k);
    if (!x) return;
    if (x->type == Json::STRING) v = x->s;
    else log_warn(std::string("Failed when parsing ") + k);
}
static void load_vec3(const Json& j, const char* k, float v[3]) {
    const Json* x = j.find(k);
    if (!x) return;
    if (x->type != Json::ARRAY) { log_warn(std::string("Failed when parsing ") + k); return; }
    for (const Json& e : x->arr) if (!e.is_number()) { log_warn(std::string("Failed when parsing ") + k); return; }
    const size_t n = x->arr.size();
    if (n == 0) return;
    if (n < 3) { v[0] = v[1] = v[2] = (float)x->arr[0].num(); return; }
    for (int c = 0; c < 3; c++) v[c] = (float)x->arr[c].num();
}

static std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw LoadError(FM_ERR_IO, "cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

/* ---- OBJ ------------------------------------------------------------------ */
static inline const char* skip_ws(const char* p, const char* e) { while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) p++; return p; }

static void load_obj(const std::string& path, MeshData& m) {
    std::string text = read_file(path);
    const char* p = text.data();
    const char* end = p + text.size();
    std::vector<int32_t> face;
    auto fix = [](long idx, size_t n) -> int32_t {
        if (idx > 0) return (int32_t)(idx - 1);
        if (idx < 0) return (int32_t)((long)n + idx);
        return -1;
    };
    while (p < end) {
        const char* le = (const char*)memchr(p, '\n', (size_t)(end - p));
        if (!le) le = end;
        const char* q = skip_ws(p, le);
        if (q + 1 < le && q[0] == 'v' && (q[1] == ' ' || q[1] == '\t')) {
            char* e2; q += 2;
            float a = strtof(q, &e2); float b = strtof(e2, &e2); float c = strtof(e2, &e2);
            m.positions.push_back(a); m.positions.push_back(b); m.positions.push_back(c);
        } else if (q + 2 < le && q[0] == 'v' && q[1] == 'n' && (q[2] == ' ' || q[2] == '\t')) {
            char* e2; q += 3;
            float a = strtof(q, &e2); float b = strtof(e2, &e2); float c = strtof(e2, &e2);
            m.normals.push_back(a); m.normals.push_back(b); m.normals.push_back(c);
        } else if (q + 2 < le && q[0] == 'v' && q[1] == 't' && (q[2] == ' ' || q[2] == '\t')) {
            char* e2; q += 3;
            float a = strtof(q, &e2); float b = strtof(e2, &e2);
            m.texcoords.push_back(a); m.texcoords.push_back(b);
        } else if (q + 1 < le && q[0] == 'f' && (q[1] == ' ' || q[1] == '\t')) {
            q += 2;
            face.clear();
            const size_t nv = m.positions.size() / 3, nn = m.normals.size() / 3, nt = m.texcoords.size() / 2;
            for (;;) {
                q = skip_ws(q, le);
                if (q >= le) break;
                char* e2;
                long vi = strtol(q, &e2, 10);
                if (e2 == q) break;
                q = e2;
                long ti = 0, ni = 0;
                if (q < le && *q == '/') {
                    q++;
                    if (q < le && *q != '/') { ti = strtol(q, &e2, 10); q = e2; }
                    if (q < le && *q == '/') { q++; ni = strtol(q, &e2, 10); q = e2; }
                }
                face.push_back(fix(vi, nv));
                face.push_back(ni ? fix(ni, nn) : -1);
                face.push_back(ti ? fix(ti, nt) : -1);
                while (q < le && *q != ' ' && *q != '\t' && *q != '\r') q++;
            }
            const size_t corners = face.size() / 3;
            /* triangles pass through; larger polygons are fanned from corner 0 */
            for (size_t k = 1; k + 1 < corners; k++) {
                const size_t cs[3] = {0, k, k + 1};
                for (size_t c : cs) {
                    m.indices.push_back(face[3 * c]);      /* v */
                    m.indices.push_back(face[3 * c + 1]);  /* vn */
                    m.indices.push_back(face[3 * c + 2]);  /* vt */
                }
            }
        }
        p = le + 1;
    }
    const int64_t nv = (int64_t)m.positions.size() / 3, nn = (int64_t)m.normals.size() / 3, nt = (int64_t)m.texcoords.size() / 2;
    for (size_t k = 0; k < m.indices.size(); k += 3) {
        if (m.indices[k] < 0 || m.indices[k] >= nv) throw LoadError(FM_ERR_PARSE, "OBJ vertex index out of range in " + path);
        if (m.indices[k + 1] >= nn) throw LoadError(FM_ERR_PARSE, "OBJ normal index out of range in " + path);
        if (m.indices[k + 2] >= nt) throw LoadError(FM_ERR_PARSE, "OBJ texcoord index out of range in " + path);
    }
}

/* ---- images ---------------------------------------------------------------- */
/* Radiance RGBE: mantissa * 2^(e - 136), e == 0 -> black. */
static void load_hdr(const std::string& path, ImageData& im) {
    std::string t = read_file(path);
    size_t p = 0;
    auto line = [&]() {
        size_t e = t.find('\n', p);
        if (e == std::string::npos) throw LoadError(FM_ERR_PARSE, "truncated HDR header in " + path);
        std::string l = t.substr(p, e - p);
        p = e + 1;
        return l;
    };
    std::string first = line();
    if (first.rfind("#?RADIANCE", 0) != 0 && first.rfind("#?RGBE", 0) != 0) throw LoadError(FM_ERR_PARSE, "not a Radiance HDR: " + path);
    for (;;) {
        std::string l = line();
        if (l.empty()) break;
        if (l.rfind("FORMAT=", 0) == 0 && l != "FORMAT=32-bit_rle_rgbe") throw LoadError(FM_ERR_UNSUPPORTED, "HDR format " + l);
    }
    std::string res = line();
    int H = 0, W = 0;
    if (sscanf(res.c_str(), "-Y %d +X %d", &H, &W) != 2 || W <= 0 || H <= 0) throw LoadError(FM_ERR_UNSUPPORTED, "HDR orientation " + res);
    im.width = W; im.height = H; im.format = FM_IMAGE_F32_LINEAR;
    im.f32.assign((size_t)W * H * 3, 0.0f);
    std::vector<unsigned char> scan((size_t)W * 4);
    const unsigned char* d = (const unsigned char*)t.data();
    const size_t n = t.size();
    for (int y = 0; y < H; y++) {
        if (W >= 8 && W < 32768 && p + 4 <= n && d[p] == 2 && d[p + 1] == 2 && ((d[p + 2] << 8) | d[p + 3]) == W && !(d[p + 2] & 0x80)) {
            p += 4;
            for (int c = 0; c < 4; c++) {
                int x = 0;
                while (x < W) {
                    if (p >= n) throw LoadError(FM_ERR_PARSE, "truncated HDR data");
                    int cnt = d[p++];
                    if (cnt > 128) {
                        cnt -= 128;
                        if (p >= n || x + cnt > W) throw LoadError(FM_ERR_PARSE, "bad HDR run");
                        unsigned char v = d[p++];
                        for (int k = 0; k < cnt; k++) scan[(size_t)(x++) * 4 + c] = v;
                    } else {
                        if (cnt == 0 || p + cnt > n || x + cnt > W) throw LoadError(FM_ERR_PARSE, "bad HDR dump");
                        for (int k = 0; k < cnt; k++) scan[(size_t)(x++) * 4 + c] = d[p++];
                    }
                }
            }
        } else {
            if (p + (size_t)W * 4 > n) throw LoadError(FM_ERR_PARSE, "truncated flat HDR data");
            memcpy(scan.data(), d + p, (size_t)W * 4);
            p += (size_t)W * 4;
        }
        for (int x = 0; x < W; x++) {
            const unsigned char* e = &scan[(size_t)x * 4];
            float* o = &im.f32[((size_t)y * W + x) * 3];
            if (e[3] != 0) {
                const float f = (float)ldexp(1.0f, (int)e[3] - (128 + 8));
                o[0] = e[0] * f; o[1] = e[1] * f; o[2] = e[2] * f;
            }
        }
    }
}

static void load_tga(const std::string& path, ImageData& im) {
    std::string t = read_file(path);
    const unsigned char* d = (const unsigned char*)t.data();
    if (t.size() < 18) throw LoadError(FM_ERR_PARSE, "truncated TGA " + path);
    const int idlen = d[0], cmap = d[1], type = d[2], bpp = d[16], desc = d[17];
    const int W = d[12] | (d[13] << 8), H = d[14] | (d[15] << 8);
    if (cmap != 0 || (type != 2 && type != 10) || (bpp != 24 && bpp != 32) || W <= 0 || H <= 0)
        throw LoadError(FM_ERR_UNSUPPORTED, "TGA variant in " + path);
    const int bytes = bpp / 8;
    size_t p = 18 + (size_t)idlen;
    std::vector<unsigned char> px((size_t)W * H * 3);
    size_t i = 0;
    const size_t total = (size_t)W * H;
    auto put = [&](const unsigned char* s) { px[3 * i] = s[2]; px[3 * i + 1] = s[1]; px[3 * i + 2] = s[0]; i++; };
    while (i < total) {
        if (type == 2) {
            if (p + (size_t)bytes > t.size()) throw LoadError(FM_ERR_PARSE, "truncated TGA data");
            put(d + p); p += bytes;
        } else {
            if (p >= t.size()) throw LoadError(FM_ERR_PARSE, "truncated TGA data");
            const int h = d[p++];
            const int cnt = (h & 0x7F) + 1;
            if (h & 0x80) {
                if (p + (size_t)bytes > t.size()) throw LoadError(FM_ERR_PARSE, "truncated TGA run");
                for (int k = 0; k < cnt && i < total; k++) put(d + p);
                p += bytes;
            } else {
                for (int k = 0; k < cnt && i < total; k++) {
                    if (p + (size_t)bytes > t.size()) throw LoadError(FM_ERR_PARSE, "truncated TGA packet");
                    put(d + p); p += bytes;
                }
            }
        }
    }
    im.width = W; im.height = H; im.format = FM_IMAGE_U8_SRGB;
    im.u8.resize(px.size());
    const bool top_down = (desc & 0x20) != 0;
    for (int y = 0; y < H; y++) {
        const int src = top_down ? y : (H - 1 - y);
        memcpy(&im.u8[(size_t)y * W * 3], &px[(size_t)src * W * 3], (size_t)W * 3);
    }
}

static void load_ppm(const std::string& path, ImageData& im) {
    std::string t = read_file(path);
    size_t p = 0;
    auto tok = [&]() {
        for (;;) {
            while (p < t.size() && isspace((unsigned char)t[p])) p++;
            if (p < t.size() && t[p] == '#') { while (p < t.size() && t[p] != '\n') p++; continue; }
            break;
        }
        size_t s = p;
        while (p < t.size() && !isspace((unsigned char)t[p])) p++;
        return t.substr(s, p - s);
    };
    if (tok() != "P6") throw LoadError(FM_ERR_UNSUPPORTED, "PPM variant in " + path);
    const int W = atoi(tok().c_str()), H = atoi(tok().c_str()), mx = atoi(tok().c_str());
    p++;
    if (W <= 0 || H <= 0 || mx != 255 || p + (size_t)W * H * 3 > t.size()) throw LoadError(FM_ERR_PARSE, "bad PPM " + path);
    im.width = W; im.height = H; im.format = FM_IMAGE_U8_SRGB;
    im.u8.assign(t.begin() + (long)p, t.begin() + (long)p + (long)W * H * 3);
}

static bool ends_with(const std::string& s, const char* suf) {
    const size_t n = strlen(suf);
    if (s.size() < n) return false;
    for (size_t k = 0; k < n; k++) if (tolower((unsigned char)s[s.size() - n + k]) != suf[k]) return false;
    return true;
}

/* Image::load_from_file: ".hdr" anywhere in the path selects the float path. */
static void load_image(const std::string& path, ImageData& im) {
    if (path.find(".hdr") != std::string::npos) load_hdr(path, im);
    else if (ends_with(path, ".tga")) load_tga(path, im);
    else if (ends_with(path, ".ppm")) load_ppm(path, im);
    else throw LoadError(FM_ERR_UNSUPPORTED, "image format of " + path + " (supported: .hdr .tga .ppm)");
}

/* ---- loader ---------------------------------------------------------------- */
struct Loader {
    LoadedScene& out;
    std::string root;
    std::unordered_map<std::string, int> image_ids;
    std::unordered_map<std::string, int> mat_ids;     /* name -> index, -2 = unsupported type */
    std::unordered_map<std::string, int> mesh_ids;

    int image(const std::string& path) {
        auto it = image_ids.find(path);
        if (it != image_ids.end()) return it->second;
        out.image_data.emplace_back(new ImageData());
        load_image(path, *out.image_data.back());
        const int id = (int)out.image_data.size() - 1;
        image_ids[path] = id;
        return id;
    }

    /* TextureDesc::Load (config.cpp:6-27) + texture configs (texture.h:11-68) */
    void texture(const Json& j, fm_texture_desc& t) {
        std::string type = "constant";
        load_str(j, "type", type);
        std::string wu = "repeat", wv = "repeat";
        load_str(j, "wrap_u", wu);
        load_str(j, "wrap_v", wv);
        /* Texture::InitWrap: an unknown name leaves the WrapClamp default */
        t.wrap_u = wu == "repeat" ? FM_WRAP_REPEAT : FM_WRAP_CLAMP;
        t.wrap_v = wv == "repeat" ? FM_WRAP_REPEAT : FM_WRAP_CLAMP;
        if (wu != "repeat" && wu != "clamp") log_warn("invalid wrap type");
        if (wv != "repeat" && wv != "clamp") log_warn("invalid wrap type");
        if (type == "constant") {
            t.type = FM_TEX_CONSTANT;
            t.value[0] = t.value[1] = t.value[2] = 0.0f;  /* ConstantTextureConfig() = black */
            load_vec3(j, "value", t.value);
        } else if (type == "image") {
            t.type = FM_TEX_IMAGE;
            std::string path, sampler = "linear";
            load_str(j, "path", path);
            load_str(j, "sampler", sampler);
            /* ImageTexture::InitSampler: unknown names keep NearestSample */
            t.sampler = sampler == "linear" ? FM_SAMPLER_LINEAR : FM_SAMPLER_NEAREST;
            if (sampler != "linear" && sampler != "nearest") log_warn("invalid sample type");
            t.image = image(root + path);
        } else {
            throw LoadError(FM_ERR_PARSE, "texture type not support.");
        }
    }
    void load_tex(const Json& j, const char* k, fm_texture_desc& t) {
        const Json* x = j.find(k);
        if (x) texture(*x, t);
    }
    static fm_texture_desc constant(float v0, float v1, float v2) {
        fm_texture_desc t{};
        t.type = FM_TEX_CONSTANT; t.sampler = FM_SAMPLER_LINEAR;
        t.wrap_u = FM_WRAP_REPEAT; t.wrap_v = FM_WRAP_REPEAT;
        t.value[0] = v0; t.value[1] = v1; t.value[2] = v2;
        return t;
    }

    void material(const Json& j) {
        std::string type, name;
        load_str(j, "type", type);
        load_str(j, "name", name);
        fm_material_desc m{};
        for (int k = 0; k < FM_D_NPARAM; k++) m.tex[k] = constant(0, 0, 0);
        int id = -2;
        if (type == "lambert_diffuse") {
            m.type = FM_MAT_LAMBERT;
            m.tex[0] = constant(1, 1, 1);
            load_tex(j, "basecolor", m.tex[0]);
            id = (int)out.materials.size();
            out.materials.push_back(m);
        } else if (type == "disney") {
            m.type = FM_MAT_DISNEY;
            m.tex[FM_D_SPECULAR] = constant(0.5f, 0.5f, 0.5f);
            m.tex[FM_D_IOR] = constant(1.5f, 1.5f, 1.5f);
            static const char* keys[FM_D_NPARAM] = {"basecolor", "metallic", "specular", "specularTint", "roughness",
                                                     "anisotropic", "sheen", "sheenTint", "clearcoat", "clearcoatGloss",
                                                     "specTrans", "specTransRoughness", "diffTrans", "flatness", "ior"};
            for (int k = 0; k < FM_D_NPARAM; k++) load_tex(j, keys[k], m.tex[k]);
            load_bool(j, "thin", m.thin);
            id = (int)out.materials.size();
            out.materials.push_back(m);
        } else if (type == "specular_transmission") {
            m.type = FM_MAT_SPECULAR_TRANSMISSION;
            for (int c = 0; c < 3; c++) m.reflect_color[c] = m.refract_color[c] = 1.0f;
            /* eta_i / eta_t have no default in the reference (uninitialised) */
            m.eta_i = 1.0f; m.eta_t = 1.0f;
            load_vec3(j, "reflect_color", m.reflect_color);
            load_vec3(j, "refract_color", m.refract_color);
            load_real(j, "eta_i", m.eta_i);
            load_real(j, "eta_t", m.eta_t);
            id = (int)out.materials.size();
            out.materials.push_back(m);
        } else {
            log_warn("Material type not supported: " + type);
        }
        mat_ids.emplace(name, id);  /* emplace: the first definition of a name wins */
    }

    int mesh(const std::string& path) {
        auto it = mesh_ids.find(path);
        if (it != mesh_ids.end()) return it->second;
        out.mesh_data.emplace_back(new MeshData());
        load_obj(path, *out.mesh_data.back());
        const int id = (int)out.mesh_data.size() - 1;
        mesh_ids[path] = id;
        return id;
    }

    void render_object(const Json& j) {
        fm_object_desc o{};
        o.radius = 1.0f;
        o.scale[0] = o.scale[1] = o.scale[2] = 1.0f;
        o.mesh = -1;
        bool has_shape = false;
        const Json* sh = j.find("shape");
        if (sh) {
            std::string type;
            load_str(*sh, "type", type);
            if (type == "triangle_mesh") {
                std::string path;
                load_str(*sh, "path", path);
                load_vec3(*sh, "translation", o.translation);
                load_vec3(*sh, "rotation", o.rotation);
                load_vec3(*sh, "scale", o.scale);
                o.shape = FM_SHAPE_TRIANGLE_MESH;
                o.mesh = mesh(root + path);
                has_shape = true;
            } else if (type == "sphere") {
                load_real(*sh, "radius", o.radius);
                load_vec3(*sh, "translation", o.translation);
                load_vec3(*sh, "rotation", o.rotation);
                load_vec3(*sh, "scale", o.scale);
                o.shape = FM_SHAPE_SPHERE;
                has_shape = true;
            } else if (type == "rectangle") {
                load_vec3(*sh, "pa", o.pa); load_vec3(*sh, "pb", o.pb);
                load_vec3(*sh, "pc", o.pc); load_vec3(*sh, "pd", o.pd);
                o.shape = FM_SHAPE_RECTANGLE;
                has_shape = true;
            } else {
                log_warn("Shape type not supported: " + type);
            }
        }
        load_vec3(j, "emittance", o.emittance);
        std::string mname = "default_material";
        load_str(j, "material", mname);
        auto it = mat_ids.find(mname);
        if (it == mat_ids.end()) {
            log_warn("Fail to find material " + mname + ", use default lambert.");
            it = mat_ids.find("default_material");
        }
        if (has_shape && it->second == -2)
            throw LoadError(FM_ERR_UNSUPPORTED, "render object uses material '" + mname + "' of an unsupported type");
        o.material = it->second;
        if (has_shape) out.objects.push_back(o);
    }

    void lights(const Json& j) {
        std::string type;
        load_str(j, "type", type);
        if (type == "env") {
            std::string path;
            load_str(j, "path", path);
            if (path.empty()) return;
            fm_light_desc l{};
            l.type = FM_LIGHT_ENV;
            l.factor = 1.0f;
            l.image = image(root + path);
            load_vec3(j, "rotation", l.rotation);
            load_real(j, "factor", l.factor);
            out.lights.push_back(l);
        } else if (type == "directional") {
            fm_light_desc l{};
            l.type = FM_LIGHT_DIRECTIONAL;
            l.direction[0] = 1.0f;  /* default `red` = (1, 0, 0) */
            load_vec3(j, "radiance", l.radiance);
            load_vec3(j, "direction", l.direction);
            out.lights.push_back(l);
        } else {
            log_warn("Light type not supported: " + type);
        }
    }

    void run(const std::string& path) {
        size_t slash = path.find_last_of('/');
        root = slash == std::string::npos ? std::string("./") : path.substr(0, slash + 1);
        std::string text = read_file(path);
        Json cfg;
        try { cfg = JsonParser(text).parse(); }
        catch (const std::exception& e) { throw LoadError(FM_ERR_PARSE, e.what()); }
        fm_scene_desc& d = out.desc;
        std::memset(&d, 0, sizeof(d));
        d.abi_version = FM_ABI_VERSION;
        /* ParseCamera (scene_loader.h:42-63) */
        const Json* c = cfg.find("camera");
        std::string ctype = "pin_hole";
        if (c) load_str(*c, "type", ctype);
        if (!c || (ctype != "pin_hole" && ctype != "thin_lens")) throw LoadError(FM_ERR_PARSE, "Camera is not specified.");
        fm_camera_desc& cam = d.camera;
        cam.type = ctype == "pin_hole" ? FM_CAMERA_PIN_HOLE : FM_CAMERA_THIN_LENS;
        cam.up[2] = 1.0f; cam.fov = 60.0f; cam.aspect = 1.0f; cam.focal_distance = 1.0f; cam.lens_radius = 0.0f;
        load_vec3(*c, "pos", cam.pos); load_vec3(*c, "look_at", cam.look_at); load_vec3(*c, "up", cam.up);
        load_real(*c, "fov", cam.fov); load_real(*c, "aspect", cam.aspect); load_real(*c, "focal_distance", cam.focal_distance);
        if (cam.type == FM_CAMERA_THIN_LENS) load_real(*c, "lens_radius", cam.lens_radius);
        /* ParseRender (scene_loader.h:65-84) */
        const Json* r = cfg.find("renderer");
        std::string rtype = "path_tracer";
        if (r) load_str(*r, "type", rtype);
        if (!r || rtype != "path_tracer") throw LoadError(FM_ERR_PARSE, "Camera is not specified.");
        fm_renderer_desc& rd = d.renderer;
        rd.width = 1; rd.height = 1; rd.spp = 10; rd.thread_count = 8;
        rd.depth = 10; rd.direct_loop = 1; rd.rr_depth = 10; rd.rr_coef = 0.85f;
        std::string output = "result";
        load_int(*r, "width", rd.width); load_int(*r, "height", rd.height); load_int(*r, "spp", rd.spp);
        load_int(*r, "thread_count", rd.thread_count); load_str(*r, "output", output);
        load_int(*r, "depth", rd.depth); load_int(*r, "direct_loop", rd.direct_loop);
        load_int(*r, "rr_depth", rd.rr_depth); load_real(*r, "rr_coef", rd.rr_coef);
        output = root + output;
        snprintf(rd.output, sizeof(rd.output), "%s", output.c_str());
        /* ParseScene (scene_loader.h:86-146) */
        const Json* s = cfg.find("scene");
        if (!s) throw LoadError(FM_ERR_PARSE, "Scene is not specified.");
        std::string acc = "bvh";
        if (const Json* a = s->find("accelerator")) load_str(*a, "type", acc);
        d.accelerator = acc == "naive" ? FM_ACCEL_NAIVE : (acc == "embree" ? FM_ACCEL_EMBREE : FM_ACCEL_BVH);
        if (const Json* ms = s->find("materials"))
            if (ms->type == Json::ARRAY) for (const Json& m : ms->arr) material(m);
        mat_ids.emplace("default_material", -1);
        if (const Json* os = s->find("render_objects"))
            if (os->type == Json::ARRAY) for (const Json& o : os->arr) render_object(o);
        if (const Json* ls = s->find("lights"))
            if (ls->type == Json::ARRAY) for (const Json& l : ls->arr) lights(l);
        /* ParsePostProcess (scene_loader.h:254-275); an unknown type is
         * skipped with a warning, like the reference */
        if (const Json* pp = cfg.find("post_process"))
            if (pp->type == Json::ARRAY)
                for (const Json& e : pp->arr) {
                    std::string t;
                    load_str(e, "type", t);
                    fm_post_op op{};
                    op.lum = 1.0f;
                    if (t == "tonemapping") { op.type = FM_POST_TONEMAPPING; load_real(e, "lum", op.lum); }
                    else if (t == "linear_to_srgb") op.type = FM_POST_LINEAR_TO_SRGB;
                    else if (t == "srgb_to_linear") op.type = FM_POST_SRGB_TO_LINEAR;
                    else if (t == "oidn_denoise") op.type = FM_POST_OIDN_DENOISE;
                    else { fprintf(stderr, "[fumeng] Post process type not supported: %s.\n", t.c_str()); continue; }
                    if (out.post_ops.size() >= FM_MAX_POST_OPS) throw LoadError(FM_ERR_PARSE, "too many post_process entries");
                    out.post_ops.push_back(op);
                }
        out.finish();
    }
};

void LoadedScene::finish() {
    images.clear();
    for (auto& im : image_data) {
        fm_image_desc x{};
        x.width = im->width; x.height = im->height; x.format = im->format;
        x.data = im->format == FM_IMAGE_U8_SRGB ? (const void*)im->u8.data() : (const void*)im->f32.data();
        images.push_back(x);
    }
    meshes.clear();
    for (auto& m : mesh_data) {
        fm_mesh_desc x{};
        x.num_positions = (int64_t)m->positions.size() / 3; x.positions = m->positions.data();
        x.num_normals = (int64_t)m->normals.size() / 3; x.normals = m->normals.data();
        x.num_texcoords = (int64_t)m->texcoords.size() / 2; x.texcoords = m->texcoords.data();
        x.num_triangles = (int64_t)m->indices.size() / 9; x.indices = m->indices.data();
        meshes.push_back(x);
    }
    desc.num_images = (int32_t)images.size(); desc.images = images.data();
    desc.num_materials = (int32_t)materials.size(); desc.materials = materials.data();
    desc.num_meshes = (int32_t)meshes.size(); desc.meshes = meshes.data();
    desc.num_objects = (int32_t)objects.size(); desc.objects = objects.data();
    desc.num_lights = (int32_t)lights.size(); desc.lights = lights.data();
    desc.num_post_ops = (int32_t)post_ops.size(); desc.post_ops = post_ops.data();
}

void load_scene_json(const std::string& path, LoadedScene& out) {
    Loader L{out, "", {}, {}, {}};
    L.run(path);
}

}  // namespace fmh
