/* Host scene ingestion: scene JSON and the files it names -> fm_scene_desc. */
#ifndef FM_HOST_LOADER_H
#define FM_HOST_LOADER_H

#include "../../../include/fm_gpu.h"
#include "json.h"

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace fmh {

struct LoadError : std::runtime_error {
    int code;
    LoadError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct MeshData {
    std::vector<float> positions, normals, texcoords;
    std::vector<int32_t> indices;  /* 9 per triangle */
};

struct ImageData {
    int32_t width = 0, height = 0, format = FM_IMAGE_F32_LINEAR;
    std::vector<float> f32;
    std::vector<unsigned char> u8;
};

/* Owns everything an fm_scene_desc points to. */
struct LoadedScene {
    fm_scene_desc desc{};
    std::vector<std::unique_ptr<ImageData>> image_data;
    std::vector<std::unique_ptr<MeshData>> mesh_data;
    std::vector<fm_image_desc> images;
    std::vector<fm_mesh_desc> meshes;
    std::vector<fm_material_desc> materials;
    std::vector<fm_object_desc> objects;
    std::vector<fm_light_desc> lights;
    std::vector<fm_post_op> post_ops;
    void finish();
};

/* Throws LoadError. */
void load_scene_json(const std::string& path, LoadedScene& out);

}  // namespace fmh
#endif
