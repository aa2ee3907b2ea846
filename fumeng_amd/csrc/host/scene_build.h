/* Host scene construction: fm_scene_desc -> flattened device arrays. */
#ifndef FM_HOST_SCENE_BUILD_H
#define FM_HOST_SCENE_BUILD_H

#include "../../../include/fm_gpu.h"
#include "../fm_layout.h"

#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace fmh {

struct BuildError : std::runtime_error {
    int code;
    BuildError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

/* Reference binary BVH node (bvh_aggregate.cpp:14-22). */
struct RefNode {
    float mn[3], mx[3];
    bool leaf;
    int64_t a, b;
};

/* Wall time of the host build's phases. */
struct BuildReport {
    double geometry_ms = 0;    /* shapes, meshes, area lights in object order */
    double ref_bvh_ms = 0;     /* reference builder (tie order) */
    double trav_tree_ms = 0;   /* traversal tree, boxes, flattening */
    double tables_ms = 0;      /* slots, images, materials, lights, env tables */
    double total_ms = 0;
    int threads = 1;           /* host threads the build used */
};

/* Host copy of everything that goes to HBM, plus facts for fm_scene_info. */
struct HostScene {
    std::vector<RefNode> ref_nodes;
    std::vector<int32_t> order;         /* geometry index per leaf slot */
    int32_t bvh_depth = 0;          /* reference tree */
    int32_t trav_depth = 0;         /* traversal tree (binary form) */
    int32_t trav_width = 2;         /* set before the build: 2, 4 or 8 */
    int32_t builder = FM_BUILDER_HOST;  /* set before the build: FM_BUILDER_* */
    bool device_tree = false;       /* the traversal tree is left to the device builder: slots
                                       are in geometry order, leaf_boxes holds their bounds */
    std::vector<float> leaf_boxes;  /* 6 floats per slot (device builder input) */
    int64_t device_nodes = 0;       /* device-built tree: inner nodes */
    int32_t build_iterations = 0;   /* device-built tree: clustering iterations */
    double device_build_ms = 0.0;
    uint64_t trav_hash = 0;         /* of the host-built traversal nodes and slot records (0: device-built) */
    int32_t trav_stack = 1;         /* LDS stack entries the traversal needs */
    int64_t num_geoms = 0;

    std::vector<fm_dnode> nodes;        /* width 2 */
    std::vector<fm_dnode4> nodes4;      /* width 4 */
    std::vector<fm_dnode8> nodes8;      /* width 8 */
    std::vector<float> prims;            /* 12 floats per slot */
    std::vector<int32_t> meta;           /* 4 ints per slot */
    std::vector<float> tri_sh;           /* 16 floats per slot */
    std::vector<fm_drect> rects;
    std::vector<fm_dsphere> spheres;
    std::vector<fm_dmaterial> materials;
    std::vector<fm_dlight> lights;
    std::vector<fm_dlight> mesh_tris;    /* triangles of "embree" mesh lights */
    std::vector<fm_denv> envs;
    std::vector<fm_dimage> images;
    std::vector<unsigned char> texels;
    std::vector<float> env_tables;
    std::vector<double> env_aux;         /* per env: fm_denv.off_aux */
    std::vector<float> srgb_lut;
    int32_t root_ref = 0;
    float root_box[6] = {0, 0, 0, 0, 0, 0};
    int32_t env_light = -1, env = -1;
    int32_t naive = 0;
    fm_dcamera camera{};
    fm_renderer_desc renderer{};
    fm_camera_desc camera_desc{};
    std::vector<fm_post_op> post_ops;   /* colour post-process list */
    BuildReport report;
};

/* Env sampling tables of env light e from its texels (RGB per texel, linear);
 * fresh = false rewrites them in place (same image size). */
void build_env_tables(HostScene& hs, fm_denv& e, const std::function<void(int, int, double*)>& texel, bool fresh);

/* Device camera record for a camera description (per-job cameras). */
void make_camera(const fm_camera_desc& c, fm_dcamera& dc);

/* Re-collapses a host-built scene's traversal tree to width W (2, 4 or 8)
 * from the binary tree in hs.nodes: sets nodes4 / nodes8, trav_width and
 * trav_stack; the slots and their order do not change.  Throws BuildError. */
void collapse_traversal(HostScene& hs, int W);
/* The width FM_BVH_WIDTH_AUTO takes for a two-wide tree of binary_nodes
 * nodes over `slots` primitive records: 2 while nodes and records fit in
 * the device's last-level cache (FM_MALL_BYTES), else 4. */
int auto_width(size_t binary_nodes, size_t slots);

/* Throws BuildError.  only_bvh stops after the reference BVH (host tests). */
void build_host_scene(const fm_scene_desc& d, HostScene& hs, bool only_bvh = false);

}  // namespace fmh
#endif
