/*
 * Device-resident scene layout, shared by the host flattener and the HIP
 * kernels.  Plain structs, no pointers to host memory.
 *
 * HBM layout (all arrays 64-byte aligned, indexed by leaf slot p = position
 * of a geometry in the reference BVH's m_objList, bvh_aggregate.cpp:43-55):
 *   nodes    BVH2 node = both child boxes + child refs, 64 B (one line)
 *   prims    3 x float4 per slot: triangle v0,v1,v2 (w: type, aux index);
 *            rectangles store pa,pb,pc and pd in the rect table;
 *            spheres store their centre/radius and index the sphere table
 *   meta     int4 per slot {geometry id, material, light, flags}
 *   tri_sh   4 x float4 per slot: vertex normals and uvs (shading only)
 */
#ifndef FM_LAYOUT_H
#define FM_LAYOUT_H

#include <stdint.h>

#define FM_PRIM_TRI 0
#define FM_PRIM_RECT 1
#define FM_PRIM_SPHERE 2

#define FM_META_HAS_NORMAL 1

/* w of a slot's first float4: FM_PRIM_* | shading class << 8 (1 + material
 * type; both delta materials share class 3) | emissive << 16, so that a
 * closest hit is classified without reading the slot's meta and material */
#define FM_SHADE_CLASS(mat_type) ((mat_type) == 3 ? 3 : 1 + (mat_type))
#define FM_PRIM_KIND(w) ((w) & 0xff)
#define FM_PRIM_INFO(w) ((w) >> 8)

/* child ref: >= 0 inner node index; < 0 leaf: ~((start << 3) | count) */
#define FM_LEAF_REF(start, count) (~(int32_t)(((uint32_t)(start) << 3) | (uint32_t)(count)))
#define FM_REF_IS_LEAF(r) ((r) < 0)
#define FM_LEAF_START(r) ((int32_t)(((uint32_t)(~(r))) >> 3))
#define FM_LEAF_COUNT(r) ((int32_t)(((uint32_t)(~(r))) & 7u))

#define FM_STACK_DEPTH 32   /* depth limit of the binary traversal trees */
#define FM_STACK_MAX 64     /* LDS traversal stack entries per lane, at most */
#define FM_SHORT_STACK 16   /* LDS entries per lane of the short (spilling) traversal stack; a power of two */
/* MI355X's Infinity Cache (MALL), the last level before HBM: traversal
 * data that fits stays cache-resident across a frame's bounces. */
#define FM_MALL_BYTES (256ull << 20)

/* Traversal trees are two-wide (fm_dnode), four-wide (fm_dnode4) or
 * eight-wide (fm_dnode8). */
#define FM_NODE4_EMPTY 0x7fffffff   /* unused child of a four- or eight-wide node */

typedef struct fm_dnode {
    float b[12];     /* c0 min xyz, c0 max xyz, c1 min xyz, c1 max xyz */
    int32_t child[2];
    int32_t pad[2];
} fm_dnode;          /* 64 bytes */

/* Four child boxes as a structure of arrays (one float4 per bound plane,
 * lane k = child k), then the four child refs. */
typedef struct fm_dnode4 {
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    int32_t child[4];
    int32_t pad[4];
} fm_dnode4;         /* 128 bytes */

/* Eight child boxes as a structure of arrays (two float4 per bound plane,
 * lanes 0-7 = children 0-7), then the eight child refs. */
typedef struct fm_dnode8 {
    float lo_x[8], hi_x[8], lo_y[8], hi_y[8], lo_z[8], hi_z[8];
    int32_t child[8];
    int32_t pad[8];
} fm_dnode8;         /* 256 bytes */

typedef struct fm_drect {
    float pa[4], pb[4], pc[4], pd[4]; /* xyz + pad */
    float normal[4];                   /* w = area */
    float tangent[4];                  /* w = (float)(1.0 / (double)area) */
} fm_drect;          /* 96 bytes */

typedef struct fm_dsphere {
    float w2o[16];    /* column-major world_to_object */
    float o2w[16];    /* column-major object_to_world */
    float radius, radius_world, rw2, area;
    float inv_area;   /* 1 / area (float division, Geometry::Pdf) */
    float center[3];  /* object_to_world.ApplyToPoint(0) */
} fm_dsphere;         /* 160 bytes */

typedef struct fm_dtex {
    int32_t type, image, sampler, wrap_u, wrap_v;
    float value[3];
} fm_dtex;            /* 32 bytes */

typedef struct fm_dmaterial {
    int32_t type, thin, has_image, pad;
    fm_dtex tex[15];
    float refl[4], refr[4];
    float eta_i, eta_t, pad2[2];
    float c_eta2[4], c_etak2[4];   /* conductor: (eta_t / eta_i)^2, (k / eta_i)^2 per channel */
} fm_dmaterial;

#define FM_DLIGHT_TRI 0
#define FM_DLIGHT_RECT 1
#define FM_DLIGHT_SPHERE 2
#define FM_DLIGHT_ENV 3
#define FM_DLIGHT_DIR 4
#define FM_DLIGHT_MESH 5    /* emissive mesh, "embree": aux = first triangle in mesh_tris,
                               p[0][3] = 1 / area sum, int bits of p[1][0] = float
                               offset of its area CDF in env_tables, of p[1][1] =
                               triangle count */

typedef struct fm_dlight {
    int32_t kind;       /* FM_DLIGHT_* */
    int32_t aux;        /* rect / sphere table index, env index */
    int32_t has_normal; /* triangle */
    int32_t geom;       /* reference geometry index */
    float radiance[4];
    float p[3][4];      /* triangle v0 v1 v2 (w of p[0] = (float)(1.0/(double)area)) */
    float n[3][4];      /* triangle vertex normals */
    float dir[4];       /* directional: normalised direction */
} fm_dlight;            /* 144 bytes */

typedef struct fm_denv {
    int32_t image, W, H, pad;
    float factor, marg_sum, pad2[2];
    float l2w[12];      /* 3 columns (xyz + pad) of light_to_world */
    float w2l[12];
    int64_t off_marg_cdf;   /* floats: H + 1 */
    int64_t off_marg_val;   /* H */
    int64_t off_cond_cdf;   /* H * (W + 1) */
    int64_t off_cond_val;   /* H * W */
    int64_t off_cond_sum;   /* H */
    int64_t off_marg_guide; /* int32 bits, guide_m + 1: last row with cdf <= g / guide_m */
    int64_t off_cond_guide; /* int32 bits, H * (guide_c + 1): the same per row */
    int32_t guide_m, guide_c;  /* powers of two */
    int64_t off_aux;        /* doubles into env_aux: sin(theta) of each row's texel centres (H),
                               each row's raw sum of luminance x sin(theta) (H), their total (1) */
} fm_denv;

typedef struct fm_dimage {
    int32_t w, h, format, pad;   /* format: 0 float4 texels, 1 RGBA8 sRGB */
    int64_t offset;              /* bytes into the texel buffer */
} fm_dimage;

typedef struct fm_dcamera {
    int32_t type, pad[3];
    float pos[4];
    float fw, fh, fd, lens;
    float c2w[16];               /* column-major camera_to_world */
} fm_dcamera;

/* Everything a kernel needs, passed by value. */
typedef struct fm_dscene {
    const fm_dnode* nodes;       /* width 2 */
    const fm_dnode4* nodes4;     /* width 4 */
    const fm_dnode8* nodes8;     /* width 8 */
    const float* prims;          /* float4 x 3 per slot */
    const int32_t* meta;         /* int4 per slot */
    const float* tri_sh;         /* float4 x 4 per slot */
    const fm_drect* rects;
    const fm_dsphere* spheres;
    const fm_dmaterial* materials;
    const fm_dlight* lights;
    const fm_dlight* mesh_tris;  /* triangles of mesh lights (FM_DLIGHT_TRI records) */
    const fm_denv* envs;
    const fm_dimage* images;
    const unsigned char* texels;
    const float* env_tables;
    const double* env_aux;       /* FP64 row facts of the env tables (fm_denv.off_aux) */
    const float* srgb_lut;       /* 256 entries */
    int32_t root_ref;            /* node index or leaf ref; ignored if num_slots == 0 */
    int32_t num_slots;
    int32_t num_lights;
    int32_t env_light;           /* light index of the first env light or -1 */
    int32_t env;                 /* env table index or -1 */
    int32_t naive;               /* brute force in object order */
    int32_t shear_float_abs;     /* triangle test's axis choice: 0 truncating integer abs, 1 float abs */
    int32_t num_materials;
    int32_t width;               /* traversal tree: 2, 4 or 8 */
    float root_box[6];
    int32_t stack_levels;        /* LDS stack entries a traversal needs */
    uint32_t node_bytes;         /* size of nodes (two-wide), < 4 GiB */
    fm_dcamera camera;
} fm_dscene;

#endif
