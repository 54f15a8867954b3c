/*
 * layout.h - HBM-resident data layouts shared by the host builder and the
 * gfx950 kernels.  Every record is a whole number of 16-byte vectors so that
 * one lane reads it with dwordx4 loads.
 */
#pragma once
#include <stdint.h>

namespace vvpt {

constexpr uint32_t kInvalid = 0xffffffffu;

/* Wide BVH node: 8 children with 8-bit quantised boxes (the compressed wide
 * node of Ylitie, Karras and Laine, HPG 2017), 80 bytes = 5 x uint4.
 *   word0: p.x p.y p.z (float) | e.x e.y e.z (biased exponents, 8 bits each) | imask (8 bits)
 *   word1: child_base | leaf_base | meta[0..3] | meta[4..7]
 *   word2..4: qlo.x[8] qlo.y[8] | qlo.z[8] qhi.x[8] | qhi.y[8] qhi.z[8]
 * meta[slot]: 0 = empty; internal child: 0x20 | (24 + slot); leaf child with n
 * (1..3) primitives at offset k (0..23) from leaf_base: (unary(n) << 5) | k.
 * Internal children of a node are stored contiguously from child_base in slot
 * order; slot s holds the child met first by rays of direction octant s
 * (bit0: dx<0, bit1: dy<0, bit2: dz<0). */
struct alignas(16) WideNode {
    float px, py, pz;
    uint32_t e_imask;   /* e.x | e.y << 8 | e.z << 16 | imask << 24 */
    uint32_t child_base;
    uint32_t leaf_base;
    uint32_t meta[2];
    uint32_t qlo_x[2], qlo_y[2];
    uint32_t qlo_z[2], qhi_x[2];
    uint32_t qhi_y[2], qhi_z[2];
};
static_assert(sizeof(WideNode) == 80, "wide node must be 80 bytes");

/* Leaf triangle in object space, 48 bytes: v0 | prim id, e1 = v1 - v0, e2 = v2 - v0. */
struct alignas(16) LeafTri {
    float v0[3];
    uint32_t prim;
    float e1[3];
    uint32_t pad0;
    float e2[3];
    uint32_t pad1;
};
static_assert(sizeof(LeafTri) == 48, "leaf triangle must be 48 bytes");

/* Instance: mirrors InstanceData (core/Instances.hpp:17-35), 128 bytes. */
struct alignas(16) DInstance {
    float w2o[12];        /* world-to-object rows (gl_WorldToObjectEXT) */
    float o2w[12];        /* object-to-world rows (gl_ObjectToWorldEXT) */
    uint32_t blas_root;   /* node index of the mesh's BLAS root */
    uint32_t material;
    float idg, idb;       /* front / back volume material, -1 = none (InstanceData.id.g/.b) */
    uint32_t vtx_offset;  /* first vertex of the mesh in the global vertex array */
    uint32_t idx_offset;  /* first triangle of the mesh in the global index array */
    uint32_t ntri;
    uint32_t flags;       /* INST_* bits of its material and volumes */
};
static_assert(sizeof(DInstance) == 128, "instance must be 128 bytes");

/* TLAS leaf in leaf order: what entering an instance needs, in one 64-byte record (one dependent load instead of
 * a leaf-to-instance index and then the instance). */
struct alignas(16) TlasLeaf {
    float w2o[12];      /* world-to-object rows, as DInstance::w2o */
    uint32_t inst;      /* global instance index */
    uint32_t blas_root; /* as DInstance::blas_root */
    uint32_t flags;     /* as DInstance::flags */
    uint32_t pad;
};
static_assert(sizeof(TlasLeaf) == 64, "TLAS leaf must be 64 bytes");

enum : uint32_t {
    INST_TRANSPARENT = 1u,   /* material transparent >= 0.99: shadow and emitter-lookup rays pass with 1 - alpha */
    INST_ALPHA_TEST = 2u,    /* material transparent > 0: path rays test alpha stochastically */
    INST_VOL_CHANGE = 4u,    /* front and back volume differ */
    INST_EMISSIVE = 8u,      /* material has non-zero emission */
    INST_NEE_PASS = 16u      /* transparent, not emissive, no volume change: emitter lookups pass it */
};

/* Shading vertex (Vertex of core/Mesh.hpp:15-21 without the unused colour), 64 bytes. */
struct alignas(16) DVertex {
    float px, py, pz, u;
    float v, nx, ny, nz;
    float tx, ty, tz, pad0;
    float bx, by, bz, pad1;
};
static_assert(sizeof(DVertex) == 64, "vertex must be 64 bytes");

/* MaterialData (shaders/include/structs.glsl:41-52) with texture slots bound to
 * the global texture table, 128 bytes. */
struct alignas(16) DMaterial {
    float albedo[4];
    float mra[4];        /* metallic, roughness, ao, transparent */
    float emissive[4];   /* rgb, intensity */
    uint32_t tex1[4];    /* albedo, metallic, roughness, ao */
    uint32_t tex2[4];    /* emissive, normal, (unused), alpha */
    float tiling[4];     /* u, v, type, unused */
    uint32_t flags;      /* bit0 emissive-capable, bit1 transparent, bit2 volume */
    uint32_t pad[7];
};
static_assert(sizeof(DMaterial) == 128, "material must be 128 bytes");

/* Texture table entry; texels live in one RGBA8 pool and one RGBA32F pool. */
struct alignas(16) DTexture {
    uint32_t offset;  /* texel offset in the pool of its format */
    uint32_t width, height;
    uint32_t format;  /* vvpt_texture_format */
};

/* Light: point, directional or mesh (LightInstance + LightData, structs.glsl:55-69). */
struct alignas(16) DLight {
    uint32_t type;  /* 0 point, 1 directional, 2 mesh */
    uint32_t inst;  /* mesh lights: instance index */
    float intensity;
    uint32_t pad;
    float color[4];
    float position[4];
};
static_assert(sizeof(DLight) == 48, "light must be 48 bytes");

}  // namespace vvpt
