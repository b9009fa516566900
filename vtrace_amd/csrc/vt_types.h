// vt_types.h — data layout shared by the HIP kernels and the host runtime of librender_hip.
//
// HBM layout (one GPU):
//   * volumes: per volume a brick map (u32 per 8^3 brick, x-fastest brick order; empty bricks
//     hold VT_EMPTY_BRICK with their empty-space sizes, below), an occupancy-mask pool (64 B =
//     16 u32 per non-empty brick, bit x + 8y + 64z of the brick), a block-mask pool (u64 per
//     non-empty brick, bit bx + 4by + 16bz set when its 2^3 block (2bx.., 2by.., 2bz..) holds an
//     occupied voxel) and a payload pool (512 B u8 palette indices or 2048 B RGBA8 per non-empty
//     brick, voxel x + 8y + 64z), plus a 256-entry palette.
//     Replaces the reference's R8G8B8A8_SRGB 3-D images (lib/memory.c:286-385).
//   * a volume-descriptor table (72 B per volume) replacing the bindless sampler3D array
//     (lib/descriptor.c:38-95, shaders/trace.frag:31).
//   * instances: the 64 B ABI records (src/render.rs:31-35) expanded to 128 B GPU records
//     holding M and inverse(M) (rows 0..2), the texture id and the volume size.
//   * LBVH: 2n-1 nodes of 32 B (lo[3], hi[3], left, right); internal nodes 0..n-2, leaves
//     n-1..2n-2 with left = instance index and right = 0xFFFFFFFF.
//   * per pixel: a 16 B hit record and an RGBA32F radiance value.
#pragma once

#include <stdint.h>

#define VT_EMPTY 0xFFFFFFFFu
// Brick map entries: a non-empty brick holds its id (< VT_EMPTY_BRICK); an empty brick holds
//   VT_EMPTY_BRICK | r << 24 | (K_o - 1) << 3o   (o = 0..7)
// r = radius of the cube of empty bricks centred on it (L-inf brick distance to the nearest
// non-empty brick - 1, at most VT_BRICK_RADIUS_CAP); K_o = side, in bricks, of the largest cube
// of empty bricks that has this brick as its corner and extends towards octant o (bit a of o
// set: towards -axis a), at most VT_BRICK_CUBE_CAP.  Bricks outside the volume count as empty.
#define VT_EMPTY_BRICK 0xE0000000u
#define VT_BRICK_RADIUS_CAP 31u
#define VT_BRICK_CUBE_CAP 8u
#ifdef __HIPCC__
__host__ __device__
#endif
inline bool vt_brick_empty(uint32_t e) { return e >= VT_EMPTY_BRICK; }
#ifdef __HIPCC__
__host__ __device__
#endif
inline uint32_t vt_brick_radius(uint32_t e) { return (e >> 24) & 31u; }
#ifdef __HIPCC__
__host__ __device__
#endif
inline uint32_t vt_brick_cube(uint32_t e, uint32_t octant) { return ((e >> (3u * octant)) & 7u) + 1u; }

struct VolumeDesc {             // 72 B
    uint32_t dim[3];
    uint32_t bdim[3];
    uint32_t fmt;               // 0 = u8 index + palette, 1 = RGBA8
    uint32_t nbricks;           // non-empty bricks
    const uint32_t* brick_map;
    const uint32_t* masks;
    const uint8_t* payload;
    const uint32_t* palette;
    const uint64_t* blocks;     // 2^3-block masks of the non-empty bricks
};

struct InstanceGPU {            // 128 B
    float M[12];                // model matrix rows 0..2, column-major: M[c*3 + r]
    float inv[12];              // inverse(model) rows 0..2, column-major
    uint32_t tex;
    uint32_t valid;
    float size[3];
    uint32_t dim[3];
    uint32_t pad[2];
};

struct BvhNode {                // 32 B
    float lo[3];
    float hi[3];
    uint32_t left;
    uint32_t right;
};

struct HitRecord {              // 16 B, one per pixel
    uint32_t inst;              // instance index or 0xFFFFFFFF
    uint32_t voxel;             // bits 0-9 of x | y << 10 | z << 20
    uint16_t steps;             // DDA steps, saturated at 65535 (flags bit 3)
    uint8_t face;               // 0..5 = -X,+X,-Y,+Y,-Z,+Z (outward normal), 0xFF miss
    uint8_t pal;                // palette index (u8 volumes)
    uint8_t mask;               // last DDA step mask
    uint8_t flags;              // bit0 hit, bit1 primary vertex shadowed, bit2 a voxel coordinate
                                // >= 32768 (only its low 15 bits held), bit3 steps saturated
    uint16_t vhi;               // bits 10-14 of x | y << 5 | z << 10 (zero below 1024)
};

// Voxel and step fields of a hit record; returns the flag bits 2-3 they need.
#ifdef __HIPCC__
__host__ __device__
#endif
inline uint32_t vt_pack_hit_voxel(HitRecord& r, uint32_t x, uint32_t y, uint32_t z, uint32_t steps) {
    r.voxel = (x & 1023u) | ((y & 1023u) << 10) | ((z & 1023u) << 20);
    r.vhi = (uint16_t)(((x >> 10) & 31u) | (((y >> 10) & 31u) << 5) | (((z >> 10) & 31u) << 10));
    r.steps = (uint16_t)(steps < 65535u ? steps : 65535u);
    return ((x | y | z) >= 32768u ? 4u : 0u) | (steps > 65535u ? 8u : 0u);
}

// Per-frame constants, passed by value to the render kernels.
struct FrameGPU {
    float PC[16];               // projection * camera
    float AB[16];               // inverse(centered camera) * inverse(projection)
    float R[9];                 // rotation part of inverse(centered camera), column-major 3x3
    float O[3];                 // camera position
    float ip00, ip11;           // inverse(projection)[0][0], [1][1]
    float W, Hv;                // viewport width and height (Hv = W for the reference quirk)
    uint32_t width, height;
    uint32_t mode, spp, max_bounces, jitter, seed, frame;
    float sun[3];
    float sun_rad[3];
    float sky[3];
    uint32_t stats;             // nonzero: count traversal work into the stats buffer
};

struct SceneGPU {
    const InstanceGPU* inst;
    const BvhNode* nodes;
    const VolumeDesc* vols;
    uint32_t ninst;
    uint32_t nvol;
    uint32_t bvh_depth;         // selects the traversal-stack variant
};

// One request (frame, or a chunk of its samples) in a batched launch of the path tracer.
struct ReqGPU {
    FrameGPU F;
    SceneGPU S;
    HitRecord* hits;        // per pixel, written by the chunk that holds sample 0
    float4* acc;            // per pixel running radiance sum over samples (FP32, in sample order)
    float4* rad;            // per pixel final radiance, written by the last chunk (acc / spp)
    float4* gbuf;           // optional, 2 per pixel averaged over samples: (albedo rgb, distance),
                            // (normal, hit fraction); needs the wave buffers' gsamp
    float4* gacc;           // with gbuf over several chunks: per pixel running G-buffer sum (2 float4)
    uint32_t s0, s1;        // samples [s0, s1) are traced in this launch
    uint32_t final_chunk;   // 1: s1 == spp
    uint32_t tile_w;        // pixel tile of 64 work items: tile_w x (64 / tile_w)
    uint32_t ty0, ty1;      // rows of tiles [ty0, ty1) are traced in this launch (ty1 = 0: all)
};

// The request's rows of tiles [*ty0, *ty1) and its tile columns.
#ifdef __HIPCC__
__host__ __device__
#endif
inline void vt_req_tiles(const ReqGPU& R, uint32_t* ty0, uint32_t* ty1, uint32_t* tiles_x) {
    const uint32_t tw = R.tile_w, th = 64u / R.tile_w;
    const uint32_t tiles_y = (R.F.height + th - 1u) / th;
    *tiles_x = (R.F.width + tw - 1u) / tw;
    *ty1 = R.ty1 ? (R.ty1 < tiles_y ? R.ty1 : tiles_y) : tiles_y;
    *ty0 = R.ty0 < *ty1 ? R.ty0 : *ty1;
}

#define VT_MAX_BATCH 64
// Work and pixel spaces of a batch: request k owns work items [work_base[k], work_base[k+1])
// (multiples of 64) and accumulate threads [pix_base[k], pix_base[k+1]).
struct BatchGPU {
    uint32_t nreq;
    uint32_t max_bounces;   // bounce stages launched (max over the requests)
    uint32_t work_base[VT_MAX_BATCH + 1];
    uint32_t pix_base[VT_MAX_BATCH + 1];
};

// Launch stages of one frame (stats and timings are kept per stage):
// mode 0: [render0]; modes 1/2: [primary, shadow d0, bounce d0, ..., shadow dB, accumulate].
#define VT_MAX_STAGES 64
enum VtStageKind {
    VT_STAGE_RENDER0 = 0,
    VT_STAGE_PRIMARY,
    VT_STAGE_SHADOW,      // (unused: sun rays run in the secondary stage)
    VT_STAGE_BOUNCE,      // (unused: bounce rays run in the secondary stage)
    VT_STAGE_ACCUMULATE,
    VT_STAGE_SECONDARY    // sun and bounce rays of one depth
};

// stats buffer layout (u64 each, VT_STAT_COUNT per stage)
enum VtStat {
    VT_STAT_RAYS = 0,
    VT_STAT_NODES,
    VT_STAT_INSTANCES,
    VT_STAT_BRICKS,
    VT_STAT_BRICKS_NONEMPTY,
    VT_STAT_HITS,
    VT_STAT_STEPS,
    VT_STAT_COUNT
};
