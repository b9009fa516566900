// trace.hip — ray generation, LBVH traversal, brick DDA and shading (wavefront form).
//
// Replaces the reference's raster pass (lib/command.c:50-109: one proxy cube per
// instance, back-face cull, depth LESS) and its fragment shader (shaders/trace.frag:41-90):
//   * a pixel sample's camera ray finds, through the LBVH, every instance box whose front
//     face it enters (the proxy fragments of that pixel, inside the clip volume);
//   * for each such box it rebuilds the fragment's inputs (model_position on the entry
//     face, screen_position = P*C*(M*model_position)) and runs trace.frag's ray setup and
//     DDA op for op in FP32 (trace.frag:46-87), on the bricked volume instead of a texture;
//   * the winner is the box with the smallest entry depth whose DDA hits (depth LESS with
//     discard, pipeline.c:146-152), ties to the lower instance index (draw order).
// Modes: 0 = primary colour blended over the sky (trace.frag:79, pipeline.c:129-137);
//        1 = primary + one shadow ray to a directional sun;
//        2 = path tracing with cosine-weighted diffuse bounces (hash RNG per pixel/sample).
// The secondary-ray rules (origin voxel face, cosine sampling, sun/sky terms) have no
// counterpart in the reference and follow this project's specification.
//
// Execution (modes 1 and 2) is a wavefront: k_primary traces every pixel sample and shades
// its first hit; then one k_secondary launch per bounce depth traces, side by side, the sun
// rays and the diffuse bounce rays of the surfaces queued at that depth, appending the
// bounce hits (wave-aggregated atomics) to a compacted queue for the next depth;
// k_accumulate sums the samples of each pixel.
// The kernels are persistent: a k_primary wave fetches 64 consecutive work items at a time
// from an atomic counter; a k_secondary wave refills its lanes as their rays finish, so
// neither terminated paths nor the rest of a wave waiting on one long walk occupy lanes.
// Each path's arithmetic is independent of queue order, so results are deterministic.
//
// DDA: primary rays walk the reference's voxel DDA (empty 8^3 bricks crossed by a tight loop
// without memory access); secondary rays cross boxes known to be empty (brick cubes, 4^3
// octants, 2^3 blocks) one per step.  In non-empty bricks the 64 B occupancy mask is staged
// once into the thread's LDS slot and tested from LDS.  The LBVH stack lives in LDS as well.
//
// All FP32 expressions are compiled with -ffp-contract=off and evaluated in the order
// written (correctly rounded division and square root), which keeps the integer outputs
// (instance, voxel, face, palette index, steps) bit-exact with the CPU oracle.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstddef>

#include <cstring>

#include "vt_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr int kStackLds = 16;   // LBVH stack entries per thread held in LDS
// Wavefront counter slots: [0] queue of the primary vertices, [2] primary (and mode 0) fetch,
// [3, 64) fetch counters of the secondary launches, [64 + d + 1] the queue built at depth d.
// Device-scope atomics on one address serialise (about 12 ns each on MI355X), which would bound
// a launch at (items / 64) x 12 ns, so every slot holds one counter per part, kPartStride u32
// apart: workgroups are dispatched round-robin over the 8 XCDs, workgroup b starts on part
// b % kParts, and a launch's atomics spread over kParts addresses.  A queue is kParts
// segments of seg_cap slots; part p of a queue's counters fills segment p.
constexpr uint32_t kSlots = 128;
constexpr uint32_t kQueueCounters = 64;
constexpr uint32_t kParts = 8;
constexpr uint32_t kPartStride = 64;   // 256 B between the counters of one slot
constexpr uint32_t kCounters = kSlots * kParts * kPartStride;
// Counter slot of the sun list of the queue whose counters are in slot qi: [1] for the queue of
// the primary vertices, qi + 32 for the queue built at depth d (slot 64 + d + 1).
__host__ __device__ constexpr uint32_t sun_slot(uint32_t qi) { return qi == 0 ? 1u : qi + 32u; }
__host__ __device__ constexpr uint32_t seg_cap(uint32_t npaths) { return ((npaths + kParts - 1) / kParts + 63u) & ~63u; }

// GLSL min/max as "y if y < x else x" (also fixes the NaN cases the DDA can produce)
__device__ __forceinline__ float gmin(float x, float y) { return (y < x) ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return (x < y) ? y : x; }
__device__ __forceinline__ float gsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

__device__ __forceinline__ uint32_t pcg_hash(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ float u01(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

struct Rng {
    uint32_t s;
    __device__ __forceinline__ float next() {
        s = pcg_hash(s);
        return u01(s);
    }
};

template <bool ON>
struct Ctr {
    __device__ __forceinline__ void ray() {}
    __device__ __forceinline__ void node() {}
    __device__ __forceinline__ void instance() {}
    __device__ __forceinline__ void brick(bool) {}
    __device__ __forceinline__ void hit() {}
    __device__ __forceinline__ void steps(uint32_t) {}
    __device__ __forceinline__ void flush(unsigned long long*) {}
};
template <>
struct Ctr<true> {
    uint32_t c[VT_STAT_COUNT] = {0, 0, 0, 0, 0, 0, 0};
    __device__ __forceinline__ void ray() { c[VT_STAT_RAYS]++; }
    __device__ __forceinline__ void node() { c[VT_STAT_NODES]++; }
    __device__ __forceinline__ void instance() { c[VT_STAT_INSTANCES]++; }
    __device__ __forceinline__ void brick(bool ne) {
        c[VT_STAT_BRICKS]++;
        c[VT_STAT_BRICKS_NONEMPTY] += ne ? 1u : 0u;
    }
    __device__ __forceinline__ void hit() { c[VT_STAT_HITS]++; }
    __device__ __forceinline__ void steps(uint32_t n) { c[VT_STAT_STEPS] += n; }
    __device__ void flush(unsigned long long* out) {
        for (int k = 0; k < VT_STAT_COUNT; ++k) {
            unsigned long long v = c[k];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if ((threadIdx.x & 63) == 0 && v) atomicAdd(&out[k], v);
        }
    }
};

// ---- per-thread LDS slots ------------------------------------------------------------
// Interleaved [entry][thread] so that lanes reading their own slot never share a bank.
struct Lds {
    uint32_t* base;  // &smem[threadIdx.x]
    __device__ __forceinline__ uint32_t& at(int k) const { return base[k * kBlock]; }
};

// LBVH traversal stack: in LDS (interleaved per thread) when the tree is shallow enough,
// otherwise a private array (trees deeper than kStackLds - 1 levels).
template <bool DEEP>
struct TravStack {
    Lds lds;
    int sp;
    __device__ __forceinline__ void push(uint32_t x) { lds.at(sp++) = x; }
    __device__ __forceinline__ uint32_t pop() { return lds.at(--sp); }
};
template <>
struct TravStack<true> {
    Lds lds;
    int sp;
    uint32_t a[64];
    __device__ __forceinline__ void push(uint32_t x) { a[sp++] = x; }
    __device__ __forceinline__ uint32_t pop() { return a[--sp]; }
};

// Loads through the global address space (pointers held in structs are otherwise generic
// and compile to flat accesses, which also wait on LDS traffic).
typedef unsigned int vt_u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint32_t gld(const uint32_t* p) { return *(const __attribute__((address_space(1))) uint32_t*)(p); }
__device__ __forceinline__ float gld(const float* p) { return *(const __attribute__((address_space(1))) float*)(p); }
__device__ __forceinline__ uint4 gld(const uint4* p) {
    const vt_u32x4 x = *(const __attribute__((address_space(1))) vt_u32x4*)(p);
    return make_uint4(x.x, x.y, x.z, x.w);
}
__device__ __forceinline__ BvhNode gld(const BvhNode* p) {
    const uint4 a = gld(reinterpret_cast<const uint4*>(p)), b = gld(reinterpret_cast<const uint4*>(p) + 1);
    BvhNode n;
    n.lo[0] = __uint_as_float(a.x); n.lo[1] = __uint_as_float(a.y); n.lo[2] = __uint_as_float(a.z);
    n.hi[0] = __uint_as_float(a.w); n.hi[1] = __uint_as_float(b.x); n.hi[2] = __uint_as_float(b.y);
    n.left = b.z;
    n.right = b.w;
    return n;
}

__device__ __forceinline__ uint64_t gld64(const uint64_t* p) {
    return *(const __attribute__((address_space(1))) uint64_t*)(p);
}
__device__ __forceinline__ uint8_t gld8(const uint8_t* p) { return *(const __attribute__((address_space(1))) uint8_t*)(p); }

// A volume descriptor read through the global address space (its fields at their offsets).
static_assert(offsetof(VolumeDesc, fmt) == 24 && offsetof(VolumeDesc, brick_map) == 32 &&
                  offsetof(VolumeDesc, blocks) == 64 && sizeof(VolumeDesc) == 72,
              "VolumeDesc layout");
__device__ __forceinline__ VolumeDesc gld_vol(const VolumeDesc* p) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(p);
    const uint64_t* q = reinterpret_cast<const uint64_t*>(w + 8);
    VolumeDesc v;
    v.dim[0] = gld(w + 0); v.dim[1] = gld(w + 1); v.dim[2] = gld(w + 2);
    v.bdim[0] = gld(w + 3); v.bdim[1] = gld(w + 4); v.bdim[2] = gld(w + 5);
    v.fmt = gld(w + 6);
    v.nbricks = gld(w + 7);
    v.brick_map = reinterpret_cast<const uint32_t*>(gld64(q + 0));
    v.masks = reinterpret_cast<const uint32_t*>(gld64(q + 1));
    v.payload = reinterpret_cast<const uint8_t*>(gld64(q + 2));
    v.palette = reinterpret_cast<const uint32_t*>(gld64(q + 3));
    v.blocks = reinterpret_cast<const uint64_t*>(gld64(q + 4));
    return v;
}
// The volume of instance I.
__device__ __forceinline__ VolumeDesc inst_vol(const SceneGPU& S, const InstanceGPU* I) {
    return gld_vol(&S.vols[gld(&I->tex)]);
}

// Stage a brick's 64 B occupancy mask (16 u32: z-slice z is words 2z, 2z+1) in the LDS slot.
__device__ __forceinline__ void stage_mask(const uint32_t* masks, uint32_t entry, const Lds& mslot) {
    const uint4* src = reinterpret_cast<const uint4*>(masks + (size_t)entry * 16);
    const uint4 w0 = gld(src + 0), w1 = gld(src + 1), w2 = gld(src + 2), w3 = gld(src + 3);
    mslot.at(0) = w0.x; mslot.at(1) = w0.y; mslot.at(2) = w0.z; mslot.at(3) = w0.w;
    mslot.at(4) = w1.x; mslot.at(5) = w1.y; mslot.at(6) = w1.z; mslot.at(7) = w1.w;
    mslot.at(8) = w2.x; mslot.at(9) = w2.y; mslot.at(10) = w2.z; mslot.at(11) = w2.w;
    mslot.at(12) = w3.x; mslot.at(13) = w3.y; mslot.at(14) = w3.z; mslot.at(15) = w3.w;
}
__device__ __forceinline__ uint64_t mask_slice(const Lds& mslot, int z) {
    return (uint64_t)mslot.at(2 * z) | ((uint64_t)mslot.at(2 * z + 1) << 32);
}

// ---- small vector helpers (explicit evaluation order) -------------------------------
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ float comp(const V3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// rows 0..2 of a column-major 3x4 (m[c*3 + r]) applied to (v, w)
__device__ __forceinline__ V3 xf34(const float* m, V3 v, float w) {
    V3 o;
    o.x = ((m[0] * v.x + m[3] * v.y) + m[6] * v.z) + m[9] * w;
    o.y = ((m[1] * v.x + m[4] * v.y) + m[7] * v.z) + m[10] * w;
    o.z = ((m[2] * v.x + m[5] * v.y) + m[8] * v.z) + m[11] * w;
    return o;
}
// column-major 4x4 times vec4
__device__ __forceinline__ float4 mv4(const float* m, float4 v) {
    float4 o;
    o.x = ((m[0] * v.x + m[4] * v.y) + m[8] * v.z) + m[12] * v.w;
    o.y = ((m[1] * v.x + m[5] * v.y) + m[9] * v.z) + m[13] * v.w;
    o.z = ((m[2] * v.x + m[6] * v.y) + m[10] * v.z) + m[14] * v.w;
    o.w = ((m[3] * v.x + m[7] * v.y) + m[11] * v.z) + m[15] * v.w;
    return o;
}
__device__ __forceinline__ V3 normalize3(V3 v) {
    float l = sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z);
    return V3{v.x / l, v.y / l, v.z / l};
}

// ---- DDA on the bricked volume (trace.frag:66-87) -----------------------------------
struct Dda {
    int v[3];
    uint32_t steps;
    uint32_t mask;
    float t_last;
    float q[3];       // start of the voxel walk that ended at v (hit point = q + t_last * unit dir)
    uint32_t entry;   // brick id of the hit voxel
    uint32_t li;      // voxel index inside the brick
    bool hit;
};

// State of one reference DDA walk (trace.frag:66-87).
struct DdaWalk {
    int vx, vy, vz;
    float sx, sy, sz;            // side distances
    float dx, dy, dz;            // deltas
    float zx, zy, zz;            // 0 * delta: what float(mask) * delta adds for an axis not stepped
    int stx, sty, stz;
    uint32_t steps, entry, li;
    float t_last;
    uint32_t lm;                 // axes of the last step (bit a: axis a stepped)
    bool hit;
};

// The walk over the bricked volume.  FAST: every delta is finite, so 0 * delta is 0, no side
// distance can become NaN and GLSL's min(y, x) = (y < x ? y : x) in "side <= min(a, b)" is
// "side <= min3(sx, sy, sz)" (one min3 for the three axes); otherwise the selects reproduce
// min's NaN behaviour exactly.  A brick is left
// through one voxel per stepping axis (the first voxel outside it on that side), which is
// what the loops test.
template <bool ST, bool FAST>
__device__ __forceinline__ void dda_walk(const VolumeDesc& V, DdaWalk& w, Ctr<ST>& ctr, const Lds& mslot) {
    const int nx = (int)V.dim[0], ny = (int)V.dim[1], nz = (int)V.dim[2];
    const uint32_t max_steps = V.dim[0] + V.dim[1] + V.dim[2];
    const uint32_t bdx = V.bdim[0], bdy = V.bdim[1];
    int vx = w.vx, vy = w.vy, vz = w.vz;
    float sx = w.sx, sy = w.sy, sz = w.sz;
    const float dx = w.dx, dy = w.dy, dz = w.dz, zx = w.zx, zy = w.zy, zz = w.zz;
    const int stx = w.stx, sty = w.sty, stz = w.stz;
    // exit voxel of a brick along a stepping axis: min(b8 + c, l) (c = 8, l = n when stepping
    // up; c = -1, l = INT_MAX when stepping down)
    const int cx = stx > 0 ? 8 : -1, cy = sty > 0 ? 8 : -1, cz = stz > 0 ? 8 : -1;
    const int lx = stx > 0 ? nx : INT_MAX, ly = sty > 0 ? ny : INT_MAX, lz_ = stz > 0 ? nz : INT_MAX;
    uint32_t steps = 0, entry = VT_EMPTY, li = 0;
    float t_last = 0.0f;
    uint32_t lm = 0;
    bool hit = false;
// One DDA step (trace.frag:83-86): axes whose side distance is the minimum step together;
// float(mask) * delta is the select between delta and 0 * delta.
#define VT_DDA_STEP()                                                                   \
    {                                                                                   \
        bool mx, my, mz;                                                                \
        if (FAST) {                                                                     \
            const float tm = fminf(fminf(sx, sy), sz);                                  \
            mx = sx <= tm;                                                              \
            my = sy <= tm;                                                              \
            mz = sz <= tm;                                                              \
        } else {                                                                        \
            mx = sx <= gmin(sy, sz);                                                    \
            my = sy <= gmin(sz, sx);                                                    \
            mz = sz <= gmin(sx, sy);                                                    \
        }                                                                               \
        t_last = mx ? sx : (my ? sy : sz);                                              \
        sx += mx ? dx : zx;                                                             \
        sy += my ? dy : zy;                                                             \
        sz += mz ? dz : zz;                                                             \
        vx += mx ? stx : 0;                                                             \
        vy += my ? sty : 0;                                                             \
        vz += mz ? stz : 0;                                                             \
        lm = (mx ? 1u : 0u) | (my ? 2u : 0u) | (mz ? 4u : 0u);                          \
        ++steps;                                                                        \
    }
    // The brick loop tests its condition at the bottom (continue goes there) and a hit ends it
    // through the condition, so the loop has one exit.
    if ((steps < max_steps) & ((uint32_t)vx < (uint32_t)nx) & ((uint32_t)vy < (uint32_t)ny) &
        ((uint32_t)vz < (uint32_t)nz)) do {
        // entering a brick: the current voxel is inside the volume
        const int bx8 = vx & ~7, by8 = vy & ~7, bz8 = vz & ~7;
        // the voxel through which each stepping axis leaves this brick (or the volume); in the
        // FAST walk every axis steps
        const int ox = FAST ? min(bx8 + cx, lx) : (stx > 0 ? min(bx8 + 8, nx) : (stx < 0 ? bx8 - 1 : INT_MIN));
        const int oy = FAST ? min(by8 + cy, ly) : (sty > 0 ? min(by8 + 8, ny) : (sty < 0 ? by8 - 1 : INT_MIN));
        const int oz = FAST ? min(bz8 + cz, lz_) : (stz > 0 ? min(bz8 + 8, nz) : (stz < 0 ? bz8 - 1 : INT_MIN));
        entry = gld(V.brick_map + (uint32_t)(bx8 >> 3) + bdx * ((uint32_t)(by8 >> 3) + bdy * (uint32_t)(bz8 >> 3)));
        ctr.brick(!vt_brick_empty(entry));
        if (vt_brick_empty(entry)) {
            if (FAST) {
                // Every direction component is nonzero, so every axis steps towards its exit
                // voxel and each step moves at least one axis: the walk leaves the volume
                // before max_steps.  rx/ry/rz count the steps each axis still needs to leave
                // this brick; the loop stops before the step that leaves it, which is taken
                // below with its bookkeeping.  "s <= min3(sx, sy, sz)" is the step test.
                uint32_t rx = (uint32_t)__mul24(stx, ox - vx);
                uint32_t ry = (uint32_t)__mul24(sty, oy - vy);
                uint32_t rz = (uint32_t)__mul24(stz, oz - vz);
                // (the exit test at the bottom: the loop body is one block, without the exec
                // saves a test in its middle costs)
                float t = fminf(fminf(sx, sy), sz);
                bool mx = sx <= t, my = sy <= t, mz = sz <= t;
                rx -= mx ? 1u : 0u;
                ry -= my ? 1u : 0u;
                rz -= mz ? 1u : 0u;
                while (min(rx, min(ry, rz)) != 0u) {
                    sx += mx ? dx : zx;
                    sy += my ? dy : zy;
                    sz += mz ? dz : zz;
                    ++steps;
                    t = fminf(fminf(sx, sy), sz);
                    mx = sx <= t;
                    my = sy <= t;
                    mz = sz <= t;
                    rx -= mx ? 1u : 0u;
                    ry -= my ? 1u : 0u;
                    rz -= mz ? 1u : 0u;
                }
                const bool lmx = sx <= t, lmy = sy <= t, lmz = sz <= t;
                lm = (lmx ? 1u : 0u) | (lmy ? 2u : 0u) | (lmz ? 4u : 0u);
                t_last = lmx ? sx : (lmy ? sy : sz);
                sx += lmx ? dx : zx;
                sy += lmy ? dy : zy;
                sz += lmz ? dz : zz;
                ++steps;
                vx = ox - __mul24(stx, (int)rx);
                vy = oy - __mul24(sty, (int)ry);
                vz = oz - __mul24(stz, (int)rz);
                continue;
            }
            do {
                VT_DDA_STEP();
            } while (steps < max_steps && vx != ox && vy != oy && vz != oz);
            continue;
        }
        stage_mask(V.masks, entry, mslot);
        if (FAST) {
            // The same remaining-step counts; the current voxel's brick-local coordinates are
            // (ox, oy, oz) - (bx8, by8, bz8) - st * (rx, ry, rz).  The voxel a step moves into is
            // tested before the step's side update, so when the loop ends (hit, or the step that
            // leaves the brick) the sides are those before the last step, which gives its axes
            // and side exactly as the step test does.
            uint32_t rx = (uint32_t)__mul24(stx, ox - vx);
            uint32_t ry = (uint32_t)__mul24(sty, oy - vy);
            uint32_t rz = (uint32_t)__mul24(stz, oz - vz);
            const int bit0 = (ox - bx8) + 8 * (oy - by8), az = oz - bz8;
            int lz = vz - bz8;
            uint64_t slice = mask_slice(mslot, lz);
            uint32_t bit = (uint32_t)(vx - bx8) + 8u * (uint32_t)(vy - by8);
            if (((uint32_t)(slice >> bit) & 1u) == 0u) {
                float t;
                for (;;) {
                    t = fminf(fminf(sx, sy), sz);
                    const bool mx = sx <= t, my = sy <= t, mz = sz <= t;
                    rx -= mx ? 1u : 0u;
                    ry -= my ? 1u : 0u;
                    rz -= mz ? 1u : 0u;
                    if (min(rx, min(ry, rz)) == 0u) break;
                    ++steps;
                    const int nlz = az - __mul24(stz, (int)rz);
                    if (nlz != lz) {
                        lz = nlz;
                        slice = mask_slice(mslot, lz);
                    }
                    bit = (uint32_t)(bit0 - __mul24(stx, (int)rx) - __mul24(8 * sty, (int)ry));
                    if (((uint32_t)(slice >> bit) & 1u) != 0u) break;
                    sx += mx ? dx : zx;
                    sy += my ? dy : zy;
                    sz += mz ? dz : zz;
                }
                const bool lmx = sx <= t, lmy = sy <= t, lmz = sz <= t;
                lm = (lmx ? 1u : 0u) | (lmy ? 2u : 0u) | (lmz ? 4u : 0u);
                t_last = lmx ? sx : (lmy ? sy : sz);
                if (min(rx, min(ry, rz)) == 0u) {   // the step that leaves the brick
                    sx += lmx ? dx : zx;
                    sy += lmy ? dy : zy;
                    sz += lmz ? dz : zz;
                    ++steps;
                }
            }
            vx = ox - __mul24(stx, (int)rx);
            vy = oy - __mul24(sty, (int)ry);
            vz = oz - __mul24(stz, (int)rz);
            if (min(rx, min(ry, rz)) != 0u) {
                li = bit + 64u * (uint32_t)lz;
                hit = true;
            }
            continue;
        }
        int cur_z = -1;
        uint64_t slice = 0;
        for (;;) {
            const int lz = vz - bz8;
            if (lz != cur_z) {
                cur_z = lz;
                slice = mask_slice(mslot, lz);
            }
            const uint32_t bit = (uint32_t)(vx - bx8) + 8u * (uint32_t)(vy - by8);
            if ((slice >> bit) & 1ull) {
                li = bit + 64u * (uint32_t)lz;
                hit = true;
                break;
            }
            VT_DDA_STEP();
            if (!(steps < max_steps && vx != ox && vy != oy && vz != oz)) break;
        }
    } while (!hit & (steps < max_steps) & ((uint32_t)vx < (uint32_t)nx) & ((uint32_t)vy < (uint32_t)ny) &
             ((uint32_t)vz < (uint32_t)nz));
#undef VT_DDA_STEP
    w.vx = vx; w.vy = vy; w.vz = vz;
    w.steps = steps; w.entry = entry; w.li = li; w.t_last = t_last;
    w.hit = hit; w.lm = lm;
}

template <bool ST>
__device__ __forceinline__ void dda_run(const VolumeDesc& V, const float mpos[3], const int v0[3], V3 mdir, Dda& r,
                                        Ctr<ST>& ctr, const Lds& mslot) {
    DdaWalk w;
    w.vx = v0[0]; w.vy = v0[1]; w.vz = v0[2];
    const float len = sqrtf((mdir.x * mdir.x + mdir.y * mdir.y) + mdir.z * mdir.z);
    const float sgx = gsign(mdir.x), sgy = gsign(mdir.y), sgz = gsign(mdir.z);
    w.stx = (int)sgx; w.sty = (int)sgy; w.stz = (int)sgz;
    w.dx = fabsf(len / mdir.x); w.dy = fabsf(len / mdir.y); w.dz = fabsf(len / mdir.z);
    // float(mask) * delta of trace.frag:84 as a select: delta, or 0 * delta (0, or NaN for inf)
    w.zx = 0.0f * w.dx; w.zy = 0.0f * w.dy; w.zz = 0.0f * w.dz;
    w.sx = ((sgx * ((float)w.vx - mpos[0]) + sgx * 0.5f) + 0.5f) * w.dx;
    w.sy = ((sgy * ((float)w.vy - mpos[1]) + sgy * 0.5f) + 0.5f) * w.dy;
    w.sz = ((sgz * ((float)w.vz - mpos[2]) + sgz * 0.5f) + 0.5f) * w.dz;
    if (isfinite(w.dx) && isfinite(w.dy) && isfinite(w.dz) && isfinite(w.sx) && isfinite(w.sy) && isfinite(w.sz))
        dda_walk<ST, true>(V, w, ctr, mslot);
    else
        dda_walk<ST, false>(V, w, ctr, mslot);
    ctr.steps(w.steps);
    const uint32_t mask = w.lm;
    r.v[0] = w.vx; r.v[1] = w.vy; r.v[2] = w.vz;
    r.steps = w.steps;
    r.mask = w.hit && w.steps > 0 ? mask : 0u;
    r.t_last = w.t_last;
    r.q[0] = mpos[0]; r.q[1] = mpos[1]; r.q[2] = mpos[2];
    r.entry = w.entry;
    r.li = w.li;
    r.hit = w.hit;
}

// Secondary-ray walk (no reference counterpart; this project's rule).  From position q in
// voxel v the ray crosses, one box per step, the largest box known to be empty that holds v:
//   * in an empty brick, the cube of side K_o anchored at it towards the ray's octant o when
//     K_o - 1 > r (its far faces are then beyond those of the centred cube on every axis, so
//     its exit is further along the ray), else the cube of radius r centred on the brick (both
//     from the brick map entry);
//   * in a non-empty brick, the brick's 4^3 octant holding v if its 8 blocks are empty, else the
//     2^3 block holding v if it is empty (block mask), else v itself -- unless v is occupied
//     (the LDS-staged occupancy mask), which ends the walk with a hit.
// A step moves q to the box's nearest exit plane along the ray (ties: every axis at that
// distance): exit axes step into the voxel beyond the plane, the other axes take floor(q)
// clamped to the box.  Every step runs the same instructions whatever the box (values chosen
// by selects), so lanes of a wave walking empty space and lanes walking voxels share each
// iteration.
// One walk, held in registers between the pool rounds of k_secondary: box_walk_begin, then
// box_walk_step until it returns false (the walk left the volume, hit a voxel, or reached the
// guard of 2 (dx + dy + dz) + 64 steps).
struct BoxWalk {
    const uint32_t* brick_map;
    const uint32_t* masks;
    const uint64_t* blocks;
    int n[3];
    uint32_t bdx, bdy, guard, oct;
    float u[3], rinv[3];
    int st[3];
    int v[3];
    float q[3];
    uint32_t steps, mask, entry, cur, li;
    uint64_t blk;
    bool hit;
};

__device__ __forceinline__ void box_walk_begin(const VolumeDesc& V, const float p[3], const int v0[3], V3 m,
                                               BoxWalk& W) {
    const float len = sqrtf((m.x * m.x + m.y * m.y) + m.z * m.z);
    const float mm[3] = {m.x, m.y, m.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        W.st[a] = mm[a] > 0.0f ? 1 : (mm[a] < 0.0f ? -1 : 0);
        W.u[a] = mm[a] / len;
        W.rinv[a] = len / mm[a];
        W.v[a] = v0[a];
        W.q[a] = p[a];
        W.n[a] = (int)V.dim[a];
    }
    W.oct = (W.st[0] < 0 ? 1u : 0u) | (W.st[1] < 0 ? 2u : 0u) | (W.st[2] < 0 ? 4u : 0u);
    W.brick_map = V.brick_map;
    W.masks = V.masks;
    W.blocks = V.blocks;
    W.bdx = V.bdim[0];
    W.bdy = V.bdim[1];
    W.guard = 2u * (V.dim[0] + V.dim[1] + V.dim[2]) + 64u;
    W.steps = 0;
    W.mask = 0;
    W.entry = VT_EMPTY;
    W.cur = VT_EMPTY;
    W.li = 0;
    W.blk = 0;
    W.hit = false;
}

template <bool ST>
__device__ __forceinline__ bool box_walk_step(BoxWalk& W, Ctr<ST>& ctr, const Lds& mslot) {
    // One exit: the step is computed for every lane and committed only where the walk goes on
    // (inside the volume, under the guard, v not occupied); only the brick load is guarded.
    const int* v = W.v;
    const bool inb = (W.steps < W.guard) & ((uint32_t)v[0] < (uint32_t)W.n[0]) &
                     ((uint32_t)v[1] < (uint32_t)W.n[1]) & ((uint32_t)v[2] < (uint32_t)W.n[2]);
    const int b8[3] = {v[0] & ~7, v[1] & ~7, v[2] & ~7};
    const uint32_t bi = (uint32_t)(b8[0] >> 3) + W.bdx * ((uint32_t)(b8[1] >> 3) + W.bdy * (uint32_t)(b8[2] >> 3));
    if (inb & (bi != W.cur)) {
        W.cur = bi;
        W.entry = gld(W.brick_map + bi);
        ctr.brick(!vt_brick_empty(W.entry));
        if (!vt_brick_empty(W.entry)) {
            W.blk = gld64(W.blocks + W.entry);
            stage_mask(W.masks, W.entry, mslot);
        }
    }
    const uint32_t entry = W.entry;
    const uint64_t blk = W.blk;
    const bool emp = vt_brick_empty(entry);
    const int lx = v[0] - b8[0], ly = v[1] - b8[1], lz = v[2] - b8[2];
    // (32-bit halves: the block mask's half lz >> 2, the occupancy word of row group ly >> 2 of
    // slice lz; an empty octant's blocks are all empty, so sz = 1 + blk_empty + 2 oct_empty)
    const uint32_t bh = lz >= 4 ? (uint32_t)(blk >> 32) : (uint32_t)blk;
    const bool oct_empty = ((bh >> (uint32_t)(2 * (lx >> 2) + 8 * (ly >> 2))) & 0x330033u) == 0u;
    const bool blk_empty = ((bh >> (uint32_t)((lx >> 1) + 4 * (ly >> 1) + 16 * ((lz >> 1) & 1))) & 1u) == 0u;
    const int sz = 1 + (int)blk_empty + 2 * (int)oct_empty;
    const uint32_t word = mslot.at(2 * lz + (ly >> 2));   // (lx, ly, lz in 0..7 for any v)
    const bool hit = inb & !emp & (sz == 1) & (((word >> (uint32_t)(lx + 8 * (ly & 3))) & 1u) != 0u);
    const bool go = inb & !hit;
    const int rad = 8 * (int)vt_brick_radius(entry);
    const int k8 = 8 * (int)vt_brick_cube(entry, W.oct);
    const bool anc = k8 - 8 > rad;
    // The box's side is the same on every axis (anchored cube k8, centred cube 8 + 2 rad, or
    // sz); in an empty brick its low face is b8 + an offset that depends only on whether the
    // axis steps down (anchored: 8 - k8, else 0; centred: -rad either way).
    const int side = emp ? (anc ? k8 : 8 + 2 * rad) : sz;
    const int off_up = anc ? 0 : -rad, off_down = anc ? 8 - k8 : -rad;
    int lo[3], hi[3];
    float tt[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int st = W.st[a];
        lo[a] = emp ? b8[a] + (st < 0 ? off_down : off_up) : (v[a] & -sz);
        hi[a] = lo[a] + side;
        const int e = st > 0 ? hi[a] : lo[a];
        const float t = ((float)e - W.q[a]) * W.rinv[a];
        tt[a] = st != 0 ? t : INFINITY;
    }
    const float tmin = gmin(gmin(gmin(INFINITY, tt[0]), tt[1]), tt[2]);
    uint32_t mk = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int st = W.st[a];
        const int l = lo[a], h = hi[a];
        const bool ex = (st != 0) & (tt[a] <= tmin);
        const float x = W.q[a] + tmin * W.u[a];
        const int f0 = (int)floorf(x);
        const int f1 = f0 < l ? l : f0;
        const int f = f1 > h - 1 ? h - 1 : f1;
        const int pe = st > 0 ? h : l;
        const float qe = (float)pe;
        const int ve = st > 0 ? h : l - 1;
        const float qn = ex ? qe : x;
        const int vn = ex ? ve : f;
        W.q[a] = go ? qn : W.q[a];
        W.v[a] = go ? vn : W.v[a];
        mk |= ex ? 1u << a : 0u;
    }
    W.mask = go ? mk : W.mask;
    W.steps += go ? 1u : 0u;
    W.li = hit ? (uint32_t)lx + 8u * (uint32_t)ly + 64u * (uint32_t)lz : W.li;
    W.hit = hit;
    return go;
}

struct HitG {
    bool hit;
    uint32_t inst;
    float tn;
    Dda d;
    float mpos[3];
    V3 mdir;
    uint8_t face;
};

__device__ __forceinline__ uint8_t face_from(int axis, float dir_a) { return (uint8_t)(2 * axis + (dir_a > 0.0f ? 0 : 1)); }
__device__ __forceinline__ int lowest_axis(uint32_t m) { return (m & 1u) ? 0 : ((m & 2u) ? 1 : 2); }

struct SlabR {
    float tn, tf;
    int axis;
    bool ok;
};
// instance slab in model space: exact IEEE divisions (defines the entry point)
__device__ __forceinline__ SlabR slab_unit(V3 o, V3 d) {
    SlabR s{-INFINITY, INFINITY, 0, true};
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
    // (both cases computed and chosen by selects; the quotients of a parallel axis are unused)
    for (int a = 0; a < 3; ++a) {
        const float t1 = (-0.5f - oo[a]) / dd[a];
        const float t2 = (0.5f - oo[a]) / dd[a];
        const bool mv = dd[a] != 0.0f;
        const float tn = mv ? gmin(t1, t2) : -INFINITY;
        const float tf = mv ? gmax(t1, t2) : INFINITY;
        s.ok = s.ok & (mv | !((oo[a] < -0.5f) | (oo[a] > 0.5f)));
        const bool nt = tn > s.tn;
        s.tn = nt ? tn : s.tn;
        s.axis = nt ? a : s.axis;
        s.tf = tf < s.tf ? tf : s.tf;
    }
    s.ok = s.ok && s.tn <= s.tf;
    return s;
}
// LBVH node slab: reciprocal direction, (bound - o) * inv_d
__device__ __forceinline__ bool slab_node(V3 o, V3 d, V3 id, const BvhNode& N, float* tn_out, float* tf_out) {
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z}, ii[3] = {id.x, id.y, id.z};
    float tn = -INFINITY, tf = INFINITY;
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float t1 = (N.lo[a] - oo[a]) * ii[a];
        const float t2 = (N.hi[a] - oo[a]) * ii[a];
        const bool mv = dd[a] != 0.0f;
        const float n_ = mv ? gmin(t1, t2) : -INFINITY;
        const float f_ = mv ? gmax(t1, t2) : INFINITY;
        ok = ok & (mv | !((oo[a] < N.lo[a]) | (oo[a] > N.hi[a])));
        tn = n_ > tn ? n_ : tn;
        tf = f_ < tf ? f_ : tf;
    }
    *tn_out = tn;
    *tf_out = tf;
    return ok & (tn <= tf);
}

// Primary ray: enter instance i from outside along (O, D), rebuilding the fragment inputs of
// trace.frag (proxy point, clip test, camera ray through it in model space), and walk it.
template <bool ST>
__device__ bool enter_instance(const SceneGPU& S, const FrameGPU& F, uint32_t i, V3 O, V3 D, HitG& h, Ctr<ST>& ctr,
                               const Lds& mslot) {
    const InstanceGPU* I = &S.inst[i];
    if (!gld(&I->valid)) return false;
    float inv[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) inv[k] = gld(&I->inv[k]);
    const V3 o = xf34(inv, O, 1.0f);
    const V3 d = xf34(inv, D, 0.0f);
    SlabR s = slab_unit(o, d);
    if (!s.ok || !(s.tn > 0.0f)) return false;
    float mp[3];
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float face = dd[a] > 0.0f ? -0.5f : 0.5f;
        const float x = oo[a] + s.tn * dd[a];
        const float xc = x < -0.5f ? -0.5f : (x > 0.5f ? 0.5f : x);
        mp[a] = a == s.axis ? face : xc;
    }
    V3 mdir;
    {
        float M[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) M[k] = gld(&I->M[k]);
        const V3 w = xf34(M, V3{mp[0], mp[1], mp[2]}, 1.0f);
        const float4 sp = mv4(F.PC, make_float4(w.x, w.y, w.z, 1.0f));
        if (!(sp.z >= 0.0f && sp.z <= sp.w)) return false;
        const float4 q = mv4(F.AB, sp);
        const V3 dir = normalize3(V3{q.x, q.y, q.z});
        mdir = xf34(inv, dir, 0.0f);
    }
    ctr.instance();
    const VolumeDesc V = inst_vol(S, I);
    float mpos[3];
    int v0[3];
    const float size[3] = {gld(&I->size[0]), gld(&I->size[1]), gld(&I->size[2])};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        mpos[a] = (mp[a] + 0.5f) * size[a];
        v0[a] = (int)floorf(gmin(mpos[a], size[a] - 1.0f));
    }
    Dda r;
    dda_run<ST>(V, mpos, v0, mdir, r, ctr, mslot);
    if (!r.hit) return false;
    h.hit = true;
    h.inst = i;
    h.tn = s.tn;
    h.d = r;
    h.mpos[0] = mpos[0]; h.mpos[1] = mpos[1]; h.mpos[2] = mpos[2];
    h.mdir = mdir;
    if (r.steps == 0) {
        h.face = face_from(s.axis, dd[s.axis]);
    } else {
        int a = lowest_axis(r.mask);
        h.face = face_from(a, comp(mdir, a));
    }
    return true;
}

__device__ __forceinline__ bool better(const HitG& c, const HitG& b) {
    return c.tn < b.tn || (c.tn == b.tn && c.inst < b.inst);
}
__device__ __forceinline__ float prune_bound(float t) { return t + fabsf(t) * 0.0009765625f + 0.0009765625f; }

// LBVH nodes a workgroup stages in LDS when its launch traces one scene of at most 64
// instances (a node fetch is then an LDS read instead of a dependent global load).
constexpr uint32_t kLdsNodes = 127;

// A workgroup's LDS copy of the tree (explicit LDS address space, so node reads are ds_reads).
struct TreeLds {
    const __attribute__((address_space(3))) BvhNode* nodes;
    bool on;
};

__device__ __forceinline__ BvhNode lds_node(const TreeLds& t, uint32_t i) {
    typedef const __attribute__((address_space(3))) uint32_t LdsU32;
    LdsU32* p = (LdsU32*)(t.nodes + i);
    BvhNode n;
    n.lo[0] = __uint_as_float(p[0]); n.lo[1] = __uint_as_float(p[1]); n.lo[2] = __uint_as_float(p[2]);
    n.hi[0] = __uint_as_float(p[3]); n.hi[1] = __uint_as_float(p[4]); n.hi[2] = __uint_as_float(p[5]);
    n.left = p[6];
    n.right = p[7];
    return n;
}

// Stages the batch's tree into `tree` when it has one request whose tree fits (on = false:
// traversal reads the nodes from global memory).  Block-uniform.
__device__ __forceinline__ TreeLds stage_tree(const ReqGPU* reqs, const BatchGPU* B, BvhNode* tree) {
    TreeLds t{(const __attribute__((address_space(3))) BvhNode*)(tree), false};
    if (B->nreq != 1) return t;
    const uint32_t ninst = reqs[0].S.ninst;
    const uint32_t nn = ninst ? 2u * ninst - 1u : 0u;
    if (nn == 0 || nn > kLdsNodes) return t;
    const BvhNode* src = reqs[0].S.nodes;
    for (uint32_t i = threadIdx.x; i < nn; i += blockDim.x) tree[i] = gld(src + i);
    __syncthreads();
    t.on = true;
    return t;
}

// Primary ray: closest hit over all instances through the LBVH (nodes from `tree` in LDS when
// given).  Secondary rays have their own traversal (in k_secondary).
template <bool ST, bool DEEP>
__device__ HitG query(const SceneGPU& S, const FrameGPU& F, V3 O, V3 D, Ctr<ST>& ctr, const Lds& stk_lds,
                      const Lds& mslot, TreeLds tree) {
    HitG best;
    best.hit = false;
    best.inst = VT_EMPTY;
    best.tn = INFINITY;
    if (S.ninst == 0) return best;
    const V3 inv_d{1.0f / D.x, 1.0f / D.y, 1.0f / D.z};
    TravStack<DEEP> stk;
    stk.lds = stk_lds;
    stk.sp = 0;
    stk.push(0u);
    while (stk.sp > 0) {
        const uint32_t ni = stk.pop();
        const BvhNode N = tree.on ? lds_node(tree, ni) : gld(S.nodes + ni);
        ctr.node();
        float tn, tf;
        if (!slab_node(O, D, inv_d, N, &tn, &tf) || tf < 0.0f) continue;
        if (best.hit && tn > prune_bound(best.tn)) continue;
        if (N.right == VT_EMPTY) {
            HitG h;
            if (enter_instance<ST>(S, F, N.left, O, D, h, ctr, mslot) && better(h, best)) best = h;
        } else {
            stk.push(N.right);
            stk.push(N.left);
        }
    }
    return best;
}

// Surface of a hit, as shading needs it (registers only).
struct SurfaceRec {
    float p[3];        // hit point, voxel units of the hit instance
    uint32_t inst;
    float X[3];        // world position
    uint32_t path;
    float n[3];        // world unit normal
    uint32_t rgba;     // texel colour (RGBA8)
    int vstart[3];     // empty neighbour voxel on the face side
    uint32_t face;
};

// Queue element: a shaded surface with its two secondary rays (48 B, written once by the launch
// that found the surface, read by the next depth's bounce ray and sun ray).  Its world
// position and normal are rebuilt from (instance, p, face) with
// the operations surface_of uses, so they are the same bits; its bounce direction is drawn
// from `rng` when the bounce ray starts (sun rays draw nothing, so every path consumes its
// random numbers in the same order as when the direction was drawn at the hit).
struct QueueRec {
    float p[3];        // hit point, voxel units of the hit instance
    float c[3];        // direct sun light added if the sun ray is clear
    uint32_t inst;
    uint32_t path;
    uint16_t v[3];     // hit voxel (the start voxel is its neighbour across the face)
    uint8_t face;
    uint8_t pad;
    uint32_t rgba;     // texel colour: a first bounce's throughput is its albedo
    uint32_t rng;      // RNG state at the surface
};
static_assert(sizeof(QueueRec) == 48, "queue record is 48 bytes");

__device__ __forceinline__ void put_rec(QueueRec* dst, const SurfaceRec& sf, const float4& contrib, uint32_t path,
                                        uint32_t rng) {
    const int fa = (int)(sf.face >> 1);
    const int off = (sf.face & 1u) == 0u ? -1 : 1;
    const uint32_t v0 = (uint32_t)(sf.vstart[0] - (fa == 0 ? off : 0));
    const uint32_t v1 = (uint32_t)(sf.vstart[1] - (fa == 1 ? off : 0));
    const uint32_t v2 = (uint32_t)(sf.vstart[2] - (fa == 2 ? off : 0));
    uint4* d = reinterpret_cast<uint4*>(dst);
    d[0] = make_uint4(__float_as_uint(sf.p[0]), __float_as_uint(sf.p[1]), __float_as_uint(sf.p[2]),
                      __float_as_uint(contrib.x));
    d[1] = make_uint4(__float_as_uint(contrib.y), __float_as_uint(contrib.z), sf.inst, path);
    d[2] = make_uint4((v0 & 0xFFFFu) | (v1 << 16), (v2 & 0xFFFFu) | (sf.face << 16), sf.rgba, rng);
}

__device__ __forceinline__ QueueRec get_rec(const QueueRec* src) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
    const uint4 a = gld(s), b = gld(s + 1), c = gld(s + 2);
    QueueRec r;
    r.p[0] = __uint_as_float(a.x); r.p[1] = __uint_as_float(a.y); r.p[2] = __uint_as_float(a.z);
    r.c[0] = __uint_as_float(a.w); r.c[1] = __uint_as_float(b.x); r.c[2] = __uint_as_float(b.y);
    r.inst = b.z;
    r.path = b.w;
    r.v[0] = (uint16_t)(c.x & 0xFFFFu); r.v[1] = (uint16_t)(c.x >> 16); r.v[2] = (uint16_t)(c.y & 0xFFFFu);
    r.face = (uint8_t)(c.y >> 16);
    r.pad = 0;
    r.rgba = c.z;
    r.rng = c.w;
    return r;
}

// World position and normal of the surface point p (voxel units) on face `face` of instance I.
__device__ __forceinline__ void surface_frame(const InstanceGPU* I, const float p[3], uint32_t face, V3& X, V3& n) {
    float M[12], inv[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) { M[k] = gld(&I->M[k]); inv[k] = gld(&I->inv[k]); }
    const V3 mp{p[0] / gld(&I->size[0]) - 0.5f, p[1] / gld(&I->size[1]) - 0.5f, p[2] / gld(&I->size[2]) - 0.5f};
    X = xf34(M, mp, 1.0f);
    const int fa = (int)(face >> 1);
    const float nm = (face & 1u) == 0u ? -1.0f : 1.0f;
    // row fa of inverse(M): inv[c*3 + fa] (inverse-transpose applied to the model normal)
    n = normalize3(V3{inv[0 * 3 + fa] * nm, inv[1 * 3 + fa] * nm, inv[2 * 3 + fa] * nm});
}

__device__ void surface_of(const SceneGPU& S, const HitG& h, SurfaceRec& sf) {
    const InstanceGPU* I = &S.inst[h.inst];
    const float len = sqrtf((h.mdir.x * h.mdir.x + h.mdir.y * h.mdir.y) + h.mdir.z * h.mdir.z);
    const int fa = h.face >> 1;
    const bool neg = (h.face & 1) == 0;
    const float md[3] = {h.mdir.x, h.mdir.y, h.mdir.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float u = md[a] / len;
        float x = h.d.q[a] + h.d.t_last * u;
        const float lo = (float)h.d.v[a], hi = (float)(h.d.v[a] + 1);
        if (a == fa) x = neg ? lo : hi;
        else x = x < lo ? lo : (x > hi ? hi : x);
        sf.p[a] = x;
        sf.vstart[a] = h.d.v[a] + (a == fa ? (neg ? -1 : 1) : 0);
    }
    V3 X, n;
    surface_frame(I, sf.p, h.face, X, n);
    sf.X[0] = X.x; sf.X[1] = X.y; sf.X[2] = X.z;
    sf.n[0] = n.x; sf.n[1] = n.y; sf.n[2] = n.z;
    sf.inst = h.inst;
    sf.face = h.face;
}

__device__ void cosine_dir(Rng& rng, V3 n, V3& w) {
    for (int t = 0; t < 16; ++t) {
        const float u = 2.0f * rng.next() - 1.0f;
        const float v = 2.0f * rng.next() - 1.0f;
        const float s = u * u + v * v;
        if (s >= 1.0f || s == 0.0f) continue;
        const float q = sqrtf(1.0f - s);
        const V3 r{2.0f * u * q, 2.0f * v * q, 1.0f - 2.0f * s};
        const V3 a{n.x + r.x, n.y + r.y, n.z + r.z};
        const float dd = (a.x * a.x + a.y * a.y) + a.z * a.z;
        const float cn = (a.x * n.x + a.y * n.y) + a.z * n.z;
        if (!(dd > 1e-6f) || !(cn > 1e-3f * sqrtf(dd))) continue;
        w = normalize3(a);
        return;
    }
    w = n;
}

__device__ __forceinline__ uint32_t hit_rgba(const SceneGPU& S, const HitG& h, uint8_t* pal) {
    const VolumeDesc V = inst_vol(S, &S.inst[h.inst]);
    if (V.fmt == 0) {
        const uint32_t idx = gld8(V.payload + (size_t)h.d.entry * 512 + h.d.li);
        *pal = (uint8_t)idx;
        return gld(V.palette + idx);
    }
    *pal = 0;
    return gld(reinterpret_cast<const uint32_t*>(V.payload) + (size_t)h.d.entry * 512 + h.d.li);
}

__device__ __forceinline__ uint32_t pixel_seed(uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t sample) {
    return pcg_hash(pixel ^ pcg_hash(sample ^ pcg_hash(frame ^ pcg_hash(seed))));
}

__device__ __forceinline__ V3 camera_ray(const FrameGPU& F, float sx, float sy) {
    const float nxd = (2.0f * sx) / F.W - 1.0f;
    const float nyd = (2.0f * sy) / F.Hv - 1.0f;
    const V3 vv{nxd * F.ip00, nyd * F.ip11, -1.0f};
    return V3{(F.R[0] * vv.x + F.R[3] * vv.y) + F.R[6] * vv.z, (F.R[1] * vv.x + F.R[4] * vv.y) + F.R[7] * vv.z,
              (F.R[2] * vv.x + F.R[5] * vv.y) + F.R[8] * vv.z};
}

// ---- persistent work distribution -----------------------------------------------------
__host__ __device__ __forceinline__ uint32_t* slot_counters(uint32_t* counters, uint32_t slot) {
    return counters + slot * kParts * kPartStride;
}

// Each wave takes 64 consecutive items of a part from that part's counter (fetch + p *
// kPartStride) until the part is exhausted, starting with part blockIdx.x % kParts and then
// taking the others in turn.  range(p, &first, &count): part p's items are first + [0, count);
// body(p, item).  `gone` (LDS, zero at start) marks the parts the block has seen exhausted, so
// only one wave per block spends an atomic on finding that out.
template <typename Range, typename Body>
__device__ __forceinline__ void for_each_item(uint32_t* fetch, uint32_t* gone, Range&& range, Body&& body) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t p0 = blockIdx.x % kParts;
    for (uint32_t k = 0; k < kParts; ++k) {
        const uint32_t p = (p0 + k) % kParts;
        uint32_t first = 0, count = 0;
        range(p, &first, &count);
        const uint32_t tiles = (count + 63u) / 64u;
        for (;;) {
            if (__hip_atomic_load(gone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & (1u << p)) break;
            uint32_t t = tiles;
            if (tiles > 0 && lane == 0) t = atomicAdd(fetch + p * kPartStride, 1u);
            t = __shfl(t, 0, 64);
            if (t >= tiles) {
                if (lane == 0) __hip_atomic_fetch_or(gone, 1u << p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                break;
            }
            const uint32_t i = t * 64u + lane;
            if (i < count) body(p, first + i);
        }
    }
}

// Items [0, nwork) in kParts contiguous ranges of whole 64-item tiles.
__device__ __forceinline__ void static_part(uint32_t nwork, uint32_t p, uint32_t* first, uint32_t* count) {
    const uint32_t tiles = (nwork + 63u) / 64u;
    const uint32_t t0 = (uint32_t)((uint64_t)tiles * p / kParts), t1 = (uint32_t)((uint64_t)tiles * (p + 1u) / kParts);
    *first = t0 * 64u;
    *count = min(t1 * 64u, nwork) - min(t0 * 64u, nwork);
}

// Slots used in segment p of a queue (its counter can pass seg_cap, see append_slot).
__device__ __forceinline__ uint32_t queue_count(uint32_t* q, uint32_t p, uint32_t cap) {
    return min(q[p * kPartStride], cap);
}

// Wave-aggregated append to queue counters q (segments of `cap` slots): the wave's appending
// lanes take consecutive slots of segment blockIdx.x % kParts in lane order with one atomic;
// what does not fit there goes on to the next segments (a counter past cap marks its segment
// full).  The segments together hold every path, so the loop always places all lanes.  k0 > 0
// starts at segment (blockIdx.x + k0) % kParts (the earlier ones known to be full).
__device__ __forceinline__ uint32_t append_slot(uint32_t* q, uint32_t cap, uint32_t k0 = 0) {
    const uint64_t m = __ballot(1);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
    const uint32_t n = (uint32_t)__popcll(m);
    const uint64_t lower = lane == 0 ? 0ull : (m & (~0ull >> (64u - lane)));
    const uint32_t rank = (uint32_t)__popcll(lower);
    uint32_t slot = 0, placed = 0, p = (blockIdx.x + k0) % kParts;
    for (uint32_t k = k0; k < kParts && placed < n; ++k, p = (p + 1u) % kParts) {
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(q + p * kPartStride, n - placed);
        base = __shfl(base, (int)leader, 64);
        const uint32_t take = base < cap ? min(cap - base, n - placed) : 0u;
        if (rank >= placed && rank < placed + take) slot = p * cap + base + (rank - placed);
        placed += take;
    }
    return slot;
}

// Two appends in one round trip: every active lane to queue qa (slot *sa), the lanes with b also
// to queue qb (slot *sb); the leader issues both atomics before either result is used.  Lanes
// that do not fit segment blockIdx.x % kParts go on through append_slot from the next segment.
__device__ __forceinline__ void append_pair(uint32_t* qa, uint32_t* qb, bool b, uint32_t cap, uint32_t* sa,
                                            uint32_t* sb) {
    const uint64_t ma = __ballot(1), mb = __ballot(b);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)ma) - 1u;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64u - lane));
    const uint32_t na = (uint32_t)__popcll(ma), nb = (uint32_t)__popcll(mb);
    const uint32_t ra = (uint32_t)__popcll(ma & below), rb = (uint32_t)__popcll(mb & below);
    const uint32_t p = blockIdx.x % kParts;
    uint32_t ba = 0, bb = 0;
    if (lane == leader) {
        ba = atomicAdd(qa + p * kPartStride, na);
        if (nb > 0u) bb = atomicAdd(qb + p * kPartStride, nb);
    }
    ba = __shfl(ba, (int)leader, 64);
    bb = __shfl(bb, (int)leader, 64);
    const uint32_t ta = ba < cap ? min(cap - ba, na) : 0u;
    const uint32_t tb = bb < cap ? min(cap - bb, nb) : 0u;
    *sa = p * cap + ba + ra;
    *sb = p * cap + bb + rb;
    if (ra >= ta) *sa = append_slot(qa, cap, 1u);
    if (b && rb >= tb) *sb = append_slot(qb, cap, 1u);
}

// Wavefront buffers (device pointers), see WaveBuffers in vt_internal.h.
struct WF {
    float4* L;            // per path: sky terms; w = 1.0f bits (1u) if the primary ray hit
    float4* Ls;           // per path with a primary hit: sun terms (sun rays only); w bits 1u if the
                          // primary vertex is shadowed (facing away, or its sun ray blocked)
    float4* br;           // per path past its first bounce: throughput (xyz)
    QueueRec* rec[2];     // per queue slot
    uint32_t* sun[2];     // sun list: the queue slots whose surfaces face the sun (one sun ray each)
    uint32_t* counters;   // kSlots x kParts counters (queue sizes and fetch counters)
    uint32_t seg_cap;     // slots per queue segment
    float4* gsamp;        // optional: per path (albedo rgb, distance), (normal, hit)
};

// Shade a surface at `depth`: direct-light contribution for the sun ray and, if the path
// continues, the cosine-sampled bounce direction (throughput times albedo).
__device__ __forceinline__ void shade(const FrameGPU& F, const float* srgb_dec, const SurfaceRec& sf, uint32_t depth,
                                      uint32_t max_b, float beta[3], float4& contrib, bool& no_sun) {
    const float rho[3] = {srgb_dec[sf.rgba & 255u], srgb_dec[(sf.rgba >> 8) & 255u], srgb_dec[(sf.rgba >> 16) & 255u]};
    const float cosl = (sf.n[0] * F.sun[0] + sf.n[1] * F.sun[1]) + sf.n[2] * F.sun[2];
    contrib = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    no_sun = !(cosl > 0.0f);
    if (!no_sun) {
        const float c = cosl * 0.3183098861837907f;
        contrib = make_float4(beta[0] * (rho[0] * (F.sun_rad[0] * c)), beta[1] * (rho[1] * (F.sun_rad[1] * c)),
                              beta[2] * (rho[2] * (F.sun_rad[2] * c)), 1.0f);
    }
    // a bounce follows: its throughput (the direction is drawn when the bounce ray starts)
    if (depth < max_b)
        for (int k = 0; k < 3; ++k) beta[k] *= rho[k];
}

// Work index -> (pixel, sample) for mode 0: waves cover an 8 x 8 pixel tile.
__device__ __forceinline__ bool work_pixel(const FrameGPU& F, uint32_t w, uint32_t spp, uint32_t* px, uint32_t* py,
                                           uint32_t* s) {
    const uint32_t tiles_x = (F.width + 7u) / 8u;
    const uint32_t per_tile = 64u * spp;
    const uint32_t tile = w / per_tile, rem = w % per_tile;
    *s = rem / 64u;
    const uint32_t pt = rem % 64u;
    *px = (tile % tiles_x) * 8u + (pt & 7u);
    *py = (tile / tiles_x) * 8u + (pt >> 3);
    return *px < F.width && *py < F.height;
}

// ---- batched wavefront ----------------------------------------------------------------------
// Request k of a batch owns the global work items [work_base[k], work_base[k+1]); work item
// w is pixel sample (pixel, s) of that request and also the index of its path state, so a
// wave's 64 consecutive items are one pixel tile at one sample of one request.

// request owning global index x in a prefix table base[0..n]
__device__ __forceinline__ uint32_t find_req(const uint32_t* base, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Work items of a request's launch: the tiles of 64 pixels (tw x th) in its rows of tiles
// [ty0, ty1), row-major, each tile's 64 * ns items its pixels' samples [s0, s1), a pixel's ns
// samples side by side; the pixels of an 8 x 8 tile in Morton order, so a wave's 64 items cover
// a compact block of 64 / ns pixels (8 x 4 at 2 samples, 4 x 4 at 4) whose rays stay close
// together.
struct TileMap {
    uint32_t tw, th, tiles_x, ns, t0;   // t0: first tile of the launch (row ty0)
    __device__ __forceinline__ TileMap(const ReqGPU& R) {
        tw = R.tile_w;
        th = 64u / tw;
        uint32_t ty0, ty1;
        vt_req_tiles(R, &ty0, &ty1, &tiles_x);
        t0 = ty0 * tiles_x;
        ns = R.s1 - R.s0;
    }
    // Morton order of an 8 x 8 tile: x in bits 0, 2, 4 and y in bits 1, 3, 5
    __device__ __forceinline__ void tile_xy(uint32_t pt, uint32_t* x, uint32_t* y) const {
        if (tw == 8u) {
            *x = (pt & 1u) | ((pt >> 1) & 2u) | ((pt >> 2) & 4u);
            *y = ((pt >> 1) & 1u) | ((pt >> 2) & 2u) | ((pt >> 3) & 4u);
        } else {
            *x = pt % tw;
            *y = pt / tw;
        }
    }
    __device__ __forceinline__ uint32_t tile_pt(uint32_t x, uint32_t y) const {
        if (tw == 8u)
            return (x & 1u) | ((y & 1u) << 1) | ((x & 2u) << 1) | ((y & 2u) << 2) | ((x & 4u) << 2) | ((y & 4u) << 3);
        return y * tw + x;
    }
    // local work index -> pixel and sample
    __device__ __forceinline__ void pixel(uint32_t lw, uint32_t s0, uint32_t* px, uint32_t* py, uint32_t* s) const {
        const uint32_t per_tile = 64u * ns;
        const uint32_t tile = t0 + lw / per_tile, rem = lw % per_tile;
        *s = s0 + rem % ns;
        uint32_t x, y;
        tile_xy(rem / ns, &x, &y);
        *px = (tile % tiles_x) * tw + x;
        *py = (tile / tiles_x) * th + y;
    }
    // pixel and sample -> local work index
    __device__ __forceinline__ uint32_t index(uint32_t px, uint32_t py, uint32_t s, uint32_t s0) const {
        const uint32_t tile = (py / th) * tiles_x + px / tw - t0;
        return tile * (64u * ns) + tile_pt(px % tw, py % th) * ns + (s - s0);
    }
};

template <bool ST, bool DEEP>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_primary(const ReqGPU* __restrict__ reqs, const BatchGPU* __restrict__ B,
                                                    const float* __restrict__ srgb_dec, WF wf,
                                                    unsigned long long* stats) {
    __shared__ uint32_t smem[(kStackLds + 16) * kBlock];
    __shared__ BvhNode tree_lds[kLdsNodes];
    const Lds stk{&smem[threadIdx.x]}, mslot{&smem[kStackLds * kBlock + threadIdx.x]};
    __shared__ uint32_t gone;
    if (threadIdx.x == 0) gone = 0u;
    const TreeLds tree = stage_tree(reqs, B, tree_lds);
    __syncthreads();
    Ctr<ST> ctr;
    const uint32_t nreq = B->nreq;
    const uint32_t nwork = B->work_base[nreq];
    for_each_item(slot_counters(wf.counters, 2), &gone,
                  [&](uint32_t p, uint32_t* first, uint32_t* count) { static_part(nwork, p, first, count); },
                  [&](uint32_t, uint32_t w) {
        // the 64 items of a wave share a request (work ranges are 64-aligned)
        const uint32_t r = __builtin_amdgcn_readfirstlane(find_req(B->work_base, nreq, w));
        const ReqGPU& Rq = reqs[r];
        const FrameGPU& F = Rq.F;
        const SceneGPU S = Rq.S;
        const TileMap tm(Rq);
        uint32_t px, py, smp;
        tm.pixel(w - B->work_base[r], Rq.s0, &px, &py, &smp);
        if (px >= F.width || py >= F.height) return;
        const uint32_t pix = py * F.width + px;
        const uint32_t max_b = F.mode == 2 ? F.max_bounces : 0u;
        Rng rng{pixel_seed(F.seed, F.frame, pix, smp)};
        float jx = 0.5f, jy = 0.5f;
        if (F.jitter) {
            jx = rng.next();
            jy = rng.next();
        }
        const V3 D = camera_ray(F, (float)px + jx, (float)py + jy);
        ctr.ray();
        HitG h = query<ST, DEEP>(S, F, V3{F.O[0], F.O[1], F.O[2]}, D, ctr, stk, mslot, tree);
        float4 L = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float beta[3] = {1.0f, 1.0f, 1.0f};
        if (!h.hit) {
            L = make_float4(1.0f * F.sky[0], 1.0f * F.sky[1], 1.0f * F.sky[2], 0.0f);
            if (smp == 0) {
                HitRecord rec;
                rec.inst = VT_EMPTY; rec.voxel = VT_EMPTY; rec.steps = 0; rec.face = 0xFF;
                rec.pal = 0; rec.mask = 0; rec.flags = 0; rec.vhi = 0;
                Rq.hits[pix] = rec;
            }
            if (Rq.gbuf && wf.gsamp) {
                wf.gsamp[2 * (size_t)w] = make_float4(F.sky[0], F.sky[1], F.sky[2], 0.0f);
                wf.gsamp[2 * (size_t)w + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        } else {
            ctr.hit();
            uint8_t pal = 0;
            SurfaceRec sf;
            sf.rgba = hit_rgba(S, h, &pal);
            if (smp == 0) {
                HitRecord rec;
                rec.inst = h.inst;
                const uint32_t fl = vt_pack_hit_voxel(rec, h.d.v[0], h.d.v[1], h.d.v[2], h.d.steps);
                rec.face = h.face;
                rec.pal = pal;
                rec.mask = (uint8_t)h.d.mask;
                rec.flags = (uint8_t)(1u | fl);
                Rq.hits[pix] = rec;
            }
            surface_of(S, h, sf);
            sf.path = w;
            if (Rq.gbuf && wf.gsamp) {
                const float ex = sf.X[0] - F.O[0], ey = sf.X[1] - F.O[1], ez = sf.X[2] - F.O[2];
                wf.gsamp[2 * (size_t)w] = make_float4(srgb_dec[sf.rgba & 255u], srgb_dec[(sf.rgba >> 8) & 255u],
                                                      srgb_dec[(sf.rgba >> 16) & 255u],
                                                      sqrtf((ex * ex + ey * ey) + ez * ez));
                wf.gsamp[2 * (size_t)w + 1] = make_float4(sf.n[0], sf.n[1], sf.n[2], 1.0f);
            }
            float4 contrib;
            bool no_sun;
            shade(F, srgb_dec, sf, 0u, max_b, beta, contrib, no_sun);
            uint32_t slot, ls;
            append_pair(slot_counters(wf.counters, 0), slot_counters(wf.counters, sun_slot(0)), !no_sun, wf.seg_cap, &slot,
                        &ls);
            put_rec(&wf.rec[0][slot], sf, contrib, w, rng.s);
            if (!no_sun) wf.sun[0][ls] = slot;
            // a hit path's state: sun terms (w: shadowed if facing away; its sun ray may set it);
            // the first bounce's throughput is the albedo and its RNG state is in the record
            L.w = __uint_as_float(1u);
            wf.Ls[w] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(no_sun ? 1u : 0u));
        }
        wf.L[w] = L;
    });
    if (ST) ctr.flush(stats);
}

// Lanes of a wave that are still walking when fewer than kPoolLow of its 64 lanes walk (and
// the launch has items left) pause; the others finish their rays and take new items.  (Swept
// over 1..56 on the 16 x 256^3 1080p frame: 8-12 is fastest; larger values refill more often
// and spend the time saved in the walk on more rounds of ray set-up and instance search.)
constexpr uint32_t kPoolLow = 8;

// The secondary rays of one depth in one launch.  Part p covers segment p of this depth's queue
// (count slots) and of its sun list (ns entries): items [0, nb) are the bounce rays of its slots
// 0..nb-1 (nb = count when paths continue, else 0), items [nb, nb + ns) the sun rays of the
// slots its sun list names (surfaces facing away from the sun have none, so no lane idles).  The two kinds
// write different per-path state (bounce rays: L, throughput, the next queue; sun
// rays: Ls and the shadow flag), so they run side by side and fill each other's tail.  in_idx / out_idx:
// counter slots of this depth's queue and the next one.  ONE: the batch holds one request (its
// scene and frame are then wave-uniform, read with scalar loads).
//
// Lanes are a pool refilled as their rays finish: each lane holds one item through rounds of
// (take items for the idle lanes; set up their rays; step every walking lane until fewer than
// kPoolLow walk; for the lanes whose walk ended: keep the hit, search the tree for the next
// instance, or finish the ray: sky or sun term, or shade and queue the hit).  A refill takes
// consecutive items of one part for its lanes with one atomic (item counters fetch + p *
// kPartStride).  The walk loop so keeps most lanes busy instead of running to the longest walk
// of a fixed 64-item tile (walk-loop lane use 0.18 -> 0.39 on the 1080p frame).
// Each ray's query: the origin instance from the surface point first, then (if that walk
// misses) the other instances the ray enters, in the tree's traversal order with the
// closest-hit pruning of query(); a sun ray stops at its first hit.
template <bool ST, bool DEEP, bool ONE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_secondary(
    const ReqGPU* __restrict__ reqs, const BatchGPU* __restrict__ B, const float* __restrict__ srgb_dec, WF wf,
    uint32_t q, uint32_t depth, uint32_t in_idx, uint32_t out_idx, uint32_t fetch_idx, uint32_t bounce,
    unsigned long long* stats) {
    __shared__ uint32_t smem[(kStackLds + 16) * kBlock];
    __shared__ BvhNode tree_lds[kLdsNodes];
    const Lds stk_lds{&smem[threadIdx.x]}, mslot{&smem[kStackLds * kBlock + threadIdx.x]};
    __shared__ uint32_t gone;
    if (threadIdx.x == 0) gone = 0u;
    const TreeLds tree = stage_tree(reqs, B, tree_lds);
    __syncthreads();
    Ctr<ST> ctr;
    uint32_t* const qin = slot_counters(wf.counters, in_idx);
    uint32_t* const qsun = slot_counters(wf.counters, sun_slot(in_idx));
    uint32_t* const fetch = slot_counters(wf.counters, fetch_idx);
    const uint32_t cap = wf.seg_cap;
    const uint32_t nq = 1u - q;
    const uint32_t nreq = B->nreq;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64u - lane));
    const uint32_t p0 = blockIdx.x % kParts;
    uint32_t kp = 0;   // parts this wave has found exhausted, in the order p0, p0 + 1, ... (uniform)

    // per-lane item and ray state
    bool busy = false, walking = false, is_bounce = false;
    uint32_t path = 0, origin = 0, ri = 0, rng = 0, inst = 0;
    uint8_t face0 = 0;
    float val[3] = {0.0f, 0.0f, 0.0f};   // bounce ray: throughput; sun ray: its light if clear
    float tn = 0.0f;
    V3 O{0.0f, 0.0f, 0.0f}, w{0.0f, 0.0f, 0.0f}, inv_d{0.0f, 0.0f, 0.0f}, mdir{0.0f, 0.0f, 0.0f};
    HitG best;
    best.hit = false;
    best.inst = VT_EMPTY;
    best.tn = INFINITY;
    TravStack<DEEP> stk;
    stk.lds = stk_lds;
    stk.sp = 0;
    BoxWalk W;
    W.steps = 0;

    for (;;) {
        // -- take items for the idle lanes (consecutive items of one part in lane order) --
        const uint64_t idle = __ballot(!busy);
        bool fresh = false;
        uint32_t item = 0, ipart = 0, inb_count = 0;
        if (idle != 0ull && kp < kParts) {
            const uint32_t need = (uint32_t)__popcll(idle);
            const uint32_t rank = (uint32_t)__popcll(idle & below);
            uint32_t placed = 0;
            while (placed < need && kp < kParts) {
                const uint32_t p = (p0 + kp) % kParts;
                const bool g = (__hip_atomic_load(&gone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & (1u << p)) != 0u;
                uint32_t take = 0, base = 0, nb_p = 0;
                if (!g) {
                    // (both counts loaded before either is used: one round trip)
                    const uint32_t cb = qin[p * kPartStride], cs = qsun[p * kPartStride];
                    nb_p = bounce ? min(cb, cap) : 0u;
                    const uint32_t count = nb_p + min(cs, cap);
                    if (lane == 0 && count > 0) base = atomicAdd(fetch + p * kPartStride, need - placed);
                    base = __shfl(base, 0, 64);
                    take = base < count ? min(count - base, need - placed) : 0u;
                }
                if (!busy && rank >= placed && rank < placed + take) {
                    fresh = true;
                    item = base + (rank - placed);
                    ipart = p;
                    inb_count = nb_p;
                }
                placed += take;
                if (placed < need) {   // part p has no items left
                    if (lane == 0 && !g) __hip_atomic_fetch_or(&gone, 1u << p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    ++kp;
                }
            }
        }
        // -- set up the new items' rays: the origin walk first --
        if (fresh) {
            const uint32_t nb = inb_count;   // bounce items of part ipart (read at the refill)
            is_bounce = item < nb;
            const uint32_t slot = is_bounce ? ipart * cap + item : wf.sun[q][ipart * cap + (item - nb)];
            const QueueRec sf = get_rec(&wf.rec[q][slot]);
            path = sf.path;
            origin = sf.inst;
            ri = ONE ? 0u : find_req(B->work_base, nreq, path);
            const ReqGPU& Rq = reqs[ri];
            const SceneGPU& S = Rq.S;
            const FrameGPU& F = Rq.F;
            const InstanceGPU* I = &S.inst[origin];
            const VolumeDesc V0 = inst_vol(S, I);   // (its loads issued before the sampling below)
            V3 X, n;
            surface_frame(&S.inst[sf.inst], sf.p, sf.face, X, n);
            Rng r{sf.rng};
            w = V3{F.sun[0], F.sun[1], F.sun[2]};
            if (is_bounce) {
                if (depth == 0) {
                    val[0] = 1.0f * srgb_dec[sf.rgba & 255u];
                    val[1] = 1.0f * srgb_dec[(sf.rgba >> 8) & 255u];
                    val[2] = 1.0f * srgb_dec[(sf.rgba >> 16) & 255u];
                } else {
                    const float4 b4 = wf.br[path];
                    val[0] = b4.x; val[1] = b4.y; val[2] = b4.z;
                }
                cosine_dir(r, n, w);
            } else {
                val[0] = sf.c[0]; val[1] = sf.c[1]; val[2] = sf.c[2];
            }
            rng = r.s;
            O = X;
            inv_d = V3{1.0f / w.x, 1.0f / w.y, 1.0f / w.z};
            ctr.ray();
            best.hit = false;
            best.inst = VT_EMPTY;
            best.tn = INFINITY;
            stk.sp = 0;
            inst = origin;
            tn = 0.0f;
            face0 = (uint8_t)sf.face;
            const int fa0 = (int)(sf.face >> 1), off0 = (sf.face & 1u) == 0u ? -1 : 1;
            const int v0[3] = {(int)sf.v[0] + (fa0 == 0 ? off0 : 0), (int)sf.v[1] + (fa0 == 1 ? off0 : 0),
                               (int)sf.v[2] + (fa0 == 2 ? off0 : 0)};
            float inv[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) inv[k] = gld(&I->inv[k]);
            mdir = xf34(inv, w, 0.0f);
            ctr.instance();
            box_walk_begin(V0, sf.p, v0, mdir, W);
            busy = true;
            walking = true;
        }
        if (__ballot(busy) == 0ull) break;   // no items left and none in flight
        // -- the walk loop: every walking lane steps until too few walk --
        for (;;) {
            if (walking) walking = box_walk_step<ST>(W, ctr, mslot);
            const uint32_t nw = (uint32_t)__popcll(__ballot(walking));
            if (nw == 0u || (nw < kPoolLow && kp < kParts)) break;
        }
        // -- lanes whose walk ended: keep its hit, then the next instance or the ray's end --
        bool fin = false;
        if (busy && !walking) {
            const SceneGPU& S = reqs[ri].S;
            ctr.steps(W.steps);
            bool done = false;
            if (W.hit) {
                HitG h;
                h.hit = true;
                h.inst = inst;
                h.tn = tn;
                h.d.v[0] = W.v[0]; h.d.v[1] = W.v[1]; h.d.v[2] = W.v[2];
                h.d.steps = W.steps;
                h.d.mask = W.steps > 0 ? W.mask : 0u;
                h.d.t_last = 0.0f;
                h.d.q[0] = W.q[0]; h.d.q[1] = W.q[1]; h.d.q[2] = W.q[2];
                h.d.entry = W.entry;
                h.d.li = W.li;
                h.d.hit = true;
                h.mpos[0] = 0.0f; h.mpos[1] = 0.0f; h.mpos[2] = 0.0f;
                h.mdir = mdir;
                if (W.steps > 0) {
                    const int a = lowest_axis(h.d.mask);
                    h.face = face_from(a, comp(mdir, a));
                } else {
                    h.face = face0;
                }
                if (inst == origin) {
                    best = h;
                    done = true;
                } else if (better(h, best)) {
                    best = h;
                    done = !is_bounce;   // a sun ray stops at any hit
                }
            }
            if (!done) {
                if (inst == origin && S.ninst > 0) stk.push(0u);   // the scene query follows the origin walk
                // The search alternates two loops so that lanes at inner nodes and lanes at
                // leaves do not pay for each other's work in one iteration: node tests until each
                // lane holds its next leaf, then the leaf's instance entry test for all of them.
                bool found = false;
                for (;;) {
                    uint32_t cand = VT_EMPTY;
                    while (stk.sp > 0) {
                        const uint32_t ni = stk.pop();
                        const BvhNode N = tree.on ? lds_node(tree, ni) : gld(S.nodes + ni);
                        ctr.node();
                        float a0, a1;
                        // (the tests combined without short-circuits: two branches per node)
                        const bool crossed = slab_node(O, w, inv_d, N, &a0, &a1);
                        const bool ok = crossed & !(a1 < 0.0f) & !(best.hit & (a0 > prune_bound(best.tn)));
                        const bool inner = N.right != VT_EMPTY;
                        if (ok & inner) {
                            stk.push(N.right);
                            stk.push(N.left);
                        }
                        if (ok & !inner & (N.left != origin)) {
                            cand = N.left;
                            break;
                        }
                    }
                    if (cand == VT_EMPTY) break;
                    const InstanceGPU* I = &S.inst[cand];
                    if (!gld(&I->valid)) continue;
                    const VolumeDesc Vc = inst_vol(S, I);   // (loads issued before the entry test)
                    float inv[12];
#pragma unroll
                    for (int k = 0; k < 12; ++k) inv[k] = gld(&I->inv[k]);
                    const V3 o = xf34(inv, O, 1.0f);
                    const V3 d = xf34(inv, w, 0.0f);
                    const SlabR sl = slab_unit(o, d);
                    if (!sl.ok || !(sl.tn > 0.0f)) continue;
                    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
                    const float size[3] = {gld(&I->size[0]), gld(&I->size[1]), gld(&I->size[2])};
                    float mpos[3];
                    int v0[3];
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        const float face = dd[a] > 0.0f ? -0.5f : 0.5f;
                        const float x = oo[a] + sl.tn * dd[a];
                        const float xc = x < -0.5f ? -0.5f : (x > 0.5f ? 0.5f : x);
                        const float mp = a == sl.axis ? face : xc;
                        mpos[a] = (mp + 0.5f) * size[a];
                        v0[a] = (int)floorf(gmin(mpos[a], size[a] - 1.0f));
                    }
                    mdir = d;
                    tn = sl.tn;
                    face0 = face_from(sl.axis, dd[sl.axis]);
                    inst = cand;
                    ctr.instance();
                    box_walk_begin(Vc, mpos, v0, mdir, W);
                    found = true;
                    break;
                }
                walking = found;
                done = !found;
            }
            fin = done;
        }
        // -- rays that ended: sky or sun term, or shade the hit and queue it for the next depth --
        if (fin) {
            const ReqGPU& Rq = reqs[ri];
            const SceneGPU& S = Rq.S;
            const FrameGPU& F = Rq.F;
            busy = false;
            if (best.hit) ctr.hit();
            if (!is_bounce) {
                if (!best.hit) {
                    float4 L = wf.Ls[path];
                    L.x = L.x + val[0];
                    L.y = L.y + val[1];
                    L.z = L.z + val[2];
                    wf.Ls[path] = L;
                } else if (depth == 0) {
                    float4 L = wf.Ls[path];
                    L.w = __uint_as_float(1u);
                    wf.Ls[path] = L;
                }
            } else if (!best.hit) {
                float4 L = wf.L[path];
                L.x = L.x + val[0] * F.sky[0];
                L.y = L.y + val[1] * F.sky[1];
                L.z = L.z + val[2] * F.sky[2];
                wf.L[path] = L;
            } else {
                SurfaceRec ns;
                uint8_t pal;
                ns.rgba = hit_rgba(S, best, &pal);
                surface_of(S, best, ns);
                float4 contrib;
                bool no_sun;
                const uint32_t max_b = F.mode == 2 ? F.max_bounces : 0u;
                shade(F, srgb_dec, ns, depth + 1u, max_b, val, contrib, no_sun);
                uint32_t ns_slot, ls;
                append_pair(slot_counters(wf.counters, out_idx), slot_counters(wf.counters, sun_slot(out_idx)), !no_sun,
                            cap, &ns_slot, &ls);
                put_rec(&wf.rec[nq][ns_slot], ns, contrib, path, rng);
                if (!no_sun) wf.sun[nq][ls] = ns_slot;
                if (depth + 1u < max_b) wf.br[path] = make_float4(val[0], val[1], val[2], 0.0f);
            }
        }
    }
    if (ST) ctr.flush(stats);
}

// Per pixel of every request: add this launch's samples to the running sum in sample order;
// the last chunk writes radiance = sum / spp and the shadow flag of sample 0.
__device__ void accumulate_pixel(const ReqGPU* __restrict__ reqs, const BatchGPU* __restrict__ B, WF wf, uint32_t t) {
    const uint32_t nreq = B->nreq;
    const uint32_t r = find_req(B->pix_base, nreq, t);
    const ReqGPU& Rq = reqs[r];
    const uint32_t W = Rq.F.width;
    const TileMap tm(Rq);
    // p: the pixel's index in the frame (the launch's pixels start at its first row of tiles)
    const uint32_t p = t - B->pix_base[r] + tm.t0 / tm.tiles_x * tm.th * W;
    const uint32_t px = p % W, py = p / W;
    const uint32_t wb = B->work_base[r];
    float acc[3] = {0.0f, 0.0f, 0.0f};
    if (Rq.s0 > 0) {
        const float4 a = Rq.acc[p];
        acc[0] = a.x; acc[1] = a.y; acc[2] = a.z;
    }
    for (uint32_t s = Rq.s0; s < Rq.s1; ++s) {
        const size_t k = wb + tm.index(px, py, s, Rq.s0);
        const float4 L = wf.L[k];
        // sun terms exist only for paths whose primary ray hit (L.w bits 1u)
        const float4 Ls = __float_as_uint(L.w) != 0u ? wf.Ls[k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        acc[0] += L.x + Ls.x;
        acc[1] += L.y + Ls.y;
        acc[2] += L.z + Ls.z;
    }
    const float fs = (float)(Rq.F.spp ? Rq.F.spp : 1u);
    if (Rq.final_chunk) Rq.rad[p] = make_float4(acc[0] / fs, acc[1] / fs, acc[2] / fs, 1.0f);
    else Rq.acc[p] = make_float4(acc[0], acc[1], acc[2], 0.0f);
    // G-buffer: mean over the samples, summed in sample order (a running sum across chunks)
    if (Rq.gbuf && wf.gsamp && (Rq.s0 == 0 || Rq.gacc) && (Rq.final_chunk || Rq.gacc)) {
        float g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (Rq.s0 > 0) {
            const float4 a = Rq.gacc[2 * p], b = Rq.gacc[2 * p + 1];
            g[0] = a.x; g[1] = a.y; g[2] = a.z; g[3] = a.w;
            g[4] = b.x; g[5] = b.y; g[6] = b.z; g[7] = b.w;
        }
        for (uint32_t s = Rq.s0; s < Rq.s1; ++s) {
            const size_t q = 2 * (size_t)(wb + tm.index(px, py, s, Rq.s0));
            const float4 a = wf.gsamp[q], b = wf.gsamp[q + 1];
            g[0] += a.x; g[1] += a.y; g[2] += a.z; g[3] += a.w;
            g[4] += b.x; g[5] += b.y; g[6] += b.z; g[7] += b.w;
        }
        if (Rq.final_chunk) {
            Rq.gbuf[2 * p] = make_float4(g[0] / fs, g[1] / fs, g[2] / fs, g[3] / fs);
            Rq.gbuf[2 * p + 1] = make_float4(g[4] / fs, g[5] / fs, g[6] / fs, g[7] / fs);
        } else {
            Rq.gacc[2 * p] = make_float4(g[0], g[1], g[2], g[3]);
            Rq.gacc[2 * p + 1] = make_float4(g[4], g[5], g[6], g[7]);
        }
    }
    if (Rq.s0 == 0 && (Rq.hits[p].flags & 1u)) {
        const size_t k0 = wb + tm.index(px, py, 0u, 0u);
        if (__float_as_uint(wf.L[k0].w) != 0u && __float_as_uint(wf.Ls[k0].w) != 0u) Rq.hits[p].flags |= 2u;
    }
}

__global__ void k_accumulate(const ReqGPU* __restrict__ reqs, const BatchGPU* __restrict__ B, WF wf) {
    const uint32_t nreq = B->nreq;
    const uint32_t total = B->pix_base[nreq];
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x)
        accumulate_pixel(reqs, B, wf, t);
}

__global__ void k_zero_counters(uint32_t* c) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kSlots * kParts) c[i * kPartStride] = 0u;
}

// Mode 0: one thread per pixel, samples looped in-thread (primary colour only).
template <bool ST, bool DEEP>
__global__ __launch_bounds__(kBlock) void k_render0(const FrameGPU F, const SceneGPU S, const float* __restrict__ srgb_dec,
                                                    HitRecord* __restrict__ hits, float4* __restrict__ rad,
                                                    uint32_t* fetch, uint32_t nwork, unsigned long long* stats) {
    __shared__ uint32_t smem[(kStackLds + 16) * kBlock];
    const Lds stk{&smem[threadIdx.x]}, mslot{&smem[kStackLds * kBlock + threadIdx.x]};
    __shared__ uint32_t gone;
    if (threadIdx.x == 0) gone = 0u;
    __syncthreads();
    Ctr<ST> ctr;
    const uint32_t spp = F.spp ? F.spp : 1u;
    for_each_item(fetch, &gone, [&](uint32_t p, uint32_t* first, uint32_t* count) { static_part(nwork, p, first, count); },
                  [&](uint32_t, uint32_t w) {
        uint32_t px, py, s0;
        if (!work_pixel(F, w, 1u, &px, &py, &s0)) return;
        const uint32_t pix = py * F.width + px;
        float acc[3] = {0.0f, 0.0f, 0.0f};
        float acca = 0.0f;
        HitRecord rec;
        rec.inst = VT_EMPTY; rec.voxel = VT_EMPTY; rec.steps = 0; rec.face = 0xFF;
        rec.pal = 0; rec.mask = 0; rec.flags = 0; rec.vhi = 0;
        for (uint32_t smp = 0; smp < spp; ++smp) {
            Rng rng{pixel_seed(F.seed, F.frame, pix, smp)};
            float jx = 0.5f, jy = 0.5f;
            if (F.jitter) {
                jx = rng.next();
                jy = rng.next();
            }
            const V3 D = camera_ray(F, (float)px + jx, (float)py + jy);
            ctr.ray();
            HitG h = query<ST, DEEP>(S, F, V3{F.O[0], F.O[1], F.O[2]}, D, ctr, stk, mslot, TreeLds{});
            if (h.hit) {
                ctr.hit();
                uint8_t pal = 0;
                const uint32_t rgba = hit_rgba(S, h, &pal);
                if (smp == 0) {
                    rec.inst = h.inst;
                    const uint32_t fl = vt_pack_hit_voxel(rec, h.d.v[0], h.d.v[1], h.d.v[2], h.d.steps);
                    rec.face = h.face;
                    rec.pal = pal;
                    rec.mask = (uint8_t)h.d.mask;
                    rec.flags = (uint8_t)(1u | fl);
                }
                const float a = (float)(rgba >> 24) / 255.0f;
                const float rgb[3] = {srgb_dec[rgba & 255u], srgb_dec[(rgba >> 8) & 255u], srgb_dec[(rgba >> 16) & 255u]};
                for (int k = 0; k < 3; ++k) acc[k] += rgb[k] * a + F.sky[k] * (1.0f - a);
                acca += a;
            } else {
                for (int k = 0; k < 3; ++k) acc[k] += F.sky[k];
                acca += 1.0f;
            }
        }
        hits[pix] = rec;
        const float fs = (float)spp;
        rad[pix] = make_float4(acc[0] / fs, acc[1] / fs, acc[2] / fs, acca / fs);
    });
    if (ST) ctr.flush(stats);
}

__global__ void k_encode_srgb8(const float4* rad, uint32_t* out, uint32_t n, const float* thr) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 c = rad[i];
    auto enc = [&](float v) -> uint32_t {
        uint32_t lo = 0, hi = 255;  // smallest k with v < thr[k], thr has 255 entries
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (v >= thr[mid]) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    uint32_t a = (uint32_t)fminf(fmaxf(c.w * 255.0f + 0.5f, 0.0f), 255.0f);
    // B8G8R8A8 byte order: b, g, r, a
    out[i] = enc(c.z) | (enc(c.y) << 8) | (enc(c.x) << 16) | (a << 24);
}

}  // namespace

static const float* g_srgb_dec = nullptr;
static const float* g_srgb_thr = nullptr;
static uint32_t g_persistent_blocks = 2048;

void vt_trace_set_tables(const float* dec_dev, const float* thr_dev) {
    g_srgb_dec = dec_dev;
    g_srgb_thr = thr_dev;
}

void vt_trace_set_persistent_blocks(uint32_t blocks) { g_persistent_blocks = blocks ? blocks : 2048; }

size_t vt_wave_bytes(uint32_t npaths) {
    const size_t qcap = (size_t)kParts * seg_cap(npaths);
    return (size_t)npaths * (16 + 16 + 16) + 2 * qcap * (sizeof(QueueRec) + 4) + (size_t)kCounters * 4 + 16 * 256 +
           sizeof(ReqGPU) * VT_MAX_BATCH + sizeof(BatchGPU);
}

int vt_wave_bind(WaveBuffers& wb, void* mem, uint32_t npaths) {
    char* p = static_cast<char*>(mem);
    auto take = [&](size_t bytes) {
        char* r = p;
        p += (bytes + 255) & ~(size_t)255;
        return r;
    };
    wb.npaths = npaths;
    wb.L = take((size_t)npaths * 16);
    wb.Ls = take((size_t)npaths * 16);
    wb.br = take((size_t)npaths * 16);
    const size_t qcap = (size_t)kParts * seg_cap(npaths);
    for (int q = 0; q < 2; ++q) {
        wb.rec[q] = take(qcap * sizeof(QueueRec));
        wb.sun[q] = take(qcap * 4);
    }
    wb.counters = reinterpret_cast<uint32_t*>(take((size_t)kCounters * 4));
    wb.tables = take(sizeof(ReqGPU) * VT_MAX_BATCH + sizeof(BatchGPU));
    return 0;
}

static WF to_wf(const WaveBuffers& wb) {
    WF w;
    w.L = reinterpret_cast<float4*>(wb.L);
    w.Ls = reinterpret_cast<float4*>(wb.Ls);
    w.br = reinterpret_cast<float4*>(wb.br);
    for (int q = 0; q < 2; ++q) {
        w.rec[q] = reinterpret_cast<QueueRec*>(wb.rec[q]);
        w.sun[q] = reinterpret_cast<uint32_t*>(wb.sun[q]);
    }
    w.counters = wb.counters;
    w.seg_cap = seg_cap(wb.npaths);
    w.gsamp = reinterpret_cast<float4*>(wb.gsamp);
    return w;
}

// Request tables: one device copy per wave-buffer block (stream order protects it), staged
// through a ring of pinned host buffers (an event per slot guards reuse).
struct TableRing {
    static constexpr int kSlots = 8;
    void* host[kSlots] = {};
    hipEvent_t ev[kSlots] = {};
    int next = 0;
    bool ready = false;
};
static TableRing g_ring;
static constexpr size_t kTableBytes = sizeof(ReqGPU) * VT_MAX_BATCH + sizeof(BatchGPU);

static int ring_init() {
    if (g_ring.ready) return 0;
    for (int i = 0; i < TableRing::kSlots; ++i) {
        if (hipHostMalloc(&g_ring.host[i], kTableBytes, hipHostMallocDefault) != hipSuccess) return -1;
        if (hipEventCreateWithFlags(&g_ring.ev[i], hipEventDisableTiming) != hipSuccess) return -1;
    }
    g_ring.ready = true;
    return 0;
}

void vt_trace_release() {
    if (!g_ring.ready) return;
    for (int i = 0; i < TableRing::kSlots; ++i) {
        hipEventSynchronize(g_ring.ev[i]);
        hipHostFree(g_ring.host[i]);
        hipEventDestroy(g_ring.ev[i]);
    }
    g_ring = TableRing{};
}

// Work and pixel prefix tables of a batch; returns false if it does not fit `npaths`.
bool vt_batch_layout(const ReqGPU* reqs, uint32_t n, uint32_t npaths, BatchGPU& B) {
    B.nreq = n;
    B.max_bounces = 0;
    uint64_t w = 0, px = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const ReqGPU& R = reqs[k];
        const uint32_t th = 64u / R.tile_w;
        uint32_t ty0, ty1, tiles_x;
        vt_req_tiles(R, &ty0, &ty1, &tiles_x);
        const uint64_t tiles = (uint64_t)tiles_x * (ty1 - ty0);
        const uint64_t r0 = (uint64_t)ty0 * th, r1 = (uint64_t)ty1 * th < R.F.height ? (uint64_t)ty1 * th : R.F.height;
        B.work_base[k] = (uint32_t)w;
        B.pix_base[k] = (uint32_t)px;
        w += tiles * 64u * (R.s1 - R.s0);
        px += (uint64_t)R.F.width * (r1 - r0);
        const uint32_t mb = R.F.mode == 2 ? R.F.max_bounces : 0u;
        if (mb > B.max_bounces) B.max_bounces = mb;
    }
    B.work_base[n] = (uint32_t)w;
    B.pix_base[n] = (uint32_t)px;
    return w <= npaths && px < 0xFFFFFFFFull;
}

template <bool ST, bool DEEP>
static int launch_batch(const BatchGPU& Bh, const WaveBuffers& wb, unsigned long long* stats, hipEvent_t* ev,
                        uint8_t* kinds, uint32_t* nstages, hipStream_t s, hipEvent_t acc_after = nullptr,
                        bool with_acc = true, uint32_t blocks = 0, bool allow_one = true) {
    const uint32_t nblk = blocks ? blocks : g_persistent_blocks;
    const dim3 grid(nblk), block(kBlock);
    const ReqGPU* reqs = reinterpret_cast<const ReqGPU*>(wb.tables);
    const BatchGPU* B = reinterpret_cast<const BatchGPU*>(static_cast<const char*>(wb.tables) + sizeof(ReqGPU) * VT_MAX_BATCH);
    WF wf = to_wf(wb);
    uint32_t stage = 0;
    auto begin = [&](uint8_t kind) {
        if (ev) hipEventRecord(ev[2 * stage], s);
        if (kinds) kinds[stage] = kind;
    };
    auto end = [&]() {
        if (ev) hipEventRecord(ev[2 * stage + 1], s);
        ++stage;
    };
    auto st_at = [&](uint32_t k) { return stats ? stats + (size_t)k * VT_STAT_COUNT : nullptr; };
    hipLaunchKernelGGL(k_zero_counters, dim3(kSlots * kParts / 256), dim3(256), 0, s, wb.counters);
    begin(VT_STAGE_PRIMARY);
    hipLaunchKernelGGL((k_primary<ST, DEEP>), grid, block, 0, s, reqs, B, g_srgb_dec, wf, st_at(stage));
    end();
    uint32_t fetch = 3, in_idx = 0;
    for (uint32_t d = 0; d <= Bh.max_bounces && fetch < kQueueCounters && kQueueCounters + d + 1 < kSlots; ++d) {
        const uint32_t out_idx = kQueueCounters + d + 1;
        begin(VT_STAGE_SECONDARY);
        if (allow_one && Bh.nreq == 1)
            hipLaunchKernelGGL((k_secondary<ST, DEEP, true>), grid, block, 0, s, reqs, B, g_srgb_dec, wf, d & 1u, d,
                               in_idx, out_idx, fetch, d < Bh.max_bounces ? 1u : 0u, st_at(stage));
        else
            hipLaunchKernelGGL((k_secondary<ST, DEEP, false>), grid, block, 0, s, reqs, B, g_srgb_dec, wf, d & 1u, d,
                               in_idx, out_idx, fetch, d < Bh.max_bounces ? 1u : 0u, st_at(stage));
        ++fetch;
        end();
        in_idx = out_idx;
    }
    if (with_acc) {
        if (acc_after) hipStreamWaitEvent(s, acc_after, 0);  // another chunk's running sum comes first
        begin(VT_STAGE_ACCUMULATE);
        hipLaunchKernelGGL(k_accumulate, dim3(nblk), dim3(256), 0, s, reqs, B, wf);
        end();
    }
    if (nstages) *nstages = stage;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Uploads the request table (stream-ordered) and launches the batched pipeline.  host_tables
// (optional, kTableBytes): stage the table there instead of the pinned ring, which then is
// never allocated; the caller keeps it unchanged until the stream has passed the launch.
int vt_trace_batch(const ReqGPU* reqs, uint32_t n, const WaveBuffers& wb, unsigned long long* stats, hipEvent_t* ev,
                   uint8_t* kinds, uint32_t* nstages, hipStream_t s, hipEvent_t acc_after, uint32_t blocks,
                   void* host_tables) {
    if (!g_srgb_dec || n == 0 || n > VT_MAX_BATCH || !wb.tables) return -1;
    int slot = -1;
    char* h = static_cast<char*>(host_tables);
    if (!h) {
        if (ring_init() != 0) return -1;
        slot = g_ring.next;
        g_ring.next = (g_ring.next + 1) % TableRing::kSlots;
        hipEventSynchronize(g_ring.ev[slot]);
        h = static_cast<char*>(g_ring.host[slot]);
    }
    ReqGPU* rq = reinterpret_cast<ReqGPU*>(h);
    BatchGPU* Bh = reinterpret_cast<BatchGPU*>(h + sizeof(ReqGPU) * VT_MAX_BATCH);
    std::memcpy(rq, reqs, sizeof(ReqGPU) * n);
    if (!vt_batch_layout(reqs, n, wb.npaths, *Bh)) return -1;
    if (Bh->max_bounces > 29) return -1;  // stages and fetch counters available per launch
    bool deep = false, st = false;
    for (uint32_t k = 0; k < n; ++k) {
        if (reqs[k].F.mode == 0 || reqs[k].tile_w == 0 || 64u % reqs[k].tile_w != 0 || reqs[k].s1 <= reqs[k].s0)
            return -1;
        deep = deep || reqs[k].S.bvh_depth + 1u > (uint32_t)kStackLds;
        st = st || reqs[k].F.stats != 0;
    }
    st = st && stats != nullptr;
    if (hipMemcpyAsync(wb.tables, h, kTableBytes, hipMemcpyHostToDevice, s) != hipSuccess) return -1;
    if (slot >= 0) hipEventRecord(g_ring.ev[slot], s);
    if (deep) return st ? launch_batch<true, true>(*Bh, wb, stats, ev, kinds, nstages, s, acc_after, true, blocks)
                        : launch_batch<false, true>(*Bh, wb, stats, ev, kinds, nstages, s, acc_after, true, blocks);
    return st ? launch_batch<true, false>(*Bh, wb, stats, ev, kinds, nstages, s, acc_after, true, blocks)
              : launch_batch<false, false>(*Bh, wb, stats, ev, kinds, nstages, s, acc_after, true, blocks);
}

// Captures the launch sequence of a batch (fixed grids, tables read from device memory)
// into a graph; the table upload stays outside the graph.
int vt_trace_capture(uint32_t max_bounces, bool deep, const WaveBuffers& wb, hipStream_t s, hipGraphExec_t* out,
                     bool with_acc, uint32_t blocks) {
    BatchGPU Bh{};
    Bh.max_bounces = max_bounces;
    hipGraph_t graph;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return -1;
    int rc = deep ? launch_batch<false, true>(Bh, wb, nullptr, nullptr, nullptr, nullptr, s, nullptr, with_acc, blocks,
                                              false)
                  : launch_batch<false, false>(Bh, wb, nullptr, nullptr, nullptr, nullptr, s, nullptr, with_acc, blocks,
                                               false);
    if (hipStreamEndCapture(s, &graph) != hipSuccess || rc != 0) return -1;
    rc = hipGraphInstantiate(out, graph, nullptr, nullptr, 0) == hipSuccess ? 0 : -1;
    hipGraphDestroy(graph);
    return rc;
}

// The accumulation launch of a batch whose tables are in wb (after a graph captured without it).
int vt_trace_accumulate(const WaveBuffers& wb, hipStream_t s, uint32_t blocks) {
    const ReqGPU* reqs = reinterpret_cast<const ReqGPU*>(wb.tables);
    const BatchGPU* B = reinterpret_cast<const BatchGPU*>(static_cast<const char*>(wb.tables) + sizeof(ReqGPU) * VT_MAX_BATCH);
    hipLaunchKernelGGL(k_accumulate, dim3(blocks ? blocks : g_persistent_blocks), dim3(256), 0, s, reqs, B, to_wf(wb));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Uploads a batch's request table only (for a captured graph launched afterwards).
int vt_trace_upload_tables(const ReqGPU* reqs, uint32_t n, const WaveBuffers& wb, hipStream_t s, BatchGPU* layout) {
    if (n == 0 || n > VT_MAX_BATCH || ring_init() != 0) return -1;
    const int slot = g_ring.next;
    g_ring.next = (g_ring.next + 1) % TableRing::kSlots;
    hipEventSynchronize(g_ring.ev[slot]);
    char* h = static_cast<char*>(g_ring.host[slot]);
    std::memcpy(h, reqs, sizeof(ReqGPU) * n);
    BatchGPU* Bh = reinterpret_cast<BatchGPU*>(h + sizeof(ReqGPU) * VT_MAX_BATCH);
    if (!vt_batch_layout(reqs, n, wb.npaths, *Bh)) return -1;
    if (layout) *layout = *Bh;
    if (hipMemcpyAsync(wb.tables, h, kTableBytes, hipMemcpyHostToDevice, s) != hipSuccess) return -1;
    hipEventRecord(g_ring.ev[slot], s);
    return 0;
}

template <bool ST, bool DEEP>
static int launch_render0(const FrameGPU& F, const SceneGPU& S, HitRecord* hits, float4* rad, const WaveBuffers& wb,
                          unsigned long long* stats, hipEvent_t* ev, uint8_t* kinds, uint32_t* nstages,
                          hipStream_t s) {
    const uint32_t tiles = ((F.width + 7u) / 8u) * ((F.height + 7u) / 8u);
    hipLaunchKernelGGL(k_zero_counters, dim3(kSlots * kParts / 256), dim3(256), 0, s, wb.counters);
    if (ev) hipEventRecord(ev[0], s);
    if (kinds) kinds[0] = VT_STAGE_RENDER0;
    hipLaunchKernelGGL((k_render0<ST, DEEP>), dim3(g_persistent_blocks), dim3(kBlock), 0, s, F, S, g_srgb_dec, hits,
                       rad, slot_counters(wb.counters, 2), tiles * 64u, stats);
    if (ev) hipEventRecord(ev[1], s);
    if (nstages) *nstages = 1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

size_t vt_trace_table_bytes() { return kTableBytes; }

int vt_trace_launch(const FrameGPU& F, const SceneGPU& S, HitRecord* hits, float4* rad, float4* gbuf,
                    const WaveBuffers& wb, unsigned long long* stats, hipEvent_t* ev, uint8_t* kinds, uint32_t* nstages,
                    hipStream_t s, void* host_tables) {
    if (!g_srgb_dec || F.mode > 2) return -1;
    if (F.mode == 0) {
        const bool st = F.stats != 0 && stats != nullptr;
        const bool deep = S.bvh_depth + 1u > (uint32_t)kStackLds;
        if (deep) return st ? launch_render0<true, true>(F, S, hits, rad, wb, stats, ev, kinds, nstages, s)
                            : launch_render0<false, true>(F, S, hits, rad, wb, stats, ev, kinds, nstages, s);
        return st ? launch_render0<true, false>(F, S, hits, rad, wb, stats, ev, kinds, nstages, s)
                  : launch_render0<false, false>(F, S, hits, rad, wb, stats, ev, kinds, nstages, s);
    }
    // modes 1 and 2: a batch of one request holding all samples of the frame
    ReqGPU R;
    std::memset(&R, 0, sizeof(R));
    R.F = F;
    R.S = S;
    R.hits = hits;
    R.acc = nullptr;
    R.rad = rad;
    R.gbuf = gbuf;
    R.s0 = 0;
    R.s1 = F.spp ? F.spp : 1u;
    R.final_chunk = 1;
    R.tile_w = 8;
    return vt_trace_batch(&R, 1, wb, stats, ev, kinds, nstages, s, nullptr, 0, host_tables);
}

int vt_encode_srgb8(const float4* rad, uint32_t* bgra, uint32_t n, hipStream_t s) {
    if (!g_srgb_thr) return -1;
    hipLaunchKernelGGL(k_encode_srgb8, dim3((n + 255) / 256), dim3(256), 0, s, rad, bgra, n, g_srgb_thr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
