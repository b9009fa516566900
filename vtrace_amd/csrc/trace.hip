// trace.hip — per-pixel ray generation, LBVH traversal, brick DDA and shading.
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
// All FP32 expressions are compiled with -ffp-contract=off and evaluated in the order
// written (correctly rounded division and square root), which keeps the integer outputs
// (instance, voxel, face, palette index, steps) bit-exact with the CPU oracle.
#include <hip/hip_runtime.h>

#include "vt_internal.h"

namespace {

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
    uint32_t entry;   // brick id of the hit voxel
    uint32_t li;      // voxel index inside the brick
    bool hit;
};

template <bool ST>
__device__ __forceinline__ void dda_run(const VolumeDesc& V, const float mpos[3], const int v0[3], V3 mdir, Dda& r,
                                        Ctr<ST>& ctr) {
    int vx = v0[0], vy = v0[1], vz = v0[2];
    const float len = sqrtf((mdir.x * mdir.x + mdir.y * mdir.y) + mdir.z * mdir.z);
    const float sgx = gsign(mdir.x), sgy = gsign(mdir.y), sgz = gsign(mdir.z);
    const int stx = (int)sgx, sty = (int)sgy, stz = (int)sgz;
    const float dx = fabsf(len / mdir.x), dy = fabsf(len / mdir.y), dz = fabsf(len / mdir.z);
    float sx = ((sgx * ((float)vx - mpos[0]) + sgx * 0.5f) + 0.5f) * dx;
    float sy = ((sgy * ((float)vy - mpos[1]) + sgy * 0.5f) + 0.5f) * dy;
    float sz = ((sgz * ((float)vz - mpos[2]) + sgz * 0.5f) + 0.5f) * dz;
    const int nx = (int)V.dim[0], ny = (int)V.dim[1], nz = (int)V.dim[2];
    const uint32_t max_steps = V.dim[0] + V.dim[1] + V.dim[2];
    const uint32_t bdx = V.bdim[0], bdy = V.bdim[1];
    uint32_t steps = 0, mask = 0;
    float t_last = 0.0f;
    uint32_t cur_b = VT_EMPTY - 1u, entry = VT_EMPTY;
    bool hit = false;
    uint32_t li = 0;
    while (steps < max_steps && (uint32_t)vx < (uint32_t)nx && (uint32_t)vy < (uint32_t)ny &&
           (uint32_t)vz < (uint32_t)nz) {
        const uint32_t bl = (uint32_t)(vx >> 3) + bdx * ((uint32_t)(vy >> 3) + bdy * (uint32_t)(vz >> 3));
        if (bl != cur_b) {
            cur_b = bl;
            entry = V.brick_map[bl];
            ctr.brick(entry != VT_EMPTY);
        }
        if (entry != VT_EMPTY) {
            li = (uint32_t)(vx & 7) + 8u * (uint32_t)(vy & 7) + 64u * (uint32_t)(vz & 7);
            const uint32_t w = V.masks[(size_t)entry * 16 + (li >> 5)];
            if ((w >> (li & 31u)) & 1u) {
                hit = true;
                break;
            }
        }
        const bool mx = sx <= gmin(sy, sz);
        const bool my = sy <= gmin(sz, sx);
        const bool mz = sz <= gmin(sx, sy);
        t_last = mx ? sx : (my ? sy : sz);
        sx += (mx ? 1.0f : 0.0f) * dx;
        sy += (my ? 1.0f : 0.0f) * dy;
        sz += (mz ? 1.0f : 0.0f) * dz;
        vx += (mx ? 1 : 0) * stx;
        vy += (my ? 1 : 0) * sty;
        vz += (mz ? 1 : 0) * stz;
        mask = (mx ? 1u : 0u) | (my ? 2u : 0u) | (mz ? 4u : 0u);
        ++steps;
    }
    ctr.steps(steps);
    r.v[0] = vx; r.v[1] = vy; r.v[2] = vz;
    r.steps = steps;
    r.mask = hit && steps > 0 ? mask : 0u;
    r.t_last = t_last;
    r.entry = entry;
    r.li = li;
    r.hit = hit;
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
__device__ __forceinline__ SlabR slab(V3 o, V3 d, const float lo[3], const float hi[3]) {
    SlabR s{-INFINITY, INFINITY, 0, true};
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float tn, tf;
        if (dd[a] != 0.0f) {
            float t1 = (lo[a] - oo[a]) / dd[a];
            float t2 = (hi[a] - oo[a]) / dd[a];
            tn = gmin(t1, t2);
            tf = gmax(t1, t2);
        } else {
            if (oo[a] < lo[a] || oo[a] > hi[a]) s.ok = false;
            tn = -INFINITY;
            tf = INFINITY;
        }
        if (tn > s.tn) { s.tn = tn; s.axis = a; }
        if (tf < s.tf) s.tf = tf;
    }
    s.ok = s.ok && s.tn <= s.tf;
    return s;
}

__device__ const float kLo[3] = {-0.5f, -0.5f, -0.5f};
__device__ const float kHi[3] = {0.5f, 0.5f, 0.5f};

// Enter instance i from outside along (O, D).  PRIMARY rebuilds the fragment inputs.
template <bool PRIMARY, bool ST>
__device__ bool enter_instance(const SceneGPU& S, const FrameGPU& F, uint32_t i, V3 O, V3 D, HitG& h, Ctr<ST>& ctr) {
    const InstanceGPU* I = &S.inst[i];
    if (!I->valid) return false;
    float inv[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) inv[k] = I->inv[k];
    const V3 o = xf34(inv, O, 1.0f);
    const V3 d = xf34(inv, D, 0.0f);
    const float lo[3] = {-0.5f, -0.5f, -0.5f}, hi[3] = {0.5f, 0.5f, 0.5f};
    SlabR s = slab(o, d, lo, hi);
    if (!s.ok || !(s.tn > 0.0f)) return false;
    float mp[3];
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a == s.axis) {
            mp[a] = dd[a] > 0.0f ? -0.5f : 0.5f;
        } else {
            float x = oo[a] + s.tn * dd[a];
            mp[a] = x < -0.5f ? -0.5f : (x > 0.5f ? 0.5f : x);
        }
    }
    V3 mdir;
    if (PRIMARY) {
        float M[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) M[k] = I->M[k];
        const V3 w = xf34(M, V3{mp[0], mp[1], mp[2]}, 1.0f);
        const float4 sp = mv4(F.PC, make_float4(w.x, w.y, w.z, 1.0f));
        if (!(sp.z >= 0.0f && sp.z <= sp.w)) return false;
        const float4 q = mv4(F.AB, sp);
        const V3 dir = normalize3(V3{q.x, q.y, q.z});
        mdir = xf34(inv, dir, 0.0f);
    } else {
        mdir = d;
    }
    ctr.instance();
    const VolumeDesc V = S.vols[I->tex];
    float mpos[3];
    int v0[3];
    const float size[3] = {I->size[0], I->size[1], I->size[2]};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        mpos[a] = (mp[a] + 0.5f) * size[a];
        v0[a] = (int)floorf(gmin(mpos[a], size[a] - 1.0f));
    }
    Dda r;
    dda_run<ST>(V, mpos, v0, mdir, r, ctr);
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

// Closest (or any) hit over all instances except `skip`, through the LBVH.
template <bool PRIMARY, bool ANY, bool ST>
__device__ HitG query(const SceneGPU& S, const FrameGPU& F, V3 O, V3 D, uint32_t skip, Ctr<ST>& ctr) {
    HitG best;
    best.hit = false;
    best.inst = VT_EMPTY;
    best.tn = INFINITY;
    if (S.ninst == 0) return best;
    uint32_t stack[64];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const uint32_t ni = stack[--sp];
        const BvhNode N = S.nodes[ni];
        ctr.node();
        SlabR s = slab(O, D, N.lo, N.hi);
        if (!s.ok || s.tf < 0.0f) continue;
        if (best.hit && s.tn > prune_bound(best.tn)) continue;
        if (N.right == VT_EMPTY) {
            const uint32_t i = N.left;
            if (i == skip) continue;
            HitG h;
            if (enter_instance<PRIMARY, ST>(S, F, i, O, D, h, ctr) && better(h, best)) {
                best = h;
                if (ANY) return best;
            }
        } else {
            stack[sp++] = N.right;
            stack[sp++] = N.left;
        }
    }
    return best;
}

struct Surface {
    float p[3];
    int vstart[3];
    V3 X;
    V3 n;
};

__device__ void surface_of(const SceneGPU& S, const HitG& h, Surface& sf) {
    const InstanceGPU* I = &S.inst[h.inst];
    const float len = sqrtf((h.mdir.x * h.mdir.x + h.mdir.y * h.mdir.y) + h.mdir.z * h.mdir.z);
    const int fa = h.face >> 1;
    const bool neg = (h.face & 1) == 0;
    const float md[3] = {h.mdir.x, h.mdir.y, h.mdir.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float u = md[a] / len;
        float x = h.mpos[a] + h.d.t_last * u;
        const float lo = (float)h.d.v[a], hi = (float)(h.d.v[a] + 1);
        if (a == fa) x = neg ? lo : hi;
        else x = x < lo ? lo : (x > hi ? hi : x);
        sf.p[a] = x;
        sf.vstart[a] = h.d.v[a] + (a == fa ? (neg ? -1 : 1) : 0);
    }
    float M[12], inv[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) { M[k] = I->M[k]; inv[k] = I->inv[k]; }
    const V3 mp{sf.p[0] / I->size[0] - 0.5f, sf.p[1] / I->size[1] - 0.5f, sf.p[2] / I->size[2] - 0.5f};
    sf.X = xf34(M, mp, 1.0f);
    const float nm = neg ? -1.0f : 1.0f;
    // row fa of inverse(M): inv[c*3 + fa]
    const V3 nw{inv[0 * 3 + fa] * nm, inv[1 * 3 + fa] * nm, inv[2 * 3 + fa] * nm};
    sf.n = normalize3(nw);
}

template <bool ANY, bool ST>
__device__ HitG secondary(const SceneGPU& S, const FrameGPU& F, uint32_t origin, const Surface& sf, uint8_t origin_face,
                          V3 w, Ctr<ST>& ctr) {
    ctr.ray();
    HitG h;
    h.hit = false;
    const InstanceGPU* I = &S.inst[origin];
    float inv[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) inv[k] = I->inv[k];
    const V3 mdir = xf34(inv, w, 0.0f);
    ctr.instance();
    const VolumeDesc V = S.vols[I->tex];
    Dda r;
    dda_run<ST>(V, sf.p, sf.vstart, mdir, r, ctr);
    if (r.hit) {
        h.hit = true;
        h.inst = origin;
        h.tn = 0.0f;
        h.d = r;
        h.mpos[0] = sf.p[0]; h.mpos[1] = sf.p[1]; h.mpos[2] = sf.p[2];
        h.mdir = mdir;
        if (r.steps > 0) {
            int a = lowest_axis(r.mask);
            h.face = face_from(a, comp(mdir, a));
        } else {
            h.face = origin_face;
        }
        ctr.hit();
        return h;
    }
    h = query<false, ANY, ST>(S, F, sf.X, w, origin, ctr);
    if (h.hit) ctr.hit();
    return h;
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
    const InstanceGPU* I = &S.inst[h.inst];
    const VolumeDesc V = S.vols[I->tex];
    if (V.fmt == 0) {
        const uint32_t idx = V.payload[(size_t)h.d.entry * 512 + h.d.li];
        *pal = (uint8_t)idx;
        return V.palette[idx];
    }
    *pal = 0;
    return reinterpret_cast<const uint32_t*>(V.payload)[(size_t)h.d.entry * 512 + h.d.li];
}

__device__ __forceinline__ uint32_t pixel_seed(uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t sample) {
    return pcg_hash(pixel ^ pcg_hash(sample ^ pcg_hash(frame ^ pcg_hash(seed))));
}

// XCD-aware tile order: consecutive workgroups land on different XCDs (round robin), so
// workgroup b is given tile (b % 8) * (tiles / 8) + b / 8 to keep each XCD's L2 on one
// contiguous band of the image.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t ntiles) {
    const uint32_t per = ntiles / 8u;
    if (per == 0 || b >= per * 8u) return b;
    return (b % 8u) * per + b / 8u;
}

constexpr int kTile = 16;  // 16 x 16 pixels per 256-thread workgroup (4 waves of 8 x 8)

template <int MODE, bool ST>
__global__ __launch_bounds__(256) void k_render(const FrameGPU F, const SceneGPU S, const float* __restrict__ srgb_dec,
                                                HitRecord* __restrict__ hits, float4* __restrict__ rad,
                                                unsigned long long* stats) {
    const uint32_t tiles_x = (F.width + kTile - 1) / kTile;
    const uint32_t tiles_y = (F.height + kTile - 1) / kTile;
    const uint32_t tile = xcd_tile(blockIdx.x, tiles_x * tiles_y);
    const uint32_t t = threadIdx.x, wv = t >> 6, ln = t & 63;
    const uint32_t px = (tile % tiles_x) * kTile + (wv & 1) * 8 + (ln & 7);
    const uint32_t py = (tile / tiles_x) * kTile + (wv >> 1) * 8 + (ln >> 3);
    Ctr<ST> ctr;
    if (px < F.width && py < F.height) {
        const uint32_t pix = py * F.width + px;
        const uint32_t spp = F.spp ? F.spp : 1u;
        float acc[3] = {0.0f, 0.0f, 0.0f};
        float acca = 0.0f;
        HitRecord rec;
        rec.inst = VT_EMPTY;
        rec.voxel = VT_EMPTY;
        rec.steps = 0;
        rec.face = 0xFF;
        rec.pal = 0;
        rec.mask = 0;
        rec.flags = 0;
        rec.pad = 0;
        const V3 O{F.O[0], F.O[1], F.O[2]};
        const V3 sun{F.sun[0], F.sun[1], F.sun[2]};
        for (uint32_t smp = 0; smp < spp; ++smp) {
            Rng rng{pixel_seed(F.seed, F.frame, pix, smp)};
            float jx = 0.5f, jy = 0.5f;
            if (F.jitter) {
                jx = rng.next();
                jy = rng.next();
            }
            const float sx = (float)px + jx, sy = (float)py + jy;
            const float nxd = (2.0f * sx) / F.W - 1.0f;
            const float nyd = (2.0f * sy) / F.Hv - 1.0f;
            const V3 vv{nxd * F.ip00, nyd * F.ip11, -1.0f};
            const V3 D{(F.R[0] * vv.x + F.R[3] * vv.y) + F.R[6] * vv.z, (F.R[1] * vv.x + F.R[4] * vv.y) + F.R[7] * vv.z,
                       (F.R[2] * vv.x + F.R[5] * vv.y) + F.R[8] * vv.z};
            ctr.ray();
            HitG h = query<true, false, ST>(S, F, O, D, VT_EMPTY, ctr);
            if (h.hit) ctr.hit();
            uint8_t pal = 0;
            uint32_t rgba = h.hit ? hit_rgba(S, h, &pal) : 0u;
            if (smp == 0 && h.hit) {
                rec.inst = h.inst;
                rec.voxel = (uint32_t)h.d.v[0] | ((uint32_t)h.d.v[1] << 10) | ((uint32_t)h.d.v[2] << 20);
                rec.steps = (uint16_t)h.d.steps;
                rec.face = h.face;
                rec.pal = pal;
                rec.mask = (uint8_t)h.d.mask;
                rec.flags = 1;
            }
            if (MODE == 0) {
                if (h.hit) {
                    const float a = (float)(rgba >> 24) / 255.0f;
                    const float rgb[3] = {srgb_dec[rgba & 255u], srgb_dec[(rgba >> 8) & 255u], srgb_dec[(rgba >> 16) & 255u]};
                    for (int k = 0; k < 3; ++k) acc[k] += rgb[k] * a + F.sky[k] * (1.0f - a);
                    acca += a;
                } else {
                    for (int k = 0; k < 3; ++k) acc[k] += F.sky[k];
                    acca += 1.0f;
                }
                continue;
            }
            float beta[3] = {1.0f, 1.0f, 1.0f};
            float L[3] = {0.0f, 0.0f, 0.0f};
            const uint32_t max_b = MODE == 2 ? F.max_bounces : 0u;
            HitG cur = h;
            uint32_t cur_rgba = rgba;
            for (uint32_t depth = 0;; ++depth) {
                if (!cur.hit) {
                    for (int k = 0; k < 3; ++k) L[k] += beta[k] * F.sky[k];
                    break;
                }
                Surface sf;
                surface_of(S, cur, sf);
                const float rho[3] = {srgb_dec[cur_rgba & 255u], srgb_dec[(cur_rgba >> 8) & 255u],
                                      srgb_dec[(cur_rgba >> 16) & 255u]};
                const float cosl = (sf.n.x * sun.x + sf.n.y * sun.y) + sf.n.z * sun.z;
                if (cosl > 0.0f) {
                    HitG sh = secondary<true, ST>(S, F, cur.inst, sf, cur.face, sun, ctr);
                    if (!sh.hit) {
                        const float c = cosl * 0.3183098861837907f;
                        for (int k = 0; k < 3; ++k) L[k] += beta[k] * (rho[k] * (F.sun_rad[k] * c));
                    } else if (smp == 0 && depth == 0) {
                        rec.flags |= 2;
                    }
                } else if (smp == 0 && depth == 0) {
                    rec.flags |= 2;
                }
                if (depth >= max_b) break;
                V3 w;
                cosine_dir(rng, sf.n, w);
                for (int k = 0; k < 3; ++k) beta[k] *= rho[k];
                cur = secondary<false, ST>(S, F, cur.inst, sf, cur.face, w, ctr);
                if (cur.hit) {
                    uint8_t p2;
                    cur_rgba = hit_rgba(S, cur, &p2);
                }
            }
            for (int k = 0; k < 3; ++k) acc[k] += L[k];
        }
        hits[pix] = rec;
        const float fs = (float)spp;
        if (MODE == 0) rad[pix] = make_float4(acc[0] / fs, acc[1] / fs, acc[2] / fs, acca / fs);
        else rad[pix] = make_float4(acc[0] / fs, acc[1] / fs, acc[2] / fs, 1.0f);
    }
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

void vt_trace_set_tables(const float* dec_dev, const float* thr_dev) {
    g_srgb_dec = dec_dev;
    g_srgb_thr = thr_dev;
}

template <int MODE, bool ST>
static void launch(const FrameGPU& F, const SceneGPU& S, HitRecord* hits, float4* rad, unsigned long long* stats,
                   hipStream_t s) {
    const uint32_t tiles = ((F.width + kTile - 1) / kTile) * ((F.height + kTile - 1) / kTile);
    hipLaunchKernelGGL((k_render<MODE, ST>), dim3(tiles), dim3(256), 0, s, F, S, g_srgb_dec, hits, rad, stats);
}

int vt_trace_launch(const FrameGPU& F, const SceneGPU& S, HitRecord* hits, float4* rad, unsigned long long* stats,
                    hipStream_t s) {
    if (!g_srgb_dec) return -1;
    const bool st = F.stats != 0 && stats != nullptr;
    switch (F.mode) {
        case 0: st ? launch<0, true>(F, S, hits, rad, stats, s) : launch<0, false>(F, S, hits, rad, stats, s); break;
        case 1: st ? launch<1, true>(F, S, hits, rad, stats, s) : launch<1, false>(F, S, hits, rad, stats, s); break;
        case 2: st ? launch<2, true>(F, S, hits, rad, stats, s) : launch<2, false>(F, S, hits, rad, stats, s); break;
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int vt_encode_srgb8(const float4* rad, uint32_t* bgra, uint32_t n, hipStream_t s) {
    if (!g_srgb_thr) return -1;
    hipLaunchKernelGGL(k_encode_srgb8, dim3((n + 255) / 256), dim3(256), 0, s, rad, bgra, n, g_srgb_thr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
