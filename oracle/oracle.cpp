// oracle.cpp — CPU restatement of vtrace's voxel ray-traversal path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (vtrace_amd/) links, loads or
// calls this file.  It is used by tests/ as the checker, by __graft_entry__.smoke()
// as the checker and by bench.py's cpu_baseline leg as the timed CPU baseline.
//
// What it restates (reference files as mounted, file:line):
//   * per-fragment ray setup + DDA:   shaders/trace.frag:41-90
//   * proxy transform / texture id:   shaders/trace.vert:32-47
//   * visibility of proxies:          lib/pipeline.c:120-121,146-152 (cull back, CW front,
//                                     depth LESS), lib/command.c:77-102 (W x W viewport,
//                                     draw order = instance order), lib/memory.c:22-40 (cube)
//   * sampler:                        lib/descriptor.c:97-125 (NEAREST, CLAMP_TO_EDGE)
//   * texel order / formats:          lib/memory.c:286-385 (x-fastest RGBA8, _SRGB),
//                                     lib/command.c:57-60 (sky clear colour),
//                                     lib/pipeline.c:129-137 (src-alpha blend)
//   * instance record:                src/render.rs:31-35,74-78 (texture id in m[3][3])
// The reference has no LBVH, bricks, RNG, shading or path tracing; those parts follow
// this project's own specification (stated at each function below) and are pinned in
// tests by brute force, closed forms and invariants rather than by the reference.
//
// Numerics: compiled with -ffp-contract=off so every FP32 expression below is evaluated
// as written, one IEEE-rounded operation at a time (GLSL evaluation order, left to right).
//
// Parity status per function (see the tests named):
//   orc_pcg_hash            pinned: bijectivity + committed vectors (tests/test_oracle.py)
//   orc_morton30            pinned: closed-form known answers (tests/test_oracle.py)
//   orc_inverse4            pinned: M * inv(M) == I to 1e-6 on random affine matrices
//   orc_instance_prep       pinned: AABB contains all transformed corners; texture-id bitcast
//   orc_lbvh_build          pinned: sort == std::stable_sort, topology invariants, AABB
//                           containment, traversal result == brute force
//   orc_brick               pinned: dense <-> brick round trip, popcount == dense count,
//                           empty radius == brute-force L-inf distance, octant cube sides ==
//                           brute-force largest anchored empty cube (tests/test_oracle.py)
//   orc_brick_blocks        pinned: 2^3-block bits == brute-force block occupancy
//   orc_render (primary)    pinned: brute-force nearest-voxel test on tiny volumes,
//                           axis-aligned closed forms, mirror symmetry, max_steps cap
//   orc_render (path)       pinned: white-furnace / single-plane closed forms, FP32-vs-FP64
//                           geometry replay
//   cosine_dir              pinned (orc_cosine_dir_samples): n.w follows P(cos <= c) = c^2
//                           (KS test), mean cos = 2/3 within its CI, uniform azimuth, unit
//                           length; the same checks reject uniform-hemisphere sampling
//   orc_srgb_encode         pinned: encode(decode(k)) == k for all 256 codes, monotone,
//                           equals round(255 * published sRGB OETF) away from code boundaries

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

extern "C" {

// ----------------------------------------------------------------------------------
// Public structures (mirrored by ctypes in tests/oracle_ffi.py)
// ----------------------------------------------------------------------------------
typedef struct orc_volume {
    uint32_t dim[3];
    uint32_t fmt;             // 0 = u8 palette index + 256-entry RGBA8 palette, 1 = RGBA8 texels
    const uint8_t* data;      // dense texels, x fastest: i + w*(j + h*k)
    const uint32_t* palette;  // fmt 0 only: 256 x RGBA8 (r in the low byte)
} orc_volume;

typedef struct orc_params {
    uint32_t width, height;
    uint32_t viewport_square;   // 1: the reference's W x W viewport (command.c:77-84)
    uint32_t mode;              // 0 primary colour, 1 primary + sun shadow ray, 2 path tracing
    uint32_t spp;
    uint32_t max_bounces;
    uint32_t jitter;            // 0: pixel centres, 1: RNG jitter inside the pixel
    uint32_t seed;
    uint32_t frame;
    float sun_dir[3];           // unit vector towards the sun (world space)
    float sun_radiance[3];
    float sky[3];
} orc_params;

typedef struct orc_hit {
    uint32_t inst;    // instance index (gl_InstanceIndex), 0xFFFFFFFF for sky
    uint32_t voxel;   // bits 0-9 of x | y << 10 | z << 20
    uint16_t steps;   // DDA steps taken in the winning instance, saturated at 65535 (flags bit 3)
    uint8_t face;     // 0..5 = -X,+X,-Y,+Y,-Z,+Z outward normal of the hit face; 0xFF miss
    uint8_t pal;      // palette index of the hit texel (fmt 0), 0 for RGBA8 volumes
    uint8_t mask;     // last DDA step mask (bit0 x, bit1 y, bit2 z), 0 if steps == 0
    uint8_t flags;    // bit0 hit, bit1 primary vertex shadowed (mode >= 1), bit2 a voxel
                      // coordinate >= 32768 (low 15 bits held), bit3 steps saturated
    uint16_t vhi;     // bits 10-14 of x | y << 5 | z << 10
} orc_hit;

typedef struct orc_stats {
    uint64_t rays;             // traversal queries (primary + shadow + bounce)
    uint64_t nodes;            // LBVH nodes visited
    uint64_t instances;        // instance boxes entered (DDA started)
    uint64_t bricks;           // 8^3 bricks entered
    uint64_t bricks_nonempty;  // of which non-empty
    uint64_t hits;             // queries that ended on a voxel
    uint64_t steps;            // DDA steps (all queries)
} orc_stats;

}  // extern "C"

namespace {

constexpr uint32_t kMiss = 0xFFFFFFFFu;
// brick map: non-empty bricks hold their id; empty ones kEmptyBrick | radius << 24 |
// (cube side - 1) << 3o for the 8 octants o
constexpr uint32_t kEmptyBrick = 0xE0000000u;
constexpr uint32_t kCubeCap = 8;
constexpr float kInf = std::numeric_limits<float>::infinity();

// ---- GLSL built-ins with the semantics this restatement fixes ---------------------
// min(x, y): GLSL.std.450 leaves NaN operands undefined; the GLSL spec text defines it
// as "y if y < x, otherwise x", which is what both this oracle and the GPU use.
inline float gmin(float x, float y) { return (y < x) ? y : x; }
inline float gmax(float x, float y) { return (x < y) ? y : x; }
inline float gsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// column-major 4x4, m[c*4 + r] (GLSL m[c][r])
inline void mv4(const float* m, const float v[4], float o[4]) {
    for (int r = 0; r < 4; ++r)
        o[r] = ((m[0 * 4 + r] * v[0] + m[1 * 4 + r] * v[1]) + m[2 * 4 + r] * v[2]) + m[3 * 4 + r] * v[3];
}
inline void mm4(const float* a, const float* b, float* o) {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            o[c * 4 + r] = ((a[0 * 4 + r] * b[c * 4 + 0] + a[1 * 4 + r] * b[c * 4 + 1]) +
                            a[2 * 4 + r] * b[c * 4 + 2]) + a[3 * 4 + r] * b[c * 4 + 3];
}

// 4x4 inverse by Laplace expansion over 2x2 minors (Eberly, "The Laplace Expansion
// Theorem: Computing the Determinants and Inverses of Matrices", Geometric Tools 2008),
// evaluated in FP64 and rounded to FP32.  Returns false if singular / non-finite.
bool inverse4(const float* mf, float* out) {
    auto A = [&](int r, int c) -> double { return (double)mf[c * 4 + r]; };
    double s0 = A(0, 0) * A(1, 1) - A(1, 0) * A(0, 1);
    double s1 = A(0, 0) * A(1, 2) - A(1, 0) * A(0, 2);
    double s2 = A(0, 0) * A(1, 3) - A(1, 0) * A(0, 3);
    double s3 = A(0, 1) * A(1, 2) - A(1, 1) * A(0, 2);
    double s4 = A(0, 1) * A(1, 3) - A(1, 1) * A(0, 3);
    double s5 = A(0, 2) * A(1, 3) - A(1, 2) * A(0, 3);
    double c5 = A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3);
    double c4 = A(2, 1) * A(3, 3) - A(3, 1) * A(2, 3);
    double c3 = A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2);
    double c2 = A(2, 0) * A(3, 3) - A(3, 0) * A(2, 3);
    double c1 = A(2, 0) * A(3, 2) - A(3, 0) * A(2, 2);
    double c0 = A(2, 0) * A(3, 1) - A(3, 0) * A(2, 1);
    double det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    if (!(det != 0.0) || !std::isfinite(det)) return false;
    double b[4][4];  // b[r][c]
    b[0][0] = (A(1, 1) * c5 - A(1, 2) * c4 + A(1, 3) * c3) / det;
    b[0][1] = (-A(0, 1) * c5 + A(0, 2) * c4 - A(0, 3) * c3) / det;
    b[0][2] = (A(3, 1) * s5 - A(3, 2) * s4 + A(3, 3) * s3) / det;
    b[0][3] = (-A(2, 1) * s5 + A(2, 2) * s4 - A(2, 3) * s3) / det;
    b[1][0] = (-A(1, 0) * c5 + A(1, 2) * c2 - A(1, 3) * c1) / det;
    b[1][1] = (A(0, 0) * c5 - A(0, 2) * c2 + A(0, 3) * c1) / det;
    b[1][2] = (-A(3, 0) * s5 + A(3, 2) * s2 - A(3, 3) * s1) / det;
    b[1][3] = (A(2, 0) * s5 - A(2, 2) * s2 + A(2, 3) * s1) / det;
    b[2][0] = (A(1, 0) * c4 - A(1, 1) * c2 + A(1, 3) * c0) / det;
    b[2][1] = (-A(0, 0) * c4 + A(0, 1) * c2 - A(0, 3) * c0) / det;
    b[2][2] = (A(3, 0) * s4 - A(3, 1) * s2 + A(3, 3) * s0) / det;
    b[2][3] = (-A(2, 0) * s4 + A(2, 1) * s2 - A(2, 3) * s0) / det;
    b[3][0] = (-A(1, 0) * c3 + A(1, 1) * c1 - A(1, 2) * c0) / det;
    b[3][1] = (A(0, 0) * c3 - A(0, 1) * c1 + A(0, 2) * c0) / det;
    b[3][2] = (-A(3, 0) * s3 + A(3, 1) * s1 - A(3, 2) * s0) / det;
    b[3][3] = (A(2, 0) * s3 - A(2, 1) * s1 + A(2, 2) * s0) / det;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            float v = (float)b[r][c];
            if (!std::isfinite(v)) return false;
            out[c * 4 + r] = v;
        }
    return true;
}

// ---- PCG hash (Jarzynski & Olano, "Hash Functions for GPU Rendering", JCGT 9(3), 2020,
// the pcg variant: one PCG-RXS-M-XS 32-bit round).
inline uint32_t pcg_hash(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
inline float u01(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

struct Rng {
    uint32_t s;
    float next() { s = pcg_hash(s); return u01(s); }
};
// Per (seed, frame, pixel, sample) stream start.
inline uint32_t rng_seed(uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t sample) {
    return pcg_hash(pixel ^ pcg_hash(sample ^ pcg_hash(frame ^ pcg_hash(seed))));
}

// ---- sRGB (R8G8B8A8_SRGB texel decode, memory.c:317; B8G8R8A8_SRGB store, swapchain.c:88)
double srgb_to_linear_d(uint32_t c8) {
    double c = (double)c8 / 255.0;
    return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4);
}
struct SrgbTables {
    float dec[256];
    float enc_thr[255];  // linear threshold between code k and k+1
    SrgbTables() {
        for (int i = 0; i < 256; ++i) dec[i] = (float)srgb_to_linear_d(i);
        for (int k = 0; k < 255; ++k) {
            double c = (k + 0.5) / 255.0;
            double lin = c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4);
            enc_thr[k] = (float)lin;
        }
    }
};
const SrgbTables& srgb() { static SrgbTables t; return t; }

// ---- Morton 30 (Karras 2012 convention: m = 4*expand(x) + 2*expand(y) + expand(z))
inline uint32_t expand10(uint32_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}
inline uint32_t morton30(uint32_t x, uint32_t y, uint32_t z) {
    return (expand10(x) << 2) | (expand10(y) << 1) | expand10(z);
}

// ---- scene ------------------------------------------------------------------------
struct Inst {
    float M[16];     // recovered model matrix ([3][3] = 1)
    float inv[16];   // inverse(M) rounded to FP32
    float lo[3], hi[3];
    uint32_t tex;
    bool valid;
};

struct Vol {
    const orc_volume* v;
    uint32_t dim[3];
    float size[3];
    std::vector<uint16_t> remap[3];  // texel index for DDA voxel v per axis
    std::vector<uint8_t> brick_nonempty;
    std::vector<uint8_t> brick_radius;  // empty bricks: radius of the cube of empty bricks around them
    std::vector<uint8_t> brick_cube[8]; // empty bricks: side of the empty cube anchored there, per octant
    std::vector<uint64_t> brick_blocks; // non-empty bricks: bit bx + 4by + 16bz per occupied 2^3 block
    uint32_t bdim[3];
    const SrgbTables* s;

    uint32_t texel_index(int x, int y, int z) const {
        return (uint32_t)remap[0][x] + dim[0] * ((uint32_t)remap[1][y] + dim[1] * (uint32_t)remap[2][z]);
    }
    // RGBA8 of the texel the reference samples for voxel (x,y,z)
    uint32_t rgba(uint32_t ti, uint8_t* pal_idx) const {
        if (v->fmt == 0) {
            uint8_t i = v->data[ti];
            *pal_idx = i;
            return v->palette[i];
        }
        *pal_idx = 0;
        uint32_t c;
        std::memcpy(&c, v->data + 4ull * ti, 4);
        return c;
    }
    bool occupied_texel(uint32_t ti) const {
        uint8_t p;
        return (rgba(ti, &p) >> 24) != 0u;
    }
};

// texel index selected by NEAREST + CLAMP_TO_EDGE at normalized coordinate voxel/size
// (trace.frag:77: texture(tex, model_ray_voxel / model_size)).
uint16_t nearest_texel(int v, float size, uint32_t dim) {
    float u = (float)v / size;
    float x = u * size;
    float f = std::floor(x);
    int t = (int)f;
    if (t < 0) t = 0;
    if (t > (int)dim - 1) t = (int)dim - 1;
    return (uint16_t)t;
}

// Empty-space radius of every brick (this project's rule, no reference counterpart): the
// L-infinity distance d, in bricks, from an empty brick to the nearest non-empty brick of the
// volume, capped at kRadiusCap + 1; radius = d - 1, so the cube of bricks within `radius` of it
// is empty (bricks outside the volume count as empty).  Computed as breadth-first layers of
// 26-neighbour dilation.  Non-empty bricks get 0.
constexpr uint32_t kRadiusCap = 31;
void brick_radius(Vol& V) {
    const int bx = (int)V.bdim[0], by = (int)V.bdim[1], bz = (int)V.bdim[2];
    std::vector<uint8_t> d(V.brick_nonempty.size());
    for (size_t b = 0; b < d.size(); ++b) d[b] = V.brick_nonempty[b] ? 0 : 255;
    for (uint32_t k = 1; k <= kRadiusCap + 1; ++k)
        for (int z = 0; z < bz; ++z)
            for (int y = 0; y < by; ++y)
                for (int x = 0; x < bx; ++x) {
                    uint8_t& c = d[x + bx * (y + by * z)];
                    if (c != 255) continue;
                    bool near = false;
                    for (int dz = -1; dz <= 1 && !near; ++dz)
                        for (int dy = -1; dy <= 1 && !near; ++dy)
                            for (int dx = -1; dx <= 1 && !near; ++dx) {
                                const int X = x + dx, Y = y + dy, Z = z + dz;
                                if (X < 0 || Y < 0 || Z < 0 || X >= bx || Y >= by || Z >= bz) continue;
                                near = d[X + bx * (Y + by * Z)] == k - 1;
                            }
                    if (near) c = (uint8_t)k;
                }
    V.brick_radius.assign(d.size(), 0);
    for (size_t b = 0; b < d.size(); ++b)
        V.brick_radius[b] = V.brick_nonempty[b] ? 0 : (uint8_t)(std::min<uint32_t>(d[b], kRadiusCap + 1) - 1);
}

// Per empty brick and octant o (bit a of o set: towards -axis a), the side K_o, in bricks, of
// the largest cube of empty bricks that has the brick as its corner and extends towards o,
// capped at kCubeCap (bricks outside the volume count as empty): K = 1 + the smallest side
// of the 7 neighbours towards o (outside: kCubeCap), a sweep from the far corner of the octant.
void brick_cubes(Vol& V) {
    const int bx = (int)V.bdim[0], by = (int)V.bdim[1], bz = (int)V.bdim[2];
    for (int o = 0; o < 8; ++o) {
        const int sx = (o & 1) ? -1 : 1, sy = (o & 2) ? -1 : 1, sz = (o & 4) ? -1 : 1;
        std::vector<uint8_t>& K = V.brick_cube[o];
        K.assign((size_t)bx * by * bz, 0);
        auto at = [&](int x, int y, int z) -> int {
            if (x < 0 || y < 0 || z < 0 || x >= bx || y >= by || z >= bz) return (int)kCubeCap;
            return K[x + (size_t)bx * (y + (size_t)by * z)];
        };
        for (int zi = 0; zi < bz; ++zi)
            for (int yi = 0; yi < by; ++yi)
                for (int xi = 0; xi < bx; ++xi) {
                    const int x = sx > 0 ? bx - 1 - xi : xi, y = sy > 0 ? by - 1 - yi : yi, z = sz > 0 ? bz - 1 - zi : zi;
                    const size_t b = x + (size_t)bx * (y + (size_t)by * z);
                    if (V.brick_nonempty[b]) continue;
                    int m = (int)kCubeCap;
                    for (int nb = 1; nb < 8; ++nb)
                        m = std::min(m, at(x + ((nb & 1) ? sx : 0), y + ((nb & 2) ? sy : 0), z + ((nb & 4) ? sz : 0)));
                    K[b] = (uint8_t)std::min<int>((int)kCubeCap, 1 + m);
                }
    }
}

void init_vol(Vol& V, const orc_volume* v) {
    V.v = v;
    V.s = &srgb();
    for (int a = 0; a < 3; ++a) {
        V.dim[a] = v->dim[a];
        V.size[a] = (float)v->dim[a];
        V.remap[a].resize(v->dim[a]);
        for (uint32_t i = 0; i < v->dim[a]; ++i) V.remap[a][i] = nearest_texel((int)i, V.size[a], v->dim[a]);
        V.bdim[a] = (v->dim[a] + 7) / 8;
    }
    V.brick_nonempty.assign((size_t)V.bdim[0] * V.bdim[1] * V.bdim[2], 0);
    V.brick_blocks.assign(V.brick_nonempty.size(), 0);
    for (uint32_t z = 0; z < V.dim[2]; ++z)
        for (uint32_t y = 0; y < V.dim[1]; ++y)
            for (uint32_t x = 0; x < V.dim[0]; ++x)
                if (V.occupied_texel(V.texel_index(x, y, z))) {
                    const size_t b = (x >> 3) + V.bdim[0] * ((y >> 3) + V.bdim[1] * (z >> 3));
                    V.brick_nonempty[b] = 1;
                    V.brick_blocks[b] |= 1ull << (((x & 7) >> 1) + 4 * ((y & 7) >> 1) + 16 * ((z & 7) >> 1));
                }
    brick_radius(V);
    brick_cubes(V);
}

// world AABB padding: keeps the LBVH conservative with respect to the model-space slab
inline float pad_of(float v) { return std::fabs(v) * 9.5367431640625e-07f + 9.5367431640625e-07f; }

void prep_instance(const float* rec, uint32_t nvol, Inst& I) {
    std::memcpy(I.M, rec, 64);
    uint32_t tex;
    std::memcpy(&tex, &rec[15], 4);
    I.tex = tex;
    I.M[15] = 1.0f;
    bool ok = inverse4(I.M, I.inv);
    I.valid = ok && tex < nvol;
    float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
    for (int c = 0; c < 8; ++c) {
        float p[4] = {(c & 1) ? 0.5f : -0.5f, (c & 2) ? 0.5f : -0.5f, (c & 4) ? 0.5f : -0.5f, 1.0f};
        float w[4];
        mv4(I.M, p, w);
        for (int a = 0; a < 3; ++a) {
            lo[a] = gmin(lo[a], w[a]);
            hi[a] = gmax(hi[a], w[a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        I.lo[a] = lo[a] - pad_of(lo[a]);
        I.hi[a] = hi[a] + pad_of(hi[a]);
    }
    if (!ok) {
        for (int a = 0; a < 3; ++a) { I.lo[a] = 0.0f; I.hi[a] = 0.0f; }
    }
}

// ---- LBVH (Karras, "Maximizing Parallelism in the Construction of BVHs, Octrees, and
// k-d Trees", HPG 2012).  Node layout (32 B): lo[3], hi[3], left, right.  Internal nodes
// 0..n-2, leaf k (k-th in Morton order) is node n-1+k with left = instance, right = kMiss.
struct Node {
    float lo[3], hi[3];
    uint32_t left, right;
};
static_assert(sizeof(Node) == 32, "node is 32 bytes");

struct Bvh {
    std::vector<Node> nodes;
    std::vector<uint32_t> order;   // sorted instance ids
    std::vector<uint32_t> codes;   // sorted morton codes
    uint32_t n = 0;
};

inline int delta(const std::vector<uint32_t>& k, int i, int j, int n) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __builtin_clz((uint32_t)i ^ (uint32_t)j);
    return __builtin_clz(a ^ b);
}

void build_bvh(const std::vector<Inst>& inst, Bvh& B) {
    uint32_t n = (uint32_t)inst.size();
    B.n = n;
    B.nodes.clear();
    B.order.clear();
    B.codes.clear();
    if (n == 0) return;
    // scene bounds of the centroids
    float cmin[3] = {kInf, kInf, kInf}, cmax[3] = {-kInf, -kInf, -kInf};
    std::vector<float> cen(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            float c = (inst[i].lo[a] + inst[i].hi[a]) * 0.5f;
            cen[3 * i + a] = c;
            cmin[a] = gmin(cmin[a], c);
            cmax[a] = gmax(cmax[a], c);
        }
    std::vector<uint32_t> code(n);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int a = 0; a < 3; ++a) {
            float ext = cmax[a] - cmin[a];
            float u = ext > 0.0f ? (cen[3 * i + a] - cmin[a]) / ext : 0.0f;
            float s = std::floor(u * 1024.0f);
            int qi = (int)s;
            if (qi < 0) qi = 0;
            if (qi > 1023) qi = 1023;
            q[a] = (uint32_t)qi;
        }
        code[i] = morton30(q[0], q[1], q[2]);
    }
    std::vector<uint32_t> idx(n);
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    B.order = idx;
    B.codes.resize(n);
    for (uint32_t i = 0; i < n; ++i) B.codes[i] = code[idx[i]];
    B.nodes.assign(2 * (size_t)n - 1, Node{});
    std::vector<int> parent(2 * (size_t)n - 1, -1);
    for (uint32_t k = 0; k < n; ++k) {
        Node& L = B.nodes[n - 1 + k];
        const Inst& I = inst[idx[k]];
        for (int a = 0; a < 3; ++a) { L.lo[a] = I.lo[a]; L.hi[a] = I.hi[a]; }
        L.left = idx[k];
        L.right = kMiss;
    }
    const std::vector<uint32_t>& K = B.codes;
    int ni = (int)n;
    for (int i = 0; i < ni - 1; ++i) {
        int d = (delta(K, i, i + 1, ni) - delta(K, i, i - 1, ni)) >= 0 ? 1 : -1;
        int dmin = delta(K, i, i - d, ni);
        int lmax = 2;
        while (delta(K, i, i + lmax * d, ni) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (delta(K, i, i + (l + t) * d, ni) > dmin) l += t;
        int j = i + l * d;
        int dnode = delta(K, i, j, ni);
        int s = 0;
        // binary search for the split: largest s with delta(i, i + s*d) > dnode
        for (int div = 2;; div *= 2) {
            int tt = (l + div - 1) / div;
            if (delta(K, i, i + (s + tt) * d, ni) > dnode) s += tt;
            if (tt <= 1) break;
        }
        int gamma = i + s * d + (d < 0 ? d : 0);
        int first = d > 0 ? i : j;
        int last = d > 0 ? j : i;
        uint32_t left = (first == gamma) ? (uint32_t)(ni - 1 + gamma) : (uint32_t)gamma;
        uint32_t right = (last == gamma + 1) ? (uint32_t)(ni - 1 + gamma + 1) : (uint32_t)(gamma + 1);
        B.nodes[i].left = left;
        B.nodes[i].right = right;
        parent[left] = i;
        parent[right] = i;
    }
    // bottom-up refit (post-order over the topology; same result as the atomic GPU refit)
    std::vector<int> visits(n > 1 ? n - 1 : 1, 0);
    for (uint32_t k = 0; k < n; ++k) {
        int node = parent[n - 1 + k];
        while (node >= 0) {
            if (++visits[node] < 2) break;
            Node& N = B.nodes[node];
            const Node& a = B.nodes[N.left];
            const Node& b = B.nodes[N.right];
            for (int ax = 0; ax < 3; ++ax) {
                N.lo[ax] = gmin(a.lo[ax], b.lo[ax]);
                N.hi[ax] = gmax(a.hi[ax], b.hi[ax]);
            }
            node = parent[node];
        }
    }
}

// ---- ray / box ---------------------------------------------------------------------
struct Slab {
    float tn, tf;
    int axis;
    bool ok;
};
// Slab test of the ray o + t*d against [lo, hi]; tn/tf exact per-axis IEEE divisions.
inline Slab slab(const float o[3], const float d[3], const float lo[3], const float hi[3]) {
    Slab s{-kInf, kInf, 0, true};
    for (int a = 0; a < 3; ++a) {
        float tn, tf;
        if (d[a] != 0.0f) {
            float t1 = (lo[a] - o[a]) / d[a];
            float t2 = (hi[a] - o[a]) / d[a];
            tn = gmin(t1, t2);
            tf = gmax(t1, t2);
        } else {
            if (o[a] < lo[a] || o[a] > hi[a]) { s.ok = false; return s; }
            tn = -kInf;
            tf = kInf;
        }
        if (tn > s.tn) { s.tn = tn; s.axis = a; }
        if (tf < s.tf) s.tf = tf;
    }
    s.ok = s.tn <= s.tf;
    return s;
}

// LBVH node test: the ray's reciprocal direction is computed once per query and the
// slab distances are (bound - o) * inv_d; zero direction components test containment.
inline Slab slab_node(const float o[3], const float d[3], const float inv_d[3], const float lo[3], const float hi[3]) {
    Slab s{-kInf, kInf, 0, true};
    for (int a = 0; a < 3; ++a) {
        float tn, tf;
        if (d[a] != 0.0f) {
            float t1 = (lo[a] - o[a]) * inv_d[a];
            float t2 = (hi[a] - o[a]) * inv_d[a];
            tn = gmin(t1, t2);
            tf = gmax(t1, t2);
        } else {
            if (o[a] < lo[a] || o[a] > hi[a]) { s.ok = false; return s; }
            tn = -kInf;
            tf = kInf;
        }
        if (tn > s.tn) { s.tn = tn; s.axis = a; }
        if (tf < s.tf) s.tf = tf;
    }
    s.ok = s.tn <= s.tf;
    return s;
}

const float kUnitLo[3] = {-0.5f, -0.5f, -0.5f};
const float kUnitHi[3] = {0.5f, 0.5f, 0.5f};

// ---- DDA (trace.frag:66-87) --------------------------------------------------------
struct DdaResult {
    bool hit;
    int v[3];
    uint32_t steps;
    uint32_t mask;     // last step mask
    float t_last;      // side value of the last step (distance along the unit model dir from q)
    float q[3];        // start of the voxel walk that ended at v (hit point = q + t_last * unit dir)
    uint32_t rgba;
    uint8_t pal;
};

struct Counters {
    orc_stats* st;
    void brick(const Vol& V, int x, int y, int z) {
        if (!st) return;
        st->bricks++;
        if (V.brick_nonempty[(x >> 3) + V.bdim[0] * ((y >> 3) + V.bdim[1] * (z >> 3))]) st->bricks_nonempty++;
    }
};

// mpos: ray position in voxel units; v0: start voxel; mdir: model-space direction
DdaResult dda(const Vol& V, const float mpos[3], const int v0[3], const float mdir[3], Counters& C) {
    DdaResult r{};
    int voxel[3] = {v0[0], v0[1], v0[2]};
    int step[3];
    float delta_[3], side[3];
    float len = std::sqrt((mdir[0] * mdir[0] + mdir[1] * mdir[1]) + mdir[2] * mdir[2]);
    for (int a = 0; a < 3; ++a) {
        float sg = gsign(mdir[a]);
        step[a] = (int)sg;
        delta_[a] = std::fabs(len / mdir[a]);
        side[a] = ((sg * ((float)voxel[a] - mpos[a]) + sg * 0.5f) + 0.5f) * delta_[a];
    }
    uint32_t steps = 0;
    uint32_t max_steps = V.dim[0] + V.dim[1] + V.dim[2];
    uint32_t mask = 0;
    float t_last = 0.0f;
    int cur_b[3] = {-1, -1, -1};
    while (steps < max_steps && voxel[0] >= 0 && voxel[1] >= 0 && voxel[2] >= 0 && voxel[0] < (int)V.dim[0] &&
           voxel[1] < (int)V.dim[1] && voxel[2] < (int)V.dim[2]) {
        int b[3] = {voxel[0] >> 3, voxel[1] >> 3, voxel[2] >> 3};
        if (b[0] != cur_b[0] || b[1] != cur_b[1] || b[2] != cur_b[2]) {
            C.brick(V, voxel[0], voxel[1], voxel[2]);
            cur_b[0] = b[0]; cur_b[1] = b[1]; cur_b[2] = b[2];
        }
        uint32_t ti = V.texel_index(voxel[0], voxel[1], voxel[2]);
        uint8_t pal;
        uint32_t c = V.rgba(ti, &pal);
        if ((c >> 24) != 0u) {
            r.hit = true;
            r.rgba = c;
            r.pal = pal;
            break;
        }
        bool mx = side[0] <= gmin(side[1], side[2]);
        bool my = side[1] <= gmin(side[2], side[0]);
        bool mz = side[2] <= gmin(side[0], side[1]);
        t_last = mx ? side[0] : (my ? side[1] : side[2]);
        side[0] += (mx ? 1.0f : 0.0f) * delta_[0];
        side[1] += (my ? 1.0f : 0.0f) * delta_[1];
        side[2] += (mz ? 1.0f : 0.0f) * delta_[2];
        voxel[0] += (mx ? 1 : 0) * step[0];
        voxel[1] += (my ? 1 : 0) * step[1];
        voxel[2] += (mz ? 1 : 0) * step[2];
        mask = (mx ? 1u : 0u) | (my ? 2u : 0u) | (mz ? 4u : 0u);
        ++steps;
    }
    if (C.st) C.st->steps += steps;
    r.v[0] = voxel[0]; r.v[1] = voxel[1]; r.v[2] = voxel[2];
    r.steps = steps;
    r.mask = r.hit ? mask : 0;
    if (r.hit && steps == 0) r.mask = 0;
    r.t_last = t_last;
    r.q[0] = mpos[0]; r.q[1] = mpos[1]; r.q[2] = mpos[2];
    return r;
}

// Secondary-ray walk (no reference counterpart; this project's rule).  From position q in voxel v
// the ray crosses, one box per step, the largest box known to be empty that holds v: in an empty
// brick the cube of side K_o anchored at it towards the ray's octant o when K_o - 1 > r (its exit
// then lies further along the ray on every axis), else the cube of radius r centred on it; in a
// non-empty brick the 4^3 octant holding v if all its 2^3 blocks are empty, else the 2^3 block
// holding v if empty, else v itself unless v is occupied (the hit).  A step moves q to the
// box's nearest exit plane along the ray (ties: every axis at that distance): exit axes step
// into the voxel beyond the plane, the other axes take floor(q) clamped to the box.  Steps
// count boxes crossed; a brick is counted when the walk enters it.
DdaResult dda2(const Vol& V, const float p[3], const int v0[3], const float m[3], Counters& C) {
    DdaResult r{};
    const float len = std::sqrt((m[0] * m[0] + m[1] * m[1]) + m[2] * m[2]);
    float u[3], rinv[3];
    int st[3];
    for (int a = 0; a < 3; ++a) {
        st[a] = m[a] > 0.0f ? 1 : (m[a] < 0.0f ? -1 : 0);
        u[a] = m[a] / len;
        rinv[a] = len / m[a];
    }
    const int oct = (st[0] < 0 ? 1 : 0) | (st[1] < 0 ? 2 : 0) | (st[2] < 0 ? 4 : 0);
    int v[3] = {v0[0], v0[1], v0[2]};
    float q[3] = {p[0], p[1], p[2]};
    const int n[3] = {(int)V.dim[0], (int)V.dim[1], (int)V.dim[2]};
    uint32_t steps = 0, mask = 0;
    int64_t cur = -1;
    const uint32_t guard = 2 * (V.dim[0] + V.dim[1] + V.dim[2]) + 64;
    for (uint32_t it = 0; it < guard; ++it) {
        if (!(v[0] >= 0 && v[1] >= 0 && v[2] >= 0 && v[0] < n[0] && v[1] < n[1] && v[2] < n[2])) break;
        const int b8[3] = {v[0] & ~7, v[1] & ~7, v[2] & ~7};
        const uint32_t bi = (uint32_t)(b8[0] >> 3) + V.bdim[0] * ((uint32_t)(b8[1] >> 3) + V.bdim[1] * (uint32_t)(b8[2] >> 3));
        if ((int64_t)bi != cur) {
            cur = bi;
            C.brick(V, v[0], v[1], v[2]);
        }
        int lo[3], hi[3];
        if (!V.brick_nonempty[bi]) {
            // the anchored cube when K_o - 1 > r: its far faces then lie beyond the centred
            // cube's on every axis the ray moves along (K_o = r + 1: the same faces)
            const int rad = 8 * (int)V.brick_radius[bi];
            const int k8 = 8 * (int)V.brick_cube[oct][bi];
            const bool anchored = k8 - 8 > rad;
            for (int a = 0; a < 3; ++a) {
                lo[a] = anchored ? (st[a] < 0 ? b8[a] + 8 - k8 : b8[a]) : b8[a] - rad;
                hi[a] = anchored ? (st[a] < 0 ? b8[a] + 8 : b8[a] + k8) : b8[a] + 8 + rad;
            }
        } else {
            const uint64_t blk = V.brick_blocks[bi];
            const int lx = v[0] - b8[0], ly = v[1] - b8[1], lz = v[2] - b8[2];
            int sz;
            if (((blk >> (2 * (lx >> 2) + 8 * (ly >> 2) + 32 * (lz >> 2))) & 0x330033ull) == 0ull) {
                sz = 4;
            } else if (((blk >> ((lx >> 1) + 4 * (ly >> 1) + 16 * (lz >> 1))) & 1ull) == 0ull) {
                sz = 2;
            } else {
                uint8_t pal;
                const uint32_t c = V.rgba(V.texel_index(v[0], v[1], v[2]), &pal);
                if ((c >> 24) != 0u) {
                    r.hit = true;
                    r.rgba = c;
                    r.pal = pal;
                    break;
                }
                sz = 1;
            }
            for (int a = 0; a < 3; ++a) {
                lo[a] = v[a] & -sz;
                hi[a] = lo[a] + sz;
            }
        }
        float t[3];
        float tmin = kInf;
        for (int a = 0; a < 3; ++a) {
            t[a] = st[a] != 0 ? ((float)(st[a] > 0 ? hi[a] : lo[a]) - q[a]) * rinv[a] : kInf;
            tmin = gmin(tmin, t[a]);
        }
        uint32_t mk = 0;
        for (int a = 0; a < 3; ++a) {
            if (st[a] != 0 && t[a] <= tmin) {
                mk |= 1u << a;
                q[a] = (float)(st[a] > 0 ? hi[a] : lo[a]);
                v[a] = st[a] > 0 ? hi[a] : lo[a] - 1;
            } else {
                const float x = q[a] + tmin * u[a];
                q[a] = x;
                int f = (int)std::floor(x);
                f = f < lo[a] ? lo[a] : (f > hi[a] - 1 ? hi[a] - 1 : f);
                v[a] = f;
            }
        }
        mask = mk;
        ++steps;
    }
    if (C.st) C.st->steps += steps;
    r.v[0] = v[0]; r.v[1] = v[1]; r.v[2] = v[2];
    r.steps = steps;
    r.mask = (r.hit && steps > 0) ? mask : 0;
    r.t_last = 0.0f;
    for (int a = 0; a < 3; ++a) r.q[a] = q[a];
    return r;
}

// face from the last step: lowest stepped axis a; outward normal is -step[a]
inline uint8_t face_from_step(int axis, float dir_a) { return (uint8_t)(2 * axis + (dir_a > 0.0f ? 0 : 1)); }
inline int lowest_axis(uint32_t mask) { return (mask & 1u) ? 0 : ((mask & 2u) ? 1 : 2); }

// ---- frame constants ----------------------------------------------------------------
struct Frame {
    float P[16], C[16];
    float invP[16], invC[16], invCc[16], AB[16], PC[16];
    float O[3];
    float W, Hv;
    bool ok;
};

bool setup_frame(const float* P, const float* C, const orc_params* prm, Frame& F) {
    std::memcpy(F.P, P, 64);
    std::memcpy(F.C, C, 64);
    float Cc[16];
    std::memcpy(Cc, C, 64);
    Cc[12] = Cc[13] = Cc[14] = 0.0f;  // centered_camera[3][0..2] = 0 (trace.frag:55-58)
    bool ok = inverse4(P, F.invP) && inverse4(C, F.invC) && inverse4(Cc, F.invCc);
    mm4(F.invCc, F.invP, F.AB);  // inverse(centered_camera) * inverse_projection
    mm4(F.P, F.C, F.PC);         // push.projection * push.camera (trace.vert:45)
    float z[4] = {0, 0, 0, 1}, o[4];
    mv4(F.invC, z, o);           // cam_pos (trace.frag:51)
    F.O[0] = o[0]; F.O[1] = o[1]; F.O[2] = o[2];
    F.W = (float)prm->width;
    F.Hv = prm->viewport_square ? (float)prm->width : (float)prm->height;
    F.ok = ok;
    return ok;
}

// camera ray direction for the raster-proxy emulation (which proxy surface point the
// pixel sample sees); the shader's own ray_dir is recomputed from screen_position.
inline void camera_ray(const Frame& F, float sx, float sy, float D[3]) {
    float nx = (2.0f * sx) / F.W - 1.0f;
    float ny = (2.0f * sy) / F.Hv - 1.0f;
    float v[3] = {nx * F.invP[0], ny * F.invP[5], -1.0f};
    for (int r = 0; r < 3; ++r) D[r] = (F.invCc[0 * 4 + r] * v[0] + F.invCc[1 * 4 + r] * v[1]) + F.invCc[2 * 4 + r] * v[2];
}

inline void xform_point(const float* m, const float p[3], float o[3]) {
    float v[4] = {p[0], p[1], p[2], 1.0f}, r[4];
    mv4(m, v, r);
    o[0] = r[0]; o[1] = r[1]; o[2] = r[2];
}
inline void xform_dir(const float* m, const float d[3], float o[3]) {
    float v[4] = {d[0], d[1], d[2], 0.0f}, r[4];
    mv4(m, v, r);
    o[0] = r[0]; o[1] = r[1]; o[2] = r[2];
}
inline void normalize3(const float v[3], float o[3]) {
    float l = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    o[0] = v[0] / l; o[1] = v[1] / l; o[2] = v[2] / l;
}

struct Scene {
    std::vector<Vol> vols;
    std::vector<Inst> inst;
    Bvh bvh;
    const Frame* F;
};

// A hit of one query, enough to shade it and spawn the next ray.
struct Hit {
    bool hit = false;
    uint32_t inst = kMiss;
    float tn = kInf;        // ordering key: box entry parameter of the query ray
    DdaResult d{};
    float mpos[3]{};        // DDA start position (voxel units)
    float mdir[3]{};        // DDA model direction
    uint8_t face = 0xFF;
};

// Run one instance for a query entering from outside (primary or LBVH-found secondary).
// Returns true if the instance's DDA hits; fills h.  `primary` adds the rasterizer clip
// and recomputes the shader ray from screen_position.
bool enter_instance(const Scene& S, uint32_t i, const float O[3], const float D[3], bool primary, Hit& h,
                    Counters& C) {
    const Inst& I = S.inst[i];
    if (!I.valid) return false;
    float o[3], d[3];
    xform_point(I.inv, O, o);
    xform_dir(I.inv, D, d);
    Slab s = slab(o, d, kUnitLo, kUnitHi);
    if (!s.ok || !(s.tn > 0.0f)) return false;   // camera/origin inside -> back faces only (culled)
    float mp[3];
    for (int a = 0; a < 3; ++a) {
        if (a == s.axis) {
            mp[a] = d[a] > 0.0f ? -0.5f : 0.5f;
        } else {
            float x = o[a] + s.tn * d[a];
            mp[a] = x < -0.5f ? -0.5f : (x > 0.5f ? 0.5f : x);
        }
    }
    const Frame& F = *S.F;
    float mdir[3];
    if (primary) {
        float w[3];
        xform_point(I.M, mp, w);                      // world_position (trace.vert:44)
        float w4[4] = {w[0], w[1], w[2], 1.0f}, sp[4];
        mv4(F.PC, w4, sp);                            // screen_position (trace.vert:45)
        if (!(sp[2] >= 0.0f && sp[2] <= sp[3])) return false;   // clip volume 0 <= z <= w
        float q[4];
        mv4(F.AB, sp, q);                             // inverse(centered_camera)*inverse_projection*sp
        float dir[3];
        float qv[3] = {q[0], q[1], q[2]};
        normalize3(qv, dir);                          // trace.frag:60
        xform_dir(I.inv, dir, mdir);                  // trace.frag:66
    } else {
        mdir[0] = d[0]; mdir[1] = d[1]; mdir[2] = d[2];
    }
    if (C.st) C.st->instances++;
    const Vol& V = S.vols[I.tex];
    float mpos[3];
    int v0[3];
    for (int a = 0; a < 3; ++a) {
        mpos[a] = (mp[a] + 0.5f) * V.size[a];                       // trace.frag:67
        v0[a] = (int)std::floor(gmin(mpos[a], V.size[a] - 1.0f));   // trace.frag:69
    }
    DdaResult r = primary ? dda(V, mpos, v0, mdir, C) : dda2(V, mpos, v0, mdir, C);
    if (!r.hit) return false;
    h.hit = true;
    h.inst = i;
    h.tn = s.tn;
    h.d = r;
    for (int a = 0; a < 3; ++a) { h.mpos[a] = mpos[a]; h.mdir[a] = mdir[a]; }
    if (r.steps == 0) h.face = face_from_step(s.axis, d[s.axis]);
    else {
        int a = lowest_axis(r.mask);
        h.face = face_from_step(a, mdir[a]);
    }
    return true;
}

inline bool better(const Hit& c, const Hit& best) {
    return c.tn < best.tn || (c.tn == best.tn && c.inst < best.inst);
}

// pruning bound of the LBVH: a node is skipped only when its entry parameter exceeds the
// best box entry by a generous relative margin (keeps the result equal to brute force).
inline float prune_bound(float t) { return t + std::fabs(t) * 0.0009765625f + 0.0009765625f; }

// Query over all instances except `skip`: closest (any_hit = false) or any hit.
// use_bvh = false: brute force in instance order; use_bvh = true: LBVH traversal in the
// GPU's order (counts nodes).
Hit query(const Scene& S, const float O[3], const float D[3], bool primary, uint32_t skip, bool any_hit,
          bool use_bvh, Counters& C) {
    Hit best;
    if (!use_bvh) {
        for (uint32_t i = 0; i < (uint32_t)S.inst.size(); ++i) {
            if (i == skip) continue;
            Hit h;
            if (enter_instance(S, i, O, D, primary, h, C) && better(h, best)) {
                best = h;
                if (any_hit) return best;
            }
        }
        return best;
    }
    const Bvh& B = S.bvh;
    if (B.n == 0) return best;
    const float inv_d[3] = {1.0f / D[0], 1.0f / D[1], 1.0f / D[2]};
    uint32_t stack[64];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        uint32_t ni = stack[--sp];
        const Node& N = B.nodes[ni];
        if (C.st) C.st->nodes++;
        Slab s = slab_node(O, D, inv_d, N.lo, N.hi);
        if (!s.ok || s.tf < 0.0f) continue;
        if (best.hit && s.tn > prune_bound(best.tn)) continue;
        if (N.right == kMiss) {
            uint32_t i = N.left;
            if (i == skip) continue;
            Hit h;
            if (enter_instance(S, i, O, D, primary, h, C) && better(h, best)) {
                best = h;
                if (any_hit) return best;
            }
        } else {
            stack[sp++] = N.right;
            stack[sp++] = N.left;
        }
    }
    return best;
}

// Continue a query inside instance `i` from a voxel face (bounce / shadow origin).
bool continue_in_instance(const Scene& S, uint32_t i, const float p[3], const int vstart[3], const float wdir[3],
                          Hit& h, Counters& C) {
    const Inst& I = S.inst[i];
    const Vol& V = S.vols[I.tex];
    float mdir[3];
    xform_dir(I.inv, wdir, mdir);
    if (C.st) C.st->instances++;
    DdaResult r = dda2(V, p, vstart, mdir, C);
    if (!r.hit) return false;
    h.hit = true;
    h.inst = i;
    h.tn = 0.0f;
    h.d = r;
    for (int a = 0; a < 3; ++a) { h.mpos[a] = p[a]; h.mdir[a] = mdir[a]; }
    h.face = 0xFF;  // filled by caller when steps == 0
    if (r.steps > 0) {
        int a = lowest_axis(r.mask);
        h.face = face_from_step(a, mdir[a]);
    }
    return true;
}

struct Surface {
    float p[3];      // hit point, voxel units of the hit instance
    int vstart[3];   // empty neighbour voxel on the face side
    float X[3];      // world position
    float n[3];      // world unit normal
    uint32_t rgba;
};

// Hit point and frame of a hit (shared by shading and the next query).
void surface_of(const Scene& S, const Hit& h, Surface& sf) {
    const Inst& I = S.inst[h.inst];
    const Vol& V = S.vols[I.tex];
    float len = std::sqrt((h.mdir[0] * h.mdir[0] + h.mdir[1] * h.mdir[1]) + h.mdir[2] * h.mdir[2]);
    int fa = h.face >> 1;
    bool neg = (h.face & 1) == 0;  // outward normal along -axis
    for (int a = 0; a < 3; ++a) {
        float u = h.mdir[a] / len;
        float x = h.d.q[a] + h.d.t_last * u;
        float lo = (float)h.d.v[a], hi = (float)(h.d.v[a] + 1);
        if (a == fa) x = neg ? lo : hi;
        else x = x < lo ? lo : (x > hi ? hi : x);
        sf.p[a] = x;
        sf.vstart[a] = h.d.v[a] + (a == fa ? (neg ? -1 : 1) : 0);
    }
    float mp[3];
    for (int a = 0; a < 3; ++a) mp[a] = sf.p[a] / V.size[a] - 0.5f;
    xform_point(I.M, mp, sf.X);
    // world normal: inverse-transpose of the model matrix applied to the model normal
    float nm = neg ? -1.0f : 1.0f;
    float nw[3] = {I.inv[0 * 4 + fa] * nm, I.inv[1 * 4 + fa] * nm, I.inv[2 * 4 + fa] * nm};
    normalize3(nw, sf.n);
    sf.rgba = h.d.rgba;
}

// secondary query from a surface: continue in the origin instance, then the rest of the scene
Hit secondary(const Scene& S, uint32_t origin, const Surface& sf, uint8_t origin_face, const float w[3], bool any_hit,
              bool use_bvh, Counters& C) {
    if (C.st) C.st->rays++;
    Hit h;
    if (continue_in_instance(S, origin, sf.p, sf.vstart, w, h, C)) {
        if (h.face == 0xFF) h.face = origin_face;
        if (C.st) C.st->hits++;
        return h;
    }
    h = query(S, sf.X, w, false, origin, any_hit, use_bvh, C);
    if (h.hit && C.st) C.st->hits++;
    return h;
}

// Cosine-weighted direction about n: n + uniform unit vector (Marsaglia 1972 disc
// rejection, at most 16 tries), normalized; falls back to n.
void cosine_dir(Rng& rng, const float n[3], float w[3]) {
    for (int t = 0; t < 16; ++t) {
        float u = 2.0f * rng.next() - 1.0f;
        float v = 2.0f * rng.next() - 1.0f;
        float s = u * u + v * v;
        if (s >= 1.0f || s == 0.0f) continue;
        float q = std::sqrt(1.0f - s);
        float r[3] = {2.0f * u * q, 2.0f * v * q, 1.0f - 2.0f * s};
        float a[3] = {n[0] + r[0], n[1] + r[1], n[2] + r[2]};
        float dd = (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2];
        float cn = (a[0] * n[0] + a[1] * n[1]) + a[2] * n[2];
        if (!(dd > 1e-6f) || !(cn > 1e-3f * std::sqrt(dd))) continue;
        normalize3(a, w);
        return;
    }
    w[0] = n[0]; w[1] = n[1]; w[2] = n[2];
}

inline void decode_rgb_d(uint32_t c, double o[3]) {
    o[0] = srgb_to_linear_d(c & 255u);
    o[1] = srgb_to_linear_d((c >> 8) & 255u);
    o[2] = srgb_to_linear_d((c >> 16) & 255u);
}

void write_hit(const Hit& h, orc_hit& out) {
    std::memset(&out, 0, sizeof(out));
    if (!h.hit) {
        out.inst = kMiss;
        out.voxel = kMiss;
        out.face = 0xFF;
        return;
    }
    // 15-bit coordinates split 10 + 5 over voxel and vhi; steps saturate instead of wrapping
    const uint32_t x = h.d.v[0], y = h.d.v[1], z = h.d.v[2];
    out.inst = h.inst;
    out.voxel = (x % 1024u) + (y % 1024u) * 1024u + (z % 1024u) * 1048576u;
    out.vhi = (uint16_t)((x / 1024u) % 32u + ((y / 1024u) % 32u) * 32u + ((z / 1024u) % 32u) * 1024u);
    out.steps = (uint16_t)std::min<uint32_t>(h.d.steps, 65535u);
    out.face = h.face;
    out.pal = h.d.pal;
    out.mask = (uint8_t)h.d.mask;
    out.flags = 1;
    if (x >= 32768u || y >= 32768u || z >= 32768u) out.flags |= 4;
    if (h.d.steps > 65535u) out.flags |= 8;
}

}  // namespace

// =====================================================================================
extern "C" {

uint32_t orc_pcg_hash(uint32_t v) { return pcg_hash(v); }
uint32_t orc_rng_seed(uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t sample) {
    return rng_seed(seed, frame, pixel, sample);
}
uint32_t orc_morton30(uint32_t x, uint32_t y, uint32_t z) { return morton30(x, y, z); }
int orc_inverse4(const float* m, float* out) { return inverse4(m, out) ? 1 : 0; }
float orc_srgb_decode(uint32_t c8) { return srgb().dec[c8 & 255u]; }
uint32_t orc_srgb_encode(float lin) {
    const float* t = srgb().enc_thr;
    uint32_t k = 0;
    while (k < 255 && lin >= t[k]) ++k;  // count of thresholds <= lin
    return k;
}
uint32_t orc_nearest_texel(int v, uint32_t dim) { return nearest_texel(v, (float)dim, dim); }

// Bounce sampler test hook: n directions about the unit normal nrm, sample i drawn by
// cosine_dir from the stream a path of pixel i at sample 0 would start with (rng_seed(seed,
// 0, i, 0)).  out: 3 floats per sample.
void orc_cosine_dir_samples(uint32_t seed, const float* nrm, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        Rng rng{rng_seed(seed, 0u, i, 0u)};
        cosine_dir(rng, nrm, out + 3 * (size_t)i);
    }
}

// DDA test hooks: run the reference DDA (level 0) or the two-level secondary DDA (level 1)
// on one volume from position mpos / start voxel v0 along mdir.  out: hit, x, y, z, steps, mask.
int orc_dda_test(const orc_volume* v, int level, const float* mpos, const int* v0, const float* mdir, int* out,
                 float* t_last_q) {
    Vol V;
    init_vol(V, v);
    Counters C{nullptr};
    DdaResult r = level == 0 ? dda(V, mpos, v0, mdir, C) : dda2(V, mpos, v0, mdir, C);
    out[0] = r.hit ? 1 : 0;
    out[1] = r.v[0]; out[2] = r.v[1]; out[3] = r.v[2];
    out[4] = (int)r.steps;
    out[5] = (int)r.mask;
    if (t_last_q) {
        t_last_q[0] = r.t_last;
        t_last_q[1] = r.q[0]; t_last_q[2] = r.q[1]; t_last_q[3] = r.q[2];
    }
    return 0;
}

// Instance preparation: records (16 floats each, render.rs:74-78 layout) ->
// inverse matrices, padded world AABBs (lo,hi), texture ids, valid flags.
void orc_instance_prep(uint32_t n, const float* records, uint32_t nvol, float* inv_out, float* aabb_out,
                       uint32_t* tex_out, uint8_t* valid_out) {
    for (uint32_t i = 0; i < n; ++i) {
        Inst I;
        prep_instance(records + 16 * (size_t)i, nvol, I);
        if (inv_out) std::memcpy(inv_out + 16 * (size_t)i, I.inv, 64);
        if (aabb_out)
            for (int a = 0; a < 3; ++a) { aabb_out[6 * i + a] = I.lo[a]; aabb_out[6 * i + 3 + a] = I.hi[a]; }
        if (tex_out) tex_out[i] = I.tex;
        if (valid_out) valid_out[i] = I.valid ? 1 : 0;
    }
}

// LBVH from instance records.  nodes_out: (2n-1) x 32 B; order_out / codes_out: n each.
int orc_lbvh_build(uint32_t n, const float* records, uint32_t nvol, void* nodes_out, uint32_t* order_out,
                   uint32_t* codes_out) {
    std::vector<Inst> inst(n);
    for (uint32_t i = 0; i < n; ++i) prep_instance(records + 16 * (size_t)i, nvol, inst[i]);
    Bvh B;
    build_bvh(inst, B);
    if (n == 0) return 0;
    if (nodes_out) std::memcpy(nodes_out, B.nodes.data(), B.nodes.size() * sizeof(Node));
    if (order_out) std::memcpy(order_out, B.order.data(), 4 * (size_t)n);
    if (codes_out) std::memcpy(codes_out, B.codes.data(), 4 * (size_t)n);
    return (int)B.nodes.size();
}

// Dense -> 8^3 bricks in the product's layout: brick map (u32 per brick, x-fastest brick
// order; empty = 0xE0000000 | radius << 24 | (octant cube side - 1) << 3o, others numbered in
// brick order), occupancy masks (64 B per non-empty brick, bit x + 8y + 64z of the brick), payload (u8 index or RGBA8 per voxel,
// voxel x + 8y + 64z).  Voxel v of the brick volume holds the texel the reference samples
// for v (nearest_texel remap).  Returns the non-empty brick count.
uint32_t orc_brick(const orc_volume* v, uint32_t* map_out, uint64_t* masks_out, uint8_t* payload_out) {
    Vol V;
    init_vol(V, v);
    uint32_t nb = V.bdim[0] * V.bdim[1] * V.bdim[2];
    uint32_t bpv = v->fmt == 0 ? 1 : 4;
    uint32_t next = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        uint32_t bx = b % V.bdim[0], by = (b / V.bdim[0]) % V.bdim[1], bz = b / (V.bdim[0] * V.bdim[1]);
        if (!V.brick_nonempty[b]) {
            if (map_out) {
                uint32_t e = kEmptyBrick | ((uint32_t)V.brick_radius[b] << 24);
                for (int o = 0; o < 8; ++o) e |= (uint32_t)(V.brick_cube[o][b] - 1) << (3 * o);
                map_out[b] = e;
            }
            continue;
        }
        uint32_t id = next++;
        if (map_out) map_out[b] = id;
        uint64_t m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t z = 0; z < 8; ++z)
            for (uint32_t y = 0; y < 8; ++y)
                for (uint32_t x = 0; x < 8; ++x) {
                    uint32_t gx = bx * 8 + x, gy = by * 8 + y, gz = bz * 8 + z;
                    uint32_t li = x + 8 * y + 64 * z;
                    uint32_t val = 0;
                    bool occ = false;
                    if (gx < V.dim[0] && gy < V.dim[1] && gz < V.dim[2]) {
                        uint32_t ti = V.texel_index(gx, gy, gz);
                        occ = V.occupied_texel(ti);
                        if (v->fmt == 0) val = v->data[ti];
                        else std::memcpy(&val, v->data + 4ull * ti, 4);
                    }
                    if (occ) m[z] |= 1ull << (x + 8 * y);
                    if (payload_out) std::memcpy(payload_out + (size_t)id * 512 * bpv + (size_t)li * bpv, &val, bpv);
                }
        if (masks_out) std::memcpy(masks_out + 8 * (size_t)id, m, 64);
    }
    return next;
}

// 2^3-block masks of every brick (0 for empty bricks), brick order.
void orc_brick_blocks(const orc_volume* v, uint64_t* out) {
    Vol V;
    init_vol(V, v);
    std::memcpy(out, V.brick_blocks.data(), V.brick_blocks.size() * 8);
}

// Persistent oracle scene: volumes prepared once (remaps, brick occupancy), instances, LBVH.
struct OrcScene {
    Frame F;
    Scene S;
    orc_params prm;
    bool use_bvh;
};

static void render_rows(OrcScene& OS, uint32_t y0, uint32_t y1, orc_hit* hits_out, float* rad_f32, double* rad_f64,
                        orc_stats* stats) {
    const Scene& S = OS.S;
    const Frame& F = OS.F;
    const orc_params* prm = &OS.prm;
    const int use_bvh = OS.use_bvh ? 1 : 0;
    Counters Cn{stats};
    const SrgbTables& T = srgb();
    const uint32_t W = prm->width;
    const uint32_t spp = prm->spp ? prm->spp : 1;
    const double inv_pi = 1.0 / 3.14159265358979323846;
    const float sun[3] = {prm->sun_dir[0], prm->sun_dir[1], prm->sun_dir[2]};
    for (uint32_t py = y0; py < y1; ++py)
        for (uint32_t px = 0; px < W; ++px) {
            size_t pix = (size_t)py * W + px;
            double acc[3] = {0, 0, 0};
            float accf[4] = {0, 0, 0, 0};
            Hit first;
            bool first_shadowed = false;
            for (uint32_t s = 0; s < spp; ++s) {
                Rng rng{rng_seed(prm->seed, prm->frame, (uint32_t)pix, s)};
                float jx = 0.5f, jy = 0.5f;
                if (prm->jitter) { jx = rng.next(); jy = rng.next(); }
                float D[3];
                camera_ray(F, (float)px + jx, (float)py + jy, D);
                if (stats) stats->rays++;
                Hit h = query(S, F.O, D, true, kMiss, false, use_bvh != 0, Cn);
                if (h.hit && stats) stats->hits++;
                if (s == 0) first = h;
                if (prm->mode == 0) {
                    // trace.frag:79 colour blended over the sky clear (pipeline.c:129-137)
                    if (h.hit) {
                        uint32_t c = h.d.rgba;
                        float a = (float)(c >> 24) / 255.0f;
                        float rgb[3] = {T.dec[c & 255u], T.dec[(c >> 8) & 255u], T.dec[(c >> 16) & 255u]};
                        for (int k = 0; k < 3; ++k) accf[k] += rgb[k] * a + prm->sky[k] * (1.0f - a);
                        accf[3] += a;
                    } else {
                        for (int k = 0; k < 3; ++k) accf[k] += prm->sky[k];
                        accf[3] += 1.0f;
                    }
                    continue;
                }
                // modes 1/2: FP32 geometry, FP64 radiance
                double beta[3] = {1, 1, 1};
                double L[3] = {0, 0, 0};
                uint32_t max_b = prm->mode == 2 ? prm->max_bounces : 0;
                Hit cur = h;
                for (uint32_t depth = 0;; ++depth) {
                    if (!cur.hit) {
                        for (int k = 0; k < 3; ++k) L[k] += beta[k] * (double)prm->sky[k];
                        break;
                    }
                    Surface sf;
                    surface_of(S, cur, sf);
                    double rho[3];
                    decode_rgb_d(sf.rgba, rho);
                    float cosl = (sf.n[0] * sun[0] + sf.n[1] * sun[1]) + sf.n[2] * sun[2];
                    if (cosl > 0.0f) {
                        Hit sh = secondary(S, cur.inst, sf, cur.face, sun, true, use_bvh != 0, Cn);
                        if (!sh.hit) {
                            for (int k = 0; k < 3; ++k)
                                L[k] += beta[k] * (rho[k] * ((double)prm->sun_radiance[k] * ((double)cosl * inv_pi)));
                        } else if (s == 0 && depth == 0) {
                            first_shadowed = true;
                        }
                    } else if (s == 0 && depth == 0) {
                        first_shadowed = true;
                    }
                    if (depth >= max_b) break;
                    float w[3];
                    cosine_dir(rng, sf.n, w);
                    for (int k = 0; k < 3; ++k) beta[k] *= rho[k];
                    cur = secondary(S, cur.inst, sf, cur.face, w, false, use_bvh != 0, Cn);
                }
                for (int k = 0; k < 3; ++k) acc[k] += L[k];
            }
            if (hits_out) {
                write_hit(first, hits_out[pix]);
                if (prm->mode >= 1 && first.hit && first_shadowed) hits_out[pix].flags |= 2;
            }
            if (prm->mode == 0 && rad_f32) {
                for (int k = 0; k < 4; ++k) rad_f32[4 * pix + k] = accf[k] / (float)spp;
            }
            if (prm->mode != 0 && rad_f64) {
                for (int k = 0; k < 3; ++k) rad_f64[4 * pix + k] = acc[k] / (double)spp;
                rad_f64[4 * pix + 3] = 1.0;
            }
        }
}

void* orc_scene_create(uint32_t nvol, const orc_volume* vols, uint32_t ninst, const float* records, const float* P,
                       const float* C, const orc_params* prm, int use_bvh) {
    OrcScene* OS = new OrcScene();
    OS->prm = *prm;
    OS->use_bvh = use_bvh != 0;
    if (!setup_frame(P, C, prm, OS->F)) {
        delete OS;
        return nullptr;
    }
    OS->S.F = &OS->F;
    OS->S.vols.resize(nvol);
    for (uint32_t i = 0; i < nvol; ++i) init_vol(OS->S.vols[i], &vols[i]);
    OS->S.inst.resize(ninst);
    for (uint32_t i = 0; i < ninst; ++i) prep_instance(records + 16 * (size_t)i, nvol, OS->S.inst[i]);
    if (use_bvh) build_bvh(OS->S.inst, OS->S.bvh);
    return OS;
}

void orc_scene_destroy(void* h) { delete static_cast<OrcScene*>(h); }

// Renders rows [y0, y1) with `threads` host threads (interleaved rows; counters summed).
int orc_scene_render(void* h, uint32_t y0, uint32_t y1, uint32_t threads, orc_hit* hits_out, float* rad_f32,
                     double* rad_f64, orc_stats* stats) {
    OrcScene& OS = *static_cast<OrcScene*>(h);
    if (y1 > OS.prm.height) y1 = OS.prm.height;
    if (threads <= 1 || y1 - y0 < 2) {
        render_rows(OS, y0, y1, hits_out, rad_f32, rad_f64, stats);
        return 0;
    }
    std::vector<orc_stats> part(threads);
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < threads; ++t) {
        part[t] = orc_stats{};
        th.emplace_back([&, t]() {
            for (uint32_t y = y0 + t; y < y1; y += threads)
                render_rows(OS, y, y + 1, hits_out, rad_f32, rad_f64, stats ? &part[t] : nullptr);
        });
    }
    for (auto& x : th) x.join();
    if (stats)
        for (auto& p : part) {
            stats->rays += p.rays; stats->nodes += p.nodes; stats->instances += p.instances;
            stats->bricks += p.bricks; stats->bricks_nonempty += p.bricks_nonempty; stats->hits += p.hits;
            stats->steps += p.steps;
        }
    return 0;
}

// Renders an explicit list of rows with `threads` host threads (rows dealt round robin).
int orc_scene_render_rows(void* h, const uint32_t* rows, uint32_t nrows, uint32_t threads, orc_hit* hits_out,
                          float* rad_f32, double* rad_f64, orc_stats* stats) {
    OrcScene& OS = *static_cast<OrcScene*>(h);
    if (threads == 0) threads = 1;
    std::vector<orc_stats> part(threads);
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < threads; ++t) {
        part[t] = orc_stats{};
        th.emplace_back([&, t]() {
            for (uint32_t i = t; i < nrows; i += threads)
                if (rows[i] < OS.prm.height)
                    render_rows(OS, rows[i], rows[i] + 1, hits_out, rad_f32, rad_f64, stats ? &part[t] : nullptr);
        });
    }
    for (auto& x : th) x.join();
    if (stats)
        for (auto& p : part) {
            stats->rays += p.rays; stats->nodes += p.nodes; stats->instances += p.instances;
            stats->bricks += p.bricks; stats->bricks_nonempty += p.bricks_nonempty; stats->hits += p.hits;
            stats->steps += p.steps;
        }
    return 0;
}

// Full frame in one call (create, render rows [y0, y1) single-threaded, destroy).
int orc_render(uint32_t nvol, const orc_volume* vols, uint32_t ninst, const float* records, const float* P,
               const float* C, const orc_params* prm, int use_bvh, uint32_t y0, uint32_t y1, orc_hit* hits_out,
               float* rad_f32, double* rad_f64, orc_stats* stats) {
    void* h = orc_scene_create(nvol, vols, ninst, records, P, C, prm, use_bvh);
    if (!h) return -1;
    int rc = orc_scene_render(h, y0, y1, 1, hits_out, rad_f32, rad_f64, stats);
    orc_scene_destroy(h);
    return rc;
}

}  // extern "C"
