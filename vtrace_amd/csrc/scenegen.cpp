// scenegen.cpp — seeded synthetic scene content (host C++).
//
// Generates the voxel volumes and palettes the renderer is benchmarked and tested on.
// The reference renders MagicaVoxel assets and OpenSimplex terrain (src/world.rs:141-198,
// src/gen/terrain.rs:27-65); the workloads here replace them with seeded value noise so
// every run is reproducible without assets.  Output layout is the reference's texel
// order for add_texture (lib/memory.c:286-385): x fastest, index x + w*(y + h*z).
//
// Value noise: lattice values u01(pcg_hash(...)) at integer points, smoothstep-weighted
// trilinear interpolation at voxel centres, fBm over `octaves` octaves with lacunarity 2 and
// gain 0.5, base period = min(dims)/4 voxels.  Occupancy takes exactly round(fill * N)
// voxels: the ones with the smallest noise value, ties broken by linear index (a radix
// select over the FP32 bit patterns, which order like the values for non-negative floats).
// Occupied voxels get a palette index from U{1..255}; empty voxels are 0.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint32_t pcg_hash(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
inline float u01(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

template <class F>
void parallel_for(size_t n, F&& f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    std::atomic<size_t> next{0};
    const size_t chunk = std::max<size_t>(1, n / (nt * 16));
    auto worker = [&]() {
        for (;;) {
            size_t b = next.fetch_add(chunk);
            if (b >= n) break;
            size_t e = std::min(n, b + chunk);
            f(b, e);
        }
    };
    for (unsigned t = 1; t < nt; ++t) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
}

struct Lattice {
    int n[3];
    float period;
    std::vector<float> v;
    float at(int x, int y, int z) const { return v[(size_t)x + (size_t)n[0] * ((size_t)y + (size_t)n[1] * z)]; }
};

Lattice make_lattice(const uint32_t dim[3], float period, uint32_t seed, uint32_t octave) {
    Lattice L;
    L.period = period;
    for (int a = 0; a < 3; ++a) L.n[a] = (int)std::ceil((double)dim[a] / period) + 2;
    L.v.resize((size_t)L.n[0] * L.n[1] * L.n[2]);
    uint32_t os = pcg_hash(seed ^ (octave * 0x9E3779B9u));
    for (int z = 0; z < L.n[2]; ++z)
        for (int y = 0; y < L.n[1]; ++y)
            for (int x = 0; x < L.n[0]; ++x)
                L.v[(size_t)x + (size_t)L.n[0] * ((size_t)y + (size_t)L.n[1] * z)] =
                    u01(pcg_hash((uint32_t)x + pcg_hash((uint32_t)y + pcg_hash((uint32_t)z + os))));
    return L;
}

inline float smooth(float f) { return f * f * (3.0f - 2.0f * f); }
inline float lerp(float a, float b, float t) { return a + (b - a) * t; }

float noise_at(const Lattice& L, uint32_t x, uint32_t y, uint32_t z) {
    float p[3] = {((float)x + 0.5f) / L.period, ((float)y + 0.5f) / L.period, ((float)z + 0.5f) / L.period};
    int i[3];
    float w[3];
    for (int a = 0; a < 3; ++a) {
        float fl = std::floor(p[a]);
        i[a] = (int)fl;
        w[a] = smooth(p[a] - fl);
    }
    float c00 = lerp(L.at(i[0], i[1], i[2]), L.at(i[0] + 1, i[1], i[2]), w[0]);
    float c10 = lerp(L.at(i[0], i[1] + 1, i[2]), L.at(i[0] + 1, i[1] + 1, i[2]), w[0]);
    float c01 = lerp(L.at(i[0], i[1], i[2] + 1), L.at(i[0] + 1, i[1], i[2] + 1), w[0]);
    float c11 = lerp(L.at(i[0], i[1] + 1, i[2] + 1), L.at(i[0] + 1, i[1] + 1, i[2] + 1), w[0]);
    return lerp(lerp(c00, c10, w[1]), lerp(c01, c11, w[1]), w[2]);
}

// Occupy exactly k voxels with the smallest values (ties: lowest linear index first).
void select_smallest(const std::vector<float>& val, size_t k, uint32_t seed, uint8_t* out) {
    const size_t n = val.size();
    if (k == 0) { std::memset(out, 0, n); return; }
    if (k >= n) k = n;
    auto bits = [&](size_t i) { uint32_t b; std::memcpy(&b, &val[i], 4); return b; };
    // pass 1: high 16 bits
    std::vector<uint64_t> h1(65536, 0);
    {
        std::vector<std::vector<uint64_t>> part;
        unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        part.assign(nt, std::vector<uint64_t>(65536, 0));
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t]() {
                size_t b = n * t / nt, e = n * (t + 1) / nt;
                auto& h = part[t];
                for (size_t i = b; i < e; ++i) h[bits(i) >> 16]++;
            });
        for (auto& t : th) t.join();
        for (auto& h : part)
            for (int i = 0; i < 65536; ++i) h1[i] += h[i];
    }
    size_t acc = 0;
    uint32_t hb = 0;
    for (; hb < 65536; ++hb) {
        if (acc + h1[hb] >= k) break;
        acc += h1[hb];
    }
    std::vector<uint64_t> h2(65536, 0);
    for (size_t i = 0; i < n; ++i) {
        uint32_t b = bits(i);
        if ((b >> 16) == hb) h2[b & 0xFFFFu]++;
    }
    uint32_t lb = 0;
    for (; lb < 65536; ++lb) {
        if (acc + h2[lb] >= k) break;
        acc += h2[lb];
    }
    const uint32_t T = (hb << 16) | lb;
    size_t take_equal = k - acc;  // number of voxels with value bits == T to occupy
    // write: occupied if bits < T, or bits == T among the first take_equal in index order
    std::vector<size_t> eq;
    for (size_t i = 0; i < n && eq.size() < take_equal; ++i)
        if (bits(i) == T) eq.push_back(i);
    parallel_for(n, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            uint32_t v = bits(i);
            bool occ = v < T;
            out[i] = occ ? (uint8_t)(1u + pcg_hash((uint32_t)i ^ pcg_hash(seed + 0x5bd1e995u)) % 255u) : 0;
        }
    });
    for (size_t i : eq) out[i] = (uint8_t)(1u + pcg_hash((uint32_t)i ^ pcg_hash(seed + 0x5bd1e995u)) % 255u);
}

}  // namespace

extern "C" {

uint32_t vt_pcg_hash(uint32_t v) { return pcg_hash(v); }

// FP32 fBm value at every voxel (for tests / inspection); out has dx*dy*dz floats.
int vt_gen_noise_field(uint32_t dx, uint32_t dy, uint32_t dz, uint32_t seed, uint32_t octaves, float* out) {
    if (!dx || !dy || !dz || !out || octaves == 0 || octaves > 8) return -1;
    const uint32_t dim[3] = {dx, dy, dz};
    float base = (float)std::max(1u, std::min(dx, std::min(dy, dz)) / 4u);
    std::vector<Lattice> lat;
    for (uint32_t o = 0; o < octaves; ++o) lat.push_back(make_lattice(dim, base / (float)(1u << o), seed, o));
    parallel_for((size_t)dy * dz, [&](size_t b, size_t e) {
        for (size_t yz = b; yz < e; ++yz) {
            uint32_t y = (uint32_t)(yz % dy), z = (uint32_t)(yz / dy);
            for (uint32_t x = 0; x < dx; ++x) {
                float s = 0.0f, amp = 1.0f;
                for (uint32_t o = 0; o < octaves; ++o) {
                    s += amp * noise_at(lat[o], x, y, z);
                    amp *= 0.5f;
                }
                out[(size_t)x + (size_t)dx * yz] = s;
            }
        }
    });
    return 0;
}

// u8 palette-index volume: exactly round(fill * N) occupied voxels (smallest fBm values).
// boxes (optional): nbox x 6 u32 (x0,y0,z0,x1,y1,z1 half-open) forced occupied first.
int vt_gen_volume(uint32_t dx, uint32_t dy, uint32_t dz, uint32_t seed, double fill, uint32_t octaves,
                  uint32_t nbox, const uint32_t* boxes, uint8_t* out) {
    if (!dx || !dy || !dz || !out || !(fill >= 0.0 && fill <= 1.0)) return -1;
    const size_t n = (size_t)dx * dy * dz;
    std::vector<float> val(n);
    if (vt_gen_noise_field(dx, dy, dz, seed, octaves, val.data()) != 0) return -1;
    // non-box voxels are shifted above every box voxel (box voxels = 0.0)
    parallel_for(n, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) val[i] += 1.0f;
    });
    for (uint32_t bi = 0; bi < nbox; ++bi) {
        const uint32_t* B = boxes + 6 * (size_t)bi;
        for (uint32_t z = B[2]; z < std::min(B[5], dz); ++z)
            for (uint32_t y = B[1]; y < std::min(B[4], dy); ++y)
                for (uint32_t x = B[0]; x < std::min(B[3], dx); ++x) val[(size_t)x + (size_t)dx * (y + (size_t)dy * z)] = 0.0f;
    }
    size_t k = (size_t)std::llround(fill * (double)n);
    select_smallest(val, k, seed, out);
    return 0;
}

// 256-entry RGBA8 palette (r in the low byte).  opaque: alpha 255 for entries 1..255 and
// 0 for entry 0 (occupancy == index != 0); otherwise alpha is uniform random too.
void vt_gen_palette(uint32_t seed, int opaque, uint32_t* out) {
    for (uint32_t i = 0; i < 256; ++i) {
        uint32_t h = pcg_hash(i ^ pcg_hash(seed ^ 0xA511E9B3u));
        uint32_t rgb = h & 0x00FFFFFFu;
        uint32_t a = opaque ? (i == 0 ? 0u : 255u) : (pcg_hash(h) >> 24);
        out[i] = rgb | (a << 24);
    }
}

// u8 index volume -> RGBA8 texels through a palette (the reference's add_texture format).
void vt_expand_palette(const uint8_t* idx, size_t n, const uint32_t* palette, uint32_t* out) {
    parallel_for(n, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) out[i] = palette[idx[i]];
    });
}

}  // extern "C"
