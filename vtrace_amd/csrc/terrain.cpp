// terrain.cpp — streamed terrain chunks with the structure of src/gen/terrain.rs.
//
// terrain.rs:37-65 colours world voxel (x, y, z): empty outside the ball x^2+y^2+z^2 <= 50^2;
// "dirt" (255, 200, 100) where (x, y - 4, z) lies outside that ball, "stone" (150, 150, 150)
// elsewhere; occupied where a noise signal at 0.1 * voxel is > 0, with colour
// (u8)(tone * base), tone = 0.5 * billow + 0.5, alpha 255.  terrain.rs:67-94 fills chunk
// (cx, cy, cz) with voxels 16c .. 16c+15 per axis, stored x fastest (at_mut(z, y, x)), and
// returns no chunk when none is occupied.
//
// The reference's signal and billow are noise-rs 0.7.0 OpenSimplex and Billow (seed 0), a
// crate not available here; both are replaced by seeded value noise on the integer lattice,
// evaluated in FP64 like the reference's samples:
//   value(p, s) = trilinear interpolation, smoothstep weights, of lattice values
//                 2 * u01(pcg(X + pcg(Y + pcg(Z + pcg(s))))) - 1 at the 8 corners around p
//   signal(p)   = value(p, seed)
//   billow(p)   = sum_{o<6} 0.5^o * (2 |value(2^o p, seed + 1 + o)| - 1) / sum_{o<6} 0.5^o
// The chunk layout, ball, colour bands and tone mapping are the reference's.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr int kChunk = 16;         // CHUNK_VOXEL_SIZE (pager.rs:24)
constexpr double kRadius2 = 2500;  // 50^2 (terrain.rs:41)

inline uint32_t pcg(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

inline double lattice(int64_t x, int64_t y, int64_t z, uint32_t s) {
    const uint32_t h = pcg((uint32_t)x + pcg((uint32_t)y + pcg((uint32_t)z + pcg(s))));
    return 2.0 * ((double)(h >> 8) * (1.0 / 16777216.0)) - 1.0;
}

inline double smooth(double f) { return f * f * (3.0 - 2.0 * f); }
inline double lerp(double a, double b, double t) { return a + (b - a) * t; }

double value(double px, double py, double pz, uint32_t s) {
    const double fx = std::floor(px), fy = std::floor(py), fz = std::floor(pz);
    const int64_t X = (int64_t)fx, Y = (int64_t)fy, Z = (int64_t)fz;
    const double wx = smooth(px - fx), wy = smooth(py - fy), wz = smooth(pz - fz);
    const double c00 = lerp(lattice(X, Y, Z, s), lattice(X + 1, Y, Z, s), wx);
    const double c10 = lerp(lattice(X, Y + 1, Z, s), lattice(X + 1, Y + 1, Z, s), wx);
    const double c01 = lerp(lattice(X, Y, Z + 1, s), lattice(X + 1, Y, Z + 1, s), wx);
    const double c11 = lerp(lattice(X, Y + 1, Z + 1, s), lattice(X + 1, Y + 1, Z + 1, s), wx);
    return lerp(lerp(c00, c10, wy), lerp(c01, c11, wy), wz);
}

double billow(double px, double py, double pz, uint32_t seed) {
    double sum = 0.0, amp = 1.0, norm = 0.0, f = 1.0;
    for (uint32_t o = 0; o < 6; ++o) {
        const double v = value(px * f, py * f, pz * f, seed + 1u + o);
        sum = sum + amp * (2.0 * std::fabs(v) - 1.0);
        norm = norm + amp;
        amp = amp * 0.5;
        f = f * 2.0;
    }
    return sum / norm;
}

// (u8)(x) of Rust's `as u8` on an f64 in [0, 255]: truncation toward zero (saturating).
inline uint32_t to_u8(double x) {
    if (!(x > 0.0)) return 0u;
    if (x >= 255.0) return 255u;
    return (uint32_t)x;
}

// terrain.rs:37-65 with the replaced noise; RGBA8 with r in the low byte
uint32_t voxel(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    const double wx = x, wy = y, wz = z;
    if (wx * wx + wy * wy + wz * wz > kRadius2) return 0u;
    const double ay = (double)(y - 4);
    const bool surface = wx * wx + ay * ay + wz * wz > kRadius2;
    const double px = wx * 0.1, py = wy * 0.1, pz = wz * 0.1;
    if (!(value(px, py, pz, seed) > 0.0)) return 0u;
    const double tone = 0.5 * billow(px, py, pz, seed) + 0.5;
    const double c0 = surface ? 255.0 : 150.0, c1 = surface ? 200.0 : 150.0, c2 = surface ? 100.0 : 150.0;
    return to_u8(tone * c0) | (to_u8(tone * c1) << 8) | (to_u8(tone * c2) << 16) | (255u << 24);
}

// true if some voxel of the chunk can lie inside the ball (cheap reject of far chunks)
bool may_touch_ball(int32_t cx, int32_t cy, int32_t cz) {
    const int32_t c[3] = {cx, cy, cz};
    double d2 = 0.0;
    for (int a = 0; a < 3; ++a) {
        const double lo = (double)c[a] * kChunk, hi = lo + (kChunk - 1);
        const double q = lo > 0.0 ? lo : (hi < 0.0 ? hi : 0.0);
        d2 += q * q;
    }
    return d2 <= kRadius2;
}

int chunk(int32_t cx, int32_t cy, int32_t cz, uint32_t seed, uint32_t* out) {
    if (!may_touch_ball(cx, cy, cz)) {
        std::memset(out, 0, sizeof(uint32_t) * kChunk * kChunk * kChunk);
        return 0;
    }
    int any = 0;
    for (int z = 0; z < kChunk; ++z)
        for (int y = 0; y < kChunk; ++y)
            for (int x = 0; x < kChunk; ++x) {
                const uint32_t c = voxel(x + cx * kChunk, y + cy * kChunk, z + cz * kChunk, seed);
                any |= c != 0u;
                out[x + kChunk * (y + kChunk * z)] = c;
            }
    return any;
}

}  // namespace

extern "C" {

uint32_t vt_terrain_voxel(int32_t x, int32_t y, int32_t z, uint32_t seed) { return voxel(x, y, z, seed); }

int32_t vt_terrain_chunk(int32_t cx, int32_t cy, int32_t cz, uint32_t seed, uint32_t* rgba) {
    if (!rgba) return -1;
    return chunk(cx, cy, cz, seed, rgba);
}

int32_t vt_terrain_chunks(const int32_t* coords, uint32_t n, uint32_t seed, uint32_t* rgba, uint8_t* filled) {
    if (!coords || !rgba || !filled) return -1;
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if (nt > n) nt = n ? n : 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([=]() {
            for (uint32_t i = t; i < n; i += nt)
                filled[i] = (uint8_t)chunk(coords[3 * i], coords[3 * i + 1], coords[3 * i + 2], seed,
                                           rgba + (size_t)i * kChunk * kChunk * kChunk);
        });
    for (auto& x : th) x.join();
    return 0;
}

}  // extern "C"
