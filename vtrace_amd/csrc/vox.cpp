// vox.cpp — MagicaVoxel .vox reader producing the bytes the reference hands to add_texture.
//
// The reference loads assets with the dot_vox crate (4.1.0, not vendored) in
// src/voxel/magica_voxel.rs:18-44: one RawDynamicChunk of size (x, y, z) per model, file voxel
// (x, y, z, i) stored at chunk (x, size.y - 1 - z, y) (magica_voxel.rs:31-36), colour =
// palette[i - 1] (the format's colour indices run 1..255 over palette entries 0..254; index 0
// takes entry 0, as a saturating subtraction does).  render.rs:246-262 passes the chunk's raw
// array (index z + dim_z * (y + dim_y * x), rawchunk.rs:287-296) to add_texture as
// (w, h, d) = (dim_x, dim_y, dim_z).  This reader returns that array as u8 palette indices
// plus a 256-entry RGBA8 palette (entry 0 transparent), which the index+palette upload
// expands to the same texels.
//
// File layout (MagicaVoxel format, version 150): "VOX " magic, u32 version, then chunk "MAIN"
// whose children include, per model, a SIZE chunk (3 x i32) followed by an XYZI chunk (u32 n,
// then n x 4 u8), and one RGBA chunk (256 x 4 u8).  Every chunk is id[4], u32 content bytes,
// u32 children bytes, content, children (little endian).  Other chunks are skipped.
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Model {
    int32_t size[3];
    uint64_t xyzi;   // offset of the voxel records
    uint32_t count;
};

struct VoxFile {
    std::vector<Model> models;
    bool has_palette = false;
    uint64_t palette = 0;
};

bool u32_at(const uint8_t* b, uint64_t n, uint64_t off, uint32_t* v) {
    if (off > n || n - off < 4) return false;
    *v = (uint32_t)b[off] | ((uint32_t)b[off + 1] << 8) | ((uint32_t)b[off + 2] << 16) | ((uint32_t)b[off + 3] << 24);
    return true;
}

// 0 = ok, -1 = malformed
int parse(const uint8_t* b, uint64_t n, VoxFile& f) {
    uint32_t version, cn, cm;
    if (!b || n < 20 || std::memcmp(b, "VOX ", 4) != 0 || !u32_at(b, n, 4, &version)) return -1;
    if (std::memcmp(b + 8, "MAIN", 4) != 0 || !u32_at(b, n, 12, &cn) || !u32_at(b, n, 16, &cm)) return -1;
    uint64_t off = 20 + (uint64_t)cn;
    const uint64_t end = off + cm;
    if (end > n) return -1;
    bool pending = false;
    int32_t size[3] = {0, 0, 0};
    while (off < end) {
        if (end - off < 12) return -1;
        uint32_t ccn = 0, ccm = 0;
        u32_at(b, n, off + 4, &ccn);
        u32_at(b, n, off + 8, &ccm);
        const uint64_t content = off + 12;
        const uint64_t next = content + (uint64_t)ccn + ccm;
        if (next > end) return -1;
        if (std::memcmp(b + off, "SIZE", 4) == 0) {
            if (ccn < 12) return -1;
            for (int a = 0; a < 3; ++a) {
                uint32_t v = 0;
                u32_at(b, n, content + 4 * a, &v);
                size[a] = (int32_t)v;
                if (size[a] <= 0 || size[a] > 2048) return -1;
            }
            pending = true;
        } else if (std::memcmp(b + off, "XYZI", 4) == 0) {
            uint32_t cnt = 0;
            if (!pending || ccn < 4 || !u32_at(b, n, content, &cnt) || (uint64_t)cnt * 4 > (uint64_t)ccn - 4) return -1;
            Model m;
            std::memcpy(m.size, size, sizeof(size));
            m.xyzi = content + 4;
            m.count = cnt;
            f.models.push_back(m);
            pending = false;
        } else if (std::memcmp(b + off, "RGBA", 4) == 0) {
            if (ccn < 1024) return -1;
            f.has_palette = true;
            f.palette = content;
        }
        off = next;
    }
    return 0;
}

}  // namespace

extern "C" {

int32_t vt_vox_models(const uint8_t* buf, uint64_t n) {
    VoxFile f;
    if (parse(buf, n, f) != 0) return -1;
    return (int32_t)f.models.size();
}

int32_t vt_vox_model(const uint8_t* buf, uint64_t n, uint32_t k, uint32_t* dims3, uint8_t* indices,
                     uint32_t* palette256) {
    VoxFile f;
    if (parse(buf, n, f) != 0 || k >= f.models.size()) return -1;
    const Model& m = f.models[k];
    if (dims3)
        for (int a = 0; a < 3; ++a) dims3[a] = (uint32_t)m.size[a];
    if (!f.has_palette) return -3;  // the format's built-in default palette is not supported
    if (palette256) {
        palette256[0] = 0u;
        for (uint32_t j = 1; j < 256; ++j) u32_at(buf, n, f.palette + 4ull * (j - 1), &palette256[j]);
    }
    if (!indices) return 0;
    const int64_t dx = m.size[0], dy = m.size[1], dz = m.size[2];
    std::memset(indices, 0, (size_t)(dx * dy * dz));
    for (uint32_t v = 0; v < m.count; ++v) {
        const uint8_t* r = buf + m.xyzi + 4ull * v;
        const int64_t X = r[0], Y = dy - 1 - (int64_t)r[2], Z = r[1];
        if (X >= dx || Y < 0 || Y >= dy || Z >= dz) return -2;  // outside the chunk (non-cubic model)
        indices[Z + dz * (Y + dy * X)] = r[3] ? r[3] : 1;
    }
    return 0;
}

}  // extern "C"
