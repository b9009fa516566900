/* vtscene.h — host-side utilities of libvtscene (no GPU): the seeded scene generator used by
 * the workloads and the .vox asset reader.  Not part of the reference's FFI surface. */
#ifndef VTSCENE_H
#define VTSCENE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

uint32_t vt_pcg_hash(uint32_t v);
/* Value-noise fBm field (4 octaves, lacunarity 2, gain 0.5, base period n/4) in [0, 1). */
int vt_gen_noise_field(uint32_t dx, uint32_t dy, uint32_t dz, uint32_t seed, uint32_t octaves, float* out);
/* u8 index volume with exactly round(fill * n) occupied voxels (indices 1..255), optional boxes
 * (6 u32 each: x0 y0 z0 x1 y1 z1) forced in. */
int vt_gen_volume(uint32_t dx, uint32_t dy, uint32_t dz, uint32_t seed, double fill, uint32_t octaves, uint32_t nbox,
                  const uint32_t* boxes, uint8_t* out);
void vt_gen_palette(uint32_t seed, int opaque, uint32_t* out);
void vt_expand_palette(const uint8_t* idx, size_t n, const uint32_t* palette, uint32_t* out);

/* MagicaVoxel .vox (replaces the dot_vox load in src/voxel/magica_voxel.rs:18-44).  Returns the
 * number of models, or -1 for a malformed buffer. */
int32_t vt_vox_models(const uint8_t* buf, uint64_t n);
/* Model k: dims3 = (x, y, z) size; indices (x*y*z bytes, may be NULL) = the array the reference
 * passes to add_texture as (w, h, d) = (x, y, z), as palette indices; palette256 (may be NULL)
 * = RGBA8 palette, r in the low byte, entry 0 transparent.  0 = ok, -1 = malformed or no such
 * model, -2 = a voxel falls outside the chunk (models that are not cubes; the reference
 * panics there), -3 = no RGBA chunk (the format's default palette is not supported). */
int32_t vt_vox_model(const uint8_t* buf, uint64_t n, uint32_t k, uint32_t* dims3, uint8_t* indices,
                     uint32_t* palette256);

/* Streamed terrain (src/gen/terrain.rs:37-94 structure; its noise-rs OpenSimplex/Billow are
 * replaced by seeded FP64 value noise).  Colour of world voxel (x, y, z) as RGBA8, r in the low
 * byte, 0 = empty. */
uint32_t vt_terrain_voxel(int32_t x, int32_t y, int32_t z, uint32_t seed);
/* Chunk (cx, cy, cz) = voxels 16c .. 16c+15 per axis into rgba (4096 u32, x fastest).  Returns 1
 * if a voxel is occupied, 0 if the chunk is empty (terrain.rs returns None), -1 on error. */
int32_t vt_terrain_chunk(int32_t cx, int32_t cy, int32_t cz, uint32_t seed, uint32_t* rgba);
/* n chunks (coords = 3n ints) on host threads: rgba = n x 4096 u32, filled[i] = chunk i occupied. */
int32_t vt_terrain_chunks(const int32_t* coords, uint32_t n, uint32_t seed, uint32_t* rgba, uint8_t* filled);

#ifdef __cplusplus
}
#endif

#endif /* VTSCENE_H */
