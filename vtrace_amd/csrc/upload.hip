// upload.hip — dense voxel volume -> 8^3 bricks (the GPU half of add_texture).
//
// Replaces the reference's staging copy into an R8G8B8A8_SRGB 3-D image
// (lib/memory.c:286-385, lib/command.c:132-134).  The dense input is the reference's
// texel order (x fastest).  Voxel v of the bricked volume holds the texel that
// NEAREST/CLAMP sampling at v/size selects (shaders/trace.frag:77, lib/descriptor.c:100-105);
// the per-axis remap tables carry that choice (identity for power-of-two sizes).
//
// Several volumes of one size can be bricked together (the streamed 16^3 terrain chunks):
// their bricks are numbered volume after volume, one scan numbers the non-empty bricks of all
// of them into shared pools, and each volume's map holds ids relative to its first brick in
// the pools, so every volume is laid out exactly as if it had been uploaded alone.
//
// Kernels:
//   k_brick_classify  one 512-thread block per brick: occupancy bit per voxel
//                     (alpha > 0, trace.frag:78), one ballot per z-slice -> 8 x u64 mask,
//                     and the 2^3-block mask (one bit per 2x2x2 block holding a voxel);
//   k_scan_*          exclusive scan of the non-empty flags -> brick ids in brick order;
//   k_radius_*        empty-space radius of every empty brick: breadth-first layers of
//                     26-neighbour dilation from the non-empty bricks (in place, layer k marks
//                     the still-unreached bricks next to layer k-1);
//   k_cube_*          per empty brick and octant, the side of the largest empty cube with the
//                     brick as its corner: layer k grows the side from k to k + 1 where the
//                     7 neighbours towards the octant all have sides >= k (in place: sides only
//                     grow, so a neighbour updated during the layer decides the same way);
//   k_brick_scatter   brick map (id, or VT_EMPTY_BRICK with radius and cube sides), masks,
//                     block masks and payload of non-empty bricks into the pools.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "vt_internal.h"

namespace {

__device__ inline uint32_t texel_rgba(const uint8_t* dense, uint32_t fmt, const uint32_t* palette, size_t ti,
                                      uint32_t* raw) {
    if (fmt == 0) {
        uint32_t i = dense[ti];
        *raw = i;
        return palette[i];
    }
    uint32_t c = reinterpret_cast<const uint32_t*>(dense)[ti];
    *raw = c;
    return c;
}

struct BrickArgs {
    const uint8_t* dense;      // nvol volumes, one after another
    const uint32_t* palette;
    const uint16_t* remap[3];
    uint32_t dim[3];
    uint32_t bdim[3];
    uint32_t fmt;
    uint32_t nbv;              // bricks per volume
    size_t vol_texels;         // texels per volume
};

// brick b of the batch -> (texel offset of its volume, brick coordinates inside the volume)
__device__ inline size_t brick_coords(const BrickArgs& a, uint32_t b, uint32_t* bx, uint32_t* by, uint32_t* bz) {
    const uint32_t k = b / a.nbv, lb = b - k * a.nbv;
    *bx = lb % a.bdim[0];
    *by = (lb / a.bdim[0]) % a.bdim[1];
    *bz = lb / (a.bdim[0] * a.bdim[1]);
    return (size_t)k * a.vol_texels;
}

// 2x2 footprints of a z-slice's occupancy (bit x + 8y) -> 16 bits, bit bx + 4by for the
// 2x2 column (2bx.., 2by..) of the slice
__device__ inline uint32_t slice_blocks(uint64_t m) {
    uint64_t t = m | (m >> 1);
    t |= t >> 8;
    uint32_t r = 0;
    for (uint32_t by = 0; by < 4; ++by)
        for (uint32_t bx = 0; bx < 4; ++bx) r |= (uint32_t)((t >> (2u * bx + 16u * by)) & 1u) << (bx + 4u * by);
    return r;
}

__global__ __launch_bounds__(512) void k_brick_classify(BrickArgs a, uint64_t* masks_all, uint64_t* blocks_all,
                                                        uint32_t* flags) {
    const uint32_t b = blockIdx.x;
    uint32_t bx, by, bz;
    const size_t vbase = brick_coords(a, b, &bx, &by, &bz);
    const uint32_t t = threadIdx.x;  // x + 8y + 64z inside the brick
    const uint32_t gx = bx * 8 + (t & 7), gy = by * 8 + ((t >> 3) & 7), gz = bz * 8 + (t >> 6);
    bool occ = false;
    if (gx < a.dim[0] && gy < a.dim[1] && gz < a.dim[2]) {
        size_t ti = vbase + (size_t)a.remap[0][gx] +
                    (size_t)a.dim[0] * ((size_t)a.remap[1][gy] + (size_t)a.dim[1] * a.remap[2][gz]);
        uint32_t raw;
        occ = (texel_rgba(a.dense, a.fmt, a.palette, ti, &raw) >> 24) != 0u;
    }
    uint64_t m = __ballot(occ);  // wave t>>6 holds z-slice t>>6
    __shared__ uint32_t any;
    __shared__ uint32_t blk[2];   // 2^3-block mask: slices 2bz, 2bz+1 -> bits 16bz.. (bz = 0, 1 | 2, 3)
    if (t == 0) { any = 0; blk[0] = 0; blk[1] = 0; }
    __syncthreads();
    if ((t & 63) == 0) {
        masks_all[(size_t)b * 8 + (t >> 6)] = m;
        if (m) {
            atomicOr(&any, 1u);
            const uint32_t bz = (t >> 6) >> 1;
            atomicOr(&blk[bz >> 1], slice_blocks(m) << (16u * (bz & 1u)));
        }
    }
    __syncthreads();
    if (t == 0) {
        flags[b] = any;
        blocks_all[b] = (uint64_t)blk[0] | ((uint64_t)blk[1] << 32);
    }
}

// cube sides of all 8 octants packed 4 bits each (0 for a non-empty brick, 1 for an empty one)
__global__ void k_cube_init(const uint32_t* flags, uint32_t* cube, uint32_t nb) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) cube[b] = flags[b] ? 0u : 0x11111111u;
}

// layer k over nvol volumes of bx*by*bz bricks each (neighbours within the same volume; outside
// the volume counts as a side of VT_BRICK_CUBE_CAP)
__global__ void k_cube_step(uint32_t* cube, uint32_t bx, uint32_t by, uint32_t bz, uint32_t nvol, uint32_t k) {
    const uint32_t gb = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nbv = bx * by * bz;
    if (gb >= nbv * nvol) return;
    const uint32_t cur = cube[gb];
    if (cur == 0u) return;
    const uint32_t base = (gb / nbv) * nbv, b = gb - base;
    const int x = (int)(b % bx), y = (int)((b / bx) % by), z = (int)(b / (bx * by));
    uint32_t out = cur;
    for (uint32_t o = 0; o < 8; ++o) {
        if (((cur >> (4u * o)) & 15u) != k) continue;
        const int dx = (o & 1u) ? -1 : 1, dy = (o & 2u) ? -1 : 1, dz = (o & 4u) ? -1 : 1;
        bool grow = true;
        for (int n = 1; n < 8 && grow; ++n) {
            const int X = x + ((n & 1) ? dx : 0), Y = y + ((n & 2) ? dy : 0), Z = z + ((n & 4) ? dz : 0);
            if (X < 0 || Y < 0 || Z < 0 || X >= (int)bx || Y >= (int)by || Z >= (int)bz) continue;
            const uint32_t nk = (cube[base + (uint32_t)X + bx * ((uint32_t)Y + by * (uint32_t)Z)] >> (4u * o)) & 15u;
            grow = nk >= k;
        }
        if (grow) out += 1u << (4u * o);
    }
    if (out != cur) cube[gb] = out;
}

__global__ void k_radius_init(const uint32_t* flags, uint32_t* dist, uint32_t nb) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) dist[b] = flags[b] ? 0u : 255u;
}

// layer k over nvol volumes of bx*by*bz bricks each (neighbours within the same volume)
__global__ void k_radius_step(uint32_t* dist, uint32_t bx, uint32_t by, uint32_t bz, uint32_t nvol, uint32_t k) {
    const uint32_t gb = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nbv = bx * by * bz;
    if (gb >= nbv * nvol || dist[gb] != 255u) return;
    const uint32_t base = (gb / nbv) * nbv, b = gb - base;
    const int x = (int)(b % bx), y = (int)((b / bx) % by), z = (int)(b / (bx * by));
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int X = x + dx, Y = y + dy, Z = z + dz;
                if (X < 0 || Y < 0 || Z < 0 || X >= (int)bx || Y >= (int)by || Z >= (int)bz) continue;
                if (dist[base + (uint32_t)X + bx * ((uint32_t)Y + by * (uint32_t)Z)] == k - 1u) {
                    dist[gb] = k;
                    return;
                }
            }
}

__global__ __launch_bounds__(512) void k_brick_scatter(BrickArgs a, const uint64_t* masks_all,
                                                       const uint64_t* blocks_all, const uint32_t* ids,
                                                       const uint32_t* flags, const uint32_t* dist, const uint32_t* cube,
                                                       uint32_t* brick_map, uint64_t* masks, uint64_t* blocks,
                                                       uint8_t* payload) {
    const uint32_t b = blockIdx.x;
    const uint32_t t = threadIdx.x;
    const bool nonempty = flags[b] != 0;
    const uint32_t id = ids[b];                            // pool slot
    const uint32_t first = ids[(b / a.nbv) * a.nbv];        // pool slot of the volume's first brick
    if (t == 0) {
        uint32_t e = id - first;
        if (!nonempty) {
            e = VT_EMPTY_BRICK | ((min(dist[b], VT_BRICK_RADIUS_CAP + 1u) - 1u) << 24);
            for (uint32_t o = 0; o < 8; ++o) e |= (((cube[b] >> (4u * o)) & 15u) - 1u) << (3u * o);
        }
        brick_map[b] = e;
    }
    if (!nonempty) return;
    if (t < 8) masks[(size_t)id * 8 + t] = masks_all[(size_t)b * 8 + t];
    if (t == 8) blocks[id] = blocks_all[b];
    uint32_t bx, by, bz;
    const size_t vbase = brick_coords(a, b, &bx, &by, &bz);
    const uint32_t gx = bx * 8 + (t & 7), gy = by * 8 + ((t >> 3) & 7), gz = bz * 8 + (t >> 6);
    uint32_t raw = 0;
    if (gx < a.dim[0] && gy < a.dim[1] && gz < a.dim[2]) {
        size_t ti = vbase + (size_t)a.remap[0][gx] +
                    (size_t)a.dim[0] * ((size_t)a.remap[1][gy] + (size_t)a.dim[1] * a.remap[2][gz]);
        (void)texel_rgba(a.dense, a.fmt, a.palette, ti, &raw);  // payload keeps the texel as stored
    }
    if (a.fmt == 0) payload[(size_t)id * 512 + t] = (uint8_t)raw;
    else reinterpret_cast<uint32_t*>(payload)[(size_t)id * 512 + t] = raw;
}

// ---- exclusive scan of u32 flags (3 phases, 1024 elements per block) ----------------
constexpr int kScanBlock = 1024;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[kScanBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
        uint32_t s = lane < kScanBlock / 64 ? wsum[lane] : 0;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(s, o, 64);
            if (lane >= o) s += y;
        }
        if (lane < kScanBlock / 64) wsum[lane] = s;
    }
    __syncthreads();
    uint32_t incl = x + (w > 0 ? wsum[w - 1] : 0u);
    *total = wsum[kScanBlock / 64 - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t* in, uint32_t n, uint32_t* block_sums) {
    uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    uint32_t v = i < n ? (in[i] != 0u ? 1u : 0u) : 0u;
    uint32_t total;
    block_exclusive_scan(v, &total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_blocks(uint32_t* block_sums, uint32_t nb, uint32_t* grand_total) {
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += kScanBlock) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < nb ? block_sums[i] : 0u;
        uint32_t total;
        uint32_t ex = block_exclusive_scan(v, &total);
        uint32_t c = carry;
        if (i < nb) block_sums[i] = ex + c;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *grand_total = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const uint32_t* in, uint32_t n, const uint32_t* block_sums,
                                                            uint32_t* out) {
    uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    uint32_t v = i < n ? (in[i] != 0u ? 1u : 0u) : 0u;
    uint32_t total;
    uint32_t ex = block_exclusive_scan(v, &total);
    if (i < n) out[i] = ex + block_sums[blockIdx.x];
}

}  // namespace

// Exclusive scan of (flag != 0) over n entries; *total_dev receives the count.
hipError_t vt_scan_flags(const uint32_t* flags, uint32_t n, uint32_t* out, uint32_t* scratch_blocks, uint32_t* total_dev,
                         hipStream_t s) {
    uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
    if (nb == 0) nb = 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanBlock), 0, s, flags, n, scratch_blocks);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanBlock), 0, s, scratch_blocks, nb, total_dev);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanBlock), 0, s, flags, n, scratch_blocks, out);
    return hipGetLastError();
}

// Brick nvol dense device volumes of one size (packed one after another).  Allocates one set
// of pools for all of them (after one sync to learn the count); out[0] owns the allocations.
int vt_brick_volume(const uint8_t* dense_dev, uint32_t fmt, const uint32_t* palette_dev, const uint16_t* remap_dev[3],
                    const uint32_t dim[3], uint32_t nvol, VolumeStore* out, hipStream_t s) {
    if (nvol == 0) return -1;
    BrickArgs a{};
    a.dense = dense_dev;
    a.palette = palette_dev;
    for (int k = 0; k < 3; ++k) {
        a.remap[k] = remap_dev[k];
        a.dim[k] = dim[k];
        a.bdim[k] = (dim[k] + 7) / 8;
    }
    a.fmt = fmt;
    a.nbv = a.bdim[0] * a.bdim[1] * a.bdim[2];
    a.vol_texels = (size_t)dim[0] * dim[1] * dim[2];
    const uint64_t nb64 = (uint64_t)a.nbv * nvol;
    if (nb64 >= 0x7FFFFFFFull) return -1;
    const uint32_t nb = (uint32_t)nb64;
    uint64_t *masks_all = nullptr, *blocks_all = nullptr;
    uint32_t *flags = nullptr, *ids = nullptr, *blocks = nullptr, *total = nullptr, *dist = nullptr, *cube = nullptr;
    const uint32_t nblk = (nb + kScanBlock - 1) / kScanBlock + 1;
    auto free_temps = [&]() {
        hipFree(masks_all); hipFree(blocks_all); hipFree(flags); hipFree(ids); hipFree(blocks); hipFree(total);
        hipFree(dist); hipFree(cube);
    };
    if (hipMalloc(&masks_all, (size_t)nb * 64) != hipSuccess || hipMalloc(&blocks_all, (size_t)nb * 8) != hipSuccess ||
        hipMalloc(&flags, (size_t)nb * 4) != hipSuccess || hipMalloc(&ids, (size_t)nb * 4) != hipSuccess ||
        hipMalloc(&blocks, (size_t)nblk * 4) != hipSuccess || hipMalloc(&total, 4) != hipSuccess ||
        hipMalloc(&dist, (size_t)nb * 4) != hipSuccess || hipMalloc(&cube, (size_t)nb * 4) != hipSuccess) {
        free_temps();
        return -1;
    }
    hipLaunchKernelGGL(k_brick_classify, dim3(nb), dim3(512), 0, s, a, masks_all, blocks_all, flags);
    vt_scan_flags(flags, nb, ids, blocks, total, s);
    hipLaunchKernelGGL(k_radius_init, dim3((nb + 255) / 256), dim3(256), 0, s, flags, dist, nb);
    const uint32_t layers = std::min<uint32_t>(VT_BRICK_RADIUS_CAP + 1u,
                                               std::max(a.bdim[0], std::max(a.bdim[1], a.bdim[2])));
    for (uint32_t k = 1; k <= layers; ++k)
        hipLaunchKernelGGL(k_radius_step, dim3((nb + 255) / 256), dim3(256), 0, s, dist, a.bdim[0], a.bdim[1],
                           a.bdim[2], nvol, k);
    hipLaunchKernelGGL(k_cube_init, dim3((nb + 255) / 256), dim3(256), 0, s, flags, cube, nb);
    for (uint32_t k = 1; k < VT_BRICK_CUBE_CAP; ++k)
        hipLaunchKernelGGL(k_cube_step, dim3((nb + 255) / 256), dim3(256), 0, s, cube, a.bdim[0], a.bdim[1],
                           a.bdim[2], nvol, k);
    // pool slot of each volume's first brick (ids[k * nbv]) and the total
    std::vector<uint32_t> first(nvol + 1, 0);
    hipMemcpy2DAsync(first.data(), 4, ids, (size_t)a.nbv * 4, 4, nvol, hipMemcpyDeviceToHost, s);
    hipMemcpyAsync(&first[nvol], total, 4, hipMemcpyDeviceToHost, s);
    hipError_t e = hipStreamSynchronize(s);
    int rc = 0;
    if (e != hipSuccess) rc = -1;
    const uint32_t count = first[nvol];
    const size_t bpv = fmt == 0 ? 1 : 4;
    uint32_t* pool_map = nullptr;
    uint32_t* pool_masks = nullptr;
    uint64_t* pool_blocks = nullptr;
    uint8_t* pool_payload = nullptr;
    if (rc == 0) {
        const size_t cnt = count ? count : 1;
        if (hipMalloc(&pool_map, (size_t)nb * 4) != hipSuccess || hipMalloc(&pool_masks, cnt * 64) != hipSuccess ||
            hipMalloc(&pool_blocks, cnt * 8) != hipSuccess || hipMalloc(&pool_payload, cnt * 512 * bpv) != hipSuccess) {
            hipFree(pool_map); hipFree(pool_masks); hipFree(pool_blocks); hipFree(pool_payload);
            rc = -1;
        }
    }
    if (rc == 0) {
        hipLaunchKernelGGL(k_brick_scatter, dim3(nb), dim3(512), 0, s, a, masks_all, blocks_all, ids, flags, dist, cube,
                           pool_map, reinterpret_cast<uint64_t*>(pool_masks), pool_blocks, pool_payload);
        if (hipGetLastError() != hipSuccess) rc = -1;
        if (hipStreamSynchronize(s) != hipSuccess) rc = -1;
        if (rc != 0) { hipFree(pool_map); hipFree(pool_masks); hipFree(pool_blocks); hipFree(pool_payload); }
    }
    if (rc == 0) {
        for (uint32_t k = 0; k < nvol; ++k) {
            VolumeStore& o = out[k];
            o.owner = k == 0;
            o.nbricks = first[k + 1] - first[k];
            o.fmt = fmt;
            for (int c = 0; c < 3; ++c) { o.dim[c] = dim[c]; o.bdim[c] = a.bdim[c]; }
            o.brick_map = pool_map + (size_t)k * a.nbv;
            o.masks = pool_masks + (size_t)first[k] * 16;
            o.payload = pool_payload + (size_t)first[k] * 512 * bpv;
            o.blocks = pool_blocks + first[k];
            o.palette = const_cast<uint32_t*>(palette_dev);
            o.bytes = (size_t)a.nbv * 4 + (size_t)o.nbricks * (64 + 8 + 512 * bpv);
        }
    }
    free_temps();
    return rc;
}

namespace {
__global__ void k_debrick(VolumeDesc v, uint8_t* dense) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)v.dim[0] * v.dim[1] * v.dim[2];
    if (i >= n) return;
    uint32_t x = (uint32_t)(i % v.dim[0]);
    uint32_t y = (uint32_t)((i / v.dim[0]) % v.dim[1]);
    uint32_t z = (uint32_t)(i / ((size_t)v.dim[0] * v.dim[1]));
    uint32_t b = (x >> 3) + v.bdim[0] * ((y >> 3) + v.bdim[1] * (z >> 3));
    uint32_t e = v.brick_map[b];
    uint32_t li = (x & 7) + 8 * (y & 7) + 64 * (z & 7);
    if (v.fmt == 0) dense[i] = vt_brick_empty(e) ? 0 : v.payload[(size_t)e * 512 + li];
    else reinterpret_cast<uint32_t*>(dense)[i] = vt_brick_empty(e) ? 0u : reinterpret_cast<const uint32_t*>(v.payload)[(size_t)e * 512 + li];
}
}  // namespace

// Bricked -> dense (voxel order, after the texel remap) for round-trip checks.
int vt_debrick(const VolumeDesc& v, uint8_t* dense_dev, hipStream_t s) {
    size_t n = (size_t)v.dim[0] * v.dim[1] * v.dim[2];
    uint32_t blocks = (uint32_t)((n + 255) / 256);
    hipLaunchKernelGGL(k_debrick, dim3(blocks), dim3(256), 0, s, v, dense_dev);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
