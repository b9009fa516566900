// abi.hip — the C-ABI boundary of librender_hip (include/vtrace.h) and its host runtime.
//
// Mirrors the reference's exported interface (src/render.rs:110-128, lib/entry.c,
// lib/memory.c) with a HIP runtime underneath: one device, one library stream (default
// priority), device-resident bricked volumes, a device instance buffer and an LBVH
// rebuilt on the stream whenever instances are committed.  The Vulkan objects of the
// reference (swapchain, descriptor sets, render pass, fences) have no counterpart here;
// stream order replaces the secondary command queue and its ordering passes
// (lib/command.c:111-251).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vtrace.h"
#include "vt_internal.h"

// Scratch of vt_render_into (the allocation-free frame entry point): the wavefront state,
// queues, counters and request table of one frame, in a file-scope device array that the
// code object brings with it (each device holds its own copy until the process ends).  Host
// code never passes `scratch` itself to a kernel (that name is a host-side stand-in): entry()
// resolves its device address with hipGetSymbolAddress and every use goes through that.
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr uint32_t kMaxTextures = 65536;  // lib/common.h:35
constexpr uint32_t kMaxChunks = 4;         // chunks (streams) per frame
constexpr uint64_t kSplitMinSamples = 1ull << 21;   // smaller frames: one stream by default

// host staging of vt_render_into's request table (instead of the pinned ring the other
// entry points allocate on first use); the call waits for its frame before returning
alignas(64) unsigned char noalloc_tables[sizeof(ReqGPU) * VT_MAX_BATCH + sizeof(BatchGPU)];

struct State {
    bool inited = false;
    int device = 0;
    hipStream_t stream = nullptr;
    user_input input{};
    double cursor[2] = {0.0, 0.0};    // headless cursor position (vt_set_cursor)
    // instances (lib/memory.c:221-267)
    float* inst_host = nullptr;    // pinned staging
    float* inst_dev = nullptr;
    uint32_t inst_capacity = 0;
    uint32_t inst_count = 0;
    uint32_t bvh_depth = 0;
    // the depth of the last build is copied to pinned memory behind the build; the host reads
    // it when the next frame picks its traversal-stack variant (no wait in the update itself)
    uint32_t* depth_host = nullptr;
    hipEvent_t depth_ready = nullptr;
    bool depth_pending = false;
    hipEvent_t inst_copied = nullptr;
    std::vector<float> inst_prev;     // records of the last build
    size_t inst_prev_vols = 0;        // textures known at the last build (validity of the records)
    uint64_t lbvh_builds = 0;
    LbvhWork bvh;
    // volumes (lib/memory.c:286-385)
    std::vector<VolumeStore> vols;
    VolumeDesc* vol_table = nullptr;  // kMaxTextures descriptors
    uint8_t* staging = nullptr;
    size_t staging_size = 0;
    uint32_t* palette_rgba_identity = nullptr;
    uint16_t* remap = nullptr;        // 3 tables
    size_t remap_cap = 0;
    // frame
    vt_render_params params{};
    HitRecord* hits = nullptr;
    float4* rad = nullptr;
    uint32_t* bgra = nullptr;
    float4* gbuf = nullptr;     // optional G-buffer, 2 x float4 per pixel
    float4* gsamp = nullptr;    // its per-sample state, 2 x float4 per path
    size_t gsamp_cap = 0;
    bool gbuf_on = false;
    size_t frame_cap = 0;
    void* wave_mem = nullptr;   // wavefront path state + queues
    size_t wave_cap = 0;        // pixel samples it holds
    WaveBuffers wave;
    // frames in chunks of tile rows on several streams: chunk k > 0 on xstream[k] with its own
    // wavefront block; ev_chunk[k] marks chunk k accumulated
    // chunks per path-traced frame (1 = one stream).  Three: on the 1080p 4-spp frame, 3 chunks
    // of tile rows at 2 workgroups per CU took 2.58 ms, 2 at 3 per CU 2.68, 4 at 2 per CU 2.69,
    // 3 at 3 per CU 2.70, 2 at 4 per CU 2.72 (chunks of samples, the best split before: 2.79).
    // Until vt_set_frame_split sets a count, frames under kSplitMinSamples pixel samples run on
    // one stream (640x480 at 1 spp: 0.22 ms on one stream, 0.59 ms in three chunks).
    uint32_t split = 3;
    bool split_auto = true;
    uint32_t cus = 0;             // compute units (persistent grids are sized per CU)
    hipStream_t xstream[kMaxChunks] = {};
    hipEvent_t ev_chunk[kMaxChunks] = {};
    void* xwave_mem[kMaxChunks] = {};
    size_t xwave_cap[kMaxChunks] = {};
    WaveBuffers xwave[kMaxChunks];
    float4* xgsamp[kMaxChunks] = {};  // with the G-buffer: per-sample state of chunks 1..n-1
    size_t xgsamp_cap[kMaxChunks] = {};
    unsigned long long* stats = nullptr;  // VT_MAX_STAGES x VT_STAT_COUNT
    uint8_t stage_kinds[VT_MAX_STAGES] = {};
    uint32_t nstages = 0;
    float* srgb_dec = nullptr;
    float* srgb_thr = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool have_frame = false;
    void* scratch_dev = nullptr;   // device address of `scratch` (hipGetSymbolAddress at entry)
    FrameGPU last_frame{};
    float last_ms = 0.0f;
    uint64_t frames = 0;
    std::string err;
};

State g;
std::mutex g_lock;

void set_err(const char* what, hipError_t e = hipSuccess) {
    g.err = what;
    if (e != hipSuccess) {
        g.err += ": ";
        g.err += hipGetErrorString(e);
    }
}

// ---- host math (frame constants) -------------------------------------------------
// FP32 products in GLSL order (m[c*4 + r]); compiled with -ffp-contract=off.
void mm4(const float* a, const float* b, float* o) {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            o[c * 4 + r] = ((a[0 * 4 + r] * b[c * 4 + 0] + a[1 * 4 + r] * b[c * 4 + 1]) + a[2 * 4 + r] * b[c * 4 + 2]) +
                           a[3 * 4 + r] * b[c * 4 + 3];
}
void mv4(const float* m, const float v[4], float o[4]) {
    for (int r = 0; r < 4; ++r)
        o[r] = ((m[0 * 4 + r] * v[0] + m[1 * 4 + r] * v[1]) + m[2 * 4 + r] * v[2]) + m[3 * 4 + r] * v[3];
}

double srgb_lin(double c) { return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4); }

void default_params(vt_render_params& p) {
    std::memset(&p, 0, sizeof(p));
    p.width = 1000;  // lib/entry.c:64-65
    p.height = 1000;
    p.viewport_square = 1;
    p.mode = 0;
    p.spp = 1;
    p.max_bounces = 0;
    p.jitter = 0;
    p.sun_dir[0] = 0.0f; p.sun_dir[1] = -1.0f; p.sun_dir[2] = 0.0f;
    p.sun_radiance[0] = p.sun_radiance[1] = p.sun_radiance[2] = 2.5f;
    p.sky[0] = 0.53f; p.sky[1] = 0.81f; p.sky[2] = 0.92f;  // lib/command.c:57-60
}

// Path-state capacity of a frame: pixel samples padded to whole 64-pixel tiles of any shape
// (tw x th = 64 pads at most 64 * (W + H) + 64 pixels).
size_t padded_paths(uint32_t w, uint32_t h, uint32_t spp) {
    return ((size_t)w * h + 64 * ((size_t)w + h) + 64) * (spp ? spp : 1u);
}

int ensure_wave_buffers() {
    const size_t paths = padded_paths(g.params.width, g.params.height, g.params.spp);
    if (g.params.mode == 0) return 0;
    if (paths <= g.wave_cap && g.wave_mem) return 0;
    if (paths > 0xFFFFFFFFull / 2) {
        set_err("frame too large (width * height * spp)");
        return -1;
    }
    hipFree(g.wave_mem);
    g.wave_mem = nullptr;
    g.wave_cap = 0;
    hipError_t e = hipMalloc(&g.wave_mem, vt_wave_bytes((uint32_t)paths));
    if (e != hipSuccess) {
        set_err("wavefront buffer allocation", e);
        return -1;
    }
    vt_wave_bind(g.wave, g.wave_mem, (uint32_t)paths);
    g.wave_cap = paths;
    g.wave.gsamp = g.gbuf_on ? g.gsamp : nullptr;
    return 0;
}

// wavefront blocks of chunks 1..n-1 of a split frame (chunk k: tile rows [k T / n, (k + 1) T / n)
// of the T rows of 8 x 8 tiles, every sample)
int ensure_split_buffers(uint32_t spp, uint32_t n) {
    const uint32_t tiles_y = (g.params.height + 7) / 8;
    for (uint32_t k = 1; k < n; ++k) {
        const uint32_t rows = 8 * ((k + 1) * tiles_y / n - k * tiles_y / n);
        const size_t paths = padded_paths(g.params.width, rows, spp);
        if (paths > g.xwave_cap[k] || !g.xwave_mem[k]) {
            hipFree(g.xwave_mem[k]);
            g.xwave_mem[k] = nullptr;
            g.xwave_cap[k] = 0;
            hipError_t e = hipMalloc(&g.xwave_mem[k], vt_wave_bytes((uint32_t)paths));
            if (e != hipSuccess) {
                set_err("wavefront buffer allocation", e);
                return -1;
            }
            vt_wave_bind(g.xwave[k], g.xwave_mem[k], (uint32_t)paths);
            g.xwave_cap[k] = paths;
        }
        g.xwave[k].gsamp = nullptr;
        if (g.gbuf_on) {
            if (paths > g.xgsamp_cap[k] || !g.xgsamp[k]) {
                hipFree(g.xgsamp[k]);
                g.xgsamp[k] = nullptr;
                g.xgsamp_cap[k] = 0;
                hipError_t e = hipMalloc(&g.xgsamp[k], paths * 2 * sizeof(float4));
                if (e != hipSuccess) {
                    set_err("G-buffer sample state allocation", e);
                    return -1;
                }
                g.xgsamp_cap[k] = paths;
            }
            g.xwave[k].gsamp = g.xgsamp[k];
        }
    }
    return 0;
}

// per-sample G-buffer state for the current wave capacity (only while the G-buffer is on)
int ensure_gsamp() {
    if (!g.gbuf_on || g.params.mode == 0) {
        g.wave.gsamp = nullptr;
        return 0;
    }
    if (g.gsamp_cap < g.wave_cap || !g.gsamp) {
        hipFree(g.gsamp);
        g.gsamp = nullptr;
        g.gsamp_cap = 0;
        hipError_t e = hipMalloc(&g.gsamp, g.wave_cap * 2 * sizeof(float4));
        if (e != hipSuccess) {
            set_err("G-buffer sample allocation", e);
            return -1;
        }
        g.gsamp_cap = g.wave_cap;
    }
    g.wave.gsamp = g.gsamp;
    return 0;
}

int ensure_frame_buffers() {
    if (ensure_wave_buffers() != 0 || ensure_gsamp() != 0) return -1;
    if (!g.wave_mem) {  // mode 0 still needs the counters block
        if (hipMalloc(&g.wave_mem, vt_wave_bytes(1)) != hipSuccess) return -1;
        vt_wave_bind(g.wave, g.wave_mem, 1);
        g.wave_cap = 1;
    }
    size_t n = (size_t)g.params.width * g.params.height;
    if (n <= g.frame_cap && g.hits && (g.gbuf || !g.gbuf_on)) return 0;
    hipFree(g.hits); hipFree(g.rad); hipFree(g.bgra); hipFree(g.gbuf);
    g.hits = nullptr; g.rad = nullptr; g.bgra = nullptr; g.gbuf = nullptr;
    hipError_t e = hipSuccess;
    if ((e = hipMalloc(&g.hits, n * sizeof(HitRecord))) != hipSuccess ||
        (e = hipMalloc(&g.rad, n * sizeof(float4))) != hipSuccess || (e = hipMalloc(&g.bgra, n * 4)) != hipSuccess ||
        (g.gbuf_on && (e = hipMalloc(&g.gbuf, n * 2 * sizeof(float4))) != hipSuccess)) {
        set_err("frame buffer allocation", e);
        g.frame_cap = 0;
        return -1;
    }
    g.frame_cap = n;
    return 0;
}

int grow_instances(uint32_t n) {
    if (n <= g.inst_capacity && g.inst_host) return 0;
    uint32_t cap = 1;
    while (cap < n + 1) cap *= 2;  // round_up_p2(n + 1), lib/memory.c:191
    if (g.inst_copied) hipEventSynchronize(g.inst_copied);
    hipHostFree(g.inst_host);
    hipFree(g.inst_dev);
    g.inst_host = nullptr;
    g.inst_dev = nullptr;
    hipError_t e;
    if ((e = hipHostMalloc(&g.inst_host, (size_t)cap * 64, hipHostMallocDefault)) != hipSuccess ||
        (e = hipMalloc(&g.inst_dev, (size_t)cap * 64)) != hipSuccess) {
        set_err("instance buffer allocation", e);
        g.inst_capacity = 0;
        return -1;
    }
    std::memset(g.inst_host, 0, (size_t)cap * 64);
    g.inst_capacity = cap;
    return 0;
}

// Nearest-texel remap per axis: the texel NEAREST + CLAMP_TO_EDGE sampling of voxel/size
// selects (shaders/trace.frag:77, lib/descriptor.c:100-105), FP32.
uint16_t nearest_texel(int v, uint32_t dim) {
    float size = (float)dim;
    float u = (float)v / size;
    float x = u * size;
    int t = (int)std::floor(x);
    if (t < 0) t = 0;
    if (t > (int)dim - 1) t = (int)dim - 1;
    return (uint16_t)t;
}

// count volumes of one size, packed one after another (bytes each), bricked in one pass;
// ids are consecutive from the returned first id.
int upload_volume(const void* data, size_t bytes, uint32_t fmt, const uint32_t dim[3], const uint32_t* palette,
                  uint32_t count = 1) {
    if (count == 0 || bytes > ((size_t)1 << 40) / count) {
        set_err("invalid volume batch");
        return -1;
    }
    if (g.vols.size() + count > kMaxTextures) {
        std::fprintf(stderr, "ERROR: Tried allocating too many textures\n");  // lib/memory.c:288
        set_err("MAX_TEXTURES reached");
        return -1;
    }
    if (!data || dim[0] == 0 || dim[1] == 0 || dim[2] == 0 || dim[0] > 65535 || dim[1] > 65535 || dim[2] > 65535) {
        set_err("invalid volume");
        return -1;
    }
    hipError_t e;
    const size_t total = bytes * count;
    if (total > g.staging_size) {  // grow to a power of two (lib/memory.c:297-302)
        size_t sz = 1;
        while (sz < total) sz *= 2;
        hipFree(g.staging);
        g.staging = nullptr;
        g.staging_size = 0;
        if ((e = hipMalloc(&g.staging, sz)) != hipSuccess) {
            set_err("staging allocation", e);
            return -1;
        }
        g.staging_size = sz;
    }
    if ((e = hipMemcpyAsync(g.staging, data, total, hipMemcpyHostToDevice, g.stream)) != hipSuccess) {
        set_err("volume copy", e);
        return -1;
    }
    size_t rc = (size_t)dim[0] + dim[1] + dim[2];
    if (rc > g.remap_cap) {
        hipFree(g.remap);
        if ((e = hipMalloc(&g.remap, rc * 2)) != hipSuccess) {
            set_err("remap allocation", e);
            g.remap_cap = 0;
            return -1;
        }
        g.remap_cap = rc;
    }
    std::vector<uint16_t> rm(rc);
    size_t off = 0;
    for (int a = 0; a < 3; ++a)
        for (uint32_t i = 0; i < dim[a]; ++i) rm[off++] = nearest_texel((int)i, dim[a]);
    hipMemcpyAsync(g.remap, rm.data(), rc * 2, hipMemcpyHostToDevice, g.stream);
    VolumeStore vs;
    if (fmt == 0) {
        if ((e = hipMalloc(&vs.palette, 1024)) != hipSuccess) {
            set_err("palette allocation", e);
            return -1;
        }
        hipMemcpyAsync(vs.palette, palette, 1024, hipMemcpyHostToDevice, g.stream);
    }
    const uint16_t* remaps[3] = {g.remap, g.remap + dim[0], g.remap + dim[0] + dim[1]};
    std::vector<VolumeStore> out(count);
    if (vt_brick_volume(g.staging, fmt, vs.palette, remaps, dim, count, out.data(), g.stream) != 0) {
        hipFree(vs.palette);
        set_err("bricking failed");
        return -1;
    }
    std::vector<VolumeDesc> descs(count);
    for (uint32_t k = 0; k < count; ++k) {
        const VolumeStore& v = out[k];
        VolumeDesc& d = descs[k];
        d = VolumeDesc{};
        for (int a = 0; a < 3; ++a) { d.dim[a] = v.dim[a]; d.bdim[a] = v.bdim[a]; }
        d.fmt = v.fmt;
        d.nbricks = v.nbricks;
        d.brick_map = v.brick_map;
        d.masks = v.masks;
        d.payload = v.payload;
        d.palette = v.palette;
        d.blocks = v.blocks;
    }
    const int32_t id = (int32_t)g.vols.size();
    hipMemcpyAsync(&g.vol_table[id], descs.data(), sizeof(VolumeDesc) * count, hipMemcpyHostToDevice, g.stream);
    if ((e = hipStreamSynchronize(g.stream)) != hipSuccess) {
        set_err("upload", e);
        return -1;
    }
    g.vols.insert(g.vols.end(), out.begin(), out.end());
    return id;
}

}  // namespace

bool vt_build_frame(const float* P, const float* C, const vt_render_params& p, FrameGPU& F) {
    float invP[16], invC[16], invCc[16], Cc[16];
    std::memcpy(Cc, C, 64);
    Cc[12] = Cc[13] = Cc[14] = 0.0f;  // centered_camera (shaders/trace.frag:55-58)
    if (!vt_inverse4(P, invP) || !vt_inverse4(C, invC) || !vt_inverse4(Cc, invCc)) return false;
    mm4(P, C, F.PC);
    mm4(invCc, invP, F.AB);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) F.R[c * 3 + r] = invCc[c * 4 + r];
    float z[4] = {0.0f, 0.0f, 0.0f, 1.0f}, o[4];
    mv4(invC, z, o);  // cam_pos (shaders/trace.frag:51)
    F.O[0] = o[0]; F.O[1] = o[1]; F.O[2] = o[2];
    F.ip00 = invP[0];
    F.ip11 = invP[5];
    F.W = (float)p.width;
    F.Hv = p.viewport_square ? (float)p.width : (float)p.height;
    F.width = p.width;
    F.height = p.height;
    F.mode = p.mode;
    F.spp = p.spp ? p.spp : 1;
    F.max_bounces = p.max_bounces;
    F.jitter = p.jitter;
    F.seed = p.seed;
    F.frame = p.frame;
    for (int k = 0; k < 3; ++k) {
        F.sun[k] = p.sun_dir[k];
        F.sun_rad[k] = p.sun_radiance[k];
        F.sky[k] = p.sky[k];
    }
    F.stats = p.stats;
    return true;
}

namespace {

bool build_frame(const float* P, const float* C, FrameGPU& F) { return vt_build_frame(P, C, g.params, F); }

SceneGPU scene_gpu() {
    if (g.depth_pending) {  // the last build's depth (its kernels ran before this frame's)
        hipEventSynchronize(g.depth_ready);
        g.bvh_depth = *g.depth_host;
        g.depth_pending = false;
    }
    SceneGPU S;
    S.inst = g.bvh.inst;
    S.nodes = g.bvh.nodes;
    S.vols = g.vol_table;
    S.ninst = g.inst_count;
    S.nvol = (uint32_t)g.vols.size();
    S.bvh_depth = g.bvh_depth;
    return S;
}

// A path-traced frame runs as n = min(g.split, rows of tiles) chunks of rows of 8 x 8 pixel tiles
// on as many streams (chunk 0 on the library stream, chunk k on xstream[k], each with its own
// wavefront block), so each chunk's launches fill the GPU while another's drain.  A chunk traces
// every sample of its pixels and sums them in sample order, so the frame is the one-stream
// frame bit for bit; chunk k accumulates after chunk k-1 and the library stream waits for the
// last.  (Chunks of samples over every pixel were slower: a wave then holds one or two samples
// of 32-64 pixels instead of all samples of 16, and its rays spread further.)
int render_split(const FrameGPU& F, const SceneGPU& S, uint32_t n) {
    const uint32_t spp = F.spp ? F.spp : 1u;
    const uint32_t tiles_y = (F.height + 7) / 8;
    if (ensure_split_buffers(spp, n) != 0) return -1;
    ReqGPU R;
    std::memset(&R, 0, sizeof(R));
    R.F = F;
    R.S = S;
    R.hits = g.hits;
    R.acc = nullptr;   // every chunk holds all samples of its pixels: no running sums
    R.rad = reinterpret_cast<float4*>(g.rad);
    R.gbuf = g.gbuf_on ? g.gbuf : nullptr;
    R.gacc = nullptr;
    R.tile_w = 8;
    R.s0 = 0;
    R.s1 = spp;
    R.final_chunk = 1;
    // persistent workgroups per launch: 4 per CU fill the GPU alone; two chunks in flight take 3
    // per CU each, 3-4 chunks 2 per CU, so the launches of the chunks share the GPU (measured on
    // the 1080p 4-spp frame)
    const uint32_t blocks = g.cus ? g.cus * (n == 2 ? 3u : 2u) : 0u;
    int rc = 0;
    for (uint32_t k = 0; k < n && rc == 0; ++k) {
        R.ty0 = k * tiles_y / n;
        R.ty1 = (k + 1) * tiles_y / n;
        hipStream_t s = k == 0 ? g.stream : g.xstream[k];
        if (k > 0) hipStreamWaitEvent(s, g.ev0, 0);
        uint8_t kinds[VT_MAX_STAGES];
        uint32_t ns = 0;
        rc = vt_trace_batch(&R, 1, k == 0 ? g.wave : g.xwave[k], g.stats, nullptr, k == 0 ? g.stage_kinds : kinds,
                            k == 0 ? &g.nstages : &ns, s, k == 0 ? nullptr : g.ev_chunk[k - 1], blocks);
        hipEventRecord(g.ev_chunk[k], s);
    }
    hipStreamWaitEvent(g.stream, g.ev_chunk[n - 1], 0);
    return rc;
}

int render_frame(const FrameGPU& F) {
    if (ensure_frame_buffers() != 0) return -1;
    if (F.stats) hipMemsetAsync(g.stats, 0, (size_t)VT_MAX_STAGES * VT_STAT_COUNT * 8, g.stream);
    hipEventRecord(g.ev0, g.stream);
    int rc;
    const uint64_t samples = (uint64_t)F.width * F.height * (F.spp ? F.spp : 1u);
    const uint32_t want = g.split_auto && samples < kSplitMinSamples ? 1u : g.split;
    const uint32_t n = std::min(want, std::min((F.height + 7) / 8, kMaxChunks));
    if (n >= 2 && F.mode >= 1)
        rc = render_split(F, scene_gpu(), n);
    else
        rc = vt_trace_launch(F, scene_gpu(), g.hits, reinterpret_cast<float4*>(g.rad), g.gbuf_on ? g.gbuf : nullptr,
                             g.wave, g.stats, nullptr, g.stage_kinds, &g.nstages, g.stream);
    hipEventRecord(g.ev1, g.stream);
    if (rc != 0) {
        set_err("trace launch failed", hipGetLastError());
        return -1;
    }
    g.last_frame = F;
    g.have_frame = true;
    return 0;
}

}  // namespace

// ---- runtime accessors used by the serving loop (serve.hip) ----------------------------
std::mutex& vt_rt_lock() { return g_lock; }
bool vt_rt_inited() { return g.inited; }
int vt_rt_device() { return g.device; }
uint32_t vt_rt_cus() { return g.cus; }
hipStream_t vt_rt_stream() { return g.stream; }
hipStream_t vt_rt_chunk_stream(uint32_t k) { return k == 0 ? g.stream : k < kMaxChunks ? g.xstream[k] : nullptr; }
const VolumeDesc* vt_rt_vol_table() { return g.vol_table; }
uint32_t vt_rt_nvol() { return (uint32_t)g.vols.size(); }
void vt_rt_set_error(const char* what) { set_err(what); }
bool vt_rt_frame(const float4** rad, const float4** gbuf, uint32_t* width, uint32_t* height, uint32_t* spp) {
    if (!g.inited || !g.have_frame || g.last_frame.mode == 0 || !g.gbuf_on || !g.gbuf) return false;
    *rad = reinterpret_cast<const float4*>(g.rad);
    *gbuf = g.gbuf;
    *width = g.last_frame.width;
    *height = g.last_frame.height;
    *spp = g.last_frame.spp;
    return true;
}

namespace {

struct Guard {
    std::lock_guard<std::mutex> lk;
    Guard() : lk(g_lock) {
        if (g.inited) hipSetDevice(g.device);
    }
};

}  // namespace

// ---- host FP64 inverse (Laplace expansion over 2x2 minors, Eberly 2008) -------------
bool vt_inverse4(const float* mf, float* out) {
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
    double b[16];  // b[r*4 + c]
    b[0] = (A(1, 1) * c5 - A(1, 2) * c4 + A(1, 3) * c3) / det;
    b[1] = (-A(0, 1) * c5 + A(0, 2) * c4 - A(0, 3) * c3) / det;
    b[2] = (A(3, 1) * s5 - A(3, 2) * s4 + A(3, 3) * s3) / det;
    b[3] = (-A(2, 1) * s5 + A(2, 2) * s4 - A(2, 3) * s3) / det;
    b[4] = (-A(1, 0) * c5 + A(1, 2) * c2 - A(1, 3) * c1) / det;
    b[5] = (A(0, 0) * c5 - A(0, 2) * c2 + A(0, 3) * c1) / det;
    b[6] = (-A(3, 0) * s5 + A(3, 2) * s2 - A(3, 3) * s1) / det;
    b[7] = (A(2, 0) * s5 - A(2, 2) * s2 + A(2, 3) * s1) / det;
    b[8] = (A(1, 0) * c4 - A(1, 1) * c2 + A(1, 3) * c0) / det;
    b[9] = (-A(0, 0) * c4 + A(0, 1) * c2 - A(0, 3) * c0) / det;
    b[10] = (A(3, 0) * s4 - A(3, 1) * s2 + A(3, 3) * s0) / det;
    b[11] = (-A(2, 0) * s4 + A(2, 1) * s2 - A(2, 3) * s0) / det;
    b[12] = (-A(1, 0) * c3 + A(1, 1) * c1 - A(1, 2) * c0) / det;
    b[13] = (A(0, 0) * c3 - A(0, 1) * c1 + A(0, 2) * c0) / det;
    b[14] = (-A(3, 0) * s3 + A(3, 1) * s1 - A(3, 2) * s0) / det;
    b[15] = (A(2, 0) * s3 - A(2, 1) * s1 + A(2, 2) * s0) / det;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            float v = (float)b[r * 4 + c];
            if (!std::isfinite(v)) return false;
            out[c * 4 + r] = v;
        }
    return true;
}

extern "C" {

uint64_t entry(void) {
    std::lock_guard<std::mutex> lk(g_lock);
    if (g.inited) return 0;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        set_err("no HIP device", e);
        return ((uint64_t)(uint32_t)(e == hipSuccess ? hipErrorNoDevice : e) << 32) | 1u;
    }
    g.device = 0;
    if ((e = hipSetDevice(g.device)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&g.ev0)) != hipSuccess || (e = hipEventCreate(&g.ev1)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&g.inst_copied, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&g.depth_ready, hipEventDisableTiming)) != hipSuccess ||
        (e = hipHostMalloc(&g.depth_host, sizeof(uint32_t), hipHostMallocDefault)) != hipSuccess ||
        (e = hipMalloc(&g.vol_table, (size_t)kMaxTextures * sizeof(VolumeDesc))) != hipSuccess ||
        (e = hipMalloc(&g.stats, (size_t)VT_MAX_STAGES * VT_STAT_COUNT * 8)) != hipSuccess ||
        (e = hipMalloc(&g.srgb_dec, 256 * 4)) != hipSuccess || (e = hipMalloc(&g.srgb_thr, 255 * 4)) != hipSuccess) {
        set_err("init", e);
        return ((uint64_t)(uint32_t)e << 32) | 1u;
    }
    for (uint32_t k = 0; k < kMaxChunks; ++k) {  // default priority, like the library stream
        if ((k > 0 && (e = hipStreamCreateWithFlags(&g.xstream[k], hipStreamNonBlocking)) != hipSuccess) ||
            (e = hipEventCreateWithFlags(&g.ev_chunk[k], hipEventDisableTiming)) != hipSuccess) {
            set_err("init", e);
            return ((uint64_t)(uint32_t)e << 32) | 1u;
        }
    }
    float dec[256], thr[255];
    for (int i = 0; i < 256; ++i) dec[i] = (float)srgb_lin((double)i / 255.0);
    for (int k = 0; k < 255; ++k) thr[k] = (float)srgb_lin((k + 0.5) / 255.0);
    hipMemcpy(g.srgb_dec, dec, sizeof(dec), hipMemcpyHostToDevice);
    hipMemcpy(g.srgb_thr, thr, sizeof(thr), hipMemcpyHostToDevice);
    hipMemset(g.stats, 0, (size_t)VT_MAX_STAGES * VT_STAT_COUNT * 8);
    vt_trace_set_tables(g.srgb_dec, g.srgb_thr);
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, g.device) == hipSuccess && prop.multiProcessorCount > 0) {
            g.cus = (uint32_t)prop.multiProcessorCount;
            vt_trace_set_persistent_blocks(g.cus * 4u);
        }
    }
    if ((e = hipGetSymbolAddress(&g.scratch_dev, HIP_SYMBOL(scratch))) != hipSuccess) {
        set_err("init: scratch address", e);
        g.scratch_dev = nullptr;
        return ((uint64_t)(uint32_t)e << 32) | 1u;
    }
    default_params(g.params);
    g.inited = true;
    if (grow_instances(0) != 0) return ((uint64_t)hipErrorOutOfMemory << 32) | 1u;
    return 0;
}

int32_t render_tick(int32_t* window_width, int32_t* window_height, const render_tick_info* info) {
    Guard G;
    if (!g.inited || !info || !info->perspective || !info->camera) {
        set_err("render_tick: not initialised or null info");
        return -1;
    }
    // lib/entry.c:157-163: the previous cursor becomes last_mouse, the current cursor is read
    // (here: the position the host set with vt_set_cursor); on the first frame both are equal
    g.input.last_mouse_x = g.input.mouse_x;
    g.input.last_mouse_y = g.input.mouse_y;
    g.input.mouse_x = g.cursor[0];
    g.input.mouse_y = g.cursor[1];
    if (g.frames == 0) {
        g.input.last_mouse_x = g.input.mouse_x;
        g.input.last_mouse_y = g.input.mouse_y;
    }
    FrameGPU F;
    if (!build_frame(static_cast<const float*>(info->perspective), static_cast<const float*>(info->camera), F)) {
        set_err("render_tick: singular projection or camera");
        return -1;
    }
    if (render_frame(F) != 0) return -1;
    hipEventSynchronize(g.ev1);
    hipEventElapsedTime(&g.last_ms, g.ev0, g.ev1);
    ++g.frames;
    if (window_width) *window_width = (int32_t)g.params.width;
    if (window_height) *window_height = (int32_t)g.params.height;
    return 0;
}

int32_t vt_render_into(const render_tick_info* info, void* hits_dev, uint64_t hits_bytes, float* radiance_dev,
                       uint64_t radiance_bytes) {
    Guard G;
    if (!g.inited || !g.scratch_dev || !info || !info->perspective || !info->camera || !hits_dev || !radiance_dev) {
        set_err("vt_render_into: not initialised, null info or null output");
        return -1;
    }
    if (vt_trace_table_bytes() != sizeof(noalloc_tables)) {
        set_err("vt_render_into: request table size mismatch");
        return -1;
    }
    FrameGPU F;
    if (!build_frame(static_cast<const float*>(info->perspective), static_cast<const float*>(info->camera), F)) {
        set_err("vt_render_into: singular projection or camera");
        return -1;
    }
    F.stats = 0;
    const size_t npix = (size_t)F.width * F.height;
    if (hits_bytes < npix * sizeof(HitRecord) || radiance_bytes < npix * sizeof(float4)) {
        set_err("vt_render_into: output buffer too small");
        return -1;
    }
    const size_t paths = F.mode == 0 ? 1 : padded_paths(F.width, F.height, F.spp);
    if (paths > 0xFFFFFFFFull / 2 || vt_wave_bytes((uint32_t)paths) > kNoAllocScratchBytes) {
        set_err("vt_render_into: the frame's wavefront state exceeds the 128 MiB scratch");
        return -1;
    }
    WaveBuffers wb;
    vt_wave_bind(wb, g.scratch_dev, (uint32_t)paths);
    wb.gsamp = nullptr;
    uint8_t kinds[VT_MAX_STAGES];
    uint32_t ns = 0;
    int rc = vt_trace_launch(F, scene_gpu(), static_cast<HitRecord*>(hits_dev), reinterpret_cast<float4*>(radiance_dev),
                             nullptr, wb, nullptr, nullptr, kinds, &ns, g.stream, noalloc_tables);
    // the table staging and the scratch are reused by the next call: the frame completes here
    const hipError_t e = hipStreamSynchronize(g.stream);
    if (rc != 0 || e != hipSuccess) {
        set_err("vt_render_into: trace launch failed", e != hipSuccess ? e : hipGetLastError());
        return -1;
    }
    return 0;
}

user_input* get_input_data_pointer(void) { return &g.input; }

void vt_set_cursor(double x, double y) {
    Guard G;
    g.cursor[0] = x;
    g.cursor[1] = y;
}

void cleanup(void) {
    std::lock_guard<std::mutex> lk(g_lock);
    if (!g.inited) return;
    hipSetDevice(g.device);
    hipStreamSynchronize(g.stream);
    vt_serve_release();
    vt_dn_release();
    for (auto& v : g.vols) {
        if (!v.owner) continue;  // shares the pools of the first volume of its batch
        hipFree(v.brick_map); hipFree(v.masks); hipFree(v.payload); hipFree(v.palette); hipFree(v.blocks);
    }
    g.vols.clear();
    vt_lbvh_free(g.bvh);
    hipHostFree(g.inst_host);
    hipFree(g.inst_dev); hipFree(g.vol_table); hipFree(g.staging); hipFree(g.remap);
    hipFree(g.hits); hipFree(g.rad); hipFree(g.bgra); hipFree(g.gbuf); hipFree(g.gsamp); hipFree(g.stats); hipFree(g.srgb_dec); hipFree(g.srgb_thr);
    hipFree(g.wave_mem);
    for (uint32_t k = 0; k < kMaxChunks; ++k) {
        if (g.xstream[k]) hipStreamSynchronize(g.xstream[k]);
        hipFree(g.xwave_mem[k]);
        if (g.ev_chunk[k]) hipEventDestroy(g.ev_chunk[k]);
        if (g.xstream[k]) hipStreamDestroy(g.xstream[k]);
    }
    for (uint32_t k = 0; k < kMaxChunks; ++k) hipFree(g.xgsamp[k]);
    hipEventDestroy(g.ev0); hipEventDestroy(g.ev1); hipEventDestroy(g.inst_copied); hipEventDestroy(g.depth_ready);
    hipHostFree(g.depth_host);
    hipStreamDestroy(g.stream);
    vt_trace_set_tables(nullptr, nullptr);
    vt_trace_release();
    g = State{};
}

int32_t add_texture(const uint8_t* data, uint32_t width, uint32_t height, uint32_t depth) {
    Guard G;
    if (!g.inited) return -1;
    const uint32_t dim[3] = {width, height, depth};
    return upload_volume(data, (size_t)4 * width * height * depth, 1, dim, nullptr);
}

int32_t vt_add_textures(const uint8_t* rgba8, uint32_t count, uint32_t width, uint32_t height, uint32_t depth,
                        int32_t* ids_out) {
    Guard G;
    if (!g.inited || count == 0) return -1;
    const uint32_t dim[3] = {width, height, depth};
    const int32_t first = upload_volume(rgba8, (size_t)4 * width * height * depth, 1, dim, nullptr, count);
    if (first < 0) return -1;
    if (ids_out)
        for (uint32_t k = 0; k < count; ++k) ids_out[k] = first + (int32_t)k;
    return 0;
}

float* start_update_instances(uint32_t instance_count) {
    Guard G;
    if (!g.inited) return nullptr;
    if (instance_count == 0) instance_count = 1;  // lib/memory.c:236
    if (instance_count > g.inst_capacity && grow_instances(instance_count) != 0) return nullptr;
    hipEventSynchronize(g.inst_copied);  // previous commit has left the staging buffer
    return g.inst_host;
}

int32_t end_update_instances(uint32_t instance_count) {
    Guard G;
    if (!g.inited) return -1;
    if (instance_count == 0) instance_count = 1;  // lib/memory.c:251
    if (instance_count > g.inst_capacity) {
        set_err("end_update_instances: more instances than started");
        return -1;
    }
    // the same records over the same textures as the last build: the device tables and the
    // tree are current (a streamed world re-sends an unchanged list on most updates)
    const size_t nrec = (size_t)instance_count * 16;
    if (g.inst_count == instance_count && g.inst_prev.size() == nrec && g.inst_prev_vols == g.vols.size() &&
        std::memcmp(g.inst_prev.data(), g.inst_host, nrec * 4) == 0)
        return 0;
    g.inst_prev.clear();
    hipError_t e = hipMemcpyAsync(g.inst_dev, g.inst_host, (size_t)instance_count * 64, hipMemcpyHostToDevice, g.stream);
    if (e != hipSuccess) {
        set_err("instance copy", e);
        return -1;
    }
    hipEventRecord(g.inst_copied, g.stream);
    if (vt_lbvh_build(g.bvh, g.inst_dev, instance_count, g.vol_table, (uint32_t)g.vols.size(), g.stream) != 0) {
        set_err("lbvh build", hipGetLastError());
        return -1;
    }
    // the tree depth selects the traversal-stack variant (LDS stack or a deep private stack);
    // it is read by the next frame (scene_gpu), so the update does not wait for the build
    e = hipMemcpyAsync(g.depth_host, g.bvh.depth, 4, hipMemcpyDeviceToHost, g.stream);
    if (e == hipSuccess) e = hipEventRecord(g.depth_ready, g.stream);
    if (e != hipSuccess) {
        set_err("lbvh depth", e);
        return -1;
    }
    g.depth_pending = true;
    g.inst_count = instance_count;
    g.inst_prev.assign(g.inst_host, g.inst_host + nrec);
    g.inst_prev_vols = g.vols.size();
    ++g.lbvh_builds;
    return 0;
}

int32_t vt_refit_instances(uint32_t instance_count) {
    Guard G;
    if (!g.inited) return -1;
    if (instance_count == 0) instance_count = 1;  // as end_update_instances
    if (instance_count != g.inst_count || g.inst_prev.size() != (size_t)instance_count * 16) {
        set_err("vt_refit_instances: the instance count differs from the last build");
        return -1;
    }
    hipError_t e = hipMemcpyAsync(g.inst_dev, g.inst_host, (size_t)instance_count * 64, hipMemcpyHostToDevice, g.stream);
    if (e != hipSuccess) {
        set_err("instance copy", e);
        return -1;
    }
    hipEventRecord(g.inst_copied, g.stream);
    if (vt_lbvh_refit(g.bvh, g.inst_dev, instance_count, g.vol_table, (uint32_t)g.vols.size(), g.stream) != 0) {
        set_err("lbvh refit", hipGetLastError());
        return -1;
    }
    g.inst_prev.assign(g.inst_host, g.inst_host + (size_t)instance_count * 16);
    g.inst_prev_vols = g.vols.size();
    return 0;
}

uint64_t vt_lbvh_builds(void) {
    Guard G;
    return g.lbvh_builds;
}

int32_t vt_set_params(const vt_render_params* p) {
    Guard G;
    if (!g.inited || !p || p->width == 0 || p->height == 0 || p->width > 16384 || p->height > 16384 || p->mode > 2 ||
        p->spp == 0 || p->spp > 65536) {
        set_err("vt_set_params: invalid parameters");
        return -1;
    }
    g.params = *p;
    return ensure_frame_buffers();
}

int32_t vt_get_params(vt_render_params* p) {
    Guard G;
    if (!g.inited || !p) return -1;
    *p = g.params;
    return 0;
}

int32_t vt_add_volume_indexed(const uint8_t* indices, uint32_t width, uint32_t height, uint32_t depth,
                              const uint32_t* palette) {
    Guard G;
    if (!g.inited || !palette) return -1;
    const uint32_t dim[3] = {width, height, depth};
    return upload_volume(indices, (size_t)width * height * depth, 0, dim, palette);
}

static int32_t read_frame(const void* src, void* dst, uint64_t bytes, size_t elem) {
    Guard G;
    if (!g.inited || !g.have_frame || !dst) return -1;
    size_t need = (size_t)g.last_frame.width * g.last_frame.height * elem;
    if (bytes < need) {
        set_err("read-back buffer too small");
        return -1;
    }
    hipError_t e = hipMemcpyAsync(dst, src, need, hipMemcpyDeviceToHost, g.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
    if (e != hipSuccess) {
        set_err("read-back", e);
        return -1;
    }
    return 0;
}

int32_t vt_read_hits(void* dst, uint64_t bytes) { return read_frame(g.hits, dst, bytes, sizeof(HitRecord)); }
int32_t vt_read_radiance(float* dst, uint64_t bytes) { return read_frame(g.rad, dst, bytes, sizeof(float4)); }

int32_t vt_enable_gbuffer(uint32_t on) {
    Guard G;
    if (!g.inited) return -1;
    g.gbuf_on = on != 0;
    return 0;
}

int32_t vt_read_gbuffer(float* dst, uint64_t bytes) {
    {
        Guard G;
        if (!g.gbuf_on || !g.gbuf || !g.have_frame || g.last_frame.mode == 0) {
            set_err("no G-buffer: enable it and render a frame in mode 1 or 2");
            return -1;
        }
    }
    return read_frame(g.gbuf, dst, bytes, 2 * sizeof(float4));
}

int32_t vt_read_bgra8(uint32_t* dst, uint64_t bytes) {
    {
        Guard G;
        if (!g.inited || !g.have_frame) return -1;
        uint32_t n = g.last_frame.width * g.last_frame.height;
        if (vt_encode_srgb8(reinterpret_cast<const float4*>(g.rad), g.bgra, n, g.stream) != 0) return -1;
    }
    return read_frame(g.bgra, dst, bytes, 4);
}

int32_t vt_read_stats(uint64_t* dst7) {
    Guard G;
    if (!g.inited || !dst7) return -1;
    std::vector<uint64_t> all((size_t)VT_MAX_STAGES * VT_STAT_COUNT);
    hipMemcpyAsync(all.data(), g.stats, all.size() * 8, hipMemcpyDeviceToHost, g.stream);
    if (hipStreamSynchronize(g.stream) != hipSuccess) return -1;
    for (int k = 0; k < VT_STAT_COUNT; ++k) {
        uint64_t t = 0;
        for (uint32_t st = 0; st < VT_MAX_STAGES; ++st) t += all[(size_t)st * VT_STAT_COUNT + k];
        dst7[k] = t;
    }
    return 0;
}

int32_t vt_read_stage_stats(uint64_t* dst, uint32_t max_stages, uint8_t* kinds) {
    Guard G;
    if (!g.inited || !dst) return -1;
    const uint32_t n = g.nstages < max_stages ? g.nstages : max_stages;
    hipMemcpyAsync(dst, g.stats, (size_t)n * VT_STAT_COUNT * 8, hipMemcpyDeviceToHost, g.stream);
    if (hipStreamSynchronize(g.stream) != hipSuccess) return -1;
    if (kinds) std::memcpy(kinds, g.stage_kinds, n);
    return (int32_t)n;
}

float vt_last_frame_ms(void) { return g.last_ms; }

// Whole frames as render_tick renders them (two streams for path-traced frames of 2+ samples,
// each frame complete before the next starts): kernel_ms = mean GPU time per frame (HIP events
// around each frame), wall_ms = mean host time per frame including the per-frame wait.
int32_t vt_bench_frames(uint32_t warmup, uint32_t iters, float* kernel_ms, float* wall_ms) {
    Guard G;
    if (!g.inited || !g.have_frame || iters == 0) return -1;
    FrameGPU F = g.last_frame;
    F.stats = 0;
    for (uint32_t i = 0; i < warmup; ++i) {
        if (render_frame(F) != 0) return -1;
        hipEventSynchronize(g.ev1);
    }
    double gpu = 0.0;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 0; i < iters; ++i) {
        if (render_frame(F) != 0) return -1;
        hipEventSynchronize(g.ev1);
        float ms = 0.0f;
        hipEventElapsedTime(&ms, g.ev0, g.ev1);
        gpu += ms;
    }
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (kernel_ms) *kernel_ms = (float)(gpu / iters);
    if (wall_ms) *wall_ms = (float)(wall / iters);
    return 0;
}

int32_t vt_set_frame_split(uint32_t chunks) {
    Guard G;
    if (!g.inited || chunks > kMaxChunks) return -1;
    hipStreamSynchronize(g.stream);
    g.split = chunks ? chunks : 1;
    g.split_auto = false;
    return 0;
}

int32_t vt_set_persistent_blocks(uint32_t blocks) {
    Guard G;
    if (!g.inited) return -1;
    hipStreamSynchronize(g.stream);
    vt_trace_set_persistent_blocks(blocks ? blocks : g.cus * 4u);
    return 0;
}

int32_t vt_bench_stages(uint32_t warmup, uint32_t iters, float* kernel_ms, float* wall_ms, float* stage_ms,
                        uint32_t max_stages, uint8_t* kinds) {
    Guard G;
    if (!g.inited || !g.have_frame || iters == 0) return -1;
    FrameGPU F = g.last_frame;
    F.stats = 0;
    for (uint32_t i = 0; i < warmup; ++i)
        if (render_frame(F) != 0) return -1;
    hipStreamSynchronize(g.stream);
    const size_t per = 2 * (size_t)VT_MAX_STAGES;
    std::vector<hipEvent_t> ev(per * iters);
    for (auto& e : ev) hipEventCreate(&e);
    hipEvent_t start, stop;
    hipEventCreate(&start);
    hipEventCreate(&stop);
    uint8_t kk[VT_MAX_STAGES] = {};
    uint32_t ns = 0;
    hipEventRecord(start, g.stream);
    int rc = 0;
    for (uint32_t i = 0; i < iters && rc == 0; ++i)
        rc = vt_trace_launch(F, scene_gpu(), g.hits, reinterpret_cast<float4*>(g.rad), g.gbuf_on ? g.gbuf : nullptr,
                             g.wave, g.stats,
                             &ev[per * i], kk, &ns, g.stream);
    hipEventRecord(stop, g.stream);
    hipEventSynchronize(stop);
    float total = 0.0f, ksum = 0.0f;
    hipEventElapsedTime(&total, start, stop);
    std::vector<float> sm(VT_MAX_STAGES, 0.0f);
    for (uint32_t i = 0; i < iters && rc == 0; ++i)
        for (uint32_t st = 0; st < ns; ++st) {
            float t = 0.0f;
            hipEventElapsedTime(&t, ev[per * i + 2 * st], ev[per * i + 2 * st + 1]);
            sm[st] += t;
            ksum += t;
        }
    for (auto& e : ev) hipEventDestroy(e);
    hipEventDestroy(start);
    hipEventDestroy(stop);
    if (kernel_ms) *kernel_ms = ksum / (float)iters;
    if (wall_ms) *wall_ms = total / (float)iters;
    for (uint32_t st = 0; st < ns && st < max_stages; ++st) {
        if (stage_ms) stage_ms[st] = sm[st] / (float)iters;
        if (kinds) kinds[st] = kk[st];
    }
    return rc == 0 ? (int32_t)ns : -1;
}

int32_t vt_volume_info(int32_t id, uint32_t* info8) {
    Guard G;
    if (!g.inited || id < 0 || (size_t)id >= g.vols.size() || !info8) return -1;
    const VolumeStore& v = g.vols[id];
    info8[0] = v.dim[0]; info8[1] = v.dim[1]; info8[2] = v.dim[2];
    info8[3] = v.fmt;
    info8[4] = v.nbricks;
    info8[5] = (uint32_t)(v.bytes & 0xFFFFFFFFu);
    info8[6] = (uint32_t)(v.bytes >> 32);
    info8[7] = 0;
    return 0;
}

int32_t vt_read_volume(int32_t id, void* dst, uint64_t bytes) {
    Guard G;
    if (!g.inited || id < 0 || (size_t)id >= g.vols.size() || !dst) return -1;
    const VolumeStore& v = g.vols[id];
    size_t need = (size_t)v.dim[0] * v.dim[1] * v.dim[2] * (v.fmt == 0 ? 1 : 4);
    if (bytes < need) return -1;
    uint8_t* tmp = nullptr;
    if (hipMalloc(&tmp, need) != hipSuccess) return -1;
    VolumeDesc d{};
    for (int a = 0; a < 3; ++a) { d.dim[a] = v.dim[a]; d.bdim[a] = v.bdim[a]; }
    d.fmt = v.fmt;
    d.nbricks = v.nbricks;
    d.brick_map = v.brick_map;
    d.masks = v.masks;
    d.payload = v.payload;
    d.palette = v.palette;
    d.blocks = v.blocks;
    int rc = vt_debrick(d, tmp, g.stream);
    if (rc == 0) hipMemcpyAsync(dst, tmp, need, hipMemcpyDeviceToHost, g.stream);
    if (hipStreamSynchronize(g.stream) != hipSuccess) rc = -1;
    hipFree(tmp);
    return rc;
}

int32_t vt_read_brick_map(int32_t id, uint32_t* dst, uint64_t bytes) {
    Guard G;
    if (!g.inited || id < 0 || (size_t)id >= g.vols.size() || !dst) return -1;
    const VolumeStore& v = g.vols[id];
    const size_t need = (size_t)v.bdim[0] * v.bdim[1] * v.bdim[2] * 4;
    if (bytes < need) return -1;
    hipMemcpyAsync(dst, v.brick_map, need, hipMemcpyDeviceToHost, g.stream);
    return hipStreamSynchronize(g.stream) == hipSuccess ? 0 : -1;
}

int32_t vt_read_lbvh(void* nodes, uint64_t node_bytes, void* instances, uint64_t inst_bytes) {
    Guard G;
    if (!g.inited || g.inst_count == 0) return -1;
    size_t nn = 2 * (size_t)g.inst_count - 1;
    if (nodes) {
        if (node_bytes < nn * sizeof(BvhNode)) return -1;
        hipMemcpyAsync(nodes, g.bvh.nodes, nn * sizeof(BvhNode), hipMemcpyDeviceToHost, g.stream);
    }
    if (instances) {
        if (inst_bytes < g.inst_count * sizeof(InstanceGPU)) return -1;
        hipMemcpyAsync(instances, g.bvh.inst, g.inst_count * sizeof(InstanceGPU), hipMemcpyDeviceToHost, g.stream);
    }
    return hipStreamSynchronize(g.stream) == hipSuccess ? 0 : -1;
}

int32_t vt_instance_count(void) { return (int32_t)g.inst_count; }

int32_t vt_synchronize(void) {
    Guard G;
    if (!g.inited) return -1;
    return hipStreamSynchronize(g.stream) == hipSuccess ? 0 : -1;
}

const char* vt_last_error(void) { return g.err.c_str(); }

}  // extern "C"
