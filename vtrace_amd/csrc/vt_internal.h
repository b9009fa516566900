// vt_internal.h — declarations shared by the translation units of librender_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>

#include "../../include/vtrace.h"
#include "vt_types.h"

// Device storage of one bricked volume.
struct VolumeStore {
    uint32_t dim[3] = {0, 0, 0};
    uint32_t bdim[3] = {0, 0, 0};
    uint32_t fmt = 0;
    uint32_t nbricks = 0;
    uint32_t* brick_map = nullptr;
    uint32_t* masks = nullptr;
    uint8_t* payload = nullptr;
    uint32_t* palette = nullptr;
    uint64_t* blocks = nullptr;
    size_t bytes = 0;
    bool owner = true;          // frees brick_map/masks/payload/palette (volumes bricked together share them)
};

// upload.hip
hipError_t vt_scan_flags(const uint32_t* flags, uint32_t n, uint32_t* out, uint32_t* scratch_blocks, uint32_t* total_dev,
                         hipStream_t s);
int vt_brick_volume(const uint8_t* dense_dev, uint32_t fmt, const uint32_t* palette_dev, const uint16_t* remap_dev[3],
                    const uint32_t dim[3], uint32_t nvol, VolumeStore* out, hipStream_t s);
int vt_debrick(const VolumeDesc& v, uint8_t* dense_dev, hipStream_t s);

// lbvh.hip: builds InstanceGPU records and the LBVH from n ABI records on the device.
struct LbvhWork {
    uint32_t cap = 0;           // instances the buffers hold
    InstanceGPU* inst = nullptr;
    BvhNode* nodes = nullptr;   // 2*cap - 1
    float* cen = nullptr;       // 3*cap centroids
    uint32_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    int32_t* parent = nullptr;  // 2*cap - 1
    uint32_t* counters = nullptr;  // cap
    float* bounds = nullptr;    // 6
    uint32_t* depth = nullptr;  // tree depth (edges on the longest root-leaf path)
    uint32_t* sort_hist = nullptr;  // 256 digit counts per 1024-key chunk (multi-workgroup sort)
};
int vt_lbvh_reserve(LbvhWork& w, uint32_t n);
void vt_lbvh_free(LbvhWork& w);
int vt_lbvh_build(LbvhWork& w, const float* records_dev, uint32_t n, const VolumeDesc* vols_dev, uint32_t nvol,
                  hipStream_t s);
int vt_lbvh_refit(LbvhWork& w, const float* records_dev, uint32_t n, const VolumeDesc* vols_dev, uint32_t nvol,
                  hipStream_t s);

// trace.hip: wavefront path state and queues (one block of device memory, see vt_wave_bind)
struct WaveBuffers {
    uint32_t npaths = 0;        // capacity in pixel samples
    void* L = nullptr;          // float4 per path: sky terms, w = primary-hit flag
    void* Ls = nullptr;         // float4 per path with a primary hit: sun terms, w = shadow flag
    void* br = nullptr;         // float4 per path past its first bounce: throughput
    void* rec[2] = {nullptr, nullptr};      // 48 B queue records per queue slot
    void* sun[2] = {nullptr, nullptr};      // u32 per slot: sun list (queue slots facing the sun)
    uint32_t* counters = nullptr;           // 128 u32: queue sizes and fetch counters
    void* tables = nullptr;                 // ReqGPU[VT_MAX_BATCH] then BatchGPU
    void* gsamp = nullptr;                  // optional, owned by the caller: 2 x float4 per path
                                            // (G-buffer of each sample, averaged per pixel)
};
size_t vt_wave_bytes(uint32_t npaths);
int vt_wave_bind(WaveBuffers& wb, void* mem, uint32_t npaths);
void vt_trace_set_tables(const float* dec_dev, const float* thr_dev);
void vt_trace_set_persistent_blocks(uint32_t blocks);
// Launches one frame.  stats: VT_MAX_STAGES x VT_STAT_COUNT counters (stage k at k*VT_STAT_COUNT).
// ev (optional): 2 events per stage recorded around each launch.  kinds/nstages describe the
// stages launched.
// host_tables (optional, vt_trace_table_bytes()): host staging of the request table instead of
// the library's pinned ring (then nothing is allocated); unchanged until the launch has run.
int vt_trace_launch(const FrameGPU& F, const SceneGPU& S, HitRecord* hits, float4* rad, float4* gbuf,
                    const WaveBuffers& wb,
                    unsigned long long* stats, hipEvent_t* ev, uint8_t* kinds, uint32_t* nstages, hipStream_t s,
                    void* host_tables = nullptr);
size_t vt_trace_table_bytes();
int vt_encode_srgb8(const float4* rad, uint32_t* bgra, uint32_t n, hipStream_t s);
// Batched path tracing (modes 1/2): n requests, their sample chunks traced in one pipeline.
bool vt_batch_layout(const ReqGPU* reqs, uint32_t n, uint32_t npaths, BatchGPU& B);
// acc_after (optional): the accumulation launch waits for this event (a previous sample chunk
// of the same pixels accumulating on another stream).
// blocks (optional): persistent workgroups per launch (0 = the library's default, 4 per CU)
int vt_trace_batch(const ReqGPU* reqs, uint32_t n, const WaveBuffers& wb, unsigned long long* stats, hipEvent_t* ev,
                   uint8_t* kinds, uint32_t* nstages, hipStream_t s, hipEvent_t acc_after = nullptr,
                   uint32_t blocks = 0, void* host_tables = nullptr);
// with_acc = false: the graph stops before the accumulation (vt_trace_accumulate launches it)
int vt_trace_capture(uint32_t max_bounces, bool deep, const WaveBuffers& wb, hipStream_t s, hipGraphExec_t* out,
                     bool with_acc = true, uint32_t blocks = 0);
int vt_trace_accumulate(const WaveBuffers& wb, hipStream_t s, uint32_t blocks = 0);
int vt_trace_upload_tables(const ReqGPU* reqs, uint32_t n, const WaveBuffers& wb, hipStream_t s, BatchGPU* layout);
void vt_trace_release();

// host math shared by the runtime (frame setup); FP64 inverse rounded to FP32.
bool vt_inverse4(const float* m, float* out);

// abi.hip: runtime state shared with the serving loop
std::mutex& vt_rt_lock();
bool vt_rt_inited();
int vt_rt_device();
uint32_t vt_rt_cus();   // compute units of the device (0 if unknown)
hipStream_t vt_rt_stream();
// the plain path's stream of sample chunk k (0 = the library stream; nullptr if k is too large)
hipStream_t vt_rt_chunk_stream(uint32_t k);
const VolumeDesc* vt_rt_vol_table();
uint32_t vt_rt_nvol();
void vt_rt_set_error(const char* what);
// last plain-path frame's radiance, G-buffer (device), size and spp; false if there is none
// (caller holds the lock)
bool vt_rt_frame(const float4** rad, const float4** gbuf, uint32_t* width, uint32_t* height, uint32_t* spp);
// frame constants from the projection / camera matrices and render parameters
bool vt_build_frame(const float* P, const float* C, const vt_render_params& p, FrameGPU& F);
// serve.hip: frees the serving loop (caller holds the runtime lock)
void vt_serve_release();
// denoise.hip: frees the denoiser (caller holds the runtime lock)
void vt_dn_release();
