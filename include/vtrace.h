/* vtrace.h — C ABI of librender_hip, the MI355X (gfx950) drop-in for vtrace's C library.
 *
 * The first seven entry points are the exports the reference's Rust host binds over FFI
 * (src/render.rs:110-128, declared in lib/common.h:253-357) with the same signatures,
 * argument meaning and return conventions.  The vt_* functions are additive headless
 * extensions (resolution, render mode, index+palette volumes, read-back, statistics).
 *
 * Threading: like the reference, the library keeps global state and is meant to be
 * called by one thread at a time; calls may come from different OS threads (the
 * reference's host spawns a render thread per frame, src/main.rs:44-54).  Every export
 * selects the library's device itself and serialises on an internal lock.
 */
#ifndef VTRACE_H
#define VTRACE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* lib/common.h:89-92 — pointers to 16 floats each (column-major mat4), read during the call */
typedef struct render_tick_info {
    void* perspective;
    void* camera;
} render_tick_info;

/* lib/common.h:145-151 (== src/render.rs:37-51) */
typedef struct user_input {
    uint8_t keys[6];
    double mouse_x;
    double mouse_y;
    double last_mouse_x;
    double last_mouse_y;
} user_input;

/* Replaces lib/entry.c:54-57 (entry -> init, lib/entry.c:59-98).  0 on success, else
 * ((uint64_t)hipError << 32) | custom_code, the reference's (VkResult << 32) | custom. */
uint64_t entry(void);

/* Replaces lib/entry.c:150-248.  Renders one frame with info->perspective and
 * info->camera (16 floats each, read during the call only).  Writes the frame size to
 * *window_width / *window_height.  0 = OK, -1 = error (the reference also returns -1
 * when its window closes; a headless library has no window). */
int32_t render_tick(int32_t* window_width, int32_t* window_height, const render_tick_info* render_tick_info);

/* Replaces lib/entry.c:146-148.  Library-owned, valid until cleanup(); headless, so the
 * fields stay zero unless the caller writes them. */
user_input* get_input_data_pointer(void);

/* Replaces lib/entry.c:100-144.  Waits for the device and frees everything. */
void cleanup(void);

/* Replaces lib/memory.c:286-385.  Uploads a width x height x depth RGBA8 volume (x fastest,
 * alpha > 0 = occupied) into 8^3 bricks.  The data is copied during the call.  Returns the
 * texture id (ids are consecutive from 0) or -1 (MAX_TEXTURES = 65536 reached, zero size,
 * allocation failure). */
int32_t add_texture(const uint8_t* data, uint32_t width, uint32_t height, uint32_t depth);

/* Replaces lib/memory.c:235-248.  Returns a host-writable buffer for instance_count records
 * of 16 floats (column-major model matrix, texture id bit-cast into element 15), NULL on
 * failure.  instance_count == 0 is treated as 1, like the reference. */
float* start_update_instances(uint32_t instance_count);

/* Replaces lib/memory.c:250-267.  Commits instance_count records (0 -> 1) and rebuilds the
 * LBVH on the device; when the records equal those of the last build and no texture was
 * added since, the resident tables and tree are kept (no copy, no rebuild).  0 or -1. */
int32_t end_update_instances(uint32_t instance_count);

/* ---------------- headless extensions (no reference counterpart) ---------------------- */

typedef struct vt_render_params {
    uint32_t width, height;     /* frame size (default 1000 x 1000, lib/entry.c:64-65) */
    uint32_t viewport_square;   /* 1: the reference's W x W viewport (lib/command.c:77-84) */
    uint32_t mode;              /* 0 primary colour, 1 primary + sun shadow ray, 2 path tracing */
    uint32_t spp;
    uint32_t max_bounces;
    uint32_t jitter;            /* 0: pixel centres, 1: hash-RNG jitter inside the pixel */
    uint32_t seed;
    uint32_t frame;
    float sun_dir[3];           /* unit vector towards the sun */
    float sun_radiance[3];
    float sky[3];               /* sky radiance; default = the reference clear colour */
    uint32_t stats;             /* 1: count traversal work (nodes, bricks, ...) */
} vt_render_params;

int32_t vt_set_params(const vt_render_params* params);
int32_t vt_get_params(vt_render_params* params);

/* Batched add_texture for streamed chunks (src/render.rs:246-268 uploads one texture per
 * render_tick; src/gen/pager.rs:52-78 queues a 16^3 chunk per newly paged position): count
 * RGBA8 volumes of one size width x height x depth, packed one after another (each x fastest,
 * width*height*depth*4 bytes), copied and bricked together (one copy, one set of launches,
 * shared pools).  Each volume is laid out exactly as add_texture lays it out alone.
 * ids_out[k] (may be NULL) = texture id of volume k; ids are consecutive.  0, or -1 with
 * nothing added. */
int32_t vt_add_textures(const uint8_t* rgba8, uint32_t count, uint32_t width, uint32_t height, uint32_t depth,
                        int32_t* ids_out);
/* Headless cursor: the position render_tick reads into user_input.mouse_x/y (after moving the
 * previous one to last_mouse_x/y, lib/entry.c:157-163); keys are written into the struct from
 * get_input_data_pointer directly, as the reference's key callback does (lib/entry.c:31-52). */
void vt_set_cursor(double x, double y);
/* Path-traced frames run as min(chunks, rows of 8 x 8 tiles) chunks of tile rows on as many
 * streams (chunks <= 4; 0 or 1: one stream).  Until this is called, frames of 2^21 or more
 * pixel samples run as 3 chunks and smaller ones on one stream.  Results are identical: a
 * chunk traces every sample of its pixels and sums them in sample order (radiance, G-buffer). */
int32_t vt_set_frame_split(uint32_t chunks);
/* Workgroups per persistent launch of a one-stream frame (0: 4 per CU, the default).  Results
 * are identical for any value: work is fetched from counters and queued per path, and an append
 * that does not fit a queue segment goes on to the next one; fewer workgroups put the appends
 * in fewer segments (1: every append starts at segment 0, which overflows).  No counterpart in
 * the reference; for tests and launch-size sweeps.  Returns 0, or -1 before entry(). */
int32_t vt_set_persistent_blocks(uint32_t blocks);
/* Like end_update_instances for a list of the same length as the last committed one (moved
 * instances): the records are committed and the LBVH keeps its topology, its boxes refitted
 * bottom-up (no Morton sort or rebuild).  Hits are those of a rebuilt tree; the traversal
 * order, and so the node counters, can differ.  0, or -1 if the count differs. */
int32_t vt_refit_instances(uint32_t instance_count);
/* Number of LBVH builds so far (end_update_instances calls that rebuilt the tree). */
uint64_t vt_lbvh_builds(void);

/* u8 palette-index volume + 256 x RGBA8 palette (alpha > 0 = occupied).  Same id space as
 * add_texture. */
int32_t vt_add_volume_indexed(const uint8_t* indices, uint32_t width, uint32_t height, uint32_t depth,
                              const uint32_t* palette);

/* Allocation-free frame: no counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Renders one frame of the resident scene with the current parameters
 * (vt_set_params) and the matrices of `info`, like render_tick, into the caller's DEVICE
 * buffers: hits_dev (width*height*16 B, the vt_read_hits layout) and radiance_dev
 * (width*height*16 B, RGBA32F).  The frame's wavefront state, queues, counters and request
 * table live in a 128 MiB device array of the library (one per device, held until the
 * process ends), so the call makes no hipMalloc / hipFree / pinned-host allocation.  Frames
 * whose state exceeds it return -1 (path-traced frames need about 152 B per pixel sample,
 * padded to whole 64-pixel tiles: 640x480 at 1 spp fits, 1920x1080 at 1 spp does not).
 * Results equal render_tick's bit for bit; no G-buffer, no counters; one stream.  The call
 * waits for its frame.  Every call on a device shares that device's array: calls that overlap
 * on one device (other threads bypassing the library lock, other libraries' copies) race.
 * 0, or -1 (vt_last_error says why). */
int32_t vt_render_into(const render_tick_info* info, void* hits_dev, uint64_t hits_bytes, float* radiance_dev,
                       uint64_t radiance_bytes);

/* Read-back of the last frame (synchronises).  bytes must be >= width*height*16 (hits and
 * radiance) or width*height*4 (B8G8R8A8 sRGB, the reference's swapchain format).
 * Hit record (16 B per pixel, sample 0): u32 instance (0xFFFFFFFF: sky); u32 bits 0-9 of the
 * hit voxel's x | y << 10 | z << 20; u16 DDA steps, saturated at 65535; u8 face (0..5 =
 * -X,+X,-Y,+Y,-Z,+Z outward normal, 0xFF miss); u8 palette index; u8 last step's axis mask;
 * u8 flags (bit 0 hit, bit 1 shadowed, bit 2 a voxel coordinate >= 32768 of which only the
 * low 15 bits are held, bit 3 steps saturated); u16 bits 10-14 of x | y << 5 | z << 10. */
int32_t vt_read_hits(void* dst, uint64_t bytes);
int32_t vt_read_radiance(float* dst, uint64_t bytes);
int32_t vt_read_bgra8(uint32_t* dst, uint64_t bytes);
/* Optional G-buffer of frames in modes 1 and 2 (denoiser features), the mean over each
 * pixel's samples of: (albedo rgb linear, distance to the camera) and (world unit normal,
 * 1) for a hit, (sky rgb, 0) and (0, 0, 0, 0) for a miss -- 2 x float4 per pixel.  Enabling
 * it adds 32 B per pixel sample of device memory and traffic.  bytes >= width*height*32. */
int32_t vt_enable_gbuffer(uint32_t on);
int32_t vt_read_gbuffer(float* dst, uint64_t bytes);
/* 7 counters of the last frame rendered with stats = 1: rays, LBVH nodes, instances,
 * bricks, non-empty bricks, hits, DDA steps. */
int32_t vt_read_stats(uint64_t* dst7);
/* Per-launch-stage counters of the last counted frame: dst holds max_stages x 7 u64, kinds
 * receives the stage kinds (0 render, 1 primary, 4 accumulate, 5 secondary: the sun and bounce
 * rays of one depth).  Returns the number of stages. */
int32_t vt_read_stage_stats(uint64_t* dst, uint32_t max_stages, uint8_t* kinds);
/* Frames on the library stream alone (one sample chunk), with the mean GPU time of every
 * launch stage (HIP events around each launch).  Returns the number of stages or -1. */
int32_t vt_bench_stages(uint32_t warmup, uint32_t iters, float* kernel_ms, float* wall_ms, float* stage_ms,
                        uint32_t max_stages, uint8_t* kinds);
/* GPU time of the last render_tick's trace kernel (HIP events on the library stream). */
float vt_last_frame_ms(void);
/* Renders `iters` frames after `warmup` with the last camera, as render_tick renders them
 * (vt_set_frame_split; each frame finished before the next).  kernel_ms = mean GPU time per
 * frame (HIP events), wall_ms = mean host time per frame including the wait. */
int32_t vt_bench_frames(uint32_t warmup, uint32_t iters, float* kernel_ms, float* wall_ms);

/* Dense read-back of a bricked volume in voxel order (u8 indices or RGBA8), for round
 * trips; info8 = {w, h, d, fmt, non-empty bricks, bytes lo, bytes hi, 0}. */
int32_t vt_read_volume(int32_t id, void* dst, uint64_t bytes);
int32_t vt_volume_info(int32_t id, uint32_t* info8);
/* The volume's brick map (one u32 per 8^3 brick, x-fastest brick order): a non-empty brick's
 * id, or 0xFFFFFF00 | r for an empty brick, r = radius of the cube of empty bricks around it
 * (L-inf brick distance to the nearest non-empty brick - 1, at most 31). */
int32_t vt_read_brick_map(int32_t id, uint32_t* dst, uint64_t bytes);
/* LBVH of the current instances: (2n-1) x 32 B nodes and n x 128 B instance records. */
int32_t vt_read_lbvh(void* nodes, uint64_t node_bytes, void* instances, uint64_t inst_bytes);
int32_t vt_instance_count(void);
int32_t vt_synchronize(void);
const char* vt_last_error(void);

/* ---------------- serving loop (continuous batching over resident scenes) ---------------- */

typedef struct vt_serve_config {
    uint32_t max_slots;         /* requests admitted at once (<= 64) */
    uint32_t spp_chunk;         /* samples of one request traced per iteration */
    uint64_t memory_budget;     /* bytes for admitted requests' buffers (48 B per pixel) */
    uint32_t max_work;          /* pixel samples traced per iteration (all requests) */
    uint32_t tile_w;            /* pixel tile of 64 samples: tile_w x (64 / tile_w); 0 = 8 */
    uint32_t max_bounces;       /* bounce stages per iteration (largest request max_bounces) */
    uint32_t use_graph;         /* 1: replay the iteration's launch sequence from a hipGraph */
    uint32_t policy;            /* VT_SERVE_ROUND_ROBIN, VT_SERVE_HEAD_RIDERS, VT_SERVE_SHORTEST_FIRST */
    uint32_t pipeline;          /* iterations in flight: 1, or 2 (queue the next before waiting) */
    uint32_t streams;           /* an iteration of 2^18+ pixel samples is spread over up to this many
                                   streams (1..4, 0 = 1): sub-batches of about equal work, a request's
                                   samples in consecutive pieces accumulated in order */
    uint32_t poll;              /* 1: vt_serve_step never waits for the GPU: it retires the iterations
                                   that have finished and launches when one may start, so requests
                                   submitted between calls join the next iteration */
} vt_serve_config;

/* round robin: every admitted request traces its next spp_chunk samples per iteration;
 * head + riders: the oldest request takes the work budget, requests whose whole remainder
 * fits a quarter of it ride along and finish in the same iteration; shortest first: the
 * budget goes to the requests with the least remaining work. */
#define VT_SERVE_ROUND_ROBIN 0u
#define VT_SERVE_HEAD_RIDERS 1u
#define VT_SERVE_SHORTEST_FIRST 2u

typedef struct vt_request {
    uint32_t scene;             /* id from vt_serve_add_scene */
    uint32_t width, height, spp, max_bounces, seed, frame;
    uint32_t viewport_square;
    float perspective[16];      /* column-major, as render_tick_info */
    float camera[16];
    float sun_dir[3];
    float sun_radiance[3];
    float sky[3];
} vt_request;

typedef struct vt_serve_stats {
    uint32_t queued, active, done_unread, tile_w;
    uint64_t memory_in_use;
    uint64_t iterations;
    uint64_t samples;           /* pixel samples traced so far */
    float last_iteration_ms;
    uint32_t sub_batches;       /* sub-batch launch sequences so far (>= iterations) */
} vt_serve_stats;

/* Sets up the serving loop (replaces any previous one). */
int32_t vt_serve_init(const vt_serve_config* config);
/* A scene = n instance records (render.rs layout) over volumes already uploaded with
 * add_texture / vt_add_volume_indexed; its LBVH stays resident.  Returns the scene id. */
int32_t vt_serve_add_scene(const float* records, uint32_t n);
/* Queues a request; returns its id (> 0) or -1. */
int64_t vt_serve_submit(const vt_request* request);
/* One iteration: admit queued requests (slots, memory budget), launch the policy's sample
 * chunks of the admitted requests in one batched launch, and retire the requests whose last
 * chunk has finished (with pipeline 2 that is the previous iteration's).  Writes up to
 * max_completed finished ids; returns the number finished or -1. */
int32_t vt_serve_step(int64_t* completed, uint32_t max_completed);
/* Copies a finished request's RGBA32F radiance and/or hit records (16 B per pixel each)
 * and releases its buffers. */
int32_t vt_serve_read(int64_t id, float* rgba, uint64_t bytes, void* hits, uint64_t hit_bytes);
int32_t vt_serve_status(vt_serve_stats* status);
void vt_serve_shutdown(void);

/* ---------------- learned denoiser (5-level conv U-Net, BF16 MFMA, FP32 accumulate) -------- */

typedef struct vt_dn_config {
    uint32_t enc[5];            /* encoder widths per level (level 4 = bottleneck) */
    uint32_t dec[4];            /* decoder widths per level; all multiples of 16, <= 96 */
} vt_dn_config;

/* Builds the network (parameters zero until vt_dn_set_params). */
int32_t vt_dn_init(const vt_dn_config* config);
/* Layer table: 4 u32 per layer {cin, cout, weight offset, bias offset} into the FP32
 * parameter vector (weights [cout][3][3][cin] row-major).  Returns the layer count (19). */
int32_t vt_dn_layers(uint32_t* out4, uint32_t max_layers);
uint64_t vt_dn_param_count(void);
/* Sets / reads the FP32 parameters (set also resets the Adam state). */
int32_t vt_dn_set_params(const float* params, uint64_t n);
int32_t vt_dn_get_params(float* params, uint64_t n);
/* One forward + backward pass without an update: x = n x h x w x 16 FP32 features
 * (rounded to BF16), target = n x h x w x 3 radiance; writes the network output
 * (n x h x w x 3 irradiance), the FP32 gradient of the mean relMSE loss of the remodulated
 * output for every parameter, and the loss.  h, w % 16 == 0. */
int32_t vt_dn_grad(const float* x, const float* target, uint32_t n, uint32_t h, uint32_t w, float* y, float* grad,
                   float* loss);
/* The 16 input features of a frame rendered with spp samples per pixel (radiance RGBA32F +
 * vt_read_gbuffer layout), as FP32 values of the BF16 tensor the network reads:
 * log(1 + radiance / (albedo + 0.01)) x3, albedo x3, normal x3, log(1 + distance) / 8,
 * log2(spp) / 8, zeros.  The network predicts irradiance; output = (albedo + 0.01) x that. */
int32_t vt_dn_features(const float* rad_rgba, const float* gbuf8, uint32_t h, uint32_t w, uint32_t spp, float* out16);
/* Training frames kept in device memory: input radiance (spp samples) + G-buffer, target. */
int32_t vt_dn_add_frame(const float* rad_rgba, const float* gbuf8, const float* target_rgba, uint32_t h, uint32_t w,
                        uint32_t spp);
int32_t vt_dn_clear_frames(void);
/* iters Adam steps on random crop x crop patches (batch per step, seeded); losses[iters]. */
int32_t vt_dn_train(uint32_t iters, uint32_t batch, uint32_t crop, float lr, uint32_t seed, float* losses);
/* Denoises an h x w frame of spp samples (host buffers) into RGBA32F; gpu_ms = device time. */
int32_t vt_dn_denoise(const float* rad_rgba, const float* gbuf8, uint32_t h, uint32_t w, uint32_t spp,
                      float* out_rgba, float* gpu_ms);
/* Denoises the last frame of render_tick (mode 1/2 with the G-buffer enabled) in place on
 * the device and copies the result out. */
int32_t vt_dn_denoise_frame(float* out_rgba, float* gpu_ms);
/* Output blend: denoise calls return alpha * network output + (1 - alpha) * input radiance,
 * alpha = alpha[k] for spp = 2^k (linear in log2 spp between entries, clamped to the ends);
 * n = 0 removes the table (alpha 1).  Weights in [0, 1], n <= 32. */
int32_t vt_dn_set_blend(const float* alpha, uint32_t n);
void vt_dn_shutdown(void);

#ifdef __cplusplus
}
#endif

#endif /* VTRACE_H */
