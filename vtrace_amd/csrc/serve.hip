// serve.hip — continuous-batching serving loop over resident scenes (no reference
// counterpart: the reference renders one interactive frame per render_tick, lib/entry.c:150).
//
// Scenes (instance lists over the library's resident volumes, each with its own LBVH) stay
// in HBM.  Requests (camera, resolution, spp, seed) queue FIFO; vt_serve_step() runs one
// iteration:
//   1. admission: the queue head is admitted while a slot is free and its buffers (hit
//      records, radiance running sum, output: 48 B per pixel) fit the memory budget
//      (first-fit arena);
//   2. batching: admitted requests with samples left contribute a chunk of samples (all
//      pixels x some samples) within the iteration's work budget, chosen by the policy:
//      round robin (every request its next spp_chunk samples), head + riders (the oldest
//      request gets the budget, requests whose whole remainder fits a quarter of it ride
//      along and finish), or shortest remaining work first;
//   3. one launch of the batched wavefront pipeline traces all chunks together (optionally
//      replayed from a captured hipGraph), so requests that arrived while others were in
//      flight share the device with them.  With pipeline depth 2 the next iteration is
//      queued before the previous one is waited for, so host work overlaps the GPU;
//   4. requests whose last chunk has finished are complete; their output waits for
//      vt_serve_read, which copies on a separate stream (overlapping the next iteration).
// A request's pixels depend only on its own scene, camera, seed and sample indices and
// its samples are summed in sample order, so its result is bit-identical whether it is
// rendered alone, in any mix, or by the plain single-frame path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "vt_internal.h"

namespace {

struct Arena {
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    void reset(size_t n) {
        size = n;
        free_.clear();
        if (n) free_[0] = n;
    }
    bool alloc(size_t n, size_t* off) {
        n = (n + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n, at = it->first + n;
                free_.erase(it);
                if (rest) free_[at] = rest;
                return true;
            }
        return false;
    }
    void release(size_t off, size_t n) {
        n = (n + 255) & ~(size_t)255;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t in_use() const {
        size_t f = 0;
        for (auto& kv : free_) f += kv.second;
        return size - f;
    }
};

struct ServeScene {
    LbvhWork bvh;
    float* records = nullptr;
    uint32_t n = 0;
    uint32_t depth = 0;
};

enum { kQueued = 0, kActive = 1, kDone = 2 };

struct ServeReq {
    vt_request q{};
    FrameGPU F{};
    size_t off = 0, bytes = 0;
    uint32_t next_s = 0;   // samples launched so far
    int state = kQueued;
};

// one launched iteration not yet waited for
struct Inflight {
    hipEvent_t done = nullptr, t0 = nullptr, t1 = nullptr;
    std::vector<int64_t> finishing;   // requests whose last samples are in it
    uint64_t work = 0;
};

constexpr uint32_t kSubBatches = 4;           // streams an iteration is spread over
constexpr uint64_t kMinSubWork = 1ull << 18;  // pixel samples below which an iteration is not split

struct Serve {
    bool on = false;
    vt_serve_config cfg{};
    char* arena_mem = nullptr;
    Arena arena;
    void* wave_mem = nullptr;
    WaveBuffers wave;
    // sub-batches 1..streams-1: own wavefront block (work budget + riders' quarter), graphs,
    // and the plain path's chunk streams (not new ones: a process gets few hardware queues,
    // and streams beyond them share one, so extra streams would run sub-batches one after
    // another); acc_done[j] marks sub-batch j accumulated
    void* sub_mem[kSubBatches] = {};
    WaveBuffers sub_wave[kSubBatches];
    hipStream_t sub[kSubBatches] = {};
    hipGraphExec_t sub_graph[kSubBatches][2][kSubBatches] = {};  // [sub-batch][deep][sub-batches - 1]
    hipEvent_t acc_done[kSubBatches] = {}, start = nullptr;
    std::vector<ServeScene> scenes;
    std::deque<int64_t> queue;
    std::vector<int64_t> active;
    std::map<int64_t, ServeReq> reqs;
    int64_t next_id = 1;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    hipStream_t copy = nullptr;
    Inflight fl[2];
    uint32_t fl_head = 0, fl_count = 0;
    uint64_t iterations = 0, samples = 0;
    uint32_t sub_batches = 0;
    float last_ms = 0.0f;
};

Serve sv;

size_t req_bytes(const vt_request& q) { return (size_t)q.width * q.height * 48; }

uint32_t tiles_of(uint32_t w, uint32_t h, uint32_t tw) {
    const uint32_t th = 64u / tw;
    return ((w + tw - 1) / tw) * ((h + th - 1) / th);
}

void free_all() {
    for (auto& s : sv.scenes) {
        vt_lbvh_free(s.bvh);
        hipFree(s.records);
    }
    sv.scenes.clear();
    for (auto& gx : sv.graph)
        if (gx) hipGraphExecDestroy(gx);
    for (uint32_t j = 0; j < kSubBatches; ++j) {
        if (sv.sub[j]) hipStreamSynchronize(sv.sub[j]);
        for (auto& per_deep : sv.sub_graph[j])
            for (auto& gx : per_deep)
                if (gx) hipGraphExecDestroy(gx);
        hipFree(sv.sub_mem[j]);
        if (sv.acc_done[j]) hipEventDestroy(sv.acc_done[j]);
    }
    if (sv.start) hipEventDestroy(sv.start);
    hipFree(sv.arena_mem);
    hipFree(sv.wave_mem);
    for (auto& f : sv.fl) {
        if (f.done) hipEventDestroy(f.done);
        if (f.t0) hipEventDestroy(f.t0);
        if (f.t1) hipEventDestroy(f.t1);
    }
    if (sv.copy) hipStreamDestroy(sv.copy);
    sv = Serve{};
}

uint64_t per_sample(const ServeReq& r, uint32_t tile_w) {
    return (uint64_t)tiles_of(r.q.width, r.q.height, tile_w) * 64u;
}

// The iteration's chunks: (request id, samples) by policy, within the work budget and the
// wave buffers' capacity.
void choose(std::vector<std::pair<int64_t, uint32_t>>& out, uint64_t& work) {
    const vt_serve_config& c = sv.cfg;
    out.clear();
    work = 0;
    std::vector<int64_t> cand;
    for (int64_t id : sv.active) {
        const ServeReq& r = sv.reqs[id];
        if (r.next_s < r.q.spp) cand.push_back(id);
    }
    if (cand.empty()) return;
    const uint64_t W = c.max_work, cap = sv.wave.npaths;
    auto remaining = [&](int64_t id) {
        const ServeReq& r = sv.reqs[id];
        return per_sample(r, c.tile_w) * (r.q.spp - r.next_s);
    };
    auto add = [&](int64_t id, uint32_t ns) {
        out.emplace_back(id, ns);
        work += per_sample(sv.reqs[id], c.tile_w) * ns;
    };
    if (c.policy == VT_SERVE_ROUND_ROBIN) {
        for (int64_t id : cand) {
            if (out.size() >= VT_MAX_BATCH) break;
            const ServeReq& r = sv.reqs[id];
            const uint64_t ps = per_sample(r, c.tile_w);
            uint32_t ns = std::min(c.spp_chunk, r.q.spp - r.next_s);
            if (work + ps * ns > W) {
                const uint64_t fit = work < W ? (W - work) / ps : 0;
                if (fit == 0 && !out.empty()) continue;
                ns = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ns, fit));
            }
            if (work + ps * ns > cap) continue;
            add(id, ns);
        }
        return;
    }
    if (c.policy == VT_SERVE_SHORTEST_FIRST)
        std::stable_sort(cand.begin(), cand.end(), [&](int64_t a, int64_t b) { return remaining(a) < remaining(b); });
    // head: the first candidate (oldest, or shortest) takes what the budget allows; alone
    // (nothing else admitted or queued) it takes what the wave buffers hold, so a lone
    // request runs like a plain frame
    {
        const ServeReq& r = sv.reqs[cand[0]];
        const uint64_t ps = per_sample(r, c.tile_w);
        const bool alone = cand.size() == 1 && sv.queue.empty();
        const uint64_t fit = std::max<uint64_t>(1, (alone ? cap : W) / ps);
        add(cand[0], (uint32_t)std::min<uint64_t>(r.q.spp - r.next_s, fit));
    }
    if (c.policy == VT_SERVE_SHORTEST_FIRST) {
        for (size_t k = 1; k < cand.size() && out.size() < VT_MAX_BATCH; ++k) {
            const ServeReq& r = sv.reqs[cand[k]];
            const uint64_t ps = per_sample(r, c.tile_w);
            const uint64_t fit = work < W ? (W - work) / ps : 0;
            const uint32_t ns = (uint32_t)std::min<uint64_t>(r.q.spp - r.next_s, fit);
            if (ns > 0) add(cand[k], ns);
        }
        return;
    }
    // riders: requests whose whole remainder fits a quarter of the budget, smallest first
    std::vector<int64_t> rest(cand.begin() + 1, cand.end());
    std::stable_sort(rest.begin(), rest.end(), [&](int64_t a, int64_t b) { return remaining(a) < remaining(b); });
    const uint64_t limit = work + std::max<uint64_t>(W > work ? W - work : 0, W / 4);
    for (int64_t id : rest) {
        if (out.size() >= VT_MAX_BATCH) break;
        const uint64_t need = remaining(id);
        if (work + need > limit || work + need > cap) break;
        add(id, sv.reqs[id].q.spp - sv.reqs[id].next_s);
    }
}

// waits for the oldest in-flight iteration and retires its finished requests
uint32_t retire_oldest(int64_t* completed, uint32_t max_completed, uint32_t n_done, int* rc) {
    Inflight& f = sv.fl[sv.fl_head];
    if (hipEventSynchronize(f.done) != hipSuccess) {
        *rc = -1;
        return n_done;
    }
    hipEventElapsedTime(&sv.last_ms, f.t0, f.t1);
    ++sv.iterations;
    sv.samples += f.work;
    for (int64_t id : f.finishing) {
        ServeReq& r = sv.reqs[id];
        r.state = kDone;
        sv.active.erase(std::find(sv.active.begin(), sv.active.end(), id));
        if (completed && n_done < max_completed) completed[n_done] = id;
        ++n_done;
    }
    f.finishing.clear();
    sv.fl_head = (sv.fl_head + 1) % 2;
    --sv.fl_count;
    return n_done;
}

}  // namespace

extern "C" {

int32_t vt_serve_init(const vt_serve_config* cfg) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!vt_rt_inited() || !cfg) return -1;
    hipSetDevice(vt_rt_device());
    if (sv.on) free_all();
    vt_serve_config c = *cfg;
    if (c.max_slots == 0 || c.max_slots > VT_MAX_BATCH || c.memory_budget < 48 * 64 || c.max_work < 64 ||
        c.spp_chunk == 0 || c.max_bounces > 29 || (c.tile_w && 64u % c.tile_w != 0) || c.policy > 2 ||
        c.pipeline > 2 || c.streams > kSubBatches) {
        vt_rt_set_error("vt_serve_init: invalid configuration");
        return -1;
    }
    if (c.tile_w == 0) c.tile_w = 8;
    if (c.pipeline == 0) c.pipeline = 1;
    if (c.streams == 0) c.streams = 1;
    // path state: the work budget plus the riders' quarter, or one sample of the largest
    // frame the budget admits
    const size_t max_px = c.memory_budget / 48;
    const size_t one_spp = max_px + 64 * 2 * 16384 + 64;
    const size_t paths = std::max<size_t>((size_t)c.max_work + c.max_work / 4, one_spp);
    if (paths > 0x7FFFFFFFull) {
        vt_rt_set_error("vt_serve_init: work budget too large");
        return -1;
    }
    bool ok = hipMalloc(&sv.arena_mem, c.memory_budget) == hipSuccess &&
              hipMalloc(&sv.wave_mem, vt_wave_bytes((uint32_t)paths)) == hipSuccess &&
              hipStreamCreateWithFlags(&sv.copy, hipStreamNonBlocking) == hipSuccess;
    for (auto& f : sv.fl)
        ok = ok && hipEventCreateWithFlags(&f.done, hipEventDisableTiming) == hipSuccess &&
             hipEventCreate(&f.t0) == hipSuccess && hipEventCreate(&f.t1) == hipSuccess;
    const size_t sub_paths = (size_t)c.max_work + c.max_work / 4;
    ok = ok && hipEventCreateWithFlags(&sv.start, hipEventDisableTiming) == hipSuccess;
    for (uint32_t j = 0; j < c.streams && ok; ++j) {
        ok = hipEventCreateWithFlags(&sv.acc_done[j], hipEventDisableTiming) == hipSuccess;
        if (j > 0)
            ok = ok && (sv.sub[j] = vt_rt_chunk_stream(j)) != nullptr &&
                 hipMalloc(&sv.sub_mem[j], vt_wave_bytes((uint32_t)sub_paths)) == hipSuccess;
        if (ok && j > 0) vt_wave_bind(sv.sub_wave[j], sv.sub_mem[j], (uint32_t)sub_paths);
    }
    if (!ok) {
        vt_rt_set_error("vt_serve_init: allocation failed");
        free_all();
        return -1;
    }
    vt_wave_bind(sv.wave, sv.wave_mem, (uint32_t)paths);
    sv.arena.reset(c.memory_budget);
    sv.cfg = c;
    sv.on = true;
    return 0;
}

int32_t vt_serve_add_scene(const float* records, uint32_t n) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!sv.on || !records || n == 0) return -1;
    hipSetDevice(vt_rt_device());
    hipStream_t s = vt_rt_stream();
    ServeScene sc;
    sc.n = n;
    if (hipMalloc(&sc.records, (size_t)n * 64) != hipSuccess) return -1;
    hipMemcpyAsync(sc.records, records, (size_t)n * 64, hipMemcpyHostToDevice, s);
    if (vt_lbvh_build(sc.bvh, sc.records, n, vt_rt_vol_table(), vt_rt_nvol(), s) != 0) {
        hipFree(sc.records);
        vt_lbvh_free(sc.bvh);
        return -1;
    }
    hipMemcpyAsync(&sc.depth, sc.bvh.depth, 4, hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    sv.scenes.push_back(sc);
    return (int32_t)sv.scenes.size() - 1;
}

int64_t vt_serve_submit(const vt_request* q) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!sv.on || !q || q->scene >= sv.scenes.size() || q->width == 0 || q->height == 0 || q->width > 16384 ||
        q->height > 16384 || q->spp == 0 || q->max_bounces > sv.cfg.max_bounces || req_bytes(*q) > sv.cfg.memory_budget) {
        vt_rt_set_error("vt_serve_submit: invalid request");
        return -1;
    }
    ServeReq r;
    r.q = *q;
    vt_render_params p;
    std::memset(&p, 0, sizeof(p));
    p.width = q->width;
    p.height = q->height;
    p.viewport_square = q->viewport_square;
    p.mode = 2;
    p.spp = q->spp;
    p.max_bounces = q->max_bounces;
    p.jitter = 1;
    p.seed = q->seed;
    p.frame = q->frame;
    for (int k = 0; k < 3; ++k) {
        p.sun_dir[k] = q->sun_dir[k];
        p.sun_radiance[k] = q->sun_radiance[k];
        p.sky[k] = q->sky[k];
    }
    if (!vt_build_frame(q->perspective, q->camera, p, r.F)) {
        vt_rt_set_error("vt_serve_submit: singular projection or camera");
        return -1;
    }
    const int64_t id = sv.next_id++;
    sv.reqs[id] = r;
    sv.queue.push_back(id);
    return id;
}

int32_t vt_serve_step(int64_t* completed, uint32_t max_completed) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!sv.on) return -1;
    hipSetDevice(vt_rt_device());
    hipStream_t s = vt_rt_stream();
    const vt_serve_config& c = sv.cfg;
    // 1. admission (FIFO: the head waits for memory, later requests wait behind it)
    while (!sv.queue.empty() && sv.active.size() < c.max_slots) {
        ServeReq& r = sv.reqs[sv.queue.front()];
        size_t off;
        if (!sv.arena.alloc(req_bytes(r.q), &off)) break;
        r.off = off;
        r.bytes = req_bytes(r.q);
        r.state = kActive;
        sv.active.push_back(sv.queue.front());
        sv.queue.pop_front();
    }
    // 2. batching by policy
    std::vector<std::pair<int64_t, uint32_t>> chunks;
    uint64_t work = 0;
    if (sv.fl_count < c.pipeline) choose(chunks, work);
    int rc = 0;
    if (!chunks.empty()) {
        Inflight& f = sv.fl[(sv.fl_head + sv.fl_count) % 2];
        f.finishing.clear();
        // 3. split the iteration into sub-batches of about equal work: requests in order, a
        // request's samples cut at a boundary into consecutive pieces of consecutive
        // sub-batches; sub-batch j accumulates after sub-batch j-1, so each request's pieces
        // add up in sample order
        uint64_t max_ps = 0;
        for (auto& ch : chunks) max_ps = std::max(max_ps, per_sample(sv.reqs[ch.first], c.tile_w));
        uint32_t nsub = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(c.streams, work / kMinSubWork));
        const uint64_t sub_cap = (uint64_t)c.max_work + c.max_work / 4;
        uint64_t target = (work + nsub - 1) / nsub;
        if (nsub > 1 && target + max_ps > sub_cap) nsub = 1, target = work;
        std::vector<std::vector<ReqGPU>> batch(nsub);
        uint32_t j = 0;
        uint64_t load = 0;
        for (auto& ch : chunks) {
            ServeReq& r = sv.reqs[ch.first];
            const uint64_t ps = per_sample(r, c.tile_w);
            ReqGPU R;
            std::memset(&R, 0, sizeof(R));
            R.F = r.F;
            const ServeScene& sc = sv.scenes[r.q.scene];
            R.S.inst = sc.bvh.inst;
            R.S.nodes = sc.bvh.nodes;
            R.S.vols = vt_rt_vol_table();
            R.S.ninst = sc.n;
            R.S.nvol = vt_rt_nvol();
            R.S.bvh_depth = sc.depth;
            const size_t npx = (size_t)r.q.width * r.q.height;
            char* base = sv.arena_mem + r.off;
            R.hits = reinterpret_cast<HitRecord*>(base);
            R.acc = reinterpret_cast<float4*>(base + npx * 16);
            R.rad = reinterpret_cast<float4*>(base + npx * 32);
            R.tile_w = c.tile_w;
            uint32_t left = ch.second;
            while (left > 0) {
                if (j + 1 < nsub && load >= target) {
                    ++j;
                    load = 0;
                }
                uint32_t take = left;
                if (j + 1 < nsub)
                    take = (uint32_t)std::min<uint64_t>(left, std::max<uint64_t>(1, (target - load + ps - 1) / ps));
                R.s0 = r.next_s;
                R.s1 = r.next_s + take;
                R.final_chunk = R.s1 == r.q.spp ? 1u : 0u;
                batch[j].push_back(R);
                load += ps * take;
                r.next_s = R.s1;
                left -= take;
                if (R.final_chunk) f.finishing.push_back(ch.first);
            }
        }
        while (nsub > 1 && batch[nsub - 1].empty()) --nsub;
        for (uint32_t k = 0; k < nsub; ++k)
            if (batch[k].size() > VT_MAX_BATCH) nsub = 0;  // cannot happen: pieces <= chunks + nsub - 1
        if (nsub == 0) {
            vt_rt_set_error("vt_serve_step: batch too large");
            return -1;
        }
        // 4. one launch sequence per sub-batch, sub-batch 0 on the library stream
        bool deep = false;
        for (auto& ch : chunks) deep = deep || sv.scenes[sv.reqs[ch.first].q.scene].depth + 1u > 16u;
        hipEventRecord(f.t0, s);
        hipEventRecord(sv.start, s);
        // persistent workgroups per launch: 4 per CU alone, 2 per CU with 3-4 sub-batches in flight
        const uint32_t cus = vt_rt_cus();
        const uint32_t blocks = (nsub > 1 && cus) ? cus * std::min(4u, std::max(2u, 8u / nsub)) : 0u;
        for (uint32_t k = 0; k < nsub && rc == 0; ++k) {
            hipStream_t sk = k == 0 ? s : sv.sub[k];
            const WaveBuffers& wk = k == 0 ? sv.wave : sv.sub_wave[k];
            if (k > 0) hipStreamWaitEvent(sk, sv.start, 0);
            hipEvent_t after = k == 0 ? nullptr : sv.acc_done[k - 1];
            if (c.use_graph) {
                hipGraphExec_t& gx = nsub == 1 ? sv.graph[deep ? 1 : 0] : sv.sub_graph[k][deep ? 1 : 0][nsub - 1];
                if (!gx) rc = vt_trace_capture(c.max_bounces, deep, wk, sk, &gx, false, blocks);
                if (rc == 0) rc = vt_trace_upload_tables(batch[k].data(), (uint32_t)batch[k].size(), wk, sk, nullptr);
                if (rc == 0) rc = hipGraphLaunch(gx, sk) == hipSuccess ? 0 : -1;
                if (rc == 0 && after) hipStreamWaitEvent(sk, after, 0);
                if (rc == 0) rc = vt_trace_accumulate(wk, sk, blocks);
            } else {
                rc = vt_trace_batch(batch[k].data(), (uint32_t)batch[k].size(), wk, nullptr, nullptr, nullptr, nullptr,
                                    sk, after, blocks);
            }
            hipEventRecord(sv.acc_done[k], sk);
        }
        if (nsub > 1) hipStreamWaitEvent(s, sv.acc_done[nsub - 1], 0);
        hipEventRecord(f.t1, s);
        hipEventRecord(f.done, s);
        if (rc != 0) {
            vt_rt_set_error("vt_serve_step: launch failed");
            return -1;
        }
        f.work = work;
        sv.sub_batches += nsub;
        ++sv.fl_count;
    }
    // 4. retire: keep at most pipeline - 1 iterations in flight after a launch; with nothing
    // to launch, wait for the oldest one.  Polling: retire only what has finished, never wait.
    uint32_t n_done = 0;
    if (c.poll) {
        while (sv.fl_count > 0 && rc == 0) {
            const hipError_t q = hipEventQuery(sv.fl[sv.fl_head].done);
            if (q == hipErrorNotReady) break;
            if (q != hipSuccess) rc = -1;
            else n_done = retire_oldest(completed, max_completed, n_done, &rc);
        }
    } else if (sv.fl_count > 0 && (sv.fl_count >= c.pipeline || chunks.empty())) {
        n_done = retire_oldest(completed, max_completed, n_done, &rc);
    }
    if (rc != 0) {
        vt_rt_set_error("vt_serve_step: iteration failed");
        return -1;
    }
    return (int32_t)n_done;
}

int32_t vt_serve_read(int64_t id, float* rgba, uint64_t bytes, void* hits, uint64_t hit_bytes) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!sv.on) return -1;
    auto it = sv.reqs.find(id);
    if (it == sv.reqs.end() || it->second.state != kDone) return -1;
    ServeReq& r = it->second;
    const size_t npx = (size_t)r.q.width * r.q.height;
    if ((rgba && bytes < npx * 16) || (hits && hit_bytes < npx * 16)) return -1;
    hipSetDevice(vt_rt_device());
    // the request's last iteration has completed (its event was waited for), so the copy may
    // run beside the iteration now in flight
    const char* base = sv.arena_mem + r.off;
    if (rgba) hipMemcpyAsync(rgba, base + npx * 32, npx * 16, hipMemcpyDeviceToHost, sv.copy);
    if (hits) hipMemcpyAsync(hits, base, npx * 16, hipMemcpyDeviceToHost, sv.copy);
    if (hipStreamSynchronize(sv.copy) != hipSuccess) return -1;
    sv.arena.release(r.off, r.bytes);
    sv.reqs.erase(it);
    return 0;
}

int32_t vt_serve_status(vt_serve_stats* st) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    if (!sv.on || !st) return -1;
    std::memset(st, 0, sizeof(*st));
    st->queued = (uint32_t)sv.queue.size();
    st->active = (uint32_t)sv.active.size();
    uint32_t done = 0;
    for (auto& kv : sv.reqs) done += kv.second.state == kDone ? 1u : 0u;
    st->done_unread = done;
    st->memory_in_use = sv.arena.in_use();
    st->iterations = sv.iterations;
    st->samples = sv.samples;
    st->last_iteration_ms = sv.last_ms;
    st->sub_batches = sv.sub_batches;
    st->tile_w = sv.cfg.tile_w;
    return 0;
}

void vt_serve_shutdown(void) {
    std::lock_guard<std::mutex> lk(vt_rt_lock());
    vt_serve_release();
}

}  // extern "C"

// Frees the serving state; the caller holds the runtime lock (cleanup() calls it too).
void vt_serve_release() {
    if (!sv.on) return;
    hipSetDevice(vt_rt_device());
    hipStreamSynchronize(vt_rt_stream());
    free_all();
}
