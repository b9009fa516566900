// lbvh.hip — GPU-built LBVH over instance boxes (replaces lib/raytrace.c's one-AABB BLAS,
// lib/raytrace.c:20-96, which the reference builds but never binds or writes).
//
// Pipeline (one stream, no host round trip):
//   k_instance_prep   ABI record (src/render.rs:74-78: column-major mat4, texture id bits
//                     in m[3][3], shaders/trace.vert:37-40) -> M, inverse(M) (FP64, rounded),
//                     padded world AABB of the unit proxy cube (lib/memory.c:22-40), centroid;
//   k_bounds          centroid bounds (one workgroup);
//   k_morton          30-bit Morton code of the normalized centroid;
//   k_radix_sort      stable LSD radix sort (4 x 8-bit digits) in one workgroup, chunked
//                     through LDS: per-wave digit match by ballots, per-wave counts, prefix;
//                     above kSortOneGroupMax keys the same passes run one workgroup per
//                     1024-key chunk (k_sort_hist, k_sort_scan, k_sort_scatter);
//   k_karras          internal nodes (Karras, HPG 2012), leaves, parent links;
//   k_refit           bottom-up AABB union, second arrival at a node continues (agent-scope
//                     atomics and loads, so no stale L1 data is read).
// Integer parts (codes, order, topology) and the FP32 min/max unions are bit-exact with the
// CPU oracle; the FP32 arithmetic (corner transform, padding, centroid) follows the same
// operation order.
#include <hip/hip_runtime.h>

#include "vt_internal.h"

namespace {

constexpr int kSortThreads = 1024;
constexpr int kSortWaves = kSortThreads / 64;
// keys sorted by the one-workgroup kernel; more keys: one workgroup per chunk per pass
constexpr uint32_t kSortOneGroupMax = 4096;

__device__ inline float gmin(float x, float y) { return (y < x) ? y : x; }
__device__ inline float gmax(float x, float y) { return (x < y) ? y : x; }
__device__ inline float pad_of(float v) { return fabsf(v) * 9.5367431640625e-07f + 9.5367431640625e-07f; }

// FP64 Laplace-expansion inverse (same formulas as the host vt_inverse4).
__device__ bool inverse4_d(const float* mf, float* out) {
#define A(r, c) ((double)mf[(c) * 4 + (r)])
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
    if (!(det != 0.0) || !isfinite(det)) return false;
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
#undef A
    bool ok = true;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            float v = (float)b[r * 4 + c];
            if (!isfinite(v)) ok = false;
            out[c * 4 + r] = v;
        }
    return ok;
}

__global__ void k_instance_prep(const float* records, uint32_t n, const VolumeDesc* vols, uint32_t nvol,
                                InstanceGPU* inst, BvhNode* leaves_tmp, float* cen) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float M[16];
    for (int k = 0; k < 16; ++k) M[k] = records[16 * (size_t)i + k];
    uint32_t tex = __float_as_uint(M[15]);
    M[15] = 1.0f;
    float inv[16];
    bool ok = inverse4_d(M, inv);
    InstanceGPU g;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 3; ++r) {
            g.M[c * 3 + r] = M[c * 4 + r];
            g.inv[c * 3 + r] = inv[c * 4 + r];
        }
    g.tex = tex;
    g.valid = (ok && tex < nvol) ? 1u : 0u;
    for (int a = 0; a < 3; ++a) {
        uint32_t d = g.valid ? vols[tex].dim[a] : 1u;
        g.dim[a] = d;
        g.size[a] = (float)d;
    }
    g.pad[0] = g.pad[1] = 0;
    inst[i] = g;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < 8; ++c) {
        float p[3] = {(c & 1) ? 0.5f : -0.5f, (c & 2) ? 0.5f : -0.5f, (c & 4) ? 0.5f : -0.5f};
        for (int a = 0; a < 3; ++a) {
            float w = ((M[0 * 4 + a] * p[0] + M[1 * 4 + a] * p[1]) + M[2 * 4 + a] * p[2]) + M[3 * 4 + a] * 1.0f;
            lo[a] = gmin(lo[a], w);
            hi[a] = gmax(hi[a], w);
        }
    }
    BvhNode L;
    for (int a = 0; a < 3; ++a) {
        L.lo[a] = ok ? lo[a] - pad_of(lo[a]) : 0.0f;
        L.hi[a] = ok ? hi[a] + pad_of(hi[a]) : 0.0f;
        cen[3 * (size_t)i + a] = (L.lo[a] + L.hi[a]) * 0.5f;
    }
    L.left = i;
    L.right = VT_EMPTY;
    leaves_tmp[i] = L;
}

__global__ __launch_bounds__(1024) void k_bounds(const float* cen, uint32_t n, float* bounds) {
    __shared__ float smin[3][1024], smax[3][1024];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        for (int a = 0; a < 3; ++a) {
            float c = cen[3 * (size_t)i + a];
            lo[a] = gmin(lo[a], c);
            hi[a] = gmax(hi[a], c);
        }
    for (int a = 0; a < 3; ++a) { smin[a][threadIdx.x] = lo[a]; smax[a][threadIdx.x] = hi[a]; }
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int a = 0; a < 3; ++a) {
                smin[a][threadIdx.x] = gmin(smin[a][threadIdx.x], smin[a][threadIdx.x + s]);
                smax[a][threadIdx.x] = gmax(smax[a][threadIdx.x], smax[a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        bounds[threadIdx.x] = smin[threadIdx.x][0];
        bounds[3 + threadIdx.x] = smax[threadIdx.x][0];
    }
}

__device__ inline uint32_t expand10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_morton(const float* cen, uint32_t n, const float* bounds, uint32_t* keys, uint32_t* vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
        float ext = bounds[3 + a] - bounds[a];
        float u = ext > 0.0f ? (cen[3 * (size_t)i + a] - bounds[a]) / ext : 0.0f;
        float s = floorf(u * 1024.0f);
        int qi = (int)s;
        qi = qi < 0 ? 0 : (qi > 1023 ? 1023 : qi);
        q[a] = (uint32_t)qi;
    }
    keys[i] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
    vals[i] = i;
}

// Rank of a key among the keys of its 1024-key chunk with the same digit d that come before
// it (a stable counting sort's offset within the digit): the wave's keys of digit d are
// matched by 8 ballots, counted per wave, and the counts prefixed over the waves.  On return
// wcount[w][d] holds the keys of digit d in waves before w; tot[d] (thread d < 256 writes it)
// the chunk's keys of digit d.  Called by every thread of the workgroup.
__device__ __forceinline__ uint32_t chunk_rank(uint32_t d, bool valid, uint32_t (*wcount)[256], uint32_t* tot) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint32_t j = tid; j < kSortWaves * 256; j += kSortThreads) (&wcount[0][0])[j] = 0;
    __syncthreads();
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
        uint64_t bal = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    const uint64_t lower = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t rank = (uint32_t)__popcll(peers & lower);
    if (valid && rank == 0) wcount[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (tid < 256) {  // prefix over waves, per digit
        uint32_t acc = 0;
        for (int w = 0; w < kSortWaves; ++w) {
            uint32_t c = wcount[w][tid];
            wcount[w][tid] = acc;
            acc += c;
        }
        tot[tid] = acc;
    }
    __syncthreads();
    return wcount[wave][d] + rank;
}

// Stable LSD radix sort, one workgroup, chunks of 1024 keys through LDS.
__global__ __launch_bounds__(kSortThreads) void k_radix_sort(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                                                             uint32_t n) {
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcount[kSortWaves][256];
    __shared__ uint32_t hist[256];
    const uint32_t tid = threadIdx.x;
    uint32_t *kin = k0, *vin = v0, *kout = k1, *vout = v1;
    for (int pass = 0; pass < 4; ++pass) {
        const uint32_t shift = pass * 8;
        if (tid < 256) hist[tid] = 0;
        __syncthreads();
        for (uint32_t i = tid; i < n; i += kSortThreads) atomicAdd(&hist[(kin[i] >> shift) & 255u], 1u);
        __syncthreads();
        if (tid == 0) {
            uint32_t acc = 0;
            for (int d = 0; d < 256; ++d) {
                base[d] = acc;
                acc += hist[d];
            }
        }
        __syncthreads();
        for (uint32_t c0 = 0; c0 < n; c0 += kSortThreads) {
            const uint32_t i = c0 + tid;
            const bool valid = i < n;
            uint32_t key = valid ? kin[i] : 0u, val = valid ? vin[i] : 0u;
            uint32_t d = (key >> shift) & 255u;
            const uint32_t r = chunk_rank(d, valid, wcount, hist);   // hist: chunk total per digit
            if (valid) {
                uint32_t pos = base[d] + r;
                kout[pos] = key;
                vout[pos] = val;
            }
            __syncthreads();
            if (tid < 256) base[tid] += hist[tid];
            __syncthreads();
        }
        uint32_t* t;
        t = kin; kin = kout; kout = t;
        t = vin; vin = vout; vout = t;
        __syncthreads();
    }
    // 4 passes: result is back in (k0, v0)
}

// The same passes with one workgroup per 1024-key chunk.  k_sort_hist: the chunk's digit counts,
// stored digit-major (ghist[d * nch + c]), so an exclusive scan of ghist gives each chunk's first
// position for each digit, chunks in order.
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint32_t* kin, uint32_t n, uint32_t shift,
                                                            uint32_t* ghist) {
    __shared__ uint32_t hist[256];
    const uint32_t tid = threadIdx.x, c = blockIdx.x, nch = gridDim.x;
    if (tid < 256) hist[tid] = 0;
    __syncthreads();
    const uint32_t i = c * kSortThreads + tid;
    if (i < n) atomicAdd(&hist[(kin[i] >> shift) & 255u], 1u);
    __syncthreads();
    if (tid < 256) ghist[tid * nch + c] = hist[tid];
}

// Exclusive scan of ghist[0, total) in place, one workgroup: each thread sums a contiguous
// segment, the segment sums are scanned in LDS, then each segment is rewritten.
__global__ __launch_bounds__(kSortThreads) void k_sort_scan(uint32_t* ghist, uint32_t total) {
    __shared__ uint32_t part[kSortThreads];
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (total + kSortThreads - 1) / kSortThreads;
    const uint32_t b = min(tid * per, total), e = min(b + per, total);
    uint32_t s = 0;
    for (uint32_t j = b; j < e; ++j) s += ghist[j];
    part[tid] = s;
    __syncthreads();
    for (uint32_t off = 1; off < (uint32_t)kSortThreads; off <<= 1) {
        const uint32_t v = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t acc = part[tid] - s;
    for (uint32_t j = b; j < e; ++j) {
        const uint32_t c = ghist[j];
        ghist[j] = acc;
        acc += c;
    }
}

// Scatter of chunk c: position = the chunk's first slot for the digit + the key's stable rank.
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                                               uint32_t* vout, uint32_t n, uint32_t shift,
                                                               const uint32_t* ghist) {
    __shared__ uint32_t wcount[kSortWaves][256];
    __shared__ uint32_t tot[256];
    const uint32_t tid = threadIdx.x, c = blockIdx.x, nch = gridDim.x;
    const uint32_t i = c * kSortThreads + tid;
    const bool valid = i < n;
    const uint32_t key = valid ? kin[i] : 0u, val = valid ? vin[i] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    const uint32_t r = chunk_rank(d, valid, wcount, tot);
    if (valid) {
        const uint32_t pos = ghist[d * nch + c] + r;
        kout[pos] = key;
        vout[pos] = val;
    }
}

__device__ inline int delta_k(const uint32_t* k, int i, int j, int n) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

__global__ void k_leaves(const uint32_t* order, uint32_t n, const BvhNode* leaves_tmp, BvhNode* nodes) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    nodes[n - 1 + k] = leaves_tmp[order[k]];
}

__global__ void k_karras(const uint32_t* K, int n, BvhNode* nodes, int32_t* parent, uint32_t* counters) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) parent[0] = -1;
    if (i >= n - 1) return;
    counters[i] = 0;
    int d = (delta_k(K, i, i + 1, n) - delta_k(K, i, i - 1, n)) >= 0 ? 1 : -1;
    int dmin = delta_k(K, i, i - d, n);
    int lmax = 2;
    while (delta_k(K, i, i + lmax * d, n) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta_k(K, i, i + (l + t) * d, n) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta_k(K, i, j, n);
    int s = 0;
    for (int div = 2;; div *= 2) {
        int tt = (l + div - 1) / div;
        if (delta_k(K, i, i + (s + tt) * d, n) > dnode) s += tt;
        if (tt <= 1) break;
    }
    int gamma = i + s * d + (d < 0 ? d : 0);
    int first = d > 0 ? i : j;
    int last = d > 0 ? j : i;
    uint32_t left = (first == gamma) ? (uint32_t)(n - 1 + gamma) : (uint32_t)gamma;
    uint32_t right = (last == gamma + 1) ? (uint32_t)(n - 1 + gamma + 1) : (uint32_t)(gamma + 1);
    nodes[i].left = left;
    nodes[i].right = right;
    parent[left] = i;
    parent[right] = i;
}

__device__ inline float load_f32_agent(const float* p) {
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT));
}
__device__ inline void store_f32_agent(float* p, float v) {
    __hip_atomic_store(reinterpret_cast<uint32_t*>(p), __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_refit(uint32_t n, BvhNode* nodes, const int32_t* parent, uint32_t* counters) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int node = parent[n - 1 + k];
    while (node >= 0) {
        uint32_t old = __hip_atomic_fetch_add(&counters[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) return;  // first arrival: the sibling's thread continues
        BvhNode* N = &nodes[node];
        const uint32_t l = N->left, r = N->right;
        for (int a = 0; a < 3; ++a) {
            float llo = load_f32_agent(&nodes[l].lo[a]), rlo = load_f32_agent(&nodes[r].lo[a]);
            float lhi = load_f32_agent(&nodes[l].hi[a]), rhi = load_f32_agent(&nodes[r].hi[a]);
            store_f32_agent(&N->lo[a], gmin(llo, rlo));
            store_f32_agent(&N->hi[a], gmax(lhi, rhi));
        }
        __atomic_thread_fence(__ATOMIC_RELEASE);
        node = parent[node];
    }
}

__global__ void k_zero_counters(uint32_t* counters, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counters[i] = 0u;
}

// Depth of the tree (edges on the longest root-leaf path): one thread per leaf walks up.
__global__ void k_depth(uint32_t n, const int32_t* parent, uint32_t* depth_out) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t d = 0;
    int node = parent[n - 1 + k];
    while (node >= 0 && d < 4096u) {
        ++d;
        node = parent[node];
    }
    atomicMax(depth_out, d);
}

}  // namespace

int vt_lbvh_reserve(LbvhWork& w, uint32_t n) {
    if (n <= w.cap && w.inst) return 0;
    vt_lbvh_free(w);
    uint32_t cap = 1;
    while (cap < n) cap *= 2;
    size_t nn = 2 * (size_t)cap;
    bool ok = hipMalloc(&w.inst, cap * sizeof(InstanceGPU)) == hipSuccess &&
              hipMalloc(&w.nodes, nn * sizeof(BvhNode)) == hipSuccess &&
              hipMalloc(&w.cen, 3 * (size_t)cap * sizeof(float) + cap * sizeof(BvhNode)) == hipSuccess &&
              hipMalloc(&w.keys[0], cap * 4) == hipSuccess && hipMalloc(&w.keys[1], cap * 4) == hipSuccess &&
              hipMalloc(&w.vals[0], cap * 4) == hipSuccess && hipMalloc(&w.vals[1], cap * 4) == hipSuccess &&
              hipMalloc(&w.parent, nn * 4) == hipSuccess && hipMalloc(&w.counters, cap * 4) == hipSuccess &&
              hipMalloc(&w.bounds, 6 * 4) == hipSuccess && hipMalloc(&w.depth, 4) == hipSuccess &&
              hipMalloc(&w.sort_hist, 256 * 4 * ((size_t)(cap + kSortThreads - 1) / kSortThreads)) == hipSuccess;
    if (!ok) {
        vt_lbvh_free(w);
        return -1;
    }
    w.cap = cap;
    return 0;
}

void vt_lbvh_free(LbvhWork& w) {
    hipFree(w.inst); hipFree(w.nodes); hipFree(w.cen);
    hipFree(w.keys[0]); hipFree(w.keys[1]); hipFree(w.vals[0]); hipFree(w.vals[1]);
    hipFree(w.parent); hipFree(w.counters); hipFree(w.bounds); hipFree(w.depth); hipFree(w.sort_hist);
    w = LbvhWork{};
}

// Refit: new records for the same n instances keep the last build's topology (leaf order and
// parents); instance data and leaf boxes are recomputed, internal boxes rebuilt bottom-up.
int vt_lbvh_refit(LbvhWork& w, const float* records_dev, uint32_t n, const VolumeDesc* vols_dev, uint32_t nvol,
                  hipStream_t s) {
    if (n == 0 || n > w.cap || !w.inst) return -1;
    const uint32_t tb = 256, nb = (n + tb - 1) / tb;
    BvhNode* leaves_tmp = reinterpret_cast<BvhNode*>(w.cen + 3 * (size_t)w.cap);
    hipLaunchKernelGGL(k_instance_prep, dim3(nb), dim3(tb), 0, s, records_dev, n, vols_dev, nvol, w.inst, leaves_tmp,
                       w.cen);
    hipLaunchKernelGGL(k_leaves, dim3(nb), dim3(tb), 0, s, w.vals[0], n, leaves_tmp, w.nodes);
    if (n > 1) {
        hipLaunchKernelGGL(k_zero_counters, dim3(nb), dim3(tb), 0, s, w.counters, n - 1);
        hipLaunchKernelGGL(k_refit, dim3(nb), dim3(tb), 0, s, n, w.nodes, w.parent, w.counters);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int vt_lbvh_build(LbvhWork& w, const float* records_dev, uint32_t n, const VolumeDesc* vols_dev, uint32_t nvol,
                  hipStream_t s) {
    if (n == 0) return 0;
    if (vt_lbvh_reserve(w, n) != 0) return -1;
    const uint32_t tb = 256, nb = (n + tb - 1) / tb;
    BvhNode* leaves_tmp = reinterpret_cast<BvhNode*>(w.cen + 3 * (size_t)w.cap);
    hipLaunchKernelGGL(k_instance_prep, dim3(nb), dim3(tb), 0, s, records_dev, n, vols_dev, nvol, w.inst, leaves_tmp,
                       w.cen);
    hipLaunchKernelGGL(k_bounds, dim3(1), dim3(1024), 0, s, w.cen, n, w.bounds);
    hipLaunchKernelGGL(k_morton, dim3(nb), dim3(tb), 0, s, w.cen, n, w.bounds, w.keys[0], w.vals[0]);
    if (n <= kSortOneGroupMax) {
        hipLaunchKernelGGL(k_radix_sort, dim3(1), dim3(kSortThreads), 0, s, w.keys[0], w.vals[0], w.keys[1], w.vals[1],
                           n);
    } else {
        const uint32_t nch = (n + kSortThreads - 1) / kSortThreads;   // 256 * nch <= sort_hist's size
        for (uint32_t pass = 0; pass < 4; ++pass) {   // pass p reads buffer p & 1: the result ends in (keys[0], vals[0])
            const uint32_t a = pass & 1u, b = a ^ 1u, shift = 8 * pass;
            hipLaunchKernelGGL(k_sort_hist, dim3(nch), dim3(kSortThreads), 0, s, w.keys[a], n, shift, w.sort_hist);
            hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(kSortThreads), 0, s, w.sort_hist, 256 * nch);
            hipLaunchKernelGGL(k_sort_scatter, dim3(nch), dim3(kSortThreads), 0, s, w.keys[a], w.vals[a], w.keys[b],
                               w.vals[b], n, shift, w.sort_hist);
        }
    }
    hipLaunchKernelGGL(k_leaves, dim3(nb), dim3(tb), 0, s, w.vals[0], n, leaves_tmp, w.nodes);
    hipMemsetAsync(w.depth, 0, 4, s);
    if (n > 1) {
        hipLaunchKernelGGL(k_karras, dim3(nb), dim3(tb), 0, s, w.keys[0], (int)n, w.nodes, w.parent, w.counters);
        hipLaunchKernelGGL(k_refit, dim3(nb), dim3(tb), 0, s, n, w.nodes, w.parent, w.counters);
        hipLaunchKernelGGL(k_depth, dim3(nb), dim3(tb), 0, s, n, w.parent, w.depth);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
