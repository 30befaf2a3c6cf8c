// pfh_kernels.hip - gfx950 kernels of the D3D11-level compute path.
//
// Stage map (reference shader -> kernels here):
//   dice.cs.glsl      -> k_dice_count, scan, k_dice_emit
//   bound.cs.glsl     -> k_bound
//   bin.cs.glsl       -> k_bin<count>, scan, k_bin<emit>, k_fill_link
//   propagate.cs.glsl -> k_propagate<count>, scan, k_propagate<emit>
//   sort.cs.glsl      -> k_row_paths, k_sort_tiles (also builds the z-buffer and
//                        the first-tile map that propagate builds with atomics)
//   fill.cs.glsl      -> k_fill
//   tile.cs.glsl      -> k_tile
//
// The reference orders microlines, fills, per-tile fill lists, alpha tiles and
// pre-sort tile lists with global atomics, so its output depends on the order
// invocations happen to run. Here every such order is the sequential one
// (invocation i before i+1): counts, exclusive scans, then writes at the
// scanned offsets; per-tile fill lists come from a per-path sort; tile lists
// come from an ordered walk of the paths covering each framebuffer tile. The
// only atomics left are integer adds of backdrop deltas (order-independent)
// and one exchange per (tile, sort chunk) in k_fill_link.
//
// Floating point: built with -ffp-contract=off; division and sqrt are
// correctly rounded (HIP defaults). Every formula follows the reference
// shader's evaluation order with mix(a, b, t) = a + t * (b - a) and
// round() = round half to even.
#include "pfh_kernels.h"

#include <hip/hip_fp16.h>

namespace pfh {

namespace {

constexpr uint32_t kFlagQuad = 0x80000000u;
constexpr uint32_t kFlagCubic = 0x40000000u;
constexpr int kBlock = 256;
constexpr uint32_t kLutRows = 259;    // padded area LUT: 256 texels + an edge copy on each side
constexpr uint32_t kLutStride = 260;  // texels per padded LUT row
constexpr uint32_t kScanTile = 4096;  // elements per scan block
constexpr uint32_t kMaxGrid = 16384;

__device__ __forceinline__ float mixf(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ float min_glsl(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float max_glsl(float x, float y) { return x < y ? y : x; }
__device__ __forceinline__ float clamp_glsl(float x, float lo, float hi) {
    return min_glsl(max_glsl(x, lo), hi);
}
__device__ __forceinline__ uint32_t f2u_sat(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)x;
}
__device__ __forceinline__ int32_t f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x < -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)x;
}
__device__ __forceinline__ uint32_t unorm8(float x) {
    float c = clamp_glsl(x, 0.0f, 1.0f);
    return (uint32_t)floorf(c * 255.0f + 0.5f);
}
__device__ __forceinline__ int32_t sext24(uint32_t w) { return (int32_t)(w << 8) >> 8; }

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

inline uint32_t grid_for(uint64_t items, uint32_t per_block, uint32_t cap = kMaxGrid) {
    uint64_t g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (cap < 1) cap = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

// x of the lane dpp_ctrl names, or 0 where it names none or the lane's row
// is not in row_mask.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint32_t dpp_or_zero(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kCtrl, kRowMask, 0xf, false);
}

// Inclusive scan over the 64 lanes of a wavefront; every lane must be
// active. Within each row of 16 lanes: x += x of the lane 1, 2, 4, 8 below
// (row_shr); then lane 15's total is added to rows 1 and 3 (row_bcast:15)
// and lane 31's total, now that of rows 0-1, to rows 2 and 3 (row_bcast:31).
// Register moves between lanes, no LDS round trips.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
    x += dpp_or_zero<0x111, 0xf>(x);
    x += dpp_or_zero<0x112, 0xf>(x);
    x += dpp_or_zero<0x114, 0xf>(x);
    x += dpp_or_zero<0x118, 0xf>(x);
    x += dpp_or_zero<0x142, 0xa>(x);
    x += dpp_or_zero<0x143, 0xc>(x);
    return x;
}

// The value of lane 63 (the total of an inclusive scan), in an SGPR.
__device__ __forceinline__ uint32_t last_lane(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// Exclusive scan of one value per thread over a 256-thread block (4 waves).
__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *lds4, uint32_t *total) {
    uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    uint32_t x = wave_inclusive_scan(v);
    if (lane == 63) lds4[wave] = x;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4; w++) {
        uint32_t s = lds4[w];
        before += (w < wave) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return before + x - v;
}

__device__ uint32_t block_reduce_sum(uint32_t v, uint32_t *lds4) {
    v = last_lane(wave_inclusive_scan(v));
    if (lane_id() == 0) lds4[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t t = lds4[0] + lds4[1] + lds4[2] + lds4[3];
    __syncthreads();
    return t;
}

// ------------------------------------------------------------------ scan
// Exclusive scan of n u32 values in place (reduce, scan partials, downsweep).
// n is read from n_ptr when given (a count produced on the device) and is
// clamped to `cap`; out[n] receives the total, and so does *total_out.

__global__ void __launch_bounds__(kBlock) k_scan_reduce(const uint32_t *data,
                                                        const uint32_t *n_ptr, uint32_t n_host,
                                                        uint32_t cap, uint32_t *partials) {
    __shared__ uint32_t lds4[4];
    uint32_t n = n_ptr ? min(*n_ptr, cap) : n_host;
    uint32_t base = blockIdx.x * kScanTile;
    if (base >= n) return;
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanTile / kBlock; k++) {
        uint32_t i = base + k * kBlock + threadIdx.x;
        uint32_t v = data[min(i, n - 1u)];  // unconditional: the loads overlap
        s += i < n ? v : 0u;
    }
    uint32_t t = block_reduce_sum(s, lds4);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// The scan of the block partials (one workgroup); returns the total in every
// thread.
__device__ uint32_t scan_partials_total(uint32_t *partials, const uint32_t *n_ptr,
                                        uint32_t n_host, uint32_t cap, uint32_t *total_out,
                                        uint32_t *data_end) {
    __shared__ uint32_t lds4[4];
    uint32_t n = n_ptr ? min(*n_ptr, cap) : n_host;
    uint32_t nb = (n + kScanTile - 1) / kScanTile;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += kBlock) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < nb ? partials[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(v, lds4, &tot);
        if (i < nb) partials[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        if (total_out) *total_out = carry;
        data_end[n] = carry;
    }
    return carry;
}

// A count the scan's total gives a buffer: the total clamped to the buffer's
// capacity goes to *clamped, and overflow gets `bit` when it does not fit.
// clamped == nullptr: none.
struct ScanCount {
    uint32_t *clamped;
    uint32_t cap;
    uint32_t *overflow;
    uint32_t bit;
};

__global__ void __launch_bounds__(kBlock) k_scan_partials(uint32_t *partials,
                                                          const uint32_t *n_ptr, uint32_t n_host,
                                                          uint32_t cap, uint32_t *total_out,
                                                          uint32_t *data_end, ScanCount count) {
    uint32_t t = scan_partials_total(partials, n_ptr, n_host, cap, total_out, data_end);
    if (threadIdx.x == 0 && count.clamped) {
        *count.clamped = t < count.cap ? t : count.cap;
        if (t > count.cap) atomicOr(count.overflow, count.bit);
    }
}

__global__ void __launch_bounds__(kBlock) k_scan_downsweep(uint32_t *data, const uint32_t *n_ptr,
                                                           uint32_t n_host, uint32_t cap,
                                                           const uint32_t *partials) {
    __shared__ uint32_t lds4[4];
    uint32_t n = n_ptr ? min(*n_ptr, cap) : n_host;
    uint32_t base = blockIdx.x * kScanTile;
    if (base >= n) return;
    uint32_t carry = partials[blockIdx.x];
    for (uint32_t k = 0; k < kScanTile / (4 * kBlock); k++) {
        uint32_t i0 = base + k * 4 * kBlock + threadIdx.x * 4;
        uint32_t v[4];
        if (i0 + 3 < n) {
            uint4 q = *reinterpret_cast<const uint4 *>(data + i0);
            v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) v[j] = (i0 + j < n) ? data[i0 + j] : 0u;
        }
        uint32_t s = v[0] + v[1] + v[2] + v[3];
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(s, lds4, &tot) + carry;
        uint32_t o[4];
        o[0] = ex;
        o[1] = ex + v[0];
        o[2] = o[1] + v[1];
        o[3] = o[2] + v[2];
        if (i0 + 3 < n) {
            *reinterpret_cast<uint4 *>(data + i0) = make_uint4(o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (i0 + j < n) data[i0 + j] = o[j];
        }
        carry += tot;
    }
}

uint32_t scan_blocks(const uint32_t *n_ptr, uint32_t n_host, uint32_t cap) {
    uint32_t max_n = n_ptr ? cap : n_host;
    uint32_t blocks = (max_n + kScanTile - 1) / kScanTile;
    return blocks < 1 ? 1 : blocks;
}

void scan_exclusive(uint32_t *data, const uint32_t *n_ptr, uint32_t n_host, uint32_t cap,
                    uint32_t *partials, uint32_t *total_out, hipStream_t s,
                    ScanCount count = ScanCount{nullptr, 0, nullptr, 0}) {
    uint32_t blocks = scan_blocks(n_ptr, n_host, cap);
    hipLaunchKernelGGL(k_scan_reduce, dim3(blocks), dim3(kBlock), 0, s, data, n_ptr, n_host, cap,
                       partials);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kBlock), 0, s, partials, n_ptr, n_host, cap,
                       total_out, data, count);
    hipLaunchKernelGGL(k_scan_downsweep, dim3(blocks), dim3(kBlock), 0, s, data, n_ptr, n_host,
                       cap, partials);
}

// ------------------------------------------------------------------ batch init

__global__ void __launch_bounds__(kBlock) k_batch_init(BatchArgs a) {
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    if (gid == 0) {
        BatchCounters c = {};
        *a.counters = c;
    }
    uint32_t fbn = (uint32_t)(a.fb_width_tiles * a.fb_height_tiles);
    for (uint32_t i = gid; i < fbn + 8; i += stride) a.zbuffer[i] = 0;
    for (uint32_t i = gid; i < fbn; i += stride) a.first_tile_map[i] = -1;
    for (uint32_t i = gid; i < a.column_count * 3; i += stride) a.backdrops[i] = a.backdrops_init[i];
    for (uint32_t i = gid; i <= a.column_count; i += stride) a.col_alpha_offsets[i] = 0;
}

__global__ void k_frame_reset(FrameState *f) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        FrameState z = {};
        *f = z;
    }
}

// The largest i < n with key(i) <= t (0 when there is none) for keys that do
// not decrease, searched by a whole wavefront 64 pivots at a time (three
// rounds of parallel loads for 2^18 keys instead of a chain of 18). t and n
// must be the same in every lane.
template <class Key>
__device__ __forceinline__ uint32_t wave_last_at_or_before(uint32_t n, uint32_t t, Key key) {
    uint32_t lo = 0, len = n, lane = lane_id();
    while (len > 1) {
        uint32_t step = (len + 63u) / 64u;
        uint32_t idx = lo + (lane + 1u) * step;
        bool in = idx < lo + len;
        uint32_t k = key(min(idx, n - 1u));
        uint32_t cnt = (uint32_t)__popcll(__ballot(in && k <= t));
        uint32_t end = lo + len;
        lo += cnt * step;
        len = min(step, end - lo);
    }
    return lo;
}

// ------------------------------------------------------------------ dice

struct DiceSeg {
    float2 p0, c0, c1, p3;
    uint32_t path;
    int32_t n;
    bool curve;
};

__device__ __forceinline__ float2 get_point(const BatchArgs &a, uint32_t i) {
    float2 p = a.scene.points[i];
    return make_float2((a.xform[0] * p.x + a.xform[1] * p.y) + a.xform[4],
                       (a.xform[2] * p.x + a.xform[3] * p.y) + a.xform[5]);
}

// dice.cs.glsl:130-148: the path of segment seg, by the shader's binary
// search over the paths' first batch segments (it stops at the first probe
// equal to seg, which can be an empty path).
__device__ uint32_t reference_segment_search(const BatchArgs &a, uint32_t seg) {
    uint32_t lo = 0, hi = a.path_count;
    int it = 0;
    while (it < 1024 && lo + 1 < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        uint32_t mid_seg = a.dice_md[mid].z;
        if (seg < mid_seg) {
            hi = mid;
        } else {
            lo = mid;
            if (seg == mid_seg) break;
        }
        it++;
    }
    return lo;
}

// The path of segment s0 + threadIdx.x, for a block of consecutive segments.
// Two wavefronts find the paths of the first and last segment, the block
// stages the first segments of the paths in between in LDS and each thread
// searches there. The result is the last path starting at or before the
// segment, which is what the shader's search returns unless two paths start
// at the segment (an empty one before it); those take the shader's search.
// Every thread of the block must call it.
__device__ uint32_t block_segment_path(const BatchArgs &a, uint32_t s0, uint32_t *zs,
                                       uint32_t *qrange) {
    uint32_t seg = s0 + threadIdx.x, wave = threadIdx.x >> 6;
    uint32_t s1 = min(s0 + kBlock, a.segment_count) - 1;
    auto first_seg = [&](uint32_t i) { return a.dice_md[i].z; };
    if (wave < 2) {
        uint32_t q = wave_last_at_or_before(a.path_count, wave == 0 ? s0 : s1, first_seg);
        if (lane_id() == 0) qrange[wave] = q;
    }
    __syncthreads();
    uint32_t q0 = qrange[0], q1 = qrange[1];
    bool staged = q1 - q0 < kBlock;
    if (staged)
        for (uint32_t i = threadIdx.x; i <= q1 - q0; i += blockDim.x) zs[i] = a.dice_md[q0 + i].z;
    __syncthreads();
    uint32_t path;
    if (staged) {
        uint32_t lo = 0, hi = q1 - q0 + 1;
        while (lo + 1 < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (zs[mid] <= seg) lo = mid;
            else hi = mid;
        }
        path = q0 + lo;
        if (path > 0 && zs[lo] == seg && (lo > 0 ? zs[lo - 1] : a.dice_md[path - 1].z) == seg)
            path = reference_segment_search(a, seg);
    } else {
        path = reference_segment_search(a, seg);
    }
    __syncthreads();  // zs is rewritten by the next call
    return path;
}

// dice.cs.glsl:149-202: point fetch and segment count of segment seg of path
// `path`.
__device__ bool dice_setup(const BatchArgs &a, uint32_t seg, uint32_t path, DiceSeg &d) {
    d.path = path;
    uint4 md = a.dice_md[path];
    uint32_t gseg = seg - md.z + md.y;
    if (gseg >= a.scene.index_count) return false;
    uint2 ix = a.scene.indices[gseg];
    uint32_t from = ix.x, flags = ix.y, to = from;
    if (flags & kFlagCubic) to += 3;
    else if (flags & kFlagQuad) to += 2;
    else to += 1;
    if (to >= a.scene.point_count) return false;
    d.p0 = get_point(a, from);
    d.p3 = get_point(a, to);
    d.curve = (flags & (kFlagCubic | kFlagQuad)) != 0;
    float countf;
    if (d.curve) {
        float2 k = get_point(a, from + 1);
        if (flags & kFlagQuad) {
            const float third = 1.0f / 3.0f;
            float kx2 = k.x * 2.0f, ky2 = k.y * 2.0f;
            d.c0 = make_float2((d.p0.x + kx2) * third, (d.p0.y + ky2) * third);
            d.c1 = make_float2((d.p3.x + kx2) * third, (d.p3.y + ky2) * third);
        } else {
            d.c0 = k;
            d.c1 = get_point(a, from + 2);
        }
        float ax = fabsf((d.c1.x - 2.0f * d.c0.x) + d.p0.x);
        float ay = fabsf((d.c1.y - 2.0f * d.c0.y) + d.p0.y);
        float bx = fabsf((d.p3.x - 2.0f * d.c1.x) + d.c0.x);
        float by = fabsf((d.p3.y - 2.0f * d.c1.y) + d.c0.y);
        float mx = 6.0f * max_glsl(ax, bx), my = 6.0f * max_glsl(ay, by);
        float len = sqrtf(mx * mx + my * my);
        countf = sqrtf(len / 2.0f);  // 8 * TOLERANCE
    } else {
        d.c0 = d.c1 = make_float2(0.0f, 0.0f);
        float dx = d.p3.x - d.p0.x, dy = d.p3.y - d.p0.y;
        countf = sqrtf(dx * dx + dy * dy) / 16.0f;  // MICROLINE_LENGTH
    }
    int32_t n = f2i_sat(ceilf(countf));
    d.n = n < 1 ? 1 : n;
    // A segment this long lies far outside the +-32768 px microline range.
    return d.n <= (1 << 20);
}

__global__ void __launch_bounds__(kBlock) k_dice_count(BatchArgs a) {
    __shared__ uint32_t zs[kBlock];
    __shared__ uint32_t qrange[2];
    for (uint32_t s0 = blockIdx.x * kBlock; s0 < a.segment_count; s0 += gridDim.x * kBlock) {
        uint32_t seg = s0 + threadIdx.x;
        if (a.path_count == 0) {
            if (seg < a.segment_count) {
                atomicOr(&a.counters->error, 1u);
                a.seg_offsets[seg] = 0;
            }
            continue;
        }
        uint32_t path = block_segment_path(a, s0, zs, qrange);
        if (seg >= a.segment_count) continue;
        DiceSeg d;
        uint32_t n = 0;
        if (dice_setup(a, seg, path, d)) {
            n = (uint32_t)d.n;
        } else {
            atomicOr(&a.counters->error, 1u);
        }
        a.seg_offsets[seg] = n;
    }
}

// dice.cs.glsl:76-90 emitMicroline.
__device__ __forceinline__ uint4 pack_microline(float x0, float y0, float x1, float y1,
                                                uint32_t path) {
    float v[4] = {x0, y0, x1, y1};
    int32_t px[4], fr[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int32_t sub = (int32_t)rintf(clamp_glsl(v[i], -32768.0f, 32767.0f) * 256.0f);
        px[i] = (int32_t)floorf((float)sub / 256.0f);
        fr[i] = sub - px[i] * 256;
    }
    return make_uint4(((uint32_t)px[0] & 0xffffu) | ((uint32_t)px[1] << 16),
                      ((uint32_t)px[2] & 0xffffu) | ((uint32_t)px[3] << 16),
                      (uint32_t)fr[0] | ((uint32_t)fr[1] << 8) | ((uint32_t)fr[2] << 16) |
                          ((uint32_t)fr[3] << 24),
                      path);
}

// Point at parameter t of a diced segment: nested mix as dice.cs.glsl:117-122
// (the cubic's de Casteljau steps; a line is one mix).
__device__ __forceinline__ float2 dice_eval(const float4 &pc, const float4 &pd, bool curve,
                                            float t) {
    float2 p0 = make_float2(pc.x, pc.y), c0 = make_float2(pc.z, pc.w);
    float2 c1 = make_float2(pd.x, pd.y), p3 = make_float2(pd.z, pd.w);
    if (!curve) return make_float2(mixf(p0.x, p3.x, t), mixf(p0.y, p3.y, t));
    float2 p01 = make_float2(mixf(p0.x, c0.x, t), mixf(p0.y, c0.y, t));
    float2 p12 = make_float2(mixf(c0.x, c1.x, t), mixf(c0.y, c1.y, t));
    float2 p23 = make_float2(mixf(c1.x, p3.x, t), mixf(c1.y, p3.y, t));
    float2 p012 = make_float2(mixf(p01.x, p12.x, t), mixf(p01.y, p12.y, t));
    float2 p123 = make_float2(mixf(p12.x, p23.x, t), mixf(p12.y, p23.y, t));
    return make_float2(mixf(p012.x, p123.x, t), mixf(p012.y, p123.y, t));
}

// dice.cs.glsl:204-216 for a block of 256 segments. The shader emits a
// segment's n microlines in one thread, each starting where the previous one
// ended (at t = i / n). Here the block's microlines are shared out evenly:
// thread k takes microlines k, k + 256, ... of the block, finds its segment
// among the block's output offsets in LDS and evaluates both ends, t = i / n
// (the segment's start itself for i = 0) and t = (i + 1) / n, with the same
// expressions, so every microline is the same.
__global__ void __launch_bounds__(kBlock) k_dice_emit(BatchArgs a) {
    __shared__ uint32_t zs[kBlock];
    __shared__ uint32_t qrange[2];
    __shared__ float4 s_pc[kBlock], s_pd[kBlock];  // p0, c0 | c1, p3
    __shared__ uint32_t s_first[kBlock];           // block-relative first microline
    __shared__ uint32_t s_n[kBlock];               // n | curve << 31
    __shared__ uint32_t s_path[kBlock];
    if (a.path_count == 0) return;
    for (uint32_t s0 = blockIdx.x * kBlock; s0 < a.segment_count; s0 += gridDim.x * kBlock) {
        uint32_t seg = s0 + threadIdx.x;
        uint32_t path = block_segment_path(a, s0, zs, qrange);
        uint32_t s_end = min(s0 + kBlock, a.segment_count);
        uint32_t lo = a.seg_offsets[s0], hi = a.seg_offsets[s_end];
        DiceSeg d;
        bool ok = seg < a.segment_count && dice_setup(a, seg, path, d);
        // a segment that failed the setup counted 0 microlines: an empty range
        s_first[threadIdx.x] = (seg < a.segment_count ? a.seg_offsets[seg] : hi) - lo;
        s_n[threadIdx.x] = ok ? ((uint32_t)d.n | (d.curve ? 0x80000000u : 0u)) : 0u;
        s_pc[threadIdx.x] = ok ? make_float4(d.p0.x, d.p0.y, d.c0.x, d.c0.y)
                               : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        s_pd[threadIdx.x] = ok ? make_float4(d.c1.x, d.c1.y, d.p3.x, d.p3.y)
                               : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        s_path[threadIdx.x] = ok ? d.path : 0u;
        __syncthreads();
        uint32_t total = hi - lo, last = s_end - s0 - 1;
        for (uint32_t m = threadIdx.x; m < total; m += kBlock) {
            // the last segment of the block starting at or before m (an empty
            // segment starts where the next one does, so this is never one)
            uint32_t k = 0;
#pragma unroll
            for (uint32_t step = kBlock / 2; step != 0; step >>= 1)
                if (k + step <= last && s_first[k + step] <= m) k += step;
            uint32_t nc = s_n[k];
            if (nc == 0) continue;
            int32_t n = (int32_t)(nc & 0x7fffffffu), i = (int32_t)(m - s_first[k]);
            bool curve = (nc >> 31) != 0;
            float4 pc = s_pc[k], pd = s_pd[k];
            float2 from = i == 0 ? make_float2(pc.x, pc.y)
                                 : dice_eval(pc, pd, curve, (float)i / (float)n);
            float2 to = dice_eval(pc, pd, curve, (float)(i + 1) / (float)n);
            uint32_t o = lo + m;
            if (o < a.microline_capacity)
                a.microlines[o] = pack_microline(from.x, from.y, to.x, to.y, s_path[k]);
        }
        __syncthreads();  // the next block of segments rewrites the LDS
    }
}

// ------------------------------------------------------------------ bound

// The path search of bound.cs.glsl:53-66: binary search of the tile's path
// by first tile index, stopping early when a probed path starts at the tile.
__device__ __forceinline__ uint32_t reference_tile_search(const BatchArgs &a, uint32_t t) {
    uint32_t lo = 0, hi = a.path_count;
    int it = 0;
    while (it < 1024 && lo + 1 < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        uint32_t mid_tile = a.tile_path_info[mid].z;
        if (t < mid_tile) {
            hi = mid;
        } else {
            lo = mid;
            if (t == mid_tile) break;
        }
        it++;
    }
    return lo;
}

// bound.cs.glsl:49-84. The reference initialises every TileD3D11 here
// (next and first fill ~0, alpha -1 with a zero backdrop delta, the path's
// colour and control word). Here the only state bin and propagate need
// before propagate writes the tiles is cleared: the per-tile fill flag and
// backdrop delta word (bin adds delta << 24 to it, as the shader does to
// tile word 2). Propagate then writes all four words of every tile with the
// values bound and bin would have left plus its own.
// With the adjoint on, bound also records each tile's path: a workgroup
// takes 1024 consecutive tiles, stages the first tile indices of the paths
// they belong to in LDS and looks each tile up there. The reference search
// returns the same path, the last one starting at or before the tile,
// except when it lands on an empty path that starts at the tile itself; that
// case takes the reference search.
constexpr uint32_t kBoundItems = 4;

__global__ void __launch_bounds__(kBlock) k_bound(BatchArgs a) {
    constexpr uint32_t kPer = kBlock * kBoundItems;
    __shared__ uint32_t zs[kPer + 1];
    __shared__ uint32_t qrange[2];
    {
        uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, n = gridDim.x * blockDim.x;
        uint4 *d4 = reinterpret_cast<uint4 *>(a.tile_delta);
        for (uint32_t i = gid; i < a.tile_count / 4u; i += n) d4[i] = make_uint4(0, 0, 0, 0);
        uint4 *f4 = reinterpret_cast<uint4 *>(a.tile_flags);
        for (uint32_t i = gid; i < a.tile_count / 16u; i += n) f4[i] = make_uint4(0, 0, 0, 0);
        if (gid < 16u) {
            uint32_t t = (a.tile_count & ~15u) + gid;
            if (t < a.tile_count) a.tile_flags[t] = 0;
            if (gid < 4u && (a.tile_count & ~3u) + gid < a.tile_count)
                a.tile_delta[(a.tile_count & ~3u) + gid] = 0;
        }
    }
    if (a.path_count == 0 || !a.tile_paths) return;
    uint32_t wave = threadIdx.x >> 6;
    auto first_tile = [&](uint32_t i) { return a.tile_path_info[i].z; };
    for (uint32_t b0 = blockIdx.x * kPer; b0 < a.tile_count; b0 += gridDim.x * kPer) {
        uint32_t b1 = min(b0 + kPer, a.tile_count) - 1;
        // wavefront 0 finds the path of the first tile, wavefront 1 that of the last
        if (wave < 2) {
            uint32_t q = wave_last_at_or_before(a.path_count, wave == 0 ? b0 : b1, first_tile);
            if (lane_id() == 0) qrange[wave] = q;
        }
        __syncthreads();
        uint32_t q0 = qrange[0], q1 = qrange[1];
        bool staged = q1 - q0 < kPer;
        if (staged)
            for (uint32_t i = threadIdx.x; i <= q1 - q0; i += blockDim.x) {
                zs[i] = a.tile_path_info[q0 + i].z;
            }
        __syncthreads();
        for (uint32_t k = 0; k < kBoundItems; k++) {
            uint32_t t = b0 + k * kBlock + threadIdx.x;
            if (t > b1) break;
            uint32_t p;
            if (staged) {
                uint32_t lo = 0, hi = q1 - q0 + 1;
                while (lo + 1 < hi) {
                    uint32_t mid = (lo + hi) >> 1;
                    if (zs[mid] <= t) lo = mid;
                    else hi = mid;
                }
                p = q0 + lo;
                if (p > 0 && zs[lo] == t &&
                    (lo > 0 ? zs[lo - 1] : a.tile_path_info[p - 1].z) == t)
                    p = reference_tile_search(a, t);
            } else {
                p = reference_tile_search(a, t);
            }
            a.tile_paths[t] = p;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ bin

struct IRect {
    int32_t x0, y0, x1, y1;
};

// One fill of bin.cs.glsl:98-121 addFill: returns false when culled.
__device__ __forceinline__ bool make_fill(float x0, float y0, float x1, float y1, int32_t tx,
                                          int32_t ty, const IRect &r, uint32_t off,
                                          uint32_t tile_count, uint32_t *tile, uint2 *words) {
    if (tx < r.x0 || ty < r.y0 || tx >= r.x1 || ty >= r.y1) return false;
    *tile = off + (uint32_t)(tx - r.x0) + (uint32_t)(ty - r.y0) * (uint32_t)(r.x1 - r.x0);
    if (*tile >= tile_count) return false;  // metadata inconsistent with the tile count
    float ox = (float)(tx * 16), oy = (float)(ty * 16);
    uint32_t sx0 = f2u_sat((x0 - ox) * 256.0f), sy0 = f2u_sat((y0 - oy) * 256.0f);
    uint32_t sx1 = f2u_sat((x1 - ox) * 256.0f), sy1 = f2u_sat((y1 - oy) * 256.0f);
    if (sx0 == sx1) return false;
    *words = make_uint2(sx0 | (sy0 << 16), sx1 | (sy1 << 16));
    return true;
}

// bin.cs.glsl:154-256 for one microline. EMIT = false counts the fills;
// EMIT = true writes them from `fill_base` on and applies backdrop deltas.
template <bool EMIT>
__device__ uint32_t bin_microline(const BatchArgs &a, uint4 ml, uint32_t fill_base) {
    uint32_t path = ml.w;
    int32_t w0 = (int32_t)ml.x, w1 = (int32_t)ml.y, w2 = (int32_t)ml.z;
    float lx0 = (float)((int32_t)((uint32_t)w0 << 16) >> 16) + (float)(w2 & 0xff) / 256.0f;
    float ly0 = (float)(w0 >> 16) + (float)((w2 >> 8) & 0xff) / 256.0f;
    float lx1 = (float)((int32_t)((uint32_t)w1 << 16) >> 16) + (float)((w2 >> 16) & 0xff) / 256.0f;
    float ly1 = (float)(w1 >> 16) + (float)((w2 >> 24) & 0xff) / 256.0f;
    if (path >= a.path_count) return 0;
    const int32_t *md = a.prop_md + (size_t)path * 12;
    IRect r = {md[0], md[1], md[2], md[3]};
    uint32_t off = (uint32_t)md[4], boff = (uint32_t)md[8];

    int32_t ftx = (int32_t)floorf(lx0 / 16.0f), fty = (int32_t)floorf(ly0 / 16.0f);
    int32_t ttx = (int32_t)floorf(lx1 / 16.0f), tty = (int32_t)floorf(ly1 / 16.0f);
    float vx = lx1 - lx0, vy = ly1 - ly0;
    int32_t sx = vx < 0.0f ? -1 : 1, sy = vy < 0.0f ? -1 : 1;
    float fcx = (float)((ftx + (vx >= 0.0f ? 1 : 0)) * 16);
    float fcy = (float)((fty + (vy >= 0.0f ? 1 : 0)) * 16);
    float tmx = (fcx - lx0) / vx, tmy = (fcy - ly0) / vy;
    float tdx = fabsf(16.0f / vx), tdy = fabsf(16.0f / vy);
    float cx = lx0, cy = ly0;
    int32_t tx = ftx, ty = fty;
    int last = 0;
    uint32_t count = 0;
    for (uint32_t it = 0; it < 1024u; it++) {
        int next;
        if (tmx < tmy) next = 1;
        else if (tmx > tmy) next = 2;
        else if (sx > 0) next = 1;
        else next = 2;
        float nt = min_glsl(next == 1 ? tmx : tmy, 1.0f);
        if (tx == ttx && ty == tty) next = 0;
        float nx = mixf(lx0, lx1, nt), ny = mixf(ly0, ly1, nt);
        uint32_t tile;
        uint2 fw;
        if (make_fill(cx, cy, nx, ny, tx, ty, r, off, a.tile_count, &tile, &fw)) {
            if (EMIT) {
                uint32_t f = fill_base + count;
                if (f < a.fill_capacity) {
                    a.fill_words[f] = fw;
                    a.fill_tiles[f] = tile;
                    a.tile_flags[tile] = 1;
                }
            }
            count++;
        }
        bool aux = false;
        float ax0 = 0, ay0 = 0, ax1 = 0, ay1 = 0;
        if (sy < 0 && next == 2) {
            aux = true;
            ax0 = nx; ay0 = ny; ax1 = (float)(tx * 16); ay1 = (float)(ty * 16);
        } else if (sy > 0 && last == 2) {
            aux = true;
            ax0 = (float)(tx * 16); ay0 = (float)(ty * 16); ax1 = cx; ay1 = cy;
        }
        if (aux && make_fill(ax0, ay0, ax1, ay1, tx, ty, r, off, a.tile_count, &tile, &fw)) {
            if (EMIT) {
                uint32_t f = fill_base + count;
                if (f < a.fill_capacity) {
                    a.fill_words[f] = fw;
                    a.fill_tiles[f] = tile;
                    a.tile_flags[tile] = 1;
                }
            }
            count++;
        }
        if (EMIT) {
            int32_t delta = 0;
            if (sx < 0 && last == 1) delta = 1;
            else if (sx > 0 && next == 1) delta = -1;
            if (delta != 0) {
                // bin.cs.glsl:123-143 adjustBackdrop
                bool ox = tx < r.x0, oy = ty < r.y0, oz = tx >= r.x1, ow = ty >= r.y1;
                if (ox || oy || oz || ow) {
                    if (!ox && oy && !oz) {
                        uint32_t bi = boff + (uint32_t)(tx - r.x0);
                        if (bi < a.column_count) atomicAdd(&a.backdrops[(size_t)bi * 3], delta);
                    }
                } else {
                    uint32_t t = off + (uint32_t)(tx - r.x0) +
                                 (uint32_t)(ty - r.y0) * (uint32_t)(r.x1 - r.x0);
                    if (t < a.tile_count) atomicAdd(&a.tile_delta[t], (uint32_t)delta << 24);
                }
            }
        }
        if (next == 1) {
            tmx += tdx;
            tx += sx;
        } else if (next == 2) {
            tmy += tdy;
            ty += sy;
        } else {
            break;
        }
        cx = nx; cy = ny;
        last = next;
    }
    return count;
}

__global__ void __launch_bounds__(kBlock) k_bin_count(BatchArgs a) {
    uint32_t n = a.counters->microline_count;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < n;
         m += gridDim.x * blockDim.x)
        a.ml_fill_offsets[m] = bin_microline<false>(a, a.microlines[m], 0);
}

__global__ void __launch_bounds__(kBlock) k_bin_emit(BatchArgs a) {
    uint32_t n = a.counters->microline_count;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < n;
         m += gridDim.x * blockDim.x)
        bin_microline<true>(a, a.microlines[m], a.ml_fill_offsets[m]);
}

// First fill of each path (path_count + 1 entries, the last one the fill
// count): the fills of a path are those of its segments' microlines.
__global__ void __launch_bounds__(kBlock) k_path_fills(BatchArgs a) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > a.path_count) return;
    uint32_t mcount = a.counters->microline_count, fcount = a.counters->fill_count;
    uint32_t seg = p < a.path_count ? a.dice_md[p].z : a.segment_count;
    if (seg > a.segment_count) seg = a.segment_count;
    uint32_t m = min(a.seg_offsets[seg], mcount);
    a.path_fills[p] = min(a.ml_fill_offsets[m], fcount);
}

// Per-tile fill lists (bin.cs.glsl:113-117): the reference pushes each fill
// onto its tile's list with atomicExchange, so in sequential order fill f
// links to the previous fill of the same tile and the tile heads the last
// one. A path's fills are contiguous and only touch the path's own tiles, so
// one workgroup per path orders its fills by (tile, fill index) chunk by
// chunk and stitches the chunks. Bin emits a microline's fills tile by tile,
// so consecutive fills often share a tile: a chunk is cut into such runs and
// only the runs are sorted, by (tile, run index); a run keeps its fills in
// index order. No other workgroup touches the path's tiles and the chunks
// run in order, so the tile head and last run are read and written with
// plain loads and stores, ordered by the barrier between chunks. The sorted
// order is also kept: sorted_fills holds each tile's fills of a chunk
// contiguously, tile_runs[t] the tile's last such range (first index, count;
// bit 31 of the count set when an earlier range exists) and run_prev[first
// index of a range] the range before it.

__device__ __forceinline__ unsigned long long shfl_xor64(unsigned long long v, uint32_t m) {
    uint32_t lo = __shfl_xor((uint32_t)v, (int)m, 64);
    uint32_t hi = __shfl_xor((uint32_t)(v >> 32), (int)m, 64);
    return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t shfl_xor_key(uint32_t v, uint32_t m) {
    return __shfl_xor(v, (int)m, 64);
}
__device__ __forceinline__ unsigned long long shfl_xor_key(unsigned long long v, uint32_t m) {
    return shfl_xor64(v, m);
}

// Sort keys: (tile - path's first tile, run index). 32 bits when the path's
// tile rect has at most 2^(32 - kLinkRunBits) tiles (chunks of 2^kLinkRunBits
// fills), 64 bits otherwise (chunks of half that, so the LDS is the same).
constexpr uint32_t kLinkRunBits = 10;
constexpr uint32_t kLinkFills = 1u << kLinkRunBits;         // fills per 32-bit chunk
constexpr uint32_t kLinkRect32 = 1u << (32 - kLinkRunBits);  // largest rect for 32-bit keys
template <typename Key> struct LinkKey;
template <> struct LinkKey<uint32_t> {
    static constexpr uint32_t kChunk = kLinkFills;
    __device__ static uint32_t make(uint32_t tile, uint32_t run) {
        return tile << kLinkRunBits | run;
    }
    __device__ static uint32_t tile(uint32_t k) { return k >> kLinkRunBits; }
    __device__ static uint32_t run(uint32_t k) { return k & (kLinkFills - 1u); }
};
template <> struct LinkKey<unsigned long long> {
    static constexpr uint32_t kChunk = kLinkFills / 2;
    __device__ static unsigned long long make(uint32_t tile, uint32_t run) {
        return (unsigned long long)tile << 32 | run;
    }
    __device__ static uint32_t tile(unsigned long long k) { return (uint32_t)(k >> 32); }
    __device__ static uint32_t run(unsigned long long k) { return (uint32_t)k; }
};

struct LinkLds {
    unsigned long long keys[2][kLinkFills / 2];  // sort buffers (u32 or u64 keys)
    uint16_t run_first[kLinkFills + 2];          // first fill of each run; [runs] = n
    uint16_t rank[kLinkFills];                   // sorted position of each run
    uint16_t out[kLinkFills];                    // first output slot of each sorted run
    uint32_t starts[kLinkFills / 64];            // run starts per (element row, wave)
    uint32_t scan[4];
};

// The rest of a chunk once its sorted run keys are in sk: output slots, tile
// ranges and run-start links, every fill to its slot, and the barrier before
// the next chunk.
template <typename Key, uint32_t kPer>
__device__ __forceinline__ void link_finish(const BatchArgs &a, LinkLds &L, const Key *sk,
                                            uint32_t base, uint32_t n, uint32_t runs,
                                            uint32_t tile_base, uint32_t f0, uint32_t start_bits,
                                            const uint32_t (&rid)[kPer],
                                            const uint2 (&fw)[kPer]) {
    using K = LinkKey<Key>;
    uint32_t tid = threadIdx.x;

    // Output slots: exclusive scan of the run lengths in sorted order
    // (thread tid holds sorted runs tid * kPer .. + kPer - 1).
    uint32_t p0 = tid * kPer;
    uint32_t len[kPer];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
        uint32_t p = p0 + k;
        len[k] = 0;
        if (p < runs) {
            uint32_t rr = K::run(sk[p]);
            len[k] = (uint32_t)L.run_first[rr + 1] - L.run_first[rr];
            L.rank[rr] = (uint16_t)p;
        }
        sum += len[k];
    }
    uint32_t total;
    uint32_t o = block_exclusive_scan(sum, L.scan, &total);
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
        if (p0 + k < runs) L.out[p0 + k] = (uint16_t)o;
        o += len[k];
    }
    __syncthreads();

    // Tile ranges and the links of run starts, one sorted run per thread.
    // k_bound leaves every fill list empty, so in the path's first chunk a
    // tile has no earlier fill and no earlier range.
    bool first_chunk = base == f0;
    for (uint32_t p = tid; p < runs; p += kBlock) {
        Key cur = sk[p];
        uint32_t lt = K::tile(cur), rr = K::run(cur);
        uint32_t first = L.run_first[rr];
        uint32_t link;
        if (p == 0 || K::tile(sk[p - 1]) != lt) {
            // last sorted run of this tile: galloping, then binary search
            uint32_t lo = p, hi = runs, step = 1;
            for (;;) {
                uint32_t nx = lo + step;
                if (nx >= runs) break;
                if (K::tile(sk[nx]) != lt) {
                    hi = nx;
                    break;
                }
                lo = nx;
                step <<= 1;
            }
            while (lo + 1 < hi) {
                uint32_t mid = (lo + hi) >> 1;
                if (K::tile(sk[mid]) == lt) lo = mid;
                else hi = mid;
            }
            uint32_t rq = K::run(sk[lo]);
            uint32_t end = L.run_first[rq + 1];  // one past the tile's last fill
            uint32_t tile = tile_base + lt;
            link = 0xffffffffu;
            uint2 prev = make_uint2(0, 0);
            // flag 2: an earlier chunk of this path gave the tile a head
            // (bin set the flag to 1; tiles are not initialised otherwise)
            if (!first_chunk && a.tile_flags[tile] == 2) {
                link = a.tile_first[tile];
                prev = a.tile_runs[tile];
            }
            a.tile_first[tile] = base + end - 1;
            a.tile_flags[tile] = 2;
            uint32_t slot = L.out[p];
            uint32_t count = (uint32_t)L.out[lo] + (end - L.run_first[rq]) - slot;
            // bit 31 of the count: an earlier range of this tile exists, and
            // only then is run_prev of this range read
            a.tile_runs[tile] =
                make_uint2(base + slot, count | (link != 0xffffffffu ? 0x80000000u : 0u));
            if (link != 0xffffffffu) a.run_prev[base + slot] = prev;
        } else {
            link = base + L.run_first[K::run(sk[p - 1]) + 1] - 1;
        }
        a.fill_links[base + first] = link;
    }
    // Every fill to its slot; a fill inside a run links to the one before.
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
        uint32_t i = k * kBlock + tid;
        if (i < n) {
            uint32_t slot = (uint32_t)L.out[L.rank[rid[k]]] + (i - L.run_first[rid[k]]);
            a.sorted_fills[base + slot] = fw[k];
            if (!(start_bits >> k & 1u)) a.fill_links[base + i] = base + i - 1;
        }
    }
    __syncthreads();
}

template <typename Key>
__device__ void link_path(const BatchArgs &a, LinkLds &L, uint32_t f0, uint32_t f1,
                          uint32_t tile_base) {
    using K = LinkKey<Key>;
    constexpr uint32_t kChunk = K::kChunk;
    constexpr uint32_t kPer = kChunk / kBlock;  // fills (and keys) per thread
    Key *keys1 = reinterpret_cast<Key *>(L.keys[1]);
    const Key kPad = ~(Key)0;
    uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t base = f0; base < f1; base += kChunk) {
        uint32_t n = min(kChunk, f1 - base);
        // Fill i = k * kBlock + tid is element k of this thread. Its tile, the
        // tile of fill i - 1 and its words are loaded up front, unconditionally
        // at clamped indices, so all loads are in flight at once.
        uint32_t tl[kPer], tp[kPer];
        uint2 fw[kPer];
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            uint32_t i = min(k * kBlock + tid, n - 1);
            tl[k] = a.fill_tiles[base + i];
            tp[k] = a.fill_tiles[base + (i > 0 ? i - 1 : 0)];
            fw[k] = a.fill_words[base + i];
        }
        // Runs: a run starts where the tile changes. Its index is the number
        // of starts before it, counted by ballots per (element row, wave).
        uint32_t start_bits = 0, le[kPer];
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            uint32_t i = k * kBlock + tid;
            bool st = i < n && (i == 0 || tl[k] != tp[k]);
            unsigned long long m = __ballot(st);
            le[k] = (uint32_t)__popcll(m & below) + (st ? 1u : 0u);
            start_bits |= (uint32_t)st << k;
            if (lane == 0) L.starts[k * (kBlock / 64) + wave] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        uint32_t runs = 0, rid[kPer];
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
#pragma unroll
            for (uint32_t w = 0; w < kBlock / 64; w++) {
                uint32_t c = L.starts[k * (kBlock / 64) + w];
                if (w == wave) rid[k] = runs + le[k] - 1;  // the run of this fill
                runs += c;
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            if (start_bits >> k & 1u) {
                L.run_first[rid[k]] = (uint16_t)(k * kBlock + tid);
                keys1[rid[k]] = K::make(tl[k] - tile_base, rid[k]);
            }
        }
        if (tid == 0) L.run_first[runs] = (uint16_t)n;
        __syncthreads();

        // Bitonic sort of the runs' keys, padded to size. Up to 256 runs,
        // wavefront 0 sorts them alone in registers (element q * 64 + lane of
        // register q: lane shuffles under 64 apart, register swaps above) and
        // the other waves wait at one barrier. Otherwise key e lives in
        // register q of lane l of wave w, e = q * 64W + w * 64 + l with
        // W = min(size / 64, 4) waves: partners at distance < 64 are exchanged
        // across lanes, at distance >= 64W within the thread, and the
        // distances in between through LDS (double-buffered, one barrier per
        // step). The input is in keys1 and the first LDS step writes keys0.
        uint32_t size = 64;
        while (size < runs) size <<= 1;
        if (size <= 256) {
            if (wave == 0) {
                uint32_t nq = size >> 6;
                Key w4[4];
#pragma unroll
                for (uint32_t q = 0; q < 4; q++) {
                    uint32_t e = q * 64u + lane;
                    w4[q] = (q < nq && e < runs) ? keys1[e] : kPad;
                }
                for (uint32_t k = 2; k <= size; k <<= 1) {
                    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                        if (j < 64u) {
#pragma unroll
                            for (uint32_t q = 0; q < 4; q++) {
                                if (q < nq) {
                                    Key o = shfl_xor_key(w4[q], j);
                                    uint32_t e = q * 64u + lane;
                                    bool keep_min = ((lane & j) == 0) == ((e & k) == 0);
                                    w4[q] = keep_min ? min(w4[q], o) : max(w4[q], o);
                                }
                            }
                        } else {
                            uint32_t rj = j >> 6;
#pragma unroll
                            for (uint32_t q = 0; q < 4; q++) {
                                if ((q & rj) == 0 && (q | rj) < nq) {
                                    uint32_t e = q * 64u + lane;
                                    Key x = w4[q], y = w4[q | rj];
                                    bool up = (e & k) == 0;
                                    w4[q] = up ? min(x, y) : max(x, y);
                                    w4[q | rj] = up ? max(x, y) : min(x, y);
                                }
                            }
                        }
                    }
                }
                Key *out0 = reinterpret_cast<Key *>(L.keys[0]);
#pragma unroll
                for (uint32_t q = 0; q < 4; q++)
                    if (q < nq) out0[q * 64u + lane] = w4[q];
            }
            __syncthreads();
            link_finish<Key>(a, L, reinterpret_cast<Key *>(L.keys[0]), base, n, runs, tile_base,
                             f0, start_bits, rid, fw);
            continue;
        }
        uint32_t lw = size >= 256 ? 2u : (size == 128 ? 1u : 0u);
        uint32_t nw = 1u << lw, span = 64u << lw, nr = size >> (6 + lw);
        bool active = wave < nw;
        Key v[kPer];
#pragma unroll
        for (uint32_t q = 0; q < kPer; q++) {
            uint32_t e = q * span + wave * 64u + lane;
            v[q] = (active && q < nr && e < runs) ? keys1[e] : kPad;
        }
        uint32_t buf = 0;
        for (uint32_t k = 2; k <= size; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                if (j < 64u) {
                    if (active) {
#pragma unroll
                        for (uint32_t q = 0; q < kPer; q++) {
                            if (q < nr) {
                                Key o = shfl_xor_key(v[q], j);
                                uint32_t e = q * span + wave * 64u + lane;
                                bool keep_min = ((lane & j) == 0) == ((e & k) == 0);
                                v[q] = keep_min ? min(v[q], o) : max(v[q], o);
                            }
                        }
                    }
                } else if (j < span) {
                    Key *kb = reinterpret_cast<Key *>(L.keys[buf]);
                    if (active) {
#pragma unroll
                        for (uint32_t q = 0; q < kPer; q++)
                            if (q < nr) kb[q * span + wave * 64u + lane] = v[q];
                    }
                    __syncthreads();
                    if (active) {
#pragma unroll
                        for (uint32_t q = 0; q < kPer; q++) {
                            if (q < nr) {
                                uint32_t e = q * span + wave * 64u + lane;
                                Key o = kb[e ^ j];
                                bool keep_min = ((e & j) == 0) == ((e & k) == 0);
                                v[q] = keep_min ? min(v[q], o) : max(v[q], o);
                            }
                        }
                    }
                    buf ^= 1u;
                } else {
                    uint32_t rj = j / span;
                    if (active) {
#pragma unroll
                        for (uint32_t q = 0; q < kPer; q++) {
                            if ((q & rj) == 0 && (q | rj) < nr) {
                                uint32_t e = q * span + wave * 64u + lane;
                                Key x = v[q], y = v[q | rj];
                                bool up = (e & k) == 0;
                                v[q] = up ? min(x, y) : max(x, y);
                                v[q | rj] = up ? max(x, y) : min(x, y);
                            }
                        }
                    }
                }
            }
        }
        // The last LDS step read keys[buf ^ 1] (keys1 holds the input when
        // there was none); the sorted keys go to keys[buf].
        Key *sk = reinterpret_cast<Key *>(L.keys[buf]);
        if (active) {
#pragma unroll
            for (uint32_t q = 0; q < kPer; q++)
                if (q < nr) sk[q * span + wave * 64u + lane] = v[q];
        }
        __syncthreads();
        link_finish<Key>(a, L, sk, base, n, runs, tile_base, f0, start_bits, rid, fw);
    }
}

// Paths of at most kLinkWaveFills fills whose tile rect has at most 2^24
// tiles are linked by link_small_paths, a wavefront each; the others by
// link_large_path, a workgroup each.
constexpr uint32_t kLinkWaveFills = 256;
__device__ __forceinline__ bool link_by_wave(uint32_t n, uint64_t rect) {
    return n <= kLinkWaveFills && rect <= (1ull << 24);
}

__device__ void link_large_path(const BatchArgs &a, LinkLds &lds, uint32_t path) {
    if (path >= a.path_count) return;
    uint32_t f0 = a.path_fills[path], f1 = max(a.path_fills[path + 1], f0);
    if (f0 == f1) return;
    const int32_t *pm = a.prop_md + (size_t)path * 12;
    uint32_t width = (uint32_t)(pm[2] - pm[0]), height = (uint32_t)(pm[3] - pm[1]);
    uint64_t rect = (uint64_t)width * height;
    if (link_by_wave(f1 - f0, rect)) return;
    if (rect <= kLinkRect32)
        link_path<uint32_t>(a, lds, f0, f1, (uint32_t)pm[4]);
    else
        link_path<unsigned long long>(a, lds, f0, f1, (uint32_t)pm[4]);
}

// The fill lists of one path of at most 256 fills, a wavefront per path (four
// per workgroup, no workgroup barriers): the same runs, (tile, run) sort,
// output slots and links as link_path for a single chunk, with the run keys
// (local tile << 8 | run) sorted in registers, four per lane (element
// q * 64 + lane): lane shuffles for partners under 64 apart, register swaps
// above. k_bound leaves every fill list empty, so no tile has an earlier
// fill or range.
struct LinkWaveLds {
    uint32_t keys[kLinkWaveFills];
    uint16_t run_first[kLinkWaveFills + 2];
    uint16_t rank[kLinkWaveFills];
    uint16_t out[kLinkWaveFills];
};

__device__ __forceinline__ void link_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ void link_small_paths(const BatchArgs &a, LinkWaveLds *lds, uint32_t block) {
    constexpr uint32_t kQ = kLinkWaveFills / 64;
    uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    uint32_t path = block * (kBlock / 64) + wave;
    if (path >= a.path_count) return;
    uint32_t f0 = a.path_fills[path], f1 = max(a.path_fills[path + 1], f0);
    uint32_t n = f1 - f0;
    const int32_t *pm = a.prop_md + (size_t)path * 12;
    uint32_t width = (uint32_t)(pm[2] - pm[0]), height = (uint32_t)(pm[3] - pm[1]);
    if (n == 0 || !link_by_wave(n, (uint64_t)width * height)) return;
    uint32_t tile_base = (uint32_t)pm[4];
    LinkWaveLds &L = lds[wave];
    unsigned long long below = (1ull << lane) - 1ull;
    // fill i = k * 64 + lane: its tile, the tile of fill i - 1, its words
    uint32_t tl[kQ], tp[kQ];
    uint2 fw[kQ];
#pragma unroll
    for (uint32_t k = 0; k < kQ; k++) {
        uint32_t i = min(k * 64u + lane, n - 1);
        tl[k] = a.fill_tiles[f0 + i];
        tp[k] = a.fill_tiles[f0 + (i > 0 ? i - 1 : 0)];
        fw[k] = a.fill_words[f0 + i];
    }
    uint32_t runs = 0, rid[kQ], start_bits = 0;
#pragma unroll
    for (uint32_t k = 0; k < kQ; k++) {
        uint32_t i = k * 64u + lane;
        bool st = i < n && (i == 0 || tl[k] != tp[k]);
        unsigned long long m = __ballot(st);
        rid[k] = runs + (uint32_t)__popcll(m & below) + (st ? 1u : 0u) - 1u;
        runs += (uint32_t)__popcll(m);
        start_bits |= (uint32_t)st << k;
    }
#pragma unroll
    for (uint32_t k = 0; k < kQ; k++) {
        if (start_bits >> k & 1u) {
            L.run_first[rid[k]] = (uint16_t)(k * 64u + lane);
            L.keys[rid[k]] = (tl[k] - tile_base) << 8 | rid[k];
        }
    }
    if (lane == 0) L.run_first[runs] = (uint16_t)n;
    link_wave_sync();
    uint32_t size = 64;
    while (size < runs) size <<= 1;
    uint32_t nq = size >> 6;
    uint32_t v[kQ];
#pragma unroll
    for (uint32_t q = 0; q < kQ; q++) {
        uint32_t e = q * 64u + lane;
        v[q] = (q < nq && e < runs) ? L.keys[e] : 0xffffffffu;
    }
    for (uint32_t k = 2; k <= size; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            if (j < 64u) {
#pragma unroll
                for (uint32_t q = 0; q < kQ; q++) {
                    if (q < nq) {
                        uint32_t o = __shfl_xor(v[q], (int)j, 64);
                        uint32_t e = q * 64u + lane;
                        bool keep_min = ((lane & j) == 0) == ((e & k) == 0);
                        v[q] = keep_min ? min(v[q], o) : max(v[q], o);
                    }
                }
            } else {
                uint32_t rj = j >> 6;
#pragma unroll
                for (uint32_t q = 0; q < kQ; q++) {
                    if ((q & rj) == 0 && (q | rj) < nq) {
                        uint32_t e = q * 64u + lane;
                        uint32_t x = v[q], y = v[q | rj];
                        bool up = (e & k) == 0;
                        v[q] = up ? min(x, y) : max(x, y);
                        v[q | rj] = up ? max(x, y) : min(x, y);
                    }
                }
            }
        }
    }
    link_wave_sync();  // every lane has read the unsorted keys
#pragma unroll
    for (uint32_t q = 0; q < kQ; q++)
        if (q < nq) L.keys[q * 64u + lane] = v[q];
    link_wave_sync();
    // output slots: exclusive scan of the run lengths in sorted order
    uint32_t carry = 0;
#pragma unroll
    for (uint32_t q = 0; q < kQ; q++) {
        if (q < nq) {
            uint32_t p = q * 64u + lane;
            uint32_t len = 0;
            if (p < runs) {
                uint32_t r = v[q] & 0xffu;
                len = (uint32_t)L.run_first[r + 1] - L.run_first[r];
                L.rank[r] = (uint16_t)p;
            }
            uint32_t x = wave_inclusive_scan(len);
            if (p < runs) L.out[p] = (uint16_t)(carry + x - len);
            carry += last_lane(x);
        }
    }
    link_wave_sync();
    const uint32_t *sk = L.keys;
    // tile ranges and the links of run starts, one sorted run per lane
#pragma unroll
    for (uint32_t q = 0; q < kQ; q++) {
        uint32_t p = q * 64u + lane;
        if (q >= nq || p >= runs) continue;
        uint32_t cur = sk[p];
        uint32_t lt = cur >> 8, rr = cur & 0xffu;
        uint32_t first = L.run_first[rr];
        uint32_t link;
        if (p == 0 || (sk[p - 1] >> 8) != lt) {
            uint32_t lo = p, hi = runs;  // last sorted run of this tile
            while (lo + 1 < hi) {
                uint32_t mid = (lo + hi) >> 1;
                if ((sk[mid] >> 8) == lt) lo = mid;
                else hi = mid;
            }
            uint32_t rq = sk[lo] & 0xffu;
            uint32_t end = L.run_first[rq + 1];
            uint32_t tile = tile_base + lt;
            link = 0xffffffffu;
            a.tile_first[tile] = f0 + end - 1;
            uint32_t slot = L.out[p];
            uint32_t count = (uint32_t)L.out[lo] + (end - L.run_first[rq]) - slot;
            a.tile_runs[tile] = make_uint2(f0 + slot, count);
        } else {
            link = f0 + L.run_first[(sk[p - 1] & 0xffu) + 1] - 1;
        }
        a.fill_links[f0 + first] = link;
    }
    // every fill to its slot; a fill inside a run links to the one before
#pragma unroll
    for (uint32_t k = 0; k < kQ; k++) {
        uint32_t i = k * 64u + lane;
        if (i < n) {
            uint32_t slot = (uint32_t)L.out[L.rank[rid[k]]] + (i - L.run_first[rid[k]]);
            a.sorted_fills[f0 + slot] = fw[k];
            if (!(start_bits >> k & 1u)) a.fill_links[f0 + i] = f0 + i - 1;
        }
    }
}

// Every path's fill lists in one launch: workgroup p < path_count links path
// p if it is too large for a wavefront (link_path), the workgroups after them
// link the small paths four at a time (link_small_paths), so both kinds run
// side by side.
__global__ void __launch_bounds__(kBlock) k_fill_link(BatchArgs a) {
    __shared__ union {
        LinkLds large;
        LinkWaveLds small[kBlock / 64];
    } lds;
    if (blockIdx.x < a.path_count)
        link_large_path(a, lds.large, blockIdx.x);
    else
        link_small_paths(a, lds.small, blockIdx.x - a.path_count);
}

// ------------------------------------------------------------------ propagate

constexpr uint32_t kPropRows = 8;   // tile rows a segment walk fetches at a time
constexpr uint32_t kPropSeg = kPropSegmentRows;

// propagate.cs.glsl:103-218 for rows [y_begin, y_begin + kPropSeg) of one
// column, entered with the running backdrop of row y_begin. EMIT = false
// counts the alpha tiles the rows allocate and sums their backdrop deltas
// into *delta_sum; EMIT = true writes the tile words and alpha tiles from
// `alpha_base` (batch-local) on. Whether a tile needs an alpha tile does not
// depend on the running backdrop, so segments of a column are counted
// independently. Linking and z-buffer writes are done by k_sort_tiles.
template <bool EMIT>
__device__ uint32_t propagate_segment(const BatchArgs &a, uint32_t col, uint32_t y_begin,
                                      int32_t backdrop, uint32_t alpha_base, uint32_t first_alpha,
                                      int32_t *delta_sum) {
    int32_t tile_x = a.backdrops[(size_t)col * 3 + 1];
    uint32_t dpath = (uint32_t)a.backdrops[(size_t)col * 3 + 2];
    if (dpath >= a.path_count) {
        atomicOr(&a.counters->error, 2u);
        return 0;
    }
    const int32_t *md = a.prop_md + (size_t)dpath * 12;
    uint32_t rx0 = (uint32_t)md[0], ry0 = (uint32_t)md[1], rx1 = (uint32_t)md[2],
             ry1 = (uint32_t)md[3];
    uint32_t height = ry1 - ry0, width = rx1 - rx0;
    uint32_t dofs = (uint32_t)md[4];
    int32_t cpi = md[7];
    bool zwrite = md[6] != 0;
    uint32_t crx0 = 0, cry0 = 0, crx1 = 0, cry1 = 0, cofs = 0;
    bool clip = cpi >= 0 && a.clip_prop_md != nullptr && (uint32_t)cpi < a.clip_path_count;
    if (cpi >= 0 && !clip) atomicOr(&a.counters->error, 2u);
    if (clip) {
        const int32_t *cm = a.clip_prop_md + (size_t)cpi * 12;
        crx0 = (uint32_t)cm[0]; cry0 = (uint32_t)cm[1];
        crx1 = (uint32_t)cm[2]; cry1 = (uint32_t)cm[3];
        cofs = (uint32_t)cm[4];
    }
    bool cload = clip && a.clip_tile_count != 0;
    // tile word 3 as bound leaves it: the colour and control bits of the
    // path bound's search finds, the column's own path, except for a path's
    // first tile when empty paths start at the same index: the search may
    // stop on one of them (bound.cs.glsl:53-66)
    uint32_t word = a.tile_path_info[dpath].w & 0x00ffffffu, word0 = word;
    if (tile_x == 0 && dpath > 0 && a.tile_path_info[dpath - 1].z == dofs)
        word0 = a.tile_path_info[reference_tile_search(a, dofs)].w & 0x00ffffffu;
    uint32_t count = 0;
    int32_t dsum = 0;
    uint32_t y_end = min(y_begin + kPropSeg, height);
    // rows are fetched kPropRows at a time (the loads do not depend on the
    // running backdrop), then processed in order
    for (uint32_t y0 = y_begin; y0 < y_end; y0 += kPropRows) {
        uint32_t dl[kPropRows];
        uint2 cl[kPropRows];
        uint8_t fl[kPropRows];
        // unconditional loads at clamped indices (a load under a branch is
        // waited for before the branch joins); rows past the column or the
        // buffers are not used below
#pragma unroll
        for (uint32_t u = 0; u < kPropRows; u++) {
            uint32_t y = y0 + u;
            uint32_t t = dofs + y * width + (uint32_t)tile_x;
            dl[u] = a.tile_delta[min(t, a.tile_count - 1u)];
            fl[u] = a.tile_flags[min(t, a.tile_count - 1u)];
            if (cload) {
                uint32_t ux = (uint32_t)tile_x + rx0, uy = y + ry0;
                uint32_t ct = cofs + (uy - cry0) * (crx1 - crx0) + (ux - crx0);
                cl[u] = a.clip_tiles[min(ct, a.clip_tile_count - 1u)];  // words 2-3
            }
        }
        bool stop = false;
#pragma unroll
        for (uint32_t u = 0; u < kPropRows; u++) {
            uint32_t y = y0 + u;
            if (y >= y_end) break;
            uint32_t t = dofs + y * width + (uint32_t)tile_x;
            if (t >= a.tile_count) {
                atomicOr(&a.counters->error, 4u);
                stop = true;
                break;
            }
            int32_t draw_alpha = -1, clip_alpha = -1;
            int32_t delta = (int32_t)dl[u] >> 24;
            int32_t tile_backdrop = backdrop;
            // first_fill >= 0 in the shader: bin gave the tile a fill
            bool have_mask = fl[u] != 0;
            bool need_alpha = have_mask;
            if (clip) {
                uint32_t ux = (uint32_t)tile_x + rx0, uy = y + ry0;  // unsigned, as the shader
                if (ux >= crx0 && uy >= cry0 && ux < crx1 && uy < cry1) {
                    uint32_t ct = cofs + (uy - cry0) * (crx1 - crx0) + (ux - crx0);
                    uint2 cw = ct < a.clip_tile_count ? cl[u] : make_uint2(0, 0);
                    int32_t this_clip_alpha = sext24(cw.x);
                    int32_t clip_backdrop = (int32_t)cw.y >> 24;
                    if (this_clip_alpha >= 0) {
                        if (have_mask) {
                            clip_alpha = this_clip_alpha;
                            need_alpha = true;
                        } else if (tile_backdrop != 0) {
                            draw_alpha = this_clip_alpha;
                            clip_alpha = -1;
                            need_alpha = false;
                        } else {
                            draw_alpha = -1;
                            clip_alpha = -1;
                            need_alpha = false;
                        }
                    } else if (clip_backdrop == 0) {
                        tile_backdrop = 0;
                        need_alpha = false;
                    }
                } else {
                    tile_backdrop = 0;
                    need_alpha = false;
                }
            }
            uint32_t ai = alpha_base + count;
            if (need_alpha) {
                if (EMIT) {
                    a.alpha_tiles[ai] = make_uint2(t, (uint32_t)clip_alpha);
                    draw_alpha = (int32_t)(ai + first_alpha);
                }
                count++;
            }
            if (EMIT) {
                // next ~0 (k_sort_tiles links it later), first fill ~0 unless
                // the tile has fills (k_fill_link, running beside this
                // kernel, writes that word), alpha tile and delta, word 3
                uint32_t z = ((uint32_t)draw_alpha & 0x00ffffffu) | ((uint32_t)delta << 24);
                uint32_t w = (y == 0 ? word0 : word) | ((uint32_t)tile_backdrop << 24);
                a.tile_status[t] =
                    (uint8_t)((tile_backdrop != 0 || draw_alpha >= 0 ? kTileLinked : 0) |
                              (tile_backdrop != 0 && draw_alpha < 0 && zwrite ? kTileOccluder
                                                                               : 0) |
                              (draw_alpha >= 0 ? kTileAlpha : 0));
                // words 0 and 1 live in tile_next / tile_first; words 2-3
                // are written whole (no partial-sector writes)
                a.tiles[t] = make_uint2(z, w);
                if (need_alpha) a.alpha_words[ai] = make_uint2(z, w);
                // k_sort_tiles links the linked tiles inside the framebuffer
                // grid; the others keep an empty link
                int32_t gx = (int32_t)(rx0 + (uint32_t)tile_x), gy = (int32_t)(ry0 + y);
                if ((tile_backdrop != 0 || draw_alpha >= 0) &&
                    (gx < 0 || gy < 0 || gx >= a.fb_width_tiles || gy >= a.fb_height_tiles))
                    a.tile_next[t] = 0xffffffffu;
            }
            backdrop += delta;
            dsum += delta;
        }
        if (stop) break;
    }
    if (delta_sum) *delta_sum = dsum;
    return count;
}

// The number of kPropSeg-row segments of column col (one for a column whose
// path index is out of range, which propagate_segment reports).
__device__ __forceinline__ uint32_t column_segments(const BatchArgs &a, uint32_t col) {
    uint32_t dpath = (uint32_t)a.backdrops[(size_t)col * 3 + 2];
    if (dpath >= a.path_count) return 1;
    const int32_t *md = a.prop_md + (size_t)dpath * 12;
    int32_t height = md[3] - md[1];
    return height > 0 ? ((uint32_t)height + kPropSeg - 1) / kPropSeg : 0u;
}

// Workgroups (x, y) cover columns 256x.. of row segment y (neighbouring
// threads walk neighbouring columns); threads past their column's height
// have nothing to do.
// Per column segment: its alpha tile count and backdrop delta sum, and the
// column's alpha count (the scan's input) as the sum over its segments.
__global__ void __launch_bounds__(kBlock) k_propagate_count(BatchArgs a) {
    uint2 cs = make_uint2(blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
    if (cs.x < a.column_count && cs.y < column_segments(a, cs.x)) {
        int32_t dsum = 0;
        uint32_t n = propagate_segment<false>(a, cs.x, cs.y * kPropSeg, 0, 0, 0, &dsum);
        size_t k = (size_t)cs.y * a.column_count + cs.x;
        a.seg_delta[k] = dsum;
        a.seg_alpha[k] = n;
        if (n) atomicAdd(&a.col_alpha_offsets[cs.x], n);
    }
}

// Allocates the batch's alpha tile range in the frame (uFirstAlphaTileIndex,
// d3d11/renderer.rs:624-625, 643-646) and writes the z-buffer header the
// reference leaves behind: bin uploads [6,0,0,0,0,microlines,0,0], bin bumps
// [1], propagate_tiles zeroes the first tile_area words and propagate bumps [4].
// count: the batch's alpha tiles (the column scan's total).
__device__ void propagate_finalize(const BatchArgs &a, uint32_t count) {
    BatchCounters *c = a.counters;
    uint32_t start = a.frame->alpha_count;
    c->alpha = count;
    c->alpha_start = start;
    uint32_t end = start + count;
    a.frame->alpha_count = end;
    c->frame_alpha_end = end;
    if (end > a.mask_capacity) atomicOr(&c->overflow, 4u);
    int32_t hdr[8] = {6, (int32_t)c->fills, 0, 0, 0, (int32_t)c->microlines, 0, 0};
    uint32_t fbn = (uint32_t)(a.fb_width_tiles * a.fb_height_tiles);
    for (uint32_t i = 0; i < 8; i++) a.zbuffer[i] = i < fbn ? 0 : hdr[i];
    a.zbuffer[4] += (int32_t)count;
}

// The partials of the column scan, then propagate_finalize with its total.
__global__ void __launch_bounds__(kBlock) k_propagate_scan_partials(BatchArgs a) {
    uint32_t t = scan_partials_total(a.scan_partials, nullptr, a.column_count, a.column_count,
                                     nullptr, a.col_alpha_offsets);
    if (threadIdx.x == 0) propagate_finalize(a, t);
}

// Per column segment: the running backdrop and alpha tile id at its first
// row are the column's (after bin) plus the sums of the segments above it.
__global__ void __launch_bounds__(kBlock) k_propagate_emit(BatchArgs a) {
    uint32_t first_alpha = a.counters->alpha_start;
    uint2 cs = make_uint2(blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
    if (cs.x < a.column_count && cs.y < column_segments(a, cs.x)) {
        int32_t backdrop = a.backdrops[(size_t)cs.x * 3 + 0];
        uint32_t alpha = a.col_alpha_offsets[cs.x];
        for (uint32_t q = 0; q < cs.y; q++) {
            size_t k = (size_t)q * a.column_count + cs.x;
            backdrop += a.seg_delta[k];
            alpha += a.seg_alpha[k];
        }
        propagate_segment<true>(a, cs.x, cs.y * kPropSeg, backdrop, alpha, first_alpha, nullptr);
    }
}

// ------------------------------------------------------------------ sort

// Lists the paths whose tile rect covers each framebuffer tile row, in path
// order (one workgroup per row; a thread tests four consecutive paths per
// step, so a step covers 1024 paths with one block scan).
__global__ void __launch_bounds__(kBlock) k_row_paths(BatchArgs a) {
    constexpr uint32_t kPer = 4;
    __shared__ uint32_t lds4[4];
    int32_t y = (int32_t)blockIdx.x;
    if (y >= a.fb_height_tiles) return;
    uint32_t carry = 0;
    uint32_t *out = a.row_paths + (size_t)y * a.path_count;
    for (uint32_t base = 0; base < a.path_count; base += kBlock * kPer) {
        uint32_t p0 = base + threadIdx.x * kPer;
        uint32_t hits = 0, n = 0;
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            uint32_t p = p0 + k;
            uint4 info = a.tile_path_info[min(p, a.path_count - 1u)];
            int32_t miny = (int32_t)(info.x) >> 16, maxy = (int32_t)(info.y) >> 16;
            int32_t minx = (int32_t)(info.x << 16) >> 16, maxx = (int32_t)(info.y << 16) >> 16;
            bool hit = p < a.path_count && y >= miny && y < maxy && minx < maxx && maxx > 0 &&
                       minx < a.fb_width_tiles;
            hits |= (hit ? 1u : 0u) << k;
            n += hit ? 1u : 0u;
        }
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(n, lds4, &tot);
        uint32_t o = carry + ex;
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++)
            if (hits >> k & 1u) out[o++] = p0 + k;
        carry += tot;
    }
    if (threadIdx.x == 0) a.row_counts[y] = carry;
}

// sort.cs.glsl:60-95 together with the linking half of propagate.cs.glsl:
// 201-216. For framebuffer tile (x, y) the tiles linked into the list are the
// covering paths' tiles with a nonzero backdrop or an alpha tile; the z value
// is the largest index of a linked, z-writing, solid tile (atomicMax from 0).
// In sequential order each link pushes onto the list head, so before the sort
// a tile's next is the previous linked tile (the next lower index); the sort
// keeps tiles with index >= z in ascending order and leaves culled tiles with
// their pre-sort next. A wavefront walks the covering paths from the top
// down, 64 at a time, and writes exactly those words; of each covering tile
// it reads the status byte propagate wrote beside it.
// A workgroup takes kSortBand framebuffer tiles of one row, two adjacent
// tiles per wavefront (their status bytes share a cache line, so the second
// probe of a path is a cache hit), lists in LDS the row's paths that overlap
// the band, and stages their rects in LDS 64 at a time for all of its
// wavefronts. Workgroups go to the eight XCDs in turn (workgroup i to XCD
// i mod 8), so workgroup i takes band (i mod 8) * ceil(n / 8) + i / 8 of the
// n bands in row-major order: each XCD takes a run of whole rows, and a
// path's status bytes, read by the bands of adjacent rows, stay in one L2.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kSortWaves = 16;
constexpr uint32_t kSortPer = 2;  // framebuffer tiles per wavefront
constexpr uint32_t kSortBand = kSortWaves * kSortPer;
constexpr uint32_t kBandCap = 8192;  // band list entries held in LDS

// The walk of one framebuffer tile: the z-buffer search, the links and the
// list statistics, carried from chunk to chunk of covering paths.
struct SortWalk {
    bool z_found;
    int32_t z;
    int32_t succ;             // lowest linked tile above the current chunk
    int32_t pending;          // culled tile whose next is in a lower chunk
    int32_t lowest_survivor;
    uint32_t n_linked, n_survivors, n_survivor_alpha;
};

__global__ void __launch_bounds__(kSortWaves * 64) k_sort_tiles(BatchArgs a) {
    __shared__ int32_t s_miny[64], s_minx[64], s_maxx[64];
    __shared__ uint32_t s_first[64];
    __shared__ uint32_t s_band[kBandCap];
    __shared__ uint32_t s_wave[kSortWaves + 1];
    uint32_t bands = ((uint32_t)a.fb_width_tiles + kSortBand - 1) / kSortBand;
    uint32_t per_xcd = gridDim.x / kXcds;
    uint32_t linear = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
    int32_t y = (int32_t)(linear / bands);
    uint32_t band = linear % bands;
    if (y >= a.fb_height_tiles) return;
    uint32_t lane = lane_id();
    int32_t x0 = (int32_t)(band * kSortBand + (threadIdx.x >> 6) * kSortPer);
    const uint32_t *list = a.row_paths + (size_t)y * a.path_count;
    uint32_t len = a.row_counts[y];
    // The row's paths that overlap this band, in order, compacted into LDS.
    int32_t bx0 = (int32_t)(band * kSortBand), bx1 = bx0 + (int32_t)kSortBand;
    uint32_t band_len = 0;
    for (uint32_t base = 0; base < len; base += blockDim.x) {
        uint32_t k = base + threadIdx.x;
        bool keep = false;
        uint32_t p = 0;
        if (k < len) {
            p = list[k];
            uint4 info = a.tile_path_info[p];
            int32_t minx = (int32_t)(info.x << 16) >> 16, maxx = (int32_t)(info.y << 16) >> 16;
            keep = minx < bx1 && maxx > bx0;
        }
        unsigned long long m = __ballot(keep);
        if (lane == 0) s_wave[threadIdx.x >> 6] = __popcll(m);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t w = 0; w < kSortWaves; w++) {
            uint32_t c = s_wave[w];
            before += w < (threadIdx.x >> 6) ? c : 0u;
            total += c;
        }
        uint32_t pos = band_len + before + __popcll(m & ((1ull << lane) - 1ull));
        if (keep && pos < kBandCap) s_band[pos] = p;
        band_len += total;
        __syncthreads();
    }
    bool in_lds = band_len <= kBandCap;
    if (in_lds) len = band_len;

    SortWalk wk[kSortPer];
#pragma unroll
    for (uint32_t u = 0; u < kSortPer; u++)
        wk[u] = SortWalk{false, 0, -1, -1, -1, 0u, 0u, 0u};
    for (uint32_t end = len; end > 0;) {
        uint32_t begin = end > 64 ? end - 64 : 0;
        uint32_t k = begin + lane;
        if (threadIdx.x < 64 && k < end) {
            uint32_t p = in_lds ? s_band[k] : list[k];
            uint4 info = a.tile_path_info[p];
            s_miny[lane] = (int32_t)(info.x) >> 16;
            s_minx[lane] = (int32_t)(info.x << 16) >> 16;
            s_maxx[lane] = (int32_t)(info.y << 16) >> 16;
            s_first[lane] = info.z;
        }
        __syncthreads();
        int32_t miny = 0, minx = 0, maxx = 0;
        uint32_t first = 0;
        if (k < end) {
            miny = s_miny[lane];
            minx = s_minx[lane];
            maxx = s_maxx[lane];
            first = s_first[lane];
        }
        __syncthreads();
#pragma unroll
        for (uint32_t u = 0; u < kSortPer; u++) {
            SortWalk &w = wk[u];
            int32_t x = x0 + (int32_t)u;
            bool valid = x < a.fb_width_tiles && k < end;
            int32_t t = -1;
            bool linked = false, occl = false, alpha_tile = false;
            if (valid) {
                uint32_t ti = first + (uint32_t)(y - miny) * (uint32_t)(maxx - minx) +
                              (uint32_t)(x - minx);
                if (x >= minx && x < maxx && ti < a.tile_count) {
                    t = (int32_t)ti;
                    uint32_t st = a.tile_status[t];  // words 2-3 as propagate wrote them
                    linked = (st & kTileLinked) != 0;
                    alpha_tile = (st & kTileAlpha) != 0;
                    occl = (st & kTileOccluder) != 0;
                }
            }
            unsigned long long lmask = __ballot(linked);
            unsigned long long omask = __ballot(occl);
            unsigned long long amask = __ballot(alpha_tile);
            int32_t zlane;  // lanes >= zlane survive
            if (w.z_found) {
                zlane = 64;
            } else if (omask) {
                zlane = 63 - __clzll((long long)omask);
                w.z = __shfl(t, zlane, 64);
                w.z_found = true;
            } else {
                zlane = 0;
            }
            if (lmask) {
                // the pending culled tile's next is this chunk's highest linked tile
                int32_t hi_lane = 63 - __clzll((long long)lmask);
                int32_t hi_t = __shfl(t, hi_lane, 64);
                if (w.pending >= 0 && lane == 0) a.tile_next[w.pending] = (uint32_t)hi_t;
                if (w.pending >= 0) w.pending = -1;
                int32_t lo_lane = __ffsll((long long)lmask) - 1;
                int32_t lo_t = __shfl(t, lo_lane, 64);
                if (linked) {
                    unsigned long long above = lane == 63 ? 0ull : (lmask & (~0ull << (lane + 1)));
                    unsigned long long below = lmask & ((1ull << lane) - 1ull);
                    int32_t up_lane = above ? __ffsll((long long)above) - 1 : 64;
                    int32_t down_lane = below ? 63 - __clzll((long long)below) : -1;
                    int32_t up_t = __shfl(t, up_lane & 63, 64);
                    int32_t down_t = __shfl(t, down_lane & 63, 64);
                    if ((int32_t)lane >= zlane) {
                        a.tile_next[t] = (uint32_t)(up_lane < 64 ? up_t : w.succ);
                    } else if (down_lane >= 0) {
                        a.tile_next[t] = (uint32_t)down_t;
                    }
                } else {
                    // keep the shuffles uniform across the wave
                    (void)__shfl(t, 0, 64);
                    (void)__shfl(t, 0, 64);
                }
                // lowest linked lane culled -> its next lies in a lower chunk
                if (lo_lane < zlane) w.pending = lo_t;
                unsigned long long surv = lmask & (zlane >= 64 ? 0ull : (~0ull << zlane));
                if (surv) w.lowest_survivor = __shfl(t, __ffsll((long long)surv) - 1, 64);
                w.n_linked += __popcll(lmask);
                w.n_survivors += __popcll(surv);
                w.n_survivor_alpha += __popcll(surv & amask);
                w.succ = lo_t;
            }
        }
        end = begin;
    }
#pragma unroll
    for (uint32_t u = 0; u < kSortPer; u++) {
        int32_t x = x0 + (int32_t)u;
        if (lane == 0 && x < a.fb_width_tiles) {
            const SortWalk &w = wk[u];
            uint32_t wave = (uint32_t)y * (uint32_t)a.fb_width_tiles + (uint32_t)x;
            if (w.pending >= 0) a.tile_next[w.pending] = 0xffffffffu;
            a.first_tile_map[wave] = w.lowest_survivor;
            a.zbuffer[8 + wave] = w.z_found ? w.z : 0;
            a.list_stats[wave] = make_uint4(w.n_linked, w.n_survivors, w.n_survivor_alpha, 0);
        }
    }
}

// ------------------------------------------------------------------ fill

// One fill in tile coordinates (pixels), the values every pixel of the tile
// shares: both x ends in list order, the left end (smaller x) and its y, the
// rise from the left to the right end, the width w = right.x - left.x, 1/w
// rounded to nearest, and the slope rise / w.
struct FillRecord {
    float fx, tx, lx, ly;
    float dy, w, rw, d;
};

// texture(uAreaLUT, .) at texel coordinates (s, u) = (256 x - 0.5, 256 y - 0.5)
// with GL_LINEAR, CLAMP_TO_EDGE and fp32 weights wa = s - floor(s),
// wb = u - floor(u). Each linear interpolation is the one Mesa's llvmpipe
// emits for float textures, v0 + w * (v1 - v0) as a fused multiply-add
// (lp_build_lerp): A = lerp(wa, t00, t10), B = lerp(wa, t01, t11), then
// lerp(wb, A, B).
// lut holds, per padded texel, the texel and its difference to the next
// texel along s: 8 floats. The 256 x 256 texels are padded with a copy of
// the edge on every side: padded texel (j, i), 0 <= i, j <= 258, is texel
// (clamp(j - 1), clamp(i - 1)) with clamp to [0, 255], rows kLutStride
// texels apart. With k = clamp(index + 1, 0, 257) the texels k and k + 1 of
// the padded table are the clamped texels index and index + 1 for every
// index, so one clamp per axis replaces four.
// The table is read through a raw buffer descriptor so that a row's texel
// and difference come from one offset: +16 bytes in the instruction, +1 row
// in an SGPR.
constexpr uint32_t kLutTexelBytes = 32;
typedef __amdgpu_buffer_rsrc_t LutRef;
__device__ __forceinline__ LutRef lut_ref(const float4 *lut) {
    // word 3 of a raw (untyped) gfx9 buffer descriptor; the record count is
    // the table size in bytes, so an offset past it would read zeros
    return __builtin_amdgcn_make_buffer_rsrc((void *)lut, (short)0,
                                             (int)(kLutRows * kLutStride * kLutTexelBytes),
                                             0x00020000);
}
__device__ __forceinline__ float4 lut_texel(LutRef lut, uint32_t offset, uint32_t row) {
    auto v = __builtin_amdgcn_raw_buffer_load_b128(lut, offset, row, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                       __uint_as_float(v[3]));
}
// The bilinear blend of padded texels (jp, ip) .. (jp + 1, ip + 1) at byte
// offset o = (jp * kLutStride + ip) * kLutTexelBytes.
__device__ __forceinline__ float4 lut_blend(LutRef lut, uint32_t o, float wa, float wb) {
    constexpr uint32_t kRow = kLutStride * kLutTexelBytes;
    float4 t00 = lut_texel(lut, o, 0u);
    float4 d0 = lut_texel(lut, o + 16u, 0u);
    float4 t01 = lut_texel(lut, o, kRow);
    float4 d1 = lut_texel(lut, o + 16u, kRow);
    float4 A, B, r;
    A.x = __builtin_fmaf(wa, d0.x, t00.x); B.x = __builtin_fmaf(wa, d1.x, t01.x);
    A.y = __builtin_fmaf(wa, d0.y, t00.y); B.y = __builtin_fmaf(wa, d1.y, t01.y);
    A.z = __builtin_fmaf(wa, d0.z, t00.z); B.z = __builtin_fmaf(wa, d1.z, t01.z);
    A.w = __builtin_fmaf(wa, d0.w, t00.w); B.w = __builtin_fmaf(wa, d1.w, t01.w);
    r.x = __builtin_fmaf(wb, B.x - A.x, A.x);
    r.y = __builtin_fmaf(wb, B.y - A.y, A.y);
    r.z = __builtin_fmaf(wb, B.z - A.z, A.z);
    r.w = __builtin_fmaf(wb, B.w - A.w, A.w);
    return r;
}
__device__ __forceinline__ float4 sample_lut_texel(LutRef lut, float s, float u) {
    float fs = floorf(s), fu = floorf(u);
    float wa = s - fs, wb = u - fu;
    // |s|, |u| < 2^21 for every packed fill, so the integer conversion is exact
    int32_t is = (int32_t)fs, iu = (int32_t)fu;
    uint32_t ip = (uint32_t)min(max(is + 1, 0), 257);
    uint32_t jp = (uint32_t)min(max(iu + 1, 0), 257);
    return lut_blend(lut, (jp * kLutStride + ip) * kLutTexelBytes, wa, wb);
}

// computeCoverage (fill_area.inc.glsl:11-27) of one fill for the pixel column
// centred at fragx and the four rows from fragy, bit-identical to the plain
// statement of the function on pixel-relative coordinates:
//  - packed coordinates are multiples of 1/256 below 256, so the differences
//    right - left (x and y) and every pixel-relative coordinate are exact and
//    the pixel offset cancels in w, dy and the slope;
//  - the window middle and offset are exact; the products by 0.5 and 16 are
//    exact, so fusing them into fma changes no rounding;
//  - offset / w comes from the rounded reciprocal and one residual step,
//    which returns the correctly rounded quotient for every w = k / 256,
//    k in [1, 65535], and offset in [0, 2w] (checked for all of them); a
//    column the fill touches always has offset in [0, w].
__device__ __forceinline__ float4 coverage_record(LutRef lut,
                                                  const FillRecord &f, float fragx, float fragy) {
    float wx = __builtin_amdgcn_fmed3f(f.fx - fragx, -0.5f, 0.5f);
    float wy = __builtin_amdgcn_fmed3f(f.tx - fragx, -0.5f, 0.5f);
    float offset = __builtin_fmaf(0.5f, wy - wx, wx) - (f.lx - fragx);
    float q0 = offset * f.rw;
    float t = __builtin_fmaf(__builtin_fmaf(-q0, f.w, offset), f.rw, q0);
    float yy = (f.ly - fragy) + t * f.dy;
    float dX = wx - wy;
    // (yy + 8) / 16 * 256 - 0.5 and |d dX| / 16 * 256 - 0.5
    float4 s = sample_lut_texel(lut, __builtin_fmaf(yy + 8.0f, 16.0f, -0.5f),
                                __builtin_fmaf(fabsf(f.d * dX), 16.0f, -0.5f));
    return make_float4(s.x * dX, s.y * dX, s.z * dX, s.w * dX);
}

// Orders this wavefront's LDS accesses: what one lane wrote before the call
// is visible to every lane of the wavefront after it.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

// The pixel columns c0..c1 of the tile (0..15) that a fill's window sees; an
// empty range (c0 > c1) when it sees none.
struct FillColumns {
    uint32_t c0, c1;
};

// Builds the record of packed fill fw (Fill words 0-1) and its column range.
__device__ __forceinline__ FillColumns make_record(uint2 fw, FillRecord &f) {
    uint32_t x0 = fw.x & 0xffffu, x1 = fw.y & 0xffffu;
    uint32_t lo = min(x0, x1), hi = max(x0, x1);
    bool any = hi > lo;
    bool lt = x0 < x1;
    uint32_t y0 = fw.x >> 16, y1 = fw.y >> 16;
    f.fx = (float)x0 * (1.0f / 256.0f);
    f.tx = (float)x1 * (1.0f / 256.0f);
    f.lx = (float)lo * (1.0f / 256.0f);
    f.ly = (float)(lt ? y0 : y1) * (1.0f / 256.0f);
    float ry = (float)(lt ? y1 : y0) * (1.0f / 256.0f);
    f.dy = ry - f.ly;
    f.w = (float)(hi - lo) * (1.0f / 256.0f);
    f.rw = any ? 1.0f / f.w : 0.0f;
    f.d = any ? f.dy / f.w : 0.0f;
    FillColumns c;
    c.c0 = lo >> 8;
    c.c1 = any ? min((hi - 1u) >> 8, 15u) : 0u;
    if (!any || c.c0 > 15u) {
        c.c0 = 1u;
        c.c1 = 0u;
    }
    return c;
}

// Adds the coverage of n <= 64 fills held one per lane (fill k in lane k, in
// the order of the list walk of fill_compute.inc.glsl) to this lane's pixels.
// A fill adds exactly +-0 to a pixel column its window does not see (dX = 0:
// both ends at or left of the column's left edge, or both at or right of its
// right edge) and the sum ends in abs(), so each lane walks only the fills
// that touch its column, in list order. rec is this wavefront's 64 records
// in LDS.
__device__ __forceinline__ void accumulate_chunk(LutRef lut, FillRecord *rec,
                                                 uint2 fw, uint32_t n, uint32_t lane,
                                                 uint32_t col, float fragx, float fragy,
                                                 float4 &acc) {
    FillRecord f;
    FillColumns fc = make_record(fw, f);
    bool have = lane < n;
    uint32_t cm = (have && fc.c0 <= fc.c1) ? (2u << fc.c1) - (1u << fc.c0) : 0u;
    if (have) rec[lane] = f;
    wave_lds_sync();
    unsigned long long mine = 0;
    if (n <= 16) {
        for (uint32_t k = 0; k < n; k++) {
            uint32_t m = __builtin_amdgcn_readlane(cm, k);
            mine |= (unsigned long long)((m >> col) & 1u) << k;
        }
    } else {
#pragma unroll
        for (uint32_t q = 0; q < 16; q++) {
            unsigned long long m = __ballot((cm >> q) & 1u);
            mine = q == col ? m : mine;
        }
    }
    while (mine != 0ull) {
        uint32_t j = (uint32_t)__ffsll((long long)mine) - 1u;
        mine &= mine - 1ull;
        float4 cv = coverage_record(lut, rec[j], fragx, fragy);
        acc.x += cv.x; acc.y += cv.y; acc.z += cv.z; acc.w += cv.w;
    }
    // every lane is done with the records before the next chunk rewrites them
    wave_lds_sync();
}

// byte / 255 rounded to nearest, from the rounded reciprocal and one residual
// step (equal to the IEEE quotient for every byte).
__device__ __forceinline__ float byte_to_unit(uint32_t c) {
    const float r = 1.0f / 255.0f;
    float x = (float)c;
    float q0 = x * r;
    return __builtin_fmaf(__builtin_fmaf(-q0, 255.0f, x), r, q0);
}

// unorm8 of a value already in [0, 1]: floor(c * 255 + 0.5), the floor being
// the truncation of a positive value.
__device__ __forceinline__ uint32_t unorm8_unit(float c) { return (uint32_t)(c * 255.0f + 0.5f); }

// fill.cs.glsl:80-87: backdrop + coverage, abs, clamp, min with the clip
// mask texel, unorm8 store of the four rows of this lane's mask texel. The
// coverage sums are finite (sums of LUT texels in [0, 1] times window widths
// in [-1, 1]), so clamp(|x|, 0, 1) is med3(|x|, 0, 1), and the stored values
// need no second clamp. For the adjoint it also writes, as asked, the signed
// sums (coverage) or their 2-bit states (cov_state: the gate 0 < |C| < 1
// with the clip texel not below min(|C|, 1), and the sign), and in a
// recompute pass (states_only) only the states.
__device__ __forceinline__ void store_mask(const BatchArgs &a, uint32_t alpha_index,
                                           int32_t backdrop, int32_t clip, uint32_t lane,
                                           float4 acc) {
    float bd = (float)backdrop;
    float C0 = bd + acc.x, C1 = bd + acc.y, C2 = bd + acc.z, C3 = bd + acc.w;
    bool clipped = clip >= 0 && (uint32_t)clip < a.mask_capacity;
    if (!a.states_only) {
        float c0 = __builtin_amdgcn_fmed3f(fabsf(C0), 0.0f, 1.0f);
        float c1 = __builtin_amdgcn_fmed3f(fabsf(C1), 0.0f, 1.0f);
        float c2 = __builtin_amdgcn_fmed3f(fabsf(C2), 0.0f, 1.0f);
        float c3 = __builtin_amdgcn_fmed3f(fabsf(C3), 0.0f, 1.0f);
        if (clipped) {
            uint32_t ct = a.mask[(size_t)(uint32_t)clip * 64 + lane];
            c0 = min_glsl(c0, byte_to_unit(ct & 0xffu));
            c1 = min_glsl(c1, byte_to_unit((ct >> 8) & 0xffu));
            c2 = min_glsl(c2, byte_to_unit((ct >> 16) & 0xffu));
            c3 = min_glsl(c3, byte_to_unit(ct >> 24));
        }
        a.mask[(size_t)alpha_index * 64 + lane] = unorm8_unit(c0) | (unorm8_unit(c1) << 8) |
                                                  (unorm8_unit(c2) << 16) |
                                                  (unorm8_unit(c3) << 24);
    }
    if (a.cov_state) {
        uint32_t ct = clipped ? a.mask[(size_t)(uint32_t)clip * 64 + lane] : 0u;
        const float C[4] = {C0, C1, C2, C3};
        uint32_t st = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float aC = fabsf(C[k]);
            bool limited = clipped && byte_to_unit((ct >> (8 * k)) & 0xffu) < min(aC, 1.0f);
            bool open = aC > 0.0f && aC < 1.0f && !limited;
            st |= ((open ? 1u : 0u) | (C[k] < 0.0f ? 2u : 0u)) << (2 * k);
        }
        a.cov_state[(size_t)alpha_index * 64 + lane] = (uint8_t)st;
    }
    // checkpoint for the adjoint: the signed sum before abs and clamp, pixel
    // (column lane & 15, row 4 (lane >> 4) + k) at [alpha][k][lane]
    if (a.coverage) {
        float *cv = a.coverage + (size_t)alpha_index * 256 + lane;
        cv[0] = C0;
        cv[64] = C1;
        cv[128] = C2;
        cv[192] = C3;
    }
}

// The tile slot (0..63) of stream position q = pos + lane of a wavefront
// group (the last tile i with S_i <= q, S = exclusive scan of the tiles' fill
// counts n), that position's fill words, and whether it is its tile's last
// fill; nothing is loaded past total.
__device__ __forceinline__ void fill_lookup(const BatchArgs &a, uint32_t pos, uint32_t total,
                                            uint32_t S, uint32_t rx, uint32_t n, uint32_t lane,
                                            uint32_t &ts, uint2 &fw, bool &last) {
    uint32_t q = pos + lane;
    uint32_t t = 0;
#pragma unroll
    for (uint32_t step = 32; step != 0; step >>= 1) {
        uint32_t s = __shfl(S, t + step, 64);
        if (s <= q) t += step;
    }
    uint32_t trx = __shfl(rx, t, 64), tn = __shfl(n, t, 64), tS = __shfl(S, t, 64);
    ts = t;
    last = q + 1u == tS + tn;
    fw = make_uint2(0, 0);
    if (q < total) fw = a.sorted_fills[trx + tn - 1u - (q - tS)];
}

// The lanes whose pc equals this lane's index, for lanes 0..15 (0 in lanes
// 16..63), pc in 0..16 (16: none): the AND over pc's four low bits of the
// ballot of the bit or of its complement, as the lane index has the bit or
// not, with the ballot of pc < 16. Five ballots and a few bit operations
// per lane, in place of one ballot and two lane writes per column.
__device__ __forceinline__ unsigned long long column_lanes(uint32_t pc, uint32_t lane) {
    unsigned long long m = __ballot(pc < 16u);
#pragma unroll
    for (uint32_t b = 0; b < 4; b++) {
        unsigned long long bits = __ballot((pc >> b) & 1u);
        unsigned long long want = ((lane >> b) & 1u) ? ~0ull : 0ull;
        m &= ~(bits ^ want);
    }
    return lane < 16u ? m : 0ull;
}

// Alpha tiles per wavefront group: 32 (64 made the 4K fill 1.5 % slower: a
// group is the unit of work balance across waves), halved (down to 8) while
// the groups would number fewer than kFillWaves, so that batches of few
// alpha tiles with long fill runs (glyph pages) still spread over every SIMD.
constexpr uint32_t kFillGroup = 32;
constexpr uint32_t kFillGroupMin = 8;
constexpr uint32_t kFillWaves = 4096;

// computeCoverage (fill_area.inc.glsl:11-27) of one fill at the pixel column
// centred at x, for the four row groups of the tile (rows 4g .. 4g + 3, the
// four channels of a mask texel), in two halves: pair_column computes what
// depends on the column only (the x window, the quotient, the rise, the LUT
// row u and its weight), pair_group the rest for group g (yy, the LUT column
// s, the bilinear blend). Every value is bit-identical to
// coverage_record(lut, f, x, 4g + 0.5).
struct PairColumn {
    float ly, rise, wb, dX;  // fill's left y, rise to the column, LUT row weight, window
};

// pair_column also tells which row groups read saturated LUT cells. Group g
// samples padded cell (jp, ip_g) with ip_g = clamp(floor(sc_g) + 1, 0, 257);
// where that cell is all 255 (ip_g <= F[jp], lut_sat) the blend is exactly 1
// and the value exactly dX, where it is all 0 (ip_g >= Z[jp]) exactly +-0.
// sc_g equals sc_0 - 64 g to within 1.9e-4 (|ly|, |rise| <= 16, so |yy| <= 32
// and |sc| < 640: half-ulp errors of 2^-19 in yy and in yy + 8, times 16, and
// 2^-15 in sc, for both), so floor(sc_g) lies between floor(sc_0 - 4e-4) and
// floor(sc_0 + 4e-4), less 64 g, and the groups are classified on those
// bounds: groups [0, nz) certainly zero, [gf, 4) certainly dX, the groups
// between to be sampled by pair_group. row0 is the byte offset of LUT row jp.
__device__ __forceinline__ PairColumn pair_column(const FillRecord &f, float x,
                                                  const uint32_t *sat, uint32_t &row0,
                                                  uint32_t &nz, uint32_t &gf) {
    PairColumn c;
    float wx = __builtin_amdgcn_fmed3f(f.fx - x, -0.5f, 0.5f);
    float wy = __builtin_amdgcn_fmed3f(f.tx - x, -0.5f, 0.5f);
    float offset = __builtin_fmaf(0.5f, wy - wx, wx) - (f.lx - x);
    float q0 = offset * f.rw;
    float t = __builtin_fmaf(__builtin_fmaf(-q0, f.w, offset), f.rw, q0);
    c.rise = t * f.dy;
    c.dX = wx - wy;
    c.ly = f.ly;
    float u = __builtin_fmaf(fabsf(f.d * c.dX), 16.0f, -0.5f);
    float fu = floorf(u);
    c.wb = u - fu;
    uint32_t jp = (uint32_t)min(max((int32_t)fu + 1, 0), 257);
    row0 = jp * kLutStride * kLutTexelBytes;
    float yy0 = (f.ly - 0.5f) + c.rise;
    float s0 = __builtin_fmaf(yy0 + 8.0f, 16.0f, -0.5f);
    int32_t lo = (int32_t)floorf(s0 - 4e-4f) + 1, hi = (int32_t)floorf(s0 + 4e-4f) + 1;
    uint32_t w = sat[jp];
    int32_t full = (int32_t)(int16_t)(w & 0xffffu), zero = (int32_t)(w >> 16);
    // zero: lo - 64 g >= Z; dX: hi - 64 g <= F (the clamps to [0, 257] are
    // monotone, so this never claims a group the clamped index would not)
    nz = (uint32_t)min(max(((lo - zero) >> 6) + 1, 0), 4);
    gf = (uint32_t)min(max((hi - full + 63) >> 6, 0), 4);
    return c;
}

// The value of row group g of a pair from pair_column's results.
__device__ __forceinline__ float4 pair_group(LutRef lut, const PairColumn &c, uint32_t row0,
                                             uint32_t g) {
    float yy = (c.ly - ((float)(g * 4u) + 0.5f)) + c.rise;
    float sc = __builtin_fmaf(yy + 8.0f, 16.0f, -0.5f);
    float fs = floorf(sc);
    float wa = sc - fs;
    uint32_t ip = (uint32_t)min(max((int32_t)fs + 1, 0), 257);
    float4 r = lut_blend(lut, row0 + ip * kLutTexelBytes, wa, c.wb);
    return make_float4(r.x * c.dX, r.y * c.dX, r.z * c.dX, r.w * c.dX);
}

// fill.cs.glsl:61-88 with fill_compute.inc.glsl:11-25. Lane (lx, ly) of a
// wavefront owns the four pixels of column lx, rows 4ly..4ly+3 of an alpha
// tile (the four channels of mask texel (lx, ly)).
//
// A wavefront takes a group of consecutive alpha tiles (32, fewer when the
// batch has few) and streams their fills: each tile's fills from the highest
// index down (its run in sorted_fills), tile after tile, in chunks of at most
// 64. Lane k turns fill k of a chunk into a record, and the chunk's (fill,
// pixel column) pairs are numbered in stream order; a chunk ends where its
// pairs fill whole rounds (but for one straddling fill), not after 64 fills.
// Each fill sets the bit of its first pair in a 1024-bit start map, so pair
// p's fill is the number of set bits at or before p, less one (one LDS OR per
// fill instead of a byte per pair). Round by round (64 pairs), lane i evaluates computeCoverage of pair i for the
// four row groups into LDS (pair_column: the LUT is sampled only where its
// cell is not saturated), the values stored column by column in pair order.
// Then, tile by tile through the round, every lane adds its column's values
// of that tile, in order, to its four sums; a fill whose window does not see
// a column adds exactly +-0 there (dX = 0) and the sum ends in abs(), so it
// is left out. A tile's mask is stored once the stream has passed its last
// fill. A pixel thus sums the same fills in the same order as the
// reference's list walk.
// Tiles whose fills are split over several runs (rare) take the list walk
// one chunk at a time instead (accumulate_chunk).
constexpr uint32_t kZeroEntry = 3 * 65 + 64;  // the value entry that holds +0

// Five wavefronts per SIMD: at most 96 VGPRs and 32 KB of LDS per workgroup.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5)))
k_fill(BatchArgs a) {
    __shared__ FillRecord s_rec[kBlock / 64][64];
    // values: column slot s (0..63) of row group g at s + 65 g; then +0
    __shared__ float4 s_res[kBlock / 64][kZeroEntry + 1];
    __shared__ uint32_t s_cadj[kBlock / 64][64];
    __shared__ unsigned long long s_pstart[kBlock / 64][16];
    __shared__ unsigned long long s_cmask[kBlock / 64][16];
    __shared__ uint8_t s_cstart[kBlock / 64][16];
    __shared__ uint8_t s_quad[kBlock / 64][4 * 64];
    __shared__ uint32_t sat[258];
    for (uint32_t i = threadIdx.x; i < 258u; i += blockDim.x) sat[i] = a.lut_sat[i];
    __syncthreads();
    uint32_t wv = threadIdx.x >> 6, lane = lane_id();
    FillRecord *rec = s_rec[wv];
    float4 *res = s_res[wv];
    uint32_t *cadj = s_cadj[wv];
    unsigned long long *pstart = s_pstart[wv];
    unsigned long long *cmask = s_cmask[wv];
    uint8_t *cstart = s_cstart[wv];
    uint8_t *quad = s_quad[wv];
    if (lane == 0) res[kZeroEntry] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    LutRef lut = lut_ref(a.lut);
    uint32_t n_alpha = uniform(a.counters->alpha);
    uint32_t start = uniform(a.counters->alpha_start);
    uint32_t col = lane & 15u, grp = lane >> 4;
    float fragx = (float)col + 0.5f, fragy = (float)(grp * 4u) + 0.5f;
    uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    uint32_t group = kFillGroup;
    while (group > kFillGroupMin && (n_alpha + group - 1u) / group < min(waves, kFillWaves))
        group >>= 1;
    uint32_t stride = waves * group;
    for (uint32_t g0 = uniform(((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * group);
         g0 < n_alpha; g0 += stride) {
        // lane i < group: alpha tile g0 + i
        uint32_t ba = g0 + lane;
        bool valid = lane < group && ba < n_alpha && start + ba < a.mask_capacity;
        uint32_t t_clip = 0xffffffffu, t_w = 0, rx = 0, ry = 0;
        if (valid) {
            uint2 at = a.alpha_tiles[ba];
            t_clip = at.y;
            uint2 zw = a.alpha_words[ba];  // the tile's words 2-3
            valid = sext24(zw.x) >= 0;
            t_w = zw.y;
            if (valid) {
                uint2 run = a.tile_runs[at.x];
                rx = run.x;
                ry = run.y;
            }
        }
        uint32_t rn = ry & 0x7fffffffu;
        if (rx + rn > a.fill_capacity) rn = 0;
        bool chained = valid && rn != 0 && (ry & 0x80000000u) != 0u;
        uint32_t n = (valid && !chained) ? min(rn, 1024u) : 0u;
        uint32_t incl = wave_inclusive_scan(n);
        uint32_t S = incl - n;
        uint32_t total = last_lane(incl);

        // this lane's sums for the tile the stream is in
        float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        uint32_t ts;
        uint2 fw;
        bool tlast;
        fill_lookup(a, 0u, total, S, rx, n, lane, ts, fw, tlast);
        for (uint32_t pos = 0, K = 0; pos < total; pos += K) {
            bool have = pos + lane < total;
            FillRecord f;
            FillColumns fc = make_record(fw, f);
            // every fill takes at least one pair (a fill that sees no column
            // gets a pair whose value no lane reads), so each fill lies in
            // some round
            uint32_t nvis = (have && fc.c0 <= fc.c1) ? fc.c1 - fc.c0 + 1u : 0u;
            uint32_t span = have ? max(nvis, 1u) : 0u;
            uint32_t pe = wave_inclusive_scan(span);
            uint32_t pb = pe - span;
            // The chunk: all the fills left when they fit in the 64 lanes,
            // else the first K fills whose pairs end at or before the last
            // multiple of 64 of the 64 fills' pairs, so that only the round
            // of a straddling fill is not full (spans are 1..16 pairs, so K
            // >= 1); the next chunk starts at fill K
            if (pos + 64u >= total) {
                K = total - pos;
            } else {
                uint32_t whole = last_lane(pe) & ~63u;
                K = (uint32_t)__popcll(__ballot(pe <= whole));
            }
            have = lane < K;
            unsigned long long have_m = __ballot(have);
            uint32_t pairs = (uint32_t)__builtin_amdgcn_readlane((int)pe, (int)(K - 1u));
            // the chunk's start map: bit pb for each fill's first pair (the
            // clear and the ORs are this wavefront's LDS operations, in order)
            if (lane < 16u) pstart[lane] = 0ull;
            if (have) {
                rec[lane] = f;
                // first column - first pair (low 16 bits, signed), first column
                // (4 bits), visible columns (5 bits), tile slot (6 bits)
                cadj[lane] = (((uint32_t)fc.c0 - pb) & 0xffffu) | ((fc.c0 & 15u) << 16) |
                             (nvis << 20) | (ts << 25);
                atomicOr(&pstart[pb >> 6], 1ull << (pb & 63u));
            }
            // the fills that end their tile's stream
            unsigned long long fin_m = __ballot(have && tlast);
            // the next chunk's fills load meanwhile
            uint32_t nts;
            uint2 nfw;
            bool ntlast;
            fill_lookup(a, pos + K, total, S, rx, n, lane, nts, nfw, ntlast);
            wave_lds_sync();
            for (uint32_t r0 = 0; r0 < pairs; r0 += 64u) {
                uint32_t r1 = min(r0 + 64u, pairs);
                // lane i: pair r0 + i, its fill k, tile slot pt and pixel
                // column pc (16 when the fill's window sees no column: no
                // value)
                uint32_t p = r0 + lane;
                uint32_t k = 0, pc = 16u, pt = 64u;
                // pair p's fill: the last whose first pair is at or before p
                unsigned long long sb = pstart[r0 >> 6];
                uint32_t before = (uint32_t)__popcll(__ballot(have && pb < r0));
                if (p < r1) {
                    k = before + (uint32_t)__popcll(sb & ((2ull << lane) - 1ull)) - 1u;
                    uint32_t ck = cadj[k];
                    uint32_t c = p + (uint32_t)(int32_t)(int16_t)(ck & 0xffffu);
                    if (c - ((ck >> 16) & 15u) < ((ck >> 20) & 31u)) pc = c;
                    pt = ck >> 25;
                }
                // The round's values are stored column by column: column q's
                // pairs, in pair (= stream) order, from cstart[q] on. Lane
                // q < 16 gets the mask of the lanes whose pair is in column q
                // (0 in lanes 16..63).
                unsigned long long lanes_q = column_lanes(pc, lane);
                uint32_t cnt = (uint32_t)__popcll(lanes_q);
                uint32_t cs = wave_inclusive_scan(cnt) - cnt;
                if (lane < 16u) {
                    cmask[lane] = lanes_q;
                    cstart[lane] = (uint8_t)cs;
                }
                // the lanes whose pair starts a tile's run in the round
                uint32_t prev = (uint32_t)__shfl((int)pt, (int)lane - 1, 64);
                unsigned long long seg_m = __ballot(p < r1 && (lane == 0u || pt != prev));
                wave_lds_sync();
                // saturated row groups directly; the others (mixed) as a list
                // of (pair, group) items sampled 64 at a time
                uint32_t nz = 0, gf = 0, slot = 0, row0 = 0;
                PairColumn pcv = {0.0f, 0.0f, 0.0f, 0.0f};
                if (pc < 16u) {
                    uint32_t rank = (uint32_t)__popcll(cmask[pc] & ((1ull << lane) - 1ull));
                    slot = cstart[pc] + rank;
                    pcv = pair_column(rec[k], (float)pc + 0.5f, sat, row0, nz, gf);
                    float4 full = make_float4(pcv.dX, pcv.dX, pcv.dX, pcv.dX);
                    float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
                    for (uint32_t g = 0; g < 4; g++) {
                        if (g < nz) res[slot + 65u * g] = zero;
                        if (g >= gf) res[slot + 65u * g] = full;
                    }
                }
                uint32_t nmix = gf - nz;
                uint32_t qe = wave_inclusive_scan(nmix);
                uint32_t nq = last_lane(qe);
                uint32_t qb = qe - nmix;
#pragma unroll
                for (uint32_t i = 0; i < 4u; i++)
                    if (i < nmix) quad[qb + i] = (uint8_t)(lane | ((nz + i) << 6));
                wave_lds_sync();
                for (uint32_t q0 = 0; q0 < nq; q0 += 64u) {
                    uint32_t q = q0 + lane;
                    uint32_t d = quad[min(q, 255u)];
                    uint32_t src = d & 63u, g = d >> 6;
                    // the item's pair data from its pair's lane
                    PairColumn qc;
                    qc.ly = __shfl(pcv.ly, (int)src, 64);
                    qc.rise = __shfl(pcv.rise, (int)src, 64);
                    qc.wb = __shfl(pcv.wb, (int)src, 64);
                    qc.dX = __shfl(pcv.dX, (int)src, 64);
                    uint32_t qrs = (uint32_t)__shfl((int)(row0 | (slot << 24)), (int)src, 64);
                    if (q < nq) res[(qrs >> 24) + 65u * g] = pair_group(lut, qc, qrs & 0xffffffu, g);
                }
                wave_lds_sync();
                // tile by tile: this lane's column values of the tile's run
                // [L, E) of pair lanes are its values from e_prev to e
                unsigned long long my_lanes = cmask[col];
                const float4 *mine = res + 65u * grp + cstart[col];
                uint32_t e_prev = 0;
                // the round ends its last tile's stream if pair r1 - 1 is the
                // last pair of a fill that ends its tile
                uint32_t kz = (uint32_t)__popcll(__ballot(pe < r1) & have_m);
                bool fin = ((fin_m >> kz) & 1ull) &&
                           (uint32_t)__builtin_amdgcn_readlane((int)pe, (int)kz) == r1;
                while (seg_m != 0ull) {
                    uint32_t L = (uint32_t)__ffsll((long long)seg_m) - 1u;
                    seg_m &= seg_m - 1ull;
                    uint32_t E = seg_m ? (uint32_t)__ffsll((long long)seg_m) - 1u : r1 - r0;
                    uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)pt, (int)L);
                    unsigned long long below = E >= 64u ? ~0ull : (1ull << E) - 1ull;
                    uint32_t e = (uint32_t)__popcll(my_lanes & below);
                    // two values per step, read together (the second one
                    // from the zero entry past the end)
                    for (uint32_t j = e_prev; j < e; j += 2u) {
                        float4 v = mine[j];
                        float4 w = j + 1u < e ? mine[j + 1u] : res[kZeroEntry];
                        acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
                        acc.x += w.x; acc.y += w.y; acc.z += w.z; acc.w += w.w;
                    }
                    e_prev = e;
                    if (seg_m != 0ull || fin) {
                        store_mask(a, start + g0 + t,
                                   (int32_t)__builtin_amdgcn_readlane(t_w, t) >> 24,
                                   (int32_t)__builtin_amdgcn_readlane(t_clip, t), lane, acc);
                        acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    }
                }
                // every lane is done with this round's values
                wave_lds_sync();
            }
            ts = nts;
            fw = nfw;
            tlast = ntlast;
        }

        // tiles without fills: backdrop (and clip) only
        unsigned long long zero = __ballot(valid && !chained && n == 0u);
        while (zero != 0ull) {
            uint32_t i = (uint32_t)__ffsll((long long)zero) - 1u;
            zero &= zero - 1ull;
            store_mask(a, start + g0 + i, (int32_t)__builtin_amdgcn_readlane(t_w, i) >> 24,
                       (int32_t)__builtin_amdgcn_readlane(t_clip, i), lane,
                       make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        }

        // tiles with chained runs: walk the runs latest first, at most 1024
        // fills, 64 at a time
        unsigned long long chain = __ballot(chained);
        while (chain != 0ull) {
            uint32_t i = (uint32_t)__ffsll((long long)chain) - 1u;
            chain &= chain - 1ull;
            uint32_t crx = __builtin_amdgcn_readlane(rx, i);
            uint32_t cry = __builtin_amdgcn_readlane(ry, i);
            uint32_t crn = __builtin_amdgcn_readlane(rn, i);
            float4 cacc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            uint32_t budget = 1024, used = 0;
            while (budget != 0 && crn != 0) {
                uint2 fw = make_uint2(0, 0);
                uint32_t filled = 0;
                while (filled < 64 && budget != 0 && crn != 0) {
                    uint32_t take = min(min(crn - used, 64u - filled), budget);
                    if (lane >= filled && lane < filled + take)
                        fw = a.sorted_fills[crx + crn - 1 - used - (lane - filled)];
                    filled += take;
                    used += take;
                    budget -= take;
                    if (used == crn) {
                        used = 0;
                        crn = 0;
                        if (cry & 0x80000000u) {
                            uint2 prev = a.run_prev[crx];
                            crx = uniform(prev.x);
                            cry = uniform(prev.y);
                            crn = cry & 0x7fffffffu;
                            if (crx + crn > a.fill_capacity) crn = 0;
                        }
                    }
                }
                accumulate_chunk(lut, rec, fw, filled, lane, col, fragx, fragy, cacc);
            }
            store_mask(a, start + g0 + i, (int32_t)__builtin_amdgcn_readlane(t_w, i) >> 24,
                       (int32_t)__builtin_amdgcn_readlane(t_clip, i), lane, cacc);
        }
    }
}

// ------------------------------------------------------------------ paints
//
// calculateColor (tile_fragment.inc.glsl:567-614) for paints that are more
// than a solid colour: the colour texture (texture() with GL_LINEAR or
// GL_NEAREST, REPEAT or CLAMP_TO_EDGE; the bilinear form lerp(w, a, b) =
// fma(w, b - a, a) the area LUT uses too), the radial gradient, blur and
// colour matrix filters, SrcIn/DestIn combining, and the composite ops.
// Every operation is the oracle's, in the same order, without contraction.

// floor(x) as an index, saturated to +-1e9 first (NaN to -1e9): every index
// then stays in range of wrap_index and converts as on the host
__device__ __forceinline__ int32_t floor_index(float x) {
    return (int32_t)fminf(fmaxf(floorf(x), -1e9f), 1e9f);
}

__device__ __forceinline__ int32_t wrap_index(int32_t i, int32_t n, bool repeat) {
    if (repeat) {
        int32_t m = i % n;
        return m < 0 ? m + n : m;
    }
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

__device__ __forceinline__ float4 texel_unit(uint32_t v) {
    return make_float4((float)(v & 0xffu) / 255.0f, (float)((v >> 8) & 0xffu) / 255.0f,
                       (float)((v >> 16) & 0xffu) / 255.0f, (float)(v >> 24) / 255.0f);
}

__device__ float4 color_texture_sample(const TileArgs &a, float u, float v) {
    if (!a.color_texture || a.color_width <= 0 || a.color_height <= 0)
        return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    bool ru = (a.color_flags & 1u) != 0, rv = (a.color_flags & 2u) != 0;
    float W = (float)a.color_width, H = (float)a.color_height;
    if (a.color_flags & 8u) {
        int32_t i = wrap_index(floor_index(u * W), a.color_width, ru);
        int32_t j = wrap_index(floor_index(v * H), a.color_height, rv);
        return texel_unit(a.color_texture[(size_t)j * a.color_width + i]);
    }
    float sx = u * W - 0.5f, sy = v * H - 0.5f;
    float fx = floorf(sx), fy = floorf(sy);
    float wa = sx - fx, wb = sy - fy;
    int32_t ix = floor_index(sx), iy = floor_index(sy);
    int32_t i0 = wrap_index(ix, a.color_width, ru);
    int32_t i1 = wrap_index(ix + 1, a.color_width, ru);
    int32_t j0 = wrap_index(iy, a.color_height, rv);
    int32_t j1 = wrap_index(iy + 1, a.color_height, rv);
    float4 t00 = texel_unit(a.color_texture[(size_t)j0 * a.color_width + i0]);
    float4 t10 = texel_unit(a.color_texture[(size_t)j0 * a.color_width + i1]);
    float4 t01 = texel_unit(a.color_texture[(size_t)j1 * a.color_width + i0]);
    float4 t11 = texel_unit(a.color_texture[(size_t)j1 * a.color_width + i1]);
    float4 A, B, r;
    A.x = __builtin_fmaf(wa, t10.x - t00.x, t00.x); B.x = __builtin_fmaf(wa, t11.x - t01.x, t01.x);
    A.y = __builtin_fmaf(wa, t10.y - t00.y, t00.y); B.y = __builtin_fmaf(wa, t11.y - t01.y, t01.y);
    A.z = __builtin_fmaf(wa, t10.z - t00.z, t00.z); B.z = __builtin_fmaf(wa, t11.z - t01.z, t01.z);
    A.w = __builtin_fmaf(wa, t10.w - t00.w, t00.w); B.w = __builtin_fmaf(wa, t11.w - t01.w, t01.w);
    r.x = __builtin_fmaf(wb, B.x - A.x, A.x);
    r.y = __builtin_fmaf(wb, B.y - A.y, A.y);
    r.z = __builtin_fmaf(wb, B.z - A.z, A.z);
    r.w = __builtin_fmaf(wb, B.w - A.w, A.w);
    return r;
}

constexpr float kFrac6Pi = 1.9098593171027443f;
constexpr float kFracPi3 = 1.0471975511965976f;

__device__ __forceinline__ float comp_div(float n, float d) { return d != 0.0f ? n / d : 0.0f; }
__device__ __forceinline__ float glsl_mod(float x, float y) { return x - y * floorf(x / y); }

__device__ void rgb_to_hsl(const float rgb[3], float hsl[3]) {
    float v = max_glsl(max_glsl(rgb[0], rgb[1]), rgb[2]);
    float xmin = min_glsl(min_glsl(rgb[0], rgb[1]), rgb[2]);
    float c = v - xmin, l = mixf(xmin, v, 0.5f);
    float tx, ty, tz;
    if (rgb[0] == v) { tx = 0.0f; ty = rgb[1]; tz = rgb[2]; }
    else if (rgb[1] == v) { tx = 2.0f; ty = rgb[2]; tz = rgb[0]; }
    else { tx = 4.0f; ty = rgb[0]; tz = rgb[1]; }
    hsl[0] = kFracPi3 * comp_div((tx * c + ty) - tz, c);
    hsl[1] = comp_div(c, v);
    hsl[2] = l;
}

__device__ void hsl_to_rgb(const float hsl[3], float rgb[3]) {
    float a = hsl[1] * min_glsl(hsl[2], 1.0f - hsl[2]);
    const float k0[3] = {0.0f, 8.0f, 4.0f};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float ks = glsl_mod(k0[i] + hsl[0] * kFrac6Pi, 12.0f);
        float m = min_glsl(ks - 3.0f, 9.0f - ks);
        rgb[i] = hsl[2] - clamp_glsl(m, -1.0f, 1.0f) * a;
    }
}

__device__ __forceinline__ float screen1(float d, float s) { return (d + s) - d * s; }
__device__ __forceinline__ float hard_light1(float d, float s) {
    return s <= 0.5f ? (d * 2.0f) * s : screen1(d, 2.0f * s - 1.0f);
}
__device__ __forceinline__ float dodge1(float d, float s) {
    return d == 0.0f ? 0.0f : (s == 1.0f ? 1.0f : d / (1.0f - s));
}

// compositeRGB (tile_fragment.inc.glsl:465-501)
__device__ void composite_rgb(const float d[3], const float s[3], int op, float out[3]) {
    if (op >= 0xc) {
        float dh[3], sh[3], m[3];
        rgb_to_hsl(d, dh);
        rgb_to_hsl(s, sh);
        m[0] = (op == 0xc || op == 0xe) ? sh[0] : dh[0];
        m[1] = (op == 0xd || op == 0xe) ? sh[1] : dh[1];
        m[2] = op == 0xf ? sh[2] : dh[2];
        hsl_to_rgb(m, out);
        return;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float x = d[i], y = s[i], r;
        switch (op) {
        case 0x1: r = x * y; break;
        case 0x2: r = screen1(x, y); break;
        case 0x3: r = hard_light1(y, x); break;
        case 0x4: r = min_glsl(x, y); break;
        case 0x5: r = max_glsl(x, y); break;
        case 0x6: r = dodge1(x, y); break;
        case 0x7: r = 1.0f - dodge1(1.0f - x, 1.0f - y); break;
        case 0x8: r = hard_light1(x, y); break;
        case 0x9: {
            float dark = x <= 0.25f ? ((16.0f * x - 12.0f) * x + 4.0f) * x : sqrtf(x);
            float factor = y <= 0.5f ? x * (1.0f - x) : dark - x;
            r = x + (y * 2.0f - 1.0f) * factor;
            break;
        }
        case 0xa: r = fabsf(x - y); break;
        case 0xb: r = (x + y) - (2.0f * x) * y; break;
        default: r = y; break;
        }
        out[i] = r;
    }
}

// calculateColor after the mask for one pixel centre (fx, fy) and a paint
// whose metadata is m[0..9]: the premultiplied source colour.
__device__ float4 paint_color(const TileArgs &a, const float4 *m, float fx, float fy,
                              float mask_alpha) {
    float4 base = m[2];
    float col[4] = {base.x, base.y, base.z, base.w};
    int ctrl = (int)m[8].x;
    int combine = (ctrl >> 8) & 0x3;
    if (combine != 0) {
        float4 t0 = m[0], t1 = m[1];
        float u = (t0.x * fx + t0.z * fy) + t1.x;
        float v = (t0.y * fx + t0.w * fy) + t1.y;
        int filter = (ctrl >> 4) & 0xf;
        float4 p0 = m[3], p1 = m[4];
        float4 c0;
        if (filter == 1) {
            // filterRadialGradient (tile_fragment.inc.glsl:280-306)
            float dpx = u - p0.x, dpy = v - p0.y, dcx = p0.z, dcy = p0.w;
            float dr = p1.y - p1.x;
            float qa = (dcx * dcx + dcy * dcy) - dr * dr;
            float qb = (dpx * dcx + dpy * dcy) + p1.x * dr;
            float qc = (dpx * dpx + dpy * dpy) - p1.x * p1.x;
            float discrim = qb * qb - qa * qc;
            c0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (discrim != 0.0f) {
                float sq = sqrtf(discrim);
                float ta = (sq + qb) / qa, tb = (-sq + qb) / qa;
                if (ta > tb) {
                    float w = ta;
                    ta = tb;
                    tb = w;
                }
                float t = ta >= 0.0f ? ta : tb;
                c0 = color_texture_sample(a, p1.z + t, p1.w + 0.0f);
            }
        } else if (filter == 3) {
            // filterBlur (tile_fragment.inc.glsl:313-351)
            float ox = p0.x / (float)a.color_width, oy = p0.y / (float)a.color_height;
            int support = (int)p0.z;
            float gx = p1.x, gy = p1.y, gz = p1.z;
            float sum = gx;
            c0 = color_texture_sample(a, u, v);
            c0.x *= gx; c0.y *= gx; c0.z *= gx; c0.w *= gx;
            {
                float nx = gx * gy, ny = gy * gz;
                gx = nx;
                gy = ny;
            }
            for (int i = 1; i <= support; i += 2) {
                float partial = gx;
                {
                    float nx = gx * gy, ny = gy * gz;
                    gx = nx;
                    gy = ny;
                }
                partial += gx;
                float k = (float)i + gx / partial;
                float sx = ox * k, sy = oy * k;
                float4 lo = color_texture_sample(a, u - sx, v - sy);
                float4 hi = color_texture_sample(a, u + sx, v + sy);
                c0.x += (lo.x + hi.x) * partial;
                c0.y += (lo.y + hi.y) * partial;
                c0.z += (lo.z + hi.z) * partial;
                c0.w += (lo.w + hi.w) * partial;
                sum += 2.0f * partial;
                {
                    float nx = gx * gy, ny = gy * gz;
                    gx = nx;
                    gy = ny;
                }
            }
            c0.x /= sum; c0.y /= sum; c0.z /= sum; c0.w /= sum;
        } else if (filter == 4) {
            // filterColorMatrix (tile_fragment.inc.glsl:353-363)
            float4 src = color_texture_sample(a, u, v);
            float4 p2 = m[5], p3 = m[6], p4 = m[7];
            c0.x = (((p0.x * src.x + p1.x * src.y) + p2.x * src.z) + p3.x * src.w) + p4.x;
            c0.y = (((p0.y * src.x + p1.y * src.y) + p2.y * src.z) + p3.y * src.w) + p4.y;
            c0.z = (((p0.z * src.x + p1.z * src.y) + p2.z * src.z) + p3.z * src.w) + p4.z;
            c0.w = (((p0.w * src.x + p1.w * src.y) + p2.w * src.z) + p3.w * src.w) + p4.w;
        } else {
            c0 = color_texture_sample(a, u, v);
        }
        // combineColor0 (tile_fragment.inc.glsl:81-89)
        if (combine == 1) {
            float al = c0.w * col[3];
            col[0] = c0.x; col[1] = c0.y; col[2] = c0.z; col[3] = al;
        } else if (combine == 2) {
            col[3] = c0.w * col[3];
        }
    }
    col[3] *= mask_alpha;
    int op = (ctrl >> 10) & 0xf;
    if (op != 0) {
        // composite (tile_fragment.inc.glsl:508-523): the destination comes
        // from colour texture 0 at fragCoord / framebufferSize, the texture
        // d3d11/tile.cs.glsl:135-152 binds there
        float4 dv = color_texture_sample(a, fx / a.fb_width, fy / a.fb_height);
        float dest[4] = {dv.x, dv.y, dv.z, dv.w}, blended[3];
        composite_rgb(dest, col, op, blended);
        float sa = col[3];
#pragma unroll
        for (int k = 0; k < 3; k++)
            col[k] = ((sa * (1.0f - dest[3])) * col[k] + (sa * dest[3]) * blended[k]) +
                     (1.0f - sa) * dest[k];
        col[3] = 1.0f;
    }
    return make_float4(col[0] * col[3], col[1] * col[3], col[2] * col[3], col[3]);
}

// ------------------------------------------------------------------ tile

// tile.cs.glsl:71-163 with calculateColor/sampleMask (tile_fragment.inc.glsl:
// 539-614). One wavefront per framebuffer tile, lane (lx, ly) shades rows
// 4ly..4ly+3 of column lx. kPaints: some paint has a colour texture, a
// filter or a composite op other than Normal (paint_color); without it every
// paint is a solid colour composited src-over, and the kernel is built
// without paint_color's registers.
template <bool kPaints>
__global__ void __launch_bounds__(kBlock) k_tile(TileArgs a) {
    uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    uint32_t fbn = (uint32_t)(a.fb_width_tiles * a.fb_height_tiles);
    if (wave >= fbn) return;
    uint32_t lane = lane_id(), lx = lane & 15u, ly = lane >> 4;
    int32_t t = a.first_tile_map[wave];
    if (t < 0 && a.load) return;
    uint32_t gx = wave % (uint32_t)a.fb_width_tiles, gy = wave / (uint32_t)a.fb_width_tiles;
    size_t pitch = (size_t)a.fb_width_tiles * 16;
    uint32_t *px = a.framebuffer + (size_t)(gy * 16 + ly * 4) * pitch + gx * 16 + lx;
    float d[4][4];
#pragma unroll
    for (int sub = 0; sub < 4; sub++) {
        if (!a.load) {
#pragma unroll
            for (int c = 0; c < 4; c++) d[sub][c] = a.clear[c];
        } else {
            uint32_t v = px[sub * pitch];
#pragma unroll
            for (int c = 0; c < 4; c++) d[sub][c] = (float)((v >> (8 * c)) & 0xffu) / 255.0f;
        }
    }
    // the next tile's words load beside this tile's mask texel: inside the
    // loop the load is unconditional, at index 0 past the list's end (in
    // bounds: the loop runs only when tile_count > 0)
    bool valid = t >= 0 && (uint32_t)t < a.tile_count;
    uint2 w = make_uint2(0, 0);  // tile words 2-3
    uint32_t w_next = 0xffffffffu;
    if (valid) {
        w = a.tiles[t];
        w_next = a.tile_next[t];
    }
    // the slots of this list, for the adjoint
    uint32_t slot = 0, slot_end = 0;
    if (a.list_tiles) {
        slot = min(a.list_offsets[wave], a.list_capacity);
        slot_end = min(a.list_offsets[wave + 1], a.list_capacity);
    }
    for (uint32_t guard = 0; valid && guard < a.tile_count; guard++) {
        if (slot < slot_end) {
            if (lane == 0) a.list_tiles[slot] = (uint32_t)t;
            slot++;
        }
        int32_t tn = (int32_t)w_next;
        bool next_valid = tn >= 0 && (uint32_t)tn < a.tile_count;
        uint32_t tc = next_valid ? (uint32_t)tn : 0u;
        uint2 wn = a.tiles[tc];
        uint32_t wn_next = a.tile_next[tc];
        int32_t alpha = sext24(w.x);
        uint32_t color = w.y & 0xffffu;
        int32_t ctrl = (int32_t)((w.y >> 16) & 0xffu);
        float backdrop;
        if (alpha >= 0) {
            backdrop = 0.0f;
        } else {
            backdrop = (float)((int32_t)w.y >> 24);
            ctrl &= ~0x3;
        }
        // the paint's metadata: base colour (texel 2) and control word (texel 8)
        const float4 *md = color < a.paint_count ? a.metadata + (size_t)color * 10 : nullptr;
        float4 base = md ? md[2] : make_float4(0, 0, 0, 0);
        // no colour texture, Normal composite
        bool plain = !kPaints || !md || md[8].x == 0.0f;
        int32_t mask_ctrl = ctrl & 0x3;
        uint32_t texel = 0;
        if (mask_ctrl != 0 && (uint32_t)alpha < a.mask_capacity)
            texel = a.mask[(size_t)(uint32_t)alpha * 64 + lane];
#pragma unroll
        for (int sub = 0; sub < 4; sub++) {
            float mask_alpha = 1.0f;
            if (mask_ctrl != 0) {
                float cov = (float)((texel >> (8 * sub)) & 0xffu) / 255.0f + backdrop;
                if (mask_ctrl & 0x1) {
                    cov = fabsf(cov);
                } else {
                    float m = cov - 2.0f * floorf(cov / 2.0f);
                    cov = 1.0f - fabsf(1.0f - m);
                }
                mask_alpha = min_glsl(mask_alpha, cov);
            }
            if (plain) {
                float al = base.w * mask_alpha;
                float ia = 1.0f - al;
                d[sub][0] = d[sub][0] * ia + base.x * al;
                d[sub][1] = d[sub][1] * ia + base.y * al;
                d[sub][2] = d[sub][2] * ia + base.z * al;
                d[sub][3] = d[sub][3] * ia + al;
            } else if constexpr (kPaints) {
                float fx = (float)(gx * 16 + lx) + 0.5f;
                float fy = (float)(gy * 16 + ly * 4 + (uint32_t)sub) + 0.5f;
                float4 c = paint_color(a, md, fx, fy, mask_alpha);
                float ia = 1.0f - c.w;
                d[sub][0] = d[sub][0] * ia + c.x;
                d[sub][1] = d[sub][1] * ia + c.y;
                d[sub][2] = d[sub][2] * ia + c.z;
                d[sub][3] = d[sub][3] * ia + c.w;
            }
        }
        w = wn;
        w_next = wn_next;
        valid = next_valid;
        t = tn;
    }
    for (; slot < slot_end; slot++)  // a truncated list: no layer
        if (lane == 0) a.list_tiles[slot] = 0xffffffffu;
#pragma unroll
    for (int sub = 0; sub < 4; sub++)
        px[sub * pitch] = unorm8(d[sub][0]) | (unorm8(d[sub][1]) << 8) |
                          (unorm8(d[sub][2]) << 16) | (unorm8(d[sub][3]) << 24);
}

// ------------------------------------------------------------------ adjoint
//
// The gradient of the rendered image with respect to the scene's points and
// paint colours, for one draw batch composited over a clear colour. It is the
// derivative of the continuous model the kernels evaluate: per path and pixel
// the signed integral C of the winding number over the pixel (exact area, of
// which the LUT coverage sum is the fp32 approximation), mask m = clamp(|C|,
// 0, 1) (min with the clip texel), and the src-over composite of k_tile. The
// quantisations on the way (8.8 fixed point, unorm8 masks and pixels, f16
// paints) pass gradients through unchanged.

__global__ void k_fill_zero(float *dst, size_t count) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        dst[i] = 0.0f;
}

__global__ void k_list_lengths(const uint4 *list_stats, uint32_t *lengths, uint32_t fb_tiles) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < fb_tiles) lengths[i] = list_stats[i].y;
}

// The sum of v over the 64 lanes of a wavefront (every lane active), in an
// SGPR: within each row of 16 lanes v += v of the lane 1, 2, 4, 8 below
// (row_shr), then lane 15's sum is added to rows 1 and 3 (row_bcast:15) and
// lane 31's to rows 2 and 3 (row_bcast:31); lane 63 holds the total. Register
// moves between lanes, no LDS round trips (a butterfly of lane shuffles made
// the tile adjoint three times slower). The order of the additions is fixed.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ float dpp_or_zero_f(float x) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(x), kCtrl, kRowMask, 0xf, false));
}
__device__ __forceinline__ float wave_sum(float v) {
    v += dpp_or_zero_f<0x111, 0xf>(v);
    v += dpp_or_zero_f<0x112, 0xf>(v);
    v += dpp_or_zero_f<0x114, 0xf>(v);
    v += dpp_or_zero_f<0x118, 0xf>(v);
    v += dpp_or_zero_f<0x142, 0xa>(v);
    v += dpp_or_zero_f<0x143, 0xc>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// One layer of a framebuffer tile's list as k_tile reads it: the tile's words
// 2-3 and its paint's base colour (the same in every lane) and this lane's
// mask texel; for the gradients also its alpha tile's clip texel, fill
// state or signed sums, and its draw path.
struct LayerIn {
    uint2 w;
    float4 base;
    uint32_t texel;
    int32_t clip;       // clip alpha tile of the layer's alpha tile, or -1
    uint32_t ctexel;    // this lane's texel of it
    uint32_t cstate;    // this lane's 2-bit fill states (cov_state)
    float C;            // this lane's signed coverage sum in sub-row sub (coverage checkpoint)
    float T;            // this lane's T_k in sub-row sub (walk 2's output)
    uint32_t draw_path;
};

// Loads the layers of list slots k0 .. k0 + kLayerGroup - 1 (below n) in a
// few rounds of loads for the whole group (the slots' tiles, their words,
// then paints, mask texels and the rest) instead of dependent loads per
// layer. Values the same in every lane are read as such (scalar loads), so
// the vector stores and atomics of the layers before do not hold them up.
constexpr uint32_t kLayerGroup = 4;
template <bool kGrad>
__device__ __forceinline__ void load_layers(const TileGradArgs &g, uint32_t off, uint32_t k0,
                                            uint32_t n, uint32_t sub, uint32_t lane,
                                            uint32_t tt[], LayerIn L[]) {
    const TileArgs &a = g.f;
#pragma unroll
    for (uint32_t j = 0; j < kLayerGroup; j++)
        tt[j] = k0 + j < n ? uniform(g.list_tiles[off + k0 + j]) : 0u;
#pragma unroll
    for (uint32_t j = 0; j < kLayerGroup; j++)
        L[j].w = k0 + j < n ? a.tiles[tt[j]] : make_uint2(0xffffffffu, 0u);
#pragma unroll
    for (uint32_t j = 0; j < kLayerGroup; j++) {
        uint32_t color = uniform(L[j].w.y & 0xffffu);
        int32_t alpha = (int32_t)uniform((uint32_t)sext24(L[j].w.x));
        int32_t ctrl = (int32_t)((L[j].w.y >> 16) & 0xffu);
        if (alpha < 0) ctrl &= ~0x3;
        L[j].base = color < a.paint_count ? a.metadata[(size_t)color * 10 + 2]
                                          : make_float4(0, 0, 0, 0);
        L[j].texel = ((ctrl & 0x3) != 0 && (uint32_t)alpha < a.mask_capacity)
                         ? a.mask[(size_t)(uint32_t)alpha * 64 + lane] : 0u;
        if (kGrad) {
            bool alpha_tile = alpha >= 0 && (uint32_t)alpha >= g.alpha_start &&
                              (uint32_t)alpha - g.alpha_start < g.alpha_count;
            L[j].clip = alpha_tile
                            ? (int32_t)g.alpha_tiles[(uint32_t)alpha - g.alpha_start].y : -1;
            L[j].ctexel = (L[j].clip >= 0 && (uint32_t)L[j].clip < a.mask_capacity)
                              ? a.mask[(size_t)(uint32_t)L[j].clip * 64 + lane] : 0u;
            L[j].cstate = (alpha_tile && g.cov_state)
                              ? g.cov_state[(size_t)(uint32_t)alpha * 64 + lane] : 0u;
            L[j].C = (alpha_tile && !g.cov_state)
                         ? g.coverage[(size_t)(uint32_t)alpha * 256 + sub * 64 + lane] : 0.0f;
            uint32_t bp = k0 + j < n ? g.tile_paths[tt[j]] : 0xffffffffu;
            L[j].draw_path = bp < g.path_count ? g.dice_md[bp].x : 0xffffffffu;
            L[j].T = g.trans[(size_t)(off + min(k0 + j, n - 1u)) * 256 + sub * 64 + lane];
        }
    }
}

// The layer's alpha tile and its mask value in sub-row `sub` of this lane's
// texel, as k_tile computes them.
__device__ __forceinline__ int32_t layer_mask(const LayerIn &L, uint32_t sub, float &m) {
    int32_t alpha = sext24(L.w.x);
    int32_t ctrl = (int32_t)((L.w.y >> 16) & 0xffu);
    float backdrop;
    if (alpha >= 0) {
        backdrop = 0.0f;
    } else {
        backdrop = (float)((int32_t)L.w.y >> 24);
        ctrl &= ~0x3;
    }
    int32_t mask_ctrl = ctrl & 0x3;
    m = 1.0f;
    if (mask_ctrl != 0) {
        float cov = (float)((L.texel >> (8 * sub)) & 0xffu) / 255.0f + backdrop;
        if (mask_ctrl & 0x1) {
            cov = fabsf(cov);
        } else {
            float q = cov - 2.0f * floorf(cov / 2.0f);
            cov = 1.0f - fabsf(1.0f - q);
        }
        m = min_glsl(m, cov);
    }
    return alpha;
}

// The composite of k_tile, D_k = D_(k-1) (1 - a_k) + (rgb_k a_k, a_k) with
// a_k = alpha_k m_k over the list's layers k = 0..n-1, differentiated: with
// g = dL/dD_n and T_k = prod_(j>k) (1 - a_j),
//   dL/da_k   = T_k (g . ((rgb_k, 1) - D_(k-1))),
//   dL/drgb_k = g_rgb T_k a_k,   dL/dalpha_k += dL/da_k m_k   (per draw path),
//   dL/dm_k   = dL/da_k alpha_k,
// and for an alpha tile dL/dC = dL/dm sign(C) when 0 < |C| < 1 and the clip
// texel does not limit m (0 otherwise). A workgroup per framebuffer tile, a
// wavefront per sub-row: lane (lx, ly) of wavefront s owns pixel (lx, 4 ly +
// s), channel s of mask texel (lx, ly). Wavefront 0 follows the list once,
// storing each slot's tile; then every wavefront runs T_k (top down, with
// a_k) and the gradients (bottom up, D recomputed) over the slots
// kLayerGroup at a time, loading a group's layers together since their
// loads do not depend on each other. Four wavefronts per tile each carry a
// quarter of the per-layer work, so a long list (the slowest tile decides
// the launch) takes a quarter of the time. Each alpha tile written gets this
// pass's epoch in grad_stamp, so alpha tiles no list composited read as zero
// without clearing grad_cov. A layer's colour gradient (a quarter per
// wavefront) goes to its slot, for k_slot_colors.
__global__ void __launch_bounds__(kBlock) k_tile_backward(TileGradArgs g) {
    __shared__ uint32_t s_n;
    const TileArgs &a = g.f;
    uint32_t wave = blockIdx.x;  // the framebuffer tile
    uint32_t fbn = (uint32_t)(a.fb_width_tiles * a.fb_height_tiles);
    if (wave >= fbn) return;
    uint32_t sub = threadIdx.x >> 6, lane = lane_id(), lx = lane & 15u, ly = lane >> 4;
    uint32_t off = g.list_offsets[wave], end = g.list_offsets[wave + 1];
    if (end > g.slot_capacity) end = g.slot_capacity;
    uint32_t gx = wave % (uint32_t)a.fb_width_tiles, gy = wave / (uint32_t)a.fb_width_tiles;
    uint32_t px = gx * 16 + lx, py = gy * 16 + ly * 4 + sub;
    float4 q = (px < g.width && py < g.height)
                   ? ((const float4 *)g.grad_image)[(size_t)py * g.width + px]
                   : make_float4(0.0f, 0.0f, 0.0f, 0.0f);

    // walk 1 (wavefront 0): each slot's tile; one dependent load per layer,
    // unless the forward composite recorded them (then 64 slots per load,
    // up to the first that holds no layer)
    if (sub == 0) {
        uint32_t n = 0;
        if (g.lists_recorded) {
            for (uint32_t k0 = off; k0 < end; k0 += 64u) {
                uint32_t k = k0 + lane;
                uint32_t t = k < end ? g.list_tiles[k] : 0xffffffffu;
                unsigned long long none = __ballot(t >= a.tile_count);
                if (none) {
                    n = k0 - off + (uint32_t)__ffsll((long long)none) - 1u;
                    break;
                }
                n = k0 + 64u - off;
            }
            n = min(n, end - off);
        } else {
            int32_t t = (int32_t)uniform((uint32_t)a.first_tile_map[wave]);
            for (uint32_t guard = 0;
                 t >= 0 && (uint32_t)t < a.tile_count && guard < a.tile_count && off + n < end;
                 guard++) {
                g.list_tiles[off + n] = (uint32_t)t;
                n++;
                t = (int32_t)uniform(a.tile_next[t]);
            }
        }
        // slots the chase did not reach (a truncated list) hold no layer
        for (uint32_t k = off + n; k < end; k++)
            if (lane == 4u) g.trans[(size_t)k * 256 + 4] = __uint_as_float(0xffffffffu);
        if (lane == 0) s_n = n;
    }
    __syncthreads();
    uint32_t n = s_n;
    float *tr = g.trans + sub * 64 + lane;  // this lane's T_k of slot k at tr[256 (off + k)]

    // walk 2: a_k and T_k, top down (each lane reads back only what it wrote)
    float T = 1.0f;
    for (uint32_t kg = n; kg > 0;) {
        uint32_t k0 = kg >= kLayerGroup ? kg - kLayerGroup : 0u;
        uint32_t tt[kLayerGroup];
        LayerIn L[kLayerGroup];
        load_layers<false>(g, off, k0, kg, sub, lane, tt, L);
#pragma unroll
        for (uint32_t jj = kLayerGroup; jj-- > 0;) {
            uint32_t k = k0 + jj;
            if (k >= kg) continue;
            float m;
            (void)layer_mask(L[jj], sub, m);
            tr[(size_t)(off + k) * 256] = T;
            T *= 1.0f - L[jj].base.w * m;
        }
        kg = k0;
    }
    // walk 3: gradients, bottom up
    float D[4] = {a.clear[0], a.clear[1], a.clear[2], a.clear[3]};
    for (uint32_t k0 = 0; k0 < n; k0 += kLayerGroup) {
        uint32_t tt[kLayerGroup];
        LayerIn L[kLayerGroup];
        load_layers<true>(g, off, k0, n, sub, lane, tt, L);
#pragma unroll
        for (uint32_t jj = 0; jj < kLayerGroup; jj++) {
            uint32_t k = k0 + jj;
            if (k >= n) break;
            float m;
            int32_t alpha = layer_mask(L[jj], sub, m);
            float4 base = L[jj].base;
            bool alpha_tile = alpha >= 0 && (uint32_t)alpha >= g.alpha_start &&
                              (uint32_t)alpha - g.alpha_start < g.alpha_count;
            float Tk = L[jj].T;
            float al = base.w * m;
            float dal = Tk * (q.x * (base.x - D[0]) + q.y * (base.y - D[1]) +
                              q.z * (base.z - D[2]) + q.w * (1.0f - D[3]));
            float gr = q.x * Tk * al, gg = q.y * Tk * al, gb = q.z * Tk * al;
            float ga = dal * m;
            if (alpha_tile) {
                float dC = 0.0f;
                if (g.cov_state) {
                    uint32_t st = L[jj].cstate >> (2 * sub);
                    if (st & 1u) dC = dal * base.w * ((st & 2u) ? -1.0f : 1.0f);
                } else {
                    float C = L[jj].C;
                    float aC = fabsf(C);
                    bool limited = false;
                    if (L[jj].clip >= 0) {
                        float cv = (float)((L[jj].ctexel >> (8 * sub)) & 0xffu) / 255.0f;
                        limited = cv < min(aC, 1.0f);
                    }
                    if (aC > 0.0f && aC < 1.0f && !limited)
                        dC = dal * base.w * (C < 0.0f ? -1.0f : 1.0f);
                }
                g.grad_cov[(size_t)(uint32_t)alpha * 256 + sub * 64 + lane] = dC;
                if (lane == 0) g.grad_stamp[(uint32_t)alpha] = g.epoch;
            }
            float ia = 1.0f - al;
            D[0] = D[0] * ia + base.x * al;
            D[1] = D[1] * ia + base.y * al;
            D[2] = D[2] * ia + base.z * al;
            D[3] = D[3] * ia + al;
            gr = wave_sum(gr);
            gg = wave_sum(gg);
            gb = wave_sum(gb);
            ga = wave_sum(ga);
            // this wavefront's quarter of the layer's colour gradient, and the
            // draw path (of the tile's path, which k_bound recorded), to words
            // 0-4 of the slot's sub-row, which walk 3 has read
            float *out = g.trans + (size_t)(off + k) * 256 + sub * 64;
            if (lane < 5u)
                out[lane] = lane == 0u ? gr : lane == 1u ? gg : lane == 2u ? gb
                          : lane == 3u ? ga : __uint_as_float(L[jj].draw_path);
        }
    }
}

// The list slots' colour gradients (k_tile_backward's four sub-row records
// of each slot) into their draw paths' RGBA in grad_colors, a thread per
// slot. (Summing them in one workgroup's LDS first was four times slower.)
__device__ __forceinline__ bool slot_colors(const TileGradArgs &g, uint32_t s, uint32_t &dp,
                                            float4 &c) {
    const float *w = g.trans + (size_t)s * 256;
    dp = __float_as_uint(w[4]);
    c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (uint32_t sub = 0; sub < 4; sub++) {
        c.x += w[sub * 64 + 0];
        c.y += w[sub * 64 + 1];
        c.z += w[sub * 64 + 2];
        c.w += w[sub * 64 + 3];
    }
    return dp < g.color_count;
}
__global__ void __launch_bounds__(kBlock) k_slot_colors(TileGradArgs g) {
    uint32_t fbn = (uint32_t)(g.f.fb_width_tiles * g.f.fb_height_tiles);
    uint32_t total = min(g.list_offsets[fbn], g.slot_capacity);
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < total;
         s += gridDim.x * blockDim.x) {
        uint32_t dp;
        float4 c;
        if (slot_colors(g, s, dp, c)) {
            atomicAdd(&g.grad_colors[dp * 4 + 0], c.x);
            atomicAdd(&g.grad_colors[dp * 4 + 1], c.y);
            atomicAdd(&g.grad_colors[dp * 4 + 2], c.z);
            atomicAdd(&g.grad_colors[dp * 4 + 3], c.w);
        }
    }
}

// The derivative of C (the winding integral over a pixel) with respect to the
// end points of one directed edge e = p1 - p0 of a closed outline: moving the
// edge sweeps area with winding +-1, so dC/dp0 = (-e.y, e.x) W0 and dC/dp1 =
// (-e.y, e.x) W1 with W0 = integral of (1 - t) dt and W1 = integral of t dt
// over the parameters t in [ta, tb] where the edge lies inside the pixel
// (image coordinates, y down; edges running towards -x add the area below
// them). Returns W0, W1 summed over the pixels the edge crosses, each
// weighted by dL/dC of that pixel.
__device__ float2 microline_pixel_weights(const MicrolineGradArgs &g, uint32_t path, float x0,
                                          float y0, float x1, float y1) {
    const int32_t *pm = g.b.prop_md + (size_t)path * 12;
    int32_t rx0 = pm[0], ry0 = pm[1], rx1 = pm[2], ry1 = pm[3];
    uint32_t tile_base = (uint32_t)pm[4];
    int32_t rw = rx1 - rx0;
    float dx = x1 - x0, dy = y1 - y0;
    int32_t cx = (int32_t)floorf(x0), cy = (int32_t)floorf(y0);
    int32_t ex = (int32_t)floorf(x1), ey = (int32_t)floorf(y1);
    int32_t sx = dx > 0.0f ? 1 : -1, sy = dy > 0.0f ? 1 : -1;
    const float inf = __builtin_inff();
    float tdx = dx != 0.0f ? fabsf(1.0f / dx) : inf;
    float tdy = dy != 0.0f ? fabsf(1.0f / dy) : inf;
    float tmx = dx > 0.0f ? ((float)(cx + 1) - x0) * tdx : (dx < 0.0f ? (x0 - (float)cx) * tdx : inf);
    float tmy = dy > 0.0f ? ((float)(cy + 1) - y0) * tdy : (dy < 0.0f ? (y0 - (float)cy) * tdy : inf);
    float w0 = 0.0f, w1 = 0.0f, ta = 0.0f;
    int32_t steps = abs(ex - cx) + abs(ey - cy) + 2;
    for (int32_t it = 0; it < steps && it < 4096; it++) {
        float tb = min(min(tmx, tmy), 1.0f);
        if (tb > ta && cx >= 0 && cy >= 0) {
            int32_t tx = cx >> 4, ty = cy >> 4;
            if (tx >= rx0 && tx < rx1 && ty >= ry0 && ty < ry1) {
                uint32_t tile = tile_base + (uint32_t)((ty - ry0) * rw + (tx - rx0));
                if (tile < g.b.tile_count) {
                    int32_t alpha = sext24(g.b.tiles[tile].x);
                    if (alpha >= 0 && (uint32_t)alpha < g.alpha_capacity &&
                        g.grad_stamp[(uint32_t)alpha] == g.epoch) {
                        uint32_t px = (uint32_t)(cx & 15), py = (uint32_t)(cy & 15);
                        float gc = g.grad_cov[(size_t)(uint32_t)alpha * 256 + (py & 3u) * 64 +
                                              (py >> 2) * 16 + px];
                        float sq = 0.5f * (tb * tb - ta * ta);
                        w0 += gc * ((tb - ta) - sq);
                        w1 += gc * sq;
                    }
                }
            }
        }
        if (tb >= 1.0f) break;
        ta = tb;
        if (tmx < tmy) {
            cx += sx;
            tmx += tdx;
        } else {
            cy += sy;
            tmy += tdy;
        }
    }
    return make_float2(w0, w1);
}


// One microline per thread: its segment (the last with seg_offsets <= m), the
// unquantised end points as k_dice_emit computes them, the pixel weights, and
// the chain rule through de Casteljau (Bernstein weights at t = i / n and
// (i + 1) / n), the quadratic-to-cubic elevation and the transform.
// The gradients of microline m's end points carried to up to four scene
// points (index ~0: none), for k_microline_backward to add.
struct PointGrads {
    uint32_t i0, i1, i2, i3;  // scene point indices, ~0: none
    float2 g0, g1, g2, g3;
};

__device__ __forceinline__ PointGrads microline_point_grads(const MicrolineGradArgs &g,
                                                            uint32_t m) {
    const BatchArgs &a = g.b;
    PointGrads pgr = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, {}, {}, {}, {}};
    {
        uint32_t lo = 0, hi = a.segment_count + 1;  // seg_offsets[lo] <= m < seg_offsets[hi]
        while (lo + 1 < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (a.seg_offsets[mid] <= m) lo = mid;
            else hi = mid;
        }
        uint32_t seg = lo;
        if (seg >= a.segment_count) return pgr;
        uint32_t path = a.microlines[m].w;
        if (path >= a.path_count) return pgr;
        DiceSeg d;
        if (!dice_setup(a, seg, path, d)) return pgr;
        int32_t i = (int32_t)(m - a.seg_offsets[seg]);
        if (i >= d.n) return pgr;
        auto point_at = [&](int32_t k) {
            if (k == 0) return d.p0;
            float t = (float)k / (float)d.n;
            if (d.curve) {
                float2 p01 = make_float2(mixf(d.p0.x, d.c0.x, t), mixf(d.p0.y, d.c0.y, t));
                float2 p12 = make_float2(mixf(d.c0.x, d.c1.x, t), mixf(d.c0.y, d.c1.y, t));
                float2 p23 = make_float2(mixf(d.c1.x, d.p3.x, t), mixf(d.c1.y, d.p3.y, t));
                float2 p012 = make_float2(mixf(p01.x, p12.x, t), mixf(p01.y, p12.y, t));
                float2 p123 = make_float2(mixf(p12.x, p23.x, t), mixf(p12.y, p23.y, t));
                return make_float2(mixf(p012.x, p123.x, t), mixf(p012.y, p123.y, t));
            }
            return make_float2(mixf(d.p0.x, d.p3.x, t), mixf(d.p0.y, d.p3.y, t));
        };
        float2 q0 = point_at(i), q1 = point_at(i + 1);
        float2 w = microline_pixel_weights(g, path, q0.x, q0.y, q1.x, q1.y);
        if (w.x == 0.0f && w.y == 0.0f) return pgr;
        float2 nrm = make_float2(-(q1.y - q0.y), q1.x - q0.x);
        float2 g0 = make_float2(nrm.x * w.x, nrm.y * w.x);  // dL/dq0
        float2 g1 = make_float2(nrm.x * w.y, nrm.y * w.y);  // dL/dq1
        // Bernstein weights of the control points at both ends
        float ta = (float)i / (float)d.n, tb = (float)(i + 1) / (float)d.n;
        uint4 md = a.dice_md[path];
        uint2 ix = a.scene.indices[seg - md.z + md.y];
        uint32_t from = ix.x, flags = ix.y;
        if (d.curve) {
            float ua = 1.0f - ta, ub = 1.0f - tb;
            float ba[4] = {ua * ua * ua, 3.0f * ta * ua * ua, 3.0f * ta * ta * ua, ta * ta * ta};
            float bb[4] = {ub * ub * ub, 3.0f * tb * ub * ub, 3.0f * tb * tb * ub, tb * tb * tb};
            float2 gc[4];
#pragma unroll
            for (int j = 0; j < 4; j++)
                gc[j] = make_float2(ba[j] * g0.x + bb[j] * g1.x, ba[j] * g0.y + bb[j] * g1.y);
            if (flags & kFlagQuad) {
                // c0 = (p0 + 2k) / 3, c1 = (p3 + 2k) / 3
                const float third = 1.0f / 3.0f, two_thirds = 2.0f / 3.0f;
                pgr.i0 = from;
                pgr.g0 = make_float2(gc[0].x + gc[1].x * third, gc[0].y + gc[1].y * third);
                pgr.i1 = from + 1;
                pgr.g1 = make_float2((gc[1].x + gc[2].x) * two_thirds,
                                     (gc[1].y + gc[2].y) * two_thirds);
                pgr.i2 = from + 2;
                pgr.g2 = make_float2(gc[3].x + gc[2].x * third, gc[3].y + gc[2].y * third);
            } else {
                pgr.i0 = from;
                pgr.g0 = gc[0];
                pgr.i1 = from + 1;
                pgr.g1 = gc[1];
                pgr.i2 = from + 2;
                pgr.g2 = gc[2];
                pgr.i3 = from + 3;
                pgr.g3 = gc[3];
            }
        } else {
            pgr.i0 = from;
            pgr.g0 = make_float2((1.0f - ta) * g0.x + (1.0f - tb) * g1.x,
                                 (1.0f - ta) * g0.y + (1.0f - tb) * g1.y);
            pgr.i1 = from + 1;
            pgr.g1 = make_float2(ta * g0.x + tb * g1.x, ta * g0.y + tb * g1.y);
        }
    }
    return pgr;
}

// One microline per thread. The microlines of a segment are consecutive and
// send their gradients to the same few points, so each wavefront first sums
// the contributions it has for one point (a ballot per distinct point) and
// adds that sum with one atomic per coordinate, instead of one per lane:
// far fewer atomics on the same addresses. (Float sums in another order: the
// gradient is compared at a tolerance, not bit for bit.)
__global__ void __launch_bounds__(kBlock) k_microline_backward(MicrolineGradArgs g) {
    uint32_t stride = gridDim.x * blockDim.x, lane = lane_id();
    const float *xf = g.b.xform;
    for (uint32_t base = blockIdx.x * blockDim.x; base < g.microline_count; base += stride) {
        uint32_t m = base + threadIdx.x;
        PointGrads pgr = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, {}, {}, {}, {}};
        if (m < g.microline_count) pgr = microline_point_grads(g, m);
        auto add = [&](uint32_t index, float2 gq) {
            // dL/dp = M^T dL/dq for q = M p + t (get_point)
            bool valid = index < g.b.scene.point_count;
            float gx = valid ? xf[0] * gq.x + xf[2] * gq.y : 0.0f;
            float gy = valid ? xf[1] * gq.x + xf[3] * gq.y : 0.0f;
            unsigned long long left = __ballot(valid);
            while (left) {
                uint32_t leader = (uint32_t)__ffsll((long long)left) - 1u;
                uint32_t key = (uint32_t)__builtin_amdgcn_readlane((int)index, (int)leader);
                bool mine = valid && index == key;
                left &= ~__ballot(mine);
                float sx = wave_sum(mine ? gx : 0.0f), sy = wave_sum(mine ? gy : 0.0f);
                if (lane == leader) {
                    atomicAdd(&g.grad_points[2 * key + 0], sx);
                    atomicAdd(&g.grad_points[2 * key + 1], sy);
                }
            }
        };
        add(pgr.i0, pgr.g0);
        add(pgr.i1, pgr.g1);
        add(pgr.i2, pgr.g2);
        add(pgr.i3, pgr.g3);
    }
}

// RGBA8 framebuffer rows to floats (byte / 255), width x height x 4.
__global__ void k_framebuffer_to_float(const uint32_t *fb, uint32_t pitch_words, uint32_t width,
                                       uint32_t height, float *dst) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= width * height) return;
    uint32_t x = i % width, y = i / width;
    uint32_t v = fb[(size_t)y * pitch_words + x];
    float4 o = make_float4((float)(v & 0xffu) / 255.0f, (float)((v >> 8) & 0xffu) / 255.0f,
                           (float)((v >> 16) & 0xffu) / 255.0f, (float)(v >> 24) / 255.0f);
    ((float4 *)dst)[i] = o;
}

// ------------------------------------------------------------------ tables

// The padded LUT of lut_blend: padded texel (j, i) of the 259 x 259 table (row
// stride kLutStride) is byte texel (clamp(j - 1), clamp(i - 1)) / 255, and
// beside it the difference to padded texel (j, i + 1), both in float.
__device__ __forceinline__ float4 padded_texel(const uint8_t *rgba, uint32_t jp, uint32_t ip) {
    uint32_t j = (uint32_t)min(max((int32_t)jp - 1, 0), 255);
    uint32_t k = (uint32_t)min(max((int32_t)ip - 1, 0), 255);
    const uint8_t *t = rgba + (j * 256u + k) * 4u;
    return make_float4((float)t[0] / 255.0f, (float)t[1] / 255.0f, (float)t[2] / 255.0f,
                       (float)t[3] / 255.0f);
}
__global__ void k_lut_expand(const uint8_t *rgba, float4 *lut) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kLutRows * kLutStride) return;
    uint32_t jp = i / kLutStride, ip = i % kLutStride;
    float4 t = padded_texel(rgba, jp, ip), n = padded_texel(rgba, jp, ip + 1u);
    lut[2u * i] = t;
    lut[2u * i + 1u] = make_float4(n.x - t.x, n.y - t.y, n.z - t.z, n.w - t.w);
}

// ColorU::to_f32 (color/src/lib.rs:70-73) then f16::from_f32 (gpu/renderer.rs:724-727).
__global__ void k_metadata_expand(const uint16_t *halves, float4 *metadata, uint32_t texels) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= texels) return;
    const __half *h = reinterpret_cast<const __half *>(halves) + (size_t)i * 4;
    metadata[i] = make_float4(__half2float(h[0]), __half2float(h[1]), __half2float(h[2]),
                              __half2float(h[3]));
}

// ------------------------------------------------------------------ D3D9 level
//
// The raster half of RendererD3D9 (gpu/d3d9/renderer.rs:150-472) on compute:
// d3d9/fill.{vs,fs}.glsl draws each fill as a quad over pixel columns
// floor(min x) .. ceil(max x) - 1 and mask texel rows floor(floor(min y) / 4)
// .. 3 of its alpha tile, and additive blending into the RGBA16F mask adds
// each fragment's computeCoverage to the texel, rounding the sum to binary16
// in fill order. One wavefront per alpha tile: lane (x, r) owns mask texel
// r * 16 + x (pixel column x, pixel rows 4r .. 4r + 3 in its four channels)
// and walks the tile's fills in order.

__device__ __forceinline__ float h2f(uint16_t h) { return __half2float(__ushort_as_half(h)); }
__device__ __forceinline__ uint16_t f2h(float f) { return __half_as_ushort(__float2half_rn(f)); }

__global__ void __launch_bounds__(kBlock) k_d3d9_fill(D3D9FillArgs a) {
    uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wave >= a.job_count) return;
    uint32_t lane = lane_id(), x = lane & 15u, r = lane >> 4;
    uint4 job = a.jobs[wave];  // alpha tile, first fill, end
    if (job.x >= a.alpha_capacity) return;
    LutRef lut = lut_ref(a.lut);
    uint2 *texel = reinterpret_cast<uint2 *>(a.mask + (size_t)job.x * 256u) + lane;
    uint2 w = *texel;
    uint16_t m[4] = {(uint16_t)(w.x & 0xffffu), (uint16_t)(w.x >> 16), (uint16_t)(w.y & 0xffffu),
                     (uint16_t)(w.y >> 16)};
    float fragx = (float)x + 0.5f, fragy = (float)(4u * r) + 0.5f;
    for (uint32_t k = job.y; k < job.z; k++) {
        uint2 fw = a.fills[k];
        FillRecord f;
        FillColumns c = make_record(fw, f);
        float miny = fminf((float)(fw.x >> 16), (float)(fw.y >> 16)) * (1.0f / 256.0f);
        int32_t r0 = (int32_t)floorf(floorf(miny) * 0.25f);
        if (x < c.c0 || x > c.c1 || (int32_t)r < r0) continue;
        float4 cov = coverage_record(lut, f, fragx, fragy);
        m[0] = f2h(h2f(m[0]) + cov.x);
        m[1] = f2h(h2f(m[1]) + cov.y);
        m[2] = f2h(h2f(m[2]) + cov.z);
        m[3] = f2h(h2f(m[3]) + cov.w);
    }
    *texel = make_uint2((uint32_t)m[0] | ((uint32_t)m[1] << 16),
                        (uint32_t)m[2] | ((uint32_t)m[3] << 16));
}

// tile_clip_copy + tile_clip_combine (d3d9/tile_clip_*.glsl): dest texel =
// min(|dest + dest backdrop|, |src + src backdrop|). A draw path's alpha tile
// is the dest of at most one clip and no clip path's tile is a dest, so the
// sources read here are the values before any combine. One thread per
// (clip, texel).
__global__ void __launch_bounds__(kBlock) k_d3d9_clip(D3D9ClipArgs a) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.clip_count * 64u) return;
    int4 c = a.clips[i >> 6];
    uint32_t texel = i & 63u;
    uint32_t dest = (uint32_t)c.x, src = (uint32_t)c.z;
    if (dest >= a.alpha_capacity || src >= a.alpha_capacity) return;
    uint2 *d = reinterpret_cast<uint2 *>(a.mask + (size_t)dest * 256u) + texel;
    const uint2 *sp = reinterpret_cast<const uint2 *>(a.mask + (size_t)src * 256u) + texel;
    uint2 dw = *d, sw = *sp;
    uint16_t dh[4] = {(uint16_t)(dw.x & 0xffffu), (uint16_t)(dw.x >> 16),
                      (uint16_t)(dw.y & 0xffffu), (uint16_t)(dw.y >> 16)};
    uint16_t sh[4] = {(uint16_t)(sw.x & 0xffffu), (uint16_t)(sw.x >> 16),
                      (uint16_t)(sw.y & 0xffffu), (uint16_t)(sw.y >> 16)};
    uint16_t o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float dv = fabsf(h2f(dh[k]) + (float)c.y);
        float sv = fabsf(h2f(sh[k]) + (float)c.w);
        o[k] = f2h(sv < dv ? sv : dv);
    }
    *d = make_uint2((uint32_t)o[0] | ((uint32_t)o[1] << 16), (uint32_t)o[2] | ((uint32_t)o[3] << 16));
}

// d3d9/tile.{vs,fs}.glsl with calculateColor for solid paints: the batch's
// tiles over one framebuffer tile, in draw order (z-culled tiles already
// left out of the list), each blended src-over into the RGBA8 target and
// rounded to 8 bits before the next. One wavefront per framebuffer tile,
// lane (x, r) the pixels of column x, rows 4r .. 4r + 3 (its mask texel).
__global__ void __launch_bounds__(kBlock) k_d3d9_tile(D3D9TileArgs a) {
    uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    uint32_t fbn = (uint32_t)(a.fb_width_tiles * a.fb_height_tiles);
    if (wave >= fbn) return;
    uint32_t begin = a.list_offsets[wave], end = a.list_offsets[wave + 1];
    if (a.load && begin == end) return;
    uint32_t lane = lane_id(), x = lane & 15u, r = lane >> 4;
    uint32_t gx = wave % (uint32_t)a.fb_width_tiles, gy = wave / (uint32_t)a.fb_width_tiles;
    size_t pitch = (size_t)a.fb_width_tiles * 16u;
    uint32_t *px = a.framebuffer + (size_t)(gy * 16u + r * 4u) * pitch + gx * 16u + x;
    uint32_t d8[4];
#pragma unroll
    for (int sub = 0; sub < 4; sub++) {
        if (a.load) {
            d8[sub] = px[sub * pitch];
        } else {
            d8[sub] = unorm8(a.clear[0]) | (unorm8(a.clear[1]) << 8) | (unorm8(a.clear[2]) << 16) |
                      (unorm8(a.clear[3]) << 24);
        }
    }
    for (uint32_t e = begin; e < end; e++) {
        uint4 t = a.tiles[a.list[e]];  // TileObjectPrimitive words
        uint32_t alpha = t.y;
        uint32_t color = t.w & 0xffffu;
        bool even_odd = ((t.w >> 16) & 0x2u) != 0;
        float backdrop = (float)((int32_t)t.w >> 24);
        uint32_t paint = color < a.paint_count ? color : 0u;
        float4 base = a.paint_count ? a.metadata[(size_t)paint * 10u + 2u]
                                    : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        uint2 mw = make_uint2(0u, 0u);
        if (alpha < a.alpha_capacity)
            mw = reinterpret_cast<const uint2 *>(a.mask + (size_t)alpha * 256u)[lane];
        uint16_t mh[4] = {(uint16_t)(mw.x & 0xffffu), (uint16_t)(mw.x >> 16),
                          (uint16_t)(mw.y & 0xffffu), (uint16_t)(mw.y >> 16)};
#pragma unroll
        for (int sub = 0; sub < 4; sub++) {
            float cov = h2f(mh[sub]) + backdrop;
            if (!even_odd) {
                cov = fabsf(cov);
            } else {
                float m = cov - 2.0f * floorf(cov / 2.0f);
                cov = 1.0f - fabsf(1.0f - m);
            }
            float mask_alpha = cov < 1.0f ? cov : 1.0f;
            float al = base.w * mask_alpha;
            float ia = 1.0f - al;
            uint32_t v = d8[sub];
            float d0 = (float)(v & 0xffu) / 255.0f, d1 = (float)((v >> 8) & 0xffu) / 255.0f;
            float d2 = (float)((v >> 16) & 0xffu) / 255.0f, d3 = (float)(v >> 24) / 255.0f;
            d8[sub] = unorm8(base.x * al + d0 * ia) | (unorm8(base.y * al + d1 * ia) << 8) |
                      (unorm8(base.z * al + d2 * ia) << 16) | (unorm8(al + d3 * ia) << 24);
        }
    }
#pragma unroll
    for (int sub = 0; sub < 4; sub++) px[sub * pitch] = d8[sub];
}

}  // namespace

// ------------------------------------------------------------------ launchers

const char *const kKernelNames[K_COUNT] = {
    "batch_init", "dice_count", "dice_scan", "dice_emit", "bound", "bin_count", "bin_scan",
    "bin_emit", "fill_link", "propagate_count", "propagate_scan", "propagate_emit",
    "row_paths", "sort_tiles", "fill", "tile"};

namespace {
struct Scope {
    Profiler *p;
    hipStream_t s;
    int k;
    bool on;
    Scope(Profiler *p_, hipStream_t s_, int k_) : p(p_), s(s_), k(k_), on(p_ && p_->wants(k_)) {
        if (on) p->begin(s, k);
    }
    ~Scope() {
        if (on) p->end(s, k);
    }
};
}  // namespace

size_t scan_partials_words(uint32_t max_elements) {
    return (size_t)(max_elements + kScanTile - 1) / kScanTile + 1;
}

void launch_frame_reset(FrameState *frame, hipStream_t s) {
    hipLaunchKernelGGL(k_frame_reset, dim3(1), dim3(64), 0, s, frame);
}

void launch_batch_init(const BatchArgs &a, hipStream_t s, Profiler *p) {
    uint64_t n = (uint64_t)a.fb_width_tiles * a.fb_height_tiles + 8;
    if ((uint64_t)a.column_count * 3 + 1 > n) n = (uint64_t)a.column_count * 3 + 1;
    Scope t(p, s, K_BATCH_INIT);
    hipLaunchKernelGGL(k_batch_init, dim3(grid_for(n, kBlock, a.grid_cap)), dim3(kBlock), 0, s, a);
}

void launch_dice(const BatchArgs &a, hipStream_t s, Profiler *p) {
    uint32_t g = grid_for(a.segment_count, kBlock, a.grid_cap);
    {
        Scope t(p, s, K_DICE_COUNT);
        hipLaunchKernelGGL(k_dice_count, dim3(g), dim3(kBlock), 0, s, a);
    }
    {
        Scope t(p, s, K_DICE_SCAN);
        scan_exclusive(a.seg_offsets, nullptr, a.segment_count, a.segment_count, a.scan_partials,
                       &a.counters->microlines, s,
                       ScanCount{&a.counters->microline_count, a.microline_capacity,
                                 &a.counters->overflow, 1u});
    }
    Scope t(p, s, K_DICE_EMIT);
    hipLaunchKernelGGL(k_dice_emit, dim3(g), dim3(kBlock), 0, s, a);
}

void launch_bound(const BatchArgs &a, hipStream_t s, Profiler *p) {
    if (a.tile_count == 0) return;
    Scope t(p, s, K_BOUND);
    hipLaunchKernelGGL(k_bound, dim3(grid_for(a.tile_count, kBlock * kBoundItems, a.grid_cap)),
                       dim3(kBlock),
                       0, s, a);
}

void launch_bin_count(const BatchArgs &a, hipStream_t s, Profiler *p) {
    uint32_t g = grid_for(a.microline_capacity, kBlock, a.grid_cap);
    {
        Scope t(p, s, K_BIN_COUNT);
        hipLaunchKernelGGL(k_bin_count, dim3(g), dim3(kBlock), 0, s, a);
    }
    Scope t(p, s, K_BIN_SCAN);
    scan_exclusive(a.ml_fill_offsets, &a.counters->microline_count, 0, a.microline_capacity,
                   a.scan_partials, &a.counters->fills, s,
                   ScanCount{&a.counters->fill_count, a.fill_capacity, &a.counters->overflow, 2u});
}

void launch_bin_emit(const BatchArgs &a, hipStream_t s, Profiler *p) {
    uint32_t g = grid_for(a.microline_capacity, kBlock, a.grid_cap);
    {
        Scope t(p, s, K_BIN_EMIT);
        hipLaunchKernelGGL(k_bin_emit, dim3(g), dim3(kBlock), 0, s, a);
    }
}

void launch_fill_link(const BatchArgs &a, hipStream_t s, Profiler *p) {
    if (!a.path_count) return;
    Scope t(p, s, K_FILL_LINK);
    hipLaunchKernelGGL(k_path_fills, dim3((a.path_count + kBlock) / kBlock), dim3(kBlock), 0, s,
                       a);
    hipLaunchKernelGGL(k_fill_link, dim3(a.path_count + (a.path_count + 3) / 4), dim3(kBlock), 0,
                       s, a);
}

void launch_propagate(const BatchArgs &a, hipStream_t s, Profiler *p) {
    dim3 g((a.column_count + kBlock - 1) / kBlock, a.prop_seg_rows);
    bool any = a.column_count && a.prop_seg_rows;
    if (any) {
        Scope t(p, s, K_PROPAGATE_COUNT);
        hipLaunchKernelGGL(k_propagate_count, g, dim3(kBlock), 0, s, a);
    }
    {
        Scope t(p, s, K_PROPAGATE_SCAN);
        uint32_t blocks = scan_blocks(nullptr, a.column_count, a.column_count);
        hipLaunchKernelGGL(k_scan_reduce, dim3(blocks), dim3(kBlock), 0, s, a.col_alpha_offsets,
                           nullptr, a.column_count, a.column_count, a.scan_partials);
        hipLaunchKernelGGL(k_propagate_scan_partials, dim3(1), dim3(kBlock), 0, s, a);
        hipLaunchKernelGGL(k_scan_downsweep, dim3(blocks), dim3(kBlock), 0, s,
                           a.col_alpha_offsets, nullptr, a.column_count, a.column_count,
                           a.scan_partials);
    }
    if (any) {
        Scope t(p, s, K_PROPAGATE_EMIT);
        hipLaunchKernelGGL(k_propagate_emit, g, dim3(kBlock), 0, s, a);
    }
}

void launch_row_paths(const BatchArgs &a, hipStream_t s, Profiler *p) {
    uint32_t fbn = (uint32_t)(a.fb_width_tiles * a.fb_height_tiles);
    if (fbn == 0) return;
    Scope t(p, s, K_ROW_PATHS);
    hipLaunchKernelGGL(k_row_paths, dim3((uint32_t)a.fb_height_tiles), dim3(kBlock), 0, s, a);
}

void launch_sort(const BatchArgs &a, hipStream_t s, Profiler *p) {
    uint32_t fbn = (uint32_t)(a.fb_width_tiles * a.fb_height_tiles);
    if (fbn == 0) return;
    Scope t(p, s, K_SORT_TILES);
    uint32_t bands = ((uint32_t)a.fb_width_tiles + kSortBand - 1) / kSortBand;
    // a whole number of workgroups per XCD; those past the last band return
    uint32_t n = (uint32_t)a.fb_height_tiles * bands;
    hipLaunchKernelGGL(k_sort_tiles, dim3((n + kXcds - 1) / kXcds * kXcds),
                       dim3(kSortWaves * 64), 0, s, a);
}

void launch_fill(const BatchArgs &a, hipStream_t s, Profiler *p) {
    if (a.tile_count == 0) return;
    uint32_t g = grid_for((uint64_t)a.tile_count * 64, kBlock, a.fill_grid);
    Scope t(p, s, K_FILL);
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(kBlock), 0, s, a);
}

void launch_tile(const TileArgs &a, hipStream_t s, Profiler *p) {
    uint32_t fbn = (uint32_t)(a.fb_width_tiles * a.fb_height_tiles);
    if (fbn == 0) return;
    Scope t(p, s, K_TILE);
    if (a.plain_paints)
        hipLaunchKernelGGL(k_tile<false>, dim3((fbn + 3) / 4), dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL(k_tile<true>, dim3((fbn + 3) / 4), dim3(kBlock), 0, s, a);
}

void launch_fill_zero(float *dst, size_t count, hipStream_t s) {
    if (count == 0) return;
    uint32_t g = grid_for(count, kBlock);
    hipLaunchKernelGGL(k_fill_zero, dim3(g), dim3(kBlock), 0, s, dst, count);
}

void launch_list_lengths(const uint4 *list_stats, uint32_t *lengths, uint32_t fb_tiles,
                         hipStream_t s) {
    if (fb_tiles == 0) return;
    hipLaunchKernelGGL(k_list_lengths, dim3((fb_tiles + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       s, list_stats, lengths, fb_tiles);
}

void launch_scan(uint32_t *data, uint32_t n, uint32_t *partials, hipStream_t s) {
    scan_exclusive(data, nullptr, n, n, partials, nullptr, s);
}

void launch_tile_backward(const TileGradArgs &a, hipStream_t s) {
    uint32_t fbn = (uint32_t)(a.f.fb_width_tiles * a.f.fb_height_tiles);
    if (fbn == 0) return;
    hipLaunchKernelGGL(k_tile_backward, dim3(fbn), dim3(kBlock), 0, s, a);
    if (a.slot_capacity && a.color_count)
        hipLaunchKernelGGL(k_slot_colors, dim3(grid_for(a.slot_capacity, kBlock)),
                           dim3(kBlock), 0, s, a);
}

void launch_microline_backward(const MicrolineGradArgs &a, hipStream_t s) {
    if (a.microline_count == 0 || a.b.path_count == 0) return;
    hipLaunchKernelGGL(k_microline_backward, dim3(grid_for(a.microline_count, kBlock)),
                       dim3(kBlock), 0, s, a);
}

void launch_framebuffer_to_float(const uint32_t *fb, uint32_t pitch_words, uint32_t width,
                                 uint32_t height, float *dst, hipStream_t s) {
    uint32_t n = width * height;
    if (n == 0) return;
    hipLaunchKernelGGL(k_framebuffer_to_float, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       s, fb, pitch_words, width, height, dst);
}

void launch_lut_expand(const uint8_t *rgba, float4 *lut, hipStream_t s) {
    hipLaunchKernelGGL(k_lut_expand, dim3((kLutRows * kLutStride + 255) / 256), dim3(256), 0, s,
                       rgba, lut);
}

void launch_d3d9_fill(const D3D9FillArgs &a, hipStream_t s) {
    if (!a.job_count) return;
    hipLaunchKernelGGL(k_d3d9_fill, dim3((a.job_count + 3) / 4), dim3(kBlock), 0, s, a);
}

void launch_d3d9_clip(const D3D9ClipArgs &a, hipStream_t s) {
    if (!a.clip_count) return;
    hipLaunchKernelGGL(k_d3d9_clip, dim3((a.clip_count * 64u + kBlock - 1) / kBlock),
                       dim3(kBlock), 0, s, a);
}

void launch_d3d9_tile(const D3D9TileArgs &a, hipStream_t s) {
    uint32_t fbn = (uint32_t)(a.fb_width_tiles * a.fb_height_tiles);
    if (!fbn) return;
    hipLaunchKernelGGL(k_d3d9_tile, dim3((fbn + 3) / 4), dim3(kBlock), 0, s, a);
}

void launch_metadata_expand(const uint16_t *halves, float4 *metadata, uint32_t texels,
                            hipStream_t s) {
    if (texels == 0) return;
    hipLaunchKernelGGL(k_metadata_expand, dim3((texels + 255) / 256), dim3(256), 0, s, halves,
                       metadata, texels);
}

}  // namespace pfh
