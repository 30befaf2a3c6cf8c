// pfh_renderer.cpp - the renderer seam: RendererD3D11 + the RendererCore state
// it uses, behind the C ABI of include/pfhip.h.
//
// Reference: renderer/src/gpu/d3d11/renderer.rs (RendererD3D11), the parts of
// renderer/src/gpu/renderer.rs it calls (area LUT upload :207-222, texture
// metadata :702-775, mask pages :1060-1122, begin/end scene :350-460).
//
// Differences by design:
//  - no host readbacks inside a frame: counts stay on the device and every
//    kernel reads them there; pfh_end_scene waits once, and if a microline,
//    fill or mask buffer overflowed it grows it (next power of two, as
//    d3d11/renderer.rs:227, 351 does) and renders the frame again;
//  - device memory comes from a size-class pool owned by the context, so
//    steady-state frames allocate nothing (GPUMemoryAllocator's role);
//  - batches stay resident until the next pfh_begin_scene, so their
//    buffers can be read back and the frame re-rendered without uploads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "pfh_internal.h"
#include "pfh_kernels.h"

extern "C" const uint8_t pfh_area_lut_rgba[];

namespace {

// IEEE binary16 of a float, rounded to nearest even (f16::from_f32 of the
// half crate, gpu/renderer.rs:712-763): subnormals, overflow to infinity and
// NaN kept.
uint16_t f32_to_f16(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xffu, man = x & 0x7fffffu;
    if (exp == 0xffu) return (uint16_t)(sign | 0x7c00u | (man ? 0x200u | (man >> 13) : 0u));
    int32_t e = (int32_t)exp - 127 + 15;
    if (e >= 0x1f) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        uint32_t shift = (uint32_t)(14 - e);
        uint32_t half = man >> shift, rem = man & ((1u << shift) - 1u), mid = 1u << (shift - 1);
        if (rem > mid || (rem == mid && (half & 1u))) half++;
        return (uint16_t)(sign | half);
    }
    uint32_t half = ((uint32_t)e << 10) | (man >> 13), rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (half & 1u))) half++;  // may carry into the exponent
    return (uint16_t)(sign | half);
}

// BlendMode::to_composite_ctrl (gpu/blend.rs:172-202) for PFH_BLEND_MODE_*.
int composite_ctrl(int32_t blend_mode) {
    static const int kTable[PFH_BLEND_MODE_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                                     4, 5, 1, 2, 8, 3, 6, 7, 9, 0xa, 0xb, 0xc,
                                                     0xd, 0xe, 0xf};
    return blend_mode >= 0 && blend_mode < PFH_BLEND_MODE_COUNT ? kTable[blend_mode] : 0;
}

// The ten RGBA texels of one TextureMetadataEntry (gpu/renderer.rs:702-775)
// with compute_filter_params (:967-1049), before binary16 rounding. False
// for a filter this path has no shader for.
bool encode_metadata(const pfh_texture_metadata_entry &e, float t[40]) {
    for (int k = 0; k < 40; k++) t[k] = 0.0f;
    const pfh_transform2f &m = e.color_0_transform;
    t[0] = m.m00; t[1] = m.m10; t[2] = m.m01; t[3] = m.m11;  // m11 m21 m12 m22
    t[4] = m.tx; t[5] = m.ty;                                 // m13 m23
    const float inv255 = 1.0f / 255.0f;                       // ColorU::to_f32
    t[8] = (float)e.base_color.r * inv255;
    t[9] = (float)e.base_color.g * inv255;
    t[10] = (float)e.base_color.b * inv255;
    t[11] = (float)e.base_color.a * inv255;
    int ctrl = (composite_ctrl(e.blend_mode) << 10) | ((e.color_0_combine_mode & 3) << 8);
    const float *p = e.filter.params;
    switch (e.filter.kind) {
    case PFH_FILTER_NONE:
        break;
    case PFH_FILTER_RADIAL_GRADIENT:
        // p0 = (line.from, line.vector), p1 = (radii, uv_origin)
        t[12] = p[0]; t[13] = p[1]; t[14] = p[2] - p[0]; t[15] = p[3] - p[1];
        t[16] = p[4]; t[17] = p[5]; t[18] = p[6]; t[19] = p[7];
        ctrl |= 1 << 4;
        break;
    case PFH_FILTER_BLUR: {
        const float sqrt_2_pi_inv = 0.3989422804014327f;
        float sigma_inv = 1.0f / p[1];
        float gy = expf(-0.5f * sigma_inv * sigma_inv);
        bool y_dir = p[0] != 0.0f;
        t[12] = y_dir ? 0.0f : 1.0f;
        t[13] = y_dir ? 1.0f : 0.0f;
        t[14] = ceilf(1.5f * p[1]) * 2.0f;  // support
        t[16] = sqrt_2_pi_inv * sigma_inv;
        t[17] = gy;
        t[18] = gy * gy;
        ctrl |= 3 << 4;
        break;
    }
    case PFH_FILTER_COLOR_MATRIX:
        for (int k = 0; k < 20; k++) t[12 + k] = p[k];
        ctrl |= 4 << 4;
        break;
    default:
        return false;
    }
    t[32] = (float)ctrl;
    return true;
}

uint32_t next_pow2(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p > 0x80000000ull ? 0x80000000u : (uint32_t)p;
}

// Size-class device memory pool (power-of-two classes from 256 bytes).
class Pool {
  public:
    ~Pool() { release_all(); }
    void *alloc(size_t bytes, hipError_t *err) {
        size_t cls = 256;
        while (cls < bytes) cls <<= 1;
        auto &fl = free_[cls];
        if (!fl.empty()) {
            void *p = fl.back();
            fl.pop_back();
            *err = hipSuccess;
            return p;
        }
        void *p = nullptr;
        *err = hipMalloc(&p, cls);
        if (*err != hipSuccess) {
            // return cached blocks to the device and try once more
            trim();
            *err = hipMalloc(&p, cls);
            if (*err != hipSuccess) return nullptr;
        }
        size_[p] = cls;
        bytes_ += cls;
        return p;
    }
    void free(void *p) {
        if (!p) return;
        auto it = size_.find(p);
        if (it == size_.end()) return;
        free_[it->second].push_back(p);
    }
    void trim() {
        for (auto &kv : free_) {
            for (void *p : kv.second) {
                (void)hipFree(p);
                bytes_ -= kv.first;
                size_.erase(p);
            }
            kv.second.clear();
        }
    }
    void release_all() {
        for (auto &kv : size_) (void)hipFree(kv.first);
        size_.clear();
        free_.clear();
        bytes_ = 0;
    }
    uint64_t bytes() const { return bytes_; }

  private:
    std::map<size_t, std::vector<void *>> free_;
    std::map<void *, size_t> size_;
    uint64_t bytes_ = 0;
};

// Per-kernel HIP-event timing on the stream each kernel is launched on.
class EventProfiler : public pfh::Profiler {
  public:
    int mode = PFH_PROFILE_OFF;
    int only = -1;
    uint64_t launches[pfh::K_COUNT] = {};
    double total_ms[pfh::K_COUNT] = {};

    ~EventProfiler() override {
        for (auto &p : pending_) {
            (void)hipEventDestroy(p.start);
            (void)hipEventDestroy(p.stop);
        }
        for (hipEvent_t e : free_) (void)hipEventDestroy(e);
    }
    bool wants(int k) const override {
        return mode == PFH_PROFILE_ALL || mode == PFH_PROFILE_SERIAL ||
               (mode == PFH_PROFILE_ONE && k == only);
    }
    void begin(hipStream_t s, int k) override {
        Pending p;
        p.kernel = k;
        p.start = take();
        p.stop = nullptr;
        if (p.start) (void)hipEventRecord(p.start, s);
        pending_.push_back(p);
    }
    void end(hipStream_t s, int k) override {
        if (pending_.empty() || pending_.back().kernel != k || pending_.back().stop) return;
        Pending &p = pending_.back();
        p.stop = take();
        if (p.stop) (void)hipEventRecord(p.stop, s);
    }
    // Folds finished measurements into the totals (the stream must be idle).
    void collect() {
        for (auto &p : pending_) {
            float ms = 0;
            if (p.start && p.stop && hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
                launches[p.kernel]++;
                total_ms[p.kernel] += ms;
            }
            if (p.start) free_.push_back(p.start);
            if (p.stop) free_.push_back(p.stop);
        }
        pending_.clear();
    }
    void reset() {
        collect();
        for (int i = 0; i < pfh::K_COUNT; i++) {
            launches[i] = 0;
            total_ms[i] = 0;
        }
    }

  private:
    struct Pending {
        int kernel;
        hipEvent_t start, stop;
    };
    hipEvent_t take() {
        if (!free_.empty()) {
            hipEvent_t e = free_.back();
            free_.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> free_;
};

struct Batch {
    pfh::BatchData host;
    bool draw = false;
    hipEvent_t front_done = nullptr;  // the batch's front stages finished
    hipEvent_t prop_done = nullptr;   // its propagate stage finished
    hipEvent_t sort_done = nullptr;   // its sort stage finished
    hipEvent_t bound_done = nullptr;  // its tiles were initialised (k_bound on the context stream)
    hipEvent_t link_done = nullptr;   // its per-tile fill lists are built
    hipEvent_t fill_done = nullptr;   // its masks are written (fill on its side stream)
    int32_t load = 1;
    float clear[4] = {0, 0, 0, 0};
    int32_t color_page = -1;       // colour texture 0 of a draw batch
    uint32_t color_flags = 0;
    // device inputs
    pfh_dice_metadata *d_dice = nullptr;
    pfh_tile_path_info *d_tpi = nullptr;
    pfh_propagate_metadata *d_prop = nullptr;
    pfh_backdrop_info *d_backdrops_init = nullptr;
    // device working buffers
    pfh_backdrop_info *d_backdrops = nullptr;
    uint32_t prop_seg_rows = 0;        // propagate segments of the tallest column
    int32_t *d_seg_delta = nullptr;
    uint32_t *d_seg_alpha = nullptr;
    uint32_t *d_seg_offsets = nullptr;
    uint4 *d_microlines = nullptr;
    uint32_t *d_ml_fill_offsets = nullptr;
    uint2 *d_fill_words = nullptr;     // Fill words 0-1
    uint32_t *d_fill_links = nullptr;  // Fill word 2
    uint32_t *d_fill_tiles = nullptr;
    uint32_t *d_path_fills = nullptr;
    uint2 *d_sorted_fills = nullptr;
    uint2 *d_run_prev = nullptr;
    uint2 *d_tile_runs = nullptr;
    uint2 *d_tiles = nullptr;  // tile words 2-3
    uint32_t *d_tile_next = nullptr;   // tile word 0 (linked tiles)
    uint32_t *d_tile_first = nullptr;  // tile word 1 (tiles with fills)
    uint32_t *d_tile_paths = nullptr;  // with gradients: the batch path of every tile
    // with gradients, draw batches: the composite's record of every list slot's
    // tile, at the offsets of an exclusive scan of the lists' lengths
    uint32_t *d_slot_off = nullptr, *d_slot_scan = nullptr, *d_slot_tiles = nullptr;
    uint8_t *d_tile_flags = nullptr;   // 1 per tile bin gave a fill
    uint8_t *d_tile_status = nullptr;  // per tile: linked / solid / alpha bits
    uint32_t *d_tile_delta = nullptr;  // bin's backdrop deltas per tile
    uint32_t *d_col_alpha = nullptr;
    uint2 *d_alpha = nullptr;
    uint2 *d_alpha_words = nullptr;  // per alpha tile: its tile words 2-3
    int32_t *d_zbuffer = nullptr;
    int32_t *d_ftm = nullptr;
    uint32_t *d_row_paths = nullptr;
    uint32_t *d_row_counts = nullptr;
    uint4 *d_list_stats = nullptr;
    uint32_t *d_scan = nullptr;
    pfh::BatchCounters *d_counters = nullptr;
    uint32_t cap_m = 0, cap_f = 0;
    pfh::BatchCounters counters = {};
};

}  // namespace

struct pfh_context {
    int device = 0;
    // Back stages (propagate onwards) run in batch order on `stream`; each
    // batch's front stages (init, dice, bound, bin, fill links) run on one of
    // the two side streams, so they overlap the previous batch's back stages.
    hipStream_t stream = nullptr;
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t frame_start = nullptr;
    std::string error;
    Pool pool;
    EventProfiler prof;
    // tables
    float4 *d_lut = nullptr;
    uint32_t *d_lut_sat = nullptr;  // saturated LUT cells per row (lut_saturation)
    float4 *d_metadata = nullptr;  // 10 float4 per paint (TextureMetadataEntry texels)
    uint32_t paint_count = 0;
    bool paints_plain = true;      // every paint's composite control word is zero
    // colour texture pages (AllocateTexturePage): RGBA8 words, row by row
    struct Page {
        int32_t width = 0, height = 0;
        uint32_t *texels = nullptr;
    };
    std::map<uint32_t, Page> pages;
    // scene sources
    std::vector<pfh_vector2f> draw_points_host, clip_points_host;
    std::vector<pfh_segment_indices> draw_idx_host, clip_idx_host;
    float2 *d_draw_points = nullptr, *d_clip_points = nullptr;
    uint2 *d_draw_idx = nullptr, *d_clip_idx = nullptr;
    // frame
    bool in_scene = false;
    int32_t width = 0, height = 0, fbw = 0, fbh = 0;
    bool has_background = false;
    float background[4] = {0, 0, 0, 0};
    bool dest_dirty = false;
    // the D3D9 level: queued fills, the frame's alpha tiles and the binary16
    // mask (256 halves per alpha tile; the leading d9_mask_zeroed tiles are
    // cleared for this frame)
    std::vector<pfh_fill_d3d9> d9_pending;
    uint32_t d9_alpha_count = 0;
    uint32_t d9_mask_zeroed = 0;
    uint16_t *d_d9_mask = nullptr;
    uint32_t d9_mask_cap = 0;
    uint32_t *d_fb = nullptr, *d_fb_saved = nullptr;
    size_t fb_words = 0;
    uint32_t *d_mask = nullptr;
    uint32_t cap_a = 0;
    // adjoint: the fill's coverage checkpoint (256 floats per alpha tile,
    // sized with the mask) and the backward pass's scratch
    bool gradients = false;
    int32_t checkpoint = PFH_CHECKPOINT_STATES;    // what the forward keeps for the adjoint
    uint8_t *d_cov_state = nullptr;                // 64 B per alpha tile (states, recompute)
    // launch shape of the grid-stride kernels and of k_fill (pfh_autotune)
    pfh_launch_shape shape = {16384, 8192};
    // pfh_rerender_scene through a captured graph of the frame
    bool use_graphs = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    float *d_coverage = nullptr;
    float *d_grad_cov = nullptr;
    size_t grad_cov_cap = 0;
    float *d_trans = nullptr;
    size_t trans_cap = 0;
    uint32_t *d_list_tiles = nullptr;
    uint32_t *d_grad_stamp = nullptr;  // epoch of the alpha tiles grad_cov holds
    uint32_t grad_epoch = 0;
    uint32_t *d_list_off = nullptr, *d_list_scan = nullptr;
    size_t list_off_cap = 0;
    pfh::FrameState *d_frame = nullptr;
    uint32_t cap_m = 16384, cap_f = 16384;
    std::vector<std::unique_ptr<Batch>> batches;
    pfh_render_stats stats = {};
    uint32_t replays = 0;
};

namespace {

pfh_status fail(pfh_context *ctx, pfh_status s, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->error = buf;
    return s;
}

#define HIP_TRY(ctx, expr)                                                                    \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail((ctx), PFH_ERROR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

template <class T>
pfh_status dalloc(pfh_context *ctx, T **out, size_t count) {
    hipError_t e;
    void *p = ctx->pool.alloc(std::max<size_t>(count, 1) * sizeof(T), &e);
    if (!p) return fail(ctx, PFH_ERROR_OUT_OF_MEMORY, "device allocation of %zu bytes failed: %s",
                        count * sizeof(T), hipGetErrorString(e));
    *out = static_cast<T *>(p);
    return PFH_OK;
}

// The saturated cells of the padded area LUT, for k_fill. Padded texel (j, i)
// is byte texel (clamp(j - 1), clamp(i - 1)) of the 256 x 256 RGBA table
// (clamp to [0, 255], row j); the bilinear sample at padded cell (jp, ip),
// 0 <= jp, ip <= 257, blends texels (jp .. jp + 1, ip .. ip + 1). Where all
// four hold 255 in every channel the blend is exactly 1 in every channel,
// where all four hold 0 exactly 0. Word jp: low 16 bits the largest ip such
// that every cell (jp, 0 .. ip) is all 255 (as int16; -0x7fff when cell
// (jp, 0) is not), high 16 bits the smallest ip such that every cell
// (jp, ip .. 257) is all 0 (0x7fff when cell (jp, 257) is not). The sentinels
// lie far enough outside [0, 257] that no index compares past them.
constexpr uint32_t kLutSatRows = 258;
void lut_saturation(const uint8_t *rgba, uint32_t *out) {
    auto texel = [&](int j, int i) {
        j = std::min(std::max(j - 1, 0), 255);
        i = std::min(std::max(i - 1, 0), 255);
        const uint8_t *t = rgba + ((size_t)j * 256 + (size_t)i) * 4;
        return (uint32_t)t[0] | (uint32_t)t[1] << 8 | (uint32_t)t[2] << 16 | (uint32_t)t[3] << 24;
    };
    auto cell = [&](int jp, int ip, uint32_t v) {
        return texel(jp, ip) == v && texel(jp + 1, ip) == v && texel(jp, ip + 1) == v &&
               texel(jp + 1, ip + 1) == v;
    };
    for (int jp = 0; jp < (int)kLutSatRows; jp++) {
        int full = -1;
        while (full + 1 <= 257 && cell(jp, full + 1, 0xffffffffu)) full++;
        int zero = 258;
        while (zero - 1 >= 0 && cell(jp, zero - 1, 0u)) zero--;
        if (full < 0) full = -0x7fff;
        if (zero == 258) zero = 0x7fff;
        out[jp] = ((uint32_t)full & 0xffffu) | ((uint32_t)zero << 16);
    }
}

template <class T>
void dfree(pfh_context *ctx, T *&p) {
    ctx->pool.free(p);
    p = nullptr;
}

void free_batch(pfh_context *ctx, Batch &b) {
    for (hipEvent_t *e :
         {&b.front_done, &b.prop_done, &b.sort_done, &b.bound_done, &b.link_done, &b.fill_done}) {
        if (*e) (void)hipEventDestroy(*e);
        *e = nullptr;
    }
    dfree(ctx, b.d_dice); dfree(ctx, b.d_tpi); dfree(ctx, b.d_prop);
    dfree(ctx, b.d_backdrops_init); dfree(ctx, b.d_backdrops); dfree(ctx, b.d_seg_offsets);
    dfree(ctx, b.d_seg_delta); dfree(ctx, b.d_seg_alpha);
    dfree(ctx, b.d_microlines); dfree(ctx, b.d_ml_fill_offsets); dfree(ctx, b.d_fill_words);
    dfree(ctx, b.d_fill_links); dfree(ctx, b.d_fill_tiles); dfree(ctx, b.d_sorted_fills); dfree(ctx, b.d_run_prev);
    dfree(ctx, b.d_path_fills);
    dfree(ctx, b.d_tile_runs);
    dfree(ctx, b.d_tiles); dfree(ctx, b.d_tile_next); dfree(ctx, b.d_tile_first);
    dfree(ctx, b.d_tile_paths); dfree(ctx, b.d_col_alpha);
    dfree(ctx, b.d_slot_off); dfree(ctx, b.d_slot_scan); dfree(ctx, b.d_slot_tiles);
    dfree(ctx, b.d_tile_flags); dfree(ctx, b.d_tile_delta);
    dfree(ctx, b.d_tile_status);
    dfree(ctx, b.d_alpha); dfree(ctx, b.d_alpha_words); dfree(ctx, b.d_zbuffer); dfree(ctx, b.d_ftm);
    dfree(ctx, b.d_row_paths); dfree(ctx, b.d_row_counts); dfree(ctx, b.d_list_stats);
    dfree(ctx, b.d_scan); dfree(ctx, b.d_counters);
}

void free_batches(pfh_context *ctx) {
    for (auto &b : ctx->batches) free_batch(ctx, *b);
    ctx->batches.clear();
}

template <class T>
pfh_status upload(pfh_context *ctx, T **dst, const std::vector<T> &src,
                  hipStream_t stream = nullptr) {
    pfh_status s = dalloc(ctx, dst, src.size());
    if (s != PFH_OK) return s;
    if (!src.empty())
        HIP_TRY(ctx, hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T),
                                    hipMemcpyHostToDevice, stream ? stream : ctx->stream));
    return PFH_OK;
}

// (Re)allocates the capacity-dependent buffers of a batch.
pfh_status size_capacity_buffers(pfh_context *ctx, Batch &b) {
    if (b.cap_m == ctx->cap_m && b.cap_f == ctx->cap_f && b.d_microlines) return PFH_OK;
    dfree(ctx, b.d_microlines); dfree(ctx, b.d_ml_fill_offsets);
    dfree(ctx, b.d_fill_words); dfree(ctx, b.d_fill_links); dfree(ctx, b.d_fill_tiles);
    dfree(ctx, b.d_sorted_fills);
    dfree(ctx, b.d_run_prev); dfree(ctx, b.d_scan);
    b.cap_m = ctx->cap_m;
    b.cap_f = ctx->cap_f;
    pfh_status s;
    if ((s = dalloc(ctx, &b.d_microlines, b.cap_m)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_ml_fill_offsets, (size_t)b.cap_m + 1)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_fill_words, b.cap_f)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_fill_links, b.cap_f)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_fill_tiles, b.cap_f)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_sorted_fills, b.cap_f)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_run_prev, b.cap_f)) != PFH_OK) return s;
    uint32_t scan_n = std::max({b.host.segment_count, b.cap_m, (uint32_t)b.host.backdrops.size()});
    if ((s = dalloc(ctx, &b.d_scan, pfh::scan_partials_words(scan_n))) != PFH_OK) return s;
    return PFH_OK;
}

pfh_status ensure_coverage(pfh_context *ctx) {
    if (!ctx->gradients) return PFH_OK;
    if (ctx->checkpoint == PFH_CHECKPOINT_COVERAGE)
        return ctx->d_coverage ? PFH_OK : dalloc(ctx, &ctx->d_coverage, (size_t)ctx->cap_a * 256);
    return ctx->d_cov_state ? PFH_OK : dalloc(ctx, &ctx->d_cov_state, (size_t)ctx->cap_a * 64);
}

pfh_status ensure_mask(pfh_context *ctx, uint32_t tiles) {
    if (ctx->d_mask && tiles <= ctx->cap_a) return ensure_coverage(ctx);
    uint32_t cap = std::max<uint32_t>(ctx->cap_a ? ctx->cap_a : 65536, next_pow2(tiles));
    dfree(ctx, ctx->d_mask);
    dfree(ctx, ctx->d_coverage);
    ctx->d_coverage = nullptr;
    dfree(ctx, ctx->d_cov_state);
    ctx->d_cov_state = nullptr;
    ctx->cap_a = cap;
    pfh_status s = dalloc(ctx, &ctx->d_mask, (size_t)cap * 64);
    return s != PFH_OK ? s : ensure_coverage(ctx);
}

Batch *find_batch(pfh_context *ctx, uint32_t id) {
    for (auto &b : ctx->batches)
        if (b->host.batch_id == id) return b.get();
    return nullptr;
}

// Propagate works on column segments of kPropSegmentRows tile rows; the
// tallest path's height gives the segment rows of its grid.
uint32_t prop_segment_rows(const pfh::BatchData &h) {
    uint32_t rows = 0;
    for (const pfh_propagate_metadata &m : h.propagate_metadata) {
        int64_t height = (int64_t)m.tile_rect[3] - m.tile_rect[1];
        if (height > 0)
            rows = std::max(rows, (uint32_t)((height + pfh::kPropSegmentRows - 1) /
                                             pfh::kPropSegmentRows));
    }
    // a column with an out-of-range path index takes one segment (reported)
    return std::max(rows, 1u);
}

pfh::BatchArgs batch_args(pfh_context *ctx, Batch &b) {
    pfh::BatchArgs a;
    std::memset(&a, 0, sizeof(a));
    bool clip_src = b.host.path_source == PFH_PATH_SOURCE_CLIP;
    a.scene.points = clip_src ? ctx->d_clip_points : ctx->d_draw_points;
    a.scene.indices = clip_src ? ctx->d_clip_idx : ctx->d_draw_idx;
    a.scene.point_count = (uint32_t)(clip_src ? ctx->clip_points_host.size()
                                              : ctx->draw_points_host.size());
    a.scene.index_count = (uint32_t)(clip_src ? ctx->clip_idx_host.size()
                                              : ctx->draw_idx_host.size());
    a.dice_md = reinterpret_cast<const uint4 *>(b.d_dice);
    a.tile_path_info = reinterpret_cast<const uint4 *>(b.d_tpi);
    a.prop_md = reinterpret_cast<const int32_t *>(b.d_prop);
    a.backdrops_init = reinterpret_cast<const int32_t *>(b.d_backdrops_init);
    a.path_count = b.host.path_count;
    a.tile_count = b.host.tile_count;
    a.segment_count = b.host.segment_count;
    a.column_count = (uint32_t)b.host.backdrops.size();
    const pfh_transform2f &t = b.host.transform;
    a.xform[0] = t.m00; a.xform[1] = t.m01; a.xform[2] = t.m10; a.xform[3] = t.m11;
    a.xform[4] = t.tx; a.xform[5] = t.ty;
    if (b.host.clip_batch_id >= 0) {
        Batch *c = find_batch(ctx, (uint32_t)b.host.clip_batch_id);
        if (c) {
            a.clip_prop_md = reinterpret_cast<const int32_t *>(c->d_prop);
            a.clip_tiles = c->d_tiles;
            a.clip_path_count = c->host.path_count;
            a.clip_tile_count = c->host.tile_count;
        }
    }
    a.backdrops = reinterpret_cast<int32_t *>(b.d_backdrops);
    a.seg_offsets = b.d_seg_offsets;
    a.microlines = b.d_microlines;
    a.microline_capacity = b.cap_m;
    a.ml_fill_offsets = b.d_ml_fill_offsets;
    a.fill_words = b.d_fill_words;
    a.fill_links = b.d_fill_links;
    a.fill_tiles = b.d_fill_tiles;
    a.fill_capacity = b.cap_f;
    a.sorted_fills = b.d_sorted_fills;
    a.tile_runs = b.d_tile_runs;
    a.run_prev = b.d_run_prev;
    a.path_fills = b.d_path_fills;
    a.tiles = b.d_tiles;
    a.tile_next = b.d_tile_next;
    a.tile_first = b.d_tile_first;
    a.tile_flags = b.d_tile_flags;
    a.tile_status = b.d_tile_status;
    a.tile_delta = b.d_tile_delta;
    a.col_alpha_offsets = b.d_col_alpha;
    a.prop_seg_rows = b.prop_seg_rows;
    a.seg_delta = b.d_seg_delta;
    a.seg_alpha = b.d_seg_alpha;
    a.alpha_tiles = b.d_alpha;
    a.alpha_words = b.d_alpha_words;
    a.zbuffer = b.d_zbuffer;
    a.first_tile_map = b.d_ftm;
    a.row_paths = b.d_row_paths;
    a.row_counts = b.d_row_counts;
    a.list_stats = b.d_list_stats;
    a.scan_partials = b.d_scan;
    a.counters = b.d_counters;
    a.frame = ctx->d_frame;
    a.mask = ctx->d_mask;
    a.mask_capacity = ctx->cap_a;
    a.lut = ctx->d_lut;
    a.lut_sat = ctx->d_lut_sat;
    a.coverage = ctx->gradients && ctx->checkpoint == PFH_CHECKPOINT_COVERAGE ? ctx->d_coverage
                                                                             : nullptr;
    a.cov_state = ctx->gradients && ctx->checkpoint == PFH_CHECKPOINT_STATES ? ctx->d_cov_state
                                                                            : nullptr;
    a.states_only = 0;
    a.tile_paths = ctx->gradients ? b.d_tile_paths : nullptr;
    a.fb_width_tiles = ctx->fbw;
    a.fb_height_tiles = ctx->fbh;
    a.grid_cap = ctx->shape.grid_cap;
    a.fill_grid = ctx->shape.fill_grid;
    return a;
}

// prepare_tiles (d3d11/renderer.rs:427-542) and, for draw batches, draw_tiles
// (:711-782). The stages that only touch the batch's own buffers (init, dice,
// bound, the per-row path lists, bin) run on the side stream `front`;
// propagate, fill and tile (the clip batch's tiles and masks, the frame's
// alpha tile count, the framebuffer) follow in batch order on the context
// stream. Sort reads what propagate writes but not the masks, so it runs on
// `front` beside fill; tile and the next batch wait for it.
// The front stages of a batch on its side stream. k_bound (tile
// initialisation) runs here too unless the caller has put it on the context
// stream (bound_done), where it overlaps dice; only bin's emit needs it.
pfh::TileArgs tile_args(pfh_context *ctx, const Batch &b) {
    pfh::TileArgs t;
    t.tiles = b.d_tiles;
    t.tile_next = b.d_tile_next;
    t.tile_count = b.host.tile_count;
    t.mask_capacity = ctx->cap_a;
    t.first_tile_map = b.d_ftm;
    t.mask = ctx->d_mask;
    t.metadata = ctx->d_metadata;
    t.paint_count = ctx->paint_count;
    t.plain_paints = ctx->paints_plain ? 1 : 0;
    bool slots = ctx->gradients && b.d_slot_tiles;
    t.list_tiles = slots ? b.d_slot_tiles : nullptr;
    t.list_offsets = slots ? b.d_slot_off : nullptr;
    t.list_capacity = slots ? b.host.tile_count + 1 : 0;
    t.color_texture = nullptr;
    t.color_width = t.color_height = 0;
    t.color_flags = b.color_flags;
    auto page = b.color_page >= 0 ? ctx->pages.find((uint32_t)b.color_page) : ctx->pages.end();
    if (page != ctx->pages.end()) {
        t.color_texture = page->second.texels;
        t.color_width = page->second.width;
        t.color_height = page->second.height;
    }
    t.fb_width = (float)ctx->width;
    t.fb_height = (float)ctx->height;
    t.framebuffer = ctx->d_fb;
    t.fb_width_tiles = ctx->fbw;
    t.fb_height_tiles = ctx->fbh;
    t.load = b.load;
    for (int i = 0; i < 4; i++) t.clear[i] = b.clear[i];
    return t;
}

// With gradients, after a draw batch's sort: the slot offsets of its lists
// (exclusive scan of their lengths), where the composite records each
// slot's tile for the adjoint.
void list_slots(pfh_context *ctx, Batch &b, hipStream_t s) {
    if (!ctx->gradients || !b.d_slot_off) return;
    uint32_t fbn = (uint32_t)(ctx->fbw * ctx->fbh);
    pfh::launch_list_lengths(b.d_list_stats, b.d_slot_off, fbn, s);
    pfh::launch_scan(b.d_slot_off, fbn, b.d_slot_scan, s);
}

pfh_status run_front(pfh_context *ctx, Batch &b, hipStream_t front, bool bound_launched) {
    pfh::BatchArgs a = batch_args(ctx, b);
    pfh::Profiler *p = ctx->prof.mode == PFH_PROFILE_OFF ? nullptr : &ctx->prof;
    pfh::launch_batch_init(a, front, p);
    pfh::launch_dice(a, front, p);
    if (!bound_launched) pfh::launch_bound(a, front, p);
    pfh::launch_bin_count(a, front, p);
    if (bound_launched) HIP_TRY(ctx, hipStreamWaitEvent(front, b.bound_done, 0));
    pfh::launch_bin_emit(a, front, p);
    HIP_TRY(ctx, hipEventRecord(b.front_done, front));
    // the per-tile fill lists build beside propagate (which needs only bin's
    // tile flags); fill waits for them
    pfh::launch_fill_link(a, front, p);
    HIP_TRY(ctx, hipEventRecord(b.link_done, front));
    return PFH_OK;
}

pfh_status run_back(pfh_context *ctx, Batch &b, hipStream_t front) {
    pfh::BatchArgs a = batch_args(ctx, b);
    pfh::Profiler *p = ctx->prof.mode == PFH_PROFILE_OFF ? nullptr : &ctx->prof;
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, hipStreamWaitEvent(s, b.front_done, 0));
    pfh::launch_propagate(a, s, p);
    HIP_TRY(ctx, hipEventRecord(b.prop_done, s));
    // the per-row path lists are only read by sort: they build while propagate runs
    pfh::launch_row_paths(a, front, p);
    HIP_TRY(ctx, hipStreamWaitEvent(front, b.prop_done, 0));
    pfh::launch_sort(a, front, p);
    list_slots(ctx, b, front);
    HIP_TRY(ctx, hipEventRecord(b.sort_done, front));
    HIP_TRY(ctx, hipStreamWaitEvent(s, b.link_done, 0));
    pfh::launch_fill(a, s, p);
    // only the composite reads the sorted lists; a clip batch's sort joins
    // the context stream at the end of the frame (join_side_streams)
    if (b.draw) {
        HIP_TRY(ctx, hipStreamWaitEvent(s, b.sort_done, 0));
        pfh::launch_tile(tile_args(ctx, b), s, p);
    }
    HIP_TRY(ctx, hipGetLastError());
    return PFH_OK;
}

// Makes the context stream wait for every batch's sort (the last work of
// each side stream in a frame), so waiting for it waits for the frame.
pfh_status join_side_streams(pfh_context *ctx) {
    for (auto &b : ctx->batches) HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, b->sort_done, 0));
    return PFH_OK;
}

pfh_status run_batch(pfh_context *ctx, Batch &b, hipStream_t front) {
    pfh_status st = run_front(ctx, b, front, false);
    return st != PFH_OK ? st : run_back(ctx, b, front);
}

// Orders the side streams after everything enqueued on the context stream so
// far (the previous frame reads the batches' buffers until it ends).
pfh_status start_frame_on_side_streams(pfh_context *ctx) {
    HIP_TRY(ctx, hipEventRecord(ctx->frame_start, ctx->stream));
    for (hipStream_t side : ctx->side) HIP_TRY(ctx, hipStreamWaitEvent(side, ctx->frame_start, 0));
    return PFH_OK;
}

void drop_graph(pfh_context *ctx) {
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graph_exec = nullptr;
    ctx->graph = nullptr;
}

// Sizes every batch's capacity buffers (allocation happens here, never while
// a frame is being captured).
pfh_status size_all(pfh_context *ctx) {
    for (auto &b : ctx->batches) {
        pfh_status s = size_capacity_buffers(ctx, *b);
        if (s != PFH_OK) return s;
    }
    return PFH_OK;
}

// The frame's launches for batches whose buffers are sized (capturable).
pfh_status enqueue_frame(pfh_context *ctx) {
    // the side streams start as soon as the previous frame is done, beside
    // the frame state reset (only propagate, on the context stream, reads it)
    pfh_status s = start_frame_on_side_streams(ctx);
    if (s != PFH_OK) return s;
    pfh::launch_frame_reset(ctx->d_frame, ctx->stream);
    if (ctx->d_fb_saved)
        HIP_TRY(ctx, hipMemcpyAsync(ctx->d_fb, ctx->d_fb_saved, ctx->fb_words * 4,
                                    hipMemcpyDeviceToDevice, ctx->stream));
    if (ctx->prof.mode == PFH_PROFILE_SERIAL) {
        // every stage of every batch in order on the context stream, so that
        // each kernel's events time it alone
        for (auto &b : ctx->batches)
            if ((s = run_batch(ctx, *b, ctx->stream)) != PFH_OK) return s;
        return join_side_streams(ctx);
    }
    // Every batch's tiles are initialised on the context stream first (it is
    // idle until the first front ends), then every batch's front stages, the
    // one with the most segments first (its chain is the longest), then the
    // back stages in batch order.
    size_t n = ctx->batches.size();
    for (size_t i = 0; i < n; i++) {
        Batch &b = *ctx->batches[i];
        pfh::BatchArgs a = batch_args(ctx, b);
        pfh::launch_bound(a, ctx->stream, ctx->prof.mode == PFH_PROFILE_OFF ? nullptr : &ctx->prof);
        HIP_TRY(ctx, hipEventRecord(b.bound_done, ctx->stream));
    }
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        return ctx->batches[x]->host.segment_count > ctx->batches[y]->host.segment_count;
    });
    for (size_t i : order)
        if ((s = run_front(ctx, *ctx->batches[i], ctx->side[i & 1], true)) != PFH_OK) return s;
    // Back stages. Propagate runs in batch order on the context stream (it
    // allocates each batch's frame-global alpha tile range, clip batches
    // first). Every batch but the last fills its masks on its side stream
    // once its propagate is done, so the next batch's propagate does not wait
    // for it; the last batch fills on the context stream after them (its
    // clip masks), beside its sort. Composites follow in batch order.
    pfh::Profiler *p = ctx->prof.mode == PFH_PROFILE_OFF ? nullptr : &ctx->prof;
    for (size_t i = 0; i < n; i++) {
        Batch &b = *ctx->batches[i];
        pfh::BatchArgs a = batch_args(ctx, b);
        HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, b.front_done, 0));
        pfh::launch_propagate(a, ctx->stream, p);
        HIP_TRY(ctx, hipEventRecord(b.prop_done, ctx->stream));
    }
    // Fills first (the last batch's fill reads their masks: they are on the
    // critical path), then the per-row path lists and sorts, which only the
    // composites read.
    for (size_t i = 0; i + 1 < n; i++) {
        Batch &b = *ctx->batches[i];
        hipStream_t side = ctx->side[i & 1];
        HIP_TRY(ctx, hipStreamWaitEvent(side, b.prop_done, 0));
        for (size_t j = 0; j < i; j++)  // earlier batches' masks (its clip masks)
            HIP_TRY(ctx, hipStreamWaitEvent(side, ctx->batches[j]->fill_done, 0));
        pfh::launch_fill(batch_args(ctx, b), side, p);
        HIP_TRY(ctx, hipEventRecord(b.fill_done, side));
    }
    for (size_t i = 0; i < n; i++) {
        Batch &b = *ctx->batches[i];
        pfh::BatchArgs a = batch_args(ctx, b);
        hipStream_t side = ctx->side[i & 1];
        // the per-row path lists and the sort run beside the last batch's
        // fill, not beside the propagate and fill lists it waits for
        HIP_TRY(ctx, hipStreamWaitEvent(side, ctx->batches[n - 1]->prop_done, 0));
        pfh::launch_row_paths(a, side, p);
        pfh::launch_sort(a, side, p);
        list_slots(ctx, b, side);
        HIP_TRY(ctx, hipEventRecord(b.sort_done, side));
    }
    if (n) {
        Batch &last = *ctx->batches[n - 1];
        for (size_t j = 0; j + 1 < n; j++)
            HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, ctx->batches[j]->fill_done, 0));
        HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, last.link_done, 0));
        pfh::launch_fill(batch_args(ctx, last), ctx->stream, p);
    }
    for (size_t i = 0; i < n; i++) {
        Batch &b = *ctx->batches[i];
        if (!b.draw) continue;
        HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, b.sort_done, 0));
        pfh::launch_tile(tile_args(ctx, b), ctx->stream, p);
    }
    HIP_TRY(ctx, hipGetLastError());
    return join_side_streams(ctx);
}

pfh_status replay_frame(pfh_context *ctx) {
    pfh_status s = size_all(ctx);
    return s != PFH_OK ? s : enqueue_frame(ctx);
}

// Captures the frame's launches on the three streams into a graph (the side
// streams join through the frame-start event and leave through the events
// the context stream waits on) and instantiates it.
pfh_status capture_frame(pfh_context *ctx) {
    pfh_status s = size_all(ctx);
    if (s != PFH_OK) return s;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (hipStream_t side : ctx->side) HIP_TRY(ctx, hipStreamSynchronize(side));
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    s = enqueue_frame(ctx);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (s != PFH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return s;
    }
    if (e != hipSuccess) return fail(ctx, PFH_ERROR_HIP, "frame capture: %s", hipGetErrorString(e));
    ctx->graph = g;
    HIP_TRY(ctx, hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
    return PFH_OK;
}

pfh_status add_batch(pfh_context *ctx, const pfh_tile_batch *in, bool draw) {
    if (!ctx || !in) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    if (!ctx->in_scene) return fail(ctx, PFH_ERROR_STATE, "batch outside begin/end scene");
    if (find_batch(ctx, in->batch_id))
        return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "batch %u already prepared", in->batch_id);
    if (in->path_count && (!in->dice_metadata || !in->tile_path_info || !in->propagate_metadata))
        return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "batch %u: missing metadata", in->batch_id);
    if (in->backdrop_count && !in->backdrops)
        return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "batch %u: missing backdrops", in->batch_id);
    if (in->clip_batch_id >= 0 && !find_batch(ctx, (uint32_t)in->clip_batch_id))
        return fail(ctx, PFH_ERROR_STATE, "batch %u: clip batch %d not prepared", in->batch_id,
                    in->clip_batch_id);
    std::unique_ptr<Batch> bp(new (std::nothrow) Batch());
    if (!bp) return fail(ctx, PFH_ERROR_OUT_OF_MEMORY, "host allocation failed");
    Batch &b = *bp;
    pfh::BatchData &h = b.host;
    h.batch_id = in->batch_id;
    h.path_source = in->path_source;
    h.path_count = in->path_count;
    h.tile_count = in->tile_count;
    h.segment_count = in->segment_count;
    h.transform = in->transform;
    h.clip_batch_id = in->clip_batch_id;
    h.dice_metadata.assign(in->dice_metadata, in->dice_metadata + in->path_count);
    h.tile_path_info.assign(in->tile_path_info, in->tile_path_info + in->path_count);
    h.propagate_metadata.assign(in->propagate_metadata, in->propagate_metadata + in->path_count);
    h.backdrops.assign(in->backdrops, in->backdrops + in->backdrop_count);
    h.color_texture_page = in->color_texture_page;
    h.color_texture_sampling_flags = in->color_texture_sampling_flags;
    if (draw && in->color_texture_page >= 0 &&
        !ctx->pages.count((uint32_t)in->color_texture_page))
        return fail(ctx, PFH_ERROR_STATE, "batch %u: colour texture page %d is not allocated",
                    in->batch_id, in->color_texture_page);
    b.color_page = draw ? in->color_texture_page : -1;
    b.color_flags = in->color_texture_sampling_flags;
    b.draw = draw;
    if (draw) {
        // clear_color_for_draw_operation (gpu/renderer.rs:1190-1211)
        b.load = (ctx->dest_dirty || !ctx->has_background) ? 1 : 0;
        for (int i = 0; i < 4; i++) b.clear[i] = b.load ? 0.0f : ctx->background[i];
        ctx->dest_dirty = true;
    }
    // The microline count is at least the segment count; start there.
    if (h.segment_count > ctx->cap_m) ctx->cap_m = next_pow2(h.segment_count);
    if (ctx->cap_f < ctx->cap_m) ctx->cap_f = ctx->cap_m;

    size_t fbn = (size_t)ctx->fbw * ctx->fbh;
    // the batch's inputs go up on the side stream its front stages run on
    hipStream_t front = ctx->side[ctx->batches.size() & 1];
    for (hipEvent_t *e :
         {&b.front_done, &b.prop_done, &b.sort_done, &b.bound_done, &b.link_done, &b.fill_done}) {
        if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) {
            *e = nullptr;
            return fail(ctx, PFH_ERROR_HIP, "hipEventCreateWithFlags failed");
        }
    }
    pfh_status s;
    if ((s = upload(ctx, &b.d_dice, h.dice_metadata, front)) != PFH_OK) return s;
    if ((s = upload(ctx, &b.d_tpi, h.tile_path_info, front)) != PFH_OK) return s;
    if ((s = upload(ctx, &b.d_prop, h.propagate_metadata, front)) != PFH_OK) return s;
    if ((s = upload(ctx, &b.d_backdrops_init, h.backdrops, front)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_backdrops, h.backdrops.size())) != PFH_OK) return s;
    b.prop_seg_rows = prop_segment_rows(h);
    if ((s = dalloc(ctx, &b.d_seg_delta, (size_t)b.prop_seg_rows * h.backdrops.size())) != PFH_OK)
        return s;
    if ((s = dalloc(ctx, &b.d_seg_alpha, (size_t)b.prop_seg_rows * h.backdrops.size())) != PFH_OK)
        return s;
    if ((s = dalloc(ctx, &b.d_seg_offsets, (size_t)h.segment_count + 1)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_path_fills, (size_t)h.path_count + 1)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_tiles, h.tile_count)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_tile_next, h.tile_count)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_tile_first, h.tile_count)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_tile_flags, h.tile_count)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_tile_status, h.tile_count)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_tile_delta, h.tile_count)) != PFH_OK) return s;
    if (ctx->gradients && (s = dalloc(ctx, &b.d_tile_paths, h.tile_count)) != PFH_OK) return s;
    if (ctx->gradients && draw) {
        uint32_t fbn_slots = (uint32_t)(ctx->fbw * ctx->fbh);
        if ((s = dalloc(ctx, &b.d_slot_off, (size_t)fbn_slots + 1)) != PFH_OK) return s;
        if ((s = dalloc(ctx, &b.d_slot_scan, pfh::scan_partials_words(fbn_slots))) != PFH_OK)
            return s;
        if ((s = dalloc(ctx, &b.d_slot_tiles, (size_t)h.tile_count + 1)) != PFH_OK) return s;
    }
    if ((s = dalloc(ctx, &b.d_tile_runs, h.tile_count)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_col_alpha, h.backdrops.size() + 1)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_alpha, h.tile_count)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_alpha_words, h.tile_count)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_zbuffer, fbn + 8)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_ftm, fbn)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_row_paths, (size_t)ctx->fbh * std::max<uint32_t>(h.path_count, 1))) !=
        PFH_OK)
        return s;
    if ((s = dalloc(ctx, &b.d_row_counts, (size_t)ctx->fbh)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_list_stats, fbn)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &b.d_counters, 1)) != PFH_OK) return s;
    if ((s = size_capacity_buffers(ctx, b)) != PFH_OK) return s;
    ctx->batches.push_back(std::move(bp));
    return run_batch(ctx, b, front);
}

// ---------------------------------------------------------------- adjoint

// Device buffers of one backward pass beyond its inputs and outputs.
struct BackwardScratch {
    uint32_t *list_off = nullptr;    // list slots per framebuffer tile (+ total), fbn + 1
    uint32_t *list_scan = nullptr;   // scan partials of list_off
    float *trans = nullptr;          // 256 floats per list slot
    uint32_t *list_tiles = nullptr;  // the tile of every list slot
    float *grad_cov = nullptr;       // dL/dC, 256 floats per alpha tile
    uint32_t *grad_stamp = nullptr;  // per alpha tile: the epoch whose grad_cov it holds
    uint32_t epoch = 0;
};

// The checks both backward entry points share; returns the frame's draw batch.
pfh_status backward_check(pfh_context *ctx, const float *grad_image, const float *grad_points,
                          uint32_t point_count, const float *grad_colors, Batch **out) {
    if (!ctx || !grad_image || !grad_points || !grad_colors)
        return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    if (ctx->in_scene) return fail(ctx, PFH_ERROR_STATE, "backward inside a scene");
    if (!ctx->gradients ||
        !(ctx->checkpoint == PFH_CHECKPOINT_COVERAGE ? (void *)ctx->d_coverage
                                                     : (void *)ctx->d_cov_state))
        return fail(ctx, PFH_ERROR_STATE, "the frame was rendered without pfh_set_gradients");
    Batch *draw = nullptr;
    for (auto &b : ctx->batches) {
        if (!b->draw) continue;
        if (draw) return fail(ctx, PFH_ERROR_STATE, "backward supports one draw batch per frame");
        draw = b.get();
    }
    if (!draw) return fail(ctx, PFH_ERROR_STATE, "no draw batch in the frame");
    if (draw->load)
        return fail(ctx, PFH_ERROR_STATE, "backward needs a frame drawn over a clear colour");
    if (!ctx->paints_plain || draw->color_page >= 0)
        return fail(ctx, PFH_ERROR_STATE,
                    "backward differentiates solid paints only (no filter, combine, blend "
                    "mode or colour texture)");
    if (point_count < ctx->draw_points_host.size())
        return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "grad_points holds %u points, the scene %zu",
                    point_count, ctx->draw_points_host.size());
    if (!draw->d_tile_paths)
        return fail(ctx, PFH_ERROR_STATE, "the draw batch was prepared before pfh_set_gradients");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    *out = draw;
    return PFH_OK;
}

uint32_t backward_alpha_end(pfh_context *ctx) {
    uint32_t alpha_end = 0;
    for (auto &x : ctx->batches) alpha_end = std::max(alpha_end, x->counters.frame_alpha_end);
    return std::min(alpha_end, ctx->cap_a);
}

// Exclusive scan of the sorted lists' lengths into sc.list_off; returns the total.
pfh_status backward_list_slots(pfh_context *ctx, Batch &b, const BackwardScratch &sc,
                               uint32_t *slots) {
    hipStream_t s = ctx->stream;
    uint32_t fbn = (uint32_t)(ctx->fbw * ctx->fbh);
    pfh::launch_list_lengths(b.d_list_stats, sc.list_off, fbn, s);
    pfh::launch_scan(sc.list_off, fbn, sc.list_scan, s);
    HIP_TRY(ctx, hipMemcpyAsync(slots, sc.list_off + fbn, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return PFH_OK;
}

pfh_status backward_launch(pfh_context *ctx, Batch &b, const float *grad_image, float *grad_points,
                           float *grad_colors, uint32_t path_count, const BackwardScratch &sc,
                           uint32_t slots, uint32_t alpha_end) {
    hipStream_t s = ctx->stream;
    if (ctx->checkpoint == PFH_CHECKPOINT_RECOMPUTE) {
        // the forward kept nothing: the fill runs again over the draw batch
        // and writes the adjoint's states only
        pfh::BatchArgs fa = batch_args(ctx, b);
        fa.coverage = nullptr;
        fa.cov_state = ctx->d_cov_state;
        fa.states_only = 1;
        pfh::launch_fill(fa, s, nullptr);
    }
    pfh::TileGradArgs tg;
    tg.f = tile_args(ctx, b);
    tg.coverage = ctx->checkpoint == PFH_CHECKPOINT_COVERAGE ? ctx->d_coverage : nullptr;
    tg.cov_state = ctx->checkpoint == PFH_CHECKPOINT_COVERAGE ? nullptr : ctx->d_cov_state;
    tg.alpha_tiles = b.d_alpha;
    tg.alpha_start = b.counters.alpha_start;
    tg.alpha_count = b.counters.alpha;
    tg.grad_image = grad_image;
    tg.width = (uint32_t)ctx->width;
    tg.height = (uint32_t)ctx->height;
    tg.list_offsets = sc.list_off;
    tg.slot_capacity = slots;
    tg.trans = sc.trans;
    // the forward composite recorded every slot's tile (same offsets: the
    // same scan of the same lengths)
    tg.lists_recorded = b.d_slot_tiles ? 1 : 0;
    tg.list_tiles = b.d_slot_tiles ? b.d_slot_tiles : sc.list_tiles;
    tg.grad_stamp = sc.grad_stamp;
    tg.epoch = sc.epoch;
    tg.tile_paths = b.d_tile_paths;
    tg.dice_md = (const uint4 *)b.d_dice;
    tg.path_count = b.host.path_count;
    tg.grad_cov = sc.grad_cov;
    tg.grad_colors = grad_colors;
    tg.color_count = path_count;
    pfh::launch_tile_backward(tg, s);
    pfh::MicrolineGradArgs mg;
    mg.b = batch_args(ctx, b);
    mg.microline_count = b.counters.microline_count;
    mg.grad_cov = sc.grad_cov;
    mg.grad_stamp = sc.grad_stamp;
    mg.epoch = sc.epoch;
    mg.alpha_capacity = alpha_end;
    mg.grad_points = grad_points;
    pfh::launch_microline_backward(mg, s);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return PFH_OK;
}

// ------------------------------------------- scratch of the allocation-free entry point

// pfh_backward_noalloc's device buffers: one array per device, shared by every
// call on that device and kept until the process ends. Host code never uses
// the array's name as an address: each device's address comes from
// hipGetSymbolAddress (scratch_setup, at pfh_context_create).
constexpr size_t kNoAllocScratchBytes = (size_t)PFH_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
bool g_scratch_known[kMaxDevices] = {};
uint8_t *g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// Once per device, with that device current.
void scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_known[device]) return;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? static_cast<uint8_t *>(p) : nullptr;
    g_scratch_known[device] = true;
}

uint8_t *scratch_pointer(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

hipError_t scratch_error(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_known[device] ? g_scratch_err[device] : hipErrorNotInitialized;
}

}  // namespace

extern "C" {

pfh_status pfh_context_create(const pfh_context_options *options, pfh_context **out) {
    if (!out) return PFH_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    pfh_context_options opt = {};
    opt.level = PFH_RENDERER_LEVEL_D3D11;
    if (options) opt = *options;
    if (opt.level != PFH_RENDERER_LEVEL_D3D11 && opt.level != PFH_RENDERER_LEVEL_D3D9)
        return PFH_ERROR_INVALID_ARGUMENT;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= opt.device || opt.device < 0)
        return PFH_ERROR_NO_DEVICE;
    std::unique_ptr<pfh_context> ctx(new (std::nothrow) pfh_context());
    if (!ctx) return PFH_ERROR_OUT_OF_MEMORY;
    ctx->device = opt.device;
    ctx->prof.mode = opt.enable_timing ? PFH_PROFILE_ALL : PFH_PROFILE_OFF;
    if (opt.initial_microline_capacity) ctx->cap_m = next_pow2(opt.initial_microline_capacity);
    if (opt.initial_fill_capacity) ctx->cap_f = next_pow2(opt.initial_fill_capacity);
    if (hipSetDevice(opt.device) != hipSuccess) return PFH_ERROR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, opt.device) != hipSuccess) return PFH_ERROR_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return PFH_ERROR_NO_DEVICE;
    scratch_setup(opt.device);  // pfh_backward_noalloc's scratch address on this device
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
        return PFH_ERROR_HIP;
    for (hipStream_t &side : ctx->side)
        if (hipStreamCreateWithFlags(&side, hipStreamNonBlocking) != hipSuccess)
            return PFH_ERROR_HIP;
    if (hipEventCreateWithFlags(&ctx->frame_start, hipEventDisableTiming) != hipSuccess)
        return PFH_ERROR_HIP;
    // Renderer::new uploads the area LUT (gpu/renderer.rs:207-222).
    uint8_t *d_bytes = nullptr;
    if (dalloc(ctx.get(), &d_bytes, 256 * 256 * 4) != PFH_OK) return PFH_ERROR_OUT_OF_MEMORY;
    // the kernels' padded layout: 259 rows of 260 texels
    // 259 rows of 260 padded texels, each a texel and its difference to the next
    if (dalloc(ctx.get(), &ctx->d_lut, 2 * 259 * 260) != PFH_OK) return PFH_ERROR_OUT_OF_MEMORY;
    if (dalloc(ctx.get(), &ctx->d_frame, 1) != PFH_OK) return PFH_ERROR_OUT_OF_MEMORY;
    if (dalloc(ctx.get(), &ctx->d_lut_sat, kLutSatRows) != PFH_OK)
        return PFH_ERROR_OUT_OF_MEMORY;
    const uint8_t *lut = opt.area_lut_rgba ? opt.area_lut_rgba : pfh_area_lut_rgba;
    if (hipMemcpy(d_bytes, lut, 256 * 256 * 4, hipMemcpyHostToDevice) != hipSuccess)
        return PFH_ERROR_HIP;
    {
        std::vector<uint32_t> sat(kLutSatRows);
        lut_saturation(lut, sat.data());
        if (hipMemcpy(ctx->d_lut_sat, sat.data(), sat.size() * 4, hipMemcpyHostToDevice) !=
            hipSuccess)
            return PFH_ERROR_HIP;
    }
    pfh::launch_lut_expand(d_bytes, ctx->d_lut, ctx->stream);
    pfh::launch_frame_reset(ctx->d_frame, ctx->stream);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return PFH_ERROR_HIP;
    dfree(ctx.get(), d_bytes);
    *out = ctx.release();
    return PFH_OK;
}

void pfh_context_destroy(pfh_context *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (hipStream_t side : ctx->side)
        if (side) (void)hipStreamSynchronize(side);
    drop_graph(ctx);
    free_batches(ctx);
    ctx->pool.release_all();
    for (hipStream_t side : ctx->side)
        if (side) (void)hipStreamDestroy(side);
    if (ctx->frame_start) (void)hipEventDestroy(ctx->frame_start);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *pfh_last_error(const pfh_context *ctx) { return ctx ? ctx->error.c_str() : ""; }

pfh_status pfh_begin_scene(pfh_context *ctx, const pfh_frame_options *frame) {
    if (!ctx || !frame) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    if (frame->width <= 0 || frame->height <= 0 || frame->width > 32768 || frame->height > 32768)
        return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "bad framebuffer size %dx%d", frame->width,
                    frame->height);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (hipStream_t side : ctx->side) HIP_TRY(ctx, hipStreamSynchronize(side));
    drop_graph(ctx);
    free_batches(ctx);
    int32_t fbw = (frame->width + 15) / 16, fbh = (frame->height + 15) / 16;
    size_t words = (size_t)fbw * fbh * 256;
    if (!ctx->d_fb || words != ctx->fb_words) {
        dfree(ctx, ctx->d_fb);
        dfree(ctx, ctx->d_fb_saved);
        pfh_status s = dalloc(ctx, &ctx->d_fb, words);
        if (s != PFH_OK) return s;
        HIP_TRY(ctx, hipMemsetAsync(ctx->d_fb, 0, words * 4, ctx->stream));
        ctx->fb_words = words;
    }
    ctx->width = frame->width;
    ctx->height = frame->height;
    ctx->fbw = fbw;
    ctx->fbh = fbh;
    ctx->has_background = frame->has_background != 0;
    ctx->background[0] = frame->background_color.r;
    ctx->background[1] = frame->background_color.g;
    ctx->background[2] = frame->background_color.b;
    ctx->background[3] = frame->background_color.a;
    ctx->dest_dirty = false;
    ctx->d9_pending.clear();
    ctx->d9_alpha_count = 0;
    ctx->d9_mask_zeroed = 0;
    // A frame that starts by loading the destination must start from the same
    // pixels when it is rendered again.
    if (!ctx->has_background) {
        if (!ctx->d_fb_saved) {
            pfh_status s = dalloc(ctx, &ctx->d_fb_saved, words);
            if (s != PFH_OK) return s;
        }
        HIP_TRY(ctx, hipMemcpyAsync(ctx->d_fb_saved, ctx->d_fb, words * 4,
                                    hipMemcpyDeviceToDevice, ctx->stream));
    } else {
        dfree(ctx, ctx->d_fb_saved);
    }
    pfh_status s = ensure_mask(ctx, 0);
    if (s != PFH_OK) return s;
    pfh::launch_frame_reset(ctx->d_frame, ctx->stream);
    if ((s = start_frame_on_side_streams(ctx)) != PFH_OK) return s;
    std::memset(&ctx->stats, 0, sizeof(ctx->stats));
    ctx->in_scene = true;
    return PFH_OK;
}

pfh_status pfh_d3d9_add_fills(pfh_context *ctx, const pfh_fill_d3d9 *fills, uint32_t count) {
    if (!ctx || (count && !fills)) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    if (!ctx->in_scene) return fail(ctx, PFH_ERROR_STATE, "add_fills outside a scene");
    for (uint32_t i = 0; i < count; i++) {
        if (fills[i].alpha_tile_index >= (1u << 24))
            return fail(ctx, PFH_ERROR_LIMIT, "alpha tile %u beyond 2^24",
                        fills[i].alpha_tile_index);
        ctx->d9_alpha_count = std::max(ctx->d9_alpha_count, fills[i].alpha_tile_index + 1);
    }
    ctx->d9_pending.insert(ctx->d9_pending.end(), fills, fills + count);
    return PFH_OK;
}

namespace {
// Mask storage for the frame's alpha tiles: grown keeping what earlier
// flushes drew, the tiles not yet cleared this frame zeroed.
pfh_status d9_ensure_mask(pfh_context *ctx, uint32_t tiles) {
    if (tiles > ctx->d9_mask_cap) {
        uint32_t cap = std::max<uint32_t>(ctx->d9_mask_cap ? ctx->d9_mask_cap : 4096, 4096);
        while (cap < tiles) cap *= 2;
        uint16_t *m = nullptr;
        pfh_status s = dalloc(ctx, &m, (size_t)cap * 256);
        if (s != PFH_OK) return s;
        if (ctx->d9_mask_zeroed)
            HIP_TRY(ctx, hipMemcpyAsync(m, ctx->d_d9_mask, (size_t)ctx->d9_mask_zeroed * 512,
                                        hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        dfree(ctx, ctx->d_d9_mask);
        ctx->d_d9_mask = m;
        ctx->d9_mask_cap = cap;
    }
    if (tiles > ctx->d9_mask_zeroed) {
        HIP_TRY(ctx, hipMemsetAsync(ctx->d_d9_mask + (size_t)ctx->d9_mask_zeroed * 256, 0,
                                    (size_t)(tiles - ctx->d9_mask_zeroed) * 512, ctx->stream));
        ctx->d9_mask_zeroed = tiles;
    }
    return PFH_OK;
}
}  // namespace

pfh_status pfh_d3d9_flush_fills(pfh_context *ctx) {
    if (!ctx) return PFH_ERROR_INVALID_ARGUMENT;
    if (!ctx->in_scene) return fail(ctx, PFH_ERROR_STATE, "flush_fills outside a scene");
    if (ctx->d9_pending.empty()) return PFH_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    pfh_status s = d9_ensure_mask(ctx, ctx->d9_alpha_count);
    if (s != PFH_OK) return s;
    // group the fills by alpha tile, keeping their order (a stable counting
    // sort): one job per alpha tile that has fills
    const std::vector<pfh_fill_d3d9> &f = ctx->d9_pending;
    std::vector<uint32_t> count(ctx->d9_alpha_count + 1, 0);
    for (const pfh_fill_d3d9 &x : f) count[x.alpha_tile_index + 1]++;
    std::vector<uint4> jobs;
    for (uint32_t t = 0; t < ctx->d9_alpha_count; t++) {
        if (count[t + 1]) jobs.push_back(make_uint4(t, count[t], count[t] + count[t + 1], 0));
        count[t + 1] += count[t];
    }
    std::vector<uint2> grouped(f.size());
    for (const pfh_fill_d3d9 &x : f)
        grouped[count[x.alpha_tile_index]++] =
            make_uint2((uint32_t)x.from_x | ((uint32_t)x.from_y << 16),
                       (uint32_t)x.to_x | ((uint32_t)x.to_y << 16));
    uint2 *d_fills = nullptr;
    uint4 *d_jobs = nullptr;
    if ((s = upload(ctx, &d_fills, grouped)) != PFH_OK) return s;
    if ((s = upload(ctx, &d_jobs, jobs)) != PFH_OK) return s;
    pfh::D3D9FillArgs a;
    a.fills = d_fills;
    a.jobs = d_jobs;
    a.job_count = (uint32_t)jobs.size();
    a.mask = ctx->d_d9_mask;
    a.alpha_capacity = ctx->d9_mask_cap;
    a.lut = ctx->d_lut;
    pfh::launch_d3d9_fill(a, ctx->stream);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ctx, d_fills);
    dfree(ctx, d_jobs);
    ctx->stats.fill_count += (uint32_t)f.size();
    ctx->d9_pending.clear();
    return PFH_OK;
}

pfh_status pfh_d3d9_draw_tiles(pfh_context *ctx, const pfh_draw_tile_batch_d3d9 *batch) {
    if (!ctx || !batch || (batch->tile_count && !batch->tiles) ||
        (batch->clip_count && !batch->clips))
        return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    if (!ctx->in_scene) return fail(ctx, PFH_ERROR_STATE, "draw_tiles outside a scene");
    const pfh_rect_i &zr = batch->z_buffer_rect;
    int64_t zw = std::max(zr.max_x - zr.min_x, 0), zh = std::max(zr.max_y - zr.min_y, 0);
    if (zw * zh && !batch->z_buffer) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "no z-buffer");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    pfh_status s = pfh_d3d9_flush_fills(ctx);
    if (s != PFH_OK) return s;
    if ((s = d9_ensure_mask(ctx, ctx->d9_alpha_count)) != PFH_OK) return s;
    // clip_tiles
    if (batch->clip_count) {
        std::vector<int4> clips(batch->clip_count);
        std::memcpy(clips.data(), batch->clips, clips.size() * sizeof(int4));
        int4 *d_clips = nullptr;
        if ((s = upload(ctx, &d_clips, clips)) != PFH_OK) return s;
        pfh::D3D9ClipArgs c;
        c.clips = d_clips;
        c.clip_count = batch->clip_count;
        c.mask = ctx->d_d9_mask;
        c.alpha_capacity = std::min(ctx->d9_mask_cap, ctx->d9_mask_zeroed);
        pfh::launch_d3d9_clip(c, ctx->stream);
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        dfree(ctx, d_clips);
    }
    // draw_tiles: the tiles over each framebuffer tile in draw order, a tile
    // below its z value left out (tile.vs.glsl's z test)
    uint32_t fbw = (uint32_t)ctx->fbw, fbh = (uint32_t)ctx->fbh;
    size_t fbn = (size_t)fbw * fbh;
    std::vector<uint32_t> offsets(fbn + 1, 0), keep;
    keep.reserve(batch->tile_count);
    auto fb_index = [&](const pfh_tile_object_d3d9 &t, size_t *g) {
        if (t.tile_x < 0 || t.tile_y < 0 || (uint32_t)t.tile_x >= fbw || (uint32_t)t.tile_y >= fbh)
            return false;
        int64_t zx = t.tile_x - zr.min_x, zy = t.tile_y - zr.min_y;
        int32_t z = (zx >= 0 && zy >= 0 && zx < zw && zy < zh) ? batch->z_buffer[zy * zw + zx] : 0;
        if ((int64_t)t.path_id < (int64_t)z) return false;
        *g = (size_t)t.tile_y * fbw + (size_t)t.tile_x;
        return true;
    };
    for (uint32_t i = 0; i < batch->tile_count; i++) {
        size_t g;
        if (fb_index(batch->tiles[i], &g)) offsets[g + 1]++;
    }
    for (size_t g = 0; g < fbn; g++) offsets[g + 1] += offsets[g];
    std::vector<uint32_t> list(offsets[fbn]), fill_pos(offsets.begin(), offsets.end() - 1);
    for (uint32_t i = 0; i < batch->tile_count; i++) {
        size_t g;
        if (fb_index(batch->tiles[i], &g)) list[fill_pos[g]++] = i;
    }
    std::vector<uint4> tiles(batch->tile_count);
    std::memcpy(tiles.data(), batch->tiles, tiles.size() * sizeof(uint4));
    uint4 *d_tiles = nullptr;
    uint32_t *d_offsets = nullptr, *d_list = nullptr;
    if ((s = upload(ctx, &d_tiles, tiles)) != PFH_OK) return s;
    if ((s = upload(ctx, &d_offsets, offsets)) != PFH_OK) return s;
    if ((s = upload(ctx, &d_list, list)) != PFH_OK) return s;
    pfh::D3D9TileArgs t;
    t.tiles = d_tiles;
    t.list_offsets = d_offsets;
    t.list = d_list;
    t.mask = ctx->d_d9_mask;
    t.alpha_capacity = std::min(ctx->d9_mask_cap, ctx->d9_mask_zeroed);
    t.metadata = ctx->d_metadata;
    t.paint_count = ctx->paint_count;
    t.framebuffer = ctx->d_fb;
    t.fb_width_tiles = (int32_t)fbw;
    t.fb_height_tiles = (int32_t)fbh;
    t.load = (ctx->dest_dirty || !ctx->has_background) ? 1 : 0;
    for (int i = 0; i < 4; i++) t.clear[i] = t.load ? 0.0f : ctx->background[i];
    ctx->dest_dirty = true;
    pfh::launch_d3d9_tile(t, ctx->stream);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ctx, d_tiles);
    dfree(ctx, d_offsets);
    dfree(ctx, d_list);
    ctx->stats.total_tile_count += batch->tile_count;
    ctx->stats.alpha_tile_count = ctx->d9_alpha_count;
    return PFH_OK;
}

pfh_status pfh_upload_texture_metadata_entries(pfh_context *ctx,
                                               const pfh_texture_metadata_entry *entries,
                                               uint32_t count) {
    if (!ctx || (count && !entries)) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    if (count > 65536) return fail(ctx, PFH_ERROR_LIMIT, "more than 65536 paints (u16 PaintId)");
    std::vector<uint16_t> halves((size_t)std::max<uint32_t>(count, 1) * 40, 0);
    bool plain = true;
    for (uint32_t i = 0; i < count; i++) {
        float t[40];
        if (!encode_metadata(entries[i], t))
            return fail(ctx, PFH_ERROR_INVALID_ARGUMENT,
                        "paint %u: unknown filter %d (the text filter is not on this path)", i,
                        entries[i].filter.kind);
        for (int k = 0; k < 40; k++) halves[(size_t)i * 40 + k] = f32_to_f16(t[k]);
        plain = plain && t[32] == 0.0f;
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ctx, ctx->d_metadata);
    pfh_status s = dalloc(ctx, &ctx->d_metadata, (size_t)std::max<uint32_t>(count, 1) * 10);
    if (s != PFH_OK) return s;
    uint16_t *d_halves = nullptr;
    if ((s = upload(ctx, &d_halves, halves)) != PFH_OK) return s;
    pfh::launch_metadata_expand(d_halves, ctx->d_metadata, count * 10, ctx->stream);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ctx, d_halves);
    ctx->paint_count = count;
    ctx->paints_plain = plain;
    drop_graph(ctx);
    return PFH_OK;
}

pfh_status pfh_upload_texture_metadata(pfh_context *ctx, const pfh_color_u *colors,
                                       uint32_t count) {
    if (!ctx || (count && !colors)) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    // a solid paint's TextureMetadataEntry (paint.rs:641-659): identity colour
    // transform, no combine, no filter, SrcOver
    std::vector<pfh_texture_metadata_entry> e(count);
    for (uint32_t i = 0; i < count; i++) {
        std::memset(&e[i], 0, sizeof(e[i]));
        e[i].color_0_transform = pfh_transform2f{1, 0, 0, 1, 0, 0};
        e[i].base_color = colors[i];
    }
    return pfh_upload_texture_metadata_entries(ctx, e.data(), count);
}

pfh_status pfh_allocate_texture_page(pfh_context *ctx, uint32_t page_id, int32_t width,
                                     int32_t height) {
    if (!ctx) return PFH_ERROR_INVALID_ARGUMENT;
    if (width <= 0 || height <= 0 || width > 16384 || height > 16384)
        return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "bad texture page size %dx%d", width, height);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (hipStream_t side : ctx->side) HIP_TRY(ctx, hipStreamSynchronize(side));
    pfh_context::Page &pg = ctx->pages[page_id];
    dfree(ctx, pg.texels);
    pg.width = width;
    pg.height = height;
    pfh_status s = dalloc(ctx, &pg.texels, (size_t)width * height);
    if (s != PFH_OK) return s;
    HIP_TRY(ctx, hipMemsetAsync(pg.texels, 0, (size_t)width * height * 4, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    return PFH_OK;
}

pfh_status pfh_upload_texel_data(pfh_context *ctx, uint32_t page_id, const pfh_rect_i *rect,
                                 const pfh_color_u *texels) {
    if (!ctx || !rect || !texels) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    auto it = ctx->pages.find(page_id);
    if (it == ctx->pages.end())
        return fail(ctx, PFH_ERROR_STATE, "texture page %u is not allocated", page_id);
    const pfh_context::Page &pg = it->second;
    int32_t w = rect->max_x - rect->min_x, h = rect->max_y - rect->min_y;
    if (rect->min_x < 0 || rect->min_y < 0 || w < 0 || h < 0 || rect->max_x > pg.width ||
        rect->max_y > pg.height)
        return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "texel rect outside page %u", page_id);
    if (!w || !h) return PFH_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    // on the frame's stream, after the work that may still read the page
    HIP_TRY(ctx, hipMemcpy2DAsync(pg.texels + (size_t)rect->min_y * pg.width + rect->min_x,
                                  (size_t)pg.width * 4, texels, (size_t)w * 4, (size_t)w * 4,
                                  (size_t)h, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return PFH_OK;
}

pfh_status pfh_upload_scene(pfh_context *ctx, const pfh_segments *draw, const pfh_segments *clip) {
    if (!ctx || !draw || !clip) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->draw_points_host.assign(draw->points, draw->points + draw->point_count);
    ctx->draw_idx_host.assign(draw->indices, draw->indices + draw->index_count);
    ctx->clip_points_host.assign(clip->points, clip->points + clip->point_count);
    ctx->clip_idx_host.assign(clip->indices, clip->indices + clip->index_count);
    dfree(ctx, ctx->d_draw_points); dfree(ctx, ctx->d_draw_idx);
    dfree(ctx, ctx->d_clip_points); dfree(ctx, ctx->d_clip_idx);
    pfh_status s;
    if ((s = dalloc(ctx, &ctx->d_draw_points, draw->point_count)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &ctx->d_draw_idx, draw->index_count)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &ctx->d_clip_points, clip->point_count)) != PFH_OK) return s;
    if ((s = dalloc(ctx, &ctx->d_clip_idx, clip->index_count)) != PFH_OK) return s;
    if (draw->point_count)
        HIP_TRY(ctx, hipMemcpyAsync(ctx->d_draw_points, ctx->draw_points_host.data(),
                                    (size_t)draw->point_count * 8, hipMemcpyHostToDevice,
                                    ctx->stream));
    if (draw->index_count)
        HIP_TRY(ctx, hipMemcpyAsync(ctx->d_draw_idx, ctx->draw_idx_host.data(),
                                    (size_t)draw->index_count * 8, hipMemcpyHostToDevice,
                                    ctx->stream));
    if (clip->point_count)
        HIP_TRY(ctx, hipMemcpyAsync(ctx->d_clip_points, ctx->clip_points_host.data(),
                                    (size_t)clip->point_count * 8, hipMemcpyHostToDevice,
                                    ctx->stream));
    if (clip->index_count)
        HIP_TRY(ctx, hipMemcpyAsync(ctx->d_clip_idx, ctx->clip_idx_host.data(),
                                    (size_t)clip->index_count * 8, hipMemcpyHostToDevice,
                                    ctx->stream));
    // the batches' front stages (side streams) read the scene
    return start_frame_on_side_streams(ctx);
}

pfh_status pfh_prepare_tiles(pfh_context *ctx, const pfh_tile_batch *batch) {
    return add_batch(ctx, batch, false);
}

pfh_status pfh_prepare_and_draw_tiles(pfh_context *ctx, const pfh_tile_batch *batch) {
    return add_batch(ctx, batch, true);
}

static pfh_status collect(pfh_context *ctx, bool *overflow) {
    *overflow = false;
    pfh_status js = join_side_streams(ctx);
    if (js != PFH_OK) return js;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    uint64_t max_m = 0, max_f = 0, alpha_end = 0;
    for (auto &b : ctx->batches) {
        HIP_TRY(ctx, hipMemcpy(&b->counters, b->d_counters, sizeof(pfh::BatchCounters),
                               hipMemcpyDeviceToHost));
        if (b->counters.error)
            return fail(ctx, PFH_ERROR_INVALID_ARGUMENT,
                        "batch %u: segment, point or path index out of range (code %u)",
                        b->host.batch_id, b->counters.error);
        max_m = std::max<uint64_t>(max_m, b->counters.microlines);
        max_f = std::max<uint64_t>(max_f, b->counters.fills);
        alpha_end = std::max<uint64_t>(alpha_end, b->counters.frame_alpha_end);
        if (b->counters.overflow) *overflow = true;
    }
    if (alpha_end >= (1u << 23))
        return fail(ctx, PFH_ERROR_LIMIT, "%llu alpha tiles exceed the 24-bit signed alpha tile id",
                    (unsigned long long)alpha_end);
    if (*overflow) {
        if (max_m > ctx->cap_m) ctx->cap_m = next_pow2(max_m);
        if (max_f > ctx->cap_f) ctx->cap_f = next_pow2(max_f);
        pfh_status s = ensure_mask(ctx, (uint32_t)alpha_end);
        if (s != PFH_OK) return s;
    }
    return PFH_OK;
}

pfh_status pfh_end_scene(pfh_context *ctx) {
    if (!ctx) return PFH_ERROR_INVALID_ARGUMENT;
    if (!ctx->in_scene) return fail(ctx, PFH_ERROR_STATE, "end_scene without begin_scene");
    ctx->in_scene = false;
    bool overflow = false;
    pfh_status s = collect(ctx, &overflow);
    for (int round = 0; s == PFH_OK && overflow && round < 4; round++) {
        drop_graph(ctx);  // capacities grew: a captured frame is stale
        ctx->replays++;
        if ((s = replay_frame(ctx)) != PFH_OK) break;
        s = collect(ctx, &overflow);
    }
    if (s != PFH_OK) return s;
    if (overflow) return fail(ctx, PFH_ERROR_OUT_OF_MEMORY, "buffers kept overflowing");
    pfh_render_stats &st = ctx->stats;
    std::memset(&st, 0, sizeof(st));
    for (auto &b : ctx->batches) {
        st.path_count += b->host.path_count;
        st.fill_count += b->counters.fills;
        st.alpha_tile_count += b->counters.alpha;
        st.total_tile_count += b->host.tile_count;
        st.microline_count += b->counters.microlines;
        st.segment_count += b->host.segment_count;
        st.column_count += b->host.backdrops.size();
        st.drawcall_count += b->draw ? 7 : 6;
    }
    st.gpu_bytes_allocated = ctx->pool.bytes();
    st.overflow_replays = ctx->replays;
    return PFH_OK;
}

pfh_status pfh_rerender_scene(pfh_context *ctx) {
    if (!ctx) return PFH_ERROR_INVALID_ARGUMENT;
    if (ctx->in_scene) return fail(ctx, PFH_ERROR_STATE, "rerender inside a scene");
    // per-kernel timing records events, which a graph would not replay as timings
    if (!ctx->use_graphs || ctx->prof.mode != PFH_PROFILE_OFF) return replay_frame(ctx);
    if (!ctx->graph_exec) {
        pfh_status s = capture_frame(ctx);
        if (s != PFH_OK) {
            drop_graph(ctx);
            return s;
        }
    }
    HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
    return PFH_OK;
}

pfh_status pfh_set_launch_shape(pfh_context *ctx, const pfh_launch_shape *shape) {
    if (!ctx || !shape) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    if (shape->grid_cap < 1 || shape->grid_cap > 65536 || shape->fill_grid < 1 ||
        shape->fill_grid > 65536)
        return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "launch shape out of range (1..65536)");
    if (ctx->in_scene) return fail(ctx, PFH_ERROR_STATE, "set_launch_shape inside a scene");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    ctx->shape = *shape;
    return PFH_OK;
}

pfh_status pfh_get_launch_shape(pfh_context *ctx, pfh_launch_shape *shape) {
    if (!ctx || !shape) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    *shape = ctx->shape;
    return PFH_OK;
}

// Device time of `frames` re-rendered frames, between events on the context
// stream (every side stream's work joins it before the frame ends).
static pfh_status time_frames(pfh_context *ctx, uint32_t frames, float *ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(ctx, hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    if (err != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(ctx, PFH_ERROR_HIP, "hipEventCreate: %s", hipGetErrorString(err));
    }
    pfh_status s = replay_frame(ctx);  // warm-up, and any buffer sizing
    if (s == PFH_OK) err = hipEventRecord(e0, ctx->stream);
    for (uint32_t i = 0; s == PFH_OK && err == hipSuccess && i < frames; i++) s = replay_frame(ctx);
    if (s == PFH_OK && err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
    if (s == PFH_OK && err == hipSuccess) err = hipEventSynchronize(e1);
    if (s == PFH_OK && err == hipSuccess) err = hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (s != PFH_OK) return s;
    if (err != hipSuccess) return fail(ctx, PFH_ERROR_HIP, "timing: %s", hipGetErrorString(err));
    return PFH_OK;
}

// The same through a captured graph of the frame (captured once, then
// replayed: one warm-up and `frames` timed launches).
static pfh_status time_graph_frames(pfh_context *ctx, uint32_t frames, float *ms) {
    drop_graph(ctx);
    pfh_status s = capture_frame(ctx);
    if (s != PFH_OK) return s;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(ctx, hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    if (err == hipSuccess) err = hipGraphLaunch(ctx->graph_exec, ctx->stream);
    if (err == hipSuccess) err = hipEventRecord(e0, ctx->stream);
    for (uint32_t i = 0; err == hipSuccess && i < frames; i++)
        err = hipGraphLaunch(ctx->graph_exec, ctx->stream);
    if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (err != hipSuccess) return fail(ctx, PFH_ERROR_HIP, "graph timing: %s", hipGetErrorString(err));
    return PFH_OK;
}

pfh_status pfh_autotune(pfh_context *ctx, uint32_t frames, pfh_launch_shape *chosen) {
    if (!ctx) return PFH_ERROR_INVALID_ARGUMENT;
    if (ctx->in_scene) return fail(ctx, PFH_ERROR_STATE, "autotune inside a scene");
    if (ctx->batches.empty()) return fail(ctx, PFH_ERROR_STATE, "autotune needs a rendered scene");
    if (frames < 1) frames = 1;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    drop_graph(ctx);
    int saved_prof = ctx->prof.mode;
    ctx->prof.mode = PFH_PROFILE_OFF;
    static const uint32_t kCandidates[] = {1024, 2048, 4096, 8192, 16384};
    pfh_launch_shape best = ctx->shape;
    pfh_status s = PFH_OK;
    // one parameter at a time: the fill kernel's grid, then the other kernels' cap
    for (int param = 0; param < 2 && s == PFH_OK; param++) {
        float best_ms = 0.0f;
        uint32_t best_v = param == 0 ? best.fill_grid : best.grid_cap;
        bool have = false;
        for (uint32_t v : kCandidates) {
            ctx->shape = best;
            (param == 0 ? ctx->shape.fill_grid : ctx->shape.grid_cap) = v;
            float ms = 0.0f;
            if ((s = time_frames(ctx, frames, &ms)) != PFH_OK) break;
            if (!have || ms < best_ms) {
                best_ms = ms;
                best_v = v;
                have = true;
            }
        }
        (param == 0 ? best.fill_grid : best.grid_cap) = best_v;
    }
    ctx->shape = best;
    // then eager launches against a replayed graph of the frame with that
    // shape: the faster one is kept for pfh_rerender_scene
    float eager_ms = 0.0f, graph_ms = 0.0f;
    if (s == PFH_OK) s = time_frames(ctx, frames, &eager_ms);
    if (s == PFH_OK) s = time_graph_frames(ctx, frames, &graph_ms);
    if (s == PFH_OK) {
        ctx->use_graphs = graph_ms < eager_ms;
        if (!ctx->use_graphs) drop_graph(ctx);
    } else {
        drop_graph(ctx);
    }
    ctx->prof.mode = saved_prof;
    if (s != PFH_OK) return s;
    if (chosen) *chosen = best;
    return PFH_OK;
}

pfh_status pfh_get_graphs(pfh_context *ctx, int32_t *enabled) {
    if (!ctx || !enabled) return PFH_ERROR_INVALID_ARGUMENT;
    *enabled = ctx->use_graphs ? 1 : 0;
    return PFH_OK;
}

pfh_status pfh_set_graphs(pfh_context *ctx, int32_t enable) {
    if (!ctx) return PFH_ERROR_INVALID_ARGUMENT;
    if (ctx->in_scene) return fail(ctx, PFH_ERROR_STATE, "set_graphs inside a scene");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    ctx->use_graphs = enable != 0;
    return PFH_OK;
}

pfh_status pfh_synchronize(pfh_context *ctx) {
    if (!ctx) return PFH_ERROR_INVALID_ARGUMENT;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return PFH_OK;
}

pfh_status pfh_read_framebuffer(pfh_context *ctx, uint8_t *rgba, size_t stride) {
    if (!ctx || !rgba) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    if (!ctx->d_fb) return fail(ctx, PFH_ERROR_STATE, "no framebuffer yet");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (stride == 0) stride = (size_t)ctx->width * 4;
    HIP_TRY(ctx, hipMemcpy2D(rgba, stride, ctx->d_fb, (size_t)ctx->fbw * 16 * 4,
                             (size_t)ctx->width * 4, (size_t)ctx->height, hipMemcpyDeviceToHost));
    return PFH_OK;
}

pfh_status pfh_read_debug(pfh_context *ctx, uint32_t batch_id, pfh_debug_kind kind, void *dst,
                          size_t capacity, size_t *out_len) {
    if (!ctx || !out_len) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    Batch *b = find_batch(ctx, batch_id);
    if (!b) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "no batch %u", batch_id);
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipMemcpy(&b->counters, b->d_counters, sizeof(pfh::BatchCounters),
                           hipMemcpyDeviceToHost));
    size_t fbn = (size_t)ctx->fbw * ctx->fbh;
    const void *src = nullptr;
    size_t bytes = 0;
    uint32_t range[2] = {b->counters.alpha_start, b->counters.alpha_start + b->counters.alpha};
    switch (kind) {
    case PFH_DEBUG_MICROLINES: src = b->d_microlines; bytes = (size_t)b->counters.microline_count * 16; break;
    case PFH_DEBUG_FILLS: bytes = (size_t)b->counters.fill_count * 12; break;
    case PFH_DEBUG_TILES: bytes = (size_t)b->host.tile_count * 16; break;
    case PFH_DEBUG_Z_BUFFER: src = b->d_zbuffer; bytes = (fbn + 8) * 4; break;
    case PFH_DEBUG_FIRST_TILE_MAP: src = b->d_ftm; bytes = fbn * 4; break;
    case PFH_DEBUG_ALPHA_TILES: src = b->d_alpha; bytes = (size_t)b->counters.alpha * 8; break;
    case PFH_DEBUG_BACKDROPS: src = b->d_backdrops; bytes = b->host.backdrops.size() * 12; break;
    case PFH_DEBUG_ALPHA_RANGE: bytes = 8; break;
    case PFH_DEBUG_COVERAGE:
        if (!ctx->gradients || !ctx->d_coverage)
            return fail(ctx, PFH_ERROR_STATE, "no coverage checkpoint (pfh_set_gradients)");
        src = ctx->d_coverage + (size_t)b->counters.alpha_start * 256;
        bytes = (size_t)std::min(b->counters.alpha, ctx->cap_a - std::min(b->counters.alpha_start, ctx->cap_a)) * 1024;
        break;
    case PFH_DEBUG_COVERAGE_STATES:
        if (!ctx->gradients || !ctx->d_cov_state)
            return fail(ctx, PFH_ERROR_STATE, "no coverage states (pfh_set_gradient_checkpoint)");
        src = ctx->d_cov_state + (size_t)b->counters.alpha_start * 64;
        bytes = (size_t)std::min(b->counters.alpha, ctx->cap_a - std::min(b->counters.alpha_start, ctx->cap_a)) * 64;
        break;
    default: return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "bad debug kind %d", (int)kind);
    }
    *out_len = bytes;
    if (!dst) return PFH_OK;  // size query
    if (capacity < bytes) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "buffer too small");
    if (kind == PFH_DEBUG_ALPHA_RANGE) {
        std::memcpy(dst, range, 8);
    } else if (kind == PFH_DEBUG_FILLS && bytes) {
        // the reference's 12-byte Fill from its two device arrays
        size_t n = b->counters.fill_count;
        std::vector<uint2> words(n);
        std::vector<uint32_t> links(n);
        HIP_TRY(ctx, hipMemcpy(words.data(), b->d_fill_words, n * 8, hipMemcpyDeviceToHost));
        HIP_TRY(ctx, hipMemcpy(links.data(), b->d_fill_links, n * 4, hipMemcpyDeviceToHost));
        uint32_t *out = static_cast<uint32_t *>(dst);
        for (size_t i = 0; i < n; i++) {
            out[3 * i] = words[i].x;
            out[3 * i + 1] = words[i].y;
            out[3 * i + 2] = links[i];
        }
    } else if (kind == PFH_DEBUG_TILES && bytes) {
        // TileD3D11: words 0 and 1 come from their own arrays, where the tile
        // has them (linked: a list link; fills: a fill list head), else ~0
        size_t n = b->host.tile_count;
        std::vector<uint32_t> next(n), first(n);
        std::vector<uint8_t> status(n), flags(n);
        std::vector<uint2> words(n);
        HIP_TRY(ctx, hipMemcpy(words.data(), b->d_tiles, n * 8, hipMemcpyDeviceToHost));
        HIP_TRY(ctx, hipMemcpy(next.data(), b->d_tile_next, n * 4, hipMemcpyDeviceToHost));
        HIP_TRY(ctx, hipMemcpy(first.data(), b->d_tile_first, n * 4, hipMemcpyDeviceToHost));
        HIP_TRY(ctx, hipMemcpy(status.data(), b->d_tile_status, n, hipMemcpyDeviceToHost));
        HIP_TRY(ctx, hipMemcpy(flags.data(), b->d_tile_flags, n, hipMemcpyDeviceToHost));
        uint32_t *out = static_cast<uint32_t *>(dst);
        for (size_t i = 0; i < n; i++) {
            out[4 * i] = (status[i] & pfh::kTileLinked) ? next[i] : 0xffffffffu;
            out[4 * i + 1] = flags[i] ? first[i] : 0xffffffffu;
            out[4 * i + 2] = words[i].x;
            out[4 * i + 3] = words[i].y;
        }
    } else if (bytes) {
        HIP_TRY(ctx, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    }
    return PFH_OK;
}

pfh_status pfh_read_mask(pfh_context *ctx, uint32_t *dst, size_t capacity_words,
                         size_t *out_words) {
    if (!ctx || !out_words) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t alpha = 0;
    for (auto &b : ctx->batches) {
        HIP_TRY(ctx, hipMemcpy(&b->counters, b->d_counters, sizeof(pfh::BatchCounters),
                               hipMemcpyDeviceToHost));
        alpha = std::max(alpha, b->counters.frame_alpha_end);
    }
    alpha = std::min(alpha, ctx->cap_a);
    *out_words = (size_t)alpha * 64;
    if (!dst) return PFH_OK;
    if (capacity_words < *out_words)
        return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "buffer too small");
    if (alpha) HIP_TRY(ctx, hipMemcpy(dst, ctx->d_mask, *out_words * 4, hipMemcpyDeviceToHost));
    return PFH_OK;
}

pfh_status pfh_set_gradients(pfh_context *ctx, int32_t enable) {
    if (!ctx) return PFH_ERROR_INVALID_ARGUMENT;
    if (ctx->in_scene) return fail(ctx, PFH_ERROR_STATE, "set_gradients inside a scene");
    drop_graph(ctx);
    ctx->gradients = enable != 0;
    if (!ctx->gradients) return PFH_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return ensure_mask(ctx, 0);
}

pfh_status pfh_set_gradient_checkpoint(pfh_context *ctx, int32_t mode) {
    if (!ctx) return PFH_ERROR_INVALID_ARGUMENT;
    if (ctx->in_scene) return fail(ctx, PFH_ERROR_STATE, "set_gradient_checkpoint inside a scene");
    if (mode != PFH_CHECKPOINT_COVERAGE && mode != PFH_CHECKPOINT_STATES &&
        mode != PFH_CHECKPOINT_RECOMPUTE)
        return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "unknown checkpoint mode %d", mode);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    ctx->checkpoint = mode;
    if (mode != PFH_CHECKPOINT_COVERAGE) {
        dfree(ctx, ctx->d_coverage);
        ctx->d_coverage = nullptr;
    } else {
        dfree(ctx, ctx->d_cov_state);
        ctx->d_cov_state = nullptr;
    }
    return ensure_coverage(ctx);
}

pfh_status pfh_framebuffer_to_float(pfh_context *ctx, float *dst) {
    if (!ctx || !dst) return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "null argument");
    if (!ctx->d_fb) return fail(ctx, PFH_ERROR_STATE, "no framebuffer yet");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    pfh::launch_framebuffer_to_float(ctx->d_fb, (uint32_t)ctx->fbw * 16, (uint32_t)ctx->width,
                                     (uint32_t)ctx->height, dst, ctx->stream);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return PFH_OK;
}

pfh_status pfh_backward(pfh_context *ctx, const float *grad_image, float *grad_points,
                        uint32_t point_count, float *grad_colors, uint32_t path_count) {
    Batch *draw = nullptr;
    pfh_status st = backward_check(ctx, grad_image, grad_points, point_count, grad_colors, &draw);
    if (st != PFH_OK) return st;
    Batch &b = *draw;
    hipStream_t s = ctx->stream;
    // list slots per framebuffer tile: exclusive scan of the sorted lists' lengths
    uint32_t fbn = (uint32_t)(ctx->fbw * ctx->fbh);
    if (ctx->list_off_cap < (size_t)fbn + 1) {
        HIP_TRY(ctx, hipStreamSynchronize(s));
        dfree(ctx, ctx->d_list_off);
        dfree(ctx, ctx->d_list_scan);
        ctx->d_list_off = ctx->d_list_scan = nullptr;
        if ((st = dalloc(ctx, &ctx->d_list_off, (size_t)fbn + 1)) != PFH_OK) return st;
        if ((st = dalloc(ctx, &ctx->d_list_scan, pfh::scan_partials_words(fbn))) != PFH_OK)
            return st;
        ctx->list_off_cap = (size_t)fbn + 1;
    }
    BackwardScratch sc;
    sc.list_off = ctx->d_list_off;
    sc.list_scan = ctx->d_list_scan;
    uint32_t slots = 0;
    if ((st = backward_list_slots(ctx, b, sc, &slots)) != PFH_OK) return st;
    if (ctx->trans_cap < (size_t)slots * 256 || !ctx->d_trans) {
        dfree(ctx, ctx->d_trans);
        dfree(ctx, ctx->d_list_tiles);
        ctx->d_trans = nullptr;
        ctx->d_list_tiles = nullptr;
        ctx->trans_cap = 0;
        size_t n = std::max<size_t>(slots, 1);
        if ((st = dalloc(ctx, &ctx->d_trans, n * 256)) != PFH_OK) return st;
        if ((st = dalloc(ctx, &ctx->d_list_tiles, n)) != PFH_OK) return st;
        ctx->trans_cap = n * 256;
    }
    uint32_t alpha_end = backward_alpha_end(ctx);
    if (ctx->grad_cov_cap < (size_t)alpha_end * 256 || !ctx->d_grad_cov) {
        dfree(ctx, ctx->d_grad_cov);
        dfree(ctx, ctx->d_grad_stamp);
        ctx->d_grad_cov = nullptr;
        ctx->d_grad_stamp = nullptr;
        ctx->grad_cov_cap = 0;
        size_t n = std::max<size_t>(alpha_end, 1);
        if ((st = dalloc(ctx, &ctx->d_grad_cov, n * 256)) != PFH_OK) return st;
        if ((st = dalloc(ctx, &ctx->d_grad_stamp, n)) != PFH_OK) return st;
        ctx->grad_cov_cap = n * 256;
        ctx->grad_epoch = 0;
    }
    // grad_cov entries count only where their stamp is this pass's epoch
    if (++ctx->grad_epoch <= 1) {  // fresh stamps, or the counter wrapped
        pfh::launch_fill_zero((float *)ctx->d_grad_stamp, ctx->grad_cov_cap / 256, s);
        ctx->grad_epoch = 1;
    }
    sc.trans = ctx->d_trans;
    sc.list_tiles = ctx->d_list_tiles;
    sc.grad_cov = ctx->d_grad_cov;
    sc.grad_stamp = ctx->d_grad_stamp;
    sc.epoch = ctx->grad_epoch;
    return backward_launch(ctx, b, grad_image, grad_points, grad_colors, path_count, sc, slots,
                           alpha_end);
}

pfh_status pfh_backward_noalloc(pfh_context *ctx, const float *grad_image, float *grad_points,
                                uint32_t point_count, float *grad_colors, uint32_t path_count) {
    Batch *draw = nullptr;
    pfh_status st = backward_check(ctx, grad_image, grad_points, point_count, grad_colors, &draw);
    if (st != PFH_OK) return st;
    uint8_t *base = scratch_pointer(ctx->device);
    if (!base)
        return fail(ctx, PFH_ERROR_HIP, "the scratch array's device address is unknown (%s)",
                    hipGetErrorString(scratch_error(ctx->device)));
    Batch &b = *draw;
    uint32_t fbn = (uint32_t)(ctx->fbw * ctx->fbh);
    uint32_t alpha_end = backward_alpha_end(ctx);
    // the scratch array, carved in 256-byte aligned regions
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        size_t at = off;
        off += (bytes + 255) & ~(size_t)255;
        return at;
    };
    size_t at_off = carve(((size_t)fbn + 1) * 4);
    size_t at_scan = carve(pfh::scan_partials_words(fbn) * 4);
    size_t at_cov = carve((size_t)std::max<uint32_t>(alpha_end, 1) * 1024);
    size_t at_stamp = carve((size_t)std::max<uint32_t>(alpha_end, 1) * 4);
    if (off > kNoAllocScratchBytes)
        return fail(ctx, PFH_ERROR_LIMIT,
                    "backward_noalloc: %u framebuffer tiles and %u alpha tiles need %zu bytes "
                    "of scratch before the list slots; the limit is %zu",
                    fbn, alpha_end, off, kNoAllocScratchBytes);
    BackwardScratch sc;
    sc.list_off = reinterpret_cast<uint32_t *>(base + at_off);
    sc.list_scan = reinterpret_cast<uint32_t *>(base + at_scan);
    uint32_t slots = 0;
    if ((st = backward_list_slots(ctx, b, sc, &slots)) != PFH_OK) return st;
    size_t at_trans = carve((size_t)std::max<uint32_t>(slots, 1) * 1024);
    size_t at_tiles = carve((size_t)std::max<uint32_t>(slots, 1) * 4);
    if (off > kNoAllocScratchBytes)
        return fail(ctx, PFH_ERROR_LIMIT,
                    "backward_noalloc: %u framebuffer tiles, %u list slots and %u alpha tiles "
                    "need %zu bytes of scratch; the limit is %zu",
                    fbn, slots, alpha_end, off, kNoAllocScratchBytes);
    sc.trans = reinterpret_cast<float *>(base + at_trans);
    sc.list_tiles = reinterpret_cast<uint32_t *>(base + at_tiles);
    sc.grad_cov = reinterpret_cast<float *>(base + at_cov);
    sc.grad_stamp = reinterpret_cast<uint32_t *>(base + at_stamp);
    // the array keeps the last call's contents (any context on this device):
    // every stamp starts at zero, and this pass's epoch is 1
    pfh::launch_fill_zero(reinterpret_cast<float *>(sc.grad_stamp),
                          std::max<uint32_t>(alpha_end, 1), ctx->stream);
    sc.epoch = 1;
    return backward_launch(ctx, b, grad_image, grad_points, grad_colors, path_count, sc, slots,
                           alpha_end);
}

pfh_status pfh_get_stats(pfh_context *ctx, pfh_render_stats *stats) {
    if (!ctx || !stats) return PFH_ERROR_INVALID_ARGUMENT;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    pfh_render_stats &st = ctx->stats;
    st.linked_tile_count = st.composited_tile_count = st.composited_alpha_tile_count = 0;
    st.clipped_alpha_tile_count = 0;
    size_t fbn = (size_t)ctx->fbw * ctx->fbh;
    std::vector<uint4> ls(fbn);
    for (auto &b : ctx->batches) {
        if (fbn)
            HIP_TRY(ctx, hipMemcpy(ls.data(), b->d_list_stats, fbn * sizeof(uint4),
                                   hipMemcpyDeviceToHost));
        for (const uint4 &v : ls) {
            st.linked_tile_count += v.x;
            if (b->draw) {
                st.composited_tile_count += v.y;
                st.composited_alpha_tile_count += v.z;
            }
        }
        std::vector<uint2> at(b->counters.alpha);
        if (!at.empty())
            HIP_TRY(ctx, hipMemcpy(at.data(), b->d_alpha, at.size() * sizeof(uint2),
                                   hipMemcpyDeviceToHost));
        for (const uint2 &v : at) st.clipped_alpha_tile_count += (int32_t)v.y >= 0 ? 1 : 0;
    }
    *stats = st;
    return PFH_OK;
}

uint32_t pfh_profile_kernel_count(void) { return pfh::K_COUNT; }

const char *pfh_profile_kernel_name(uint32_t kernel) {
    return kernel < (uint32_t)pfh::K_COUNT ? pfh::kKernelNames[kernel] : nullptr;
}

pfh_status pfh_profile_set(pfh_context *ctx, int32_t mode, int32_t kernel) {
    if (!ctx || mode < PFH_PROFILE_OFF || mode > PFH_PROFILE_SERIAL ||
        (mode == PFH_PROFILE_ONE && (kernel < 0 || kernel >= pfh::K_COUNT)))
        return fail(ctx, PFH_ERROR_INVALID_ARGUMENT, "bad profile mode %d / kernel %d", mode,
                    kernel);
    ctx->prof.mode = mode;
    ctx->prof.only = kernel;
    return PFH_OK;
}

pfh_status pfh_profile_reset(pfh_context *ctx) {
    if (!ctx) return PFH_ERROR_INVALID_ARGUMENT;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->prof.reset();
    return PFH_OK;
}

pfh_status pfh_profile_read(pfh_context *ctx, pfh_kernel_time *out, uint32_t capacity,
                            uint32_t *count) {
    if (!ctx || !count) return PFH_ERROR_INVALID_ARGUMENT;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->prof.collect();
    *count = pfh::K_COUNT;
    if (!out) return PFH_OK;
    for (uint32_t i = 0; i < (uint32_t)pfh::K_COUNT && i < capacity; i++) {
        out[i].name = pfh::kKernelNames[i];
        out[i].launches = ctx->prof.launches[i];
        out[i].total_ms = ctx->prof.total_ms[i];
    }
    return PFH_OK;
}

}  // extern "C"
