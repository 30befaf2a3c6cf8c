// pfh_palette.h - paints with a colour texture: gradients and image
// patterns, and the palette build that turns a scene's paints into texture
// metadata, colour texture pages and texel uploads.
//
// Follows renderer/src/paint.rs (Paint :53-146, build_paint_info :399-437,
// assign_paint_locations :456-591, calculate_texture_transforms :593-638,
// create_texture_metadata :640-659, GradientTileBuilder :789-846),
// content/src/gradient.rs (Gradient::add :143-153, sample :181-210) and
// content/src/pattern.rs, in binary32 with the reference's operation order.
#ifndef PFH_PALETTE_H
#define PFH_PALETTE_H

#include "../../include/pfhip.h"

#include <string>
#include <vector>

namespace pfh {

struct GradientDesc {
    int32_t geometry = PFH_GRADIENT_LINEAR;
    pfh_vector2f from{0, 0}, to{0, 0};
    float radius0 = 0, radius1 = 0;
    pfh_transform2f transform{1, 0, 0, 1, 0, 0};
    std::vector<pfh_color_stop> stops;  // stably sorted by offset
    int32_t wrap = PFH_GRADIENT_WRAP_CLAMP;
};

struct PatternDesc {
    int32_t width = 0, height = 0;
    std::vector<pfh_color_u> pixels;
    pfh_transform2f transform{1, 0, 0, 1, 0, 0};
    uint32_t flags = 0;
    pfh_filter filter{};
};

enum PaintOverlay { OVERLAY_NONE = 0, OVERLAY_GRADIENT = 1, OVERLAY_PATTERN = 2 };

struct PaintDesc {
    pfh_color_u base{0, 0, 0, 255};
    int32_t overlay = OVERLAY_NONE;
    int32_t composite_op = PFH_PAINT_COMPOSITE_SRC_IN;
    GradientDesc gradient;
    PatternDesc pattern;
};

// Validates and copies a gradient or pattern description; false if invalid.
bool paint_from_gradient(const pfh_color_u &base, const pfh_gradient &g, int32_t composite_op,
                         PaintDesc *out);
bool paint_from_pattern(const pfh_color_u &base, const pfh_pattern &p, int32_t composite_op,
                        PaintDesc *out);

// Palette::push_paint's cache key: equal keys are equal paints.
std::string paint_key(const PaintDesc &paint);

// Paint::is_opaque (paint.rs:162-176).
bool paint_is_opaque(const PaintDesc &paint);

// TileBatchTexture (gpu_data.rs): page -1 for a paint without colour texture.
struct TileBatchTexture {
    int32_t page = -1;
    uint32_t sampling_flags = 0;
    int32_t composite_op = 0;
    bool operator==(const TileBatchTexture &o) const {
        return page == o.page && sampling_flags == o.sampling_flags &&
               composite_op == o.composite_op;
    }
    bool operator!=(const TileBatchTexture &o) const { return !(*this == o); }
};

struct PageSize {
    int32_t width, height;
};

struct TexelUpload {
    uint32_t page;
    pfh_rect_i rect;
    std::vector<pfh_color_u> texels;
};

struct PaletteBuild {
    std::vector<pfh_texture_metadata_entry> metadata;  // one per paint
    std::vector<uint8_t> opaque;                       // Paint::is_opaque per paint
    std::vector<TileBatchTexture> textures;            // per paint
    std::vector<PageSize> pages;                       // page sizes, by page id
    std::vector<TexelUpload> uploads;                  // in command order
};

// Palette::build_paint_info for a fresh texture manager: every page is new.
void build_palette(const std::vector<PaintDesc> &paints, const pfh_transform2f &build_transform,
                   PaletteBuild *out);

// Gradient::sample (gradient.rs:181-210).
pfh_color_u gradient_sample(const std::vector<pfh_color_stop> &stops, float t);

}  // namespace pfh

#endif
