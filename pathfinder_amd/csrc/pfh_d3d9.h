// pfh_d3d9.h - the CPU tiler of the D3D9 level and the batch it feeds.
//
// renderer/src/tiler.rs (Tiler::new :39-63, generate_fills :85-93,
// prepare_tiles :95-161, process_segment :164-183, process_line_segment
// :189-308), renderer/src/builder.rs (ObjectBuilder::add_fill :510-556,
// adjust_alpha_tile_backdrop :585-607, build_paths_on_cpu :224-318, the D3D9
// arm of build_tile_batches_for_draw_path_display_item :998-1041),
// content/src/clip.rs:494-565 and content/src/segment.rs (to_cubic :171-182,
// is_flat :292-300, split :307-345), in binary32 with the reference's
// operation order. Alpha tiles are numbered path by path in build order
// (the reference draws them from an atomic counter shared by its threads).
#ifndef PFH_D3D9_H
#define PFH_D3D9_H

#include "../../include/pfhip.h"

#include <vector>

namespace pfh9 {

struct PathInput {
    uint32_t seg_begin, seg_end;  // into the packed segments
    int32_t fill_rule;            // PFH_FILL_RULE_*
    int32_t clip_path;            // clip path index or -1 (draw paths)
    uint32_t paint;
    bool occludes;                // opaque paint and SrcOver
};

struct Built {
    std::vector<std::vector<pfh_fill_d3d9>> fills;  // per path: clip paths, then draw paths
    std::vector<pfh_tile_object_d3d9> tiles;        // the draw batch, in draw order
    std::vector<pfh_clip_d3d9> clips;
    std::vector<int32_t> z;
    pfh_rect_i z_rect{0, 0, 0, 0};
    uint32_t alpha_tile_count = 0;
};

// Tiles every clip path, then every draw path (on up to `threads` threads),
// and packs the draw batch. points/indices are the packed segments of
// SegmentsD3D11; clip and draw paths index their own arrays.
void build(const pfh_vector2f *clip_points, const pfh_segment_indices *clip_indices,
           const std::vector<PathInput> &clip_paths, const pfh_vector2f *draw_points,
           const pfh_segment_indices *draw_indices, const std::vector<PathInput> &draw_paths,
           const pfh_transform2f &transform, const float view_box[4], int threads, Built *out);

}  // namespace pfh9

#endif
