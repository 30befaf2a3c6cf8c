// pfh_internal.h - host-side types shared by the scene front-end and the renderer.
#ifndef PFH_INTERNAL_H
#define PFH_INTERNAL_H

#include <cstdint>
#include <vector>

#include "../../include/pfhip.h"

namespace pfh {

// Owning storage for one TileBatchDataD3D11 (gpu_data.rs:123-168).
struct BatchData {
    uint32_t batch_id = 0;
    uint32_t path_source = PFH_PATH_SOURCE_DRAW;
    uint32_t path_count = 0, tile_count = 0, segment_count = 0;
    pfh_transform2f transform{1, 0, 0, 1, 0, 0};
    std::vector<pfh_dice_metadata> dice_metadata;
    std::vector<pfh_tile_path_info> tile_path_info;
    std::vector<pfh_propagate_metadata> propagate_metadata;
    std::vector<pfh_backdrop_info> backdrops;
    int32_t clip_batch_id = -1;
    int32_t color_texture_page = -1;  // DrawTileBatchD3D11.color_texture
    uint32_t color_texture_sampling_flags = 0;
};

inline void batch_view(const BatchData &b, pfh_tile_batch *out) {
    out->batch_id = b.batch_id;
    out->path_source = b.path_source;
    out->path_count = b.path_count;
    out->tile_count = b.tile_count;
    out->segment_count = b.segment_count;
    out->transform = b.transform;
    out->dice_metadata = b.dice_metadata.data();
    out->tile_path_info = b.tile_path_info.data();
    out->propagate_metadata = b.propagate_metadata.data();
    out->backdrops = b.backdrops.data();
    out->backdrop_count = (uint32_t)b.backdrops.size();
    out->clip_batch_id = b.clip_batch_id;
    out->color_texture_page = b.color_texture_page;
    out->color_texture_sampling_flags = b.color_texture_sampling_flags;
}

}  // namespace pfh

#endif
