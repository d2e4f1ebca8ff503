// The reference's heightmap asset path on the host: an 8- or 16-bit PNG decoded the way
// RenderEngine::load2DTexture does it (src/render-engine.cpp:1308-1322) - stb_image
// forcing RGBA and flipping rows so row 0 is the image's bottom - keeping the red
// channel the water program samples (heightmap.r, assets/shaders/water-grid.vert:106),
// as the R8 texels of a wt_heightmap.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace wavetool {

// Decodes a non-interlaced PNG of colour type grey, grey+alpha, RGB, RGBA or palette, bit
// depth 8 (16 for the non-palette types, reduced to the high byte as stb_image does).
// texels receives width x height bytes, row 0 = the bottom image row. Returns false with
// a message for a file it cannot read, a corrupt stream or an unsupported format.
bool load_heightmap_png(const std::string& path, std::vector<uint8_t>* texels, uint32_t* width, uint32_t* height,
                        std::string* error);

}  // namespace wavetool
