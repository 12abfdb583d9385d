// bm_tiff.h -- the input step of the path: multi-page 8-bit TIFF volumes and
// their calibration (bone-morphometry 0.7.0-rc.3, src/tiff_io.rs):
//   read_binary_tiff / read_roi_tiff / read_binary_tiff_impl   src/tiff_io.rs:18-147
//   read_isotropic_spacing_um and its helpers                 src/tiff_io.rs:149-317
// The reference decodes with the tiff crate 0.11.3 (Cargo.toml:18, not vendored);
// this is a decoder of its own for what that crate reads as 8-bit grayscale:
// classic and BigTIFF, either byte order, strips or tiles, no compression,
// PackBits, LZW or Deflate, with or without the horizontal predictor.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace bm {

class Engine;

// What a page may hold (src/tiff_io.rs:12-17).
enum class TiffSamples { Any, ZeroOr255, RoiMask };

struct TiffVolumeInfo {
  uint64_t width = 0, height = 0, depth = 0;
};

// Decodes every page, checks the samples and writes value > threshold as
// 0/1 into out (width*height*depth bytes).  With out null only the page
// directory is read, for the size.  Pages are decoded on up to `threads`
// host threads (0: the hardware's).  Errors: bm::Error with the reference's
// messages.
TiffVolumeInfo read_binary_tiff(const std::string& path, uint8_t threshold, TiffSamples accepted,
                                uint8_t* out, uint64_t capacity, int threads);

// The same into device memory: decoded pages go through pinned staging
// buffers to d_out on stream s while later pages are decoded; returns when
// the volume is resident.
TiffVolumeInfo read_binary_tiff_device(Engine& e, const std::string& path, uint8_t threshold,
                                       TiffSamples accepted, uint8_t* d_out, uint64_t capacity,
                                       int threads, hipStream_t s);

// read_isotropic_spacing_um (src/tiff_io.rs:149-218): [x, y, z] in um.
void read_isotropic_spacing_um(const std::string& path, double out[3]);

}  // namespace bm
