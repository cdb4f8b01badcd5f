// Host image decoders for the glTF loader: PNG (all colour types and bit
// depths, Adam7), baseline / extended-sequential Huffman JPEG, and Radiance
// RGBE (.hdr). LDR images decode to 8-bit RGBA, the layout the reference's
// glTF reader hands to its materials (4 components requested); HDR images
// decode to float RGB.
#ifndef SHB_IMAGE_DECODE_H_
#define SHB_IMAGE_DECODE_H_

#include <cstdint>
#include <string>
#include <vector>

namespace shb {

struct Image8 {
  int width = 0, height = 0;
  std::vector<uint8_t> rgba;  // width * height * 4
};

struct ImageF {
  int width = 0, height = 0;
  std::vector<float> rgb;  // width * height * 3
};

bool decode_png(const uint8_t* data, size_t n, Image8* out, std::string* err);
bool decode_jpeg(const uint8_t* data, size_t n, Image8* out, std::string* err);
bool decode_hdr(const uint8_t* data, size_t n, ImageF* out, std::string* err);
// PNG or JPEG by signature
bool decode_image8(const uint8_t* data, size_t n, Image8* out, std::string* err);
bool is_hdr(const uint8_t* data, size_t n);

}  // namespace shb

#endif  // SHB_IMAGE_DECODE_H_
