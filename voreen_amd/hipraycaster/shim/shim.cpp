// Out-of-line parts of the Voreen shim.
#include "tgt/texture.h"

namespace tgt {

std::vector<float>& Texture::lastUpload() {
    static std::vector<float> img;
    return img;
}

void Texture::uploadTexture() {
    if (dataType_ != GL_FLOAT || !data_) return;
    const float* f = reinterpret_cast<const float*>(data_);
    lastUpload().assign(f, f + (size_t)dims_.x * dims_.y * dims_.z * 4);
}

}  // namespace tgt
