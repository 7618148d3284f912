// TransFunc2DPrimitives subset
// (include/voreen/core/datastructures/transfunc/2d/2dprimitives/transfunc2dprimitives.h,
//  transfunc2d.h:93-101, transfuncbase.h:106).  The texture the editor's
// primitives are painted into is given directly (float RGBA, width x height).
#pragma once
#include <cstring>
#include <memory>

#include "tgt/texture.h"
#include "tgt/vector.h"

namespace voreen {

class TransFunc2DPrimitives {
public:
    TransFunc2DPrimitives() { domains_[0] = domains_[1] = tgt::vec2(0.f, 1.f); }
    tgt::Texture* getTexture() const { return tex_.get(); }
    tgt::vec2 getDomain(size_t dimension) const { return domains_[dimension]; }
    void setDomain(const tgt::vec2& domain, size_t dimension) { domains_[dimension] = domain; }
    // shim: the painted texture (copied)
    void setTextureData(const float* rgba, int w, int h) {
        GLubyte* d = new GLubyte[(size_t)w * h * 16];
        std::memcpy(d, rgba, (size_t)w * h * 16);
        tex_.reset(new tgt::Texture(tgt::ivec3(w, h, 1), GL_RGBA, GL_RGBA, GL_FLOAT, tgt::Texture::NEAREST,
                                    tgt::Texture::CLAMP, d, true));
    }

private:
    std::unique_ptr<tgt::Texture> tex_;
    tgt::vec2 domains_[2];
};

}  // namespace voreen
