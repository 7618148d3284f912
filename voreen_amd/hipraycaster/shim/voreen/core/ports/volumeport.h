// VolumePort / RenderPort subset (include/voreen/core/ports/volumeport.h,
// renderport.h).  A RenderPort holds an RGBA float texture of its size.
#pragma once
#include <memory>

#include "voreen/core/datastructures/volume/volumeram.h"
#include "voreen/core/shim_base.h"

namespace voreen {

class VolumePort : public Port {
public:
    VolumePort(PortDirection d, const std::string& id, const std::string& gui = "", bool allowMultiple = false,
               int inv = Processor::INVALID_PROGRAM)
        : Port(d, id, gui) { (void)allowMultiple; (void)inv; }
    bool hasData() const { return data_ != 0; }
    const VolumeBase* getData() const { return data_; }
    void setData(const VolumeBase* v) { data_ = v; setChanged(true); }

private:
    const VolumeBase* data_ = 0;
};

class RenderPort : public Port {
public:
    RenderPort(PortDirection d, const std::string& id, const std::string& gui = "", bool allowMultiple = false,
               int inv = Processor::INVALID_PROGRAM)
        : Port(d, id, gui) { (void)allowMultiple; (void)inv; }
    tgt::ivec2 getSize() const { return size_; }
    bool hasData() const { return tex_ != 0; }
    tgt::Texture* getColorTexture() { return tex_.get(); }
    const tgt::Texture* getColorTexture() const { return tex_.get(); }
    // test driver: set an RGBA float image (copied)
    void setImage(const float* rgba, int w, int h) {
        GLubyte* d = new GLubyte[(size_t)w * h * 16];
        std::memcpy(d, rgba, (size_t)w * h * 16);
        tex_.reset(new tgt::Texture(tgt::ivec3(w, h, 1), GL_RGBA, GL_RGBA, GL_FLOAT, tgt::Texture::NEAREST,
                                    tgt::Texture::CLAMP, d, true));
        size_ = tgt::ivec2(w, h);
        setConnected(true);
    }
    void activateTarget(const std::string& = "") { active_ = true; }
    void deactivateTarget() { active_ = false; }
    void clearTarget() { tgt::Texture::lastUpload().clear(); }

private:
    std::unique_ptr<tgt::Texture> tex_;
    tgt::ivec2 size_;
    bool active_ = false;
};

}  // namespace voreen
