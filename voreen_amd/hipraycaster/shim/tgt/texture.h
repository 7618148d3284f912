// tgt::Texture subset (ext/tgt/texture.h): a host buffer of width x height
// texels.  uploadTexture() of an RGBA float texture records its contents as
// the last rendered image (read by the test driver).
#pragma once
#include <cstring>
#include <vector>

#include "tgt/tgt_gl.h"
#include "tgt/vector.h"

namespace tgt {

class Texture {
public:
    enum Filter { NEAREST = GL_NEAREST, LINEAR = GL_LINEAR };
    enum Wrapping { REPEAT = GL_REPEAT, CLAMP = GL_CLAMP_TO_EDGE };

    Texture(const ivec3& dimensions, GLint format = GL_RGBA, GLint internalformat = GL_RGBA,
            GLenum dataType = GL_UNSIGNED_BYTE, Filter filter = LINEAR, Wrapping wrapping = REPEAT,
            GLubyte* data = 0, bool ownership = true)
        : dims_(dimensions), dataType_(dataType), data_(data), owned_(ownership) {
        (void)format;
        (void)internalformat;
        (void)filter;
        (void)wrapping;
    }
    ~Texture() {
        if (owned_) delete[] data_;
    }
    Texture(const Texture&) = delete;
    Texture& operator=(const Texture&) = delete;

    GLenum getDataType() const { return dataType_; }
    int getWidth() const { return dims_.x; }
    int getHeight() const { return dims_.y; }
    ivec3 getDimensions() const { return dims_; }
    GLubyte* getCpuTextureData() { return data_; }
    const GLubyte* getCpuTextureData() const { return data_; }
    size_t getSizeOnCPU() const {
        return (size_t)dims_.x * dims_.y * dims_.z * 4 * (dataType_ == GL_FLOAT ? 4 : 1);
    }
    void setCpuTextureData(GLubyte* data, bool ownership) {
        if (owned_) delete[] data_;
        data_ = data;
        owned_ = ownership;
    }
    void downloadTexture() const {}
    // Copy of the texels in the requested type (only the stored type is supported).
    GLubyte* downloadTextureToBuffer(GLenum format, GLenum dataType) const {
        (void)format;
        (void)dataType;
        const size_t n = getSizeOnCPU();
        GLubyte* out = new GLubyte[n];
        if (data_) std::memcpy(out, data_, n);
        else std::memset(out, 0, n);
        return out;
    }
    void uploadTexture();

    static std::vector<float>& lastUpload();

private:
    ivec3 dims_;
    GLenum dataType_;
    GLubyte* data_;
    bool owned_;
};

}  // namespace tgt
