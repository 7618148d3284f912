// TransFunc1DKeys subset
// (include/voreen/core/datastructures/transfunc/1d/1dkeys/transfunc1dkeys.h).
// The shim builds its RGBA8 texture and pre-integration tables with
// libhipdvr's restatements (hipdvr_tf_from_keys, hipdvr_preint_build); in a
// Voreen build the real class computes them.
#pragma once
#include <memory>
#include <vector>

#include "hipdvr.h"
#include "tgt/texture.h"
#include "voreen/core/datastructures/transfunc/1d/preintegrationtable.h"

namespace voreen {

// TransFuncMappingKey accessors (1d/1dkeys/utils/transfuncmappingkey.h:97-172)
class TransFuncMappingKey {
public:
    explicit TransFuncMappingKey(const hipdvr_tf_key& k) : k_(k) {}
    float getIntensity() const { return k_.intensity; }
    tgt::col4 getColorL() const { return tgt::col4(k_.color_l[0], k_.color_l[1], k_.color_l[2], k_.color_l[3]); }
    tgt::col4 getColorR() const { return tgt::col4(k_.color_r[0], k_.color_r[1], k_.color_r[2], k_.color_r[3]); }
    bool isSplit() const { return k_.split != 0; }

private:
    hipdvr_tf_key k_;
};

// TransFuncBase::AlphaMode (transfuncbase.h:165-169)
enum AlphaModeShim { TF_ZERO_ALPHA = 0, TF_USE_ALPHA = 1, TF_ONE_ALPHA = 2 };

class TransFunc1DKeys {
public:
    TransFunc1DKeys(int width = 1024) : width_(width), domain_(0.f, 1.f) {
        hipdvr_tf_key k0 = {0.f, 0, {0, 0, 0, 0}, {0, 0, 0, 0}};
        hipdvr_tf_key k1 = {1.f, 0, {255, 255, 255, 255}, {255, 255, 255, 255}};
        keys_ = {k0, k1};   // setToStandardFunc (transfunc1dkeys.cpp:413-421)
    }
    void setKeys(const std::vector<hipdvr_tf_key>& k) {
        keys_ = k;
        views_.clear();
        for (const hipdvr_tf_key& key : keys_) views_.emplace_back(key);
        invalidate();
    }
    // transfunc1dkeys.h:91-102, transfunc1d.h:104,148, transfuncbase.h:173
    int getNumKeys() const { return (int)keys_.size(); }
    const TransFuncMappingKey* getKey(int i) const {
        if (views_.size() != keys_.size()) {
            views_.clear();
            for (const hipdvr_tf_key& key : keys_) views_.emplace_back(key);
        }
        return &views_[(size_t)(i < 0 ? 0 : (i >= (int)views_.size() ? (int)views_.size() - 1 : i))];
    }
    float getGammaValue() const { return 1.f; }
    tgt::vec2 getThreshold() const { return tgt::vec2(0.f, 1.f); }
    AlphaModeShim getAlphaMode() const { return TF_USE_ALPHA; }
    void resize(int width) { if (width != width_) { width_ = width; invalidate(); } }
    int getWidth() const { return width_; }
    void setDomain(const tgt::vec2& d) { domain_ = d; table_.reset(); }
    tgt::vec2 getDomain() const { return domain_; }

    tgt::Texture* getTexture() {
        if (!tex_) {
            GLubyte* data = new GLubyte[4 * (size_t)width_];
            hipdvr_tf_from_keys(keys_.data(), (int32_t)keys_.size(), width_, 1.f, 0.f, 1.f,
                                HIPDVR_ALPHA_USE, data);
            tex_.reset(new tgt::Texture(tgt::ivec3(width_, 1, 1), GL_RGBA, GL_RGBA, GL_UNSIGNED_BYTE,
                                        tgt::Texture::NEAREST, tgt::Texture::CLAMP, data, true));
        }
        return tex_.get();
    }

    // dimension 0 = derive from the TF width, capped at 1024
    const PreIntegrationTable* getPreIntegrationTable(float step, size_t dimension, bool useIntegral) {
        size_t n = dimension ? dimension : (width_ < 1024 ? (size_t)width_ : 1024);
        if (!table_ || table_n_ != n || table_step_ != step || table_integral_ != useIntegral) {
            std::vector<float> t(n * n * 4);
            hipdvr_preint_build(keys_.data(), (int32_t)keys_.size(), 1.f, 0.f, 1.f, HIPDVR_ALPHA_USE,
                                (uint32_t)n, step, useIntegral ? 1 : 0, t.data());
            table_.reset(new PreIntegrationTable(n, std::move(t)));
            table_n_ = n;
            table_step_ = step;
            table_integral_ = useIntegral;
        }
        return table_.get();
    }

private:
    void invalidate() { tex_.reset(); table_.reset(); }
    int width_;
    tgt::vec2 domain_;
    std::vector<hipdvr_tf_key> keys_;
    mutable std::vector<TransFuncMappingKey> views_;
    std::unique_ptr<tgt::Texture> tex_;
    std::unique_ptr<PreIntegrationTable> table_;
    size_t table_n_ = 0;
    float table_step_ = 0.f;
    bool table_integral_ = false;
};

}  // namespace voreen
