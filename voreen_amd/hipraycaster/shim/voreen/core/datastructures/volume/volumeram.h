// VolumeRAM / VolumeBase / RealWorldMapping subset
// (include/voreen/core/datastructures/volume/volumeram.h, volumebase.h,
//  meta/realworldmappingmetadata.h).  The volume wraps a caller-owned buffer.
#pragma once
#include <string>

#include "tgt/vector.h"

namespace voreen {

class RealWorldMapping {
public:
    RealWorldMapping(float scale = 1.f, float offset = 0.f) : scale_(scale), offset_(offset) {}
    float getScale() const { return scale_; }
    float getOffset() const { return offset_; }
    float normalizedToRealWorld(float n) const { return (n * scale_) + offset_; }

private:
    float scale_, offset_;
};

class VolumeRAM {
public:
    // baseType "uint8", "uint16" or "float"; channels interleaved per voxel
    VolumeRAM(const void* data, const tgt::svec3& dims, const std::string& baseType, size_t channels = 1)
        : data_(data), dims_(dims), type_(baseType), channels_(channels) {}
    const void* getData() const { return data_; }
    tgt::svec3 getDimensions() const { return dims_; }
    std::string getBaseType() const { return type_; }
    size_t getNumChannels() const { return channels_; }
    size_t getBytesPerVoxel() const {
        return (type_ == "uint8" ? 1 : (type_ == "uint16" ? 2 : 4)) * channels_;
    }

private:
    const void* data_;
    tgt::svec3 dims_;
    std::string type_;
    size_t channels_;
};

class VolumeBase {
public:
    VolumeBase(const VolumeRAM& ram, const RealWorldMapping& rwm, const tgt::vec3& spacing,
               const tgt::vec3& offset)
        : ram_(ram), rwm_(rwm), spacing_(spacing), offset_(offset) {}
    template <typename R> const R* getRepresentation() const { return &ram_; }
    tgt::svec3 getDimensions() const { return ram_.getDimensions(); }
    RealWorldMapping getRealWorldMapping() const { return rwm_; }
    tgt::vec3 getSpacing() const { return spacing_; }
    tgt::vec3 getOffset() const { return offset_; }
    tgt::mat4 getWorldToPhysicalMatrix() const { return tgt::mat4(); }
    size_t getNumChannels() const { return ram_.getNumChannels(); }
    size_t getBytesPerVoxel() const { return ram_.getBytesPerVoxel(); }

private:
    VolumeRAM ram_;
    RealWorldMapping rwm_;
    tgt::vec3 spacing_, offset_;
};

}  // namespace voreen
