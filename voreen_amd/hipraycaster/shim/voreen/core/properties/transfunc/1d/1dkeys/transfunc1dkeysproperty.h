// TransFunc1DKeysProperty subset
// (include/voreen/core/properties/transfunc/1d/1dkeys/transfunc1dkeysproperty.h).
// setVolume resizes the table to the volume's bit depth, capped at 4096
// (transfuncpropertybase.cpp:134-174, transfunc1dkeysproperty.cpp:94-97).
#pragma once
#include "voreen/core/datastructures/transfunc/1d/1dkeys/transfunc1dkeys.h"
#include "voreen/core/datastructures/volume/volumeram.h"
#include "voreen/core/properties/transfunc/transfunctypeproperty.h"

namespace voreen {

class TransFunc1DKeysProperty : public TransFuncPropertyBase {
public:
    TransFunc1DKeysProperty(const std::string& id, const std::string& gui)
        : TransFuncPropertyBase(id, gui, TFT_1DKEYS) {}
    TransFunc1DKeys* get() { return &tf_; }
    const TransFunc1DKeys* get() const { return &tf_; }
    void setVolume(const VolumeBase* v) {
        if (!v) return;
        const size_t bits = v->getBytesPerVoxel() * 8;
        int max = 1 << (bits > 16 ? 16 : bits);
        tf_.resize(max < 4096 ? max : 4096);
    }
    std::string str() const override { return "keys"; }
    bool set(const std::string&) override { return false; }

private:
    TransFunc1DKeys tf_;
};

}  // namespace voreen
