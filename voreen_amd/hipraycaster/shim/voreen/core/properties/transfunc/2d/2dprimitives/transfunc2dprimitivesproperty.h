// TransFunc2DPrimitivesProperty stand-in (declaration only: HIPRaycaster
// exposes the property for id compatibility and reports 2D TFs as unsupported).
#pragma once
#include "voreen/core/datastructures/volume/volumeram.h"
#include "voreen/core/properties/transfunc/transfunctypeproperty.h"

namespace voreen {

class TransFunc2DPrimitivesProperty : public TransFuncPropertyBase {
public:
    TransFunc2DPrimitivesProperty(const std::string& id, const std::string& gui)
        : TransFuncPropertyBase(id, gui, TFT_2DPRIMITIVES) {}
    void setVolume(const VolumeBase*) {}
    std::string str() const override { return "primitives"; }
    bool set(const std::string&) override { return false; }
};

}  // namespace voreen
