// TransFunc2DPrimitivesProperty subset
// (include/voreen/core/properties/transfunc/2d/2dprimitives/transfunc2dprimitivesproperty.h:69).
#pragma once
#include "voreen/core/datastructures/transfunc/2d/2dprimitives/transfunc2dprimitives.h"
#include "voreen/core/datastructures/volume/volumeram.h"
#include "voreen/core/properties/transfunc/transfunctypeproperty.h"

namespace voreen {

class TransFunc2DPrimitivesProperty : public TransFuncPropertyBase {
public:
    TransFunc2DPrimitivesProperty(const std::string& id, const std::string& gui)
        : TransFuncPropertyBase(id, gui, TFT_2DPRIMITIVES), tf_(new TransFunc2DPrimitives()) {}
    TransFunc2DPrimitives* get() const { return tf_.get(); }
    void setVolume(const VolumeBase*) {}
    std::string str() const override { return "primitives"; }
    bool set(const std::string&) override { return false; }

private:
    std::unique_ptr<TransFunc2DPrimitives> tf_;
};

}  // namespace voreen
