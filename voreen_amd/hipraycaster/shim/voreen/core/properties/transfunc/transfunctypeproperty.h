// TransFuncTypeProperty subset
// (include/voreen/core/properties/transfunc/transfunctypeproperty.h).
#pragma once
#include "voreen/core/properties/optionproperty.h"

namespace voreen {

enum TransFuncType { TFT_UNSUPPORTED, TFT_1DKEYS, TFT_1DGAUSSIAN, TFT_2DPRIMITIVES };

class TransFuncPropertyBase : public Property {
public:
    TransFuncPropertyBase(const std::string& id, const std::string& gui, TransFuncType t)
        : Property(id, gui, Processor::INVALID_RESULT), type_(t) {}
    TransFuncType getTFType() const { return type_; }

private:
    TransFuncType type_;
};

class TransFuncTypeProperty : public OptionProperty<TransFuncType> {
public:
    TransFuncTypeProperty(const std::string& id, const std::string& gui, int inv = Processor::INVALID_PROGRAM)
        : OptionProperty<TransFuncType>(id, gui, inv) {}
    // option keys of TransFuncTypeProperty::addTransFuncTypeOption
    // (src/core/properties/transfunc/transfunctypeproperty.cpp:113-128, spelling included)
    void addTransFuncProperty(TransFuncPropertyBase* p) {
        const TransFuncType t = p->getTFType();
        const char* key = t == TFT_1DKEYS ? "1D Keys Color Map"
                        : (t == TFT_2DPRIMITIVES ? "2D Primitves Color Map" : "1D Gaussian Color Map");
        addOption(key, key, t);
    }
};

}  // namespace voreen
