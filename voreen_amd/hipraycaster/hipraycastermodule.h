#ifndef VRN_HIPRAYCASTERMODULE_H
#define VRN_HIPRAYCASTERMODULE_H

#include "voreen/core/voreenmodule.h"

namespace voreen {

/// Registers the HIPRaycaster processor (include/voreen/core/voreenmodule.h:230-285).
class HIPRaycasterModule : public VoreenModule {
public:
    HIPRaycasterModule(const std::string& modulePath);

    virtual std::string getDescription() const {
        return "Direct volume rendering ray march on AMD Instinct MI355X (gfx950) through libhipdvr.";
    }
};

}  // namespace voreen

#endif
