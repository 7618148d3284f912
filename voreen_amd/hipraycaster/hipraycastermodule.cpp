#include "hipraycastermodule.h"

#include "processors/hipraycaster.h"

namespace voreen {

HIPRaycasterModule::HIPRaycasterModule(const std::string& modulePath)
    : VoreenModule(modulePath)
{
    setID("HIPRaycaster");
    setGuiName("HIP Raycaster");
    registerProcessor(new HIPRaycaster());
}

}  // namespace voreen
