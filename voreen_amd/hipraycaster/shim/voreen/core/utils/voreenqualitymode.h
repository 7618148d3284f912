// VoreenQualityMode subset (include/voreen/core/utils/voreenqualitymode.h:
// isInteractionMode, the QualityMode singleton macro) for the HIPRaycaster
// test build: the driver switches interaction mode on and off.
#pragma once

namespace voreen {

class VoreenQualityMode {
public:
    bool isInteractionMode() const { return interaction_; }
    void setInteractionMode(bool on) { interaction_ = on; }
    static VoreenQualityMode& getRef() {
        static VoreenQualityMode q;
        return q;
    }

private:
    bool interaction_ = false;
};

}  // namespace voreen

#define QualityMode ::voreen::VoreenQualityMode::getRef()
