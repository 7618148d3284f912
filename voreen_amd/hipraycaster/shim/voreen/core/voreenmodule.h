// VoreenModule subset (include/voreen/core/voreenmodule.h:230-285).
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "voreen/core/shim_base.h"

namespace voreen {

class VoreenModule {
public:
    VoreenModule(const std::string& modulePath) : path_(modulePath) {}
    virtual ~VoreenModule() { for (Processor* p : processors_) delete p; }
    virtual std::string getDescription() const { return ""; }
    const std::string& getID() const { return id_; }
    const std::string& getGuiName() const { return gui_; }
    const std::vector<Processor*>& getRegisteredProcessors() const { return processors_; }

protected:
    void setID(const std::string& id) { id_ = id; }
    void setGuiName(const std::string& g) { gui_ = g; }
    void registerProcessor(Processor* p) { processors_.push_back(p); }
    void registerSerializableType(Processor* p) { registerProcessor(p); }

private:
    std::string path_, id_, gui_;
    std::vector<Processor*> processors_;
};

}  // namespace voreen
