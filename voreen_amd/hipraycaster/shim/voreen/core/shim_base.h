// Processor/Property/Port/logging subset of voreen_core
// (include/voreen/core/processors/processor.h, properties/property.h,
// ports/port.h, ext/tgt/logmanager.h) for the HIPRaycaster test build.
#pragma once
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "tgt/camera.h"
#include "tgt/texture.h"
#include "tgt/vector.h"

namespace voreen {

// Log sink of the shim: LERROR/LWARNING append here (tgt::LogManager stand-in).
struct ShimLog {
    static std::vector<std::string>& errors() { static std::vector<std::string> e; return e; }
    static std::vector<std::string>& warnings() { static std::vector<std::string> w; return w; }
};
#define LERROR(msg) do { std::ostringstream _s; _s << loggerCat_ << ": " << msg; ::voreen::ShimLog::errors().push_back(_s.str()); } while (0)
#define LWARNING(msg) do { std::ostringstream _s; _s << loggerCat_ << ": " << msg; ::voreen::ShimLog::warnings().push_back(_s.str()); } while (0)

class Processor;

class Property {
public:
    Property(const std::string& id, const std::string& gui, int inv) : id_(id), gui_(gui), inv_(inv) {}
    virtual ~Property() {}
    const std::string& getID() const { return id_; }
    const std::string& getGuiName() const { return gui_; }
    int getInvalidationLevel() const { return inv_; }
    void setDescription(const std::string& d) { desc_ = d; }
    // text form of the value, for the test driver
    virtual std::string str() const = 0;
    virtual bool set(const std::string& v) = 0;
    virtual std::vector<std::string> optionKeys() const { return std::vector<std::string>(); }
    Processor* owner_ = 0;

protected:
    void changed();

private:
    std::string id_, gui_, desc_;
    int inv_;
};

class Port {
public:
    enum PortDirection { INPORT, OUTPORT };
    Port(PortDirection d, const std::string& id, const std::string& gui) : dir_(d), id_(id), gui_(gui) {}
    virtual ~Port() {}
    const std::string& getID() const { return id_; }
    bool isInport() const { return dir_ == INPORT; }
    bool isConnected() const { return connected_; }
    void setConnected(bool c) { connected_ = c; }
    bool hasChanged() const { return changed_; }
    void setChanged(bool c) { changed_ = c; }

private:
    PortDirection dir_;
    std::string id_, gui_;
    bool connected_ = false;
    bool changed_ = false;
};

class Processor {
public:
    enum InvalidationLevel { VALID = 0, INVALID_RESULT = 1, INVALID_PATH = 10, INVALID_PROGRAM = 20 };
    enum CodeState { CODE_STATE_OBSOLETE, CODE_STATE_BROKEN, CODE_STATE_EXPERIMENTAL,
                     CODE_STATE_TESTING, CODE_STATE_STABLE };
    virtual ~Processor() {}
    virtual Processor* create() const = 0;
    virtual std::string getClassName() const = 0;
    virtual std::string getCategory() const = 0;
    virtual CodeState getCodeState() const = 0;
    virtual bool isReady() const = 0;
    virtual void invalidate(int inv = INVALID_RESULT) { if (inv > invalidation_) invalidation_ = inv; }
    int getInvalidationLevel() const { return invalidation_; }
    void setValid() { invalidation_ = VALID; }
    const std::vector<Port*>& getPorts() const { return ports_; }
    const std::vector<Property*>& getProperties() const { return props_; }
    Property* getProperty(const std::string& id) const {
        for (Property* p : props_) if (p->getID() == id) return p;
        return 0;
    }
    // NetworkEvaluator's call sequence (src/core/network/networkevaluator.cpp:322-374)
    void runProcess() { process(); setValid(); }
    void runInitialize() { initialize(); }
    void runDeinitialize() { deinitialize(); }
    void describe() { setDescriptions(); }

protected:
    virtual void process() = 0;
    virtual void initialize() {}
    virtual void deinitialize() {}
    virtual void setDescriptions() {}
    void setDescription(const std::string& d) { description_ = d; }
    void addPort(Port& p) { ports_.push_back(&p); }
    void addPort(Port* p) { ports_.push_back(p); }
    void addProperty(Property& p) { p.owner_ = this; props_.push_back(&p); }

private:
    std::vector<Port*> ports_;
    std::vector<Property*> props_;
    std::string description_;
    int invalidation_ = INVALID_PROGRAM;
};

inline void Property::changed() { if (owner_) owner_->invalidate(inv_); }

template <typename T> std::string shim_to_string(const T& v) { std::ostringstream s; s << v; return s.str(); }

template <typename T>
class NumericPropertyShim : public Property {
public:
    NumericPropertyShim(const std::string& id, const std::string& gui, T v, T lo, T hi, int inv)
        : Property(id, gui, inv), v_(v), lo_(lo), hi_(hi) {}
    T get() const { return v_; }
    void set(T v) { v_ = v < lo_ ? lo_ : (v > hi_ ? hi_ : v); changed(); }
    T getMinValue() const { return lo_; }
    T getMaxValue() const { return hi_; }
    std::string str() const override { return shim_to_string(v_); }
    bool set(const std::string& s) override { std::istringstream in(s); T v; if (!(in >> v)) return false; set(v); return true; }

private:
    T v_, lo_, hi_;
};

class FloatProperty : public NumericPropertyShim<float> {
public:
    FloatProperty(const std::string& id, const std::string& gui, float v = 0.f, float lo = 0.f,
                  float hi = 1.f, int inv = Processor::INVALID_RESULT)
        : NumericPropertyShim<float>(id, gui, v, lo, hi, inv) {}
};

class IntProperty : public NumericPropertyShim<int> {
public:
    IntProperty(const std::string& id, const std::string& gui, int v = 0, int lo = 0, int hi = 100,
                int inv = Processor::INVALID_RESULT)
        : NumericPropertyShim<int>(id, gui, v, lo, hi, inv) {}
};

class BoolProperty : public Property {
public:
    BoolProperty(const std::string& id, const std::string& gui, bool v = false, int inv = Processor::INVALID_RESULT)
        : Property(id, gui, inv), v_(v) {}
    bool get() const { return v_; }
    std::string str() const override { return v_ ? "true" : "false"; }
    bool set(const std::string& s) override { v_ = (s == "true" || s == "1"); changed(); return true; }

private:
    bool v_;
};

// FloatVec4Property / ColorProperty / LightSourceProperty stand-in
class Vec4Property : public Property {
public:
    Vec4Property(const std::string& id, const std::string& gui, const tgt::vec4& v, int inv = Processor::INVALID_RESULT)
        : Property(id, gui, inv), v_(v) {}
    const tgt::vec4& get() const { return v_; }
    std::string str() const override {
        std::ostringstream s; s << v_.x << " " << v_.y << " " << v_.z << " " << v_.w; return s.str();
    }
    bool set(const std::string& t) override {
        std::istringstream in(t); tgt::vec4 v; if (!(in >> v.x >> v.y >> v.z >> v.w)) return false;
        v_ = v; changed(); return true;
    }

private:
    tgt::vec4 v_;
};
typedef Vec4Property ColorProperty;
typedef Vec4Property LightSourceProperty;

class CameraProperty : public Property {
public:
    CameraProperty(const std::string& id, const std::string& gui, const tgt::Camera& c = tgt::Camera(),
                   bool adaptToScene = true, int inv = Processor::INVALID_RESULT)
        : Property(id, gui, inv), c_(c) { (void)adaptToScene; }
    const tgt::Camera& get() const { return c_; }
    std::string str() const override {
        tgt::vec3 p = c_.getPosition(); std::ostringstream s; s << p.x << " " << p.y << " " << p.z; return s.str();
    }
    bool set(const std::string& t) override {
        std::istringstream in(t); tgt::vec3 p; if (!(in >> p.x >> p.y >> p.z)) return false;
        c_.setPosition(p); changed(); return true;
    }

private:
    tgt::Camera c_;
};

}  // namespace voreen
