// C entry points that drive HIPRaycaster through the Voreen shim, for the
// pytest suite (tests/test_hipraycaster_module.py).  They play the roles of
// the network around the processor: VolumeSource (volume inport),
// MeshEntryExitPoints (entry/exit images), the TF editor (keys), and the
// NetworkEvaluator (isReady / process).
#include <cstring>
#include <memory>
#include <sstream>
#include <string>

#include "hipraycastermodule.h"
#include "processors/hipraycaster.h"

using namespace voreen;

namespace {

struct Exposed : public HIPRaycaster {
    VolumePort& vin() { return volumeInport_; }
    VolumePort* gin() {
        for (Port* p : getPorts())
            if (p->getID() == "volumehandle.gradientvolumehandle") return dynamic_cast<VolumePort*>(p);
        return 0;
    }
    RenderPort& ein() { return entryPort_; }
    RenderPort& xin() { return exitPort_; }
    RenderPort& out() { return outport1_; }
    float step(const VolumeBase* v) { return getSamplingStepSize(v); }
    float renderStep(const VolumeBase* v) { return renderStepSize(v->getRepresentation<VolumeRAM>()); }
};

struct Driver {
    std::unique_ptr<HIPRaycasterModule> module;
    std::unique_ptr<Exposed> proc;
    std::unique_ptr<VolumeBase> volume;
    std::unique_ptr<VolumeBase> gradient;
    bool initialized = false;
};

std::string json_escape(const std::string& s) {
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o;
}

int copy_out(const std::string& s, char* buf, int n) {
    if (!buf || n <= 0) return (int)s.size();
    const int k = (int)s.size() < n - 1 ? (int)s.size() : n - 1;
    std::memcpy(buf, s.data(), k);
    buf[k] = 0;
    return (int)s.size();
}

}  // namespace

extern "C" {

void* hrc_create() {
    Driver* d = new Driver();
    d->module.reset(new HIPRaycasterModule("."));
    d->proc.reset(new Exposed());
    ShimLog::errors().clear();
    ShimLog::warnings().clear();
    return d;
}

void hrc_destroy(void* h) {
    Driver* d = (Driver*)h;
    if (d->initialized) d->proc->runDeinitialize();
    delete d;
}

// JSON description: module, registered class, ports and properties.
int hrc_describe(void* h, char* buf, int n) {
    Driver* d = (Driver*)h;
    std::ostringstream js;
    Processor* reg = d->module->getRegisteredProcessors().at(0);
    js << "{\"module\":\"" << d->module->getID() << "\",\"registered\":\"" << reg->getClassName()
       << "\",\"class\":\"" << d->proc->getClassName() << "\",\"category\":\"" << d->proc->getCategory()
       << "\",\"ports\":[";
    bool first = true;
    for (Port* p : d->proc->getPorts()) {
        js << (first ? "" : ",") << "{\"id\":\"" << p->getID() << "\",\"in\":" << (p->isInport() ? "true" : "false") << "}";
        first = false;
    }
    js << "],\"properties\":{";
    first = true;
    for (Property* p : d->proc->getProperties()) {
        js << (first ? "" : ",") << "\"" << p->getID() << "\":{\"value\":\"" << json_escape(p->str()) << "\",\"options\":[";
        bool f2 = true;
        for (const std::string& k : p->optionKeys()) {
            js << (f2 ? "" : ",") << "\"" << k << "\"";
            f2 = false;
        }
        js << "],\"invalidation\":" << p->getInvalidationLevel() << "}";
        first = false;
    }
    js << "}}";
    return copy_out(js.str(), buf, n);
}

int hrc_set_property(void* h, const char* id, const char* value) {
    Driver* d = (Driver*)h;
    Property* p = d->proc->getProperty(id);
    if (!p) return 1;
    return p->set(value) ? 0 : 2;
}

int hrc_set_volume(void* h, const void* data, const uint32_t dims[3], int dtype, float rwm_scale,
                   float rwm_offset, const float spacing[3], const float offset[3]) {
    Driver* d = (Driver*)h;
    VolumeRAM ram(data, tgt::svec3(dims[0], dims[1], dims[2]), dtype == 0 ? "uint8" : "uint16");
    d->volume.reset(new VolumeBase(ram, RealWorldMapping(rwm_scale, rwm_offset),
                                   tgt::vec3(spacing[0], spacing[1], spacing[2]),
                                   tgt::vec3(offset[0], offset[1], offset[2])));
    d->proc->vin().setData(d->volume.get());
    d->proc->vin().setConnected(true);
    return 0;
}

// Gradient volume inport (RGB/RGBA; dtype 0 uint8, 1 uint16, 2 float).
int hrc_set_gradient_volume(void* h, const void* data, const uint32_t dims[3], int dtype, int channels) {
    Driver* d = (Driver*)h;
    VolumeRAM ram(data, tgt::svec3(dims[0], dims[1], dims[2]),
                  dtype == 0 ? "uint8" : (dtype == 1 ? "uint16" : "float"), (size_t)channels);
    d->gradient.reset(new VolumeBase(ram, RealWorldMapping(), tgt::vec3(1.f, 1.f, 1.f), tgt::vec3(0.f, 0.f, 0.f)));
    VolumePort* g = d->proc->gin();
    if (!g) return 1;
    g->setData(d->gradient.get());
    g->setConnected(true);
    return 0;
}

// The 2D TF editor: painted texture (float RGBA w x hh) and the two domains.
int hrc_set_tf2d(void* h, const float* rgba, int w, int hh, const float domain[4]) {
    Driver* d = (Driver*)h;
    TransFunc2DPrimitivesProperty* p =
        dynamic_cast<TransFunc2DPrimitivesProperty*>(d->proc->getProperty("tFunc2DPrimitives"));
    if (!p) return 1;
    p->get()->setTextureData(rgba, w, hh);
    p->get()->setDomain(tgt::vec2(domain[0], domain[1]), 0);
    p->get()->setDomain(tgt::vec2(domain[2], domain[3]), 1);
    return 0;
}

int hrc_set_tf(void* h, const hipdvr_tf_key* keys, int n, float dom_lo, float dom_hi) {
    Driver* d = (Driver*)h;
    TransFunc1DKeysProperty* p = dynamic_cast<TransFunc1DKeysProperty*>(d->proc->getProperty("tFunc1DKeys"));
    if (!p) return 1;
    p->get()->setKeys(std::vector<hipdvr_tf_key>(keys, keys + n));
    p->get()->setDomain(tgt::vec2(dom_lo, dom_hi));
    return 0;
}

int hrc_set_images(void* h, const float* entry, const float* exit_, int w, int hh) {
    Driver* d = (Driver*)h;
    d->proc->ein().setImage(entry, w, hh);
    d->proc->xin().setImage(exit_, w, hh);
    return 0;
}

int hrc_connect(void* h, int entry, int exit_) {
    Driver* d = (Driver*)h;
    d->proc->ein().setConnected(entry != 0);
    d->proc->xin().setConnected(exit_ != 0);
    return 0;
}

int hrc_is_ready(void* h) { return ((Driver*)h)->proc->isReady() ? 1 : 0; }

// QualityMode.setInteractionMode (the network evaluator's interaction flag)
void hrc_set_interaction(int on) { QualityMode.setInteractionMode(on != 0); }

// the step HIPRaycaster marches with
float hrc_render_step(void* h) {
    Driver* d = (Driver*)h;
    return d->volume ? d->proc->renderStep(d->volume.get()) : -1.f;
}

float hrc_step(void* h) {
    Driver* d = (Driver*)h;
    return d->volume ? d->proc->step(d->volume.get()) : -1.f;
}

// initialize (once) + process; copies the output image (zeros if none was
// written) and returns the number of errors the processor logged.
int hrc_process(void* h, float* out, int w, int hh) {
    Driver* d = (Driver*)h;
    if (!d->initialized) {
        d->proc->runInitialize();
        d->initialized = true;
    }
    const size_t before = ShimLog::errors().size();
    d->proc->runProcess();
    d->proc->vin().setChanged(false);
    if (d->proc->gin()) d->proc->gin()->setChanged(false);
    const std::vector<float>& img = tgt::Texture::lastUpload();
    const size_t n = (size_t)w * hh * 4;
    if (img.size() == n) std::memcpy(out, img.data(), n * sizeof(float));
    else std::memset(out, 0, n * sizeof(float));
    return (int)(ShimLog::errors().size() - before);
}

int hrc_last_error(void* h, char* buf, int n) {
    (void)h;
    return copy_out(ShimLog::errors().empty() ? "" : ShimLog::errors().back(), buf, n);
}

int hrc_stats(void* h, hipdvr_stats* st) {
    *st = ((Driver*)h)->proc->getLastStats();
    return 0;
}

// 1 if the last frame went through the learned path, with its statistics
int hrc_learned_stats(void* h, hipdvr_learned_stats* st) {
    *st = ((Driver*)h)->proc->getLastLearnedStats();
    return ((Driver*)h)->proc->lastFrameLearned() ? 1 : 0;
}

int hrc_invalidation(void* h) { return ((Driver*)h)->proc->getInvalidationLevel(); }

}  // extern "C"
