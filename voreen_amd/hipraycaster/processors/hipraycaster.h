#ifndef VRN_HIPRAYCASTER_H
#define VRN_HIPRAYCASTER_H

#include "voreen/core/processors/volumeraycaster.h"
#include "voreen/core/properties/transfunc/transfunctypeproperty.h"
#include "voreen/core/properties/transfunc/1d/1dkeys/transfunc1dkeysproperty.h"
#include "voreen/core/properties/transfunc/2d/2dprimitives/transfunc2dprimitivesproperty.h"
#include "voreen/core/datastructures/transfunc/2d/2dprimitives/transfunc2dprimitives.h"
#include "voreen/core/properties/optionproperty.h"
#include "voreen/core/properties/intproperty.h"
#include "voreen/core/properties/floatproperty.h"
#include "voreen/core/properties/boolproperty.h"
#include "voreen/core/properties/cameraproperty.h"
#include "voreen/core/ports/volumeport.h"

#include "hipdvr.h"
#include "hipdvr_learned.h"

#include <string>
#include <vector>

namespace voreen {

/**
 * Direct volume rendering on an AMD Instinct MI355X.
 *
 * Drop-in for the ray-march path of CPURaycaster
 * (modules/base/processors/render/cpuraycaster.h:48-93): the same ports
 * (volumehandle.volumehandle, image.entrypoints, image.exitpoints,
 * image.output, volumehandle.gradientvolumehandle) and the same properties
 * (samplingRate, classification, compositing, shading, gradient, light and
 * material from VolumeRaycaster/VolumeRenderer; tFuncType, tFunc1DKeys,
 * tFunc2DPrimitives, textureFilterMode_, preIntegrationTableSize from
 * CPURaycaster; camera as in SingleVolumeRaycaster).  The march runs in the
 * HIP kernels of libhipdvr.so through its C-ABI (hipdvr.h).  Properties added
 * here have ids starting with "hip", among them the learned stepping:
 * hipLearnedStepping (on/off), hipPsnrFloor (dB below which a frame is
 * rendered plainly), hipRemarchStride (the judge re-marches 1/stride^2 of
 * the pixels plainly every frame) and hipLearnedTrainingTFs (random TFs the
 * network is trained on, on the device, once per volume).
 *
 * Invalidation: a new volume (INVALID_PROGRAM on the volume inport,
 * src/core/processors/volumeraycaster.cpp:44) re-uploads the volume; a new TF
 * re-uploads the table; everything else only re-renders.
 */
class HIPRaycaster : public VolumeRaycaster {
public:
    HIPRaycaster();
    virtual ~HIPRaycaster();

    virtual Processor* create() const;

    virtual std::string getClassName() const { return "HIPRaycaster"; }
    virtual std::string getCategory() const  { return "Raycasting"; }
    virtual CodeState getCodeState() const   { return CODE_STATE_TESTING; }

    /// Same rule as CPURaycaster::isReady (cpuraycaster.cpp:82-84).
    virtual bool isReady() const;

    /// Statistics of the last frame (rays, samples, reconstructed samples,
    /// shaded samples, kernel time).
    const hipdvr_stats& getLastStats() const { return lastStats_; }

    /// Whether the last frame went through the learned path, and its
    /// statistics (reconstructed samples, the judge's PSNR, fallback).
    bool lastFrameLearned() const { return lastFrameLearned_; }
    const hipdvr_learned_stats& getLastLearnedStats() const { return lastLearned_; }

protected:
    virtual void setDescriptions() {
        setDescription("Direct volume rendering ray march on an AMD MI355X GPU (HIP). Results match "
                       "CPURaycaster sample for sample; Phong shading follows the GLSL raycaster.");
    }

    virtual void process();
    virtual void initialize();
    virtual void deinitialize();

    /// The march step: 1/(max(dim)*samplingRate), CPURaycaster's rule
    /// (cpuraycaster.cpp:104) in every quality mode.
    float renderStepSize(const VolumeRAM* vram) const;

private:
    bool ensureContext();
    bool ensureVolume(const VolumeBase* vb, const VolumeRAM* vram);
    bool ensureTransFunc(tgt::Texture* tfTexture, const tgt::vec2& domain);
    bool ensureGradientVolume(const VolumeBase* gvb, const VolumeRAM* gram);
    bool ensureTransFunc2D(const TransFunc2DPrimitives* tf);
    void releaseDevice();
    void fillParams(hipdvr_params& p, const VolumeBase* vb, float step);
    void writeOutput(const std::vector<float>& rgba, int w, int h);
    bool fail(const std::string& msg);
    // learned stepping: the options it serves, and the frame through it
    bool learnedEligible(const hipdvr_params& p, bool tf2d, const std::string& cls) const;
    bool renderLearned(const hipdvr_params& params, const float* entry, const float* exit, int w, int h,
                       float* output);
    void releaseLearned();

    VolumePort gradientVolumePort_;

    TransFuncTypeProperty tFuncType_;
    TransFunc1DKeysProperty tFunc1DKeys_;
    TransFunc2DPrimitivesProperty tFunc2DPrimitives_;
    IntOptionProperty texFilterMode_;
    IntOptionProperty preIntegrationTableSize_;
    CameraProperty camera_;

    // HIP-specific
    IntProperty hipDevice_;
    IntOptionProperty hipBrickSize_;
    StringOptionProperty hipKernel_;
    BoolProperty hipLearned_;
    FloatProperty hipPsnrFloor_;
    IntProperty hipRemarchStride_;
    IntProperty hipLearnedTFs_;

    hipdvr_ctx* ctx_;
    hipdvr_volume* dvrVolume_;
    hipdvr_tf* dvrTF_;
    hipdvr_tf* dvrTF2D_;
    const VolumeBase* uploadedVolume_;
    unsigned long volumeGeneration_;      // counts volume uploads
    const VolumeBase* uploadedGradient_;
    unsigned long gradientGeneration_;    // volume upload the gradient volume is attached to
    std::vector<float> uploadedTF2D_;
    tgt::vec4 uploadedDomain2D_;
    std::vector<uint8_t> uploadedTF_;
    tgt::vec2 uploadedDomain_;
    int uploadedBrickLog2_;
    float uploadedPreintStep_;
    std::string uploadedPreintKey_;
    unsigned long tfGeneration_;          // counts 1D TF uploads
    hipdvr_stats lastStats_;

    // learned stepping state: per volume upload, the network trained for it
    hipdvr_learned* learned_;
    hipdvr_mlp* mlp_;
    unsigned long learnedGeneration_;     // volume upload learned_ and mlp_ belong to
    int learnedTFs_;                      // training TFs of mlp_
    std::string preparedKey_;             // TF upload, step and options of the last prepare
    float* dEntry_;
    float* dExit_;
    float* dOut_;
    size_t dBytes_;
    bool lastFrameLearned_;
    hipdvr_learned_stats lastLearned_;

    static const std::string loggerCat_;
};

}  // namespace voreen

#endif
