// RenderProcessor / VolumeRenderer / VolumeRaycaster subset
// (include/voreen/core/processors/renderprocessor.h, volumerenderer.h,
// volumeraycaster.h) with the reference's port ids, property ids, defaults
// and option keys (src/core/processors/volumeraycaster.cpp:42-157,379-388,
// volumerenderer.cpp:41-54, src/core/utils/glsl.cpp:397-409,
// classificationmodes.cpp:84-90).
#pragma once
#include "voreen/core/utils/voreenqualitymode.h"
#include "voreen/core/ports/volumeport.h"
#include "voreen/core/properties/optionproperty.h"

namespace voreen {

class RenderProcessor : public Processor {
protected:
    void renderQuad() {}
};

class VolumeRenderer : public RenderProcessor {
public:
    VolumeRenderer()
        : lightPosition_("lightPosition", "Light Source Position", tgt::vec4(2.3f, 1.5f, 1.5f, 1.f))
        , lightAmbient_("lightAmbient", "Ambient Light", tgt::vec4(0.4f, 0.4f, 0.4f, 1.f))
        , lightDiffuse_("lightDiffuse", "Diffuse Light", tgt::vec4(0.8f, 0.8f, 0.8f, 1.f))
        , lightSpecular_("lightSpecular", "Specular Light", tgt::vec4(0.6f, 0.6f, 0.6f, 1.f))
        , applyLightAttenuation_("applyLightAttenuation", "Apply Attenuation", false, Processor::INVALID_PROGRAM)
        , materialShininess_("materialShininess", "Shininess", 60.f, 0.1f, 128.f)
    {}

protected:
    LightSourceProperty lightPosition_;
    ColorProperty lightAmbient_;
    ColorProperty lightDiffuse_;
    ColorProperty lightSpecular_;
    BoolProperty applyLightAttenuation_;
    FloatProperty materialShininess_;
};

class VolumeRaycaster : public VolumeRenderer {
public:
    VolumeRaycaster()
        : volumeInport_(Port::INPORT, "volumehandle.volumehandle", "Volume Input", false, Processor::INVALID_PROGRAM)
        , entryPort_(Port::INPORT, "image.entrypoints", "Entry-points Input", false, Processor::INVALID_RESULT)
        , exitPort_(Port::INPORT, "image.exitpoints", "Exit-points Input", false, Processor::INVALID_RESULT)
        , outport1_(Port::OUTPORT, "image.output", "Image Output", true, Processor::INVALID_PROGRAM)
        , samplingRate_("samplingRate", "Sampling Rate", 2.f, 0.01f, 20.f)
        , isoValue_("isoValue", "Iso Value", 0.5f, 0.0f, 1.0f)
        , gammaValue1_("gammaValue", "Gamma Value (OP1)", 0.f, -1.f, 1.f)
        , gradientMode_("gradient", "Gradient Calculation", Processor::INVALID_PROGRAM)
        , classificationMode_("classification", "Classification", Processor::INVALID_PROGRAM)
        , shadeMode_("shading", "Shading", Processor::INVALID_PROGRAM)
        , compositingMode1_("compositing", "Compositing (Output 1)", Processor::INVALID_PROGRAM)
        , interactionCoarseness_("interactionCoarseness", "Interaction Coarseness", 3, 1, 8, Processor::VALID)
        , interactionQuality_("interactionQuality", "Interaction Step Size", 1.0f, 0.01f, 1.0f, Processor::VALID)
    {
        addPort(volumeInport_);
        addPort(entryPort_);
        addPort(exitPort_);
        addPort(outport1_);
        gradientMode_.addOption("none", "none");
        gradientMode_.addOption("forward-differences", "Forward Differences");
        gradientMode_.addOption("central-differences", "Central Differences");
        gradientMode_.addOption("sobel", "Sobel");
        gradientMode_.addOption("filtered", "Filtered");
        gradientMode_.select("central-differences");
        classificationMode_.addOption("transfer-function", "Transfer Function");
        classificationMode_.addOption("pre-integrated-gpu", "Pre-Integrated TF (GPU)");
        classificationMode_.addOption("pre-integrated-cpu", "Pre-Integrated TF (CPU)");
        classificationMode_.addOption("pre-integrated-approximation", "Pre-Integrated TF (Approximation)");
        classificationMode_.select("transfer-function");
        const char* shades[] = {"none", "phong-diffuse", "phong-specular", "phong-diffuse-ambient",
                                "phong-diffuse-specular", "phong", "toon", "cook-torrance", "oren-nayar", "ward"};
        for (const char* s : shades) shadeMode_.addOption(s, s);
        shadeMode_.select("phong");
        const char* comps[] = {"dvr", "mip", "mop", "mida", "iso", "fhp", "fhn"};
        for (const char* c : comps) compositingMode1_.addOption(c, c);
        // volumeraycaster.cpp:143-153 (the quality presets are not modelled)
        addProperty(samplingRate_);
        addProperty(interactionCoarseness_);
        addProperty(interactionQuality_);
    }

    // volumeraycaster.cpp:312-322
    float getSamplingStepSize(const VolumeBase* vol) {
        const tgt::svec3 dim = vol->getDimensions();
        float samplingStepSize = 1.f / ((float)tgt::max(dim) * samplingRate_.get());
        if (QualityMode.isInteractionMode())
            samplingStepSize /= interactionQuality_.get();
        return samplingStepSize;
    }

protected:
    virtual void initialize() {}
    virtual void deinitialize() {}

    VolumePort volumeInport_;
    RenderPort entryPort_;
    RenderPort exitPort_;
    RenderPort outport1_;
    FloatProperty samplingRate_;
    FloatProperty isoValue_;
    FloatProperty gammaValue1_;
    StringOptionProperty gradientMode_;
    StringOptionProperty classificationMode_;
    StringOptionProperty shadeMode_;
    StringOptionProperty compositingMode1_;
    IntProperty interactionCoarseness_;
    FloatProperty interactionQuality_;
};

}  // namespace voreen
