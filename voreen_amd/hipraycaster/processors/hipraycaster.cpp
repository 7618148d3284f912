#include "hipraycaster.h"

#include "voreen/core/datastructures/transfunc/1d/1dkeys/transfunc1dkeys.h"
#include "voreen/core/datastructures/transfunc/1d/preintegrationtable.h"
#include "voreen/core/datastructures/volume/volumeram.h"
#include "tgt/texture.h"

#include <cstring>

namespace voreen {

const std::string HIPRaycaster::loggerCat_("voreen.hipraycaster.HIPRaycaster");

HIPRaycaster::HIPRaycaster()
    : VolumeRaycaster()
    , gradientVolumePort_(Port::INPORT, "volumehandle.gradientvolumehandle", "Gradient Volume Input")
    , tFuncType_("tFuncType", "TF Type")
    , tFunc1DKeys_("tFunc1DKeys", "1D Color Map")
    , tFunc2DPrimitives_("tFunc2DPrimitives", "2D Color Map")
    , texFilterMode_("textureFilterMode_", "Texture Filtering")
    , preIntegrationTableSize_("preIntegrationTableSize", "Width of pre-integration table")
    , camera_("camera", "Camera", tgt::Camera(tgt::vec3(0.f, 0.f, 3.5f), tgt::vec3(0.f, 0.f, 0.f),
                                              tgt::vec3(0.f, 1.f, 0.f)), true)
    , hipDevice_("hipDevice", "HIP device", 0, 0, 63, Processor::INVALID_PROGRAM)
    , hipBrickSize_("hipBrickSize", "Skipping brick size", Processor::INVALID_PROGRAM)
    , hipKernel_("hipKernel", "March kernel")
    , hipLearned_("hipLearnedStepping", "Learned stepping", false)
    , hipPsnrFloor_("hipPsnrFloor", "Learned stepping: PSNR floor (dB)", 45.f, 20.f, 99.f)
    , hipRemarchStride_("hipRemarchStride", "Learned stepping: judge re-marches 1/n^2 of the pixels", 8, 1, 64)
    , hipLearnedTFs_("hipLearnedTrainingTFs", "Learned stepping: training TFs", 200, 1, 2000,
                     Processor::INVALID_PROGRAM)
    , ctx_(0)
    , dvrVolume_(0)
    , dvrTF_(0)
    , dvrTF2D_(0)
    , uploadedVolume_(0)
    , volumeGeneration_(0)
    , uploadedGradient_(0)
    , gradientGeneration_(0)
    , uploadedDomain_(0.f, 0.f)
    , uploadedBrickLog2_(0)
    , uploadedPreintStep_(-1.f)
    , tfGeneration_(0)
    , learned_(0)
    , mlp_(0)
    , learnedGeneration_(0)
    , learnedTFs_(0)
    , dEntry_(0)
    , dExit_(0)
    , dOut_(0)
    , dBytes_(0)
    , lastFrameLearned_(false)
{
    // ports and properties of CPURaycaster (cpuraycaster.cpp:44-76)
    addPort(gradientVolumePort_);

    addProperty(tFuncType_);
    addProperty(tFunc1DKeys_);
    tFuncType_.addTransFuncProperty(&tFunc1DKeys_);
    addProperty(tFunc2DPrimitives_);
    tFuncType_.addTransFuncProperty(&tFunc2DPrimitives_);

    texFilterMode_.addOption("nearest", "Nearest", GL_NEAREST);
    texFilterMode_.addOption("linear", "Linear", GL_LINEAR);
    texFilterMode_.selectByKey("linear");
    addProperty(texFilterMode_);

    addProperty(classificationMode_);

    preIntegrationTableSize_.addOption("deriveFromBitDepth", "Derive from bit depth", 0);
    preIntegrationTableSize_.addOption("256", "256", 256);
    preIntegrationTableSize_.addOption("512", "512", 512);
    preIntegrationTableSize_.addOption("1024", "1024", 1024);
    preIntegrationTableSize_.select("deriveFromBitDepth");
    addProperty(preIntegrationTableSize_);

    // camera, shading and light as SingleVolumeRaycaster adds them
    // (modules/base/processors/render/singlevolumeraycaster.cpp:107-130); the
    // default shading is "none" so the default image is CPURaycaster's
    addProperty(camera_);
    addProperty(gradientMode_);
    addProperty(shadeMode_);
    shadeMode_.select("none");
    addProperty(compositingMode1_);
    addProperty(isoValue_);
    addProperty(gammaValue1_);
    addProperty(lightPosition_);
    addProperty(lightAmbient_);
    addProperty(lightDiffuse_);
    addProperty(lightSpecular_);
    addProperty(materialShininess_);

    // HIP-specific
    addProperty(hipDevice_);
    hipBrickSize_.addOption("8", "8^3", 3);
    hipBrickSize_.addOption("16", "16^3", 4);
    hipBrickSize_.addOption("32", "32^3", 5);
    hipBrickSize_.selectByKey("8");
    addProperty(hipBrickSize_);
    hipKernel_.addOption("auto", "Fastest exact kernel");
    hipKernel_.addOption("reference", "Reference layout, no skipping");
    hipKernel_.addOption("bricked", "Bricked, empty-space skipping");
    hipKernel_.addOption("cell", "Cell records, empty-space skipping");
    hipKernel_.addOption("slab", "Slabs staged in LDS, empty-space skipping");
    hipKernel_.selectByKey("auto");
    addProperty(hipKernel_);
    // learned stepping (include/hipdvr_learned.h): the network proposes skip
    // distances and step multiples per brick; the plain march re-marches
    // 1/hipRemarchStride^2 of the pixels every frame and the frame falls back
    // to the plain march below hipPsnrFloor
    addProperty(hipLearned_);
    addProperty(hipPsnrFloor_);
    addProperty(hipRemarchStride_);
    addProperty(hipLearnedTFs_);

    std::memset(&lastStats_, 0, sizeof(lastStats_));
    std::memset(&lastLearned_, 0, sizeof(lastLearned_));
}

HIPRaycaster::~HIPRaycaster() {
    releaseDevice();
}

Processor* HIPRaycaster::create() const {
    return new HIPRaycaster();
}

bool HIPRaycaster::isReady() const {
    return volumeInport_.hasData() && entryPort_.isConnected() && exitPort_.isConnected();
}

void HIPRaycaster::initialize() {
    VolumeRaycaster::initialize();
}

void HIPRaycaster::deinitialize() {
    releaseDevice();
    VolumeRaycaster::deinitialize();
}

void HIPRaycaster::releaseLearned() {
    if (learned_) hipdvr_learned_destroy(learned_);
    if (mlp_) hipdvr_mlp_destroy(mlp_);
    if (ctx_) {
        if (dEntry_) hipdvr_buffer_free(ctx_, dEntry_);
        if (dExit_) hipdvr_buffer_free(ctx_, dExit_);
        if (dOut_) hipdvr_buffer_free(ctx_, dOut_);
    }
    learned_ = 0;
    mlp_ = 0;
    learnedGeneration_ = 0;
    learnedTFs_ = 0;
    preparedKey_.clear();
    dEntry_ = dExit_ = dOut_ = 0;
    dBytes_ = 0;
}

void HIPRaycaster::releaseDevice() {
    releaseLearned();
    if (dvrTF_) hipdvr_tf_destroy(dvrTF_);
    if (dvrTF2D_) hipdvr_tf_destroy(dvrTF2D_);
    if (dvrVolume_) hipdvr_volume_destroy(dvrVolume_);
    if (ctx_) hipdvr_destroy(ctx_);
    dvrTF_ = 0;
    dvrTF2D_ = 0;
    uploadedGradient_ = 0;
    uploadedTF2D_.clear();
    dvrVolume_ = 0;
    ctx_ = 0;
    uploadedVolume_ = 0;
    uploadedTF_.clear();
    uploadedPreintKey_.clear();
}

bool HIPRaycaster::fail(const std::string& msg) {
    LERROR(msg);
    return false;
}

bool HIPRaycaster::ensureContext() {
    if (ctx_) return true;
    int rc = hipdvr_create(&ctx_, hipDevice_.get());
    if (rc != HIPDVR_OK) {
        ctx_ = 0;
        return fail("hipdvr_create failed on device " + std::to_string(hipDevice_.get()) +
                    " (status " + std::to_string(rc) + ")");
    }
    return true;
}

// Upload on a new volume or a new brick size (INVALID_PROGRAM).
bool HIPRaycaster::ensureVolume(const VolumeBase* vb, const VolumeRAM* vram) {
    const int brickLog2 = hipBrickSize_.getValue();
    if (dvrVolume_ && uploadedVolume_ == vb && uploadedBrickLog2_ == brickLog2 &&
        !volumeInport_.hasChanged())
        return true;
    if (dvrVolume_) {
        hipdvr_volume_destroy(dvrVolume_);
        dvrVolume_ = 0;
        uploadedVolume_ = 0;
    }
    const std::string type = vram->getBaseType();
    hipdvr_volume_desc d;
    std::memset(&d, 0, sizeof(d));
    if (type == "uint8") d.dtype = HIPDVR_U8;
    else if (type == "uint16") d.dtype = HIPDVR_U16;
    else return fail("HIPRaycaster supports uint8 and uint16 volumes, got " + type);
    if (vram->getNumChannels() != 1) return fail("HIPRaycaster supports single-channel volumes");
    const tgt::svec3 dims = vram->getDimensions();
    for (int i = 0; i < 3; ++i) d.dims[i] = static_cast<uint32_t>(dims[i]);
    const RealWorldMapping rwm = vb->getRealWorldMapping();
    d.rwm_scale = rwm.getScale();
    d.rwm_offset = rwm.getOffset();
    const tgt::vec3 sp = vb->getSpacing(), off = vb->getOffset();
    for (int i = 0; i < 3; ++i) {
        d.spacing[i] = sp[i];
        d.offset[i] = off[i];
    }
    d.brick_log2 = static_cast<uint32_t>(brickLog2);
    const int rc = hipdvr_volume_create(ctx_, &d, vram->getData(), &dvrVolume_);
    if (rc != HIPDVR_OK) {
        dvrVolume_ = 0;
        return fail(std::string("volume upload failed: ") + hipdvr_last_error(ctx_));
    }
    uploadedVolume_ = vb;
    uploadedBrickLog2_ = brickLog2;
    ++volumeGeneration_;
    return true;
}

// Attach the gradient volume of the 2D TF path to the uploaded volume
// (checks of cpuraycaster.cpp:134-141 are made by the caller).
bool HIPRaycaster::ensureGradientVolume(const VolumeBase* gvb, const VolumeRAM* gram) {
    if (uploadedGradient_ == gvb && gradientGeneration_ == volumeGeneration_ && !gradientVolumePort_.hasChanged())
        return true;
    const std::string type = gram->getBaseType();
    const int32_t dtype = type == "uint8" ? HIPDVR_U8 : type == "uint16" ? HIPDVR_U16 : type == "float" ? HIPDVR_F32 : -1;
    if (dtype < 0) return fail("HIPRaycaster supports uint8, uint16 and float gradient volumes, got " + type);
    const int rc = hipdvr_volume_set_gradient(dvrVolume_, gram->getData(), dtype,
                                              static_cast<uint32_t>(gram->getNumChannels()));
    if (rc != HIPDVR_OK) {
        uploadedGradient_ = 0;
        return fail(std::string("gradient volume upload failed: ") + hipdvr_last_error(ctx_));
    }
    uploadedGradient_ = gvb;
    gradientGeneration_ = volumeGeneration_;
    return true;
}

// Upload the texture and domains of tFunc2DPrimitives_ when they changed.
// CPURaycaster reads its texels as float RGBA (cpuraycaster.cpp:334-338); a
// texture held as bytes is converted by /255.
bool HIPRaycaster::ensureTransFunc2D(const TransFunc2DPrimitives* tf) {
    const tgt::Texture* tex = tf ? tf->getTexture() : 0;
    if (!tex || !tex->getCpuTextureData() || tex->getWidth() < 2 || tex->getHeight() < 2)
        return fail("2D transfer function has no texture data");
    const size_t n = 4 * static_cast<size_t>(tex->getWidth()) * tex->getHeight();
    std::vector<float> texels(n);
    if (tex->getDataType() == GL_FLOAT) {
        std::memcpy(texels.data(), tex->getCpuTextureData(), n * sizeof(float));
    } else {
        const GLubyte* b = tex->getCpuTextureData();
        for (size_t i = 0; i < n; ++i) texels[i] = static_cast<float>(b[i]) / 255.f;
    }
    const tgt::vec2 d0 = tf->getDomain(0), d1 = tf->getDomain(1);
    const tgt::vec4 dom(d0.x, d0.y, d1.x, d1.y);
    if (dvrTF2D_ && texels == uploadedTF2D_ && uploadedDomain2D_ == dom) return true;
    if (dvrTF2D_) {
        hipdvr_tf_destroy(dvrTF2D_);
        dvrTF2D_ = 0;
    }
    const float domain[4] = {dom.x, dom.y, dom.z, dom.w};
    const int rc = hipdvr_tf2d_create(ctx_, texels.data(), static_cast<uint32_t>(tex->getWidth()),
                                      static_cast<uint32_t>(tex->getHeight()), domain, &dvrTF2D_);
    if (rc != HIPDVR_OK) {
        dvrTF2D_ = 0;
        uploadedTF2D_.clear();
        return fail(std::string("2D transfer function upload failed: ") + hipdvr_last_error(ctx_));
    }
    uploadedTF2D_.swap(texels);
    uploadedDomain2D_ = dom;
    return true;
}

// Upload the RGBA8 lookup table of tFunc1DKeys_ when it changed.
bool HIPRaycaster::ensureTransFunc(tgt::Texture* tfTexture, const tgt::vec2& domain) {
    const int width = tfTexture->getWidth();
    const GLubyte* data = tfTexture->getCpuTextureData();
    if (!data || width < 2) return fail("transfer function has no texture data");
    const size_t bytes = 4 * static_cast<size_t>(width);
    if (dvrTF_ && uploadedTF_.size() == bytes && std::memcmp(uploadedTF_.data(), data, bytes) == 0 &&
        uploadedDomain_ == domain)
        return true;
    if (dvrTF_) {
        hipdvr_tf_destroy(dvrTF_);
        dvrTF_ = 0;
        uploadedPreintKey_.clear();
    }
    const int rc = hipdvr_tf_create(ctx_, data, static_cast<uint32_t>(width), domain.x, domain.y, &dvrTF_);
    if (rc != HIPDVR_OK) {
        dvrTF_ = 0;
        uploadedTF_.clear();
        return fail(std::string("transfer function upload failed: ") + hipdvr_last_error(ctx_));
    }
    uploadedTF_.assign(data, data + bytes);
    uploadedDomain_ = domain;
    ++tfGeneration_;
    return true;
}

// The learned path serves the plain DVR march: 1D TF classification, linear
// filter, DVR compositing, no shading or Phong with central differences.
bool HIPRaycaster::learnedEligible(const hipdvr_params& p, bool tf2d, const std::string& cls) const {
    return !tf2d && cls == "transfer-function" && p.filter == HIPDVR_FILTER_LINEAR &&
           p.compositing == HIPDVR_COMPOSITE_DVR &&
           (p.shading == HIPDVR_SHADE_NONE ||
            (p.shading == HIPDVR_SHADE_PHONG && p.gradient == HIPDVR_GRADIENT_CENTRAL));
}

// One learned frame: (re)train the network when the volume or the number of
// training TFs changed (on the device, hipdvr_learned_train, with the current
// TF among the tables traced), prepare the stepping table when the TF, step or
// options changed, then render through device buffers kept between frames.
bool HIPRaycaster::renderLearned(const hipdvr_params& params, const float* entry, const float* exit, int w, int h,
                                 float* output) {
    if (!learned_ || learnedGeneration_ != volumeGeneration_) {
        releaseLearned();
        if (hipdvr_learned_create(ctx_, dvrVolume_, &learned_) != HIPDVR_OK) {
            learned_ = 0;
            return fail(std::string("learned state: ") + hipdvr_last_error(ctx_));
        }
        learnedGeneration_ = volumeGeneration_;
    }
    if (!mlp_ || learnedTFs_ != hipLearnedTFs_.get()) {
        if (mlp_) hipdvr_mlp_destroy(mlp_);
        mlp_ = 0;
        preparedKey_.clear();
        if (hipdvr_mlp_create(ctx_, 0x5EED, &mlp_) != HIPDVR_OK) {
            mlp_ = 0;
            return fail(std::string("stepping network: ") + hipdvr_last_error(ctx_));
        }
        hipdvr_learned_train_opts topts;
        std::memset(&topts, 0, sizeof(topts));
        topts.n_tfs = static_cast<uint32_t>(hipLearnedTFs_.get());
        topts.sampling_rate = samplingRate_.get();
        topts.extra_tfs = uploadedTF_.data();
        topts.n_extra = 1;
        topts.extra_width = static_cast<uint32_t>(uploadedTF_.size() / 4);
        hipdvr_learned_train_info info;
        if (hipdvr_learned_train(learned_, mlp_, &topts, &info) != HIPDVR_OK) {
            hipdvr_mlp_destroy(mlp_);
            mlp_ = 0;
            return fail(std::string("training the stepping network failed: ") + hipdvr_last_error(ctx_));
        }
        learnedTFs_ = hipLearnedTFs_.get();
    }
    hipdvr_learned_opts lopts;
    std::memset(&lopts, 0, sizeof(lopts));
    lopts.psnr_floor = hipPsnrFloor_.get();
    lopts.remarch_stride = hipRemarchStride_.get();
    lopts.skip_logit = 0.f;
    lopts.max_log2_step = 3;
    const std::string key = std::to_string(tfGeneration_) + ":" + std::to_string(params.step) + ":" +
                            std::to_string(lopts.psnr_floor) + ":" + std::to_string(lopts.remarch_stride);
    if (key != preparedKey_) {
        if (hipdvr_learned_prepare(learned_, dvrTF_, &params, mlp_, &lopts) != HIPDVR_OK) {
            preparedKey_.clear();
            return fail(std::string("learned stepping table: ") + hipdvr_last_error(ctx_));
        }
        preparedKey_ = key;
    }
    const size_t bytes = 4 * sizeof(float) * static_cast<size_t>(w) * h;
    if (bytes > dBytes_) {
        if (dEntry_) hipdvr_buffer_free(ctx_, dEntry_);
        if (dExit_) hipdvr_buffer_free(ctx_, dExit_);
        if (dOut_) hipdvr_buffer_free(ctx_, dOut_);
        dEntry_ = dExit_ = dOut_ = 0;
        dBytes_ = 0;
        if (hipdvr_buffer_alloc(ctx_, bytes, reinterpret_cast<void**>(&dEntry_)) != HIPDVR_OK ||
            hipdvr_buffer_alloc(ctx_, bytes, reinterpret_cast<void**>(&dExit_)) != HIPDVR_OK ||
            hipdvr_buffer_alloc(ctx_, bytes, reinterpret_cast<void**>(&dOut_)) != HIPDVR_OK)
            return fail(std::string("device buffers: ") + hipdvr_last_error(ctx_));
        dBytes_ = bytes;
    }
    if (hipdvr_memcpy_h2d(ctx_, dEntry_, entry, bytes) != HIPDVR_OK ||
        hipdvr_memcpy_h2d(ctx_, dExit_, exit, bytes) != HIPDVR_OK)
        return fail(std::string("entry/exit upload: ") + hipdvr_last_error(ctx_));
    if (hipdvr_render_learned_device(learned_, dvrTF_, dEntry_, dExit_, static_cast<uint32_t>(w),
                                     static_cast<uint32_t>(h), &params, &lopts, dOut_, &lastLearned_) != HIPDVR_OK)
        return fail(std::string("learned render: ") + hipdvr_last_error(ctx_));
    if (hipdvr_memcpy_d2h(ctx_, output, dOut_, bytes) != HIPDVR_OK)
        return fail(std::string("image download: ") + hipdvr_last_error(ctx_));
    return true;
}

void HIPRaycaster::fillParams(hipdvr_params& p, const VolumeBase* vb, float step) {
    std::memset(&p, 0, sizeof(p));
    p.step = step;
    p.filter = texFilterMode_.getValue() == GL_NEAREST ? HIPDVR_FILTER_NEAREST : HIPDVR_FILTER_LINEAR;
    p.classification = HIPDVR_CLASS_TF;
    // compositing keys of volumeraycaster.cpp:379-388
    const std::string comp = compositingMode1_.getKey();
    p.compositing = comp == "mip"    ? HIPDVR_COMPOSITE_MIP
                  : comp == "mop"    ? HIPDVR_COMPOSITE_MOP
                  : comp == "mida"   ? HIPDVR_COMPOSITE_MIDA
                  : comp == "iso"    ? HIPDVR_COMPOSITE_ISO
                  : comp == "fhp"    ? HIPDVR_COMPOSITE_FHP
                  : comp == "fhn"    ? HIPDVR_COMPOSITE_FHN
                                     : HIPDVR_COMPOSITE_DVR;
    p.gamma = gammaValue1_.get();
    p.iso_value = isoValue_.get();
    // shading keys of glsl.cpp:397-409
    const std::string sh = shadeMode_.getKey();
    p.shading = sh == "phong"                  ? HIPDVR_SHADE_PHONG
              : sh == "phong-diffuse"          ? HIPDVR_SHADE_PHONG_DIFFUSE
              : sh == "phong-specular"         ? HIPDVR_SHADE_PHONG_SPECULAR
              : sh == "phong-diffuse-ambient"  ? HIPDVR_SHADE_PHONG_DIFFUSE_AMBIENT
              : sh == "phong-diffuse-specular" ? HIPDVR_SHADE_PHONG_DIFFUSE_SPECULAR
              : sh == "toon"                   ? HIPDVR_SHADE_TOON
              : sh == "cook-torrance"          ? HIPDVR_SHADE_COOK_TORRANCE
              : sh == "oren-nayar"             ? HIPDVR_SHADE_OREN_NAYAR
              : sh == "ward"                   ? HIPDVR_SHADE_WARD
                                               : HIPDVR_SHADE_NONE;
    // light and camera to the volume's physical space (src/core/utils/glsl.cpp:61-72)
    const tgt::mat4 worldToPhysical = vb->getWorldToPhysicalMatrix();
    const tgt::vec3 lp = (worldToPhysical * lightPosition_.get()).xyz();
    const tgt::vec3 cp = (worldToPhysical * tgt::vec4(camera_.get().getPositionWithOffsets(), 1.f)).xyz();
    for (int i = 0; i < 3; ++i) {
        p.light_pos[i] = lp[i];
        p.light_ambient[i] = lightAmbient_.get()[i];
        p.light_diffuse[i] = lightDiffuse_.get()[i];
        p.light_specular[i] = lightSpecular_.get()[i];
        p.camera_pos[i] = cp[i];
    }
    p.shininess = materialShininess_.get();
    p.loop_cap = 0;
    const std::string k = hipKernel_.getKey();
    p.kernel = k == "reference" ? HIPDVR_KERNEL_REFERENCE
             : k == "bricked"   ? HIPDVR_KERNEL_BRICKED
             : k == "cell"      ? HIPDVR_KERNEL_CELL
             : k == "slab"      ? HIPDVR_KERNEL_SLAB
                                : HIPDVR_KERNEL_AUTO;
    p.batch = 0;
    p.exact_shading = 0;
    // gradient keys of volumeraycaster.cpp:126-131
    const std::string g = gradientMode_.getKey();
    p.gradient = g == "forward-differences" ? HIPDVR_GRADIENT_FORWARD
               : g == "none"                ? HIPDVR_GRADIENT_NONE
               : g == "filtered"            ? HIPDVR_GRADIENT_FILTERED
               : g == "sobel"               ? HIPDVR_GRADIENT_SOBEL
                                            : HIPDVR_GRADIENT_CENTRAL;
}

// Output exactly as CPURaycaster::process (cpuraycaster.cpp:157-159,197-207).
void HIPRaycaster::writeOutput(const std::vector<float>& rgba, int w, int h) {
    outport1_.activateTarget();
    outport1_.clearTarget();
    if (!rgba.empty()) {
        tgt::Texture outputTex(tgt::ivec3(w, h, 1), GL_RGBA, GL_RGBA, GL_FLOAT, tgt::Texture::LINEAR,
                               tgt::Texture::REPEAT, (GLubyte*)(const_cast<float*>(rgba.data())), false);
        outputTex.uploadTexture();
        glClear(GL_COLOR_BUFFER_BIT | GL_DEPTH_BUFFER_BIT);
        renderQuad();
    }
    outport1_.deactivateTarget();
}

float HIPRaycaster::renderStepSize(const VolumeRAM* vram) const {
    // the sampling rate over the largest dimension, as CPURaycaster computes
    // it (cpuraycaster.cpp:104), in interaction mode too: it does not call
    // VolumeRaycaster::getSamplingStepSize, whose interaction mode divides by
    // interactionQuality (volumeraycaster.cpp:312-322)
    const tgt::svec3 volDim = vram->getDimensions();
    return 1.f / (tgt::max(volDim) * samplingRate_.get());
}

void HIPRaycaster::process() {
    const VolumeBase* vb = volumeInport_.getData();
    const VolumeRAM* vram = vb ? vb->getRepresentation<VolumeRAM>() : 0;
    const tgt::ivec2 size = entryPort_.getSize();
    if (!vram) {
        fail("no input volume");
        writeOutput(std::vector<float>(), size.x, size.y);
        return;
    }
    tFunc1DKeys_.setVolume(vb);
    tFunc2DPrimitives_.setVolume(vb);

    const bool tf2d = tFuncType_.getValue() == TFT_2DPRIMITIVES;
    if (tFuncType_.getValue() != TFT_1DKEYS && !tf2d) {
        fail("HIPRaycaster supports 1D key and 2D primitive transfer functions");
        writeOutput(std::vector<float>(), size.x, size.y);
        return;
    }
    const std::string comp = compositingMode1_.getKey();
    const std::string cls = classificationMode_.getKey();
    if (comp != "dvr" && (cls != "transfer-function" || tf2d)) {
        fail("compositing mode '" + comp + "' needs 1D transfer-function classification");
        writeOutput(std::vector<float>(), size.x, size.y);
        return;
    }
    const VolumeBase* gvb = gradientVolumePort_.getData();
    const VolumeRAM* gram = gvb ? gvb->getRepresentation<VolumeRAM>() : 0;
    if (tf2d) {
        // the checks of cpuraycaster.cpp:134-146
        if (!gram || gram->getNumChannels() < 3) {
            fail("To use 2D tfs a RGB or RGBA gradient volume is needed");
            writeOutput(std::vector<float>(), size.x, size.y);
            return;
        }
        if (gram->getDimensions() != vram->getDimensions()) {
            fail("Gradient volume dimensions differ from intensity volume dimensions");
            writeOutput(std::vector<float>(), size.x, size.y);
            return;
        }
        if (cls.compare(0, 14, "pre-integrated") == 0) {
            fail("Pre-integration cannot be used with 2D tfs.");
            writeOutput(std::vector<float>(), size.x, size.y);
            return;
        }
    }

    if (!ensureContext() || !ensureVolume(vb, vram)) {
        writeOutput(std::vector<float>(), size.x, size.y);
        return;
    }

    const float step = renderStepSize(vram);

    hipdvr_params params;
    fillParams(params, vb, step);
    hipdvr_tf* renderTF = 0;

    if (tf2d) {
        if (!ensureGradientVolume(gvb, gram) || !ensureTransFunc2D(tFunc2DPrimitives_.get())) {
            writeOutput(std::vector<float>(), size.x, size.y);
            return;
        }
        params.classification = HIPDVR_CLASS_TF2D;
        renderTF = dvrTF2D_;
    } else {
        tgt::Texture* tfTexture = tFunc1DKeys_.get()->getTexture();
        tfTexture->downloadTexture();
        if (!ensureTransFunc(tfTexture, tFunc1DKeys_.get()->getDomain())) {
            writeOutput(std::vector<float>(), size.x, size.y);
            return;
        }
        renderTF = dvrTF_;
    }

    if (tf2d) {
        // classification set above
    } else if (cls == "pre-integrated-cpu" || cls == "pre-integrated-approximation") {
        const bool approx = cls == "pre-integrated-approximation";
        const std::string key = cls + ":" + std::to_string(preIntegrationTableSize_.getValue());
        if (uploadedPreintKey_ != key || uploadedPreintStep_ != step) {
            // the reference's own table (preintegrationtable.cpp:67-204), uploaded once per TF/step
            const PreIntegrationTable* table = tFunc1DKeys_.get()->getPreIntegrationTable(
                step, static_cast<size_t>(preIntegrationTableSize_.getValue()), approx);
            if (!table || !table->getTable()) {
                fail("no pre-integration table");
                writeOutput(std::vector<float>(), size.x, size.y);
                return;
            }
            const int rc = hipdvr_tf_set_preint(dvrTF_, reinterpret_cast<const float*>(table->getTable()),
                                                static_cast<uint32_t>(table->getDimension()));
            if (rc != HIPDVR_OK) {
                fail(std::string("pre-integration table upload failed: ") + hipdvr_last_error(ctx_));
                writeOutput(std::vector<float>(), size.x, size.y);
                return;
            }
            uploadedPreintKey_ = key;
            uploadedPreintStep_ = step;
        }
        params.classification = HIPDVR_CLASS_PREINT;
    } else if (cls == "pre-integrated-gpu") {
        // the exact table computed on the device from the TF keys
        // (preintegrationtable.cpp:67-204 run per entry; :206-254 is the GL path)
        const std::string key = cls + ":" + std::to_string(preIntegrationTableSize_.getValue());
        if (uploadedPreintKey_ != key || uploadedPreintStep_ != step) {
            const TransFunc1DKeys* tf = tFunc1DKeys_.get();
            std::vector<hipdvr_tf_key> keys((size_t)tf->getNumKeys());
            for (int i = 0; i < tf->getNumKeys(); ++i) {
                const TransFuncMappingKey* k = tf->getKey(i);
                keys[(size_t)i].intensity = k->getIntensity();
                keys[(size_t)i].split = k->isSplit() ? 1 : 0;
                const tgt::col4 cl = k->getColorL(), cr = k->getColorR();
                for (int c = 0; c < 4; ++c) {
                    keys[(size_t)i].color_l[c] = cl[(size_t)c];
                    keys[(size_t)i].color_r[c] = cr[(size_t)c];
                }
            }
            const int width = tf->getWidth();
            const int dim = preIntegrationTableSize_.getValue();
            const uint32_t n = (uint32_t)(dim ? dim : (width < 1024 ? width : 1024));
            const tgt::vec2 thr = tf->getThreshold();
            const int rc = hipdvr_tf_preint_device(dvrTF_, keys.data(), (int32_t)keys.size(), tf->getGammaValue(),
                                                   thr.x, thr.y, (int32_t)tf->getAlphaMode(), n, step, 0);
            if (rc != HIPDVR_OK) {
                fail(std::string("pre-integration table on the device failed: ") + hipdvr_last_error(ctx_));
                writeOutput(std::vector<float>(), size.x, size.y);
                return;
            }
            uploadedPreintKey_ = key;
            uploadedPreintStep_ = step;
        }
        params.classification = HIPDVR_CLASS_PREINT;
    }

    // entry/exit points, as CPURaycaster downloads them (cpuraycaster.cpp:166-169)
    tgt::vec4* entryBuffer = reinterpret_cast<tgt::vec4*>(
        entryPort_.getColorTexture()->downloadTextureToBuffer(GL_RGBA, GL_FLOAT));
    tgt::vec4* exitBuffer = reinterpret_cast<tgt::vec4*>(
        exitPort_.getColorTexture()->downloadTextureToBuffer(GL_RGBA, GL_FLOAT));

    std::vector<float> output(4 * static_cast<size_t>(size.x) * size.y, 0.f);
    lastFrameLearned_ = false;
    if (hipLearned_.get()) {
        if (learnedEligible(params, tf2d, cls)) {
            lastFrameLearned_ = renderLearned(params, reinterpret_cast<const float*>(entryBuffer),
                                              reinterpret_cast<const float*>(exitBuffer), size.x, size.y,
                                              output.data());
            if (!lastFrameLearned_) LWARNING("learned stepping failed; rendering plainly");
        } else {
            LWARNING("learned stepping serves DVR with a 1D TF, the linear filter and no shading or Phong "
                     "(central differences); rendering plainly");
        }
    }
    if (!lastFrameLearned_) {
        const int rc = hipdvr_render(ctx_, dvrVolume_, renderTF, reinterpret_cast<const float*>(entryBuffer),
                                     reinterpret_cast<const float*>(exitBuffer), size.x, size.y, &params,
                                     output.data(), &lastStats_);
        if (rc != HIPDVR_OK) {
            fail(std::string("hipdvr_render failed: ") + hipdvr_last_error(ctx_));
            output.clear();
        }
    }
    delete[] entryBuffer;
    delete[] exitBuffer;
    writeOutput(output, size.x, size.y);
}

}  // namespace voreen
