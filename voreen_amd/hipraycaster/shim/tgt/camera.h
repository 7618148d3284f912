// tgt::Camera subset (ext/tgt/camera.h).
#pragma once
#include "tgt/vector.h"

namespace tgt {

class Camera {
public:
    Camera(const vec3& position = vec3(0.f, 0.f, 0.f), const vec3& focus = vec3(0.f, 0.f, -1.f),
           const vec3& up = vec3(0.f, 1.f, 0.f))
        : position_(position), focus_(focus), up_(up) {}
    vec3 getPosition() const { return position_; }
    vec3 getPositionWithOffsets() const { return position_; }
    vec3 getFocus() const { return focus_; }
    vec3 getUpVector() const { return up_; }
    void setPosition(const vec3& p) { position_ = p; }

private:
    vec3 position_, focus_, up_;
};

}  // namespace tgt
