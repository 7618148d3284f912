// tgt vector/matrix subset (ext/tgt/vector.h, ext/tgt/matrix.h).
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>

namespace tgt {

template <typename T> struct Vector2 {
    T x, y;
    Vector2() : x(0), y(0) {}
    Vector2(T a, T b) : x(a), y(b) {}
    T& operator[](size_t i) { return i == 0 ? x : y; }
    const T& operator[](size_t i) const { return i == 0 ? x : y; }
    bool operator==(const Vector2& o) const { return x == o.x && y == o.y; }
    bool operator!=(const Vector2& o) const { return !(*this == o); }
};

template <typename T> struct Vector3 {
    T x, y, z;
    Vector3() : x(0), y(0), z(0) {}
    explicit Vector3(T v) : x(v), y(v), z(v) {}
    Vector3(T a, T b, T c) : x(a), y(b), z(c) {}
    template <typename U> explicit Vector3(const Vector3<U>& o) : x(T(o.x)), y(T(o.y)), z(T(o.z)) {}
    T& operator[](size_t i) { return i == 0 ? x : (i == 1 ? y : z); }
    const T& operator[](size_t i) const { return i == 0 ? x : (i == 1 ? y : z); }
    bool operator==(const Vector3& o) const { return x == o.x && y == o.y && z == o.z; }
    bool operator!=(const Vector3& o) const { return !(*this == o); }
};

template <typename T> struct Vector4 {
    T x, y, z, w;
    Vector4() : x(0), y(0), z(0), w(0) {}
    explicit Vector4(T v) : x(v), y(v), z(v), w(v) {}
    Vector4(T a, T b, T c, T d) : x(a), y(b), z(c), w(d) {}
    Vector4(const Vector3<T>& v, T d) : x(v.x), y(v.y), z(v.z), w(d) {}
    T& operator[](size_t i) { return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w)); }
    const T& operator[](size_t i) const { return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w)); }
    Vector3<T> xyz() const { return Vector3<T>(x, y, z); }
    bool operator==(const Vector4& o) const { return x == o.x && y == o.y && z == o.z && w == o.w; }
    bool operator!=(const Vector4& o) const { return !(*this == o); }
};

typedef Vector2<float> vec2;
typedef Vector3<float> vec3;
typedef Vector4<float> vec4;
typedef Vector2<int> ivec2;
typedef Vector3<int> ivec3;
typedef Vector3<size_t> svec3;
typedef Vector4<uint8_t> col4;

template <typename T> inline T max(const Vector3<T>& v) {
    T m = v.x;
    if (v.y > m) m = v.y;
    if (v.z > m) m = v.z;
    return m;
}

struct mat4 {
    float m[4][4];
    mat4() {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) m[i][j] = (i == j) ? 1.f : 0.f;
    }
    static mat4 identity() { return mat4(); }
};

inline vec4 operator*(const mat4& a, const vec4& v) {
    vec4 r;
    for (int i = 0; i < 4; ++i) {
        float s = 0.f;
        for (int j = 0; j < 4; ++j) s += a.m[i][j] * v[j];
        r[i] = s;
    }
    return r;
}

}  // namespace tgt
