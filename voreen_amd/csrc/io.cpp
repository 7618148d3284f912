// io.cpp -- .vvd/.raw volume loading and whole-volume min/max
// (include/hipdvr_io.h).
//
// The .vvd reader follows the XML that VvdVolumeReader consumes
// (modules/core/io/vvdformat.cpp:70-90 for RawData: the "Paths" item list or
// the older "filename" attribute relative to the document, "format", "x",
// "y", "z"; meta items Offset / Spacing (Vec3MetaData) and RealWorldMapping
// (scale, offset), realworldmappingmetadata.cpp:53-62).  Only the first
// volume of a file is read; uint8 and uint16 voxels are supported.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hipdvr_io.h"
#include "ctx.h"
#include "errors.h"
#include "volume.h"

using namespace hipdvr;

namespace {

// Text of the document with <!-- --> comments removed.
std::string read_xml(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) fail(HIPDVR_ERR_INVALID, "cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    std::string s = ss.str(), out;
    out.reserve(s.size());
    size_t i = 0;
    while (i < s.size()) {
        const size_t c = s.find("<!--", i);
        if (c == std::string::npos) {
            out.append(s, i, std::string::npos);
            break;
        }
        out.append(s, i, c - i);
        const size_t e = s.find("-->", c + 4);
        if (e == std::string::npos) fail(HIPDVR_ERR_INVALID, "unterminated comment in " + path);
        i = e + 3;
    }
    return out;
}

// The start tag "<name ...>" at or after pos (name followed by a space, '/' or '>').
size_t find_tag(const std::string& s, const std::string& name, size_t pos, size_t end = std::string::npos) {
    const std::string open = "<" + name;
    while (true) {
        const size_t p = s.find(open, pos);
        if (p == std::string::npos || (end != std::string::npos && p >= end)) return std::string::npos;
        const char c = p + open.size() < s.size() ? s[p + open.size()] : '\0';
        if (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '/' || c == '>') return p;
        pos = p + 1;
    }
}

// Attribute value of the start tag at p ("" and false when absent).
bool attr(const std::string& s, size_t p, const std::string& name, std::string& value) {
    const size_t gt = s.find('>', p);
    if (gt == std::string::npos) return false;
    size_t q = p;
    while (true) {
        q = s.find(name + "=", q);
        if (q == std::string::npos || q >= gt) return false;
        const char before = s[q - 1];
        if (before == ' ' || before == '\t' || before == '\n' || before == '\r') break;
        q += name.size();
    }
    q += name.size() + 1;
    const char quote = s[q];
    if (quote != '"' && quote != '\'') return false;
    const size_t e = s.find(quote, q + 1);
    if (e == std::string::npos || e > gt) return false;
    value = s.substr(q + 1, e - q - 1);
    return true;
}

float attr_float(const std::string& s, size_t p, const std::string& name, const std::string& what) {
    std::string v;
    if (!attr(s, p, name, v)) fail(HIPDVR_ERR_INVALID, "missing attribute " + name + " of " + what);
    char* end = nullptr;
    const float f = std::strtof(v.c_str(), &end);
    if (end == v.c_str()) fail(HIPDVR_ERR_INVALID, "bad number '" + v + "' for " + what);
    return f;
}

// <MetaItem name="..."> ... <value .../> inside [lo, hi); npos when absent.
size_t meta_value(const std::string& s, const std::string& name, size_t lo, size_t hi) {
    size_t p = lo;
    while ((p = find_tag(s, "MetaItem", p, hi)) != std::string::npos) {
        std::string n;
        if (attr(s, p, "name", n) && n == name) return find_tag(s, "value", p, hi);
        ++p;
    }
    return std::string::npos;
}

std::string dir_of(const std::string& path) {
    const size_t k = path.find_last_of('/');
    return k == std::string::npos ? std::string(".") : path.substr(0, k);
}

void parse_vvd(const std::string& path, hipdvr_volume_desc& d, std::string& raw) {
    const std::string s = read_xml(path);
    const size_t vol = find_tag(s, "Volume", 0);
    if (vol == std::string::npos) fail(HIPDVR_ERR_INVALID, "no <Volume> in " + path);
    size_t vol_end = s.find("</Volume>", vol);
    if (vol_end == std::string::npos) vol_end = s.size();
    const size_t rd = find_tag(s, "RawData", vol, vol_end);
    if (rd == std::string::npos) fail(HIPDVR_ERR_INVALID, "no <RawData> in " + path);
    std::memset(&d, 0, sizeof d);
    std::string fmt;
    if (!attr(s, rd, "format", fmt)) fail(HIPDVR_ERR_INVALID, "RawData without format in " + path);
    if (fmt == "uint8") d.dtype = HIPDVR_U8;
    else if (fmt == "uint16") d.dtype = HIPDVR_U16;
    else fail(HIPDVR_ERR_UNSUPPORTED, "voxel format '" + fmt + "' is not supported (uint8, uint16)");
    const char* axes[3] = {"x", "y", "z"};
    for (int i = 0; i < 3; ++i) {
        const float v = attr_float(s, rd, axes[i], "RawData");
        if (!(v >= 1.f) || v != (float)(uint32_t)v) fail(HIPDVR_ERR_INVALID, "bad RawData dimension");
        d.dims[i] = (uint32_t)v;
    }
    std::string file;
    if (!attr(s, rd, "filename", file)) {
        const size_t rd_end = s.find("</RawData>", rd);
        const size_t item = find_tag(s, "item", rd, rd_end);
        if (item == std::string::npos || !attr(s, item, "value", file))
            fail(HIPDVR_ERR_INVALID, "RawData names no file in " + path);
    }
    raw = (!file.empty() && file[0] == '/') ? file : dir_of(path) + "/" + file;
    d.rwm_scale = 1.f;
    d.rwm_offset = 0.f;
    for (int i = 0; i < 3; ++i) {
        d.spacing[i] = 1.f;
        d.offset[i] = 0.f;
    }
    size_t p;
    if ((p = meta_value(s, "Spacing", vol, vol_end)) != std::string::npos)
        for (int i = 0; i < 3; ++i) d.spacing[i] = attr_float(s, p, axes[i], "Spacing");
    if ((p = meta_value(s, "Offset", vol, vol_end)) != std::string::npos)
        for (int i = 0; i < 3; ++i) d.offset[i] = attr_float(s, p, axes[i], "Offset");
    if ((p = meta_value(s, "RealWorldMapping", vol, vol_end)) != std::string::npos) {
        d.rwm_scale = attr_float(s, p, "scale", "RealWorldMapping");
        d.rwm_offset = attr_float(s, p, "offset", "RealWorldMapping");
    }
}

thread_local std::string g_parse_error;

}  // namespace

extern "C" {

int hipdvr_vvd_parse(const char* vvd_path, hipdvr_volume_desc* desc, char* raw_path, uint32_t raw_path_cap) {
    if (!vvd_path || !desc) return HIPDVR_ERR_INVALID;
    g_parse_error.clear();
    return ctx_guarded(nullptr, [&] {
        try {
            std::string raw;
            parse_vvd(vvd_path, *desc, raw);
            if (raw_path) {
                if (raw.size() + 1 > raw_path_cap) fail(HIPDVR_ERR_INVALID, "raw_path buffer too small");
                std::memcpy(raw_path, raw.c_str(), raw.size() + 1);
            }
        } catch (const Error& e) {
            g_parse_error = e.msg;
            throw;
        }
    });
}

const char* hipdvr_vvd_last_error(void) { return g_parse_error.c_str(); }

int hipdvr_volume_load_raw(hipdvr_ctx* ctx, const char* raw_path, const hipdvr_volume_desc* desc,
                           uint64_t byte_offset, hipdvr_volume** out) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    return ctx_guarded(ctx, [&] {
        if (!raw_path || !desc || !out) fail(HIPDVR_ERR_INVALID, "NULL argument");
        if (desc->dtype != HIPDVR_U8 && desc->dtype != HIPDVR_U16)
            fail(HIPDVR_ERR_INVALID, "dtype must be HIPDVR_U8 or HIPDVR_U16");
        const uint64_t n = (uint64_t)desc->dims[0] * desc->dims[1] * desc->dims[2];
        const uint64_t bytes = n * (desc->dtype == HIPDVR_U8 ? 1 : 2);
        std::ifstream in(raw_path, std::ios::binary);
        if (!in) fail(HIPDVR_ERR_INVALID, std::string("cannot open ") + raw_path);
        in.seekg(0, std::ios::end);
        const uint64_t size = (uint64_t)in.tellg();
        if (size < byte_offset + bytes)
            fail(HIPDVR_ERR_INVALID, std::string("raw file too small for the volume: ") + raw_path);
        // stage through pinned memory in slices onto the device, then build
        void* d_lin = nullptr;
        hip_check(hipMalloc(&d_lin, bytes ? bytes : 1), "hipMalloc(raw volume)");
        void* pinned = nullptr;
        const uint64_t slice = 64ull << 20;
        try {
            hip_check(hipHostMalloc(&pinned, slice, hipHostMallocDefault), "hipHostMalloc");
            in.seekg((std::streamoff)byte_offset);
            hipStream_t st = ctx_stream(ctx);
            for (uint64_t off = 0; off < bytes; off += slice) {
                const uint64_t k = bytes - off < slice ? bytes - off : slice;
                hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
                in.read((char*)pinned, (std::streamsize)k);
                if ((uint64_t)in.gcount() != k) fail(HIPDVR_ERR_INVALID, "short read of the raw file");
                hip_check(hipMemcpyAsync((char*)d_lin + off, pinned, k, hipMemcpyHostToDevice, st), "upload");
            }
            hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
            *out = volume_create(ctx, st, desc, d_lin, /*on_device=*/true);
        } catch (...) {
            (void)hipFree(d_lin);
            if (pinned) (void)hipHostFree(pinned);
            throw;
        }
        (void)hipFree(d_lin);
        (void)hipHostFree(pinned);
    });
}

int hipdvr_volume_load_vvd(hipdvr_ctx* ctx, const char* vvd_path, uint32_t brick_log2, hipdvr_volume** out,
                           hipdvr_volume_desc* desc_out) {
    if (!ctx) return HIPDVR_ERR_INVALID;
    hipdvr_volume_desc d;
    std::string raw;
    int rc = ctx_guarded(ctx, [&] {
        if (!vvd_path || !out) fail(HIPDVR_ERR_INVALID, "NULL argument");
        parse_vvd(vvd_path, d, raw);
        d.brick_log2 = brick_log2;
    });
    if (rc != HIPDVR_OK) return rc;
    rc = hipdvr_volume_load_raw(ctx, raw.c_str(), &d, 0, out);
    if (rc == HIPDVR_OK && desc_out) *desc_out = d;
    return rc;
}

int hipdvr_volume_minmax(hipdvr_volume* vol, float out[4]) {
    if (!vol) return HIPDVR_ERR_INVALID;
    return ctx_guarded(vol->ctx, [&] {
        if (!out) fail(HIPDVR_ERR_INVALID, "NULL argument");
        ctx_set_device(vol->ctx);
        uint32_t mm[2];
        volume_raw_minmax(vol, ctx_stream(vol->ctx), mm);
        const float d = vol->desc.dtype == HIPDVR_U8 ? 255.0f : 65535.0f;
        // VolumeAtomic::getVoxelNormalized (volumeatomic.h:531-536), then the
        // real-world mapping (realworldmappingmetadata.cpp:65-67)
        out[0] = (float)mm[0] / d;
        out[1] = (float)mm[1] / d;
        out[2] = (out[0] * vol->desc.rwm_scale) + vol->desc.rwm_offset;
        out[3] = (out[1] * vol->desc.rwm_scale) + vol->desc.rwm_offset;
    });
}

}  // extern "C"
