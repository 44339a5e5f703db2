// binary_protocol.cpp — PCLS frames (include/astraea/binary_protocol.hpp), the wire format of
// server/binary_protocol.cpp:75-219: fixed little-endian fields written byte by byte, so the
// encoding does not depend on the host's byte order.
#include "../../include/astraea/binary_protocol.hpp"

#include <zlib.h>

#include <cstring>
#include <limits>

namespace astraea::server {
namespace {

template <class T>
void putLE(uint8_t* dst, T v) {
    uint64_t bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    for (std::size_t b = 0; b < sizeof(T); ++b) dst[b] = (uint8_t)(bits >> (8 * b));
}

template <class T>
T getLE(const uint8_t* src) {
    uint64_t bits = 0;
    for (std::size_t b = 0; b < sizeof(T); ++b) bits |= (uint64_t)src[b] << (8 * b);
    T v;
    std::memcpy(&v, &bits, sizeof(T));
    return v;
}

}  // namespace

void encodeHeader40(const BinaryFrameHeader& h, uint8_t out[kPclsHeaderBytes]) {
    putLE(out + 0, h.magic);
    putLE(out + 4, h.version);
    putLE(out + 6, h.flags);
    putLE(out + 8, h.simTimeSec);
    putLE(out + 16, h.frameSeq);
    putLE(out + 20, h.dtSec);
    putLE(out + 24, h.particleCount);
    putLE(out + 28, h.tracerCount);
    putLE(out + 32, h.tracerGridRes);
    putLE(out + 36, h.payloadBytes);
}

bool decodeHeader40(const uint8_t* d, std::size_t len, BinaryFrameHeader& h) {
    if (!d || len < kPclsHeaderBytes) return false;
    h.magic = getLE<uint32_t>(d + 0);
    h.version = getLE<uint16_t>(d + 4);
    h.flags = getLE<uint16_t>(d + 6);
    h.simTimeSec = getLE<double>(d + 8);
    h.frameSeq = getLE<uint32_t>(d + 16);
    h.dtSec = getLE<float>(d + 20);
    h.particleCount = getLE<uint32_t>(d + 24);
    h.tracerCount = getLE<uint32_t>(d + 28);
    h.tracerGridRes = getLE<uint32_t>(d + 32);
    h.payloadBytes = getLE<uint32_t>(d + 36);
    return h.magic == kPclsMagic && h.version == kPclsVersion;
}

bool zlibCompress(const uint8_t* src, std::size_t n, std::vector<uint8_t>& out) {
    if (n > (std::size_t)std::numeric_limits<uLong>::max()) return false;
    static const uint8_t kEmpty = 0;
    uLongf cap = compressBound((uLong)n);
    out.resize(cap);
    if (compress2(out.data(), &cap, (src && n) ? src : &kEmpty, (uLong)n, Z_BEST_SPEED) != Z_OK) {
        out.clear();
        return false;
    }
    out.resize(cap);
    return true;
}

bool zlibDecompress(const uint8_t* src, std::size_t n, std::size_t expected, std::vector<uint8_t>& out) {
    if ((!src && n) || expected > (std::size_t)std::numeric_limits<uLongf>::max()) return false;
    out.resize(expected);
    if (expected == 0 && n == 0) return true;
    static uint8_t scratch = 0;
    uLongf got = (uLongf)expected;
    const int rc = uncompress(expected ? out.data() : &scratch, &got, n ? src : &scratch, (uLong)n);
    if (rc != Z_OK || got != expected) {
        out.clear();
        return false;
    }
    return true;
}

std::vector<uint8_t> buildBinaryFrame(const TelemetrySnapshot& snap, const std::vector<uint8_t>& particles,
                                      const std::vector<uint8_t>& tracers, const std::vector<uint8_t>* speed,
                                      bool compress) {
    const uint32_t np = (uint32_t)(particles.size() / kParticleRecordBytes);
    const uint32_t nt = (uint32_t)(tracers.size() / kTracerRecordBytes);
    const bool withSpeed = speed && np > 0 && speed->size() == np;
    std::vector<uint8_t> payload;
    payload.reserve(particles.size() + tracers.size() + (withSpeed ? speed->size() : 0));
    payload.insert(payload.end(), particles.begin(), particles.end());
    payload.insert(payload.end(), tracers.begin(), tracers.end());
    if (withSpeed) payload.insert(payload.end(), speed->begin(), speed->end());

    BinaryFrameHeader h;
    h.flags = (uint16_t)((nt ? kFlagHasTracers : 0) | (snap.livingBodyPresent ? kFlagHasLivingCom : 0) |
                         (withSpeed ? kFlagHasSpeedArray : 0));
    h.simTimeSec = snap.simTimeSec;
    h.frameSeq = (uint32_t)(snap.frameSeq & 0xffffffffull);
    h.dtSec = snap.dtSec;
    h.particleCount = np;
    h.tracerCount = nt;
    h.tracerGridRes = nt ? snap.tracerGridRes : 0u;
    h.payloadBytes = (uint32_t)payload.size();

    std::vector<uint8_t> packed;
    const bool zipped = compress && zlibCompress(payload.data(), payload.size(), packed);
    const std::vector<uint8_t>& body = zipped ? packed : payload;
    if (zipped) h.flags |= kFlagZlib;
    std::vector<uint8_t> frame(kPclsHeaderBytes + body.size());
    encodeHeader40(h, frame.data());
    if (!body.empty()) std::memcpy(frame.data() + kPclsHeaderBytes, body.data(), body.size());
    return frame;
}

}  // namespace astraea::server
