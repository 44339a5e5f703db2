// binary_protocol.hpp — the PCLS render-frame wire format of the reference's server
// (server/binary_protocol.hpp:1-61), same declarations so the server compiles unchanged.
//
// Frame = 40-byte little-endian header, then the payload: particle records (20 B each), tracer
// records (40 B each) and, when flagged, one speed byte per particle; the payload is zlib
// compressed when kFlagZlib is set (payloadBytes is always the uncompressed size).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "types.hpp"

namespace astraea::server {

constexpr uint32_t kPclsMagic = 0x50434C53u;  // "SLCP" little-endian
constexpr uint16_t kPclsVersion = 1u;
constexpr std::size_t kPclsHeaderBytes = 40u;
constexpr std::size_t kParticleRecordBytes = 20u;
constexpr std::size_t kTracerRecordBytes = 40u;

enum FrameFlags : uint16_t {
    kFlagZlib = 1u << 0,
    kFlagHasTracers = 1u << 1,
    kFlagHasLivingCom = 1u << 2,
    kFlagHasSpeedArray = 1u << 3
};

struct BinaryFrameHeader {
    uint32_t magic = kPclsMagic;
    uint16_t version = kPclsVersion;
    uint16_t flags = 0;
    double simTimeSec = 0.0;
    uint32_t frameSeq = 0;
    float dtSec = 0.0f;
    uint32_t particleCount = 0;
    uint32_t tracerCount = 0;
    uint32_t tracerGridRes = 0;
    uint32_t payloadBytes = 0;
};

/// Encode the 40-byte header: magic u32 @0, version u16 @4, flags u16 @6, simTimeSec f64 @8,
/// frameSeq u32 @16, dtSec f32 @20, particleCount u32 @24, tracerCount u32 @28,
/// tracerGridRes u32 @32, payloadBytes u32 @36.
void encodeHeader40(const BinaryFrameHeader& hdr, uint8_t out[kPclsHeaderBytes]);

/// Decode a header; false when shorter than 40 bytes or magic / version differ.
bool decodeHeader40(const uint8_t* data, std::size_t len, BinaryFrameHeader& outHdr);

/// zlib at Z_BEST_SPEED.
bool zlibCompress(const uint8_t* src, std::size_t srcLen, std::vector<uint8_t>& outCompressed);

/// Inflate exactly expectedSize bytes; false on any zlib error or size mismatch.
bool zlibDecompress(const uint8_t* src, std::size_t srcLen, std::size_t expectedSize,
                    std::vector<uint8_t>& outDecompressed);

/// Header + payload of one render frame from Simulation::extractRenderFrame's output.
std::vector<uint8_t> buildBinaryFrame(const TelemetrySnapshot& snapshot, const std::vector<uint8_t>& particleBytes,
                                      const std::vector<uint8_t>& tracerBytes,
                                      const std::vector<uint8_t>* speedBytes, bool compressZlib);

}  // namespace astraea::server
