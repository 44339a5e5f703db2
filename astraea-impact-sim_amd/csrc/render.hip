// render.hip — render frames: the 20-byte particle records, the speed bytes and the 40-byte crust
// tracer records of Simulation::extractRenderFrame (engine/sph_gravity.cu:1554-1743,
// engine/simulation.cpp:1805-1867), packed from the resident state in one pass.
//
// Particle record (little-endian): x, y, z km about the system centre (f32 x3), log temperature
// u16 = round(65535 clamp(ln(max(T,10)/10) / ln 20000, 0, 1)), log pressure u16 = 0 at P <= 1e-3
// GPa else round(1 + 65534 clamp(ln(P / 1e-3 GPa) / ln 1e7, 0, 1)), material u8, body << 4 | phase
// u8, damage u8 = round(255 clamp(D, 0, 1)), log radius u8 of the particle's volume-equivalent
// sphere = round(255 clamp(ln(max(r km, 1)) / ln 2000, 0, 1)).  Speed byte = round(255 clamp(
// ln(max(|v| km/s, 0.05) / 0.05) / ln 2000, 0, 1)).
//
// Records are staged in LDS so that a block writes its 256 x 20 bytes as contiguous dwords.
// Output order is the context's particle order, segments concatenated without padding.
#include "eos.hpp"
#include "launch.hpp"

namespace astraea_hip {
namespace {

__device__ __forceinline__ float finiteOr0(float v) { return isfinite(v) ? v : 0.0f; }
__device__ __forceinline__ uint32_t quant(float norm, float scale) {
    return (uint32_t)lrintf(scale * fminf(1.0f, fmaxf(0.0f, norm)));
}

__global__ __launch_bounds__(kBlock) void kPackRender(DeviceCtx c, const float* com, const uint32_t* outBase,
                                                      uint32_t* records, uint8_t* speed) {
    __shared__ uint32_t lds[kBlock * 5];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t block0 = blockIdx.x * kBlock;
    const uint32_t sg = c.seg.slotSeg[block0];
    const uint32_t off = c.seg.off[sg], n = c.seg.n[sg];
    if (block0 - off >= n) return;  // a padding block: uniform
    const uint32_t count = min((uint32_t)kBlock, n - (block0 - off));
    const uint32_t o0 = outBase[sg] + (block0 - off);
    if (threadIdx.x < count) {
        constexpr float kInvLn20000 = 1.0f / 9.9034875525361280f;
        constexpr float kInvLn1e7 = 1.0f / 16.118095650958320f;
        constexpr float kInvLn2000 = 1.0f / 7.6009024595420823f;
        const float4 p = c.cur.pos[i];
        const float4 v = c.cur.vel[i];
        const float4 eo = c.d.eos[i];
        const uint32_t meta = c.cur.meta[i];
        const float cx = com[3 * sg], cy = com[3 * sg + 1], cz = com[3 * sg + 2];
        const float x = finiteOr0((p.x - cx) * 1000.0f);
        const float y = finiteOr0((p.y - cy) * 1000.0f);
        const float z = finiteOr0((p.z - cz) * 1000.0f);
        const float T = fmaxf(10.0f, finiteOr0(eo.z));
        const uint32_t tq = quant(logf(T * 0.1f) * kInvLn20000, 65535.0f);
        const float pGPa = finiteOr0(eo.x) * 1000.0f;
        const uint32_t pq = pGPa > 1.0e-3f ? 1u + quant(logf(pGPa * 1000.0f) * kInvLn1e7, 65534.0f) : 0u;
        const uint32_t mat = meta & 0xffu;
        const uint32_t pb = (((meta >> 8) & 0x0fu) << 4) | (c.d.phase[i] & 0x0fu);
        const uint32_t dq = quant(finiteOr0(c.cur.hu[i].w), 255.0f);
        const float vol = fmaxf(p.w, 0.0f) / fmaxf(v.w, 1.0e-9f);
        const float rKm = cbrtf(3.0f * vol / (4.0f * kPiF)) * 1000.0f;
        const uint32_t rq = quant(logf(fmaxf(rKm, 1.0f)) * kInvLn2000, 255.0f);
        uint32_t* r = lds + 5 * threadIdx.x;
        r[0] = __float_as_uint(x);
        r[1] = __float_as_uint(y);
        r[2] = __float_as_uint(z);
        r[3] = tq | (pq << 16);
        r[4] = mat | (pb << 8) | (dq << 16) | (rq << 24);
        if (speed) {
            const float s = finiteOr0(sqrtf(v.x * v.x + v.y * v.y + v.z * v.z));
            speed[o0 + threadIdx.x] = (uint8_t)quant(logf(fmaxf(s, 0.05f) * 20.0f) * kInvLn2000, 255.0f);
        }
    }
    __syncthreads();
    uint32_t* dst = records + 5ull * o0;
    for (uint32_t k = threadIdx.x; k < 5 * count; k += kBlock) dst[k] = lds[k];
}

// Crust tracer records (sph_gravity.cu:1586-1668): the particles flagged as tracer anchors look
// their tracer up by original id and write record t.  Rest radius and normal, ocean flag and the
// original id of each tracer live in tables indexed by tracer, never permuted.
__global__ __launch_bounds__(kBlock) void kExtractTracers(DeviceCtx c, const int32_t* tracerOfOid, uint32_t nOid,
                                                          const float4* rest, const uint8_t* ocean, uint32_t nTracers,
                                                          const float* livingCom, const float* sysCom,
                                                          uint32_t* out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = c.seg.slotSeg[i];
    if (i - c.seg.off[sg] >= c.seg.n[sg]) return;
    const uint32_t meta = c.cur.meta[i];
    if (((meta >> 16) & 0xffu) == 0u) return;
    const uint32_t oid = c.cur.oid[i];
    if (oid >= nOid) return;
    const int32_t t = tracerOfOid[oid];
    if (t < 0 || (uint32_t)t >= nTracers) return;
    const float4 p = c.cur.pos[i];
    const float4 hu = c.cur.hu[i];
    const float4 rs = rest[t];  // (rest radius, normal x, y, z)
    const float dx = p.x - livingCom[3 * sg], dy = p.y - livingCom[3 * sg + 1], dz = p.z - livingCom[3 * sg + 2];
    const float r = sqrtf(dx * dx + dy * dy + dz * dz);
    const float disp = (r - rs.x) * 1000.0f;
    float nd = 0.0f;
    if (r > 1.0e-9f) {
        const float inv = 1.0f / r;
        nd = acosf(fminf(1.0f, fmaxf(-1.0f, dx * inv * rs.y + dy * inv * rs.z + dz * inv * rs.w)));
    }
    const float dam = fminf(1.0f, fmaxf(0.0f, finiteOr0(hu.w)));
    const float T = fmaxf(3.0f, finiteOr0(c.d.eos[i].z));
    const uint8_t ph = c.d.phase[i];
    const bool isOcean = ocean[t] != 0u;
    const bool disrupted = dam >= 0.85f || ph == kVapor || ph == kPlasma || fabsf(r - rs.x) > 0.5f * fmaxf(rs.x, 1.0e-3f);
    const float dvShock = fmaxf(0.0f, -finiteOr0(c.d.p1a[i].z)) * fmaxf(hu.x, 0.0f) * 1000.0f;
    const float overKpa = fmaxf(0.0f, 1.225f * 340.0f * dvShock * 1.0e-3f);
    const float tsunami = isOcean ? fmaxf(0.0f, finiteOr0(disp) * 1000.0f) : 0.0f;
    uint32_t* o = out + 10ull * t;
    o[0] = (uint32_t)t | ((isOcean ? 1u : 0u) | (disrupted ? 2u : 0u)) << 16;
    o[1] = __float_as_uint(finiteOr0((p.x - sysCom[3 * sg]) * 1000.0f));
    o[2] = __float_as_uint(finiteOr0((p.y - sysCom[3 * sg + 1]) * 1000.0f));
    o[3] = __float_as_uint(finiteOr0((p.z - sysCom[3 * sg + 2]) * 1000.0f));
    o[4] = __float_as_uint(finiteOr0(disp));
    o[5] = __float_as_uint(finiteOr0(nd));
    o[6] = __float_as_uint(T);
    o[7] = __float_as_uint(dam);
    o[8] = __float_as_uint(finiteOr0(overKpa));
    o[9] = __float_as_uint(finiteOr0(tsunami));
}

// Tracer anchor flags from the table: meta byte 2 = 1 where tracerOfOid[oid] >= 0, else 0.
__global__ __launch_bounds__(kBlock) void kMarkTracers(DeviceCtx c, const int32_t* tracerOfOid, uint32_t nOid) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = c.seg.slotSeg[i];
    if (i - c.seg.off[sg] >= c.seg.n[sg]) return;
    const uint32_t oid = c.cur.oid[i];
    const uint32_t flag = (oid < nOid && tracerOfOid[oid] >= 0) ? 1u : 0u;
    const uint32_t meta = c.cur.meta[i];
    c.cur.meta[i] = (meta & ~(0xffu << 16)) | (flag << 16);
}

}  // namespace

void launchPackRender(const DeviceCtx& c, hipStream_t s, const float* com, const uint32_t* outBase, uint32_t* records,
                      uint8_t* speed) {
    kPackRender<<<c.blocks, kBlock, 0, s>>>(c, com, outBase, records, speed);
}
void launchExtractTracers(const DeviceCtx& c, hipStream_t s, const int32_t* tracerOfOid, uint32_t nOid,
                          const float4* rest, const uint8_t* ocean, uint32_t nTracers, const float* livingCom,
                          const float* sysCom, uint32_t* out) {
    kExtractTracers<<<c.blocks, kBlock, 0, s>>>(c, tracerOfOid, nOid, rest, ocean, nTracers, livingCom, sysCom, out);
}
void launchMarkTracers(const DeviceCtx& c, hipStream_t s, const int32_t* tracerOfOid, uint32_t nOid) {
    kMarkTracers<<<c.blocks, kBlock, 0, s>>>(c, tracerOfOid, nOid);
}

}  // namespace astraea_hip
