/*
 * astraea_hip.h — C ABI of the MI355X (gfx950) per-step physics library.
 *
 * This is the drop-in boundary under the reference's `astraea::Simulation`
 * (engine/simulation.hpp:12-47).  It replaces the reference's CUDA launch
 * layer `astraea::cuda::*` (engine/sph_gravity.cuh:61-92) and the device-buffer
 * management of `Simulation::Impl` (engine/simulation.cpp:345-536).  Every
 * signature uses plain pointers, sizes and status codes so that C, C++ and
 * Python ctypes can bind it.
 *
 * Conventions
 *  - Every function returns 0 on success or a negative ASTRAEA_E_* code.  The
 *    message of the last failure on the calling thread is returned by
 *    astraea_last_error().  The reference throws std::runtime_error /
 *    std::invalid_argument instead (engine/sph_gravity.cuh:12-22,
 *    engine/sph_gravity.cu:2001-2003, engine/simulation.cpp:1296-1298,
 *    1712-1714); the C++ Simulation layer above this ABI rethrows.
 *  - A context owns one HIP stream and all device memory.  Calls on one
 *    context must be serialised by the caller (the reference's contract:
 *    one host thread drives one default stream, simulation.cpp:1711-1753).
 *  - Particles are grouped in segments (independent scenarios).  A single
 *    scenario is one segment.  Host arrays passed to upload/download are
 *    the segments' particles concatenated in segment order, without padding.
 *  - Units are the reference's PSU system (engine/units.hpp:13-27):
 *    length 1e6 m, mass 1e24 kg, time 1e3 s.
 */
#ifndef ASTRAEA_HIP_H
#define ASTRAEA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASTRAEA_ABI_VERSION 1

/* Status codes */
#define ASTRAEA_OK 0
#define ASTRAEA_E_INVALID (-1)   /* bad argument (reference: std::invalid_argument) */
#define ASTRAEA_E_HIP (-2)       /* HIP runtime failure (reference: ASTRAEA_CUDA_CHECK) */
#define ASTRAEA_E_STATE (-3)     /* call order (reference: step before loadScenario) */
#define ASTRAEA_E_CAPACITY (-4)  /* capacity < count (sph_gravity.cu:2001-2003) */
#define ASTRAEA_E_OVERFLOW (-5)  /* a device-side capacity (neighbour table, stack) overflowed */

/* Physics knobs.  Defaults (astraea_default_params) reproduce the reference's
 * compile-time constants; the benchmark configurations change them. */
typedef struct astraea_params {
    float theta;              /* BH opening angle; reference literal 0.5 (simulation.cpp:533,1729) */
    float soft_factor;        /* eps_i = soft_factor * h0_i; reference 0.8 (sph_gravity.cu:842,873) */
    float cfl_factor;         /* CFL coefficient; reference 0.25 (sph_gravity.cu:1510) */
    float gamma_hhe;          /* H/He ideal-gas gamma; reference 1.4 (eos.cu:93) */
    float cooling_min_temp_k; /* radiative cooling only where T > this; reference 0 (sph_gravity.cu:1222-1235) */
    uint32_t strength_mask;   /* bit m: material m carries deviatoric stress and damage;
                                 reference: every material with mu > 0 (sph_gravity.cu:42-73) = 0x17 */
    uint32_t precision;       /* ASTRAEA_PREC_FP32 (reference layout) or ASTRAEA_PREC_FP64_ACCUM */
    uint32_t group_size;      /* particles sharing one gravity opening decision: 32 = the
                                 reference's warp groups (sph_gravity.cu:932-933), its interaction
                                 lists (the default); 0 or 64 = one wavefront, half the
                                 traversals, more interactions */
    uint32_t reserved[4];
} astraea_params;

#define ASTRAEA_PREC_FP32 0u       /* FP32 positions and velocities, as the reference */
#define ASTRAEA_PREC_FP64_ACCUM 1u /* FP64 master positions/velocities for kick and drift */

/* Traversal group of group_size 0: the gravity opening decision is shared by one wavefront. */
#define ASTRAEA_GROUP_SIZE 64
/* Each segment starts at a multiple of this many particle slots. */
#define ASTRAEA_SEGMENT_ALIGN 256
/* Depth of the per-wavefront traversal stack; the reference uses 192 (sph_gravity.cu:22). */
#define ASTRAEA_STACK_DEPTH 192

/* Per-particle fields for upload/download. */
enum astraea_field {
    /* persistent state, float32 */
    ASTRAEA_F_POS_X = 0, ASTRAEA_F_POS_Y, ASTRAEA_F_POS_Z,
    ASTRAEA_F_VEL_X, ASTRAEA_F_VEL_Y, ASTRAEA_F_VEL_Z,
    ASTRAEA_F_MASS, ASTRAEA_F_H, ASTRAEA_F_H0, ASTRAEA_F_RHO, ASTRAEA_F_U,
    ASTRAEA_F_SXX, ASTRAEA_F_SYY, ASTRAEA_F_SZZ, ASTRAEA_F_SXY, ASTRAEA_F_SXZ, ASTRAEA_F_SYZ,
    ASTRAEA_F_DAMAGE, ASTRAEA_F_DAMAGE_CBRT, ASTRAEA_F_EPS_FAIL,
    /* persistent state, uint8 / uint32 */
    ASTRAEA_F_MATERIAL, ASTRAEA_F_BODY, ASTRAEA_F_TRACER, ASTRAEA_F_ORIGINAL_ID,
    /* persistent state, float64 (ASTRAEA_PREC_FP64_ACCUM only) */
    ASTRAEA_F_POS64_X, ASTRAEA_F_POS64_Y, ASTRAEA_F_POS64_Z,
    ASTRAEA_F_VEL64_X, ASTRAEA_F_VEL64_Y, ASTRAEA_F_VEL64_Z,
    /* derived, float32 (download only) */
    ASTRAEA_F_ACC_X, ASTRAEA_F_ACC_Y, ASTRAEA_F_ACC_Z,
    ASTRAEA_F_ACC_GRAV_X, ASTRAEA_F_ACC_GRAV_Y, ASTRAEA_F_ACC_GRAV_Z, ASTRAEA_F_PHI,
    ASTRAEA_F_DRHO_DT, ASTRAEA_F_DU_DT, ASTRAEA_F_PRESSURE, ASTRAEA_F_SOUND_SPEED, ASTRAEA_F_TEMP_K,
    ASTRAEA_F_SHEPARD, ASTRAEA_F_SURF_FACTOR, ASTRAEA_F_DIV_V,
    ASTRAEA_F_CURL_X, ASTRAEA_F_CURL_Y, ASTRAEA_F_CURL_Z,
    ASTRAEA_F_DS_XX, ASTRAEA_F_DS_YY, ASTRAEA_F_DS_ZZ, ASTRAEA_F_DS_XY, ASTRAEA_F_DS_XZ, ASTRAEA_F_DS_YZ,
    ASTRAEA_F_DDAMAGE_CBRT_DT, ASTRAEA_F_RAD_POWER_W,
    /* derived, uint8 */
    ASTRAEA_F_PHASE,
    /* derived, uint32: neighbour count of the last pass-1 (self included, as the reference) */
    ASTRAEA_F_NGB_COUNT,
    ASTRAEA_F_COUNT_
};

/* Element size in bytes of a field (4, 1, 8), or 0 for an unknown id. */
size_t astraea_field_size(int field);

const char* astraea_last_error(void);
int astraea_abi_version(void);
void astraea_default_params(astraea_params* out);

typedef struct astraea_ctx astraea_ctx;

/* Allocate a context on `device` for up to `capacity` particles in up to
 * `max_segments` segments.  Replaces allocateParticleBuffers/allocateRadixBvh
 * (sph_gravity.cuh:83-86) and Simulation::Impl::ensureCapacity (simulation.cpp:428-454).
 * Slots are addressed with 28 bits in the neighbour walk, so capacity is at most
 * ASTRAEA_MAX_CAPACITY (far beyond what 288 GB holds with the neighbour table). */
#define ASTRAEA_MAX_CAPACITY (1u << 27)
int astraea_create(astraea_ctx** out, int device, uint32_t capacity, uint32_t max_segments);
int astraea_destroy(astraea_ctx* ctx);
/* The context's HIP stream (a hipStream_t), for callers that time or order work on it. */
void* astraea_stream(astraea_ctx* ctx);

int astraea_set_params(astraea_ctx* ctx, const astraea_params* params);
int astraea_get_params(astraea_ctx* ctx, astraea_params* out);

/* Declare the segment layout: `counts[s]` particles in segment s.  Clears all
 * state.  Replaces the count bookkeeping of uploadHostStateToDevice
 * (simulation.cpp:469-475). */
int astraea_set_layout(astraea_ctx* ctx, uint32_t num_segments, const uint32_t* counts);
int astraea_num_particles(astraea_ctx* ctx, uint32_t* out_total);

/* Copy one field between host (concatenated segments) and device, in the context's
 * current particle order: the input order until the first sort, the sorted order after
 * it (ORIGINAL_ID holds each particle's input index).  Upload is valid for persistent
 * fields; download for every field.  Fields that are not uploaded start as in
 * uploadHostStateToDevice (simulation.cpp:469-520): zeros, except EPS_FAIL = 1,
 * ORIGINAL_ID = index in the concatenation.  Uploading positions, masses or smoothing
 * lengths invalidates the tree; other fields leave it valid. */
int astraea_upload_field(astraea_ctx* ctx, int field, const void* host_src);
int astraea_download_field(astraea_ctx* ctx, int field, void* host_dst);

/* ---- stage entry points (reference launch layer, sph_gravity.cuh:61-92) ---- */

/* launchClearStats (sph_gravity.cu:1986-1994): zero the derived accumulators. */
int astraea_clear_stats(astraea_ctx* ctx);
/* launchSortAndBuildBvh (sph_gravity.cu:1996-2061): per-segment AABB, 63-bit Morton
 * keys, stable radix sort, permutation of the particle state, Karras radix
 * tree, bottom-up node moments, and the neighbour table of the new order. */
int astraea_sort_and_build_bvh(astraea_ctx* ctx);
/* launchSphPass1Neighbors (sph_gravity.cu:2063-2074), plus the EOS of the same
 * particle (launchEvaluateEos) fused into its epilogue. */
int astraea_sph_pass1(astraea_ctx* ctx);
/* launchEvaluateEos (sph_gravity.cu:2076-2083). */
int astraea_evaluate_eos(astraea_ctx* ctx);
/* launchGravityBarnesHut (sph_gravity.cu:2085-2097); theta from the params. */
int astraea_gravity(astraea_ctx* ctx);
/* launchSphPass2Forces (sph_gravity.cu:2099-2111). */
int astraea_sph_pass2(astraea_ctx* ctx);
/* launchReduceDtPsu (sph_gravity.cu:2136-2161) for every segment.  dt stays on
 * the device; `out_dt` (num_segments floats, may be NULL) forces a copy-out. */
int astraea_reduce_dt(astraea_ctx* ctx, float max_dt_psu, float* out_dt);
/* launchLeapfrogKickHalf (sph_gravity.cu:2113-2120) with half of the device dt. */
int astraea_kick_half(astraea_ctx* ctx, int clamp_thermal);
/* launchLeapfrogDrift (sph_gravity.cu:2122-2134) with the device dt. */
int astraea_drift(astraea_ctx* ctx);

/* Simulation::Impl::evaluateInitialForcesOnDevice (simulation.cpp:525-536):
 * clear -> sort/BVH -> pass 1 (+EOS) -> gravity -> pass 2, then the dt reduction
 * input state is ready. */
int astraea_evaluate_initial_forces(astraea_ctx* ctx);

/* `steps` full KDK steps, the kernel sequence of Simulation::step
 * (simulation.cpp:1721-1732): dt -> kick -> drift -> sort/BVH -> pass 1 -> EOS ->
 * gravity -> pass 2 -> kick -> EOS.  Asynchronous on the context stream; with
 * `use_graph` the sequence is captured once in a hipGraph and replayed. */
int astraea_step(astraea_ctx* ctx, float max_dt_psu, uint32_t steps, int use_graph);

/* Block until the context stream is idle.  A step whose neighbour table came out incomplete
 * (see astraea_set_neighbour_stride) holds the rest of the queued steps on the device; this call
 * (and every other entry point except astraea_step, which only queues) then widens the table,
 * finishes that step and runs the held steps, so the result is the run a wide enough table gives.
 * ASTRAEA_E_OVERFLOW only when the table cannot hold what a particle needs (more than 65536
 * entries, or a wavefront union above 65536) or the neighbour walk's stack overflowed: the state
 * then stays at the step after its sort, and the forces must be evaluated again. */
int astraea_synchronize(astraea_ctx* ctx);

/* Per-segment scalars after the last dt reduction: dt (PSU) of the last step. */
int astraea_get_dt(astraea_ctx* ctx, float* out_dt);

/* Per-segment FP64 reductions on the device (Simulation telemetry,
 * simulation.cpp:985-1087).  `out` receives ASTRAEA_ENERGY_STRIDE doubles per
 * segment, in the order of the ASTRAEA_EN_* indices (PSU units). */
#define ASTRAEA_EN_MASS 0
#define ASTRAEA_EN_KINETIC 1
#define ASTRAEA_EN_POTENTIAL 2   /* 1/2 sum m phi */
#define ASTRAEA_EN_INTERNAL 3
#define ASTRAEA_EN_MOM_X 4
#define ASTRAEA_EN_MOM_Y 5
#define ASTRAEA_EN_MOM_Z 6
#define ASTRAEA_EN_ANG_X 7
#define ASTRAEA_EN_ANG_Y 8
#define ASTRAEA_EN_ANG_Z 9
#define ASTRAEA_EN_RADIATED 10   /* radiated energy integral (PSU energy), trapezoid as simulation.cpp:1741-1744 */
#define ASTRAEA_EN_RAD_POWER_W 11 /* sum of radPowerW (W) */
#define ASTRAEA_EN_MAX_TEMP 12
#define ASTRAEA_EN_MAX_PRESSURE 13
#define ASTRAEA_EN_COM_X 14
#define ASTRAEA_EN_COM_Y 15
#define ASTRAEA_EN_COM_Z 16
#define ASTRAEA_EN_SIM_TIME 17   /* accumulated sum of dt (PSU time) */
#define ASTRAEA_ENERGY_STRIDE 18
int astraea_energy(astraea_ctx* ctx, double* out);

/* Largest-remnant classification of the telemetry, per segment, as FP64 reductions on
 * the device (replaces the host particle loops of Simulation::computeTelemetryFromDevice,
 * simulation.cpp:1000-1013, 1036-1039, 1094-1112 and 1115-1227): the sums of body 0 and
 * of the other bodies; the remnant seed, the particle of lowest finite phi (lowest
 * ORIGINAL_ID on ties); up to five iterations of the particles bound to the remnant
 * guess, each stopping the segment when none is bound; and the final split into bound
 * planet / disk / ejecta, with the bound mass of bodies other than 0, the largest damage
 * and the Y spin of body 0 within 1.5 target radii of its centre.  `target_radius_psu`
 * holds one radius per segment.  `out` receives ASTRAEA_CLASS_STRIDE doubles per segment
 * in the order of the ASTRAEA_CL_* indices (PSU units; centres and velocities already
 * divided by their mass, zero for a body without mass).  Synchronous. */
#define ASTRAEA_CL_BODY0_MASS 0
#define ASTRAEA_CL_BODY0_COM 1       /* x, y, z */
#define ASTRAEA_CL_BODY0_VEL 4       /* vx, vy, vz */
#define ASTRAEA_CL_BODY1_MASS 7      /* every body other than 0 */
#define ASTRAEA_CL_BODY1_COM 8
#define ASTRAEA_CL_BODY1_VEL 11
#define ASTRAEA_CL_SEED_X 14         /* x, y, z, vx, vy, vz of the seed particle */
#define ASTRAEA_CL_REM_MASS 20       /* remnant after the iterations */
#define ASTRAEA_CL_REM_POS 21
#define ASTRAEA_CL_REM_VEL 24
#define ASTRAEA_CL_REM_RADIUS 27
#define ASTRAEA_CL_BOUND_PLANET 28   /* masses of the final split */
#define ASTRAEA_CL_DISK 29
#define ASTRAEA_CL_EJECTA 30
#define ASTRAEA_CL_IMPACTOR_BOUND 31 /* bound mass of bodies other than 0 */
#define ASTRAEA_CL_SPIN_L 32         /* sum m (dz dvx - dx dvz) about body 0's centre */
#define ASTRAEA_CL_SPIN_I 33         /* sum m (dx^2 + dz^2) */
#define ASTRAEA_CL_MAX_DAMAGE 34
#define ASTRAEA_CL_ITERATIONS 35     /* remnant iterations that found bound mass */
#define ASTRAEA_CLASS_STRIDE 36
int astraea_classify(astraea_ctx* ctx, const double* target_radius_psu, double* out);

/* Render frame particle records (packRenderRecordsKernel, sph_gravity.cu:1671-1743, as
 * Simulation::extractRenderFrame uses it, simulation.cpp:1805-1867): 20 bytes per particle,
 * little-endian: x, y, z km about com_psu of the particle's segment (3 doubles per segment,
 * PSU) as f32; u16 log temperature; u16 log pressure (0 at P <= 1e-3 GPa); u8 material;
 * u8 body << 4 | phase; u8 damage * 255; u8 log radius of the volume-equivalent sphere.
 * `speed` (NULL or one byte per particle): u8 log speed.  Records are in the context's
 * current particle order (ORIGINAL_ID gives the input index), segments concatenated. */
int astraea_pack_render(astraea_ctx* ctx, const double* com_psu, uint8_t* records, uint8_t* speed);

/* Allocation-free variant of astraea_pack_render.  It has no counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops such as a render stream), where astraea_pack_render allocates its staging buffers on the
 * first call.  `records_dev` (20 bytes per particle, 4-byte aligned) and `speed_dev` (NULL or one
 * byte per particle) are the caller's device buffers; they receive exactly what astraea_pack_render
 * returns, in the same order.  The per-segment centres and output offsets go through a fixed table
 * in the library's device memory (ASTRAEA_NOALLOC_MAX_SEGMENTS segments, 4 KiB per GPU, held until
 * the process ends); a layout with more segments is rejected with ASTRAEA_E_INVALID.  No device
 * memory is allocated or freed.  A queued step held by an incomplete neighbour table is not
 * finished here (that widens the table): ASTRAEA_E_STATE asks for astraea_synchronize first.
 * Every call on one GPU uses that GPU's table, which keeps the last call's contents: calls that
 * overlap in time on one GPU (from several threads or contexts) race.  Returns when the records
 * are written. */
#define ASTRAEA_NOALLOC_MAX_SEGMENTS 256
int astraea_pack_render_noalloc(astraea_ctx* ctx, const double* com_psu, uint8_t* records_dev, uint8_t* speed_dev);

/* Crust tracers (the tracer fields of the reference's particle buffers, types.hpp:150-156,
 * filled by planet_builder.cpp:737-906).  Tracer t sits on the particle whose ORIGINAL_ID is
 * particle_id[t]; rest = 4 floats per tracer (rest radius PSU, rest unit normal x, y, z about
 * the living body's centre); ocean = 1 for an ocean tracer.  Sets the TRACER flag of those
 * particles (which drives the living-body test of the cooling) and clears it elsewhere.  At most
 * 65535 tracers.  set_layout clears the tracers. */
int astraea_set_tracers(astraea_ctx* ctx, uint32_t count, const uint32_t* particle_id, const float* rest,
                        const uint8_t* ocean);
/* Crust tracer records (extractTracersKernel, sph_gravity.cu:1586-1668): 40 bytes per tracer,
 * at its index: u16 index, u16 flags (1 ocean, 2 disrupted), f32 x, y, z km about sys_com_psu,
 * radial displacement km and normal change (rad) about living_com_psu, T (K), damage,
 * overpressure (kPa), tsunami height (m).  Centres: 3 doubles per segment. */
int astraea_extract_tracers(astraea_ctx* ctx, const double* living_com_psu, const double* sys_com_psu,
                            uint8_t* records);

/* Damped relaxation of the loaded bodies, one body per segment, all segments at once
 * (replaces Simulation::relaxLoadedSingleBody, simulation.cpp:585-781, which relaxes one body
 * with a host round trip per step).  Evaluates the initial forces, then `steps` times: a KDK
 * step with dt_max = 0.02 PSU and the thermal clamp in both kicks; the segment's centre of mass
 * and mean velocity removed, velocities damped by 0.85, u set to u_cold(rho) + c_v T_anchor,
 * stress, damage and their rates cleared; forces of the corrected state.  With steps > 0, then
 * four undamped steps.  `out[s]` = v_rms / v_esc of segment s with v_esc = sqrt(2 G M / r),
 * r = max(nominal_radius_psu[s], largest |x|) (positions about the origin).  With steps > 0,
 * u is re-anchored and stress and damage cleared once more at the end.
 * `anchor_temp_k` holds one temperature (K) per particle, indexed by ORIGINAL_ID (an id past
 * the end reads 288 K); NULL takes each particle's EOS temperature after the initial forces.
 * The forces are left stale: evaluate the initial forces again before stepping. */
int astraea_relax(astraea_ctx* ctx, uint32_t steps, const float* anchor_temp_k,
                  const double* nominal_radius_psu, double* out_vrms_over_vesc);

/* Counters: `out[0]` = sum over active particles of the particle-node and
 * particle-particle interactions of the gravity evaluations since the last call,
 * `out[1]` = traversal-stack overflows of those (forced acceptances,
 * sph_gravity.cu:933), `out[2]` = neighbour candidates of the last table (all
 * particles), `out[3]` = table entries per particle, `out[4]` = sum of pass-1
 * neighbour counts (self excluded), `out[5]` = sum of pass-2 pairs, `out[6]` =
 * largest candidate count of one particle in the last table, `out[7]` = steps held by an
 * incomplete neighbour table and run again after it was widened (since the context was created),
 * `out[8]` = wavefronts of the last table built in the spill pool, `out[9]` = pool bytes they use.
 * Up to `n` values. */
int astraea_counters(astraea_ctx* ctx, uint64_t* out, int n);

/* Width of the neighbour table's regular region, in entries per particle, rounded up to a
 * multiple of 32 (no counterpart in the reference, which walks the tree in every pass).  Each
 * wavefront of 64 particles keeps the union of its particles' candidates (up to
 * min(32 * stride, 65536) ids, 4 bytes each, in four banks) and each particle 16-bit
 * positions into them, stride / 4 per bank, so the regular region takes 4 * stride * capacity
 * bytes.  A wavefront with a particle that has more candidates in a bank than stride / 4, or
 * more candidates than its union holds, is built again in a spill pool exactly as wide as it
 * needs, so memory follows the candidates the particles have.  The pool grows when a table does
 * not fit it: at a load, or for queued steps at the next synchronising call, which holds the
 * steps until then and runs them (astraea_synchronize). */
#define ASTRAEA_DEFAULT_NEIGHBOUR_STRIDE 320
int astraea_set_neighbour_stride(astraea_ctx* ctx, uint32_t stride);

/* Tree of segment `seg` from the last sort, in the reference's local encoding
 * (sph_gravity.cuh:26-59): internal nodes 0..n-2, a negative child c is the leaf
 * ~c.  Any pointer may be NULL.  `keys`/`perm` have n entries (sorted Morton
 * keys and indicesSorted, local to the segment); `left`, `right`,
 * `parent_internal` have n-1; `parent_leaf` has n; `node_cm` (com.xyz, mass),
 * `node_min` (aabb min.xyz, extent), `node_max` (aabb max.xyz, max support)
 * have 4*(n-1) floats. */
int astraea_download_tree(astraea_ctx* ctx, uint32_t seg,
                          uint64_t* keys, uint32_t* perm,
                          int32_t* left, int32_t* right,
                          int32_t* parent_internal, int32_t* parent_leaf,
                          float* node_cm, float* node_min, float* node_max);

/* Stage timing: when enabled, astraea_step and the stage calls record HIP events
 * around each stage; astraea_stage_times returns the accumulated milliseconds
 * per ASTRAEA_STAGE_* and the number of samples, and resets them. */
#define ASTRAEA_STAGE_DT 0
#define ASTRAEA_STAGE_KICK1 1
#define ASTRAEA_STAGE_DRIFT 2
#define ASTRAEA_STAGE_MORTON 3
#define ASTRAEA_STAGE_SORT 4
#define ASTRAEA_STAGE_PERMUTE_TREE 5
#define ASTRAEA_STAGE_MOMENTS 6
#define ASTRAEA_STAGE_NEIGHBOURS 7
#define ASTRAEA_STAGE_PASS1 8
#define ASTRAEA_STAGE_GRAVITY 9
#define ASTRAEA_STAGE_PASS2 10
#define ASTRAEA_STAGE_KICK2 11
#define ASTRAEA_STAGE_COUNT_ 12
int astraea_enable_stage_timing(astraea_ctx* ctx, int enable);
int astraea_stage_times(astraea_ctx* ctx, double* out_ms, uint64_t* out_samples);

/* Launch variants (no counterpart in the reference, whose block sizes, stack depth and sort are
 * compile-time constants or a library's, sph_gravity.cu:19-22, 2028-2046): the union records each
 * neighbour pass stages in LDS per round, the radix sort's digit width, and the wavefronts per
 * block and per-wave traversal stack entries of the gravity walk and of the neighbour build.
 * Every variant gives bit-identical results (a particle's pairs and records run in the same order;
 * every sort is the stable sort by key) as long as no stack fills; they differ in LDS use, and so
 * in how many wavefronts share a SIMD or a CU's cache, and in passes.  The (waves, stack) pairs
 * are (2, 1024), (4, 1024), (4, 512) and (8, 512). */
typedef struct astraea_tuning {
    uint32_t pass1_round;   /* 128 or 256 (default 256) */
    uint32_t pass2_round;   /* 128 or 160 (default 160) */
    uint32_t drift_round;   /* 128, 160, 192 or 256 (default 160) */
    uint32_t sort_bits;     /* 8 (8 passes) or 9 (7 passes) (default 8) */
    uint32_t gravity_waves; /* wavefronts per block of the gravity walk (default 4) */
    uint32_t gravity_stack; /* its per-wave stack entries (default 1024) */
    uint32_t build_waves;   /* wavefronts per block of the neighbour build (default 4) */
    uint32_t build_stack;   /* its per-wave stack entries (default 1024) */
} astraea_tuning;
int astraea_get_tuning(astraea_ctx* ctx, astraea_tuning* out);
int astraea_set_tuning(astraea_ctx* ctx, const astraea_tuning* tuning);
/* The autotuner: advances the loaded state by 4 + 8 * steps_per_variant eager steps (one untimed
 * step with each variant first, then the timed ones), timing every variant of each stage with the
 * stage events (in palindromic orders, so that the state's slow change of cost cancels), and keeps
 * the fastest for each.  The simulation does not depend
 * on the variant, so these are ordinary steps.  Then it launches the gravity walk and the
 * neighbour build of the resulting state again in each (waves, stack) variant, once untimed and
 * twice timed in palindromic order, timed with events; those launches rewrite what the step computed, bit for
 * bit, and a variant whose stack filled (a forced acceptance, an incomplete table) is not chosen
 * (the table is then built again with the chosen variant).  `chosen` and `ms` (mean milliseconds:
 * pass 1 at 128 and 256, pass 2 at 128 and 160, drift at 128, 160, 192 and 256, the sort at 8 and 9
 * bits, then gravity and the build at (2, 1024), (4, 1024), (4, 512) and (8, 512); 18 values, a
 * variant whose stack filled reads -1) may be NULL.  On an error the previous tuning is
 * restored. */
int astraea_autotune(astraea_ctx* ctx, float max_dt_psu, uint32_t steps_per_variant, astraea_tuning* chosen,
                     double* ms);

/* ---- host-side initial conditions (no device work) ---- */

/* A differentiated body on a seed-jittered hexagonal-close-packed lattice:
 * uniform density per layer, u from the closed-form hydrostatic pressure of
 * piecewise-constant density (see astraea_body_spec).  Writes `n` particles
 * into the caller's arrays (any may be NULL).  Positions are centred on the
 * body's centre of mass; velocities are zero. */
#define ASTRAEA_MAX_LAYERS 4
typedef struct astraea_body_spec {
    double radius_m;                        /* outer radius */
    uint32_t num_layers;                    /* 1..ASTRAEA_MAX_LAYERS, innermost first */
    uint8_t layer_material[ASTRAEA_MAX_LAYERS];
    double layer_mass_frac[ASTRAEA_MAX_LAYERS];
    double layer_density_kgm3[ASTRAEA_MAX_LAYERS]; /* <= 0: the material's rho0 */
    double h_over_spacing;                  /* smoothing length / lattice spacing (1.3) */
    double jitter;                          /* jitter amplitude / spacing (0.1) */
    double surface_temp_k;                  /* temperature where P = 0 (reference T_ref); <= 0: material tRefK */
    uint32_t strength_flaws;                /* 1: draw Weibull flaws eps_fail */
    double weibull_m;                       /* Weibull modulus (9) */
    double weibull_k;                       /* Weibull k (m^-3), 4e29 */
    uint64_t seed;
} astraea_body_spec;

typedef struct astraea_body_out {
    float *x, *y, *z, *m, *h, *rho, *u, *eps_fail;
    uint8_t* mat;
} astraea_body_out;

void astraea_default_body_spec(astraea_body_spec* out);
/* Returns the body's total mass (kg) in *out_mass_kg and lattice spacing (m). */
int astraea_build_body(const astraea_body_spec* spec, uint32_t n, astraea_body_out* out,
                       double* out_mass_kg, double* out_spacing_m);

/* Layered planet in hydrostatic equilibrium with this library's EOS (replaces
 * PlanetBuilder::solveHydrostatic1D and buildPlanet, planet_builder.cpp:20-265, 398-735):
 * RK4 outward from a central pressure, density from the EOS at the pressure and at a
 * temperature linear in pressure between core and surface, material switched at each
 * layer's cumulative mass fraction, surface where P reaches surface_pressure_pa; the
 * central pressure is shot on until the mass is reached.  Particles on Fibonacci-sphere
 * shells about one spacing (m/rho)^(1/3) thick, each shell turned by a random rotation,
 * h0 = h_factor (m/rho)^(1/3), u = u_cold(rho) + c_v T, failure strains
 * eps_min * clamp((-ln U)^(1/9), 1, 3).  Positions about the centre, PSU units. */
typedef struct astraea_planet_spec {
    double mass_kg;
    uint32_t num_layers;                    /* 1..ASTRAEA_MAX_LAYERS, innermost first; H/He only outermost */
    uint8_t layer_material[ASTRAEA_MAX_LAYERS];
    double layer_mass_frac[ASTRAEA_MAX_LAYERS];
    double surface_temp_k;
    double core_temp_k;
    double surface_pressure_pa;             /* 1e5 for gas giants, 0 otherwise (planet_builder.cpp:407) */
    double gamma_hhe;                       /* H/He ideal-gas gamma of the profile (1.4) */
    double h_factor;                        /* 1.2 (planet_builder.cpp:681) */
    uint64_t seed;
} astraea_planet_spec;

typedef struct astraea_planet_info {
    double mass_kg;                         /* mass of the solved profile */
    double radius_m;
    double central_pressure_pa;
    uint32_t profile_samples;
} astraea_planet_info;

void astraea_default_planet_spec(astraea_planet_spec* out);
/* `n` particles into `out` (any pointer may be NULL); `temp_k` (NULL or n floats) receives each
 * particle's profile temperature, the relaxation anchor of the reference. */
int astraea_build_planet(const astraea_planet_spec* spec, uint32_t n, astraea_body_out* out, float* temp_k,
                         astraea_planet_info* info);
/* The solved radial profile: up to `cap` samples of radius (m), density (kg/m^3), pressure (Pa),
 * enclosed mass (kg), temperature (K) and material, centre first (any pointer may be NULL). */
int astraea_planet_profile(const astraea_planet_spec* spec, uint32_t cap, double* r, double* rho, double* p,
                           double* m, double* t, uint8_t* mat, astraea_planet_info* info);

/* Crust tracers of a LivingTerra body (planet_builder.cpp:737-906): of the n particles
 * (positions about the body's centre, PSU), the outermost shell (within 1.8 h0 of the largest
 * radius, plus any water ice, at least 3 x 6 G^2 candidates) is binned on the cube-sphere grid
 * of 6 G^2 cells; each cell takes the candidate nearest its centre, then empty cells the nearest
 * unused candidate of the pool; indices are compacted when fewer than 6 G^2 were found.  Per
 * particle: tracer_index (-1 for none), rest_radius, rest_normal (3 floats), ocean (water ice).
 * `out_count` receives the number of tracers (0 when n < 6 G^2 or 6 G^2 > 65535). */
int astraea_select_tracers(uint32_t n, const float* x, const float* y, const float* z, const float* h0,
                           const uint8_t* material, uint32_t grid_res, int32_t* tracer_index,
                           float* rest_radius, float* rest_normal, uint8_t* ocean, uint32_t* out_count);
/* Cube-sphere mapping of the tracer grid (planet_builder.cpp:329-395): face 0..5 = +X, -X, +Y,
 * -Y, +Z, -Z; (u, v) in [0, 1]^2. */
void astraea_cube_face_uv_to_normal(uint32_t face, float u, float v, float* nx, float* ny, float* nz);
void astraea_normal_to_cube_face_uv(float nx, float ny, float nz, uint32_t* face, float* u, float* v);

/* Outer radius (m) of the body `spec` describes (its radius_m is ignored) when its total mass
 * is `mass_kg`: uniform layers at their densities; for an H/He outermost layer, the surface of
 * the n = 1 polytrope envelope over the core layers that astraea_build_body builds (no
 * counterpart in the reference, whose planet_builder.cpp:398-471 integrates the profile to the
 * surface pressure). */
int astraea_body_radius_for_mass(const astraea_body_spec* spec, double mass_kg, double* radius_m);

/* Two-body placement on a Keplerian approach orbit in the XZ plane, the
 * formulas of Simulation::loadScenario (simulation.cpp:1389-1436), followed by
 * the zero-COM / zero-momentum shift of mergeTwoBodies (simulation.cpp:919-944).
 * Adjusts x and vx/vz in place for n_t target then n_i impactor particles. */
int astraea_place_two_bodies(uint32_t n_t, uint32_t n_i,
                             float* x, float* y, float* z,
                             float* vx, float* vy, float* vz, const float* m,
                             double r_target_psu, double r_impactor_psu,
                             double v_ratio_esc, double impact_param_b, double sep_radii,
                             double* out_v_esc_psu);

/* splitmix64: the seed sequence used for per-scenario seeds. */
uint64_t astraea_splitmix64(uint64_t x);

#ifdef __cplusplus
}
#endif

#endif /* ASTRAEA_HIP_H */
