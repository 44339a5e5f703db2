/*
 * oracle.h — FP64 CPU restatement of the reference's per-step physics path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the C++
 * Simulation layer, bench.py's GPU leg) may link, load or call this code.  It
 * is the checker for the GPU path (tests/, __graft_entry__.smoke()) and the
 * timed CPU baseline of bench.py.
 *
 * Arithmetic is FP64.  Every discrete decision the GPU makes on FP32 data is
 * replicated in FP32 with the same operation sequence, so that integer results
 * (Morton keys, sort permutation, tree topology, gravity opening decisions,
 * pass-1 neighbour counts) are bit-exact and the FP64 sums run over the same
 * interaction sets:
 *   - scene AABB, Morton keys            engine/sph_gravity.cu:140-171, 273-361
 *   - stable (key, index) radix sort     engine/sph_gravity.cu:2028-2046 (CUB SortPairs, bits [0,63))
 *   - Karras radix tree                  engine/sph_gravity.cu:173-187, 478-528
 *   - bottom-up node moments (FP32)      engine/sph_gravity.cu:531-620
 *   - grouped Barnes-Hut opening         engine/sph_gravity.cu:810-961 (group size G parameter)
 *   - pass-1 neighbour count (r < H_i)   engine/sph_gravity.cu:645-788
 *
 * Parity status of each function is listed in the header of oracle.cpp.
 * Field ids and the params struct are those of include/astraea_hip.h.
 */
#ifndef ASTRAEA_ORACLE_H
#define ASTRAEA_ORACLE_H

#include <stdint.h>

#include "../include/astraea_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

or_ctx* or_create(void);
void or_destroy(or_ctx* c);
/* group_size: particles sharing one gravity opening decision (reference: 32;
 * the HIP path: ASTRAEA_GROUP_SIZE). */
int or_set_params(or_ctx* c, const astraea_params* p, int group_size);
int or_set_layout(or_ctx* c, uint32_t num_segments, const uint32_t* counts);
int or_upload_field(or_ctx* c, int field, const void* src);
/* float fields are written as float32 (rounded from FP64), u8/u32 as is */
int or_download_field(or_ctx* c, int field, void* dst);
/* float fields as FP64 (the oracle's own values) */
int or_download_field_f64(or_ctx* c, int field, double* dst);

int or_clear_stats(or_ctx* c);
int or_sort_and_build(or_ctx* c);
int or_pass1(or_ctx* c);
int or_eos(or_ctx* c);
int or_gravity(or_ctx* c);
int or_pass2(or_ctx* c);
int or_reduce_dt(or_ctx* c, double max_dt_psu, double* out_dt);
int or_kick_half(or_ctx* c, int clamp_thermal);
int or_drift(or_ctx* c);
int or_initial_forces(or_ctx* c);
int or_step(or_ctx* c, double max_dt_psu, uint32_t steps);
int or_energy(or_ctx* c, double* out);
int or_counters(or_ctx* c, uint64_t* out, int n);
int or_neighbour_stats(or_ctx* c, uint64_t* out);
int or_set_dt(or_ctx* c, const double* dt);
int or_get_dt(or_ctx* c, double* dt);
int or_download_tree(or_ctx* c, uint32_t seg, uint64_t* keys, uint32_t* perm,
                     int32_t* left, int32_t* right, int32_t* parent_internal, int32_t* parent_leaf,
                     float* node_cm, float* node_min, float* node_max);
int or_num_threads(void);

/* EOS entry points (engine/eos.cu:283-388, 391-534) */
double or_cold_energy_psu(int mat, double rho_psu);
void or_eos_eval(int mat, float rho_psu, float u_psu, float damage, double gamma_hhe,
                 float* p, float* cs, float* t, uint8_t* phase);
void or_eos_from_rho_temp_si(int mat, double rho_kgm3, double temp_k, double* u_si, double* p_si);
double or_invert_density_si(int mat, double p_si, double temp_k, double rho_guess);
/* Tillotson table entry k of material mat (order of astraea TillotsonParams, eos.hpp:18-37) */
double or_material_param(int mat, int k);

/* Morton key of one point for given FP32 bounds {minx,miny,minz,maxx,maxy,maxz} */
uint64_t or_morton_key(float x, float y, float z, const float* bounds6);

/* Damped relaxation of the loaded bodies, one per segment (engine/simulation.cpp:585-781);
 * anchor_temp (n_anchor doubles, K, by original id) or NULL for the EOS temperatures after the
 * initial forces.  out[s] = v_rms / v_esc of segment s. */
int or_relax(or_ctx* c, uint32_t steps, const double* anchor_temp, uint32_t n_anchor, const double* nominal_radius,
             double* out);

/* Largest-remnant classification of one body pair's arrays (engine/simulation.cpp:1000-1227),
 * in input order; `out` receives ASTRAEA_CLASS_STRIDE doubles (ASTRAEA_CL_* order). */
void or_classify_arrays(uint32_t n, const double* x, const double* y, const double* z, const double* vx,
                        const double* vy, const double* vz, const double* m, const float* phi, const uint8_t* body,
                        const float* damage, double target_radius, double* out);

#ifdef __cplusplus
}
#endif

#endif
