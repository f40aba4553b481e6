/*
 * nbl_oracle.h -- C API of the CPU oracle (liborc.so).
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++ restatement of the
 * reference's electron-transport hot path, used as the checker of the HIP
 * library and as the CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Types (nbl_triangle, nbl_particle, nbl_material_desc, ...) are the public
 * ones of include/nbl_hip.h.
 */
#ifndef NBL_ORACLE_H
#define NBL_ORACLE_H

#include "../include/nbl_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry ----------------------------------------------------------- */
typedef struct orc_geometry_s* orc_geometry;
orc_geometry orc_geometry_create(const nbl_triangle* tris, uint32_t n);
void orc_geometry_destroy(orc_geometry g);
void orc_geometry_info(orc_geometry g, float* center3, float* halfsize3,
    float* max_extent, uint64_t* linearized_size, uint32_t* n_unique,
    uint32_t* n_nodes, uint32_t* n_leaves, uint32_t* max_depth);
void orc_geometry_export(orc_geometry g, int32_t* octree_data, uint32_t* tri_order);
void orc_in_domain_batch(orc_geometry g, const nbl_vec3* pos, uint32_t n, uint8_t* out);
/* node_steps / tri_tests may be NULL; otherwise per-ray counters. */
void orc_propagate_batch(orc_geometry g, const nbl_vec3* pos, const nbl_vec3* dir,
    const float* dist, const uint32_t* ignore_tri, const int32_t* ignore_mat,
    uint32_t n, float* out_dist, uint32_t* out_tri,
    uint32_t* node_steps, uint32_t* tri_tests);
/* Brute force over all triangles (geometry/trilist.inl:52-85). */
void orc_trilist_propagate_batch(const nbl_triangle* tris, uint32_t ntri,
    const nbl_vec3* pos, const nbl_vec3* dir, const float* dist,
    const uint32_t* ignore_tri, const int32_t* ignore_mat, uint32_t n,
    float* out_dist, uint32_t* out_tri);
void orc_intersect_ray_batch(const nbl_triangle* tris, const nbl_vec3* pos,
    const nbl_vec3* dir, uint32_t n, float* out_t);

/* ---- materials and tables ---------------------------------------------- */
typedef struct orc_materials_s* orc_materials;
int orc_kieft_prepare(const nbl_kieft_raw* raw, float* el_log_imfp, float* el_icdf,
    float* in_log_imfp, float* in_log_w0_icdf, float* ionisation, nbl_material_desc* desc);
orc_materials orc_materials_create(const nbl_material_desc* mats, int32_t n);
void orc_materials_destroy(orc_materials m);
float orc_materials_max_energy(orc_materials m);
void orc_table1d_get_batch(orc_materials m, int32_t material, int32_t table,
    const float* x, uint32_t n, float* out);
void orc_table2d_get_batch(orc_materials m, int32_t material, int32_t table,
    int32_t rounddown, const float* x, const float* y, uint32_t n, float* out);
/* full_penn::create's table prep (physics/full_penn.h:296-399). */
int orc_penn_prepare(const nbl_penn_raw* raw, float* log_imfp, float* log_omega_icdf,
    float* q_icdf, float* ionisation, nbl_material_desc* desc);
/* n draws of full Penn's plasmon-decay x = E_i / omega at f = E_F / omega
 * (full_penn.h:141-187) from one XORWOW stream. */
void orc_penn_plasmon_x(float f, uint32_t n, uint64_t seed, float* out);
void orc_table3d_get_batch(orc_materials m, int32_t material, int32_t table,
    int32_t rounddown, const float* x, const float* y, const float* z, uint32_t n, float* out);
/* Same contract as nbl_physics_batch. */
int orc_physics_batch(orc_materials m, orc_geometry g, int32_t event,
    const int32_t* material, const uint32_t* last_tri, const float* uniforms,
    uint32_t n, nbl_particle* io, int32_t* out_material, uint8_t* out_status,
    nbl_particle* out_secondary, uint8_t* out_has_secondary,
    float* out_scalar, uint8_t* out_type);

/* ---- XORWOW (Marsaglia 2003) with the project's seeding ----------------- */
void orc_rng_seed(uint64_t seed, uint64_t first_sequence, uint32_t n, uint32_t* states);
void orc_rng_uniform_batch(uint32_t* states, uint32_t n, uint32_t draws, float* out);

/* ---- GPU-driver semantics on the CPU (gpu_driver.inl:312-567) ------------ */
typedef struct orc_slots_s* orc_slots;
orc_slots orc_slots_create(uint32_t capacity, uint64_t seed);
void orc_slots_destroy(orc_slots s);
void orc_slots_read(orc_slots s, const nbl_slot_view* v);
void orc_slots_write(orc_slots s, const nbl_slot_view* v);
uint32_t orc_slots_push(orc_slots s, const nbl_particle* p, const uint32_t* tags, uint32_t n);
/* One do_iteration: init, stable 2-bit sort, intersect, elastic, inelastic. */
void orc_slots_iteration(orc_slots s, orc_materials m, orc_geometry g,
    float min_energy, float max_energy);
void orc_sort_status(const uint8_t* keys, uint32_t n, uint32_t* out_index);

/* ---- CPU driver runs (simple_cpu_driver + cpu_mt_main) ------------------- */
#define ORC_RNG_MT19937 0 /* std::mt19937 + generate_canonical (random.inl:7-30) */
#define ORC_RNG_XORWOW  1 /* one XORWOW stream per driver thread                 */
typedef struct {
    uint64_t primaries, events, detected, detected_se, detected_bse;
    uint64_t lookups_1d, lookups_2d, rounddowns, node_steps, tri_tests;
    uint64_t boundary_events, secondaries;
    double seconds;          /* wall time of the simulation phase */
    uint32_t threads;
} orc_run_stats;
/* cpu_mt_main.cpp:158-219: n_threads drivers pull one primary at a time and
 * run it to the end.  Detected particles go to det/det_tags (up to max_det;
 * stats->detected counts all).  hist (NBL_HIST_BINS, 1-eV bins) may be NULL. */
int orc_cpu_mt_run(orc_geometry g, orc_materials m, const nbl_particle* primaries,
    const uint32_t* tags, uint64_t n, float min_energy, uint32_t seed,
    int32_t n_threads, int32_t rng_kind, double time_limit_s,
    nbl_particle* det, uint32_t* det_tags, uint64_t max_det,
    uint64_t* hist, orc_run_stats* stats);
/* The same run (XORWOW streams, no time limit) with the energy-deposition
 * tally of energydeposit_cpu_driver (energydeposit_cpu_driver.h:72-94) summed
 * into an nx x ny x nz grid over [lo, hi] (energy[(iz*ny + iy)*nx + ix], eV);
 * *total and *count cover every deposit, inside the grid or not. */
int orc_cpu_mt_run_deposit(orc_geometry g, orc_materials m, const nbl_particle* primaries,
    const uint32_t* tags, uint64_t n, float min_energy, uint32_t seed, int32_t n_threads,
    uint64_t* hist, orc_run_stats* stats, int32_t nx, int32_t ny, int32_t nz, const float* lo,
    const float* hi, double* energy, double* total, uint64_t* count);

#ifdef __cplusplus
}
#endif
#endif
