/*
 * nbl_hip.h -- C ABI of the MI355X electron-transport library (libnbl_hip.so).
 *
 * The reference (Nebula 1.0.2) exposes this hot path as compile-time C++
 * templates, not as an FFI.  Each entry point below replaces one member of
 * that template API; the replaced reference interface is cited per function
 * (paths relative to the reference's source/ directory).
 *
 * Conventions
 *   - Every function returns an int status: NBL_OK (0) or a negative NBL_ERR_*.
 *     nbl_last_error() returns a thread-local message for the last failure.
 *   - No C++ exception crosses this boundary.
 *   - Pointers to particles, triangles, tables and results are HOST pointers
 *     owned by the caller for the duration of the call.  The library owns all
 *     device memory it allocates and its own HIP streams.
 *   - Handles are not thread-safe; use one driver per host thread and device
 *     (as gpu_main.cu:77-82 does).
 *   - All arithmetic is IEEE float32 (real=float, config/config.h:13).
 */
#ifndef NBL_HIP_H
#define NBL_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define NBL_OK             0
#define NBL_ERR_INVALID   (-1) /* bad argument or handle */
#define NBL_ERR_ALLOC     (-2) /* device/host allocation failed (bad_alloc) */
#define NBL_ERR_HIP       (-3) /* HIP runtime error */
#define NBL_ERR_RUNTIME   (-4) /* runtime_error: malformed input, limits */
#define NBL_ERR_NODEVICE  (-5) /* no HIP device visible */

/* Triangle id meaning "no triangle" (the reference's nullptr triangle*). */
#define NBL_NO_TRIANGLE 0xFFFFFFFFu

/* Special (non-physical) material indices: core/cpu_material_manager.h:82-91 */
#define NBL_MAT_NOP            (-128)
#define NBL_MAT_TERMINATOR     (-127)
#define NBL_MAT_DETECTOR       (-126)
#define NBL_MAT_DETECTOR_LT50  (-125)
#define NBL_MAT_DETECTOR_GE50  (-124)
#define NBL_MAT_VACUUM         (-123)
#define NBL_MAT_MIRROR         (-122)

/* ---- plain data types, layout-identical to the reference --------------- */
typedef struct { float x, y, z; } nbl_vec3;                          /* vec3, config/data_types.h:38-40 */
typedef struct { nbl_vec3 pos, dir; float kin_energy; } nbl_particle; /* particle, core/particle.h:5-10 (28 B) */
typedef struct {                                                     /* triangle, core/triangle.h:8-77 (44 B) */
    int32_t material_in;  /* material on the side the normal points away from */
    int32_t material_out; /* material on the side the normal points towards   */
    nbl_vec3 r0, e1, e2;  /* vertex 0 and the edges r1-r0, r2-r0             */
} nbl_triangle;

/* 1D table: data[i] at x = x_min + i*(x_max-x_min)/(n-1)  (common/util/table_1D.h:46-48) */
typedef struct {
    int32_t n;
    float x_min, x_max;
    const float* data;
} nbl_table1d_desc;

/* 2D table: value(i,j) = data[i*height + j]; i along x (width), j along y
 * (height).  Same storage order as table_2D<real,false> (table_2D.inl:151-157). */
typedef struct {
    int32_t width, height;
    float x_min, x_max, y_min, y_max;
    const float* data;
} nbl_table2d_desc;

/* 3D table: value(i,j,k) = data[(i*height + j)*depth + k]; i along x, j along
 * y, k along z (table_3D.inl:70-74 storage order). */
typedef struct {
    int32_t width, height, depth;
    float x_min, x_max, y_min, y_max, z_min, z_max;
    const float* data;
} nbl_table3d_desc;

#define NBL_INELASTIC_KIEFT 0   /* kieft_inelastic (physics/kieft/inelastic.h)            */
#define NBL_INELASTIC_PENN  1   /* full_penn (physics/full_penn.h), the reference default */

/* A material with Kieft-Bosch elastic physics and Kieft-Bosch or full Penn
 * inelastic physics, tables already prepared as the reference's create()
 * functions leave them (physics/kieft/elastic.h:120-166,
 * physics/kieft/inelastic.h:212-327, physics/full_penn.h:296-399).  For full
 * Penn, inelastic_log_imfp and inelastic_log_w0_icdf hold its log IMFP and log
 * omega ICDF (same roles, same prep), penn_q_icdf and penn_ionisation its 3D
 * tables, and ionisation is unused.  All materials of a set use one model. */
typedef struct {
    float barrier;      /* E_F + work function, eV (core/material.h:52-57)            */
    float fermi;        /* eV (inelastic.h:220)                                       */
    float band_gap;     /* eV; <0 for metals (inelastic.h:221)                        */
    float phonon_loss;  /* eV (elastic.h:160)                                         */
    float recoil_const; /* 2*(2 m_e)/M_atom (elastic.h:161-163)                       */
    nbl_table1d_desc elastic_log_imfp;       /* log(imfp/nm^-1) vs log(K/eV)          */
    nbl_table2d_desc elastic_costheta_icdf;  /* cos(theta) vs (log K, P), clamped     */
    nbl_table1d_desc inelastic_log_imfp;     /* log(imfp/nm^-1) vs log(K/eV)          */
    nbl_table2d_desc inelastic_log_w0_icdf;  /* log(omega'/eV) vs (log K, P)          */
    nbl_table2d_desc ionisation;             /* binding/eV vs (log omega', P), rounddown */
    int32_t inelastic_model;                 /* NBL_INELASTIC_KIEFT or NBL_INELASTIC_PENN */
    nbl_table3d_desc penn_q_icdf;            /* hbar q/sqrt(2m) / eV^1/2 vs (log K, omega/K, P) */
    nbl_table3d_desc penn_ionisation;        /* binding/eV vs (log K, log(omega/K), P), rounddown; -1 = none */
} nbl_material_desc;

/* Raw full Penn data (the /full_penn and /ionization groups of a material
 * file) before full_penn::create's post-processing.  Energy axes are
 * log-spaced, probability and omega/K (q table) axes linear, the omega/K axis
 * of the binding table log-spaced. */
typedef struct {
    float fermi, band_gap;
    int32_t n_k; double k_min, k_max;        /* IMFP and omega ICDF energy grid, eV   */
    int32_t n_p; double p_min, p_max;        /* omega ICDF probability grid           */
    const float* imfp;                       /* [n_k]         nm^-1                   */
    const float* omega_icdf;                 /* [n_k][n_p]    eV                      */
    int32_t q_nk, q_nw, q_np;
    double q_k_min, q_k_max, q_w_min, q_w_max, q_p_min, q_p_max;
    const float* q_icdf;                     /* [q_nk][q_nw][q_np]  q in nm^-1        */
    int32_t i_nk, i_nw, i_np;
    double i_k_min, i_k_max, i_w_min, i_w_max, i_p_min, i_p_max;
    const float* binding_icdf;               /* [i_nk][i_nw][i_np]  eV                */
} nbl_penn_raw;

/* Raw cross-section data as a cross-section tool writes it, before the
 * reference's create() post-processing.  All energy grids are log-spaced
 * from k_min to k_max (eV); all probability grids linear from p_min to p_max. */
typedef struct {
    float barrier, fermi, band_gap, phonon_loss;
    double effective_a_grams;       /* mass of the effective atom, g (elastic.h:161-163) */
    int32_t n_k;                    /* energy points, shared by every table             */
    double k_min, k_max;            /* eV                                               */
    int32_t n_p;                    /* probability points, shared by every 2D table     */
    double p_min, p_max;
    const float* elastic_imfp;      /* [n_k]        nm^-1                               */
    const float* elastic_costheta;  /* [n_k][n_p]   cos(theta) ICDF                     */
    const float* inelastic_imfp;    /* [n_k]        nm^-1                               */
    const float* inelastic_w0;      /* [n_k][n_p]   omega' ICDF, eV                     */
    const float* binding_icdf;      /* [n_k][n_p]   binding energy ICDF, eV (K axis = omega') */
    int32_t n_outer_shells;
    const double* outer_shells;     /* eV, any order                                    */
} nbl_kieft_raw;

/* ---- library ------------------------------------------------------------ */
const char* nbl_last_error(void);
int nbl_version(int* major, int* minor, int* patch);
/* Name of device 0 and its compute-unit count. */
int nbl_device_info(char* name, size_t name_len, int* compute_units);
/* No reference counterpart: a plain device-to-device copy kernel of `bytes`
 * (multiple of 16) run `reps` times, for scaling hardware byte counters and
 * as an achievable-bandwidth yardstick; *ms_per_copy = mean duration. */
int nbl_bench_copy(uint64_t bytes, int32_t reps, double* ms_per_copy);

/* ---- geometry: replaces geometry::octree<true> (geometry/octree.h:31-111) --- */
typedef struct nbl_geometry_s* nbl_geometry;

typedef struct {
    nbl_vec3 aabb_min, aabb_max;  /* simulation domain (octree.h:93-97)              */
    float max_extent;             /* 2*|halfsize| (octree.inl:300)                    */
    uint32_t n_triangles_in;      /* triangles given to create()                       */
    uint32_t n_triangles_unique;  /* triangles kept after duplicate removal            */
    uint32_t n_nodes;             /* internal octree nodes (root included)             */
    uint32_t n_leaves;
    uint32_t n_leaf_refs;         /* leaf->triangle references                         */
    uint32_t max_depth;
    uint64_t linearized_size;     /* size of the reference-format octree_data array    */
    uint64_t device_bytes;        /* device memory held by the geometry                */
} nbl_geometry_info;

/* octree<true>::create(std::vector<triangle>) (octree.h:42-49, octree.inl:8-21):
 * builds the octree with the reference's subdivision rules on the host, then
 * uploads it.  Triangle ids returned by propagate are indices into tris. */
int nbl_geometry_create(const nbl_triangle* tris, uint32_t n, nbl_geometry* out);
/* octree<true>::destroy (octree.h:54) */
int nbl_geometry_destroy(nbl_geometry g);
int nbl_geometry_get_info(nbl_geometry g, nbl_geometry_info* info);
/* The octree in the reference's linearized form (octree_builder.h:209-248):
 * octree_data (linearized_size ints) and, for each position of triangle_data,
 * the index of that triangle in the create() input.  Either pointer may be NULL. */
int nbl_geometry_export_linearized(nbl_geometry g, int32_t* octree_data, uint32_t* triangle_order);
/* linearized_octree(octree_root(triangles)) (octree_builder.cpp:171-352) on
 * the host only, no device needed: the reference-format octree_data
 * (*octree_size ints; pass octree_data = NULL to query the size), the input
 * index of each triangle_data position (*n_unique entries) and the root box.
 * Any output pointer may be NULL. */
int nbl_octree_linearize(const nbl_triangle* tris, uint32_t n, int32_t* octree_data,
    uint64_t* octree_size, uint32_t* triangle_order, uint32_t* n_unique,
    float* center3, float* halfsize3);
/* octree::in_domain (octree.inl:24-30), batched on the device. */
int nbl_geometry_in_domain_batch(nbl_geometry g, const nbl_vec3* pos, uint32_t n, uint8_t* out);
/* octree::propagate (octree.inl:32-192), batched on the device.  One ray per
 * element; ignore_tri = NBL_NO_TRIANGLE for none.  out_tri = NBL_NO_TRIANGLE
 * when no triangle is hit; out_dist as the reference's isect_distance. */
int nbl_geometry_propagate_batch(nbl_geometry g,
    const nbl_vec3* pos, const nbl_vec3* dir, const float* dist,
    const uint32_t* ignore_tri, const int32_t* ignore_mat, uint32_t n,
    float* out_dist, uint32_t* out_tri);
/* Same, with pointers to DEVICE memory and no host staging (the timed form). */
int nbl_geometry_propagate_batch_device(nbl_geometry g,
    const nbl_vec3* pos, const nbl_vec3* dir, const float* dist,
    const uint32_t* ignore_tri, const int32_t* ignore_mat, uint32_t n,
    float* out_dist, uint32_t* out_tri);

/* ---- materials: replaces gpu_material_manager (core/gpu_material_manager.h:19-79) */
typedef struct nbl_materials_s* nbl_materials;

/* kieft_elastic::create + kieft_inelastic::create table post-processing
 * (elastic.h:120-166, inelastic.h:212-327) on the host.  The five output
 * arrays must hold n_k, n_k*n_p, n_k, n_k*n_p and n_k*n_p floats; desc is
 * filled to point at them. */
int nbl_kieft_prepare(const nbl_kieft_raw* raw,
    float* el_log_imfp, float* el_icdf, float* in_log_imfp,
    float* in_log_w0_icdf, float* ionisation, nbl_material_desc* desc);
/* gpu_material_manager::create(cpu_material_manager const&)
 * (gpu_material_manager.inl:5-26): uploads n materials. */
int nbl_materials_create(const nbl_material_desc* mats, int32_t n, nbl_materials* out);
/* full_penn::create (physics/full_penn.h:296-399): prepares the full Penn
 * tables into the caller's arrays (log_imfp [n_k], log_omega_icdf [n_k*n_p],
 * q_icdf [q_nk*q_nw*q_np], ionisation [i_nk*i_nw*i_np]) and switches d
 * (elastic part already filled by nbl_kieft_prepare) to the full Penn model. */
int nbl_penn_prepare(const nbl_penn_raw* raw, float* log_imfp, float* log_omega_icdf, float* q_icdf,
    float* ionisation, nbl_material_desc* d);
int nbl_materials_destroy(nbl_materials m);
/* Maximal energy accepted by all materials (cpu_material_manager::get_max_energy,
 * scatter_list.h:64-69, elastic.h:46-49). */
int nbl_materials_get_max_energy(nbl_materials m, float* out);

/* Table selectors for the lookup entry points. */
#define NBL_TAB_ELASTIC_IMFP      0 /* 1D */
#define NBL_TAB_INELASTIC_IMFP    1 /* 1D */
#define NBL_TAB_ELASTIC_ICDF      2 /* 2D */
#define NBL_TAB_INELASTIC_W0_ICDF 3 /* 2D */
#define NBL_TAB_IONISATION        4 /* 2D */
/* table_1D::get (table_1D.inl:65-81), batched on the device. */
int nbl_table1d_get_batch(nbl_materials m, int32_t material, int32_t table,
    const float* x, uint32_t n, float* out);
/* table_2D::get (table_2D.inl:73-94) or, with rounddown != 0,
 * table_2D::get_rounddown (table_2D.inl:96-104), batched on the device. */
int nbl_table2d_get_batch(nbl_materials m, int32_t material, int32_t table,
    int32_t rounddown, const float* x, const float* y, uint32_t n, float* out);
#define NBL_TAB_PENN_Q_ICDF       5 /* 3D */
#define NBL_TAB_PENN_IONISATION   6 /* 3D */
/* table_3D::get (table_3D.inl:78-106) or, with rounddown != 0,
 * table_3D::get_rounddown (table_3D.inl:108-118), batched on the device
 * (full Penn materials only). */
int nbl_table3d_get_batch(nbl_materials m, int32_t material, int32_t table,
    int32_t rounddown, const float* x, const float* y, const float* z, uint32_t n, float* out);

/* ---- single-event physics with injected uniforms (unit-parity entry points) */
/* Event kinds. */
#define NBL_EVENT_SAMPLE_PATH 0 /* material::sample_path (scatter_list.h:106-112)        */
#define NBL_EVENT_ELASTIC     1 /* kieft_elastic::execute (elastic.h:67-115)              */
#define NBL_EVENT_INELASTIC   2 /* kieft_inelastic::execute (inelastic.h:68-206), or
                                   full_penn::execute (full_penn.h:66-288) for full Penn materials */
#define NBL_EVENT_INTERSECT   3 /* boundary_intersect::execute (boundary_intersect.h:39-156) */
/* Outcome codes written to out_status. */
#define NBL_OUT_ALIVE       0
#define NBL_OUT_DETECTED    1
#define NBL_OUT_TERMINATED  2
#define NBL_NUM_UNIFORMS    8 /* uniforms reserved per particle; draws are taken in order */
/* Runs one event for each of n particles on the device.  For particle i the
 * j-th uniform draw returns uniforms[i*NBL_NUM_UNIFORMS + j].
 *   material[i]      current material index; last_tri[i] the intersected triangle
 *                    (NBL_EVENT_INTERSECT only; g may be NULL otherwise)
 *   io               particles, updated in place
 *   out_material     new material index (changes only on transmission)
 *   out_status       NBL_OUT_*
 *   out_secondary    secondary electron (NBL_EVENT_INELASTIC); out_has_secondary 0/1
 *   out_scalar       NBL_EVENT_SAMPLE_PATH: free path; out_type: event type 0/1/2
 *                    (scatter_event, core/events.h:9-18) */
int nbl_physics_batch(nbl_materials m, nbl_geometry g, int32_t event,
    const int32_t* material, const uint32_t* last_tri, const float* uniforms,
    uint32_t n, nbl_particle* io, int32_t* out_material, uint8_t* out_status,
    nbl_particle* out_secondary, uint8_t* out_has_secondary,
    float* out_scalar, uint8_t* out_type);

/* ---- driver: replaces drivers::gpu_driver (drivers/gpu/gpu_driver.h:25-229) - */
typedef struct nbl_driver_s* nbl_driver;

typedef struct {
    uint32_t threads_per_block; /* 0 = library default (256)                      */
    uint32_t reserved[7];
} nbl_driver_opts;

/* gpu_driver::gpu_driver (gpu_driver.inl:51-89).  seed seeds the per-slot
 * XORWOW streams (Marsaglia 2003; project seeding, not cuRAND's). */
int nbl_driver_create(uint32_t capacity, nbl_geometry g, nbl_materials m,
    float min_energy, float max_energy, uint64_t seed,
    const nbl_driver_opts* opts, nbl_driver* out);
/* gpu_driver::~gpu_driver (gpu_driver.inl:95-117) */
int nbl_driver_destroy(nbl_driver d);
/* gpu_driver::push (gpu_driver.inl:123-130, gpu_particle_manager.inl:105-164) */
int nbl_driver_push(nbl_driver d, const nbl_particle* particles,
    const uint32_t* tags, uint32_t n, uint32_t* pushed);
/* gpu_driver::do_iteration (gpu_driver.inl:312-317): init, sort, events. */
int nbl_driver_do_iteration(nbl_driver d);
/* gpu_driver::get_running_count / get_detected_count (gpu_driver.inl:143-160) */
int nbl_driver_get_running_count(nbl_driver d, uint32_t* out);
int nbl_driver_get_detected_count(nbl_driver d, uint32_t* out);
/* Detected-particle callback: (particle, tag of the primary, user). */
typedef void (*nbl_detect_fn)(const nbl_particle* p, uint32_t tag, void* user);
/* gpu_driver::flush_detected (gpu_driver.inl:132-141, gpu_particle_manager.inl:206-235) */
int nbl_driver_flush_detected(nbl_driver d, nbl_detect_fn fn, void* user);
/* gpu_driver::allocate_input_buffers (gpu_driver.inl:163-212) */
int nbl_driver_allocate_input_buffers(nbl_driver d, uint32_t n);
/* work_pool::get_work(bool*, particle*, uint32_t*, N) (common/work_pool.h:69-88):
 * fill the free entries (data[i]==0) of the n-long input buffer, set data[i]=1. */
typedef void (*nbl_pool_fill_fn)(uint8_t* data, nbl_particle* particles,
    uint32_t* tags, uint32_t n, void* user);
/* gpu_driver::push_to_buffer (gpu_driver.inl:214-242) */
int nbl_driver_push_to_buffer(nbl_driver d, nbl_pool_fill_fn fill, void* user);
/* gpu_driver::push_to_simulation (gpu_driver.inl:244-253) */
int nbl_driver_push_to_simulation(nbl_driver d);
/* gpu_driver::buffer_detected (gpu_driver.inl:255-269) */
int nbl_driver_buffer_detected(nbl_driver d);
/* gpu_driver::flush_buffered (gpu_driver.inl:271-306); *running = active count */
int nbl_driver_flush_buffered(nbl_driver d, nbl_detect_fn fn, void* user, uint32_t* running);

/* Slot-state access for parity tests and tools (no reference counterpart:
 * the reference reads these arrays through cuda_mem_scope round trips,
 * common/cuda/cuda_mem_scope.h:21-38).  Arrays are capacity long; any
 * pointer may be NULL.  rng holds 6 uint32 per slot. */
typedef struct {
    uint8_t* status; uint32_t* particle_idx; nbl_particle* particles;
    uint32_t* tags; int32_t* material; uint32_t* last_tri; uint32_t* rng;
} nbl_slot_view;
int nbl_driver_read_slots(nbl_driver d, const nbl_slot_view* v);
int nbl_driver_write_slots(nbl_driver d, const nbl_slot_view* v);
/* Status codes of a slot (gpu_particle_manager.h:93-103). */
#define NBL_ST_PENDING         0x0
#define NBL_ST_INELASTIC_EVENT 0x4
#define NBL_ST_ELASTIC_EVENT   0x1
#define NBL_ST_INTERSECT_EVENT 0x2
#define NBL_ST_NO_EVENT        0x6
#define NBL_ST_DETECTED        0xA
#define NBL_ST_NEW_SECONDARY   0xE
#define NBL_ST_TERMINATED      0x3

/* Stable sort of n one-byte keys on bits [0,2), payload = index: the
 * in-house replacement of cub::DeviceRadixSort::SortPairs(..., 0, 2)
 * (gpu_particle_manager.inl:237-246).  Host arrays. */
int nbl_sort_status_batch(const uint8_t* keys, uint32_t n, uint32_t* out_index);

/* ---- XORWOW streams (replaces cuRAND, common/util/random.inl:35-62) ------- */
/* Seed n states for sequences first_sequence .. first_sequence+n-1; 6 words each. */
int nbl_rng_seed(uint64_t seed, uint64_t first_sequence, uint32_t n, uint32_t* states);
/* Draw `draws` uniforms in (0,1] from each of n states, on the device. */
int nbl_rng_uniform_batch(const uint32_t* states, uint32_t n, uint32_t draws, float* out);

/* ---- persistent batched transport (no reference counterpart: replaces the
 * prescan + frame loop of gpu_main.cu:108-213 with one launch) ------------ */
/* A scan pixel: n primaries entering at (x,y,z) + beam spread, direction dir. */
typedef struct {
    float x, y, z;           /* beam centre, nm                                     */
    float energy;            /* eV                                                  */
    float sigma;             /* Gaussian beam sigma, nm (profile 0)                 */
    float half_x, half_y;    /* half widths of a uniform rectangle, nm (profile 1)  */
    uint32_t profile;        /* 0 = Gaussian, 1 = uniform rectangle                 */
    uint32_t n_primaries;
    uint32_t reserved;
} nbl_pixel;
typedef struct {
    const nbl_pixel* pixels;
    uint32_t n_pixels;
    nbl_vec3 dir;            /* common beam direction (normalised by the kernel)   */
    uint64_t seed;           /* primary-sampling seed                              */
} nbl_pixel_batch;

#define NBL_HIST_BINS 1024

/* Energy-deposition tally (the GPU counterpart of energydeposit_cpu_driver,
 * drivers/cpu/energydeposit_cpu_driver.h:44-94): every elastic or inelastic
 * event that changes the electron's energy deposits (energy before - energy
 * after) at the event position; the deposits are summed into a voxel grid
 * instead of written as 28-byte records (cpu_energydep.cpp:184-192). */
typedef struct {
    uint32_t nx, ny, nz;     /* voxels along x, y, z                                  */
    nbl_vec3 lo, hi;         /* grid box, nm                                          */
    double* energy;          /* [nz][ny][nx] out: deposited energy per voxel, eV      */
    double total;            /* out: all deposited energy, eV (inside the box or not) */
    uint64_t count;          /* out: deposit events                                    */
} nbl_deposit_grid;

/* Detected electrons of a persistent run as records (the reference's output,
 * gpu_main.cu:199-206, before the pixel lookup), in detection order: the
 * particle at the detector and its primary's tag.  count is every detection;
 * only the first min(count, capacity) are stored (capacity 0: count only). */
typedef struct {
    nbl_particle* particles; /* [capacity] out                                        */
    uint32_t* tags;          /* [capacity] out                                        */
    uint64_t capacity;
    uint64_t count;          /* out                                                    */
} nbl_detection_buffer;

/* Energy-deposition events as records, as cpu_energydep.cpp:184-192 writes
 * them before the pixel lookup: the event position (the particle after the
 * event), its energy before and the energy it lost, in event order of
 * completion; tags are the primaries' tags.  count is every deposit event;
 * only the first min(count, capacity) are stored (capacity 0: count only). */
typedef struct {
    nbl_vec3 pos;            /* event position, nm                                     */
    float energy;            /* kinetic energy before the event, eV                    */
    float deposit;           /* energy before - energy after, eV (> 0)                 */
} nbl_deposit_record;        /* 20 B */
typedef struct {
    nbl_deposit_record* records; /* [capacity] out                                     */
    uint32_t* tags;              /* [capacity] out                                     */
    uint64_t capacity;
    uint64_t count;              /* out                                                 */
} nbl_deposit_buffer;

typedef struct {
    uint64_t primaries;      /* primaries whose cascades ended                    */
    uint64_t detected_se;    /* detected with E < 50 eV                            */
    uint64_t detected_bse;   /* detected with E >= 50 eV                           */
    uint64_t events;         /* init + event steps (one per active particle step);
                                this and the work counters below need
                                nbl_persistent_opts.count_work in persistent runs  */
    uint64_t lookups_1d, lookups_2d, rounddowns;
    uint64_t node_steps, tri_tests, boundary_events, secondaries, pending_retries;
    uint64_t stopped;        /* electrons dropped: 2^20-step cap reached, or a
                                secondary with no stack space left (normally 0;
                                nonzero is an error unless opts->allow_drops) */
    double   kernel_ms;      /* device time of the transport launch(es)            */
    uint32_t* pixel_se;      /* optional [n_pixels] out (may be NULL)              */
    uint32_t* pixel_bse;     /* optional [n_pixels] out (may be NULL)              */
    uint32_t* hist;          /* optional [NBL_HIST_BINS] detected-energy histogram, 1-eV bins */
    nbl_deposit_grid* deposit; /* optional energy-deposition tally (may be NULL)      */
    nbl_detection_buffer* detected; /* optional detected-electron records (may be NULL) */
    nbl_deposit_buffer* deposits;   /* optional energy-deposition records (may be NULL)  */
} nbl_tallies;

typedef struct {
    uint32_t block_size;       /* threads per workgroup (64..1024, multiple of 64); 0 = default
                                  (128; 256 with a learned sampler)                       */
    uint32_t blocks_per_cu;    /* resident workgroups per CU; 0 = default (10: 5 wavefronts
                                  per SIMD; 4 with a learned sampler, 3 with it and a
                                  deposit tally)                                          */
    uint32_t stack_depth;      /* per-wavefront secondary stack entries in LDS; 0 = default
                                  (96; 128 with a learned sampler)                        */
    uint32_t refill_threshold; /* refill a wavefront when >= this many lanes idle; 0 = default (4) */
    uint32_t overflow_depth;   /* further per-wavefront stack entries in HBM; 0 = default (16384) */
    uint32_t inelastic_batch;  /* run inelastic events once this many lanes of a wavefront
                                  want one (1 = at once); 0 = default (8)                  */
    uint32_t inelastic_wait;   /* ... or once one has waited this many loop iterations;
                                  0 = default (8)                                          */
    uint32_t allow_drops;      /* 0 (default): a run that stopped any electron (a
                                  secondary with no stack room left, or the 2^20-step
                                  cap) returns NBL_ERR_RUNTIME, tallies filled but
                                  biased; 1: such a run returns NBL_OK and counts them
                                  in nbl_tallies.stopped                                   */
    uint32_t count_work;       /* 1: fill the work counters of the tallies: events,
                                  lookups_*, rounddowns, node_steps, tri_tests,
                                  boundary_events, secondaries, pending_retries;
                                  0 (default): they stay 0 and the launch runs about 4 %
                                  faster.  Runs with a deposit tally always count.
                                  primaries, detected_* and stopped are always
                                  filled.                                                 */
} nbl_persistent_opts;

/* Runs every primary of the batch to the end of its cascade in one
 * persistent launch; per-pixel SE/BSE tallies and counters go to *out.
 * An inelastic event waits (PENDING, gpu_driver.inl:523-528) while its
 * wavefront's secondary stack is full; only a wavefront whose 64 lanes all
 * wait at once goes ahead and drops what does not fit.  Unless
 * opts->allow_drops, a run that dropped or step-capped any electron returns
 * NBL_ERR_RUNTIME (with *out filled) instead of silently biased tallies. */
int nbl_driver_run_persistent(nbl_driver d, const nbl_pixel_batch* batch,
    const nbl_persistent_opts* opts, nbl_tallies* out);
/* The same transport over an explicit list of n primaries with tags (the
 * primaries of gpu_main.cu's work pool, pushed by push_to_buffer in the
 * reference); each starts in vacuum.  Per-pixel tallies are not available
 * (out->pixel_se and pixel_bse must be NULL): use out->detected. */
int nbl_driver_run_persistent_list(nbl_driver d, const nbl_particle* primaries, const uint32_t* tags, uint64_t n,
    const nbl_persistent_opts* opts, nbl_tallies* out);

/* HBM secondary-stack space of the allocation-free variant below, per GPU. */
#define NBL_NOALLOC_STACK_BYTES (128u << 20)
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory while they run (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the transport of
 * nbl_driver_run_persistent_list that allocates nothing.  Primaries, tags and
 * the detection records are DEVICE arrays owned by the caller (records in
 * detection order, the first min(count, detected_capacity) stored; count =
 * out->detected_se + out->detected_bse).  The kernel's scratch (primary
 * queues, work counters, histogram, per-wavefront HBM secondary stacks) is one
 * fixed library array per GPU, NBL_NOALLOC_STACK_BYTES of stack plus 10 KiB,
 * held until the process ends.  opts->overflow_depth defaults to what fits
 * (at most 16384); a depth x wavefront count beyond the stack space is
 * refused (NBL_ERR_INVALID).  out->hist is a host array or NULL;
 * out->pixel_se, pixel_bse, deposit, detected and deposits must be NULL.  Invalid
 * primaries are refused as by the list entry point (checked on the device).
 * Calls that overlap on one GPU share the array and race; run them one at a
 * time per device.  The driver's stream runs the work; the call returns when
 * it has finished. */
int nbl_driver_run_persistent_list_noalloc(nbl_driver d, const nbl_particle* dev_primaries,
    const uint32_t* dev_tags, uint64_t n, nbl_particle* dev_detected, uint32_t* dev_detected_tags,
    uint64_t detected_capacity, const nbl_persistent_opts* opts, nbl_tallies* out);

/* ---- file formats (io/load_tri_file.cpp, io/load_pri_file.cpp), host only ---- */
typedef struct { int32_t x, y; } nbl_int2;     /* int2 of io/load_pri_file.h:10 (pixel index) */

/* load_tri_file (load_tri_file.cpp:8-60): text, one triangle per line with 11
 * columns "material_in material_out Ax Ay Az Bx By Bz Cx Cy Cz" (nm); blank
 * lines and lines starting with '#' are skipped.  Two calls: with out == NULL
 * only *n is set.  A file that cannot be opened gives *n = 0 (the reference
 * returns an empty list); a line with another column count is an
 * NBL_ERR_RUNTIME error naming the line. */
int nbl_tri_file_load(const char* path, nbl_triangle* out, uint32_t capacity, uint32_t* n);

/* Primaries that load_pri_file (load_pri_file.cpp:122-154) warns about; they are kept. */
typedef struct {
    uint64_t outside_geometry;   /* position not strictly inside [aabb_min, aabb_max] */
    uint64_t low_energy;         /* kin_energy <= EPSILON                              */
    uint64_t high_energy;        /* kin_energy > max_energy                             */
    uint64_t bad_direction;      /* every |dir component| <= EPSILON                    */
} nbl_pri_warnings;

/* load_pri_file (load_pri_file.cpp:72-187): binary records of 36 bytes,
 * 7 float32 (pos, dir, energy) + 2 int32 (pixel x, y).  Two calls: with out,
 * pixels and warnings all NULL only *n is set.  A file that cannot be opened
 * gives *n = 0; a size that is not a multiple of 36 B is an error. */
int nbl_pri_file_load(const char* path, nbl_vec3 aabb_min, nbl_vec3 aabb_max, float max_energy,
    nbl_particle* out, nbl_int2* pixels, uint64_t capacity, uint64_t* n, nbl_pri_warnings* warnings);
/* sort_pri_file (load_pri_file.cpp:189-229): in place, by the Morton code of
 * the positions quantised to 16 bits per axis over their bounding box. */
int nbl_pri_sort_morton(nbl_particle* particles, nbl_int2* pixels, uint64_t n);
/* prescan_shuffle (load_pri_file.cpp:231-243): in place, the first
 * prescan_size places get a uniformly random pick of the rest. */
int nbl_pri_prescan_shuffle(nbl_particle* particles, nbl_int2* pixels, uint64_t n, uint64_t prescan_size);

/* ---- learned ICDF sampler (no reference counterpart: stands in for the
 * elastic cos(theta) table of kieft_elastic, physics/kieft/elastic.h:67-87,
 * and the inelastic omega' table of kieft_inelastic, inelastic.h:68-72) ---- */
/* Two BF16 MLPs (elastic, inelastic) shared by all materials of a material
 * set: features (log E, u, log u, log(1-u), 8-dim learned material
 * embedding) -> 3 x 64 ReLU -> 1, FP32 accumulation on the matrix cores.
 * Trained at creation on samples drawn from the material tables. */
typedef struct nbl_sampler_s* nbl_sampler;
#define NBL_SAMPLER_ELASTIC    0   /* value: cos(theta), as the elastic table */
#define NBL_SAMPLER_INELASTIC  1   /* value: log(omega'), as the inelastic table */
/* FP32 parameter block of one network, in floats:
 *   W1 [64][16] @0, b1 [64] @1024, W2 [64][64] @1088, b2 @5184, W3 [64][64] @5248,
 *   b3 @9344, w4 [64] @9408, b4 @9472 (+3 pad), embeddings [64 materials][8] @9476 */
#define NBL_SAMPLER_PARAMS 9988
typedef struct {
    uint64_t samples;        /* training samples per network; 0 = 1e8            */
    uint32_t batch;          /* samples per step (multiple of 256); 0 = 4096      */
    float lr;                /* Adam learning rate (cosine decay); 0 = 4e-3       */
    float e_min, e_max;      /* kinetic-energy range of the samples, eV; 0 = the tables' */
    uint64_t seed;
    uint32_t kinds;          /* tables the networks replace: bit NBL_SAMPLER_ELASTIC,
                                bit NBL_SAMPLER_INELASTIC; 0 = both.  The others stay
                                tables in the transport (and are not trained). */
    uint32_t reserved[3];
} nbl_sampler_opts;
#define NBL_SAMPLER_KIND_BIT(k) (1u << (k))
typedef struct {
    double loss_elastic, loss_inelastic;   /* final mean squared error, standardised targets */
    double train_ms;
    uint64_t steps, samples;               /* per network */
    double e_min, e_max;                   /* energy range trained on, eV */
} nbl_sampler_report;

/* Trains the networks of opts->kinds (default both) for the materials of m
 * (1..64 materials); the report's loss of a kind not trained is 0. */
int nbl_sampler_create(nbl_materials m, const nbl_sampler_opts* opts, nbl_sampler* out);
int nbl_sampler_destroy(nbl_sampler s);
int nbl_sampler_get_report(nbl_sampler s, nbl_sampler_report* r);
/* FP32 parameters (n = NBL_SAMPLER_PARAMS) and consts[4] = feature-0 shift,
 * feature-0 scale, output shift, output scale. */
int nbl_sampler_get_params(nbl_sampler s, int kind, float* params, uint32_t n, float* consts);
/* The network's value for n samples (host arrays), through the matrix-core
 * path the transport uses: cos(theta) or log(omega') (trained kinds only). */
int nbl_sampler_eval_batch(nbl_sampler s, int kind, const int32_t* mat, const float* energy, const float* u,
    uint32_t n, float* out);
/* Known-answer test of the BF16 matrix-core operand layout (one 16x16x16
 * product); *max_abs_err is 0 on a correct device. */
int nbl_sampler_selftest(double* max_abs_err);
/* Use sampler s (NULL = the tables) in the driver's persistent transport.
 * s must have been created for the driver's material set. */
int nbl_driver_set_sampler(nbl_driver d, nbl_sampler s);

#ifdef __cplusplus
}
#endif

#endif /* NBL_HIP_H */
