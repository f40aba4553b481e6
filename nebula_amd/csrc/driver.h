// driver.h -- state of an nbl_driver (the gpu_driver replacement).
#pragma once

#include "nbl_host.h"
#include "transport_device.h"

namespace nbl {
const dgeom& geometry_view(nbl_geometry g);
int geometry_max_material(nbl_geometry g);
uint32_t geometry_n_input(nbl_geometry g);
bool geometry_has_record(nbl_geometry g, uint32_t id);
const dmat* materials_device(nbl_materials m);
int materials_count(nbl_materials m);
const dmat& materials_view(nbl_materials m, int i);
float materials_max_energy(nbl_materials m);

// Device view of the slot arrays (gpu_particle_manager.h:79-91), passed by value.
struct dslots {
    uint32_t cap;
    uint8_t* status;
    uint32_t* pidx;        // slot indices sorted by status bits [0,2)
    dparticle* parts;
    uint32_t* tags;
    int32_t* mat;
    uint32_t* last_tri;
    rng_state* rng;
    uint32_t* bucket;      // [4] slots per status bucket after the last sort
};

// In-house stable counting sort of status bits [0,2) with slot payload,
// replacing cub::DeviceRadixSort::SortPairs (gpu_particle_manager.inl:237-246).
struct sorter {
    static constexpr unsigned kBlock = 256, kItems = 16;     // 4096 keys per block
    dbuf<uint32_t> block_counts;                              // [4][n_blocks]
    dbuf<uint32_t> bucket;                                    // [4]
    unsigned n_blocks = 0;
    void init(uint32_t n);
    void run(const uint8_t* keys, uint32_t n, uint32_t* out, hipStream_t s);
};

}  // namespace nbl

struct nbl_driver_s {
    uint32_t cap = 0;
    nbl_geometry geom = nullptr;
    nbl_materials mats = nullptr;
    float min_energy = 0, max_energy = 0;
    uint64_t seed = 0;
    uint64_t runs = 0;                  // persistent launches so far (stream offset)
    unsigned tpb = 256;
    hipStream_t stream = nullptr, copy_stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;   // persistent launch timing

    nbl::dbuf<uint8_t> status;
    nbl::dbuf<uint32_t> pidx;
    nbl::dbuf<nbl::dparticle> parts;
    nbl::dbuf<uint32_t> tags;
    nbl::dbuf<int32_t> mat;
    nbl::dbuf<uint32_t> last_tri;
    nbl::dbuf<nbl::rng_state> rng;
    nbl::sorter sort;
    nbl::dbuf<float> overflow;          // persistent kernel: per-wavefront HBM stacks
    nbl_sampler sampler = nullptr;      // persistent kernel: learned ICDF sampler (NULL = tables)

    // input buffers (gpu_driver.h:211-217)
    uint32_t in_size = 0;
    nbl::dbuf<uint8_t> din_data;
    nbl::dbuf<nbl::dparticle> din_parts;
    nbl::dbuf<uint32_t> din_tags;
    uint8_t* hin_data = nullptr;
    nbl::dparticle* hin_parts = nullptr;
    uint32_t* hin_tags = nullptr;
    // output snapshot (gpu_driver.h:220-225)
    nbl::dbuf<uint8_t> dout_status;
    nbl::dbuf<nbl::dparticle> dout_parts;
    nbl::dbuf<uint32_t> dout_tags;
    uint8_t* hout_status = nullptr;
    nbl::dparticle* hout_parts = nullptr;
    uint32_t* hout_tags = nullptr;

    nbl::dslots view() const
    {
        return { cap, status.p, pidx.p, parts.p, tags.p, mat.p, last_tri.p, rng.p, sort.bucket.p };
    }
    ~nbl_driver_s();
};
