// io.cpp -- the reference's file formats and primary ordering, host only:
// .tri geometry text, .pri primary records, the optional Morton sort of the
// primaries and the prescan shuffle (io/load_tri_file.cpp,
// io/load_pri_file.cpp).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "nbl_host.h"

using namespace nbl;

namespace {

constexpr float EPS = 1.1920928955078125e-06f;    // 10*FLT_EPSILON, config/data_types.h:94
constexpr size_t PRI_BYTES = 7 * sizeof(float) + 2 * sizeof(int32_t);

// load_tri_file.cpp:8-60: whitespace-separated columns, 11 per line
// (material_in material_out Ax Ay Az Bx By Bz Cx Cy Cz); empty lines and
// lines whose first non-blank character is '#' are skipped; numbers are
// read with std::stoi / std::stod and the vertices rounded to float;
// stored as r0 = A, e1 = B - A, e2 = C - A in float.  A file that cannot
// be opened gives no triangles, as in the reference.
std::vector<nbl_triangle> read_tri(const char* path)
{
    std::vector<nbl_triangle> out;
    std::ifstream f(path);
    if (!f.is_open()) return out;
    std::string line;
    size_t line_no = 0;
    while (std::getline(f, line)) {
        ++line_no;
        if (line.empty()) continue;
        const size_t i = line.find_first_not_of(" \t");
        if (i < line.size() && line[i] == '#') continue;
        std::istringstream ss(line);
        std::vector<std::string> col;
        for (std::string t; ss >> t;) col.push_back(t);
        if (col.size() != 11)
            throw std::runtime_error("nbl_tri_file_load: invalid number of columns in line " + std::to_string(line_no) +
                                     " of '" + path + "'");
        nbl_triangle t{};
        t.material_in = std::stoi(col[0]);
        t.material_out = std::stoi(col[1]);
        float v[9];
        for (int k = 0; k < 9; ++k) v[k] = (float)std::stod(col[2 + k]);
        t.r0 = { v[0], v[1], v[2] };
        t.e1 = { v[3] - v[0], v[4] - v[1], v[5] - v[2] };
        t.e2 = { v[6] - v[0], v[7] - v[1], v[8] - v[2] };
        out.push_back(t);
    }
    return out;
}

// Morton code of three 16-bit coordinates (bit i of x, y, z at 3i, 3i+1, 3i+2).
uint64_t morton3(uint16_t x, uint16_t y, uint16_t z)
{
    uint64_t m = 0;
    for (unsigned i = 0; i < 16; ++i) {
        m |= (uint64_t)((x >> i) & 1u) << (3 * i);
        m |= (uint64_t)((y >> i) & 1u) << (3 * i + 1);
        m |= (uint64_t)((z >> i) & 1u) << (3 * i + 2);
    }
    return m;
}

}  // namespace

extern "C" {

int nbl_tri_file_load(const char* path, nbl_triangle* out, uint32_t capacity, uint32_t* n)
{
    return guarded([&] {
        require(path && n, "nbl_tri_file_load: NULL argument");
        const std::vector<nbl_triangle> t = read_tri(path);
        require(t.size() <= 0xFFFFFFFFull, "nbl_tri_file_load: too many triangles");
        *n = (uint32_t)t.size();
        if (out) {
            require(capacity >= t.size(), "nbl_tri_file_load: capacity smaller than the triangle count");
            std::copy(t.begin(), t.end(), out);
        }
    });
}

int nbl_pri_file_load(const char* path, nbl_vec3 aabb_min, nbl_vec3 aabb_max, float max_energy, nbl_particle* out,
    nbl_int2* pixels, uint64_t capacity, uint64_t* n, nbl_pri_warnings* warnings)
{
    return guarded([&] {
        require(path && n, "nbl_pri_file_load: NULL argument");
        *n = 0;
        nbl_pri_warnings w{};
        std::ifstream f(path, std::ifstream::binary);
        if (!f.is_open()) {
            if (warnings) *warnings = w;
            return;                                   // as the reference: no primaries
        }
        f.seekg(0, std::ios::end);
        const uint64_t bytes = (uint64_t)f.tellg();
        f.seekg(0, std::ios::beg);
        require(bytes % PRI_BYTES == 0, "nbl_pri_file_load: unexpected end of primaries file (size is not a "
                                        "multiple of 36 bytes)");
        const uint64_t count = bytes / PRI_BYTES;
        *n = count;
        if (!out && !pixels && !warnings) return;
        require(!out || capacity >= count, "nbl_pri_file_load: capacity smaller than the primary count");
        // load_pri_file.cpp:104-154: read in chunks, count (but keep) primaries
        // outside the domain, with energy <= EPSILON or above max_energy, or
        // with a null direction
        std::vector<char> buf(PRI_BYTES * 4096);
        uint64_t done = 0;
        while (done < count) {
            const uint64_t k = std::min<uint64_t>(4096, count - done);
            f.read(buf.data(), (std::streamsize)(k * PRI_BYTES));
            require((uint64_t)f.gcount() == k * PRI_BYTES, "nbl_pri_file_load: read error");
            for (uint64_t i = 0; i < k; ++i) {
                float fl[7];
                int32_t px[2];
                std::memcpy(fl, buf.data() + i * PRI_BYTES, sizeof fl);
                std::memcpy(px, buf.data() + i * PRI_BYTES + sizeof fl, sizeof px);
                const nbl_particle p{ { fl[0], fl[1], fl[2] }, { fl[3], fl[4], fl[5] }, fl[6] };
                if (!(p.pos.x > aabb_min.x && p.pos.y > aabb_min.y && p.pos.z > aabb_min.z &&
                      p.pos.x < aabb_max.x && p.pos.y < aabb_max.y && p.pos.z < aabb_max.z))
                    ++w.outside_geometry;
                if (!(p.kin_energy > EPS)) ++w.low_energy;
                if (p.kin_energy > max_energy) ++w.high_energy;
                if (!(std::fabs(p.dir.x) > EPS || std::fabs(p.dir.y) > EPS || std::fabs(p.dir.z) > EPS))
                    ++w.bad_direction;
                if (out) out[done + i] = p;
                if (pixels) pixels[done + i] = { px[0], px[1] };
            }
            done += k;
        }
        if (warnings) *warnings = w;
    });
}

int nbl_pri_sort_morton(nbl_particle* p, nbl_int2* pixels, uint64_t n)
{
    return guarded([&] {
        require(n == 0 || (p && pixels), "nbl_pri_sort_morton: NULL argument");
        if (n == 0) return;
        // sort_pri_file, load_pri_file.cpp:189-229: positions scaled to the
        // primaries' bounding box and quantised to 16 bits per axis, in float
        nbl_vec3 lo = p[0].pos, hi = p[0].pos;
        for (uint64_t i = 0; i < n; ++i) {
            lo = { std::min(lo.x, p[i].pos.x), std::min(lo.y, p[i].pos.y), std::min(lo.z, p[i].pos.z) };
            hi = { std::max(hi.x, p[i].pos.x), std::max(hi.y, p[i].pos.y), std::max(hi.z, p[i].pos.z) };
        }
        const nbl_vec3 size{ hi.x - lo.x, hi.y - lo.y, hi.z - lo.z };
        const float q = (float)std::numeric_limits<uint16_t>::max();
        auto key = [&](const nbl_particle& a) {
            const float x = a.pos.x - lo.x, y = a.pos.y - lo.y, z = a.pos.z - lo.z;
            return morton3((uint16_t)(x / size.x * q), (uint16_t)(y / size.y * q), (uint16_t)(z / size.z * q));
        };
        // The reference sorts an index permutation with std::sort and this
        // comparison; doing the same keeps the order of equal keys identical
        // with the same standard library.
        std::vector<size_t> perm(n);
        std::iota(perm.begin(), perm.end(), 0);
        std::sort(perm.begin(), perm.end(), [&](size_t i, size_t j) { return key(p[i]) < key(p[j]); });
        std::vector<nbl_particle> sp(n);
        std::vector<nbl_int2> sx(n);
        for (uint64_t i = 0; i < n; ++i) { sp[i] = p[perm[i]]; sx[i] = pixels[perm[i]]; }
        std::copy(sp.begin(), sp.end(), p);
        std::copy(sx.begin(), sx.end(), pixels);
    });
}

int nbl_pri_prescan_shuffle(nbl_particle* p, nbl_int2* pixels, uint64_t n, uint64_t prescan_size)
{
    return guarded([&] {
        require(n == 0 || (p && pixels), "nbl_pri_prescan_shuffle: NULL argument");
        // prescan_shuffle, load_pri_file.cpp:231-243: a partial Fisher-Yates
        // shuffle of the first prescan_size places with a default-seeded
        // std::default_random_engine, so that the prescan sees a random sample
        std::default_random_engine rng;
        for (uint64_t i = 0; i < std::min<uint64_t>(prescan_size, n); ++i) {
            std::uniform_int_distribution<size_t> d((size_t)i, (size_t)(n - 1));
            const size_t j = d(rng);
            std::swap(p[i], p[j]);
            std::swap(pixels[i], pixels[j]);
        }
    });
}

}  // extern "C"
