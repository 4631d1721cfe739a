// Binary hull-pool files: a b3amd_hull_pool (the reference's b3Hull arrays, hull.h:37-64,
// flattened per pool) written as one structure-of-arrays record so that a pool built once
// (b3amd_quickhull / b3amd_gen_hull_pool) is reloaded without rebuilding.
//
// Layout, little-endian:
//   u8[8] magic "B3AMDHUL", u32 version (1), u32 hull count H,
//   u64 vertex, half-edge and face totals (NV, NE, NF),
//   u32[H][6] desc, f32[H][3] centroids, f32[NV][3] vertices, u32[NE][4] half-edges
//   (origin, twin, face, next), u32[NF] face edges, f32[NF][4] planes (normal, offset),
//   u64 FNV-1a hash of everything before it.
// Reading checks the magic, version, sizes, hash and that every hull's ranges lie inside the
// totals; the records themselves are validated when the pool is uploaded.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "bounce_amd.h"

namespace {

constexpr char kMagic[8] = {'B', '3', 'A', 'M', 'D', 'H', 'U', 'L'};
constexpr uint32_t kVersion = 1;

struct Header {
    char magic[8];
    uint32_t version;
    uint32_t count;
    uint64_t nv, ne, nf;
};
static_assert(sizeof(Header) == 40, "hull file header layout");

uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001B3ull;
    return h;
}

struct Section {
    const void* data;
    size_t bytes;
};

// The payload sections of a pool with the given totals, in file order.
std::vector<Section> sections(const b3amd_hull_pool* p, uint64_t nv, uint64_t ne, uint64_t nf) {
    return {{p->desc, 24ull * p->count}, {p->centroids, 12ull * p->count}, {p->vertices, 12 * nv},
            {p->edges, 16 * ne},         {p->faces, 4 * nf},              {p->planes, 16 * nf}};
}

int read_header(FILE* f, Header* h) {
    if (std::fread(h, sizeof(*h), 1, f) != 1) return B3AMD_ERR_ARGUMENT;
    if (std::memcmp(h->magic, kMagic, 8) != 0 || h->version != kVersion) return B3AMD_ERR_UNSUPPORTED;
    return B3AMD_OK;
}

}  // namespace

extern "C" {

int b3amd_hull_file_write(const char* path, const b3amd_hull_pool* pool) {
    if (!path || !pool || (pool->count && (!pool->desc || !pool->centroids))) return B3AMD_ERR_ARGUMENT;
    uint64_t nv = 0, ne = 0, nf = 0;
    for (uint32_t h = 0; h < pool->count; ++h) {
        const uint32_t* d = pool->desc + 6 * (size_t)h;
        nv = nv > (uint64_t)d[0] + d[1] ? nv : (uint64_t)d[0] + d[1];
        ne = ne > (uint64_t)d[2] + d[3] ? ne : (uint64_t)d[2] + d[3];
        nf = nf > (uint64_t)d[4] + d[5] ? nf : (uint64_t)d[4] + d[5];
    }
    if ((nv && !pool->vertices) || (ne && !pool->edges) || (nf && (!pool->faces || !pool->planes)))
        return B3AMD_ERR_ARGUMENT;
    FILE* f = std::fopen(path, "wb");
    if (!f) return B3AMD_ERR_ARGUMENT;
    Header h;
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.count = pool->count;
    h.nv = nv, h.ne = ne, h.nf = nf;
    uint64_t hash = fnv1a(0xCBF29CE484222325ull, &h, sizeof(h));
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
    for (const Section& s : sections(pool, nv, ne, nf)) {
        if (!ok || s.bytes == 0) continue;
        ok = std::fwrite(s.data, 1, s.bytes, f) == s.bytes;
        hash = fnv1a(hash, s.data, s.bytes);
    }
    ok = ok && std::fwrite(&hash, 8, 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? B3AMD_OK : B3AMD_ERR_ARGUMENT;
}

int b3amd_hull_file_info(const char* path, uint32_t* count, uint64_t sizes[3]) {
    if (!path || !count || !sizes) return B3AMD_ERR_ARGUMENT;
    FILE* f = std::fopen(path, "rb");
    if (!f) return B3AMD_ERR_ARGUMENT;
    Header h;
    const int rc = read_header(f, &h);
    std::fclose(f);
    if (rc) return rc;
    *count = h.count;
    sizes[0] = h.nv, sizes[1] = h.ne, sizes[2] = h.nf;
    return B3AMD_OK;
}

int b3amd_hull_file_read(const char* path, uint32_t* desc, float* centroids, float* vertices, uint32_t* edges,
                         uint32_t* faces, float* planes) {
    if (!path) return B3AMD_ERR_ARGUMENT;
    FILE* f = std::fopen(path, "rb");
    if (!f) return B3AMD_ERR_ARGUMENT;
    Header h;
    int rc = read_header(f, &h);
    b3amd_hull_pool p{h.count, desc, centroids, vertices, edges, faces, planes};
    uint64_t hash = fnv1a(0xCBF29CE484222325ull, &h, sizeof(h));
    if (rc == B3AMD_OK) {
        for (const Section& s : sections(&p, h.nv, h.ne, h.nf)) {
            if (s.bytes == 0) continue;
            if (!s.data || std::fread(const_cast<void*>(s.data), 1, s.bytes, f) != s.bytes) {
                rc = B3AMD_ERR_ARGUMENT;
                break;
            }
            hash = fnv1a(hash, s.data, s.bytes);
        }
    }
    uint64_t stored = 0;
    if (rc == B3AMD_OK && (std::fread(&stored, 8, 1, f) != 1 || stored != hash)) rc = B3AMD_ERR_ARGUMENT;
    std::fclose(f);
    if (rc) return rc;
    for (uint32_t i = 0; i < h.count; ++i) {
        const uint32_t* d = desc + 6 * (size_t)i;
        if ((uint64_t)d[0] + d[1] > h.nv || (uint64_t)d[2] + d[3] > h.ne || (uint64_t)d[4] + d[5] > h.nf)
            return B3AMD_ERR_ARGUMENT;
    }
    return B3AMD_OK;
}

}  // extern "C"
