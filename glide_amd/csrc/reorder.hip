// Particle reordering for gather locality.
//
// The fused step kernel gathers its met records per particle; when the 64
// lanes of a wavefront hold particles in unrelated cells every lane touches its
// own cache lines and the gathers become L2-bandwidth bound. Sorting the
// particle state by met cell puts each wavefront's particles in a few
// neighbouring cells, so the corner loads coalesce. The key is column-major
// (row, column, then level): particles mix vertically much faster than they
// move horizontally, so a column key keeps its locality longer and a re-sort
// moves most particles only a short way. Particles that can never be active
// again (escaped, or past their backward window) sort to the end.
//
// With several releases (n_releases > 1) the key is release-major: a
// release's particles stay together, sorted by cell, so a wavefront's
// particles share a release and a few cells and the fused fixed-point
// footprint sums long runs of one key per atomic (releases interleaved in a
// column would change the key from lane to lane).
//
// Results do not depend on the order: noise is keyed by the particle id, which
// moves with the particle.
#include "sort.h"

namespace glide {

struct ReorderKeyArgs {
    GlideParticles s;
    float lon0, lat0, sx, sy;
    int nx, ny, nz;
    const float* levels;
    float cursor_off, length_s;
    float horizon;  // particles with r_off < horizon become active only after it: second bucket
    bool terminus;  // GLIDE_F_RELEASE_TERMINUS: lo_off holds the release terminus
    uint32_t ncell, sentinel;
    uint32_t groups;  // release groups of the key (1: by cell only)
    bool by_level;    // the cell includes the level (else the column alone)
    uint32_t* keys;
    uint32_t* idx;
    uint32_t* live;  // optional count of particles that can still become active
};

constexpr int kMaxLev = 128;

// Key of particle i; returns whether it counts as live (active before the horizon).
__device__ __forceinline__ bool key_one(const ReorderKeyArgs& a, const float* lev, int64_t i) {
    const GlideParticles& s = a.s;
    // can never be active again: the cursor only decreases
    bool over;
    if (a.terminus) {
        over = s.lo_off[i] >= a.cursor_off;  // active needs cursor > te
    } else {
        const float lo = s.lo_off ? s.lo_off[i] : s.r_off[i] - a.length_s;
        over = lo > a.cursor_off;
    }
    const bool done = (s.alive[i] == 0) || over;
    // not active before the cursor passes the horizon (released further back)
    const bool later = !done && s.r_off[i] < a.horizon;
    uint32_t key = a.sentinel;
    if (!done) {
        const int64_t ps = s.pos_stride;
        int ix, iy;
        if (s.cell != nullptr) {  // a grid point plus offsets
            const int32_t c = s.cell[i];
            const int pcy = c >> 16, pcx = c & 0xffff;
            ix = static_cast<int>(cell_index(pcx, s.lon[i], a.sx, a.nx));
            iy = static_cast<int>(cell_index(pcy, s.lat[i], a.sy, a.ny));
        } else {
            ix = static_cast<int>(frac_index(s.lon[i * ps], a.lon0, a.sx, a.nx));
            iy = static_cast<int>(frac_index(s.lat[i * ps], a.lat0, a.sy, a.ny));
        }
        const int iz = a.by_level ? static_cast<int>(level_index(lev, a.nz, ladder_top(a.nz), s.z[i * ps])) : 0;
        uint32_t grp = 0;
        if (a.groups > 1) {
            const int32_t r = s.rel_idx[i];
            grp = static_cast<uint32_t>(r < 0 ? 0 : (r >= static_cast<int32_t>(a.groups) ? a.groups - 1 : r));
        }
        const uint32_t cell = a.by_level ? static_cast<uint32_t>((static_cast<int64_t>(iy) * a.nx + ix) * a.nz + iz)
                                         : static_cast<uint32_t>(static_cast<int64_t>(iy) * a.nx + ix);
        key = ((later ? a.groups : 0u) + grp) * a.ncell + cell;
    }
    a.keys[i] = key;
    a.idx[i] = static_cast<uint32_t>(i);
    return !done && !later;
}

constexpr int kKeyThreads = 1024;
constexpr int kKeyMaxBlocks = 512;  // bounded grid: few live-count atomics on the one counter

// Grid-stride over the particles. The live count goes through LDS and one
// integer atomic per workgroup: atomics from every workgroup of a launch over
// all particles queue on the single counter's address (≈10 ns each).
__global__ __launch_bounds__(kKeyThreads) void reorder_key_kernel(const ReorderKeyArgs a) {
    __shared__ float lev[kLadder];
    __shared__ uint32_t blive;
    fill_ladder(lev, a.levels, a.nz);
    if (threadIdx.x == 0) blive = 0;
    __syncthreads();
    uint32_t c = 0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.s.n; i += stride)
        c += key_one(a, lev, i) ? 1u : 0u;
    if (a.live != nullptr) {
        if (c) atomicAdd(&blive, c);
        __syncthreads();
        if (threadIdx.x == 0 && blive) atomicAdd(a.live, blive);
    }
}

struct PermuteArgs {
    GlideParticles src, dst;
    const uint32_t* idx;
};

template <typename T>
__device__ __forceinline__ void mv(const T* s, T* d, int64_t j, int64_t i) {
    if (s != nullptr && d != nullptr) d[j] = s[i];
}

__global__ __launch_bounds__(256) void permute_kernel(const PermuteArgs a) {
    // contiguous destinations per XCD: their sources, mostly nearby, stay in one L2
    const int64_t j = static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    if (j >= a.src.n) return;
    const int64_t i = a.idx[j];
    const GlideParticles& s = a.src;
    const GlideParticles& d = a.dst;
    mv(s.lon, d.lon, j, i);
    mv(s.lat, d.lat, j, i);
    mv(s.z, d.z, j, i);
    mv(s.up, d.up, j, i);
    mv(s.vp, d.vp, j, i);
    mv(s.wp, d.wp, j, i);
    mv(s.um, d.um, j, i);
    mv(s.vm, d.vm, j, i);
    mv(s.weight, const_cast<float*>(d.weight), j, i);
    mv(s.r_off, const_cast<float*>(d.r_off), j, i);
    mv(s.we_off, const_cast<float*>(d.we_off), j, i);
    mv(s.rel_idx, const_cast<int32_t*>(d.rel_idx), j, i);
    mv(s.uid, const_cast<uint32_t*>(d.uid), j, i);
    mv(s.alive, d.alive, j, i);
    mv(s.lo_off, const_cast<float*>(d.lo_off), j, i);
    mv(s.cell, d.cell, j, i);
}

}  // namespace glide

using namespace glide;

extern "C" int glide_reorder_keys(const GlideStepParams* p, const GlideParticles* s, const GlideGrid* g,
                                  double horizon_off, int32_t n_releases, uint32_t* keys, uint32_t* idx,
                                  uint32_t* live, void* stream) {
    GLIDE_REQUIRE(p && s && g && keys && idx, "glide_reorder_keys: NULL argument");
    GLIDE_REQUIRE(s->lon && s->lat && s->z && s->alive && s->r_off, "particle arrays are NULL");
    GLIDE_REQUIRE(s->pos_stride >= 1 && s->n >= 0, "bad particles");
    GLIDE_REQUIRE(g->nz >= 2 && g->nz <= kMaxLev && g->nx >= 2 && g->ny >= 2, "bad grid");
    GLIDE_REQUIRE(g->lon1 != g->lon0 && g->lat1 != g->lat0, "Grid dimension has 0 span (first == last)");
    GLIDE_REQUIRE(s->cell == nullptr || (s->pos_stride == 1 && g->nx <= 65535 && g->ny <= 65535),
                  "cell positions need pos_stride 1 and nx, ny <= 65535");
    if (s->n == 0) return GLIDE_OK;
    ReorderKeyArgs a{};
    a.s = *s;
    a.nx = g->nx;
    a.ny = g->ny;
    a.nz = g->nz;
    a.lon0 = static_cast<float>(g->lon0);
    a.lat0 = static_cast<float>(g->lat0);
    a.sx = static_cast<float>((g->nx - 1) / (g->lon1 - g->lon0));
    a.sy = static_cast<float>((g->ny - 1) / (g->lat1 - g->lat0));
    a.levels = g->levels;
    a.cursor_off = static_cast<float>(p->cursor_off);
    a.length_s = static_cast<float>(p->length_s);
    a.terminus = (p->flags & GLIDE_F_RELEASE_TERMINUS) != 0;
    GLIDE_REQUIRE(!a.terminus || s->lo_off, "GLIDE_F_RELEASE_TERMINUS needs lo_off");
    const int64_t groups = n_releases > 1 ? n_releases : 1;
    GLIDE_REQUIRE(groups == 1 || s->rel_idx, "release-major keys need rel_idx");
    // (column, level) cells while the key fits 32 bits, else columns alone
    a.by_level = 2 * groups * static_cast<int64_t>(g->nz) * g->ny * g->nx < 0xffffffffll;
    const int64_t ncell = a.by_level ? static_cast<int64_t>(g->nz) * g->ny * g->nx : static_cast<int64_t>(g->ny) * g->nx;
    GLIDE_REQUIRE(2 * groups * ncell < 0xffffffffll, "met grid x releases too large for the reorder keys");
    a.ncell = static_cast<uint32_t>(ncell);
    a.groups = static_cast<uint32_t>(groups);
    a.sentinel = static_cast<uint32_t>(2 * groups * ncell);
    a.horizon = static_cast<float>(horizon_off);
    a.keys = keys;
    a.idx = idx;
    a.live = live;
    int64_t blocks = (s->n + kKeyThreads - 1) / kKeyThreads;
    if (blocks > kKeyMaxBlocks) blocks = kKeyMaxBlocks;
    hipLaunchKernelGGL(reorder_key_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kKeyThreads), 0,
                       static_cast<hipStream_t>(stream), a);
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}

extern "C" int glide_permute_particles(const GlideParticles* src, const GlideParticles* dst, const uint32_t* idx,
                                       void* stream) {
    GLIDE_REQUIRE(src && dst && idx, "glide_permute_particles: NULL argument");
    GLIDE_REQUIRE(src->n == dst->n && src->n >= 0, "source and destination sizes differ");
    GLIDE_REQUIRE(src->pos_stride == 1 && dst->pos_stride == 1, "permutation needs structure-of-arrays positions");
    GLIDE_REQUIRE(src->lon != dst->lon, "permutation cannot run in place");
    if (src->n == 0) return GLIDE_OK;
    PermuteArgs a{*src, *dst, idx};
    hipLaunchKernelGGL(permute_kernel, dim3(static_cast<unsigned>((src->n + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), a);
    GLIDE_HIP(hipGetLastError());
    return GLIDE_OK;
}
