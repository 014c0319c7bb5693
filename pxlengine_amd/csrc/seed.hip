// Initial particles seeded on the device: the seeding loop of FluidState::new
// (crates/pxl-render/src/fluid.rs:96-149) without the host loop and the upload.
//
// The reference walks the interior cells of each fill block in k, j, i order and draws three
// LCG numbers per particle from one stream that continues across blocks.  Here one thread
// takes one cell of a block's seed range (the cells whose centre can pass the block's test):
//   flags:  1 where the cell passes the block's test and is no obstacle cell
//   scan:   exclusive scan of the flags = the cell's rank among the block's seeded cells
//   write:  particle p = (particles of earlier blocks) + rank * ppc + q draws numbers
//           3p, 3p + 1, 3p + 2 of the stream, reached by jumping the LCG ahead 3p steps
//           (the affine map s -> a s + c composed with itself by squaring)
// Every float is computed with the host loop's f32 operations in its order, so the particles
// are bitwise the host's.
#include "common.hpp"
#include "kernels.hpp"
#include "sim.hpp"
#include "spec.hpp"

namespace pxl {
namespace {

constexpr uint64_t LCG_MUL = 6364136223846793005ull;
constexpr uint64_t LCG_INC = 1442695040888963407ull;

struct SeedBlock {
    float min[3], max[3], center[3], vel[3];
    float radius;
    int sphere;
    uint32_t lo[3], ext[3];  // seed range: first index and extent per axis
    uint32_t ppc;
    uint64_t first;  // particles seeded by the earlier blocks
    uint64_t rng0;   // 0x9e3779b97f4a7c15 ^ seed (fluid.rs:120)
};

// The LCG state after m more draws.
__device__ __forceinline__ uint64_t lcg_skip(uint64_t s, uint64_t m) {
    uint64_t A = 1, C = 0, a = LCG_MUL, c = LCG_INC;
    while (m) {
        if (m & 1) {
            A = a * A;
            C = a * C + c;
        }
        c = (a + 1) * c;
        a = a * a;
        m >>= 1;
    }
    return A * s + C;
}

__device__ __forceinline__ float lcg_draw(uint64_t &s) {
    s = s * LCG_MUL + LCG_INC;
    return (float)(s >> 40) / (float)(1ull << 24);
}

__device__ __forceinline__ void range_cell(const SeedBlock &B, uint32_t e, uint32_t &i, uint32_t &j, uint32_t &k) {
    i = B.lo[0] + e % B.ext[0];
    j = B.lo[1] + (e / B.ext[0]) % B.ext[1];
    k = B.lo[2] + e / (B.ext[0] * B.ext[1]);
}

__global__ __launch_bounds__(BLOCK) void k_seed_flags(const DevSim S, const SeedBlock B, uint32_t ncell,
                                                      uint32_t *flags) {
    const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= ncell) return;
    uint32_t i, j, k;
    range_cell(B, e, i, j, k);
    const float cx = S.ox + ((float)i + 0.5f) * S.h;
    const float cy = S.oy + ((float)j + 0.5f) * S.h;
    const float cz = S.oz + ((float)k + 0.5f) * S.h;
    bool in;
    if (B.sphere) {
        const float dx = cx - B.center[0], dy = cy - B.center[1], dz = cz - B.center[2];
        const float d2 = dx * dx + dy * dy + dz * dz;
        in = d2 <= B.radius * B.radius;
    } else {
        in = !(cx < B.min[0] || cx > B.max[0] || cy < B.min[1] || cy > B.max[1] || cz < B.min[2] || cz > B.max[2]);
    }
    if (in && S.n_obst && S.obst_id[i + (uint32_t)S.nx * (j + (uint32_t)S.ny * k)]) in = false;
    flags[e] = in ? 1u : 0u;
}

__global__ __launch_bounds__(BLOCK) void k_seed_write(const DevSim S, const SeedBlock B, uint32_t ncell,
                                                      const uint32_t *flags, const uint32_t *rank) {
    const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= ncell || !flags[e]) return;
    uint32_t i, j, k;
    range_cell(B, e, i, j, k);
    const uint64_t p0 = B.first + (uint64_t)rank[e] * B.ppc;
    uint64_t s = lcg_skip(B.rng0, 3 * p0);
    for (uint32_t q = 0; q < B.ppc && p0 + q < S.np; q++) {
        const size_t p = (size_t)(p0 + q);
        const float px = S.ox + ((float)i + lcg_draw(s)) * S.h;
        const float py = S.oy + ((float)j + lcg_draw(s)) * S.h;
        const float pz = S.oz + ((float)k + lcg_draw(s)) * S.h;
        S.ix[p] = px;
        S.iy[p] = py;
        S.iz[p] = pz;
        S.ivx[p] = B.vel[0];
        S.ivy[p] = B.vel[1];
        S.ivz[p] = B.vel[2];
    }
}

}  // namespace

void seed_on_device(Sim &s, uint32_t *flags, uint32_t *rank) {
    const FluidSpec &f = s.spec;
    const DevSim &d = s.d;
    std::vector<uint32_t> obst;
    obstacle_cells(f, obst);
    const std::vector<uint64_t> counts = seed_counts(f, obst);
    uint32_t n[3] = {(uint32_t)d.nx, (uint32_t)d.ny, (uint32_t)d.nz};
    const float org[3] = {d.ox, d.oy, d.oz};
    uint64_t first = 0;
    uint32_t *tail = reinterpret_cast<uint32_t *>(s.host_scal);  // pinned: last flag and rank of a range
    for (size_t bi = 0; bi < f.fill.size(); bi++) {
        const FluidBlockSpec &b = f.fill[bi];
        const SeedRange r = seed_range(b, n, d.h, org);
        const uint64_t ncell = r.cells();
        if (counts[bi] && ncell) {
            if (ncell > (uint64_t)d.ncells) throw std::logic_error("seed range larger than the grid");
            SeedBlock B{};
            for (int a = 0; a < 3; a++) {
                B.min[a] = b.min[a];
                B.max[a] = b.max[a];
                B.center[a] = b.center[a];
                B.vel[a] = b.velocity[a];
                B.lo[a] = r.lo[a];
                B.ext[a] = r.hi[a] - r.lo[a];
            }
            B.radius = b.radius;
            B.sphere = b.sphere ? 1 : 0;
            B.ppc = f.particles_per_cell;
            B.first = first;
            B.rng0 = 0x9e3779b97f4a7c15ull ^ (uint64_t)s.seed;
            const uint32_t nc = (uint32_t)ncell, grid = (uint32_t)((ncell + BLOCK - 1) / BLOCK);
            PXL_LAUNCH(k_seed_flags, grid, BLOCK, 0, s.stream, d, B, nc, flags);
            launch_exclusive_scan(flags, nc, d.scan_tmp, rank, s.stream);
            // the device's count of seeded cells must be the host's (which sized the arrays)
            // before anything is written
            HIP_CHECK(hipMemcpyAsync(tail, flags + nc - 1, 4, hipMemcpyDeviceToHost, s.stream));
            HIP_CHECK(hipMemcpyAsync(tail + 1, rank + nc - 1, 4, hipMemcpyDeviceToHost, s.stream));
            HIP_CHECK(hipStreamSynchronize(s.stream));
            if ((uint64_t)(tail[0] + tail[1]) * f.particles_per_cell != counts[bi])
                throw std::logic_error("device seeding counted " +
                                       std::to_string((uint64_t)(tail[0] + tail[1]) * f.particles_per_cell) +
                                       " particles for fill block " + std::to_string(bi) + ", the host " +
                                       std::to_string(counts[bi]));
            PXL_LAUNCH(k_seed_write, grid, BLOCK, 0, s.stream, d, B, nc, flags, rank);
        }
        first += counts[bi];
    }
    if (first != d.np) throw std::logic_error("seeded particle count differs from the simulation's");
}

}  // namespace pxl
