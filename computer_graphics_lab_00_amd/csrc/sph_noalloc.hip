// sph_noalloc.hip — one SPH substep on caller-owned device buffers that allocates nothing.
//
// The same kernels as sph_substep (grid build, density pass, force + integrate pass), run
// on state the caller keeps in device memory (AoS float4 in particle-index order, the
// layout of FluidSolver::particles()). Every other buffer the substep needs — the hash
// buckets, the spatial bins and their scan, the sort's scratch, the sorted state and the
// launch tables — lives in one file-scope device array sized for at most
// SPH_NOALLOC_MAX_PARTICLES particles and SPH_NOALLOC_MAX_CELLS hash buckets. The host
// addresses it through hipGetSymbolAddress, looked up once per device on the first call
// made with that device current. The launch tables are written by a kernel that takes
// them as arguments, so nothing is copied from host memory either. The force pass walks
// the 27 buckets (no neighbour list: the list would need 256 B per particle).
//
// Every call on one device shares that device's array: calls that overlap on a device
// (two streams, or two host threads) race, and must be serialised by the caller.
#include <string.h>

#include <mutex>

#include "sph_engine.h"
#include "sph_internal.h"

namespace {

using sphdev::kBlock;

constexpr uint32_t kMaxN = SPH_NOALLOC_MAX_PARTICLES;
constexpr uint32_t kMaxC = SPH_NOALLOC_MAX_CELLS;
constexpr uint32_t kMaxBins = 1u << 18;                 // spatial bins (coarsened beyond this)
constexpr uint32_t kMaxBlocks = kMaxN / kBlock;
constexpr uint32_t kMaxScanTiles = (kMaxBins + 2 + sphdev::kScanTile - 1) / sphdev::kScanTile;
constexpr int kMaxDevices = 64;

constexpr size_t up256(size_t x) { return (x + 255) & ~size_t(255); }

// Byte offsets of the regions of `scratch`, each 256-byte aligned.
constexpr size_t kOffDesc = 0;
constexpr size_t kOffBounds = kOffDesc + up256(sizeof(sphdev::SceneDesc));
constexpr size_t kOffBlockScene = kOffBounds + up256(sizeof(sphdev::BoundaryConsts));
constexpr size_t kOffIds = kOffBlockScene + up256((kMaxBlocks + 1) * sizeof(uint16_t));
constexpr size_t kOffIdsSorted = kOffIds + up256(kMaxN * 4ull);
constexpr size_t kOffKeys = kOffIdsSorted + up256(kMaxN * 4ull);
constexpr size_t kOffBinOf = kOffKeys + up256(kMaxN * 4ull);
constexpr size_t kOffSlots = kOffBinOf + up256(kMaxN * 4ull);
constexpr size_t kOffKeysSorted = kOffSlots + up256(kMaxN * 4ull);
constexpr size_t kOffPairSrc = kOffKeysSorted + up256(kMaxN * 4ull);
constexpr size_t kOffPairKI = kOffPairSrc + up256(kMaxN * 4ull);
constexpr size_t kOffBuckets = kOffPairKI + up256(kMaxN * 8ull);
constexpr size_t kOffRep = kOffBuckets + up256(kMaxC * 8ull);
constexpr size_t kOffBinStart = kOffRep + up256(kMaxC * 4ull);
constexpr size_t kOffBlockSums = kOffBinStart + up256((kMaxBins + 2) * 4ull);
constexpr size_t kOffPosSorted = kOffBlockSums + up256((kMaxScanTiles + 1) * 4ull);
constexpr size_t kOffVelSorted = kOffPosSorted + up256((kMaxN + sphdev::kPosPad) * 16ull);
constexpr size_t kOffPosNext = kOffVelSorted + up256(kMaxN * 16ull);
constexpr size_t kOffVelNext = kOffPosNext + up256(kMaxN * 16ull);
constexpr size_t kScratchBytes = kOffVelNext + up256(kMaxN * 16ull);
static_assert(kScratchBytes == SPH_NOALLOC_SCRATCH_BYTES, "the header states the scratch size");

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_scratchMutex;
uint8_t* g_scratch[kMaxDevices] = {};

// The device address of this device's `scratch`, looked up on the first call per device.
int scratch_on(int device, uint8_t** out) {
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratch[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return sphhost::hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return sphhost::fail(SPH_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned null");
        g_scratch[device] = static_cast<uint8_t*>(p);
    }
    *out = g_scratch[device];
    return SPH_OK;
}

// The launch tables of a one-scene substep, written from the kernel's arguments: the
// scene description, its colliders and fields, and the block -> scene table (all zero).
__global__ void __launch_bounds__(kBlock)
    k_noalloc_tables(sphdev::SceneDesc d, sphdev::BoundaryConsts b, uint32_t blocks, sphdev::SceneDesc* dd,
                     sphdev::BoundaryConsts* db, uint16_t* blockScene) {
    static_assert(sizeof(sphdev::SceneDesc) % 4 == 0 && sizeof(sphdev::BoundaryConsts) % 4 == 0, "word copies");
    const uint32_t* sd = reinterpret_cast<const uint32_t*>(&d);
    const uint32_t* sb = reinterpret_cast<const uint32_t*>(&b);
    uint32_t* od = reinterpret_cast<uint32_t*>(dd);
    uint32_t* ob = reinterpret_cast<uint32_t*>(db);
    for (uint32_t i = threadIdx.x; i < sizeof(d) / 4; i += kBlock) od[i] = sd[i];
    for (uint32_t i = threadIdx.x; i < sizeof(b) / 4; i += kBlock) ob[i] = sb[i];
    for (uint32_t i = threadIdx.x; i <= blocks; i += kBlock) blockScene[i] = 0;
}

bool is_pow2(uint32_t v) { return v != 0 && (v & (v - 1)) == 0; }

}  // namespace

// One FluidSolver::substep on caller device buffers; the allocation-free variant of
// sph_substep (declared in sph_c_api.h).
extern "C" int sph_substep_device(const SphParams* p, float dt, uint32_t modeFlags, uint32_t eos,
                                  const float* posIn, const float* velIn, float* posOut, float* velOut,
                                  uint32_t count, void* stream) {
    sphhost::clear_error();
    if (!p) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_substep_device: params is null");
    if (count > kMaxN)
        return sphhost::fail(SPH_ERR_CAPACITY, "sph_substep_device: count above SPH_NOALLOC_MAX_PARTICLES (" +
                                                   std::to_string(kMaxN) + ")");
    if (!is_pow2(p->cellCount)) return sphhost::fail(SPH_ERR_NOT_POWER_OF_TWO, "cellCount must be a power of two");
    if (p->cellCount > kMaxC)
        return sphhost::fail(SPH_ERR_CAPACITY, "sph_substep_device: cellCount above SPH_NOALLOC_MAX_CELLS (" +
                                                   std::to_string(kMaxC) + ")");
    if (!(p->smoothingRadius > 0.0f)) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "smoothingRadius must be > 0");
    if (modeFlags & ~(uint32_t)(SPH_MODE_UNCAPPED | SPH_MODE_DEDUPE | SPH_MODE_FAST_MATH))
        return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_substep_device: unknown mode flags");
    if (eos != SPH_EOS_LINEAR && eos != SPH_EOS_TAIT)
        return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_substep_device: unknown equation of state");
    if (count == 0) return SPH_OK;
    if (!posIn || !velIn || !posOut || !velOut)
        return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "sph_substep_device: null state buffer");
    int device = 0;
    SPH_HIP(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) return sphhost::fail(SPH_ERR_INVALID_ARGUMENT, "device ordinal too high");
    uint8_t* base = nullptr;
    if (int st = scratch_on(device, &base)) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);

    // The one scene's constants, exactly as a solver of `count` particles plans them.
    sphhost::SceneHost sh;
    sh.maxN = count;
    sh.maxC = p->cellCount;
    sh.binCap = kMaxBins;
    sh.n = count;
    sh.params = *p;
    sh.hasParams = true;
    sh.mode = modeFlags;
    sh.eos = eos;
    sh.list = false;
    const sphdev::BinGrid grid = sphhost::make_bins(sh, *p);
    if (grid.bins > kMaxBins) return sphhost::fail(SPH_ERR_CAPACITY, "sph_substep_device: spatial bins overflow");
    sphdev::SceneDesc d;
    memset(&d, 0, sizeof(d));
    d.k = sphhost::make_consts(sh, *p, dt, false);
    d.k.n = count;
    d.g = grid;
    d.offset = 0;
    d.capacity = count;
    d.cellOffset = 0;
    d.binOffset = 0;
    d.binCapacity = grid.bins;
    d.blockOffset = 0;
    d.active = 1;
    sphdev::BoundaryConsts bc;
    memset(&bc, 0, sizeof(bc));
    memcpy(bc.colliders, p->colliders, sizeof(bc.colliders));
    memcpy(bc.fields, p->fields, sizeof(bc.fields));
    const uint32_t blocks = (count + kBlock - 1) / kBlock;

    sphdev::SceneDesc* dDesc = reinterpret_cast<sphdev::SceneDesc*>(base + kOffDesc);
    sphdev::BoundaryConsts* dBounds = reinterpret_cast<sphdev::BoundaryConsts*>(base + kOffBounds);
    uint16_t* dBlockScene = reinterpret_cast<uint16_t*>(base + kOffBlockScene);
    k_noalloc_tables<<<1, kBlock, 0, s>>>(d, bc, blocks, dDesc, dBounds, dBlockScene);
    SPH_HIP(hipGetLastError());

    sphdev::Launch L;
    L.scenes = dDesc;
    L.bounds = dBounds;
    L.blockScene = dBlockScene;
    L.blocks = blocks;
    L.numScenes = 1;
    L.total = count;
    L.totalCells = p->cellCount;
    L.totalBins = grid.bins + 1;   // the scene's bins and its (empty) filler bin
    L.modes = modeFlags;

    uint32_t* ids = reinterpret_cast<uint32_t*>(base + kOffIds);
    sphdev::GridBuffers g;
    g.pos = reinterpret_cast<const float4*>(posIn);
    g.vel = reinterpret_cast<const float4*>(velIn);
    g.ids = ids;
    g.keys = reinterpret_cast<uint32_t*>(base + kOffKeys);
    g.buckets = reinterpret_cast<uint2*>(base + kOffBuckets);
    g.rep = reinterpret_cast<uint32_t*>(base + kOffRep);
    g.binStart = reinterpret_cast<uint32_t*>(base + kOffBinStart);
    g.binOf = reinterpret_cast<uint32_t*>(base + kOffBinOf);
    g.slots = reinterpret_cast<uint32_t*>(base + kOffSlots);
    g.pairKI = reinterpret_cast<uint2*>(base + kOffPairKI);
    g.pairSrc = reinterpret_cast<uint32_t*>(base + kOffPairSrc);
    g.blockSums = reinterpret_cast<uint32_t*>(base + kOffBlockSums);
    g.posSorted = reinterpret_cast<float4*>(base + kOffPosSorted);
    g.velSorted = reinterpret_cast<float4*>(base + kOffVelSorted);
    g.idsSorted = reinterpret_cast<uint32_t*>(base + kOffIdsSorted);
    g.keysSorted = reinterpret_cast<uint32_t*>(base + kOffKeysSorted);
    float4* posNext = reinterpret_cast<float4*>(base + kOffPosNext);
    float4* velNext = reinterpret_cast<float4*>(base + kOffVelNext);

    // The state arrives in index order: slot i holds particle i.
    SPH_HIP(sphdev::launch_iota(ids, count, s));
    // FluidGrid::build (counts, the bins and their scan start from the clear kernel)
    SPH_HIP(sphdev::launch_grid_clear(L, g, s));
    SPH_HIP(sphdev::launch_keys(L, g, s));
    SPH_HIP(sphdev::launch_bin_slots(L, g, s));
    SPH_HIP(sphdev::launch_scan(g.binStart, L.totalBins, count, g.blockSums, s));
    SPH_HIP(sphdev::launch_scatter(L, g, s));
    SPH_HIP(sphdev::launch_rank_reorder(L, g, s));
    // density + EOS in place, then force + integrate + boundaries into the next state
    SPH_HIP(sphdev::launch_density(L, g.posSorted, g.velSorted, g.buckets, nullptr, nullptr, nullptr, s));
    SPH_HIP(sphdev::launch_force(L, g.posSorted, g.velSorted, g.buckets, nullptr, nullptr, posNext,
                                 velNext, nullptr, s));
    // back to index order, into the caller's buffers
    SPH_HIP(sphdev::launch_unpermute(posNext, g.idsSorted, reinterpret_cast<float4*>(posOut), count, s));
    SPH_HIP(sphdev::launch_unpermute(velNext, g.idsSorted, reinterpret_cast<float4*>(velOut), count, s));
    return SPH_OK;
}
