// noalloc.hip — bz_cull_noalloc: bz_cull without any device allocation.
//
// The decide and emit kernels of a cull pass need compaction scratch (per-tile counts and per-tile staging slots).
// bz_cull takes it from a per-stream buffer that grows with hipMalloc; this entry point takes it from `scratch`, a
// file-scope device array of fixed size that every GPU holds from the moment this library's code object is loaded
// until the process ends. Callers that cannot allocate (code under graph capture, tight serving loops) use it.
//
// In host code `scratch` names a host-side stand-in, not the device array: the device address is looked up once per
// device with hipGetSymbolAddress (bz_create calls noalloc_scratch_address with the device current) and only that
// pointer is ever passed to kernels.
#include <mutex>

#include "bz_ctx.h"

namespace bz {

// counts[kNoAllocTiles] and their prefix[kNoAllocTiles + 1] (u32), then staging[kNoAllocEntries] (objectId, LOD)
// pairs. Every tile size the tunings allow (512, 1024, 2048 objects) and the cached LATE pass's 4096-object
// super-tiles fit BZ_NOALLOC_MAX_OBJECTS objects: at most BZ_NOALLOC_MAX_OBJECTS / 512 tiles, and the range rounded
// up to a whole tile or super-tile of staging slots.
constexpr uint32_t kNoAllocTiles = BZ_NOALLOC_MAX_OBJECTS / 512u;
constexpr uint64_t kNoAllocEntries = (uint64_t)BZ_NOALLOC_MAX_OBJECTS + kSuperTile;
constexpr uint64_t kNoAllocCountsBytes = (counts_words(kNoAllocTiles) * 4u + 255u) / 256u * 256u;
constexpr uint64_t kNoAllocBytes = kNoAllocCountsBytes + kNoAllocEntries * sizeof(uint2);
static_assert(kNoAllocBytes == 33652992ull, "the size bz_cull_noalloc's declaration states");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_address(int device, void** out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[device] = p;
    }
    *out = g_ptr[device];
    return hipSuccess;
}

}  // namespace bz

extern "C" bz_status bz_cull_noalloc(bz_ctx* ctx, const bz_cull_args* a, void* stream)
{
    using namespace bz;
    if (!ctx || !a || !a->view) return BZ_E_INVALID;
    if (!ctx->noalloc_scratch) return ctx->fail(BZ_E_STATE, "the fixed scratch of this device was not found");
    uint32_t count = a->object_count;
    if (a->object_begin == 0 && count == 0) {
        Scene* s = ctx->scene(a->scene);
        if (!s) return ctx->fail(BZ_E_INVALID, "unknown scene handle");
        count = s->n_objects;
    }
    if (count > BZ_NOALLOC_MAX_OBJECTS)
        return ctx->fail(BZ_E_INVALID, "bz_cull_noalloc culls at most BZ_NOALLOC_MAX_OBJECTS objects per call");
    unsigned char* base = static_cast<unsigned char*>(ctx->noalloc_scratch);
    CompactScratch fixed{};
    fixed.counts = reinterpret_cast<uint32_t*>(base);
    fixed.staging = reinterpret_cast<uint2*>(base + kNoAllocCountsBytes);
    fixed.capacity_tiles = kNoAllocTiles;
    fixed.capacity_entries = kNoAllocEntries;
    set_prefix(fixed);
    return ctx->cull(a, (hipStream_t)stream, &fixed);
}
