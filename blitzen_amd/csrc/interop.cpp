// interop.cpp — buffers and semaphores shared with a Vulkan renderer on Linux.
//
// The reference renderer owns the cull outputs as Vulkan buffers: the indirect command buffer and the draw count
// that vkCmdDrawIndexedIndirectCount reads (vulkanDraw.cpp:470-471), and the depth attachment the pyramid is built
// from (vulkanDraw.cpp:554-575). Exported with VK_KHR_external_memory_fd they arrive here as file descriptors
// (dma-buf on amdgpu), are imported with hipImportExternalMemory and mapped to device pointers that bz_cull,
// bz_pyramid_build and the batched calls take like any other. Exported semaphores (VK_KHR_external_semaphore_fd)
// order the two APIs' queues on a HIP stream.
//
// Descriptor ownership: the caller's descriptor is never consumed. The library imports a duplicate; if the runtime
// has not taken that duplicate over when the import returns, the library closes it after the release.
#include <fcntl.h>
#include <unistd.h>

#include "bz_ctx.h"
#include "../../include/bzcull.h"

namespace bz {

namespace {

bool fd_open(int fd) { return fd >= 0 && fcntl(fd, F_GETFD) != -1; }

}  // namespace

ExtMem::~ExtMem()
{
    for (void* p : maps) (void)hipFree(p);
    if (mem) (void)hipDestroyExternalMemory(mem);
    if (fd >= 0) close(fd);
}

ExtSem::~ExtSem()
{
    if (sem) (void)hipDestroyExternalSemaphore(sem);
    if (fd >= 0) close(fd);
}

}  // namespace bz

using namespace bz;

extern "C" {

bz_status bz_extmem_import(bz_ctx* ctx, int fd, uint64_t size, bz_extmem* out)
{
    if (!ctx || !out) return BZ_E_INVALID;
    *out = 0;
    if (fd < 0 || !fd_open(fd)) return ctx->fail(BZ_E_INVALID, "not an open file descriptor");
    if (!size) return ctx->fail(BZ_E_INVALID, "size must be > 0");
    const int dup_fd = fcntl(fd, F_DUPFD_CLOEXEC, 0);
    if (dup_fd < 0) return ctx->fail(BZ_E_INVALID, "cannot duplicate the file descriptor");
    auto m = std::make_unique<ExtMem>();
    m->size = size;
    hipExternalMemoryHandleDesc d;
    std::memset(&d, 0, sizeof d);
    d.type = hipExternalMemoryHandleTypeOpaqueFd;
    d.handle.fd = dup_fd;
    d.size = size;
    hipError_t e = hipImportExternalMemory(&m->mem, &d);
    if (fd_open(dup_fd)) m->fd = dup_fd;  // not taken over by the runtime: closed after the release
    if (e != hipSuccess) {
        m->mem = nullptr;
        return ctx->hip_fail(e, "hipImportExternalMemory");
    }
    const bz_extmem h = ++ctx->next_handle;
    ctx->extmems[h] = std::move(m);
    *out = h;
    return BZ_OK;
}

bz_status bz_extmem_map(bz_ctx* ctx, bz_extmem mem, uint64_t offset, uint64_t size, void** dev_ptr)
{
    if (!ctx || !dev_ptr) return BZ_E_INVALID;
    *dev_ptr = nullptr;
    auto it = ctx->extmems.find(mem);
    if (it == ctx->extmems.end()) return ctx->fail(BZ_E_INVALID, "unknown external memory handle");
    ExtMem& m = *it->second;
    if (!size || offset > m.size || size > m.size - offset)
        return ctx->fail(BZ_E_INVALID, "mapped range outside the imported memory");
    hipExternalMemoryBufferDesc d;
    std::memset(&d, 0, sizeof d);
    d.offset = offset;
    d.size = size;
    void* p = nullptr;
    hipError_t e = hipExternalMemoryGetMappedBuffer(&p, m.mem, &d);
    if (e != hipSuccess) return ctx->hip_fail(e, "hipExternalMemoryGetMappedBuffer");
    m.maps.push_back(p);
    *dev_ptr = p;
    return BZ_OK;
}

bz_status bz_extmem_release(bz_ctx* ctx, bz_extmem mem)
{
    if (!ctx) return BZ_E_INVALID;
    auto it = ctx->extmems.find(mem);
    if (it == ctx->extmems.end()) return ctx->fail(BZ_E_INVALID, "unknown external memory handle");
    (void)hipDeviceSynchronize();  // no launch may still use a mapped pointer
    ctx->extmems.erase(it);
    return BZ_OK;
}

bz_status bz_extsem_import(bz_ctx* ctx, int fd, uint32_t kind, bz_extsem* out)
{
    if (!ctx || !out) return BZ_E_INVALID;
    *out = 0;
    if (kind != BZ_EXTSEM_BINARY && kind != BZ_EXTSEM_TIMELINE) return ctx->fail(BZ_E_INVALID, "unknown semaphore kind");
    if (fd < 0 || !fd_open(fd)) return ctx->fail(BZ_E_INVALID, "not an open file descriptor");
    const int dup_fd = fcntl(fd, F_DUPFD_CLOEXEC, 0);
    if (dup_fd < 0) return ctx->fail(BZ_E_INVALID, "cannot duplicate the file descriptor");
    auto s = std::make_unique<ExtSem>();
    s->kind = kind;
    hipExternalSemaphoreHandleDesc d;
    std::memset(&d, 0, sizeof d);
    d.type = kind == BZ_EXTSEM_TIMELINE ? hipExternalSemaphoreHandleTypeTimelineSemaphoreFd
                                        : hipExternalSemaphoreHandleTypeOpaqueFd;
    d.handle.fd = dup_fd;
    hipError_t e = hipImportExternalSemaphore(&s->sem, &d);
    if (fd_open(dup_fd)) s->fd = dup_fd;
    if (e != hipSuccess) {
        s->sem = nullptr;
        return ctx->hip_fail(e, "hipImportExternalSemaphore");
    }
    const bz_extsem h = ++ctx->next_handle;
    ctx->extsems[h] = std::move(s);
    *out = h;
    return BZ_OK;
}

bz_status bz_extsem_wait(bz_ctx* ctx, bz_extsem sem, uint64_t value, void* stream)
{
    if (!ctx) return BZ_E_INVALID;
    auto it = ctx->extsems.find(sem);
    if (it == ctx->extsems.end()) return ctx->fail(BZ_E_INVALID, "unknown external semaphore handle");
    hipExternalSemaphoreWaitParams p;
    std::memset(&p, 0, sizeof p);
    p.params.fence.value = value;  // timeline value; ignored for a binary semaphore
    hipError_t e = hipWaitExternalSemaphoresAsync(&it->second->sem, &p, 1, (hipStream_t)stream);
    return e == hipSuccess ? BZ_OK : ctx->hip_fail(e, "hipWaitExternalSemaphoresAsync");
}

bz_status bz_extsem_signal(bz_ctx* ctx, bz_extsem sem, uint64_t value, void* stream)
{
    if (!ctx) return BZ_E_INVALID;
    auto it = ctx->extsems.find(sem);
    if (it == ctx->extsems.end()) return ctx->fail(BZ_E_INVALID, "unknown external semaphore handle");
    hipExternalSemaphoreSignalParams p;
    std::memset(&p, 0, sizeof p);
    p.params.fence.value = value;
    hipError_t e = hipSignalExternalSemaphoresAsync(&it->second->sem, &p, 1, (hipStream_t)stream);
    return e == hipSuccess ? BZ_OK : ctx->hip_fail(e, "hipSignalExternalSemaphoresAsync");
}

bz_status bz_extsem_release(bz_ctx* ctx, bz_extsem sem)
{
    if (!ctx) return BZ_E_INVALID;
    auto it = ctx->extsems.find(sem);
    if (it == ctx->extsems.end()) return ctx->fail(BZ_E_INVALID, "unknown external semaphore handle");
    (void)hipDeviceSynchronize();
    ctx->extsems.erase(it);
    return BZ_OK;
}

}  // extern "C"
