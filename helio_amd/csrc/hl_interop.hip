// Memory of another graphics or compute API on the same GPU, imported as HIP
// external memory and mapped as device pointers for the *_d arguments of the
// other entry points.  In the reference the frame resources this path reads
// and writes are wgpu buffers and textures (tile_light_lists,
// tile_light_counts and pre_aa: crates/libhelio/src/frame.rs:228-234) handed
// to the pass by RenderGraph::replace_pass_at
// (crates/helio-core/src/graph/execution.rs:193).  Under wgpu's Vulkan
// backend such a resource's VkDeviceMemory, allocated exportable, yields a
// file descriptor (vkGetMemoryFdKHR, opaque or dma-buf); that descriptor is
// what hl_external_import takes.
//
// Ownership: the caller keeps its descriptor.  The library imports its own
// duplicate (F_DUPFD_CLOEXEC) and closes it when the import is released:
// hipDestroyExternalMemory leaves the imported descriptor open on ROCm 7.
// The import holds a reference to the memory, so the exporter may free its
// own handle while the import lives.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstring>

#include "hl_internal.hpp"

using namespace hl;

struct hl_external {
    hl_context* ctx = nullptr;
    hipExternalMemory_t mem = nullptr;
    int fd = -1;        // the library's duplicate of the caller's descriptor
    uint64_t size = 0;  // bytes imported
};

namespace hl {

hl_status release_external(hl_external* e) {
    hl_context* c = e->ctx;
    hl_status st = HL_OK;
    if (e->mem) {
        const hipError_t r = hipDestroyExternalMemory(e->mem);
        if (r != hipSuccess) st = hip_fail(c, r, "hipDestroyExternalMemory");
    }
    if (e->fd >= 0) (void)close(e->fd);
    auto& v = c->externals;
    v.erase(std::remove(v.begin(), v.end(), e), v.end());
    delete e;
    return st;
}

}  // namespace hl

extern "C" {

HL_API hl_status hl_external_import(hl_context* c, int fd, uint64_t size, uint32_t kind, hl_external** out) {
    if (!c) return HL_EINVAL;
    if (!out) return fail(c, HL_EINVAL, "out is null");
    *out = nullptr;
    if (kind != HL_EXTERNAL_FD)
        return fail(c, HL_EUNSUPPORTED, "external memory kind %u (only HL_EXTERNAL_FD is supported)", kind);
    if (fd < 0) return fail(c, HL_EINVAL, "descriptor %d is negative", fd);
    if (size == 0) return fail(c, HL_EINVAL, "external memory of 0 bytes");
    const int dup = fcntl(fd, F_DUPFD_CLOEXEC, 0);
    if (dup < 0) return fail(c, HL_EINVAL, "descriptor %d: %s", fd, strerror(errno));
    // a dma-buf reports its size through lseek; an import larger than the
    // memory would map pages that do not exist
    const off_t have = lseek(dup, 0, SEEK_END);
    if (have < 0 || lseek(dup, 0, SEEK_SET) < 0) {
        const int err = errno;
        (void)close(dup);
        return fail(c, HL_EINVAL, "descriptor %d has no size (not a dma-buf?): %s", fd, strerror(err));
    }
    if ((uint64_t)have < size) {
        (void)close(dup);
        return fail(c, HL_EINVAL, "import of %llu bytes from a descriptor of %llu bytes",
                    (unsigned long long)size, (unsigned long long)have);
    }
    hipError_t r = hipSetDevice(c->device);
    if (r != hipSuccess) {
        (void)close(dup);
        return hip_fail(c, r, "hipSetDevice");
    }
    hipExternalMemoryHandleDesc d;
    memset(&d, 0, sizeof d);
    d.type = hipExternalMemoryHandleTypeOpaqueFd;
    d.handle.fd = dup;
    d.size = size;
    hl_external* e = new hl_external();
    e->ctx = c;
    e->fd = dup;
    e->size = size;
    r = hipImportExternalMemory(&e->mem, &d);
    if (r != hipSuccess) {
        e->mem = nullptr;
        (void)close(dup);
        delete e;
        return hip_fail(c, r, "hipImportExternalMemory");
    }
    c->externals.push_back(e);
    *out = e;
    return HL_OK;
}

HL_API hl_status hl_external_map(hl_external* e, uint64_t offset, uint64_t size, void** out_d) {
    if (!e) return HL_EINVAL;
    hl_context* c = e->ctx;
    if (!out_d) return fail(c, HL_EINVAL, "out_d is null");
    *out_d = nullptr;
    if (size == 0 || offset > e->size || size > e->size - offset)
        return fail(c, HL_EINVAL, "range [%llu, +%llu) outside the %llu imported bytes", (unsigned long long)offset,
                    (unsigned long long)size, (unsigned long long)e->size);
    if (offset % HL_EXTERNAL_MAP_ALIGN)
        return fail(c, HL_EINVAL, "offset %llu is not a multiple of %u", (unsigned long long)offset,
                    (unsigned)HL_EXTERNAL_MAP_ALIGN);
    HL_HIP(c, hipSetDevice(c->device));
    hipExternalMemoryBufferDesc b;
    memset(&b, 0, sizeof b);
    b.offset = offset;
    b.size = size;
    HL_HIP(c, hipExternalMemoryGetMappedBuffer(out_d, e->mem, &b));
    return HL_OK;
}

HL_API hl_status hl_external_release(hl_external* e) {
    if (!e) return HL_EINVAL;
    hl_context* c = e->ctx;
    HL_HIP(c, hipSetDevice(c->device));
    // work queued on the context stream may still read or write the memory
    HL_HIP(c, hipStreamSynchronize(c->stream));
    return release_external(e);
}

}  // extern "C"
