// Whole-call host I/O: operands in pageable caller memory travel through
// pinned staging to HBM, results come back the same way.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>

#include "../../include/bytecrt.h"

namespace bcrt {

// Chunks of products move through a ring of slots, each with pinned input,
// pinned output and device a/b/c regions.  The calling thread packs operands
// (with a pack pool of host threads) in pieces and enqueues each piece's
// host-to-device copy right behind it, then the chunk's compute and its
// device-to-host copies in pieces; a drain thread waits for each output piece
// and unpacks it into the caller's array (with its own pool) or hands it to a
// sink.  Packing, both DMA directions, compute and unpacking of different
// chunks therefore overlap.
class HostIo {
   public:
    // Enqueues the product of n device-resident products on stream s.
    using Enqueue = std::function<void(const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, uint32_t n,
                                       hipStream_t s)>;

    // words = N*L per operand; chunks hold at most max_chunk products.
    HostIo(size_t words, uint32_t max_chunk, hipStream_t compute);
    ~HostIo();
    HostIo(const HostIo&) = delete;
    HostIo& operator=(const HostIo&) = delete;

    uint32_t chunk() const;
    // a, b: [batch][words].  Results go to c ([batch][words]) or, when c is
    // null, to sink(ctx, piece, nwords) in product order on the drain thread.
    void run(const uint32_t* a, const uint32_t* b, uint32_t batch, uint32_t* c, bcrt_sink sink, void* ctx,
             const Enqueue& enqueue);

   private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace bcrt
