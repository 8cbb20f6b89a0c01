// lsvd_workspace.hpp -- the device workspace a handle keeps for the
// temporaries of its entry points (the chip-wide route, the large polar,
// Phase-BSS, the batched consumers), so that repeated calls allocate
// nothing once the workspace has grown to their size.
//
// The workspace is a list of device chunks used as a stack.  An entry point
// opens a WsScope (first statement); every WsMem::alloc inside it takes
// the next free bytes, and the scope hands them all back when it ends.
// Scopes nest in call order (a helper's scope ends before its caller
// allocates again), and the innermost open scope of the thread is the one
// WsMem uses.  When no chunk has room, a new chunk is allocated (at least
// kWsChunkMin bytes) and kept until the handle is destroyed.  The entry
// points using it synchronise their stream before returning, so the next
// call may reuse the bytes at once.  A handle is used by one host thread
// at a time.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

struct lsvd_handle_s;

namespace lsvd {

constexpr size_t kWsChunkMin = size_t(32) << 20;
constexpr size_t kWsAlign = 256;

class Workspace {
public:
    struct Mark {
        size_t chunk, used;
    };
    ~Workspace();
    hipError_t alloc(size_t bytes, void** out);
    Mark mark() const;
    void release(const Mark& m);
    size_t reserved_bytes() const;
    size_t chunk_count() const { return chunks_.size(); }

private:
    struct Chunk {
        char* p;
        size_t cap, used;
    };
    std::vector<Chunk> chunks_;
    size_t cur_ = 0;
};

class WsScope {
public:
    explicit WsScope(lsvd_handle_s* h);
    ~WsScope();
    WsScope(const WsScope&) = delete;
    WsScope& operator=(const WsScope&) = delete;
    hipError_t alloc(size_t bytes, void** out);
    static WsScope* current();

private:
    Workspace* ws_;
    Workspace::Mark mark_;
    WsScope* outer_;
};

// A temporary device buffer from the innermost open scope (the interface of
// the per-call buffers it replaces: alloc, then .p).
struct WsMem {
    void* p = nullptr;
    hipError_t alloc(size_t bytes);
};

Workspace* handle_workspace(lsvd_handle_s* h);

}  // namespace lsvd
