// lsvd_workspace.cpp -- the handle's device workspace (chunks used as a
// stack, see lsvd_workspace.hpp).
#include "lsvd_workspace.hpp"

#include <algorithm>

#include "lsvd_api_internal.hpp"

namespace lsvd {

namespace {
thread_local WsScope* t_scope = nullptr;
}

Workspace::~Workspace() {
    for (Chunk& c : chunks_) (void)hipFree(c.p);
}

hipError_t Workspace::alloc(size_t bytes, void** out) {
    const size_t need = std::max<size_t>((bytes + kWsAlign - 1) / kWsAlign * kWsAlign, kWsAlign);
    // first fit from the current chunk on (chunks after it are empty)
    for (size_t i = cur_; i < chunks_.size(); ++i) {
        Chunk& c = chunks_[i];
        if (i > cur_) c.used = 0;
        if (c.cap - c.used >= need) {
            *out = c.p + c.used;
            c.used += need;
            cur_ = i;
            return hipSuccess;
        }
    }
    Chunk c{nullptr, std::max(need, kWsChunkMin), 0};
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&c.p), c.cap);
    if (e != hipSuccess) return e;
    c.used = need;
    chunks_.push_back(c);
    cur_ = chunks_.size() - 1;
    *out = c.p;
    return hipSuccess;
}

Workspace::Mark Workspace::mark() const {
    return Mark{cur_, cur_ < chunks_.size() ? chunks_[cur_].used : 0};
}

void Workspace::release(const Mark& m) {
    if (chunks_.empty()) return;
    cur_ = std::min(m.chunk, chunks_.size() - 1);
    chunks_[cur_].used = m.chunk < chunks_.size() ? m.used : 0;
    for (size_t i = cur_ + 1; i < chunks_.size(); ++i) chunks_[i].used = 0;
}

size_t Workspace::reserved_bytes() const {
    size_t s = 0;
    for (const Chunk& c : chunks_) s += c.cap;
    return s;
}

Workspace* handle_workspace(lsvd_handle_s* h) {
    if (!h->ws) h->ws = new Workspace();
    return h->ws;
}

WsScope::WsScope(lsvd_handle_s* h) : ws_(handle_workspace(h)), mark_(ws_->mark()), outer_(t_scope) {
    t_scope = this;
}

WsScope::~WsScope() {
    ws_->release(mark_);
    t_scope = outer_;
}

hipError_t WsScope::alloc(size_t bytes, void** out) { return ws_->alloc(bytes, out); }

WsScope* WsScope::current() { return t_scope; }

hipError_t WsMem::alloc(size_t bytes) {
    WsScope* s = WsScope::current();
    if (!s) return hipErrorNotInitialized;  // a WsScope must be open
    return s->alloc(bytes, &p);
}

}  // namespace lsvd

extern "C" int lsvd_workspace_info(lsvd_handle_t h, size_t* bytes, int* chunks) {
    if (!h) return LSVD_E_ARG;
    if (bytes) *bytes = h->ws ? h->ws->reserved_bytes() : 0;
    if (chunks) *chunks = h->ws ? (int)h->ws->chunk_count() : 0;
    return LSVD_OK;
}
