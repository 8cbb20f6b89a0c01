// lsvd_extras.cpp -- per-handle state of the paths beyond the batched SVD.
#include "lsvd_api_internal.hpp"

namespace lsvd {
struct Extras {
    int unused = 0;
};
}  // namespace lsvd

void lsvd_release_extras(lsvd_handle_t h) {
    if (!h) return;
    delete h->extras;
    h->extras = nullptr;
}
