// LDS-tiled sweep (placeholder until the tiled kernel lands).
#pragma once
#include "pf_runtime.hpp"

namespace pf {
template <class R, int MODE>
int launch_tiled(const pf_ctx* c, const Mesh<R>& m, hipStream_t s) {
    (void)c; (void)m; (void)s;
    return fail(PF_ERR_STATE, "tiled kernel not available in this build");
}
}  // namespace pf
