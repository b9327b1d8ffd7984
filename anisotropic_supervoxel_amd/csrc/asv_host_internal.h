// asv_host_internal.h -- host-side hooks shared by the library's own
// translation units (context reuse between streams); not part of the C ABI.
#pragma once

#include "../../include/asv.h"

namespace asv {

// True when context c can serve a run with parameters p (same frame size
// and run settings, at least p->depth frames and `seeds` seeds of capacity).
bool ctx_reusable(const asv_ctx* c, const asv_params* p, int32_t seeds);
int32_t ctx_seed_cap(const asv_ctx* c);

// A stream's resident context: taken out before the stream is destroyed,
// offered to a new stream before its first block.  The stream adopts it
// (true) when it fits; otherwise the caller keeps it.
asv_ctx* stream_take_ctx(asv_stream* s);
bool stream_adopt_ctx(asv_stream* s, asv_ctx* c);

}  // namespace asv
