// asv_host_internal.h -- host-side hooks shared by the library's own
// translation units (context reuse between streams); not part of the C ABI.
#pragma once

#include "../../include/asv.h"

namespace asv {

// True when context c can serve a run with parameters p (same frame size
// and run settings, at least p->depth frames and `seeds` seeds of capacity).
bool ctx_reusable(const asv_ctx* c, const asv_params* p, int32_t seeds);
int32_t ctx_seed_cap(const asv_ctx* c);

// Device bytes of the brick lists of the fused first two passes for a clip
// of W x H x D voxels with room for `seeds` seeds and a reach (first two
// steps together) of at most R voxels.
int64_t seed_brick_bytes(int W, int H, int D, int64_t seeds, int R);

// A stream's resident context: taken out before the stream is destroyed,
// offered to a new stream before its first block.  The stream adopts it
// (true) when it fits; otherwise the caller keeps it.
asv_ctx* stream_take_ctx(asv_stream* s);
bool stream_adopt_ctx(asv_stream* s, asv_ctx* c);

// A streaming block in two halves around its Lloyd loop: begin sets up the
// block's seeds and frames in the stream's resident context (on `stream`),
// the caller runs the loop on that context (stream_ctx), end reads the
// sums back, post-processes, stitches the finished frames into `out` and
// keeps the seeds for the next block.  asv_stream_block_into is begin,
// asv_run, end.
int stream_block_begin(asv_stream* s, int32_t block, const uint8_t* rgb, const float* fwd, const float* bwd,
                       void* stream);
asv_ctx* stream_ctx(asv_stream* s);
int stream_block_end(asv_stream* s, int32_t* out, int32_t* frames_done, void* stream);

// The Lloyd loops of several contexts (same iteration count) together:
// each context's prepare, seed stages, own passes, reductions and
// finalisations on its own stream (streams[i]), and per pass index and step
// one marching launch over all contexts that march it on batch_stream, with
// events where a context's work moves between the two.  Everything is
// enqueued, nothing waited for; each context's work ends on its own stream.
// Labels and sums are those of each context's own asv_run.  Keeps its launch
// descriptors in a ring of device buffers; one object per thread.
class WindowBatcher {
public:
    WindowBatcher();
    ~WindowBatcher();
    int run(asv_ctx* const* ctxs, void* const* streams, int n, void* batch_stream);
    WindowBatcher(const WindowBatcher&) = delete;
    WindowBatcher& operator=(const WindowBatcher&) = delete;

private:
    struct Impl;
    Impl* impl_;
};

}  // namespace asv
