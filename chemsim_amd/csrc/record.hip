// Frame recording (display.rs:157-185): per frame, Simulation::step (the
// speed-factor loop of State::step, main.rs:128-136), Simulation::render into
// an RGBA image (render.rs through RGB::to_rgba, display.rs:27-29, alpha 255)
// and encoder.write_frame.  The reference renders on the host and hands each
// frame to the encoder before the next step starts.  Here the steps and the
// render passes of frame k+1.. run on the lattice's stream while frame k's
// RGBA bytes travel to pinned host memory on a second stream and the caller's
// frame function (the encoder) runs on the host: a ring of `depth` device
// slots and pinned host slots, with events ordering render -> copy -> reuse.
#include <hip/hip_runtime.h>

#include <chrono>
#include <vector>

#include "cs_internal.hpp"

namespace cs {

namespace {

struct RecordRing {
    cs_lattice* l = nullptr;
    hipStream_t copy = nullptr;
    std::vector<hipEvent_t> rendered, copied;
    uint8_t* dslots = nullptr;
    uint8_t* hslots = nullptr;

    ~RecordRing()
    {
        // drain before releasing anything a queued operation still uses
        if (copy) (void)hipStreamSynchronize(copy);
        if (l) (void)hipStreamSynchronize(l->stream);
        for (hipEvent_t e : rendered)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : copied)
            if (e) (void)hipEventDestroy(e);
        if (copy) (void)hipStreamDestroy(copy);
        if (dslots) (void)hipFree(dslots);
        if (hslots) (void)hipHostFree(hslots);
    }
};

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

}  // namespace cs

using namespace cs;

extern "C" int cs_lattice_record(cs_lattice* l, int frames, int steps_per_frame, int mode, int overlay, int depth,
                                 cs_frame_fn fn, void* user, cs_record_stats* stats)
{
    return guarded([&] {
        CS_REQUIRE(l != nullptr && fn != nullptr, CS_ERR_ARG, "null argument");
        CS_REQUIRE(frames >= 0 && steps_per_frame >= 0, CS_ERR_ARG, "negative frame or step count");
        CS_REQUIRE(mode >= CS_RENDER_DENSITY && mode <= CS_RENDER_MOMENTUM, CS_ERR_ARG, "bad render mode");
        CS_REQUIRE(depth >= 1 && depth <= CS_RECORD_MAX_DEPTH, CS_ERR_ARG, "ring depth out of range");
        CS_HIP(hipSetDevice(l->ctx->device));
        const size_t fbytes = (size_t)4 * l->w * l->h;
        const size_t dstride = (fbytes + 255) / 256 * 256;  // device slots 256-byte aligned for the pixel pass
        RecordRing ring;
        ring.l = l;
        CS_HIP(hipStreamCreateWithFlags(&ring.copy, hipStreamNonBlocking));
        ring.rendered.assign(depth, nullptr);
        ring.copied.assign(depth, nullptr);
        for (int s = 0; s < depth; ++s) {
            CS_HIP(hipEventCreateWithFlags(&ring.rendered[s], hipEventDisableTiming));
            CS_HIP(hipEventCreateWithFlags(&ring.copied[s], hipEventDisableTiming));
        }
        CS_HIP(hipMalloc(&ring.dslots, dstride * depth));
        CS_HIP(hipHostMalloc((void**)&ring.hslots, fbytes * depth, hipHostMallocDefault));
        const auto t_start = std::chrono::steady_clock::now();  // set-up above is not timed

        double wait_ms = 0, cb_ms = 0;
        int64_t delivered = 0;
        bool stop = false;
        auto deliver = [&](int64_t j) {
            const int s = (int)(j % depth);
            const auto t0 = std::chrono::steady_clock::now();
            CS_HIP(hipEventSynchronize(ring.copied[s]));
            wait_ms += ms_since(t0);
            const auto t1 = std::chrono::steady_clock::now();
            const int rc = fn(user, j, ring.hslots + fbytes * s, l->w, l->h);
            cb_ms += ms_since(t1);
            if (rc != 0) stop = true;
            else ++delivered;
        };

        CS_HIP(hipEventRecord(l->ev0, l->stream));
        for (int64_t k = 0; k < frames && !stop; ++k) {
            const int s = (int)(k % depth);
            // slot s last held frame k - depth, delivered (copy finished) in an earlier iteration
            do_steps(l, steps_per_frame);
            uint8_t* dslot = ring.dslots + dstride * s;
            render_enqueue(l, mode, overlay, dslot, 4);
            CS_HIP(hipEventRecord(ring.rendered[s], l->stream));
            CS_HIP(hipStreamWaitEvent(ring.copy, ring.rendered[s], 0));
            CS_HIP(hipMemcpyAsync(ring.hslots + fbytes * s, dslot, fbytes, hipMemcpyDeviceToHost, ring.copy));
            CS_HIP(hipEventRecord(ring.copied[s], ring.copy));
            if (k - (depth - 1) >= 0) deliver(k - (depth - 1));
        }
        CS_HIP(hipEventRecord(l->ev1, l->stream));
        for (int64_t j = delivered; j < frames && !stop; ++j) deliver(j);
        CS_HIP(hipStreamSynchronize(ring.copy));
        CS_HIP(hipStreamSynchronize(l->stream));
        if (stats) {
            float gpu = 0;
            CS_HIP(hipEventElapsedTime(&gpu, l->ev0, l->ev1));
            stats->gpu_ms = gpu;
            stats->wait_ms = wait_ms;
            stats->callback_ms = cb_ms;
            stats->frames = delivered;
            stats->wall_ms = ms_since(t_start);
        }
        CS_REQUIRE(!stop, CS_ERR_CALLBACK, "frame function returned non-zero; recording stopped");
    });
}
