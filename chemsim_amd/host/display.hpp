// Render and frame-recording surface of the C++ drop-in layer for an
// lbm::State: what src/main.rs's LBMSim::render (main.rs:138-177, through
// render.rs:7-178) and src/display.rs's record (display.rs:157-185) do with
// the lattice, on the GPU through cs_lattice_render / cs_lattice_record.
// Only the image bytes leave the device (the reference reads every field
// back and colours it pixel by pixel on the host).
#pragma once

#include <chemsim.h>

#include <cstdint>
#include <exception>
#include <functional>
#include <vector>

#include "lbm.hpp"

namespace chemsim {
namespace display {

// main.rs:36-42: the field shown.
enum class DisplayMode { Density = CS_RENDER_DENSITY, Speed = CS_RENDER_SPEED, Velocity = CS_RENDER_VELOCITY,
                         MomentumDensity = CS_RENDER_MOMENTUM };

// DisplayMode::next (main.rs:44-53).
inline DisplayMode next(DisplayMode m)
{
    switch (m) {
    case DisplayMode::Density: return DisplayMode::Speed;
    case DisplayMode::Speed: return DisplayMode::Velocity;
    case DisplayMode::Velocity: return DisplayMode::MomentumDensity;
    default: return DisplayMode::Density;
    }
}

// LBMSim::render (main.rs:138-177): render_scalar_field of density or speed,
// or render_vector_field of velocity or momentum density, then (if
// `geometry`) render_geometry.  Returns h x w RGB pixels, pixel (x, y) at
// 3 (y w + x).  The state's host edits reach the device first; its host copy
// of the populations stays valid.
inline std::vector<uint8_t> render(const lbm::State<>& s, DisplayMode mode, bool geometry = true)
{
    const auto [w, h] = s.size();
    std::vector<uint8_t> rgb(3 * w * h);
    lbm::check(cs_lattice_render(s.device_view(), (int)mode, geometry ? 1 : 0, rgb.data(), nullptr));
    return rgb;
}

// encoder.write_frame of display.rs:181-182: frame index and h x w RGBA
// pixels (alpha 255), valid only during the call.  Throwing stops the
// recording; record() rethrows the exception once the frames in flight have
// finished.
using FrameFn = std::function<void(int64_t frame, const uint8_t* rgba, int w, int h)>;

namespace detail {
struct FrameSink {
    const FrameFn* fn;
    std::exception_ptr error;
};
inline int deliver(void* user, int64_t frame, const uint8_t* rgba, int w, int h)
{
    auto* sink = static_cast<FrameSink*>(user);
    try {
        (*sink->fn)(frame, rgba, w, h);
        return 0;
    } catch (...) {
        sink->error = std::current_exception();  // never let it cross the C ABI
        return 1;
    }
}
}  // namespace detail

// record (display.rs:157-185): `frames` times, LBMSim::step's speed-factor
// loop of `steps_per_frame` steps (main.rs:128-136) then the render of
// `mode` as RGBA, handed to write_frame in frame order.  `depth` frames ring
// (1: the reference's order; more overlaps the copies and write_frame with
// the next frames' steps).  The state's time advances by the steps taken; when
// write_frame throws, up to depth - 1 frames more than were delivered have
// been stepped.  Returns the number of frames delivered.
inline int64_t record(lbm::State<>& s, int64_t frames, int steps_per_frame, DisplayMode mode, bool geometry,
                      int depth, const FrameFn& write_frame)
{
    cs_lattice* lat = s.device_lattice();
    double t0 = 0, t1 = 0;
    lbm::check(cs_lattice_time(lat, &t0));
    detail::FrameSink sink{&write_frame, nullptr};
    cs_record_stats stats{};
    const int rc = cs_lattice_record(lat, (int)frames, steps_per_frame, (int)mode, geometry ? 1 : 0, depth,
                                     detail::deliver, &sink, &stats);
    lbm::check(cs_lattice_time(lat, &t1));
    s.time += (lbm::Scalar)(t1 - t0);
    if (sink.error) std::rethrow_exception(sink.error);
    lbm::check(rc);
    return stats.frames;
}

}  // namespace display
}  // namespace chemsim
