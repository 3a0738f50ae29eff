// Drives the C++ drop-in layer (chemsim_amd/host/lbm.hpp) the way main.rs
// drives lbm.rs: a lattice at compute_equilibrium of a uniform flow, a
// circular obstacle with solid borders, State::initial with BGK, steps, and
// the fields the renderer reads.  Writes raw float32 arrays into OUTDIR for
// tests/test_gpu_host_cpp.py to compare with the oracle:
//   fused.f32     20 State::step with BGK (fused device step)
//   phases.f32    20 x (stream, bounce_back, collide) called one by one
//   plugin.f32    20 State::step with a user operator evaluated on the host
//   bgk_eval.f32  BGK::evaluate of the initial populations and their equilibrium
//   moments.f32   density, ux, uy, speed, pressure after the fused run
//   edited.f32    10 fused steps, a geometry edit, 10 more
//   trt.f32       20 State::step with TRT::new(0.25, 10) (main.rs:188-189)
//   trt_eval.f32  TRT::evaluate of the initial populations
//   trt_tau.f32   (tau_plus, tau_minus)
//   reg.f32, reg_eval.f32  the same with Regularized::new(BGK::new(0.6))
//   kbc.f32, kbc_eval.f32  the same with KBC::new(0.05)
//   edit.f32      5 fused steps, populations_mut() edited, 5 more
//   render.u8     display::render of the four modes with the geometry, after
//                 the 20 fused steps (4 x h x w x 3)
//   record.u8     display::record: 3 frames of 4 steps, velocity, depth 2
//                 (3 x h x w x 4); record.f32 the populations afterwards
//   stop.txt      frames delivered before a throwing write_frame stopped a
//                 recording, and the steps the state advanced
// `lbm_host_test cpu` runs only checks that need no GPU.
#include <display.hpp>
#include <lbm.hpp>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

using namespace chemsim;
using lbm::Scalar;

static int failures = 0;
#define EXPECT(c)                                                              \
    do {                                                                       \
        if (!(c)) {                                                            \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                                        \
        }                                                                      \
    } while (0)

static uint32_t bits(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// A user-written BGK: the plugin path with host arithmetic in the reference's order.
struct HostBGK final : lbm::CollisionOperator<lbm::D2Q9> {
    Scalar tau;
    explicit HostBGK(Scalar t) : tau(t) {}
    lbm::Populations evaluate(const lbm::D2Q9& lat, const lbm::Populations& eq,
                              const lbm::Discretization& d) const override
    {
        const Scalar factor = -d.delta_t / tau;
        lbm::Populations out;
        for (size_t i = 0; i < 9; ++i) {
            const auto& f = lat.populations()[i].second;
            out.emplace_back(lat.populations()[i].first, f + (f - eq[i].second).scale(factor));
        }
        return out;
    }
    Scalar kinematic_shear_viscosity(const lbm::Discretization& d) const override
    {
        return (d.delta_x * d.delta_x / (3.0f * d.delta_t * d.delta_t)) * (tau - d.delta_t / 2.0f);
    }
};

static void cpu_checks()
{
    lbm::Discretization disc{1.0f, 1.0f};
    EXPECT(bits(disc.isothermal_speed_of_sound()) == 0x3f13cd3au);
    const Scalar cs = disc.isothermal_speed_of_sound();
    EXPECT(bits(cs * cs) == 0x3eaaaaaau);
    const auto dirs = lbm::D2Q9::directions();
    EXPECT(dirs.size() == 9);
    float wsum = 0;
    for (const auto& d : dirs) {
        wsum += d.w_scalar;
        // the stencil's single 1 sits at row 1 + cy, column 1 - cx
        const auto st = d.stencil.get_underlying();
        int ones = 0;
        for (float v : st) ones += v == 1.0f;
        EXPECT(ones == 1);
        EXPECT(d.stencil.at((size_t)(1 - (int)d.c_vector.x), (size_t)(1 + (int)d.c_vector.y)) == 1.0f);
    }
    EXPECT(bits(dirs[0].w_scalar) == bits(16.0f / 36.0f));
    EXPECT(bits(dirs[7].w_scalar) == bits(1.0f / 36.0f));
    EXPECT(std::fabs(wsum - 1.0f) < 1e-6f);
    // stencils of direction 1 and 2 (lbm.rs:233-243): row-major one-hot
    EXPECT(dirs[1].stencil.get_underlying()[3] == 1.0f);
    EXPECT(dirs[2].stencil.get_underlying()[7] == 1.0f);
    bool threw = false;
    try {
        matrix::Matrix::new_({1.f, 2.f, 3.f}, {2, 2});
    } catch (const matrix::Error&) {
        threw = true;
    }
    EXPECT(threw);
    auto m = matrix::Matrix::new_({1.f, 2.f, 3.f, 4.f, 5.f, 6.f}, {3, 2});
    EXPECT(m.get_shape() == std::make_pair(size_t(3), size_t(2)));
    EXPECT(m.at(2, 1) == 6.f && m.hadamard(m).at(1, 0) == 4.f && m.sum() == 21.0);
    lbm::BGK b(0.6f);
    EXPECT(b.as_bgk() == &b);
    EXPECT(std::fabs(b.kinematic_shear_viscosity(disc) - (0.1f / 3.0f)) < 1e-7f);
    const auto trt = lbm::TRT::new_(0.25f, 10.0f, disc);
    EXPECT(std::fabs(trt.lambda(disc) - 0.25f) < 1e-5f);
    EXPECT(std::fabs(trt.kinematic_shear_viscosity(disc) - 10.0f) < 1e-4f);
    cs_params tp{};
    EXPECT(trt.device_params(tp) && tp.collision == CS_COLL_TRT && (float)tp.tau == trt.tau_plus);
    const auto reg = lbm::Regularized<lbm::BGK>::new_(lbm::BGK(0.6f));
    EXPECT(reg.kinematic_shear_viscosity(disc) == b.kinematic_shear_viscosity(disc));
    EXPECT(reg.device_params(tp) && tp.collision == CS_COLL_REGULARIZED);
    const auto kbc = lbm::KBC::new_(0.05f);
    EXPECT(kbc.kinematic_shear_viscosity(disc) == 0.05f);
    EXPECT(kbc.device_params(tp) && tp.collision == CS_COLL_KBC && (float)tp.ks_viscosity == 0.05f);
    lbm::Geometry g(4, 3);
    const auto v0 = g.version();
    g.set(3, 2, true);
    EXPECT(g.get(3, 2) && g.flags()[2 * 4 + 3] == 1 && g.version() != v0);
}

static void write(const std::string& path, const std::vector<float>& v)
{
    std::ofstream f(path, std::ios::binary);
    f.write((const char*)v.data(), (std::streamsize)(v.size() * sizeof(float)));
}

static void write_bytes(const std::string& path, const std::vector<uint8_t>& v)
{
    std::ofstream f(path, std::ios::binary);
    f.write((const char*)v.data(), (std::streamsize)v.size());
}

static std::vector<float> flat(const lbm::Populations& p)
{
    std::vector<float> out;
    for (const auto& pr : p) out.insert(out.end(), pr.second.data().begin(), pr.second.data().end());
    return out;
}

// main.rs:180-328 with a smaller obstacle: uniform flow (0.02, 0), rho = 1.
static lbm::State<> make_state(size_t w, size_t h, std::unique_ptr<lbm::CollisionOperator<lbm::D2Q9>> op)
{
    const lbm::Discretization disc{1.0f, 1.0f};
    auto rho = matrix::Matrix::new_filled(1.0f, {w, h});
    auto vx = matrix::Matrix::new_filled(0.02f, {w, h});
    auto vy = matrix::Matrix::new_filled(0.0f, {w, h});
    const auto eq = lbm::compute_equilibrium(rho, {vx, vy}, lbm::D2Q9::directions(), disc);
    std::vector<lbm::Population> pops;
    for (const auto& pr : eq) pops.push_back(pr.second);
    std::vector<bool> solid(w * h, false);
    for (size_t x = 0; x < w; ++x)
        for (size_t y = 0; y < h; ++y) {
            const double r = std::sqrt(std::pow(x - w / 2.0, 2) + std::pow(y - h / 2.0, 2));
            solid[y * w + x] = r < 9.0 || x == 0 || y == 0 || x == w - 1 || y == h - 1;
        }
    return lbm::State<>::initial(std::make_unique<lbm::D2Q9>(pops), lbm::Geometry::from_row_major(solid, w, h),
                                 std::move(op), disc);
}

int main(int argc, char** argv)
{
    cpu_checks();
    if (argc >= 2 && std::string(argv[1]) == "cpu") {
        std::printf("cpu checks: %s\n", failures ? "FAILED" : "ok");
        return failures ? 1 : 0;
    }
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s OUTDIR | cpu\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    const size_t w = 96, h = 64;
    try {
        {
            auto s = make_state(w, h, std::make_unique<lbm::BGK>(0.6f));
            const auto eq = s.equilibrium();
            write(out + "/bgk_eval.f32", flat(lbm::BGK(0.6f).evaluate(*s.lattice, eq, s.discretization)));
            EXPECT(!s.is_unstable());
            for (int i = 0; i < 20; ++i) s.step();
            EXPECT(s.time == 20.0f);
            write(out + "/fused.f32", flat(s.populations()));
            std::vector<float> m;
            for (const auto& f : {s.density(), s.velocity().first, s.velocity().second, s.speed(), s.pressure()})
                m.insert(m.end(), f.data().begin(), f.data().end());
            write(out + "/moments.f32", m);
            std::vector<uint8_t> img;
            for (auto mode : {display::DisplayMode::Density, display::DisplayMode::Speed,
                              display::DisplayMode::Velocity, display::DisplayMode::MomentumDensity}) {
                const auto rgb = display::render(s, mode, true);
                img.insert(img.end(), rgb.begin(), rgb.end());
            }
            write_bytes(out + "/render.u8", img);
            // rendering reads the device state only: the host copy is still valid
            write(out + "/fused_after_render.f32", flat(s.populations()));
        }
        {
            // host edits between steps reach the device before the next step
            auto s = make_state(w, h, std::make_unique<lbm::BGK>(0.6f));
            for (int i = 0; i < 5; ++i) s.step();
            auto& pm = s.lattice->populations_mut();
            pm[2].second.at(30, 20) += 0.01f;
            pm[5].second.at(31, 21) = 0.0f;
            for (int i = 0; i < 5; ++i) s.step();
            write(out + "/edit.f32", flat(s.populations()));
        }
        {
            auto s = make_state(w, h, std::make_unique<lbm::BGK>(0.6f));
            std::vector<uint8_t> frames;
            const int64_t got = display::record(s, 3, 4, display::DisplayMode::Velocity, true, 2,
                                                [&](int64_t k, const uint8_t* rgba, int fw, int fh) {
                                                    EXPECT(k == (int64_t)(frames.size() / (4 * w * h)));
                                                    EXPECT(fw == (int)w && fh == (int)h);
                                                    frames.insert(frames.end(), rgba, rgba + 4 * (size_t)fw * fh);
                                                });
            EXPECT(got == 3 && s.time == 12.0f);
            write_bytes(out + "/record.u8", frames);
            write(out + "/record.f32", flat(s.populations()));
        }
        {
            // a write_frame that throws stops the recording; the exception
            // comes back once the frames in flight are done
            auto s = make_state(w, h, std::make_unique<lbm::BGK>(0.6f));
            int delivered = 0;
            bool caught = false;
            try {
                display::record(s, 6, 1, display::DisplayMode::Density, false, 3,
                                [&](int64_t k, const uint8_t*, int, int) {
                                    if (k == 2) throw std::runtime_error("encoder full");
                                    ++delivered;
                                });
            } catch (const std::runtime_error& e) {
                caught = std::string(e.what()) == "encoder full";
            }
            EXPECT(caught && delivered == 2);
            std::ofstream f(out + "/stop.txt");
            f << delivered << " " << s.time << "\n";
        }
        {
            auto s = make_state(w, h, std::make_unique<lbm::BGK>(0.6f));
            for (int i = 0; i < 20; ++i) {
                s.stream();
                s.bounce_back();
                s.collide();
            }
            write(out + "/phases.f32", flat(s.populations()));
        }
        {
            auto s = make_state(w, h, std::make_unique<HostBGK>(0.6f));
            for (int i = 0; i < 20; ++i) s.step();
            write(out + "/plugin.f32", flat(s.populations()));
        }
        {
            auto s = make_state(w, h, std::make_unique<lbm::BGK>(0.6f));
            for (int i = 0; i < 10; ++i) s.step();
            for (size_t x = 20; x < 30; ++x) s.geometry.set(x, 10, true);
            for (int i = 0; i < 10; ++i) s.step();
            write(out + "/edited.f32", flat(s.populations()));
        }
        {
            const auto trt = lbm::TRT::new_(0.25f, 10.0f, lbm::Discretization{1.0f, 1.0f});
            auto s = make_state(w, h, std::make_unique<lbm::TRT>(trt));
            write(out + "/trt_eval.f32", flat(trt.evaluate(*s.lattice, s.equilibrium(), s.discretization)));
            for (int i = 0; i < 20; ++i) s.step();
            write(out + "/trt.f32", flat(s.populations()));
            write(out + "/trt_tau.f32", {trt.tau_plus, trt.tau_minus});
        }
        {
            const auto reg = lbm::Regularized<lbm::BGK>::new_(lbm::BGK(0.6f));
            auto s = make_state(w, h, std::make_unique<lbm::Regularized<lbm::BGK>>(reg));
            write(out + "/reg_eval.f32", flat(reg.evaluate(*s.lattice, s.equilibrium(), s.discretization)));
            for (int i = 0; i < 20; ++i) s.step();
            write(out + "/reg.f32", flat(s.populations()));
        }
        {
            const auto kbc = lbm::KBC::new_(0.05f);
            auto s = make_state(w, h, std::make_unique<lbm::KBC>(kbc));
            write(out + "/kbc_eval.f32", flat(kbc.evaluate(*s.lattice, s.equilibrium(), s.discretization)));
            for (int i = 0; i < 20; ++i) s.step();
            write(out + "/kbc.f32", flat(s.populations()));
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::printf("host layer run: %s\n", failures ? "FAILED" : "ok");
    return failures ? 1 : 0;
}
