// Drives every entry point of libchemsim_hip.so's host code, and the C++
// drop-in layer (chemsim_amd/host/lbm.hpp) the way a renderer and a frame
// recorder drive it, against the CPU model of the HIP runtime in hip_stub.cpp,
// under AddressSanitizer and UBSan.  Kernels do not run, so no value is
// checked here -- only that no host or "device" buffer is read or written
// outside its bounds or after it was freed, and that the status codes of the
// calls that must succeed are CS_OK.  Test infrastructure only.
#include <lbm.hpp>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace chemsim;

static int failures = 0;

static void ok(int rc, const char* what)
{
    if (rc != CS_OK) {
        std::fprintf(stderr, "FAILED %s: rc=%d (%s)\n", what, rc, cs_last_error());
        ++failures;
    }
}
#define OK(call) ok((call), #call)
#define REJECTED(call)                                                      \
    do {                                                                    \
        if ((call) == CS_OK) {                                              \
            std::fprintf(stderr, "FAILED (accepted) %s\n", #call);          \
            ++failures;                                                     \
        }                                                                   \
    } while (0)

struct Frames {
    int w = 0, h = 0;
    std::vector<std::vector<uint8_t>> got;
    int stop_after = -1;
};

static int on_frame(void* user, int64_t frame, const uint8_t* rgba, int w, int h)
{
    auto* f = static_cast<Frames*>(user);
    if (w != f->w || h != f->h || frame != (int64_t)f->got.size()) return 1;
    f->got.emplace_back(rgba, rgba + (size_t)4 * w * h);
    return f->stop_after >= 0 && (int)f->got.size() >= f->stop_after ? 1 : 0;
}

static cs_params params(int edge, int boundary, int coll, int rounding)
{
    cs_params p{};
    p.tau = 0.6;
    p.delta_x = 1;
    p.delta_t = 1;
    p.edge = edge;
    p.boundary = boundary;
    p.lid_uy = 0.05;
    p.inlet_ux = 0.05;
    p.outlet_rho = 1;
    p.rho0 = 1;
    p.rounding = rounding;
    p.seed = 7;
    p.collision = coll;
    p.tau_minus = 0.7;
    p.ks_viscosity = 0.05;
    return p;
}

static std::vector<uint8_t> flags_for(int w, int h)
{
    std::vector<uint8_t> f((size_t)w * h, CS_FLUID);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            uint8_t& c = f[(size_t)y * w + x];
            if (y == h - 1) c = CS_WALL;
            else if (y == 0) c = CS_LID;
            else if (x == 0) c = CS_INLET;
            else if (x == w - 1) c = CS_OUTLET;
            else if ((x - w / 2) * (x - w / 2) + (y - h / 2) * (y - h / 2) < 4) c = CS_WALL;
        }
    return f;
}

static void lattice_api(cs_ctx* ctx, int w, int h, int storage)
{
    cs_lattice* l = nullptr;
    OK(cs_lattice_create(ctx, w, h, storage, &l));
    if (!l) return;
    const size_t n = (size_t)w * h;
    int sw = 0, sh = 0, ss = -1;
    OK(cs_lattice_shape(l, &sw, &sh, &ss));
    const bool full = storage == CS_F32 || storage == CS_F64;
    for (int edge : {(int)CS_EDGE_ZERO, (int)CS_EDGE_PERIODIC})
        for (int bb : {(int)CS_BB_REFERENCE, (int)CS_BB_HALFWAY}) {
            const int colls = full ? 4 : 1;
            for (int coll = 0; coll < colls; ++coll) {
                const cs_params p = params(edge, bb, coll, storage == CS_BF16DEV || storage == CS_F16DEV
                                                               ? CS_ROUND_STOCHASTIC
                                                               : CS_ROUND_NEAREST);
                OK(cs_lattice_set_params(l, &p));
                std::vector<float> rho(n, 1.0f), ux(n, 0.01f), uy(n, 0.0f);
                OK(cs_lattice_init_equilibrium(l, rho.data(), ux.data(), uy.data()));
                OK(cs_lattice_step(l, 3));
                float ms = 0;
                OK(cs_lattice_step_timed(l, 2, &ms));
            }
        }
    const cs_params p = params(CS_EDGE_ZERO, CS_BB_HALFWAY, CS_COLL_BGK, CS_ROUND_NEAREST);
    OK(cs_lattice_set_params(l, &p));
    const auto fl = flags_for(w, h);
    OK(cs_lattice_set_flags(l, fl.data()));
    {
        std::vector<float> soa(9 * n, 0.1f), back(9 * n);
        OK(cs_lattice_upload_f32(l, soa.data()));
        OK(cs_lattice_download_f32(l, back.data()));
        std::vector<double> d(9 * n, 0.1), dback(9 * n);
        OK(cs_lattice_upload_f64(l, d.data()));
        OK(cs_lattice_download_f64(l, dback.data()));
        std::vector<uint8_t> raw(9 * n * (storage == CS_F64 ? 8 : storage == CS_F32 || storage == CS_F32DEV ? 4 : 2));
        OK(cs_lattice_download_raw(l, raw.data()));
        OK(cs_lattice_upload_raw(l, raw.data()));
    }
    {
        std::vector<float> rho(n, 1.0f), ux(n, 0.01f), uy(n, 0.0f);
        OK(cs_lattice_init_nonequilibrium(l, rho.data(), ux.data(), uy.data()));
        std::vector<double> drho(n, 1.0), dux(n, 0.0), duy(n, 0.0);
        OK(cs_lattice_init_equilibrium_f64(l, drho.data(), dux.data(), duy.data()));
    }
    // runs of more than 64 steps take the graph replay on small lattices
    OK(cs_lattice_step(l, 130));
    {
        // the phases alone: reference boundary mode, full storages only
        const cs_params pr = params(CS_EDGE_ZERO, CS_BB_REFERENCE, CS_COLL_BGK, CS_ROUND_NEAREST);
        OK(cs_lattice_set_params(l, &pr));
        if (full) {
            OK(cs_lattice_stream(l));
            OK(cs_lattice_bounce_back(l));
            OK(cs_lattice_collide(l));
        } else {
            REJECTED(cs_lattice_stream(l));
        }
        OK(cs_lattice_set_params(l, &p));
    }
    double t = 0;
    OK(cs_lattice_time(l, &t));
    OK(cs_lattice_sync(l));
    void* f = nullptr;
    void* stream = nullptr;
    int pitch = 0;
    int64_t plane = 0;
    OK(cs_lattice_device_ptr(l, &f, &pitch, &plane));
    OK(cs_lattice_stream_handle(l, &stream));
    {
        std::vector<float> a(n), b(n), e(9 * n);
        OK(cs_lattice_density(l, a.data()));
        OK(cs_lattice_momentum_density(l, a.data(), b.data()));
        OK(cs_lattice_velocity(l, a.data(), b.data()));
        OK(cs_lattice_speed(l, a.data()));
        OK(cs_lattice_pressure(l, a.data()));
        OK(cs_lattice_equilibrium(l, e.data()));
        OK(cs_lattice_non_equilibrium(l, e.data()));
        int unstable = -1;
        OK(cs_lattice_is_unstable(l, &unstable));
        std::vector<double> r(n), x(n), y(n);
        OK(cs_lattice_moments_f64(l, r.data(), x.data(), y.data()));
        OK(cs_lattice_moments_f64(l, nullptr, x.data(), nullptr));
        OK(cs_lattice_moments_f64_noalloc(l, r.data(), x.data(), y.data()));
        OK(cs_lattice_moments_f64_noalloc(l, nullptr, x.data(), nullptr));
        OK(cs_lattice_moments_f64_noalloc(l, r.data(), nullptr, y.data()));
        OK(cs_lattice_moments_f64_noalloc(l, nullptr, nullptr, nullptr));
        double sums[3];
        OK(cs_lattice_sums_f64(l, sums));
        double res = 0;
        REJECTED(cs_lattice_residual_f64(l, &res));  // no snapshot yet
        OK(cs_lattice_snapshot_velocity(l));
        OK(cs_lattice_step(l, 2));
        OK(cs_lattice_residual_f64(l, &res));
        int64_t steps = 0;
        OK(cs_lattice_run_to_residual(l, 1e-6, 10, 50, &steps, &res));
    }
    for (int mode = 0; mode < 4; ++mode)
        for (int overlay = 0; overlay < 2; ++overlay) {
            std::vector<uint8_t> rgb(3 * n);
            float ms = 0;
            OK(cs_lattice_render(l, mode, overlay, rgb.data(), overlay ? &ms : nullptr));
        }
    for (int depth = 1; depth <= 4; ++depth)
        for (int frames : {0, 1, 5}) {
            Frames fr;
            fr.w = w;
            fr.h = h;
            cs_record_stats st{};
            OK(cs_lattice_record(l, frames, 2, depth % 4, depth & 1, depth, on_frame, &fr, &st));
            if ((int)fr.got.size() != frames) {
                std::fprintf(stderr, "FAILED record delivered %zu of %d frames\n", fr.got.size(), frames);
                ++failures;
            }
        }
    {
        Frames fr;
        fr.w = w;
        fr.h = h;
        fr.stop_after = 2;
        if (cs_lattice_record(l, 6, 1, 0, 0, 3, on_frame, &fr, nullptr) != CS_ERR_CALLBACK) {
            std::fprintf(stderr, "FAILED record stop\n");
            ++failures;
        }
    }
    REJECTED(cs_lattice_record(l, 1, 1, 0, 0, 0, on_frame, nullptr, nullptr));
    REJECTED(cs_lattice_render(l, 9, 0, nullptr, nullptr));
    OK(cs_lattice_destroy(l));
}

static void batch_api(cs_ctx* ctx, int storage)
{
    const std::vector<int32_t> ws = {64, 17, 130, 8, 96}, hs = {64, 33, 70, 8, 40};
    cs_batch* b = nullptr;
    OK(cs_batch_create(ctx, (int)ws.size(), ws.data(), hs.data(), storage, &b));
    if (!b) return;
    int count = 0;
    OK(cs_batch_count(b, &count));
    for (int i = 0; i < count; ++i) {
        const size_t n = (size_t)ws[i] * hs[i];
        const cs_params p = params(CS_EDGE_ZERO, CS_BB_HALFWAY, CS_COLL_BGK,
                                   storage == CS_F16DEV ? CS_ROUND_STOCHASTIC : CS_ROUND_NEAREST);
        OK(cs_batch_set_params(b, i, &p));
        const auto fl = flags_for(ws[i], hs[i]);
        OK(cs_batch_set_flags(b, i, fl.data()));
        std::vector<float> rho(n, 1.0f), ux(n, 0.0f), uy(n, 0.0f);
        OK(cs_batch_init_equilibrium(b, i, rho.data(), ux.data(), uy.data()));
        std::vector<float> soa(9 * n, 0.1f);
        OK(cs_batch_upload_f32(b, i, soa.data()));
    }
    cs_batch_tuning t{};
    OK(cs_batch_get_tuning(b, &t));
    for (int launch : {(int)CS_LAUNCH_PERSISTENT, (int)CS_LAUNCH_TILES})
        for (int bt : {256, 1024}) {
            t.launch = launch;
            t.block_threads = bt;
            t.cells_per_thread = bt == 256 ? 2 : 1;
            OK(cs_batch_set_tuning(b, &t));
            float ms = 0;
            OK(cs_batch_step(b, 3, &ms));
        }
    float ms = 0;
    OK(cs_batch_step_per_lattice(b, 2, &ms));
    cs_batch_tuning best{};
    OK(cs_batch_autotune(b, 1, &best, &ms));
    OK(cs_batch_set_active(b, 1, 0));
    OK(cs_batch_step(b, 2, nullptr));
    OK(cs_batch_set_active(b, 1, 1));
    OK(cs_batch_snapshot_velocity(b));
    OK(cs_batch_step(b, 1, nullptr));
    std::vector<double> res(count);
    OK(cs_batch_residuals(b, res.data()));
    for (int i = 0; i < count; ++i) {
        const size_t n = (size_t)ws[i] * hs[i];
        std::vector<float> a(n), c(n), d(n), soa(9 * n);
        OK(cs_batch_velocity(b, i, a.data(), c.data()));
        OK(cs_batch_moments(b, i, a.data(), c.data(), d.data()));
        OK(cs_batch_download_f32(b, i, soa.data()));
        std::vector<uint8_t> raw(9 * n * (storage == CS_F32 ? 4 : 2));
        OK(cs_batch_download_raw(b, i, raw.data()));
    }
    REJECTED(cs_batch_set_params(b, count, nullptr));
    OK(cs_batch_destroy(b));
}

static void cnn_api(cs_ctx* ctx)
{
    const int H = 32, W = 64, N = 2;
    cs_cnn* net = nullptr;
    OK(cs_cnn_create(ctx, H, W, 8, N, 1u, &net));
    if (!net) return;
    int64_t np = 0;
    OK(cs_cnn_param_count(net, &np));
    std::vector<float> prm((size_t)np), grads((size_t)np);
    OK(cs_cnn_get_params(net, prm.data()));
    OK(cs_cnn_set_params(net, prm.data()));
    const size_t px = (size_t)H * W;
    std::vector<float> in(N * px * 8, 0.5f), out(N * px * 3), tgt(N * px * 3, 0.1f), wt(N * px, 1.0f);
    OK(cs_cnn_forward(net, N, in.data(), out.data()));
    OK(cs_cnn_forward(net, 1, in.data(), out.data()));
    float loss = 0;
    OK(cs_cnn_loss_grad(net, N, in.data(), tgt.data(), wt.data(), &loss));
    OK(cs_cnn_get_grads(net, grads.data()));
    OK(cs_cnn_adam_step(net, 1e-3f, 0.9f, 0.999f, 1e-8f));
    const int count = 5;
    std::vector<float> din(count * px * 8, 0.5f), dtg(count * px * 3, 0.1f), dw(count * px, 1.0f);
    OK(cs_cnn_set_dataset(net, count, din.data(), dtg.data(), dw.data()));
    std::vector<int32_t> order = {4, 0, 3, 1, 2};
    OK(cs_cnn_train_epoch(net, order.data(), count, 1e-3f, &loss));
    float ms = 0;
    double flops = 0;
    OK(cs_cnn_time_step(net, 2, &ms, &flops));
    OK(cs_cnn_destroy(net));
    const int M = 40, Nn = 24, K = 64;
    std::vector<uint16_t> A((size_t)M * K, 0x3f80), B((size_t)Nn * K, 0x3f80);
    std::vector<float> C((size_t)M * Nn);
    OK(cs_gemm_bf16(ctx, M, Nn, K, A.data(), B.data(), C.data(), &ms));
}

// ---- the C++ layer, as a renderer / recorder of an lbm::State drives it ------

static lbm::State<> make_state(size_t w, size_t h)
{
    const lbm::Discretization disc{1.0f, 1.0f};
    auto rho = matrix::Matrix::new_filled(1.0f, {w, h});
    auto vx = matrix::Matrix::new_filled(0.02f, {w, h});
    auto vy = matrix::Matrix::new_filled(0.0f, {w, h});
    const auto eq = lbm::compute_equilibrium(rho, {vx, vy}, lbm::D2Q9::directions(), disc);
    std::vector<lbm::Population> pops;
    for (const auto& pr : eq) pops.push_back(pr.second);
    std::vector<bool> solid(w * h, false);
    for (size_t x = 0; x < w; ++x) solid[x] = solid[(h - 1) * w + x] = true;
    return lbm::State<>::initial(std::make_unique<lbm::D2Q9>(pops), lbm::Geometry::from_row_major(solid, w, h),
                                 std::make_unique<lbm::BGK>(0.6f), disc);
}

// render of main.rs:157-176 for a State: bring host edits to the device, then
// the device render pass of the lattice's current populations.
static std::vector<uint8_t> render_state(lbm::State<>& s, int mode, bool overlay)
{
    const auto [w, h] = s.size();
    cs_lattice* lat = s.lattice->handle();
    std::vector<uint8_t> rgb(3 * w * h);
    lbm::check(cs_lattice_render(lat, mode, overlay ? 1 : 0, rgb.data(), nullptr));
    return rgb;
}

static void host_layer()
{
    const size_t w = 48, h = 20;
    auto s = make_state(w, h);
    for (int i = 0; i < 3; ++i) s.step();
    // hold a reference to the host populations across device work
    const lbm::Populations& view = s.populations();
    float keep = view[1].second.data()[5];
    for (int mode = 0; mode < 4; ++mode) (void)render_state(s, mode, mode & 1);
    s.step();
    (void)s.density();
    (void)s.velocity();
    (void)s.speed();
    (void)s.pressure();
    (void)s.equilibrium();
    (void)s.non_equilibrium();
    (void)s.is_unstable();
    keep += s.populations()[1].second.data()[5];
    // host edits between steps
    auto& mut = s.lattice->populations_mut();
    mut[2].second.data()[7] = 0.5f;
    s.step();
    s.stream();
    s.bounce_back();
    s.collide();
    auto& mut2 = s.lattice->populations_mut();
    mut2[3].second.data()[1] = keep;
    (void)render_state(s, 2, true);
    // recording through the State's lattice
    {
        Frames fr;
        fr.w = (int)w;
        fr.h = (int)h;
        cs_lattice* lat = s.lattice->handle();
        s.lattice->invalidate_host();
        lbm::check(cs_lattice_record(lat, 4, 2, 1, 1, 2, on_frame, &fr, nullptr));
        double t = 0;
        lbm::check(cs_lattice_time(lat, &t));
        s.time = (lbm::Scalar)t;
    }
    (void)s.populations();
    s.geometry.set(5, 5, true);
    s.step();
    (void)lbm::BGK(0.6f).evaluate(*s.lattice, s.equilibrium(), s.discretization);
    (void)lbm::TRT::new_(0.25f, 10.0f, s.discretization).evaluate(*s.lattice, s.equilibrium(), s.discretization);
    (void)lbm::KBC(0.05f).evaluate(*s.lattice, s.equilibrium(), s.discretization);
    (void)s.lattice->swap_populations();
}

int main()
{
    cs_ctx* ctx = nullptr;
    OK(cs_ctx_create(0, &ctx));
    char name[64];
    OK(cs_ctx_device_name(ctx, name, sizeof name));
    for (int cpt : {0, 1, 2, 4, (int)CS_STEP_VECTOR}) {
        OK(cs_set_cells_per_thread(cpt));
        for (int storage : {(int)CS_F32, (int)CS_F64, (int)CS_F32DEV, (int)CS_BF16DEV, (int)CS_F16DEV}) {
            lattice_api(ctx, 37, 23, storage);
            lattice_api(ctx, 64, 64, storage);
            lattice_api(ctx, 1, 1, storage);
            lattice_api(ctx, 1, 9, storage);
        }
    }
    OK(cs_set_cells_per_thread(0));
    OK(cs_set_step_graphs(0));
    lattice_api(ctx, 64, 64, CS_F32);
    OK(cs_set_step_graphs(1));
    for (int storage : {(int)CS_F32, (int)CS_F16DEV}) batch_api(ctx, storage);
    {
        // the batch kernels use 32-bit byte offsets within a lattice's buffer
        const int32_t big[2] = {64, 8192};
        cs_batch* b = nullptr;
        if (cs_batch_create(ctx, 2, big, big, CS_F64, &b) != CS_ERR_UNSUPPORTED || b) {
            std::fprintf(stderr, "FAILED: 8192^2 FP64 batch lattice not rejected\n");
            ++failures;
        }
        const int32_t wide[1] = {8192}, tall[1] = {7283};  // 9 * 8192 * 7283 * 8 B just above 4 GiB
        if (cs_batch_create(ctx, 1, wide, tall, CS_F64, &b) != CS_ERR_UNSUPPORTED) {
            std::fprintf(stderr, "FAILED: 8192 x 7283 FP64 batch lattice not rejected\n");
            ++failures;
        }
    }
    cnn_api(ctx);
    {
        std::vector<float> f(999, 0.1f), feq(999, 0.11f), out(999);
        OK(cs_bgk_evaluate(ctx, 999, f.data(), feq.data(), 0.6f, 1.0f, out.data()));
        float tp = 0, tm = 0;
        OK(cs_trt_new(0.25f, 10.0f, 1.0f, 1.0f, &tp, &tm));
        float ms = 0;
        OK(cs_bench_copy(ctx, 1 << 20, 2, &ms));
        uint64_t mismatches = 1;
        OK(cs_check_reciprocal(ctx, &mismatches));
    }
    try {
        host_layer();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "FAILED host layer: %s\n", e.what());
        ++failures;
    }
    OK(cs_ctx_destroy(ctx));
    std::printf("sanitized host run: %s\n", failures ? "FAILED" : "ok");
    return failures ? 1 : 0;
}
