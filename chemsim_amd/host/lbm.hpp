// C++ drop-in layer over the chemsim C ABI, with the surface of src/lbm.rs:
// Scalar, Vector, Discretization, Direction, Geometry, Population(s),
// compute_equilibrium, the Lattice interface, D2Q9, the CollisionOperator
// plugin interface, BGK and State.  The populations live on the GPU (one
// cs_lattice per D2Q9); State::step runs the fused HIP step when the
// collision operator is BGK and the three separate phases otherwise, with a
// user operator evaluated on host populations.
//
// Errors: the reference's arrayfire-rust panics when an ArrayFire call fails;
// here every failing C-ABI call throws lbm::Error with cs_last_error().
#pragma once

#include <chemsim.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "matrix.hpp"

namespace chemsim {
namespace lbm {

using matrix::Matrix;
using Scalar = float;  // lbm.rs:13

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

inline void check(int rc)
{
    if (rc != CS_OK) throw Error(rc, std::string("chemsim: ") + cs_last_error());
}

// ---- device --------------------------------------------------------------

// The GPU every lattice of the process runs on (af::set_device + af::init,
// main.rs:331-332).  Call set_device before the first lattice to pick one.
class Device {
public:
    static void set_device(int index) { instance().select(index); }
    static cs_ctx* get() { return instance().ctx(); }

private:
    static Device& instance()
    {
        static Device d;
        return d;
    }
    void select(int index)
    {
        if (ctx_ && index == index_) return;
        if (ctx_) cs_ctx_destroy(ctx_);
        ctx_ = nullptr;
        index_ = index;
        check(cs_ctx_create(index_, &ctx_));
    }
    cs_ctx* ctx()
    {
        if (!ctx_) select(index_);
        return ctx_;
    }
    ~Device()
    {
        if (ctx_) cs_ctx_destroy(ctx_);
    }
    cs_ctx* ctx_ = nullptr;
    int index_ = 0;
};

inline void set_device(int index) { Device::set_device(index); }

// ---- small types (lbm.rs:17-95) ------------------------------------------

struct Vector {
    Scalar x = 0, y = 0;
    std::pair<Scalar, Scalar> to_pair() const { return {x, y}; }
    Vector operator+(Vector r) const { return {x + r.x, y + r.y}; }
    bool operator==(const Vector& r) const { return x == r.x && y == r.y; }
};

struct Discretization {
    Scalar delta_x = 1, delta_t = 1;
    // lbm.rs:81-86: dx / (sqrt(3) dt) in f32
    Scalar isothermal_speed_of_sound() const { return delta_x / (std::sqrt(Scalar(3)) * delta_t); }
};

// A lattice direction: weight, velocity and the 3x3 one-hot streaming
// stencil (lbm.rs:90-95, 202-282).  The stencil's 1 sits at row 1 + cy,
// column 1 - cx (row-major), which is what fixes the streaming orientation
// f_i(r, c) <- f_i(r + cx, c - cy).
struct Direction {
    int index = 0;
    Scalar w_scalar = 0;
    Vector c_vector;
    Matrix stencil;
};

// Solid mask, h x w, element (x, y) at [y * w + x] (lbm.rs:99, built as in
// main.rs:269-312).  set() bumps a version so a State uploads edits made
// between steps (main.rs:77-89) before its next step.
class Geometry {
public:
    Geometry() = default;
    Geometry(size_t w, size_t h) : w_(w), h_(h), cells_(w * h, 0) {}
    static Geometry from_row_major(const std::vector<bool>& solid, size_t w, size_t h)
    {
        if (solid.size() != w * h) throw matrix::Error();
        Geometry g(w, h);
        for (size_t i = 0; i < solid.size(); ++i) g.cells_[i] = solid[i] ? 1 : 0;
        return g;
    }
    bool get(size_t x, size_t y) const { return cells_.at(y * w_ + x) != 0; }
    void set(size_t x, size_t y, bool solid)
    {
        cells_.at(y * w_ + x) = solid ? 1 : 0;
        ++version_;
    }
    std::pair<size_t, size_t> size() const { return {w_, h_}; }
    const std::vector<uint8_t>& flags() const { return cells_; }
    uint64_t version() const { return version_; }

private:
    size_t w_ = 0, h_ = 0;
    std::vector<uint8_t> cells_;
    uint64_t version_ = 1;
};

using Population = Matrix;                                  // lbm.rs:103
using Populations = std::vector<std::pair<Direction, Population>>;  // lbm.rs:107

namespace detail {
constexpr int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

inline std::vector<float> flatten(const Populations& p, size_t w, size_t h)
{
    if (p.size() != 9) throw std::invalid_argument("lbm: D2Q9 needs 9 populations");
    std::vector<float> soa(9 * w * h);
    for (size_t i = 0; i < 9; ++i) {
        if (p[i].second.get_shape() != std::make_pair(w, h)) throw std::invalid_argument("lbm: population shape");
        std::copy(p[i].second.data().begin(), p[i].second.data().end(), soa.begin() + i * w * h);
    }
    return soa;
}
}  // namespace detail

// D2Q9::directions (lbm.rs:202-282).
inline std::vector<Direction> directions()
{
    std::vector<Direction> d(9);
    volatile Scalar n0 = 16, n1 = 4, n5 = 1, den = 36;  // RN(16/36), RN(4/36), RN(1/36)
    for (int i = 0; i < 9; ++i) {
        d[i].index = i;
        d[i].w_scalar = i == 0 ? n0 / den : (i < 5 ? n1 / den : n5 / den);
        d[i].c_vector = {Scalar(detail::CX[i]), Scalar(detail::CY[i])};
        std::vector<float> st(9, 0.f);
        st[3 * (1 + detail::CY[i]) + (1 - detail::CX[i])] = 1.f;
        d[i].stencil = Matrix::new_(st, {3, 3});
    }
    return d;
}

namespace detail {
inline Populations wrap(const std::vector<float>& soa, size_t w, size_t h)
{
    const auto dirs = directions();
    Populations p;
    p.reserve(9);
    for (size_t i = 0; i < 9; ++i)
        p.emplace_back(dirs[i], Matrix::new_(std::vector<float>(soa.begin() + i * w * h,
                                                                soa.begin() + (i + 1) * w * h),
                                             {w, h}));
    return p;
}

// RAII handle of one device lattice.
struct LatticeHandle {
    cs_lattice* p = nullptr;
    LatticeHandle(size_t w, size_t h, int storage)
    {
        check(cs_lattice_create(Device::get(), (int)w, (int)h, storage, &p));
    }
    ~LatticeHandle()
    {
        if (p) cs_lattice_destroy(p);
    }
    LatticeHandle(const LatticeHandle&) = delete;
    LatticeHandle& operator=(const LatticeHandle&) = delete;
};

inline cs_params params_for(const Discretization& d, Scalar tau)
{
    cs_params p{};
    p.tau = tau;
    p.delta_x = d.delta_x;
    p.delta_t = d.delta_t;
    p.edge = CS_EDGE_ZERO;
    p.boundary = CS_BB_REFERENCE;
    p.outlet_rho = 1.0;
    p.rho0 = 1.0;
    p.collision = CS_COLL_BGK;
    p.tau_minus = tau;
    return p;
}
}  // namespace detail

// compute_equilibrium (lbm.rs:43-71), evaluated on the GPU with the
// reference's f32 operation order.  `dirs` must be the D2Q9 directions.
inline Populations compute_equilibrium(const Matrix& density, const std::pair<Matrix, Matrix>& velocity,
                                       const std::vector<Direction>& dirs, Discretization disc)
{
    const auto shape = density.get_shape();
    if (velocity.first.get_shape() != shape || velocity.second.get_shape() != shape)
        throw std::invalid_argument("lbm: compute_equilibrium shape mismatch");
    if (dirs.size() != 9) throw std::invalid_argument("lbm: compute_equilibrium needs the D2Q9 directions");
    const auto [w, h] = shape;
    detail::LatticeHandle lat(w, h, CS_F32);
    const cs_params p = detail::params_for(disc, 1.0f);
    check(cs_lattice_set_params(lat.p, &p));
    check(cs_lattice_init_equilibrium(lat.p, density.data().data(), velocity.first.data().data(),
                                      velocity.second.data().data()));
    std::vector<float> soa(9 * w * h);
    check(cs_lattice_download_f32(lat.p, soa.data()));
    return detail::wrap(soa, w, h);
}

// ---- Lattice (lbm.rs:111-176) ------------------------------------------------

class Lattice {
public:
    virtual ~Lattice() = default;
    virtual std::pair<size_t, size_t> size() const = 0;
    virtual const Populations& populations() const = 0;
    // Host view for editing; the edits reach the device before the next step.
    virtual Populations& populations_mut() = 0;
    virtual Populations swap_populations() const = 0;
    virtual Matrix density() const = 0;
    virtual std::pair<Matrix, Matrix> momentum_density() const = 0;
    virtual std::pair<Matrix, Matrix> velocity() const = 0;
    virtual Matrix speed() const = 0;
    virtual Populations equilibrium(const Discretization& disc) const = 0;
    virtual Populations non_equilibrium(const Discretization& disc) const = 0;
};

class D2Q9;

// ---- CollisionOperator (lbm.rs:327-341) --------------------------------------

struct BGK;

template <class L>
class CollisionOperator {
public:
    virtual ~CollisionOperator() = default;
    virtual Populations evaluate(const L& lattice, const Populations& equilibrium,
                                 const Discretization& disc) const = 0;
    virtual Scalar kinematic_shear_viscosity(const Discretization& disc) const = 0;
    // Operators that are BGK expose it.
    virtual const BGK* as_bgk() const { return nullptr; }
    // Operators the device step implements (BGK, TRT) fill in their
    // parameters and return true; State then runs the fused device step.
    virtual bool device_params(cs_params&) const { return false; }
};

// D2Q9 with device-resident populations (lbm.rs:180-200, 285-325).
class D2Q9 final : public Lattice {
public:
    // D2Q9::new (lbm.rs:187-200): nine equally shaped populations in
    // direction order.  `storage` selects the device representation
    // (cs_storage; CS_F32 is the reference's).
    explicit D2Q9(const std::vector<Population>& pops, int storage = CS_F32)
    {
        if (pops.size() != 9) throw std::invalid_argument("lbm: D2Q9::new needs 9 populations");
        size_ = pops[0].get_shape();
        for (const auto& p : pops)
            if (p.get_shape() != size_) throw std::invalid_argument("lbm: D2Q9::new population shapes differ");
        lat_ = std::make_unique<detail::LatticeHandle>(size_.first, size_.second, storage);
        const auto dirs = directions();
        host_.clear();
        for (int i = 0; i < 9; ++i) host_.emplace_back(dirs[i], pops[i]);
        state_ = HostDirty;
        params_ = detail::params_for(Discretization{}, 1.0f);
    }
    static std::vector<Direction> directions() { return lbm::directions(); }

    std::pair<size_t, size_t> size() const override { return size_; }
    const Populations& populations() const override
    {
        if (state_ == Device) {
            std::vector<float> soa(9 * cells());
            check(cs_lattice_download_f32(lat_->p, soa.data()));
            host_ = detail::wrap(soa, size_.first, size_.second);
            state_ = HostClean;
        }
        return host_;
    }
    Populations& populations_mut() override
    {
        populations();
        state_ = HostDirty;
        return host_;
    }
    // swap_populations (lbm.rs:298-309): direction i carries population opp(i).
    Populations swap_populations() const override
    {
        const Populations& p = populations();
        Populations s;
        for (int i = 0; i < 9; ++i) s.emplace_back(p[i].first, p[detail::OPP[i]].second);
        return s;
    }
    Matrix density() const override { return field(cs_lattice_density); }
    std::pair<Matrix, Matrix> momentum_density() const override { return fields(cs_lattice_momentum_density); }
    std::pair<Matrix, Matrix> velocity() const override { return fields(cs_lattice_velocity); }
    Matrix speed() const override { return field(cs_lattice_speed); }
    // rho cs^2 with the parameters' discretization (State::pressure)
    Matrix pressure(const Discretization& disc) const
    {
        cs_params p = params_;
        p.delta_x = disc.delta_x;
        p.delta_t = disc.delta_t;
        set_params(p);
        return field(cs_lattice_pressure);
    }
    Populations equilibrium(const Discretization& disc) const override
    {
        return pops_of(disc, cs_lattice_equilibrium);
    }
    Populations non_equilibrium(const Discretization& disc) const override
    {
        return pops_of(disc, cs_lattice_non_equilibrium);
    }

    // ---- used by State -------------------------------------------------
    cs_lattice* handle() const
    {
        flush();
        return lat_->p;
    }
    void set_params(const cs_params& p) const
    {
        if (std::memcmp(&p, &params_, sizeof p) != 0 || !params_set_) {
            check(cs_lattice_set_params(lat_->p, &p));
            params_ = p;
            params_set_ = true;
        }
    }
    // The device is about to change the populations: the host copy goes stale.
    void invalidate_host() const { state_ = Device; }

private:
    enum Sync { Device, HostClean, HostDirty };
    size_t cells() const { return size_.first * size_.second; }
    void flush() const
    {
        if (state_ == HostDirty) {
            const auto soa = detail::flatten(host_, size_.first, size_.second);
            check(cs_lattice_upload_f32(lat_->p, soa.data()));
            state_ = HostClean;
        }
    }
    Matrix field(int (*fn)(cs_lattice*, float*)) const
    {
        Matrix m = Matrix::new_filled(0.f, size_);
        check(fn(handle(), m.data().data()));
        return m;
    }
    std::pair<Matrix, Matrix> fields(int (*fn)(cs_lattice*, float*, float*)) const
    {
        Matrix a = Matrix::new_filled(0.f, size_), b = Matrix::new_filled(0.f, size_);
        check(fn(handle(), a.data().data(), b.data().data()));
        return {std::move(a), std::move(b)};
    }
    Populations pops_of(const Discretization& disc, int (*fn)(cs_lattice*, float*)) const
    {
        cs_params p = params_;
        p.delta_x = disc.delta_x;
        p.delta_t = disc.delta_t;
        set_params(p);
        std::vector<float> soa(9 * cells());
        check(fn(handle(), soa.data()));
        return detail::wrap(soa, size_.first, size_.second);
    }

    std::pair<size_t, size_t> size_;
    std::unique_ptr<detail::LatticeHandle> lat_;
    mutable Populations host_;
    mutable Sync state_ = Device;
    mutable cs_params params_{};
    mutable bool params_set_ = false;
};

// BGK (lbm.rs:345-370).
struct BGK final : CollisionOperator<D2Q9> {
    Scalar tau;
    explicit BGK(Scalar t) : tau(t) {}
    // f + (f - f_eq) * (-dt / tau) on the GPU over the given populations.
    Populations evaluate(const D2Q9& lattice, const Populations& eq, const Discretization& disc) const override
    {
        const auto [w, h] = lattice.size();
        const auto f = detail::flatten(lattice.populations(), w, h);
        const auto fe = detail::flatten(eq, w, h);
        std::vector<float> out(f.size());
        check(cs_bgk_evaluate(Device::get(), (int64_t)f.size(), f.data(), fe.data(), tau, disc.delta_t, out.data()));
        return detail::wrap(out, w, h);
    }
    // lbm.rs:366-369: (dx^2 / (3 dt^2)) (tau - dt/2)
    Scalar kinematic_shear_viscosity(const Discretization& d) const override
    {
        return (d.delta_x * d.delta_x / (Scalar(3) * d.delta_t * d.delta_t)) * (tau - d.delta_t / Scalar(2));
    }
    const BGK* as_bgk() const override { return this; }
    bool device_params(cs_params& p) const override
    {
        p.collision = CS_COLL_BGK;
        p.tau = tau;
        p.tau_minus = tau;
        return true;
    }
};

// TRT (lbm.rs:374-451), run by the device step with the reference's
// arithmetic (swap_equilibrium included).
namespace detail {
// One collision of the lattice's populations on the device with the operator's
// cs_params (the evaluate() of every operator the device implements).
inline Populations collide_on_device(const CollisionOperator<D2Q9>& op, const D2Q9& lattice,
                                     const Discretization& disc)
{
    const auto [w, h] = lattice.size();
    LatticeHandle tmp(w, h, CS_F32);
    cs_params p = params_for(disc, 1.0f);
    op.device_params(p);
    check(cs_lattice_set_params(tmp.p, &p));
    const auto f = flatten(lattice.populations(), w, h);
    check(cs_lattice_upload_f32(tmp.p, f.data()));
    check(cs_lattice_collide(tmp.p));
    std::vector<float> out(f.size());
    check(cs_lattice_download_f32(tmp.p, out.data()));
    return wrap(out, w, h);
}
}  // namespace detail

struct TRT final : CollisionOperator<D2Q9> {
    Scalar tau_minus, tau_plus;
    TRT(Scalar tm, Scalar tp) : tau_minus(tm), tau_plus(tp) {}
    // TRT::new (lbm.rs:380-390)
    static TRT new_(Scalar lambda, Scalar ks_viscosity, const Discretization& disc)
    {
        float tp = 0.f, tm = 0.f;
        check(cs_trt_new(lambda, ks_viscosity, disc.delta_x, disc.delta_t, &tp, &tm));
        return TRT(tm, tp);
    }
    // lbm.rs:392-399
    Scalar lambda(const Discretization& disc) const
    {
        const Scalar dt = disc.delta_t;
        Scalar result = 1.0f;
        result *= ((tau_plus / dt) - 0.5f);
        result *= ((tau_minus / dt) - 0.5f);
        return result;
    }
    // On the device: the lattice's populations through the TRT collide kernel
    // (the passed equilibrium is unused, as in lbm.rs:411-412).
    Populations evaluate(const D2Q9& lattice, const Populations&, const Discretization& disc) const override
    {
        return detail::collide_on_device(*this, lattice, disc);
    }
    // lbm.rs:446-450: cs^2 (tau_plus / dt - 1/2)
    Scalar kinematic_shear_viscosity(const Discretization& disc) const override
    {
        const Scalar cs = disc.isothermal_speed_of_sound();
        return cs * cs * (tau_plus / disc.delta_t - 0.5f);
    }
    bool device_params(cs_params& p) const override
    {
        p.collision = CS_COLL_TRT;
        p.tau = tau_plus;
        p.tau_minus = tau_minus;
        return true;
    }
};

// KBC (lbm.rs:458-590): entropic multi-relaxation with the kinematic shear
// viscosity as its only parameter.  evaluate() runs the KBC collide kernel on
// the lattice's populations (the passed equilibrium is unused, lbm.rs:475-476);
// the reference's diagnostic print of an entropy sum (lbm.rs:574-582) is not made.
struct KBC final : CollisionOperator<D2Q9> {
    Scalar ks_viscosity;
    explicit KBC(Scalar nu) : ks_viscosity(nu) {}
    static KBC new_(Scalar nu) { return KBC(nu); }  // KBC::new, lbm.rs:462-466
    Populations evaluate(const D2Q9& lattice, const Populations&, const Discretization& disc) const override
    {
        return detail::collide_on_device(*this, lattice, disc);
    }
    Scalar kinematic_shear_viscosity(const Discretization&) const override { return ks_viscosity; }  // lbm.rs:586-588
    bool device_params(cs_params& p) const override
    {
        p.collision = CS_COLL_KBC;
        p.ks_viscosity = ks_viscosity;
        return true;
    }
};

// Regularized<C> (lbm.rs:592-666): Latt & Chopard's regularized collision.  As in
// the reference, evaluate() never calls the wrapped operator; it only supplies
// the viscosity (lbm.rs:662-664).
template <class C>
struct Regularized final : CollisionOperator<D2Q9> {
    C underlying;
    explicit Regularized(C u) : underlying(std::move(u)) {}
    static Regularized new_(C u) { return Regularized(std::move(u)); }  // lbm.rs:600-604
    Populations evaluate(const D2Q9& lattice, const Populations&, const Discretization& disc) const override
    {
        return detail::collide_on_device(*this, lattice, disc);
    }
    Scalar kinematic_shear_viscosity(const Discretization& disc) const override
    {
        return underlying.kinematic_shear_viscosity(disc);
    }
    bool device_params(cs_params& p) const override
    {
        p.collision = CS_COLL_REGULARIZED;
        return true;
    }
};

// ---- State (lbm.rs:670-819) --------------------------------------------------

template <class L = D2Q9>
class State {
public:
    Scalar time = 0;
    std::unique_ptr<L> lattice;
    Geometry geometry;
    std::unique_ptr<CollisionOperator<L>> collision;
    Discretization discretization;

    // State::initial (lbm.rs:679-692)
    static State initial(std::unique_ptr<L> lattice, Geometry geometry,
                         std::unique_ptr<CollisionOperator<L>> collision, Discretization disc)
    {
        if (geometry.size() != lattice->size()) throw std::invalid_argument("lbm: geometry and lattice sizes differ");
        State s;
        s.lattice = std::move(lattice);
        s.geometry = std::move(geometry);
        s.collision = std::move(collision);
        s.discretization = disc;
        return s;
    }

    // stream -> bounce_back -> collide, time += dt (lbm.rs:694-714); one fused
    // device pass when the operator is BGK.
    void step()
    {
        if (device_operator()) {
            prepare();
            // handle() uploads host edits first; only then is the host copy stale
            cs_lattice* h = lattice->handle();
            lattice->invalidate_host();
            check(cs_lattice_step(h, 1));
        } else {
            stream();
            bounce_back();
            collide();
        }
        time += discretization.delta_t;
    }
    // lbm.rs:716-729: f_i(r, c) <- f_i(r + cx, c - cy), zero fill.
    void stream()
    {
        prepare();
        cs_lattice* h = lattice->handle();
        lattice->invalidate_host();
        check(cs_lattice_stream(h));
    }
    // lbm.rs:741-751: f_i <- f_opp(i) at solid cells.
    void bounce_back()
    {
        prepare();
        cs_lattice* h = lattice->handle();
        lattice->invalidate_host();
        check(cs_lattice_bounce_back(h));
    }
    // lbm.rs:731-739: the operator over every cell with the lattice's equilibrium.
    void collide()
    {
        if (device_operator()) {
            prepare();
            cs_lattice* h = lattice->handle();
            lattice->invalidate_host();
            check(cs_lattice_collide(h));
        } else {
            Populations fs = collision->evaluate(*lattice, equilibrium(), discretization);
            lattice->populations_mut() = std::move(fs);
        }
    }

    std::pair<size_t, size_t> size() const { return lattice->size(); }
    Scalar delta_x() const { return discretization.delta_x; }
    Scalar delta_t() const { return discretization.delta_t; }
    const Populations& populations() const { return lattice->populations(); }
    Scalar isothermal_speed_of_sound() const { return discretization.isothermal_speed_of_sound(); }
    Matrix density() const { return lattice->density(); }
    // lbm.rs:784-787: density scaled by cs^2 (cs in f32, squared in f32)
    Matrix pressure() const { return lattice->pressure(discretization); }
    std::pair<Matrix, Matrix> momentum_density() const { return lattice->momentum_density(); }
    std::pair<Matrix, Matrix> velocity() const { return lattice->velocity(); }
    Matrix speed() const { return lattice->speed(); }
    Populations equilibrium() const { return lattice->equilibrium(discretization); }
    Populations non_equilibrium() const { return lattice->non_equilibrium(discretization); }
    // lbm.rs:815-818: min over cells of f_eq[0] < 0
    bool is_unstable() const
    {
        prepare();
        int u = 0;
        check(cs_lattice_is_unstable(lattice->handle(), &u));
        return u != 0;
    }

    // The device lattice with this state's parameters, geometry and host
    // edits uploaded, for device work that only reads the populations (the
    // render pass).  The host copy of the populations stays valid.
    cs_lattice* device_view() const
    {
        prepare();
        return lattice->handle();
    }
    // The same for device work that changes the populations (frame
    // recording): the host copy is stale from here on and is downloaded again
    // on the next host read.
    cs_lattice* device_lattice()
    {
        cs_lattice* h = device_view();
        lattice->invalidate_host();
        return h;
    }

private:
    bool device_operator() const
    {
        cs_params p{};
        return collision && collision->device_params(p);
    }
    // Parameters (the operator's when the device implements it, else BGK
    // with tau = 1 for the streaming and bounce-back passes) and geometry
    // reach the device before any device pass.
    void prepare() const
    {
        cs_params p = detail::params_for(discretization, Scalar(1));
        if (collision) collision->device_params(p);
        lattice->set_params(p);
        if (geometry.version() != uploaded_geometry_) {
            check(cs_lattice_set_flags(lattice->handle(), geometry.flags().data()));
            uploaded_geometry_ = geometry.version();
        }
    }
    mutable uint64_t uploaded_geometry_ = 0;
};

}  // namespace lbm
}  // namespace chemsim
