// pica.hpp -- C++ surface of libpica_amd with pica's type and method names.
//
// pica's hot path is a set of header-only templates instantiated in the caller
// (YeeGrid.h:191-256, YeeSolver.h:11-38, EnsembleSupercells.h:15-69, and the per-particle
// functors FieldInterpolatorCICSupercell / BorisPusher / CurrentDepositorCICSupercell).  This
// header keeps those names for the objects that live on the GPU and replaces the caller's
// per-particle loop with one call per ensemble (pushParticles / step).  Everything forwards to
// the C ABI of include/pica_amd.h; errors become pica_amd::Error exceptions on this side.
//
// Link: g++ -std=c++17 -I include your.cpp -L pica_amd -lpica_amd -Wl,-rpath,<dir of the .so>
#ifndef PICA_AMD_PICA_HPP
#define PICA_AMD_PICA_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../pica_amd.h"

namespace pica_amd {

// Vector3<T> of pica's Vectors.h:380-423, the subset this surface needs
template <class T>
struct Vector3 {
    T x, y, z;
    Vector3() : x(0), y(0), z(0) {}
    Vector3(T a, T b, T c) : x(a), y(b), z(c) {}
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
typedef Vector3<double> FP3;
typedef Vector3<int> Int3;

enum class FormFactor { CIC = PK_CIC, TSC = PK_TSC };
enum class Depositor { Direct = PK_DIRECT, Esirkepov = PK_ESIRKEPOV };
enum Component { Ex = PK_EX, Ey, Ez, Bx, By, Bz, Jx, Jy, Jz, Rho };

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

class Context {
public:
    explicit Context(int device = 0)
    {
        const int r = pk_ctx_create(device, &h_);
        if (r) throw Error(r, "pk_ctx_create failed");
    }
    ~Context() { pk_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    pk_ctx* handle() const { return h_; }
    void check(int r) const
    {
        if (r) throw Error(r, pk_last_error(h_));
    }
    void sync() const { check(pk_sync(h_)); }
    // library options (PK_OPT_PARTICLE_KERNEL, PK_OPT_CHANGE_QUEUE, PK_OPT_INTERLEAVE)
    void setOption(int option, int value) const { check(pk_ctx_set_option(h_, option, value)); }
    int getOption(int option) const
    {
        int v = 0;
        check(pk_ctx_get_option(h_, option, &v));
        return v;
    }
    // ParticleTypes::typesVector (Particle.h:26-31)
    void setSpecies(const std::vector<double>& mass, const std::vector<double>& charge) const
    {
        check(pk_species_set(h_, (int)mass.size(), mass.data(), charge.data()));
    }

private:
    pk_ctx* h_ = nullptr;
};

// the context of device 0 that pica_compat.hpp's objects share (created on first use, lives until exit)
inline Context& defaultContext()
{
    static Context ctx(0);
    return ctx;
}

// pica::YeeGrid<Three, Real>(origin, step, size) (YeeGrid.h:201-202); Real = double or float
template <class Real>
class YeeGrid {
    static_assert(std::is_same<Real, double>::value || std::is_same<Real, float>::value, "double or float");

public:
    typedef Real ValueType;
    YeeGrid(Context& ctx, const FP3& origin, const FP3& step, const Int3& size, int ghost = 0)
        : ctx_(ctx), origin_(origin), step_(step), size_(size)
    {
        const int s[3] = {size.x, size.y, size.z};
        const double o[3] = {origin.x, origin.y, origin.z}, h[3] = {step.x, step.y, step.z};
        ctx.check(pk_grid_create(ctx.handle(), std::is_same<Real, double>::value ? PK_F64 : PK_F32, s, o, h, ghost,
                                 &h_));
    }
    // ncell physical cells with two periodic ghost layers, origin -2 step (the layout the
    // supercell interpolator and depositor assume)
    static YeeGrid periodic(Context& ctx, const Int3& ncell, const FP3& step)
    {
        return YeeGrid(ctx, FP3(-2 * step.x, -2 * step.y, -2 * step.z), step,
                       Int3(ncell.x + 4, ncell.y + 4, ncell.z + 4), 2);
    }
    YeeGrid(YeeGrid&& o) noexcept : ctx_(o.ctx_), origin_(o.origin_), step_(o.step_), size_(o.size_), h_(o.h_)
    {
        o.h_ = nullptr;
    }
    ~YeeGrid()
    {
        if (h_) pk_grid_destroy(h_);
    }
    YeeGrid(const YeeGrid&) = delete;
    YeeGrid& operator=(const YeeGrid&) = delete;

    Int3 getSize() const { return size_; }
    FP3 getOrigin() const { return origin_; }
    FP3 getStep() const { return step_; }
    size_t numElements() const { return (size_t)size_.x * size_.y * size_.z; }
    // Array3d layout: z fastest (Array.h:143)
    void copyIn(Component c, const std::vector<Real>& a) { ctx_.check(pk_grid_copy(h_, c, (void*)a.data(), PK_H2D)); }
    std::vector<Real> copyOut(Component c) const
    {
        std::vector<Real> a(numElements());
        ctx_.check(pk_grid_copy(h_, c, a.data(), PK_D2H));
        return a;
    }
    void zero(Component c) { ctx_.check(pk_grid_zero(h_, c)); }
    void fillGhostsEB() { ctx_.check(pk_ghost(h_, PK_FILL_E | PK_FILL_B)); }
    // snapshot files (restart, fixtures): every component, ghosts included, bit for bit
    void save(const std::string& path) const { ctx_.check(pk_grid_save(h_, path.c_str())); }
    void load(const std::string& path) { ctx_.check(pk_grid_load(h_, path.c_str())); }
    pk_grid* handle() const { return h_; }
    Context& context() const { return ctx_; }

private:
    Context& ctx_;
    FP3 origin_, step_;
    Int3 size_;
    pk_grid* h_ = nullptr;
};

// pica::YeeSolver (YeeSolver.h:11-25)
class YeeSolver {
public:
    template <class Real>
    void updateE(YeeGrid<Real>& grid, Real dt)
    {
        grid.context().check(pk_yee_update_e(grid.handle(), (double)dt));
    }
    template <class Real>
    void updateB(YeeGrid<Real>& grid, Real dt)
    {
        grid.context().check(pk_yee_update_b(grid.handle(), (double)dt));
    }
};

// Structure-of-arrays particles on the host side of add() / getParticles()
struct ParticleArraySoA {
    std::vector<double> x, y, z, px, py, pz, factor;   // position, p/(mc) (ParticleRef::getP), weight
    size_t size() const { return factor.size(); }
    void pushBack(const FP3& pos, const FP3& p, double w)
    {
        x.push_back(pos.x); y.push_back(pos.y); z.push_back(pos.z);
        px.push_back(p.x); py.push_back(p.y); pz.push_back(p.z);
        factor.push_back(w);
    }
};

// pica::EnsembleSupercells(min, max, numCells, numCellsPerSupercell) (EnsembleSupercells.h:25-34);
// Real = double: absolute FP64 positions; Real = float: int32 cell + FP32 offset positions.
// The domain minimum must be 0 (pica's supercell classes assume grid origin -2 step).
template <class Real>
class EnsembleSupercells {
public:
    EnsembleSupercells(Context& ctx, const FP3& minPosition, const FP3& maxPosition, const Int3& numCells,
                       const Int3& numCellsPerSupercell, int typeIndex = 0)
        : ctx_(ctx), numCells_(numCells), ncps_(numCellsPerSupercell)
    {
        if (minPosition.x != 0 || minPosition.y != 0 || minPosition.z != 0)
            throw Error(PK_E_ARG, "EnsembleSupercells: the domain minimum must be 0");
        const int nc[3] = {numCells.x, numCells.y, numCells.z};
        const int np[3] = {numCellsPerSupercell.x, numCellsPerSupercell.y, numCellsPerSupercell.z};
        const double h[3] = {maxPosition.x / numCells.x, maxPosition.y / numCells.y, maxPosition.z / numCells.z};
        const bool f64 = std::is_same<Real, double>::value;
        ctx.check(pk_ens_create(ctx.handle(), f64 ? PK_F64 : PK_F32, f64 ? PK_POS_ABS : PK_POS_SPLIT, nc, h, np,
                                typeIndex, &h_));
    }
    ~EnsembleSupercells()
    {
        if (h_) pk_ens_destroy(h_);
    }
    EnsembleSupercells(const EnsembleSupercells&) = delete;
    EnsembleSupercells& operator=(const EnsembleSupercells&) = delete;

    // EnsembleSupercells::add for every particle of the array (replaces the contents)
    void add(const ParticleArraySoA& p)
    {
        ctx_.check(pk_ens_upload(h_, p.size(), p.x.data(), p.y.data(), p.z.data(), p.px.data(), p.py.data(),
                                 p.pz.data(), p.factor.data()));
    }
    void loadThermal(int ppc, double uth, double driftX, double weight, uint64_t seed)
    {
        ctx_.check(pk_ens_load_thermal(h_, ppc, uth, driftX, weight, seed));
    }
    // all particles in supercell order; offsets of the supercells in `offsets` (nsc + 1 entries)
    ParticleArraySoA getParticles(std::vector<int64_t>* offsets = nullptr)
    {
        ParticleArraySoA p;
        const int64_t n = size();
        for (auto* v : {&p.x, &p.y, &p.z, &p.px, &p.py, &p.pz, &p.factor}) v->resize((size_t)n);
        if (offsets) offsets->resize((size_t)pk_ens_num_supercells(h_) + 1);
        ctx_.check(pk_ens_download(h_, p.x.data(), p.y.data(), p.z.data(), p.px.data(), p.py.data(), p.pz.data(),
                                   p.factor.data(), offsets ? offsets->data() : nullptr));
        return p;
    }
    int64_t size()
    {
        const int64_t n = pk_ens_size(h_);
        if (n < 0) ctx_.check(PK_E_STATE);
        return n;
    }
    // reorder every supercell now (FP32: same-cell pairs in adjacent slots; FP64: no-op), as the
    // re-sort does every PK_OPT_INTERLEAVE re-sorts
    void interleave() { ctx_.check(pk_ens_interleave(h_)); }
    // snapshot files: the particles, bits and order; a load needs the same geometry and species
    void save(const std::string& path) { ctx_.check(pk_ens_save(h_, path.c_str())); }
    void load(const std::string& path) { ctx_.check(pk_ens_load(h_, path.c_str())); }
    Int3 getNumCells() const { return numCells_; }
    Int3 getNumCellsPerSupercell() const { return ncps_; }
    Int3 getNumSupercells() const
    {
        return Int3(numCells_.x / ncps_.x, numCells_.y / ncps_.y, numCells_.z / ncps_.z);
    }
    pk_ens* handle() const { return h_; }

private:
    Context& ctx_;
    Int3 numCells_, ncps_;
    pk_ens* h_ = nullptr;
};

// The caller's per-particle loop of FieldInterpolatorCICSupercell::get, BorisPusher::push and
// CurrentDepositorCICSupercell::deposit, for every particle, on the GPU; J accumulates into grid.
template <class Real>
void pushParticles(EnsembleSupercells<Real>& ens, YeeGrid<Real>& grid, double dt, FormFactor ff = FormFactor::CIC,
                   Depositor dep = Depositor::Direct)
{
    grid.context().check(pk_particle_step(ens.handle(), grid.handle(), dt, (int)ff, (int)dep));
    grid.context().check(pk_ens_resort(ens.handle()));
}

// A whole periodic PIC step: J = 0, pushParticles, fold J, B(dt/2), E(dt), B(dt/2), re-sort.
template <class Real>
void step(EnsembleSupercells<Real>& ens, YeeGrid<Real>& grid, double dt, FormFactor ff = FormFactor::CIC,
          Depositor dep = Depositor::Direct, int nsteps = 1)
{
    grid.context().check(pk_run(ens.handle(), grid.handle(), dt, (int)ff, (int)dep, nsteps));
}

// One step of several species (electrons, mobile ions, ...) on one grid: every ensemble's
// particles deposit into the same J before the field update.
template <class Real>
void step(std::vector<EnsembleSupercells<Real>*>& species, YeeGrid<Real>& grid, double dt,
          FormFactor ff = FormFactor::CIC, Depositor dep = Depositor::Direct)
{
    std::vector<pk_ens*> h;
    for (auto* e : species) h.push_back(e->handle());
    grid.context().check(pk_step_species(h.data(), (int)h.size(), grid.handle(), dt, (int)ff, (int)dep));
}

struct Diagnostics {
    double fieldEnergy, kineticEnergy, charge, gaussMax, gaussDrift, rhoMax;
};

template <class Real>
Diagnostics diagnostics(EnsembleSupercells<Real>& ens, YeeGrid<Real>& grid, FormFactor ff = FormFactor::CIC)
{
    double d[PK_DIAG_N];
    grid.context().check(pk_diag(ens.handle(), grid.handle(), (int)ff, d));
    return Diagnostics{d[PK_DIAG_FIELD_ENERGY], d[PK_DIAG_KINETIC_ENERGY], d[PK_DIAG_CHARGE], d[PK_DIAG_GAUSS_MAX],
                       d[PK_DIAG_GAUSS_DRIFT], d[PK_DIAG_RHO_MAX]};
}

}  // namespace pica_amd

#endif  // PICA_AMD_PICA_HPP
