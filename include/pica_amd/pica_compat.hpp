// pica_compat.hpp -- pica's own namespace and signatures over libpica_amd, so that a pica caller
// switches by changing its includes and replacing its per-particle loop with one call.
//
// What a pica caller writes stays as it is:
//   pica::YeeGrid<pica::Three, Real>(origin, step, size), grid.ex(i, j, k) ... jz(i, j, k),
//   grid.ex() (an Array3d with operator()), getOrigin / getStep / getSize / getCellIndex
//                                                                    (YeeGrid.h:191-256, Grid.h:12-83)
//   pica::YeeSolver::updateE / updateB                               (YeeSolver.h:11-38)
//   pica::Particle3d, ParticleTypes::typesVector / types / numTypes  (Particle.h:21-112)
//   ParticleArray<Particle3d, ParticleRepresentation_SoA>::Type with ParticleRef, pushBack,
//   operator[], size                                                 (ParticleArray.h:45-231)
//   Ensemble<PArray, EnsembleRepresentation_Supercells>::Type, i.e. EnsembleSupercells(min, max,
//   numCells, numCellsPerSupercell), add, getParticles(Int3), getSupercellIndex, getNumSupercells
//                                                                    (EnsembleSupercells.h:15-69)
// and the loop over supercells of FieldInterpolatorCICSupercell::get + BorisPusher::push +
// CurrentDepositorCICSupercell::deposit becomes
//   pica::pushParticles(ensemble, grid, dt [, form factor, depositor])
// which runs interpolation, push, deposition (J += into the grid, as pica's depositor flushes)
// and the migration of particles between supercells on the GPU.
//
// Data live on the GPU; the grid and the ensemble keep a host mirror for pica's element access.
// Coherence is automatic: element access brings the mirror up to date (one download) and marks it
// modified; the next GPU operation uploads a modified mirror first.  Access by element is
// therefore cheap in loops between GPU operations and costs one transfer per switch.
//
// The ParticleTypes table is pica's: the caller defines typesVector / types / numTypes as pica
// requires (Particle.h:26-31); it is passed to the library at every pushParticles.
//
// Not provided: the per-particle functors themselves (FieldInterpolatorCIC(Supercell),
// CurrentDepositorCIC(Supercell), BorisPusher::push) -- per-particle calls cannot cross to the GPU
// one particle at a time, so the loop that uses them is what pushParticles replaces.
// BorisPusher<Particle, Real>(dt) is provided so that callers that construct one still compile.
#ifndef PICA_AMD_PICA_COMPAT_HPP
#define PICA_AMD_PICA_COMPAT_HPP

#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "pica.hpp"

namespace pica {

enum Dimension { One = 1, Two = 2, Three = 3 };

// Vector3<T> with the arithmetic pica callers use on positions and momenta (Vectors.h:380-697)
template <class T>
struct Vector3 {
    T x, y, z;
    Vector3() : x(0), y(0), z(0) {}
    Vector3(T a, T b, T c) : x(a), y(b), z(c) {}
    template <class U>
    explicit Vector3(const Vector3<U>& o) : x((T)o.x), y((T)o.y), z((T)o.z) {}
    T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    T norm2() const { return x * x + y * y + z * z; }
    T norm() const { return std::sqrt(norm2()); }
    T volume() const { return x * y * z; }
};
template <class T> Vector3<T> operator+(const Vector3<T>& a, const Vector3<T>& b) { return Vector3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> Vector3<T> operator-(const Vector3<T>& a, const Vector3<T>& b) { return Vector3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> Vector3<T> operator*(const Vector3<T>& a, const Vector3<T>& b) { return Vector3<T>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class T> Vector3<T> operator/(const Vector3<T>& a, const Vector3<T>& b) { return Vector3<T>(a.x / b.x, a.y / b.y, a.z / b.z); }
template <class T> Vector3<T> operator*(const Vector3<T>& a, T s) { return Vector3<T>(a.x * s, a.y * s, a.z * s); }
template <class T> Vector3<T> operator*(T s, const Vector3<T>& a) { return Vector3<T>(a.x * s, a.y * s, a.z * s); }
template <class T> Vector3<T> operator/(const Vector3<T>& a, T s) { return Vector3<T>(a.x / s, a.y / s, a.z / s); }
typedef Vector3<double> FP3;
typedef Vector3<int> Int3;

// truncate / floor of a position (Vectors.h:652-679)
inline Int3 truncate(const FP3& v) { return Int3((int)v.x, (int)v.y, (int)v.z); }
inline Int3 floor(const FP3& v) { return Int3((int)std::floor(v.x), (int)std::floor(v.y), (int)std::floor(v.z)); }

namespace constants {
const double c = 29979245800.0;          // Constants.h:11
}

// ---- ParticleTypes (Particle.h:21-31): defined by the caller, as pica requires
struct ParticleType {
    double mass;
    double charge;
};
namespace ParticleTypes {
extern std::vector<ParticleType> typesVector;
extern const ParticleType* types;
extern int numTypes;
}  // namespace ParticleTypes

// ---- Particle<Three> (Particle.h:33-112): position, p = momentum / (m c), factor, type
template <Dimension dimension>
class Particle;

template <>
class Particle<Three> {
public:
    typedef double Real;
    typedef FP3 PositionType;
    typedef FP3 MomentumType;
    typedef double GammaType, MassType, ChargeType, FactorType;
    typedef short TypeIndexType;
    Particle() : factor(1), invGamma(1), typeIndex(0) {}
    Particle(const PositionType& position, const MomentumType& momentum, FactorType f = 1, TypeIndexType type = 0)
        : position(position), factor(f), typeIndex(type)
    {
        setMomentum(momentum);
    }
    PositionType getPosition() const { return position; }
    void setPosition(const PositionType& v) { position = v; }
    MomentumType getMomentum() const { return p * (constants::c * getMass()); }
    void setMomentum(const MomentumType& m) { setP(m / (constants::c * getMass())); }
    MomentumType getP() const { return p; }
    void setP(const MomentumType& v)
    {
        p = v;
        invGamma = 1.0 / std::sqrt(1.0 + p.norm2());
    }
    MomentumType getVelocity() const { return p * (constants::c * invGamma); }
    GammaType getGamma() const { return 1.0 / invGamma; }
    MassType getMass() const { return ParticleTypes::types[typeIndex].mass; }
    ChargeType getCharge() const { return ParticleTypes::types[typeIndex].charge; }
    FactorType getFactor() const { return factor; }
    void setFactor(FactorType f) { factor = f; }
    TypeIndexType getType() const { return typeIndex; }
    void setType(TypeIndexType t) { typeIndex = t; }

private:
    PositionType position;
    MomentumType p;
    FactorType factor;
    GammaType invGamma;
    TypeIndexType typeIndex;
};
typedef Particle<Three> Particle3d;

// ---- ParticleArraySoA (ParticleArray.h:165-216): host storage with ParticleRef access
template <class ParticleT>
class ParticleArraySoA {
public:
    typedef ParticleT Particle;
    typedef FP3 PositionType;
    typedef FP3 MomentumType;

    class ConstParticleRef {
    public:
        typedef FP3 PositionType;
        typedef FP3 MomentumType;
        typedef double GammaType, FactorType;
        typedef short TypeIndexType;
        ConstParticleRef(const ParticleArraySoA& a, int i) : a_(a), i_(i) {}
        PositionType getPosition() const { return FP3(a_.pos[0][i_], a_.pos[1][i_], a_.pos[2][i_]); }
        MomentumType getP() const { return FP3(a_.ps[0][i_], a_.ps[1][i_], a_.ps[2][i_]); }
        MomentumType getMomentum() const { return getP() * (constants::c * getMass()); }
        MomentumType getVelocity() const { return getP() * (constants::c * a_.invGammas[i_]); }
        GammaType getGamma() const { return 1.0 / a_.invGammas[i_]; }
        double getMass() const { return ParticleTypes::types[a_.types[i_]].mass; }
        double getCharge() const { return ParticleTypes::types[a_.types[i_]].charge; }
        FactorType getFactor() const { return a_.factors[i_]; }
        TypeIndexType getType() const { return a_.types[i_]; }

    private:
        const ParticleArraySoA& a_;
        int i_;
    };

    class ParticleRef : public ConstParticleRef {
    public:
        ParticleRef(ParticleArraySoA& a, int i) : ConstParticleRef(a, i), a_(a), i_(i) {}
        void setPosition(const FP3& v)
        {
            for (int d = 0; d < 3; d++) a_.pos[d][i_] = v[d];
            a_.touch();
        }
        void setP(const FP3& p)
        {
            for (int d = 0; d < 3; d++) a_.ps[d][i_] = p[d];
            a_.invGammas[i_] = 1.0 / std::sqrt(1.0 + p.norm2());
            a_.touch();
        }
        void setMomentum(const FP3& m) { setP(m / (constants::c * ParticleTypes::types[a_.types[i_]].mass)); }
        void setFactor(double f)
        {
            a_.factors[i_] = f;
            a_.touch();
        }
        void setType(short t)
        {
            a_.types[i_] = t;
            a_.touch();
        }

    private:
        ParticleArraySoA& a_;
        int i_;
    };

    int size() const { return (int)factors.size(); }
    ParticleRef operator[](int i) { return ParticleRef(*this, i); }
    ConstParticleRef operator[](int i) const { return ConstParticleRef(*this, i); }
    ParticleRef back() { return (*this)[size() - 1]; }
    template <class Ref>
    void pushBack(const Ref& p)
    {
        const FP3 x = p.getPosition(), u = p.getP();
        for (int d = 0; d < 3; d++) {
            pos[d].push_back(x[d]);
            ps[d].push_back(u[d]);
        }
        factors.push_back(p.getFactor());
        invGammas.push_back(1.0 / p.getGamma());
        types.push_back(p.getType());
        touch();
    }
    void popBack()
    {
        for (int d = 0; d < 3; d++) {
            pos[d].pop_back();
            ps[d].pop_back();
        }
        factors.pop_back();
        invGammas.pop_back();
        types.pop_back();
        touch();
    }

private:
    std::vector<double> pos[3], ps[3], factors, invGammas;
    std::vector<short> types;
    bool* dirty_ = nullptr;     // the owning ensemble's "mirror modified" flag, if any
    void touch()
    {
        if (dirty_) *dirty_ = true;
    }
    template <class PA>
    friend class EnsembleSupercells;
};

enum ParticleRepresentation { ParticleRepresentation_AoS, ParticleRepresentation_SoA };
template <class P, ParticleRepresentation rep>
struct ParticleArray {};
template <class P>
struct ParticleArray<P, ParticleRepresentation_SoA> {
    typedef ParticleArraySoA<P> Type;
};

// ---- YeeGrid<Three, Real> (YeeGrid.h:191-256) over a pica_amd grid with a coherent host mirror
template <class Real>
class Array3d;

template <Dimension dimension, class Real = double>
class YeeGrid;

template <class Real>
class YeeGrid<Three, Real> {
    static_assert(std::is_same<Real, double>::value || std::is_same<Real, float>::value, "double or float");

public:
    typedef Real ValueType;
    typedef FP3 PositionType;
    typedef Int3 IndexType;
    typedef pica::Array3d<Real> ArrayType;

    YeeGrid(const PositionType& origin, const PositionType& step, const IndexType& size)
        : origin_(origin), step_(step), size_(size),
          grid_(pica_amd::defaultContext(), pica_amd::FP3(origin.x, origin.y, origin.z),
                pica_amd::FP3(step.x, step.y, step.z), pica_amd::Int3(size.x, size.y, size.z), ghostOf(origin, step, size))
    {
        for (int c = 0; c < 9; c++) arrays_[c].reset(new ArrayType(this, c));
    }
    YeeGrid(const YeeGrid&) = delete;
    YeeGrid& operator=(const YeeGrid&) = delete;

    PositionType getOrigin() const { return origin_; }
    PositionType getStep() const { return step_; }
    IndexType getSize() const { return size_; }
    // Grid::getCellIndex (Grid.h:36): truncate((x - origin) / step)
    IndexType getCellIndex(const PositionType& x) const { return truncate((x - origin_) / step_); }

#define PICA_AMD_COMPONENT(name, comp)                                                                       \
    Real& name(int i, int j, int k) { return mutableHost(comp)[index(i, j, k)]; }                        \
    Real name(int i, int j, int k) const { return constHost(comp)[index(i, j, k)]; }                    \
    ArrayType& name() { return *arrays_[comp]; }                                                          \
    const ArrayType& name() const { return *arrays_[comp]; }
    PICA_AMD_COMPONENT(ex, 0)
    PICA_AMD_COMPONENT(ey, 1)
    PICA_AMD_COMPONENT(ez, 2)
    PICA_AMD_COMPONENT(bx, 3)
    PICA_AMD_COMPONENT(by, 4)
    PICA_AMD_COMPONENT(bz, 5)
    PICA_AMD_COMPONENT(jx, 6)
    PICA_AMD_COMPONENT(jy, 7)
    PICA_AMD_COMPONENT(jz, 8)
#undef PICA_AMD_COMPONENT

    // the device grid, current (a modified host mirror is uploaded first); the caller is about to
    // change it on the GPU, so the mirror goes stale
    pica_amd::YeeGrid<Real>& device()
    {
        flush();
        hostValid_ = false;
        return grid_;
    }
    // upload a modified host mirror now (also done before every GPU operation)
    void flush()
    {
        if (!hostDirty_) return;
        for (int c = 0; c < 9; c++) grid_.copyIn((pica_amd::Component)c, host_[c]);
        hostDirty_ = false;
    }

private:
    friend class pica::Array3d<Real>;
    static int ghostOf(const PositionType& origin, const PositionType& step, const IndexType& size)
    {
        // pica's supercell classes assume grid origin -2 step; such a grid gets two periodic ghost
        // layers, so that pushParticles can run on it
        const bool two = origin.x == -2 * step.x && origin.y == -2 * step.y && origin.z == -2 * step.z &&
                         size.x > 4 && size.y > 4 && size.z > 4;
        return two ? 2 : 0;
    }
    size_t index(int i, int j, int k) const { return ((size_t)i * size_.y + j) * size_.z + k; }
    void pull() const
    {
        if (hostValid_) return;
        for (int c = 0; c < 9; c++) host_[c] = grid_.copyOut((pica_amd::Component)c);
        hostValid_ = true;
    }
    const std::vector<Real>& constHost(int c) const
    {
        pull();
        return host_[c];
    }
    std::vector<Real>& mutableHost(int c)
    {
        pull();
        hostDirty_ = true;
        return host_[c];
    }

    PositionType origin_, step_;
    IndexType size_;
    pica_amd::YeeGrid<Real> grid_;
    mutable std::vector<Real> host_[9];
    mutable bool hostValid_ = false;
    bool hostDirty_ = false;
    std::unique_ptr<ArrayType> arrays_[9];
};

// Array3d view of one component (Array.h:118-167): operator()(i, j, k), getSize, assignment of a value
template <class Real>
class Array3d {
public:
    Array3d(YeeGrid<Three, Real>* g, int comp) : g_(g), comp_(comp) {}
    Real& operator()(int i, int j, int k) { return g_->mutableHost(comp_)[g_->index(i, j, k)]; }
    Real operator()(int i, int j, int k) const { return g_->constHost(comp_)[g_->index(i, j, k)]; }
    Real& operator()(const Int3& i) { return (*this)(i.x, i.y, i.z); }
    Real operator()(const Int3& i) const { return (*this)(i.x, i.y, i.z); }
    Int3 getSize() const { return g_->getSize(); }
    Array3d& operator=(Real v)
    {
        std::vector<Real>& a = g_->mutableHost(comp_);
        for (auto& x : a) x = v;
        return *this;
    }

private:
    YeeGrid<Three, Real>* g_;
    int comp_;
};

// ---- YeeSolver (YeeSolver.h:11-38)
class YeeSolver {
public:
    template <Dimension d, class Real>
    void updateE(YeeGrid<d, Real>& grid, Real dt)
    {
        pica_amd::YeeSolver().updateE(grid.device(), dt);
    }
    template <Dimension d, class Real>
    void updateB(YeeGrid<d, Real>& grid, Real dt)
    {
        pica_amd::YeeSolver().updateB(grid.device(), dt);
    }
};

// ---- BorisPusher<Particle, Real>(dt) (BorisPusher.h:13-30): the time step; push() is replaced by
// pushParticles for the whole ensemble
template <class ParticleT, class Real>
class BorisPusher {
public:
    explicit BorisPusher(Real dt) : dt_(dt) {}
    Real getDt() const { return dt_; }

private:
    Real dt_;
};

// ---- EnsembleSupercells<ParticleArraySoA<Particle3d>> (EnsembleSupercells.h:15-69) over a
// pica_amd ensemble (FP64 positions) with a per-supercell host mirror
enum EnsembleRepresentation { EnsembleRepresentation_Unordered, EnsembleRepresentation_Ordered,
                              EnsembleRepresentation_Supercells };

template <class PA>
class EnsembleSupercells {
public:
    typedef typename PA::Particle Particle;
    typedef typename PA::ParticleRef ParticleRef;
    typedef typename PA::ConstParticleRef ConstParticleRef;
    typedef FP3 PositionType;
    typedef Int3 IndexType;
    static const int dimension = 3;

    EnsembleSupercells(PositionType minPosition, PositionType maxPosition, IndexType numCells,
                       IndexType numCellsPerSupercell)
        : min_(minPosition), max_(maxPosition), numCells_(numCells), ncps_(numCellsPerSupercell)
    {
        const FP3 cell = (maxPosition - minPosition) / FP3(numCells.x, numCells.y, numCells.z);
        const FP3 sc = FP3(numCellsPerSupercell.x, numCellsPerSupercell.y, numCellsPerSupercell.z) * cell;
        invSc_ = FP3(1.0 / sc.x, 1.0 / sc.y, 1.0 / sc.z);
        const Int3 n = getNumSupercells();
        cells_.resize((size_t)n.x * n.y * n.z);
        for (auto& a : cells_) a.dirty_ = &hostDirty_;
        hostValid_ = true;        // an empty ensemble: the (empty) mirror is current
    }
    EnsembleSupercells(const EnsembleSupercells&) = delete;
    EnsembleSupercells& operator=(const EnsembleSupercells&) = delete;

    PositionType getMinPosition() const { return min_; }
    PositionType getMaxPosition() const { return max_; }
    IndexType getNumCells() const { return numCells_; }
    IndexType getNumCellsPerSupercell() const { return ncps_; }
    IndexType getNumSupercells() const
    {
        return Int3((numCells_.x + ncps_.x - 1) / ncps_.x, (numCells_.y + ncps_.y - 1) / ncps_.y,
                    (numCells_.z + ncps_.z - 1) / ncps_.z);
    }
    // floor((x - min) * inverse(supercell size)) (EnsembleSupercells.h:53-57)
    template <class Ref>
    IndexType getSupercellIndex(const Ref& particle) const
    {
        return floor((particle.getPosition() - min_) * invSc_);
    }
    template <class Ref>
    void add(const Ref& particle)
    {
        pull();
        cells_[linear(getSupercellIndex(particle))].pushBack(particle);
    }
    // the particles of one supercell (a host mirror; changes through ParticleRef reach the GPU at
    // the next pushParticles)
    PA& getParticles(IndexType idx)
    {
        pull();
        hostDirty_ = true;
        return cells_[linear(idx)];
    }
    const PA& getParticles(IndexType idx) const
    {
        pull();
        return cells_[linear(idx)];
    }
    // the device ensemble, current; the caller is about to change it on the GPU
    pica_amd::EnsembleSupercells<double>& device()
    {
        flush();
        if (!ens_) create(0);
        hostValid_ = false;
        return *ens_;
    }
    // upload a modified host mirror now (also done before every GPU operation).  One species per
    // ensemble: the library steps an ensemble with one ParticleTypes entry.
    void flush()
    {
        if (!hostDirty_ && ens_) return;
        pica_amd::ParticleArraySoA p;
        bool first = true;
        for (const PA& a : cells_)
            for (int i = 0; i < a.size(); i++) {
                if (first) type_ = a.types[i];
                first = false;
                if (a.types[i] != type_)
                    throw pica_amd::Error(PK_E_UNSUPPORTED, "EnsembleSupercells: particles of one type per ensemble");
                p.pushBack(pica_amd::FP3(a.pos[0][i], a.pos[1][i], a.pos[2][i]),
                           pica_amd::FP3(a.ps[0][i], a.ps[1][i], a.ps[2][i]), a.factors[i]);
            }
        if (!ens_ || ensType_ != type_) create(type_);
        ens_->add(p);
        hostDirty_ = false;
    }

private:
    size_t linear(const Int3& i) const
    {
        const Int3 n = getNumSupercells();
        return ((size_t)i.x * n.y + i.y) * n.z + i.z;
    }
    void create(short type)
    {
        ens_.reset(new pica_amd::EnsembleSupercells<double>(
            pica_amd::defaultContext(), pica_amd::FP3(min_.x, min_.y, min_.z), pica_amd::FP3(max_.x, max_.y, max_.z),
            pica_amd::Int3(numCells_.x, numCells_.y, numCells_.z), pica_amd::Int3(ncps_.x, ncps_.y, ncps_.z), type));
        ensType_ = type;
    }
    void pull() const
    {
        if (hostValid_) return;
        std::vector<int64_t> off;
        pica_amd::ParticleArraySoA p = ens_->getParticles(&off);
        for (size_t s = 0; s + 1 < off.size(); s++) {
            PA& a = cells_[s];
            for (int d = 0; d < 3; d++) {
                a.pos[d].clear();
                a.ps[d].clear();
            }
            a.factors.clear();
            a.invGammas.clear();
            a.types.clear();
            for (int64_t i = off[s]; i < off[s + 1]; i++) {
                const double pos[3] = {p.x[i], p.y[i], p.z[i]}, u[3] = {p.px[i], p.py[i], p.pz[i]};
                for (int d = 0; d < 3; d++) {
                    a.pos[d].push_back(pos[d]);
                    a.ps[d].push_back(u[d]);
                }
                a.factors.push_back(p.factor[i]);
                a.invGammas.push_back(1.0 / std::sqrt(1.0 + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
                a.types.push_back(type_);
            }
        }
        hostValid_ = true;
    }

    FP3 min_, max_, invSc_;
    Int3 numCells_, ncps_;
    std::unique_ptr<pica_amd::EnsembleSupercells<double>> ens_;   // created at the first upload
    mutable std::vector<PA> cells_;
    mutable bool hostValid_ = false;
    bool hostDirty_ = false;
    short type_ = 0, ensType_ = 0;
};

template <class PA, EnsembleRepresentation rep>
struct Ensemble {};
template <class PA>
struct Ensemble<PA, EnsembleRepresentation_Supercells> {
    typedef EnsembleSupercells<PA> Type;
};

// ---- the replacement of the caller's particle loop: for every particle of every supercell,
// FieldInterpolatorCICSupercell::get, BorisPusher::push and CurrentDepositorCICSupercell::deposit
// (J accumulates into the grid), then the migration of particles that left their supercell.
// The grid must have pica's supercell origin -2 step (two periodic ghost layers); the species
// table is pica's ParticleTypes.
enum class FormFactor { CIC = PK_CIC, TSC = PK_TSC };
enum class Depositor { Direct = PK_DIRECT, Esirkepov = PK_ESIRKEPOV };

template <class PA>
void pushParticles(EnsembleSupercells<PA>& ens, YeeGrid<Three, double>& grid, double dt,
                   FormFactor ff = FormFactor::CIC, Depositor dep = Depositor::Direct)
{
    std::vector<double> m, q;
    for (int t = 0; t < ParticleTypes::numTypes; t++) {
        m.push_back(ParticleTypes::types[t].mass);
        q.push_back(ParticleTypes::types[t].charge);
    }
    pica_amd::Context& ctx = pica_amd::defaultContext();
    if (!m.empty()) ctx.setSpecies(m, q);
    pica_amd::pushParticles(ens.device(), grid.device(), dt, (pica_amd::FormFactor)ff, (pica_amd::Depositor)dep);
}

}  // namespace pica

#endif  // PICA_AMD_PICA_COMPAT_HPP
