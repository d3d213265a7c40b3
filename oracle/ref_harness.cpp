// ref_harness.cpp -- exposes pica's own header-only code through extern "C" entry points so
// that the oracle restatement (pica_oracle.c) can be checked against the real thing.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile with -I<reference>/include into
// oracle/_ref/libpica_ref.so, only where the pica sources are present.  Nothing of pica is
// copied here: every numerical operation below is a call into pica's templates.
//
// The only code of our own is glue pica does not have: copying arrays in and out of
// YeeGrid, periodic ghost fill/fold, wrap-around of particle positions, and the re-sort
// rule between steps (the same definitions as pica_oracle.c).
// pica's headers rely on their includers for some declarations, so the base headers go first.
#include "pica/math/Constants.h"
#include "pica/math/Vectors.h"
#include "pica/particles/Particle.h"
#include "pica/particles/ParticleArray.h"
#include "pica/particles/Ensemble.h"
#include "pica/grid/YeeGrid.h"
#include "pica/fieldSolver/YeeSolver.h"
#include "pica/fieldInterpolation/FieldInterpolator.h"
#include "pica/currentDeposition/CurrentDepositor.h"
#include "pica/particlePush/BorisPusher.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace pica {
namespace ParticleTypes {
std::vector<ParticleType> typesVector;
const ParticleType* types = 0;
int numTypes = 0;
}  // namespace ParticleTypes
}  // namespace pica

using namespace pica;

typedef YeeGrid<Three, double> Grid3;
typedef YeeGrid<One, double> Grid1;
typedef YeeGrid<Two, double> Grid2;
typedef YeeGrid<Three, float> Grid3f;
typedef ParticleArraySoA<Particle3d> PArray;
typedef Ensemble<PArray, EnsembleRepresentation_Supercells>::Type PEnsemble;

namespace {

template<class G, class T>
void load_grid(G& g, T* const f[9])
{
    const Int3 s = g.getSize();
    for (int i = 0; i < s.x; i++)
        for (int j = 0; j < s.y; j++)
            for (int k = 0; k < s.z; k++) {
                const long c = ((long)i * s.y + j) * s.z + k;
                g.ex(i, j, k) = f[0][c]; g.ey(i, j, k) = f[1][c]; g.ez(i, j, k) = f[2][c];
                g.bx(i, j, k) = f[3][c]; g.by(i, j, k) = f[4][c]; g.bz(i, j, k) = f[5][c];
                g.jx(i, j, k) = f[6][c]; g.jy(i, j, k) = f[7][c]; g.jz(i, j, k) = f[8][c];
            }
}

template<class G, class T>
void store_grid(const G& g, T* const f[9])
{
    const Int3 s = g.getSize();
    for (int i = 0; i < s.x; i++)
        for (int j = 0; j < s.y; j++)
            for (int k = 0; k < s.z; k++) {
                const long c = ((long)i * s.y + j) * s.z + k;
                f[0][c] = g.ex(i, j, k); f[1][c] = g.ey(i, j, k); f[2][c] = g.ez(i, j, k);
                f[3][c] = g.bx(i, j, k); f[4][c] = g.by(i, j, k); f[5][c] = g.bz(i, j, k);
                f[6][c] = g.jx(i, j, k); f[7][c] = g.jy(i, j, k); f[8][c] = g.jz(i, j, k);
            }
}

// reduced-dimension grids: arrays of s[0] (1D) or s[0] x s[1] (2D, y fastest) values
void load_grid1(Grid1& g, double* const f[9])
{
    const int n = g.getSize().x;
    for (int i = 0; i < n; i++) {
        g.ex(i) = f[0][i]; g.ey(i) = f[1][i]; g.ez(i) = f[2][i];
        g.bx(i) = f[3][i]; g.by(i) = f[4][i]; g.bz(i) = f[5][i];
        g.jx(i) = f[6][i]; g.jy(i) = f[7][i]; g.jz(i) = f[8][i];
    }
}

void store_grid1(const Grid1& g, double* const f[9])
{
    const int n = g.getSize().x;
    for (int i = 0; i < n; i++) {
        f[0][i] = g.ex(i); f[1][i] = g.ey(i); f[2][i] = g.ez(i);
        f[3][i] = g.bx(i); f[4][i] = g.by(i); f[5][i] = g.bz(i);
        f[6][i] = g.jx(i); f[7][i] = g.jy(i); f[8][i] = g.jz(i);
    }
}

void load_grid2(Grid2& g, double* const f[9])
{
    const Grid2::IndexType s = g.getSize();
    for (int i = 0; i < s.x; i++)
        for (int j = 0; j < s.y; j++) {
            const long c = (long)i * s.y + j;
            g.ex(i, j) = f[0][c]; g.ey(i, j) = f[1][c]; g.ez(i, j) = f[2][c];
            g.bx(i, j) = f[3][c]; g.by(i, j) = f[4][c]; g.bz(i, j) = f[5][c];
            g.jx(i, j) = f[6][c]; g.jy(i, j) = f[7][c]; g.jz(i, j) = f[8][c];
        }
}

void store_grid2(const Grid2& g, double* const f[9])
{
    const Grid2::IndexType s = g.getSize();
    for (int i = 0; i < s.x; i++)
        for (int j = 0; j < s.y; j++) {
            const long c = (long)i * s.y + j;
            f[0][c] = g.ex(i, j); f[1][c] = g.ey(i, j); f[2][c] = g.ez(i, j);
            f[3][c] = g.bx(i, j); f[4][c] = g.by(i, j); f[5][c] = g.bz(i, j);
            f[6][c] = g.jx(i, j); f[7][c] = g.jy(i, j); f[8][c] = g.jz(i, j);
        }
}

Grid1 make_grid1(const int s[1], const double origin[1], const double step[1])
{
    return Grid1(Grid1::PositionType(origin[0]), Grid1::PositionType(step[0]), Grid1::IndexType(s[0]));
}

Grid2 make_grid2(const int s[2], const double origin[2], const double step[2])
{
    return Grid2(Grid2::PositionType(origin[0], origin[1]), Grid2::PositionType(step[0], step[1]),
                 Grid2::IndexType(s[0], s[1]));
}

void set_species(double q, double m)
{
    ParticleTypes::typesVector.resize(1);
    ParticleTypes::typesVector[0].charge = q;
    ParticleTypes::typesVector[0].mass = m;
    ParticleTypes::types = &ParticleTypes::typesVector[0];
    ParticleTypes::numTypes = 1;
}

Particle3d make_particle(double x, double y, double z, double ux, double uy, double uz, double w)
{
    Particle3d p;
    p.setPosition(FP3(x, y, z));
    p.setP(FP3(ux, uy, uz));   // exact u, no round trip through setMomentum
    p.setFactor(w);
    p.setType(0);
    return p;
}

// periodic ghosts of one Array3d, width G (same definition as or_ghost_fill / or_ghost_fold)
template<class A>
void ghost_fill(A& a, int G)
{
    const Int3 s = a.getSize();
    for (int d = 0; d < 3; d++) {
        const int N = s[d] - 2 * G;
        const int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
        for (int q = 0; q < G; q++)
            for (int p1 = 0; p1 < s[o1]; p1++)
                for (int p2 = 0; p2 < s[o2]; p2++) {
                    Int3 lo, src, hi, src2;
                    lo[d] = q; src[d] = q + N; hi[d] = G + N + q; src2[d] = G + q;
                    lo[o1] = src[o1] = hi[o1] = src2[o1] = p1;
                    lo[o2] = src[o2] = hi[o2] = src2[o2] = p2;
                    a(lo) = a(src);
                    a(hi) = a(src2);
                }
    }
}

template<class A>
void ghost_fold(A& a, int G)
{
    const Int3 s = a.getSize();
    for (int d = 0; d < 3; d++) {
        const int N = s[d] - 2 * G;
        const int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
        for (int p1 = 0; p1 < s[o1]; p1++)
            for (int p2 = 0; p2 < s[o2]; p2++) {
                for (int q = 0; q < G; q++) {
                    Int3 t, g; t[d] = q + N; g[d] = q; t[o1] = g[o1] = p1; t[o2] = g[o2] = p2;
                    a(t) += a(g);
                }
                for (int q = 0; q < G; q++) {
                    Int3 t, g; t[d] = G + q; g[d] = G + N + q; t[o1] = g[o1] = p1; t[o2] = g[o2] = p2;
                    a(t) += a(g);
                }
            }
    }
}

}  // namespace

extern "C" {

int ref_version(void) { return 1; }

// YeeSolver::updateE / updateB on a YeeGrid<Three, double> (YeeSolver.h:28-38)
void ref_yee_update_e(const int s[3], const double origin[3], const double step[3], double* f[9], double dt)
{
    Grid3 g(FP3(origin[0], origin[1], origin[2]), FP3(step[0], step[1], step[2]), Int3(s[0], s[1], s[2]));
    load_grid(g, f);
    YeeSolver solver;
    solver.updateE(g, dt);
    store_grid(g, f);
}

void ref_yee_update_b(const int s[3], const double origin[3], const double step[3], double* f[9], double dt)
{
    Grid3 g(FP3(origin[0], origin[1], origin[2]), FP3(step[0], step[1], step[2]), Int3(s[0], s[1], s[2]));
    load_grid(g, f);
    YeeSolver solver;
    solver.updateB(g, dt);
    store_grid(g, f);
}

void ref_yee_update_e_f32(const int s[3], const float origin[3], const float step[3], float* f[9], float dt)
{
    Grid3f g(Vector3<float>(origin[0], origin[1], origin[2]), Vector3<float>(step[0], step[1], step[2]),
             Int3(s[0], s[1], s[2]));
    load_grid(g, f);
    YeeSolver solver;
    solver.updateE(g, dt);
    store_grid(g, f);
}

void ref_yee_update_b_f32(const int s[3], const float origin[3], const float step[3], float* f[9], float dt)
{
    Grid3f g(Vector3<float>(origin[0], origin[1], origin[2]), Vector3<float>(step[0], step[1], step[2]),
             Int3(s[0], s[1], s[2]));
    load_grid(g, f);
    YeeSolver solver;
    solver.updateB(g, dt);
    store_grid(g, f);
}

// YeeSolver on pica's reduced-dimension grids (YeeSolver.h:41-128); op 0 = updateE, 1 = updateB
void ref_yee_update_1d(const int s[1], const double origin[1], const double step[1], double* f[9], double dt, int op)
{
    Grid1 g = make_grid1(s, origin, step);
    load_grid1(g, f);
    YeeSolver solver;
    if (op == 0) solver.updateE(g, dt);
    else solver.updateB(g, dt);
    store_grid1(g, f);
}

void ref_yee_update_2d(const int s[2], const double origin[2], const double step[2], double* f[9], double dt, int op)
{
    Grid2 g = make_grid2(s, origin, step);
    load_grid2(g, f);
    YeeSolver solver;
    if (op == 0) solver.updateE(g, dt);
    else solver.updateB(g, dt);
    store_grid2(g, f);
}

// FieldInterpolatorCIC<YeeGrid<One,double>>::get (FieldInterpolatorYeeGrid.h:51-83).  The Two
// specialisation (:86-133) does not compile: its get() builds PositionType (a Vector2) from
// coeffStaggered.z / coeffCollocated.z (:108-118), members a Vector2 lacks.
void ref_interp_cic_1d(const int s[1], const double origin[1], const double step[1], double* f[9], long n,
                       const double* pos, double* e, double* b)
{
    Grid1 g = make_grid1(s, origin, step);
    load_grid1(g, f);
    FieldInterpolatorCIC<Grid1> interp(g);
    for (long p = 0; p < n; p++) {
        Vector3<double> ev, bv;
        interp.get(Grid1::PositionType(pos[p]), ev, bv);
        e[3 * p] = ev.x; e[3 * p + 1] = ev.y; e[3 * p + 2] = ev.z;
        b[3 * p] = bv.x; b[3 * p + 1] = bv.y; b[3 * p + 2] = bv.z;
    }
}

// CurrentDepositorCIC<YeeGrid<One|Two,double>>::deposit (CurrentDepositorYeeGrid.h:52-122); cur is n x 3
void ref_deposit_cic_1d(const int s[1], const double origin[1], const double step[1], double* f[9], long n,
                        const double* pos, const double* cur)
{
    Grid1 g = make_grid1(s, origin, step);
    load_grid1(g, f);
    CurrentDepositorCIC<Grid1> dep(g);
    for (long p = 0; p < n; p++)
        dep.deposit(Grid1::PositionType(pos[p]), Vector3<double>(cur[3 * p], cur[3 * p + 1], cur[3 * p + 2]));
    store_grid1(g, f);
}

void ref_deposit_cic_2d(const int s[2], const double origin[2], const double step[2], double* f[9], long n,
                        const double* pos, const double* cur)
{
    Grid2 g = make_grid2(s, origin, step);
    load_grid2(g, f);
    CurrentDepositorCIC<Grid2> dep(g);
    for (long p = 0; p < n; p++)
        dep.deposit(Grid2::PositionType(pos[2 * p], pos[2 * p + 1]),
                    Vector3<double>(cur[3 * p], cur[3 * p + 1], cur[3 * p + 2]));
    store_grid2(g, f);
}

// FieldInterpolatorCIC<YeeGrid<Three,double>>::get for n positions
void ref_interp_cic(const int s[3], const double origin[3], const double step[3], double* f[9], long n,
                    const double* x, const double* y, const double* z, double* e, double* b)
{
    Grid3 g(FP3(origin[0], origin[1], origin[2]), FP3(step[0], step[1], step[2]), Int3(s[0], s[1], s[2]));
    load_grid(g, f);
    FieldInterpolatorCIC<Grid3> interp(g);
    for (long p = 0; p < n; p++) {
        Vector3<double> ev, bv;
        interp.get(FP3(x[p], y[p], z[p]), ev, bv);
        e[3 * p] = ev.x; e[3 * p + 1] = ev.y; e[3 * p + 2] = ev.z;
        b[3 * p] = bv.x; b[3 * p + 1] = bv.y; b[3 * p + 2] = bv.z;
    }
}

// FieldInterpolatorCICSupercell<double> (gather + get) for n positions of one supercell
void ref_interp_cic_sc(const int s[3], const double origin[3], const double step[3], double* f[9],
                       const double scmin[3], const int ncps[3], long n, const double* x, const double* y,
                       const double* z, double* e, double* b)
{
    Grid3 g(FP3(origin[0], origin[1], origin[2]), FP3(step[0], step[1], step[2]), Int3(s[0], s[1], s[2]));
    load_grid(g, f);
    FieldInterpolatorCICSupercell<double> interp(g, FP3(scmin[0], scmin[1], scmin[2]),
                                                 Int3(ncps[0], ncps[1], ncps[2]));
    for (long p = 0; p < n; p++) {
        Vector3<double> ev, bv;
        interp.get(FP3(x[p], y[p], z[p]), ev, bv);
        e[3 * p] = ev.x; e[3 * p + 1] = ev.y; e[3 * p + 2] = ev.z;
        b[3 * p] = bv.x; b[3 * p + 1] = bv.y; b[3 * p + 2] = bv.z;
    }
}

// CurrentDepositorCIC<YeeGrid<Three,double>>::deposit, cur is n x 3
void ref_deposit_cic(const int s[3], const double origin[3], const double step[3], double* f[9], long n,
                     const double* x, const double* y, const double* z, const double* cur)
{
    Grid3 g(FP3(origin[0], origin[1], origin[2]), FP3(step[0], step[1], step[2]), Int3(s[0], s[1], s[2]));
    load_grid(g, f);
    CurrentDepositorCIC<Grid3> dep(g);
    for (long p = 0; p < n; p++)
        dep.deposit(FP3(x[p], y[p], z[p]), Vector3<double>(cur[3 * p], cur[3 * p + 1], cur[3 * p + 2]));
    store_grid(g, f);
}

// CurrentDepositorCICSupercell<double>: zero tile, deposit, flush on destruction
void ref_deposit_cic_sc(const int s[3], const double origin[3], const double step[3], double* f[9],
                        const double scmin[3], const int ncps[3], long n, const double* x, const double* y,
                        const double* z, const double* cur)
{
    Grid3 g(FP3(origin[0], origin[1], origin[2]), FP3(step[0], step[1], step[2]), Int3(s[0], s[1], s[2]));
    load_grid(g, f);
    {
        CurrentDepositorCICSupercell<double> dep(g, FP3(scmin[0], scmin[1], scmin[2]),
                                                 Int3(ncps[0], ncps[1], ncps[2]));
        for (long p = 0; p < n; p++)
            dep.deposit(FP3(x[p], y[p], z[p]), Vector3<double>(cur[3 * p], cur[3 * p + 1], cur[3 * p + 2]));
    }
    store_grid(g, f);
}

// BorisPusher<Particle3d,double>::push on a ParticleArraySoA; e, b are n x 3
void ref_boris_push(long n, double q, double m, double dt, double* x, double* y, double* z, double* ux,
                    double* uy, double* uz, const double* w, const double* e, const double* b)
{
    set_species(q, m);
    PArray pa;
    for (long p = 0; p < n; p++) pa.pushBack(make_particle(x[p], y[p], z[p], ux[p], uy[p], uz[p], w[p]));
    BorisPusher<Particle3d, double> pusher(dt);
    for (long p = 0; p < n; p++) {
        PArray::ParticleRef ref = pa[(int)p];
        pusher.push(&ref, Vector3<double>(e[3 * p], e[3 * p + 1], e[3 * p + 2]),
                    Vector3<double>(b[3 * p], b[3 * p + 1], b[3 * p + 2]));
        const FP3 pos = ref.getPosition();
        const FP3 u = ref.getP();
        x[p] = pos.x; y[p] = pos.y; z[p] = pos.z;
        ux[p] = u.x; uy[p] = u.y; uz[p] = u.z;
    }
}

// Grid::getCellIndex (Grid.h:36)
void ref_cell_index(const int s[3], const double origin[3], const double step[3], long n, const double* x,
                    const double* y, const double* z, int* out)
{
    Grid3 g(FP3(origin[0], origin[1], origin[2]), FP3(step[0], step[1], step[2]), Int3(s[0], s[1], s[2]));
    for (long p = 0; p < n; p++) {
        const Int3 c = g.getCellIndex(FP3(x[p], y[p], z[p]));
        out[3 * p] = c.x; out[3 * p + 1] = c.y; out[3 * p + 2] = c.z;
    }
}

// EnsembleSupercells::getSupercellIndex (EnsembleSupercells.h:53-57)
void ref_supercell_index(const double mn[3], const double mx[3], const int ncell[3], const int ncps[3], long n,
                         const double* x, const double* y, const double* z, int* out)
{
    set_species(-4.80320427e-10, 9.10938215e-28);
    PEnsemble ens(FP3(mn[0], mn[1], mn[2]), FP3(mx[0], mx[1], mx[2]), Int3(ncell[0], ncell[1], ncell[2]),
                  Int3(ncps[0], ncps[1], ncps[2]));
    for (long p = 0; p < n; p++) {
        const Particle3d pt = make_particle(x[p], y[p], z[p], 0, 0, 0, 1);
        const Int3 c = ens.getSupercellIndex(pt);
        out[3 * p] = c.x; out[3 * p + 1] = c.y; out[3 * p + 2] = c.z;
    }
}

// A periodic PIC run assembled from pica's own supercell interpolator, Boris pusher, supercell
// CIC depositor and Yee solver.  Grid: ncell + 2*2 ghosts, origin -2 step, physical domain
// [0, ncell*step).  Particles (any order) are bucketed with EnsembleSupercells::add.  Each step:
// J = 0; colours c = 0..7 of supercells, per supercell get/push/deposit(x_new, v*q*w) and the
// depositor's flush; fold J; B(dt/2), E(dt), B(dt/2) with ghost fills; wrap + re-sort.
// On return, particles are written back in supercell order (off has nsc+1 entries) and the
// grid arrays (incl. ghosts) are stored in f.
int ref_run_cic_direct(const int ncell[3], const double step[3], const int ncps[3], double q, double m, double dt,
                       int nsteps, long np, double* x, double* y, double* z, double* ux, double* uy, double* uz,
                       double* w, long* off, double* f[9], int nthreads)
{
    const int G = 2;
    set_species(q, m);
    const Int3 nc(ncell[0], ncell[1], ncell[2]), nps(ncps[0], ncps[1], ncps[2]);
    const FP3 h(step[0], step[1], step[2]);
    const FP3 L(ncell[0] * step[0], ncell[1] * step[1], ncell[2] * step[2]);
    Grid3 g(FP3(-2 * step[0], -2 * step[1], -2 * step[2]), h, Int3(nc.x + 2 * G, nc.y + 2 * G, nc.z + 2 * G));
    load_grid(g, f);
    const Int3 nsc((nc.x + nps.x - 1) / nps.x, (nc.y + nps.y - 1) / nps.y, (nc.z + nps.z - 1) / nps.z);
    PEnsemble* ens = new PEnsemble(FP3(0, 0, 0), L, nc, nps);
    for (long p = 0; p < np; p++) ens->add(make_particle(x[p], y[p], z[p], ux[p], uy[p], uz[p], w[p]));
    const long nsct = (long)nsc.x * nsc.y * nsc.z;
    std::vector<std::vector<signed char> > cls(nsct);
    YeeSolver solver;
    BorisPusher<Particle3d, double> pusher(dt);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    int err = 0;
    for (int it = 0; it < nsteps; it++) {
        const Int3 gs = g.getSize();
        for (int i = 0; i < gs.x; i++)
            for (int j = 0; j < gs.y; j++)
                for (int k = 0; k < gs.z; k++) g.jx(i, j, k) = g.jy(i, j, k) = g.jz(i, j, k) = 0.0;
        // colours as in or_particle_stage: supercells p = ceil((n+4)/n) apart never share tile cells
        int pc[3];
        for (int d = 0; d < 3; d++) pc[d] = (nps[d] + 4 + nps[d] - 1) / nps[d];
        for (int colour = 0; colour < pc[0] * pc[1] * pc[2]; colour++) {
            #pragma omp parallel for schedule(dynamic, 4)
            for (long s = 0; s < nsct; s++) {
                const int sx = (int)(s / ((long)nsc.y * nsc.z));
                const int sy = (int)((s / nsc.z) % nsc.y);
                const int sz = (int)(s % nsc.z);
                if (((sx % pc[0]) * pc[1] + sy % pc[1]) * pc[2] + sz % pc[2] != colour) continue;
                const FP3 scMin(sx * nps.x * h.x, sy * nps.y * h.y, sz * nps.z * h.z);
                FieldInterpolatorCICSupercell<double> interp(g, scMin, nps);
                CurrentDepositorCICSupercell<double> dep(g, scMin, nps);
                PArray& pa = ens->getParticles(Int3(sx, sy, sz));
                cls[s].assign(pa.size(), 13);
                for (int i = 0; i < pa.size(); i++) {
                    PArray::ParticleRef ref = pa[i];
                    Vector3<double> ev, bv;
                    interp.get(ref.getPosition(), ev, bv);
                    pusher.push(&ref, ev, bv);
                    const Vector3<double> cur = ref.getVelocity() * (ref.getCharge() * ref.getFactor());
                    dep.deposit(ref.getPosition(), cur);
                    // glue: periodic wrap and direction class (as or_wrap_dir)
                    FP3 pos = ref.getPosition();
                    const int src[3] = {sx, sy, sz};
                    int code = 0;
                    for (int d = 0; d < 3; d++) {
                        const double cellSize = (L[d] - 0.0) / (double)nc[d];
                        const double inv = 1.0 / ((double)nps[d] * cellSize);
                        if (pos[d] < 0.0) pos[d] += L[d];
                        if (pos[d] >= L[d]) pos[d] -= L[d];
                        int dst = (int)std::floor((pos[d] - 0.0) * inv);
                        if (dst < 0) dst = 0;
                        if (dst > nsc[d] - 1) dst = nsc[d] - 1;
                        int dd = dst - src[d];
                        if (dd > 1) dd -= nsc[d];
                        if (dd < -1) dd += nsc[d];
                        if (dd < -1 || dd > 1) err = 1;
                        code = code * 3 + (dd + 1);
                    }
                    ref.setPosition(pos);
                    cls[s][i] = (signed char)code;
                }
            }
        }
        if (err) break;
        ghost_fold(g.jx(), G); ghost_fold(g.jy(), G); ghost_fold(g.jz(), G);
        ghost_fill(g.jx(), G); ghost_fill(g.jy(), G); ghost_fill(g.jz(), G);
        solver.updateB(g, dt / 2);
        ghost_fill(g.bx(), G); ghost_fill(g.by(), G); ghost_fill(g.bz(), G);
        solver.updateE(g, dt);
        ghost_fill(g.ex(), G); ghost_fill(g.ey(), G); ghost_fill(g.ez(), G);
        solver.updateB(g, dt / 2);
        ghost_fill(g.bx(), G); ghost_fill(g.by(), G); ghost_fill(g.bz(), G);
        // re-sort (as or_resort): stayers keep slots < nfinal, holes there are filled by the
        // stayers found at slots >= nfinal; then movers by direction class from s - delta(class)
        PEnsemble* next = new PEnsemble(FP3(0, 0, 0), L, nc, nps);
        for (long s = 0; s < nsct; s++) {
            const int sx = (int)(s / ((long)nsc.y * nsc.z));
            const int sy = (int)((s / nsc.z) % nsc.y);
            const int sz = (int)(s % nsc.z);
            PArray& own = ens->getParticles(Int3(sx, sy, sz));
            PArray& dstArr = next->getParticles(Int3(sx, sy, sz));
            const int cnt = own.size();
            int nmove = 0;
            for (int i = 0; i < cnt; i++) nmove += (cls[s][i] != 13);
            const int nfinal = cnt - nmove;
            std::vector<int> tail;
            for (int i = nfinal; i < cnt; i++)
                if (cls[s][i] == 13) tail.push_back(i);
            int h = 0;
            for (int i = 0; i < nfinal; i++)
                dstArr.pushBack(own[cls[s][i] == 13 ? i : tail[h++]]);
            for (int c = 0; c < 27; c++) {
                if (c == 13) continue;
                int src[3] = {sx - (c / 9 - 1), sy - ((c / 3) % 3 - 1), sz - (c % 3 - 1)};
                for (int d = 0; d < 3; d++) {
                    if (src[d] < 0) src[d] += nsc[d];
                    if (src[d] >= nsc[d]) src[d] -= nsc[d];
                }
                const long ss = ((long)src[0] * nsc.y + src[1]) * nsc.z + src[2];
                PArray& from = ens->getParticles(Int3(src[0], src[1], src[2]));
                for (int i = 0; i < from.size(); i++)
                    if (cls[ss][i] == c) dstArr.pushBack(from[i]);
            }
        }
        delete ens;
        ens = next;
    }
    long k = 0;
    off[0] = 0;
    for (long s = 0; s < nsct; s++) {
        const int sx = (int)(s / ((long)nsc.y * nsc.z));
        const int sy = (int)((s / nsc.z) % nsc.y);
        const int sz = (int)(s % nsc.z);
        const PArray& pa = ens->getParticles(Int3(sx, sy, sz));
        for (int i = 0; i < pa.size(); i++, k++) {
            const FP3 pos = pa[i].getPosition();
            const FP3 u = pa[i].getP();
            x[k] = pos.x; y[k] = pos.y; z[k] = pos.z;
            ux[k] = u.x; uy[k] = u.y; uz[k] = u.z;
            w[k] = pa[i].getFactor();
        }
        off[s + 1] = k;
    }
    delete ens;
    store_grid(g, f);
    return err ? -4 : 0;
}

}  // extern "C"
