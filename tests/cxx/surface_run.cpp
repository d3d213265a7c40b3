// Drives the C++ surface (include/pica_amd/pica.hpp) the way a pica caller would: builds a
// periodic YeeGrid and an EnsembleSupercells, adds particles, runs steps, writes the result.
// Halfway it saves both to snapshot files and continues on fresh objects loaded from them,
// stepped through the several-species call (with one species).
// usage: surface_run <in.bin> <out.bin> <nsteps> <shape 0|1> <depositor 0|1>
// in.bin : int32 ncell[3], int32 ncps[3], int64 n, then n x {x,y,z,ux,uy,uz,w} doubles
// out.bin: int64 n, n x 7 doubles (supercell order), int64 nsc+1 offsets, 9 grid components
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "pica_amd/pica.hpp"

using namespace pica_amd;

int main(int argc, char** argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s in.bin out.bin nsteps shape depositor\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    int nc[3], np[3];
    int64_t n = 0;
    if (std::fread(nc, 4, 3, f) != 3 || std::fread(np, 4, 3, f) != 3 || std::fread(&n, 8, 1, f) != 1) return 4;
    std::vector<double> buf((size_t)n * 7);
    if (std::fread(buf.data(), 8, buf.size(), f) != buf.size()) return 5;
    std::fclose(f);
    try {
        Context ctx(0);
        const Int3 ncell(nc[0], nc[1], nc[2]);
        YeeGrid<double> grid = YeeGrid<double>::periodic(ctx, ncell, FP3(1, 1, 1));
        EnsembleSupercells<double> ens(ctx, FP3(0, 0, 0), FP3(nc[0], nc[1], nc[2]), ncell, Int3(np[0], np[1], np[2]));
        ParticleArraySoA p;
        for (int64_t i = 0; i < n; i++) {
            const double* r = &buf[(size_t)i * 7];
            p.pushBack(FP3(r[0], r[1], r[2]), FP3(r[3], r[4], r[5]), r[6]);
        }
        ens.add(p);
        const double dt = 0.5 / 29979245800.0;
        const FormFactor ff = (FormFactor)std::atoi(argv[4]);
        const Depositor dep = (Depositor)std::atoi(argv[5]);
        const int nsteps = std::atoi(argv[3]), first = nsteps / 2;
        if (first > 0) step(ens, grid, dt, ff, dep, first);
        const std::string snap = std::string(argv[2]);
        grid.save(snap + ".grid");
        ens.save(snap + ".ens");
        YeeGrid<double> grid2 = YeeGrid<double>::periodic(ctx, ncell, FP3(1, 1, 1));
        EnsembleSupercells<double> ens2(ctx, FP3(0, 0, 0), FP3(nc[0], nc[1], nc[2]), ncell, Int3(np[0], np[1], np[2]));
        grid2.load(snap + ".grid");
        ens2.load(snap + ".ens");
        // options and the supercell interleave (a no-op for FP64 ensembles: the results stay the oracle's)
        ctx.setOption(PK_OPT_INTERLEAVE, ctx.getOption(PK_OPT_INTERLEAVE));
        ens2.interleave();
        std::vector<EnsembleSupercells<double>*> species{&ens2};
        for (int s = first; s < nsteps; s++) step(species, grid2, dt, ff, dep);
        std::vector<int64_t> off;
        ParticleArraySoA q = ens2.getParticles(&off);
        FILE* o = std::fopen(argv[2], "wb");
        if (!o) return 6;
        const int64_t m = (int64_t)q.size();
        std::fwrite(&m, 8, 1, o);
        for (size_t i = 0; i < q.size(); i++) {
            const double r[7] = {q.x[i], q.y[i], q.z[i], q.px[i], q.py[i], q.pz[i], q.factor[i]};
            std::fwrite(r, 8, 7, o);
        }
        std::fwrite(off.data(), 8, off.size(), o);
        for (int c = Ex; c <= Jz; c++) {
            std::vector<double> a = grid2.copyOut((Component)c);
            std::fwrite(a.data(), 8, a.size(), o);
        }
        std::fclose(o);
        const Diagnostics d = diagnostics(ens2, grid2, ff);
        std::printf("particles %lld field energy %.6e kinetic %.6e\n", (long long)m, d.fieldEnergy, d.kineticEnergy);
    } catch (const Error& e) {
        std::fprintf(stderr, "pica_amd error %d: %s\n", e.code(), e.what());
        return 1;
    }
    return 0;
}
